"""ctypes binding of the oracle library (oracle/cpu_ref.c).

Test infrastructure only: the checker for the HIP path, never the product.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {"plain": os.path.join(_HERE, "libosbl_oracle.so"), "fma": os.path.join(_HERE, "libosbl_oracle_fma.so")}

CONV = {"central_skew": 0, "central_cons": 1, "weno_js": 2, "weno_z": 3, "teno5": 4, "teno6": 5}
AVG = {"simple": 0, "roe": 1}
VISC = {"none": 0, "const": 1, "sutherland": 2}
RK = {"rk3": 0, "ls3": 1, "ls3_ssp": 2, "ls4": 3}
BC = {"periodic": 0, "extrap0": 1, "extrap1": 2, "zero_gradient": 3, "adiabatic_wall": 4,
      "isothermal_wall": 5, "dirichlet": 6, "inlet_pextrap": 7, "symmetry": 8}
BSCHEME = {"carpenter": 0, "reduced_access": 1}
FILTERS = {"weno": 1, "sfd": 2, "binomial": 4}
WF_SENSOR = {"ducros": 0, "constant": 1}


# parity bar: relative L2 difference per conservative variable (FP64)
TOL = 1e-10
# a momentum component whose norm is below this fraction of the momentum
# magnitude (e.g. the Taylor-Green vortex's rho*w, zero at t = 0) is measured
# against the momentum magnitude instead of its own norm
SMALL_COMPONENT = 0.1


def parity_errors(q, qr):
    """Relative L2 difference of q against the reference qr per conservative
    variable (leading axis: rho, rho*u_0.., rho*E; nvar = ndim + 2).  A
    momentum component whose reference norm is below SMALL_COMPONENT times the
    momentum magnitude sqrt(sum_i |rho*u_i|^2) is divided by that magnitude:
    rounding in such a component scales with the whole momentum field, not
    with the component's own (small or zero) size."""
    q = np.asarray(q)
    qr = np.asarray(qr)
    nv = qr.shape[0]
    err = np.empty(nv)
    for m in range(nv):
        den = np.linalg.norm(qr[m])
        num = np.linalg.norm(q[m] - qr[m])
        err[m] = num / den if den > 0 else num
    if nv >= 3:
        mom = np.sqrt(sum(np.linalg.norm(qr[m]) ** 2 for m in range(1, nv - 1)))
        for m in range(1, nv - 1):
            if mom > 0 and np.linalg.norm(qr[m]) < SMALL_COMPONENT * mom:
                err[m] = np.linalg.norm(q[m] - qr[m]) / mom
    return err


def filter_fields(p):
    """Filter settings of a problem dict: (bitmask, binomial order, WENO-filter
    sensor, Mach correction, sigma, chi, omega, control)."""
    f = p.get("filters", {}) or {}
    mask = 0
    for k in f:
        mask |= FILTERS[k]
    b = f.get("binomial", {}) or {}
    s = f.get("sfd", {}) or {}
    w = f.get("weno", {}) or {}
    return (mask, int(b.get("order", 2)), WF_SENSOR[w.get("sensor", "ducros")], 1 if w.get("mach_correction") else 0,
            float(b.get("sigma", 0.05)), float(s.get("chi", 0.1)), float(s.get("omega", 1.0 / 0.75)),
            float(w.get("control", 1.0)))


class RefConfig(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int32),
        ("n", ctypes.c_int32 * 3),
        ("delta", ctypes.c_double * 3),
        ("conv", ctypes.c_int32),
        ("averaging", ctypes.c_int32),
        ("viscous", ctypes.c_int32),
        ("rk", ctypes.c_int32),
        ("bc", (ctypes.c_int32 * 2) * 3),
        ("gama", ctypes.c_double),
        ("Minf", ctypes.c_double),
        ("Re", ctypes.c_double),
        ("Pr", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("eps", ctypes.c_double),
        ("teno_ct", ctypes.c_double),
        ("bscheme", (ctypes.c_int32 * 2) * 3),
        ("stretch", ctypes.c_int32 * 3),
        ("filters", ctypes.c_int32),
        ("suthT", ctypes.c_double),
        ("refT", ctypes.c_double),
        ("twall", ctypes.c_double),
        ("sensor", ctypes.c_int32),
        ("curvilinear", ctypes.c_int32),
        ("teno_a1", ctypes.c_double),
        ("teno_a2", ctypes.c_double),
        ("sensor_eps", ctypes.c_double),
        ("binomial_order", ctypes.c_int32),
        ("wf_sensor", ctypes.c_int32),
        ("wf_mach", ctypes.c_int32),
        ("pad3_", ctypes.c_int32),
        ("sigma_filt", ctypes.c_double),
        ("chi_filt", ctypes.c_double),
        ("omega_filt", ctypes.c_double),
        ("shock_filter_control", ctypes.c_double),
        ("body_force", ctypes.c_double * 3),
    ]


def build(force=False):
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not all(os.path.exists(p) for p in _LIBS.values()):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "all"], check=True, stdout=subprocess.DEVNULL)
    return _LIBS["plain"]


_libs = {}


def lib(variant="plain"):
    """The oracle library; variant "fma" is the same source built with FMA contraction."""
    if variant not in _libs:
        if not os.path.exists(_LIBS[variant]):
            build()
        L = ctypes.CDLL(_LIBS[variant])
        dp = ctypes.POINTER(ctypes.c_double)
        L.ref_create.argtypes = [ctypes.POINTER(RefConfig), ctypes.POINTER(ctypes.c_void_p)]
        L.ref_destroy.argtypes = [ctypes.c_void_p]
        L.ref_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.ref_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.ref_rhs.argtypes = [ctypes.c_void_p, dp]
        L.ref_run.argtypes = [ctypes.c_void_p, ctypes.c_int64]
        L.ref_recon_side.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, ctypes.c_double]
        L.ref_recon_side.restype = ctypes.c_double
        L.ref_smoothness.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.ref_eigensystem.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp,
                                      ctypes.c_double, dp, dp, dp]
        L.ref_roe_average.argtypes = [ctypes.c_int, ctypes.c_double, dp, dp, dp]
        L.ref_central_weights.argtypes = [dp, dp]
        L.ref_rk_coefficients.argtypes = [ctypes.c_int, dp, dp]
        L.ref_set_coordinates.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.ref_set_bc_plane.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, dp]
        L.ref_closure_weights.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.ref_get_metric.argtypes = [ctypes.c_void_p, ctypes.c_int, dp, dp]
        L.ref_get_sensor.argtypes = [ctypes.c_void_p, dp]
        L.ref_set_filter_mask.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.ref_get_sfd_field.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.ref_set_sfd_field.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.ref_weno_filter_flux.argtypes = [dp, dp, dp, dp]
        L.ref_weno_filter_flux.restype = ctypes.c_double
        L.ref_eigensystem_k.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, ctypes.c_double, dp,
                                        dp, dp, dp]
        L.ref_set_grid_coordinates.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        L.ref_get_metric_field.argtypes = [ctypes.c_void_p, ctypes.c_int, dp]
        _libs[variant] = L
    return _libs[variant]


def _ptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(p):
    """Build a RefConfig from a problem dict (the same dict the HIP path takes)."""
    c = RefConfig()
    c.ndim = p["ndim"]
    for d in range(3):
        c.n[d] = p["n"][d] if d < p["ndim"] else 1
        c.delta[d] = p["delta"][d] if d < p["ndim"] else 1.0
    c.conv = CONV[p["conv"]]
    c.averaging = AVG[p.get("averaging", "roe")]
    c.viscous = VISC[p.get("viscous", "none")]
    c.rk = RK[p.get("rk", "rk3")]
    for d in range(3):
        for s in range(2):
            c.bc[d][s] = BC[p.get("bc", [["periodic"] * 2] * 3)[d][s]] if d < p["ndim"] else 0
    c.gama = p.get("gama", 1.4)
    c.Minf = p.get("Minf", 1.0)
    c.Re = p.get("Re", 1.0)
    c.Pr = p.get("Pr", 1.0)
    c.dt = p["dt"]
    c.eps = p.get("eps", 1e-16)
    c.teno_ct = p.get("teno_ct", 1e-6)
    bs = p.get("bscheme", [["carpenter"] * 2] * 3)
    for d in range(3):
        for s in range(2):
            c.bscheme[d][s] = BSCHEME[bs[d][s]] if d < p["ndim"] else 0
        c.stretch[d] = int(p.get("stretch", [0, 0, 0])[d]) if d < p["ndim"] else 0
    c.suthT = p.get("SuthT", 110.4)
    c.refT = p.get("RefT", 288.0)
    c.twall = p.get("Twall", 1.0)
    c.sensor = 1 if p.get("sensor") else 0
    c.curvilinear = 1 if p.get("curvilinear") else 0
    c.teno_a1 = p.get("teno_a1", 10.5)
    c.teno_a2 = p.get("teno_a2", 4.5)
    c.sensor_eps = p.get("sensor_eps", 1.0e-30)
    (c.filters, c.binomial_order, c.wf_sensor, c.wf_mach, c.sigma_filt, c.chi_filt, c.omega_filt,
     c.shock_filter_control) = filter_fields(p)
    for d in range(3):
        c.body_force[d] = float(p.get("body_force", [0.0, 0.0, 0.0])[d]) if d < p["ndim"] else 0.0
    return c


def weno_filter_flux(fp, fm, cf):
    """(filter flux, rj) of one characteristic field from f+, f-, CF on points -2..3."""
    a, b, c = (np.ascontiguousarray(v, dtype=np.float64) for v in (fp, fm, cf))
    rj = np.zeros(1)
    v = lib().ref_weno_filter_flux(_ptr(a), _ptr(b), _ptr(c), _ptr(rj))
    return v, rj[0]


def eigensystem_k(gama, rho, u, a, k):
    """(L, R, lambda) of the Euler eigensystem about the metric direction k."""
    nd = len(u)
    nv = nd + 2
    uu, kk = (np.ascontiguousarray(v, dtype=np.float64) for v in (u, k))
    L, R, lam = np.zeros(nv * nv), np.zeros(nv * nv), np.zeros(nv)
    lib().ref_eigensystem_k(nd, gama, rho, _ptr(uu), a, _ptr(kk), _ptr(L), _ptr(R), _ptr(lam))
    return L.reshape(nv, nv), R.reshape(nv, nv), lam


def closure_weights(scheme, order):
    """Side-0 closure rows (rows x 6) of the first (order 1) or second (order 2) derivative."""
    w = np.zeros(24)
    rows = lib().ref_closure_weights(BSCHEME[scheme], order, _ptr(w))
    return w[: 6 * rows].reshape(rows, 6)


class RefSolver:
    """One block solved by the oracle; fields are interior arrays in (z, y, x) order."""

    def __init__(self, problem, variant="plain"):
        self.p = problem
        self.L = lib(variant)
        self.nd = problem["ndim"]
        self.nvar = self.nd + 2
        self.shape = tuple(reversed(problem["n"][: self.nd]))
        self.cfg = make_config(problem)
        h = ctypes.c_void_p()
        if self.L.ref_create(ctypes.byref(self.cfg), ctypes.byref(h)) != 0:
            raise RuntimeError("ref_create failed")
        self.h = h
        for d, x in problem.get("coords", {}).items():
            a = np.ascontiguousarray(x, dtype=np.float64)
            if self.L.ref_set_coordinates(self.h, int(d), _ptr(a)) != 0:
                raise RuntimeError(f"ref_set_coordinates({d}) failed")
        if problem.get("grid_coords") is not None:
            a = np.ascontiguousarray(problem["grid_coords"], dtype=np.float64)
            if self.L.ref_set_grid_coordinates(self.h, int(problem.get("grid_halo", 2)), _ptr(a)) != 0:
                raise RuntimeError("ref_set_grid_coordinates failed")
        for (d, side), pl in problem.get("planes", {}).items():
            a = np.ascontiguousarray(pl, dtype=np.float64)
            if self.L.ref_set_bc_plane(self.h, int(d), int(side), _ptr(a)) != 0:
                raise RuntimeError(f"ref_set_bc_plane({d}, {side}) failed")
        mask = ((problem.get("filters") or {}).get("binomial") or {}).get("mask")
        if mask is not None:
            self._mask = np.ascontiguousarray(mask, dtype=np.uint8)
            assert self._mask.shape == self.shape
            self.L.ref_set_filter_mask(self.h, self._mask.ctypes.data)

    def set_filter_mask(self, mask):
        """Grid condition of the binomial filter (one byte per interior cell);
        None applies the filter everywhere."""
        if mask is None:
            self._mask = None
            ok = self.L.ref_set_filter_mask(self.h, None)
        else:
            self._mask = np.ascontiguousarray(mask, dtype=np.uint8)
            assert self._mask.shape == self.shape
            ok = self.L.ref_set_filter_mask(self.h, self._mask.ctypes.data)
        if ok != 0:
            raise RuntimeError("no binomial filter on this block")

    def sfd_field(self):
        out = np.empty((self.nvar,) + self.shape)
        for m in range(self.nvar):
            if self.L.ref_get_sfd_field(self.h, m, _ptr(out[m])) != 0:
                raise RuntimeError("no SFD on this block")
        return out

    def set_sfd_field(self, qf):
        for m in range(self.nvar):
            a = np.ascontiguousarray(qf[m], dtype=np.float64)
            if self.L.ref_set_sfd_field(self.h, m, _ptr(a)) != 0:
                raise RuntimeError("no SFD on this block")

    def sensor(self):
        out = np.empty(self.shape)
        if self.L.ref_get_sensor(self.h, _ptr(out)) != 0:
            raise RuntimeError("no sensor on this block")
        return out

    def metric(self, d):
        n = self.p["n"][d]
        dm, sd = np.empty(n), np.empty(n)
        if self.L.ref_get_metric(self.h, d, _ptr(dm), _ptr(sd)) != 0:
            raise RuntimeError("no metric for direction %d" % d)
        return dm, sd

    def metric_field(self, idx):
        """Interior values of D_ab (idx = a*ndim + b) or detJ (idx = ndim*ndim)."""
        out = np.empty(self.shape)
        if self.L.ref_get_metric_field(self.h, int(idx), _ptr(out)) != 0:
            raise RuntimeError("no curvilinear metric array %d" % idx)
        return out

    def __del__(self):
        if getattr(self, "h", None):
            self.L.ref_destroy(self.h)
            self.h = None

    def set_state(self, q):
        for m in range(self.nvar):
            a = np.ascontiguousarray(q[m], dtype=np.float64)
            assert a.shape == self.shape
            self.L.ref_set_field(self.h, m, _ptr(a))

    def get_state(self):
        out = np.empty((self.nvar,) + self.shape)
        for m in range(self.nvar):
            a = np.empty(self.shape)
            self.L.ref_get_field(self.h, m, _ptr(a))
            out[m] = a
        return out

    def rhs(self):
        out = np.empty((self.nvar,) + self.shape)
        self.L.ref_rhs(self.h, _ptr(out))
        return out

    def run(self, niter):
        self.L.ref_run(self.h, int(niter))
