"""ctypes binding of libosbl_hip.so (the C-ABI declared in include/opensbli_amd.h).

The HIP library is the only compute path: if it is missing or cannot be
loaded, every entry point raises.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libosbl_hip.so")

CONV = {"central_skew": 0, "central_cons": 1, "weno_js": 2, "weno_z": 3, "teno5": 4, "teno6": 5}
AVG = {"simple": 0, "roe": 1}
VISC = {"none": 0, "const": 1, "sutherland": 2}
RK = {"rk3": 0, "ls3": 1, "ls3_ssp": 2, "ls4": 3}
BC = {"periodic": 0, "extrap0": 1, "extrap1": 2, "zero_gradient": 3, "adiabatic_wall": 4,
      "isothermal_wall": 5, "dirichlet": 6, "inlet_pextrap": 7, "symmetry": 8}
BSCHEME = {"carpenter": 0, "reduced_access": 1}
FILTERS = {"weno": 1, "sfd": 2, "binomial": 4}
WF_SENSOR = {"ducros": 0, "constant": 1}

# symbols the header declares (checked by the ABI test)
EXPORTS = [
    "osbl_abi_version", "osbl_last_error", "osbl_create", "osbl_destroy", "osbl_set_field", "osbl_get_field",
    "osbl_set_field_dev", "osbl_get_field_dev", "osbl_run", "osbl_rhs", "osbl_sync", "osbl_stream",
    "osbl_profile", "osbl_kernel_stats", "osbl_work_per_stage", "osbl_kernel_work", "osbl_num_stages",
    "osbl_batch_create", "osbl_batch_destroy", "osbl_batch_set_state", "osbl_batch_get_state", "osbl_batch_run",
    "osbl_batch_set_threads", "osbl_batch_set_graph_steps", "osbl_batch_save", "osbl_batch_restore",
    "osbl_rhs_vjp_dev", "osbl_step_stages_dev", "osbl_adjoint_step_dev", "osbl_adjoint_step_dev_noalloc",
    "osbl_use_graphs",
    "osbl_set_coordinates", "osbl_set_bc_plane", "osbl_set_filter_mask", "osbl_get_sfd_field",
    "osbl_set_sfd_field", "osbl_get_field_halo", "osbl_set_grid_coordinates", "osbl_get_metric_field",
    "osbl_get_sensor", "osbl_nan_count", "osbl_get_point",
]


class OsblConfig(ctypes.Structure):
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
        ("reserved3", ctypes.c_int32),
        ("sigma_filt", ctypes.c_double),
        ("chi_filt", ctypes.c_double),
        ("omega_filt", ctypes.c_double),
        ("shock_filter_control", ctypes.c_double),
        ("body_force", ctypes.c_double * 3),
    ]


class OsblError(RuntimeError):
    pass


_lib = None


def load():
    """Load libosbl_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OsblError(f"{LIB_PATH} is missing: build it with `python -m opensbli_amd.build` "
                        "or __graft_entry__.build(); there is no CPU fallback")
    # torch's wheel bundles its own libamdhip64.so.7 under the same soname as
    # the one this library's RUNPATH names; whichever is loaded first serves
    # the whole process.  Loading torch first keeps one runtime, the one torch
    # was built against, so torch tensors and this library share devices and
    # streams however the caller orders its imports.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, i64 = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int64
    L.osbl_abi_version.restype = ctypes.c_int
    L.osbl_last_error.restype = ctypes.c_char_p
    L.osbl_create.argtypes = [ctypes.POINTER(OsblConfig), ctypes.POINTER(vp)]
    L.osbl_destroy.argtypes = [vp]
    L.osbl_destroy.restype = None
    L.osbl_set_field.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_get_field.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_set_field_dev.argtypes = [vp, ctypes.c_int, vp]
    L.osbl_get_field_dev.argtypes = [vp, ctypes.c_int, vp]
    L.osbl_get_field_halo.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int32), dp]
    L.osbl_run.argtypes = [vp, i64]
    L.osbl_rhs.argtypes = [vp, dp]
    L.osbl_sync.argtypes = [vp]
    L.osbl_stream.argtypes = [vp]
    L.osbl_stream.restype = vp
    L.osbl_profile.argtypes = [vp, ctypes.c_int]
    L.osbl_kernel_stats.argtypes = [vp, ctypes.c_int, ctypes.c_char_p, ctypes.c_int, dp,
                                    ctypes.POINTER(ctypes.c_int64)]
    L.osbl_work_per_stage.argtypes = [vp, dp, dp]
    L.osbl_kernel_work.argtypes = [vp, ctypes.c_int, dp, dp]
    L.osbl_num_stages.argtypes = [vp]
    L.osbl_batch_create.argtypes = [ctypes.POINTER(OsblConfig), ctypes.c_int, vp, vp, vp, vp, ctypes.POINTER(vp)]
    L.osbl_batch_destroy.argtypes = [vp]
    L.osbl_batch_destroy.restype = None
    L.osbl_batch_set_state.argtypes = [vp, ctypes.c_int, vp]
    L.osbl_batch_get_state.argtypes = [vp, ctypes.c_int, vp]
    L.osbl_batch_run.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_batch_set_threads.argtypes = [vp, ctypes.c_int, vp, vp]
    L.osbl_batch_set_graph_steps.argtypes = [vp, ctypes.c_int]
    L.osbl_batch_save.argtypes = [vp]
    L.osbl_batch_restore.argtypes = [vp]
    L.osbl_rhs_vjp_dev.argtypes = [vp, vp, vp, vp]
    L.osbl_step_stages_dev.argtypes = [vp, vp]
    L.osbl_adjoint_step_dev.argtypes = [vp, vp, vp]
    L.osbl_adjoint_step_dev_noalloc.argtypes = [vp, vp, vp]
    L.osbl_use_graphs.argtypes = [vp, ctypes.c_int]
    L.osbl_set_coordinates.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_set_bc_plane.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
    L.osbl_set_filter_mask.argtypes = [vp, vp]
    L.osbl_get_sfd_field.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_set_sfd_field.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_set_grid_coordinates.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_get_metric_field.argtypes = [vp, ctypes.c_int, dp]
    L.osbl_get_sensor.argtypes = [vp, dp]
    L.osbl_nan_count.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64)]
    L.osbl_get_point.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int32), dp]
    _lib = L
    return L


def check(rc):
    if rc < 0:
        raise OsblError(load().osbl_last_error().decode() or f"osbl error {rc}")
    return rc


def make_config(p):
    """osbl_config from a problem description dict (see opensbli_amd.problems)."""
    c = OsblConfig()
    nd = p["ndim"]
    c.ndim = nd
    for d in range(3):
        c.n[d] = p["n"][d] if d < nd else 1
        c.delta[d] = p["delta"][d] if d < nd else 1.0
    c.conv = CONV[p["conv"]]
    c.averaging = AVG[p.get("averaging", "roe")]
    c.viscous = VISC[p.get("viscous", "none")]
    c.rk = RK[p.get("rk", "rk3")]
    bcs = p.get("bc", [["periodic", "periodic"]] * 3)
    for d in range(3):
        for s in range(2):
            c.bc[d][s] = BC[bcs[d][s]] if d < nd else 0
    c.gama = p.get("gama", 1.4)
    c.Minf = p.get("Minf", 1.0)
    c.Re = p.get("Re", 1.0)
    c.Pr = p.get("Pr", 1.0)
    c.dt = p["dt"]
    c.eps = p.get("eps", 1e-16)
    c.teno_ct = p.get("teno_ct", 1e-6)
    bs = p.get("bscheme", [["carpenter", "carpenter"]] * 3)
    for d in range(3):
        for s in range(2):
            c.bscheme[d][s] = BSCHEME[bs[d][s]] if d < nd else 0
        c.stretch[d] = int(p.get("stretch", [0, 0, 0])[d]) if d < nd else 0
    c.suthT = p.get("SuthT", 110.4)
    c.refT = p.get("RefT", 288.0)
    c.twall = p.get("Twall", 1.0)
    c.sensor = 1 if p.get("sensor") else 0
    c.curvilinear = 1 if p.get("curvilinear") else 0
    c.teno_a1 = p.get("teno_a1", 10.5)
    c.teno_a2 = p.get("teno_a2", 4.5)
    c.sensor_eps = p.get("sensor_eps", 1.0e-30)
    # filters: {"weno": {"sensor", "mach_correction", "control"}, "sfd": {"chi", "omega"},
    #           "binomial": {"order", "sigma", "mask"}}, defaults as the reference's constructors
    f = p.get("filters") or {}
    bits = 0
    for k in f:
        if k not in FILTERS:
            raise ValueError(f"unknown filter {k!r}")
        bits |= FILTERS[k]
    c.filters = bits
    b, sf, w = f.get("binomial") or {}, f.get("sfd") or {}, f.get("weno") or {}
    c.binomial_order = int(b.get("order", 2))
    c.sigma_filt = float(b.get("sigma", 0.05))
    c.chi_filt = float(sf.get("chi", 0.1))
    c.omega_filt = float(sf.get("omega", 1.0 / 0.75))
    c.wf_sensor = WF_SENSOR[w.get("sensor", "ducros")]
    c.wf_mach = 1 if w.get("mach_correction") else 0
    c.shock_filter_control = float(w.get("control", 1.0))
    for d in range(3):
        c.body_force[d] = float(p.get("body_force", [0.0, 0.0, 0.0])[d]) if d < nd else 0.0
    return c


def _dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Block:
    """One block on the GPU: the state lives in HBM between calls.

    Fields are numpy arrays of interior points in (z, y, x) order (x fastest),
    variable order rho, rhou0.., rhoE.
    """

    def __init__(self, problem):
        L = load()
        self.p = dict(problem)
        self.nd = problem["ndim"]
        self.nvar = self.nd + 2
        self.shape = tuple(reversed(list(problem["n"][: self.nd])))
        self.cfg = make_config(problem)
        h = ctypes.c_void_p()
        check(L.osbl_create(ctypes.byref(self.cfg), ctypes.byref(h)))
        self.h = h
        for d, x in problem.get("coords", {}).items():
            a = np.ascontiguousarray(x, dtype=np.float64)
            check(L.osbl_set_coordinates(self.h, int(d), _dptr(a)))
        if problem.get("grid_coords") is not None:
            self.set_grid_coordinates(problem["grid_coords"], problem.get("grid_halo", 2))
        for (d, side), pl in problem.get("planes", {}).items():
            a = np.ascontiguousarray(pl, dtype=np.float64)
            check(L.osbl_set_bc_plane(self.h, int(d), int(side), _dptr(a)))
        mask = ((problem.get("filters") or {}).get("binomial") or {}).get("mask")
        if mask is not None:
            self.set_filter_mask(mask)

    def set_grid_coordinates(self, coords, halo=2):
        """Coordinates of a curvilinear block: (ndim,) + the interior extended by
        `halo` points on every side, axes in (z, y, x) order."""
        a = np.ascontiguousarray(coords, dtype=np.float64)
        want = (self.nd,) + tuple(n + 2 * halo for n in self.shape)
        if a.shape != want:
            raise ValueError(f"coordinates have shape {a.shape}, expected {want}")
        check(load().osbl_set_grid_coordinates(self.h, int(halo), _dptr(a)))

    def sensor(self):
        """The Ducros sensor theta of the last stage evaluated (interior)."""
        out = np.empty(self.shape)
        check(load().osbl_get_sensor(self.h, _dptr(out)))
        return out

    def nan_count(self, var=-1):
        """NaN values among the interior values of conservative variable `var`
        of the current state (all variables when var < 0)."""
        n = ctypes.c_int64(0)
        check(load().osbl_nan_count(self.h, int(var), ctypes.byref(n)))
        return n.value

    def point(self, var, index):
        """Value of conservative variable `var` at interior index (i, j, k)
        (ndim entries, x first)."""
        if len(index) != self.nd:
            raise ValueError(f"index {tuple(index)} has {len(index)} entries, the block {self.nd} dimensions")
        idx = (ctypes.c_int32 * self.nd)(*[int(i) for i in index])
        v = ctypes.c_double(0.0)
        check(load().osbl_get_point(self.h, int(var), idx, ctypes.byref(v)))
        return v.value

    def metric_field(self, idx):
        """Interior values of D_ab (idx = a*ndim + b) or detJ (idx = ndim*ndim) of
        a curvilinear block; D_dd (idx = d*ndim + d) of a stretched direction."""
        out = np.empty(self.shape)
        check(load().osbl_get_metric_field(self.h, int(idx), _dptr(out)))
        return out

    def set_filter_mask(self, mask):
        """Grid condition of the binomial filter (interior cells, nonzero where it applies); None clears it."""
        if mask is None:
            check(load().osbl_set_filter_mask(self.h, None))
            return
        m = np.ascontiguousarray(mask, dtype=np.uint8)
        if m.shape != self.shape:
            raise ValueError(f"mask has shape {m.shape}, expected {self.shape}")
        check(load().osbl_set_filter_mask(self.h, m.ctypes.data))

    def sfd_field(self):
        """The SFD filtered solution vfilt, (nvar,) + shape."""
        out = np.empty((self.nvar,) + self.shape)
        for m in range(self.nvar):
            a = np.empty(self.shape)
            check(load().osbl_get_sfd_field(self.h, m, _dptr(a)))
            out[m] = a
        return out

    def set_sfd_field(self, qf):
        for m in range(self.nvar):
            a = np.ascontiguousarray(qf[m], dtype=np.float64)
            if a.shape != self.shape:
                raise ValueError(f"field {m} has shape {a.shape}, expected {self.shape}")
            check(load().osbl_set_sfd_field(self.h, m, _dptr(a)))

    def close(self):
        twin = getattr(self, "_adjoint_twin", None)  # the adjoint's recomputation context
        if twin is not None:
            twin.close()
            self._adjoint_twin = None
        if getattr(self, "h", None):
            load().osbl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, q):
        for m in range(self.nvar):
            a = np.ascontiguousarray(q[m], dtype=np.float64)
            if a.shape != self.shape:
                raise ValueError(f"field {m} has shape {a.shape}, expected {self.shape}")
            check(load().osbl_set_field(self.h, m, _dptr(a)))

    def get_state(self):
        out = np.empty((self.nvar,) + self.shape)
        for m in range(self.nvar):
            a = np.empty(self.shape)
            check(load().osbl_get_field(self.h, m, _dptr(a)))
            out[m] = a
        return out

    def halo_depths(self):
        """(d_m, d_p) per direction of get_field_halo: d_m <= 0 low side, d_p >= 0 high side."""
        dm, dp_ = (ctypes.c_int32 * 3)(), (ctypes.c_int32 * 3)()
        check(load().osbl_get_field_halo(self.h, 0, dm, dp_, None))
        nd = len(self.shape)
        return list(dm)[:nd], list(dp_)[:nd]

    def get_field_halo(self, m):
        """Variable m with the halo layers the boundary conditions fill (applied
        to the current state first): (array, d_m, d_p), the array's axes in
        reverse direction order (x last)."""
        dm, dp_ = self.halo_depths()
        nd = len(self.shape)
        shape = tuple(self.shape[k] - dm[nd - 1 - k] + dp_[nd - 1 - k] for k in range(nd))
        a = np.empty(shape)
        m_arr, p_arr = (ctypes.c_int32 * 3)(), (ctypes.c_int32 * 3)()
        check(load().osbl_get_field_halo(self.h, int(m), m_arr, p_arr, _dptr(a)))
        return a, dm, dp_

    def set_state_dev(self, ptrs):
        for m, ptr in enumerate(ptrs):
            check(load().osbl_set_field_dev(self.h, m, ctypes.c_void_p(ptr)))

    def get_state_dev(self, ptrs):
        for m, ptr in enumerate(ptrs):
            check(load().osbl_get_field_dev(self.h, m, ctypes.c_void_p(ptr)))

    def rhs(self):
        out = np.empty((self.nvar,) + self.shape)
        check(load().osbl_rhs(self.h, _dptr(out)))
        return out

    def run(self, niter):
        check(load().osbl_run(self.h, int(niter)))

    def use_graphs(self, enable):
        check(load().osbl_use_graphs(self.h, 1 if enable else 0))

    def sync(self):
        check(load().osbl_sync(self.h))

    def stream(self):
        return load().osbl_stream(self.h)

    def profile(self, enable):
        check(load().osbl_profile(self.h, 1 if enable else 0))

    def kernel_stats(self):
        L = load()
        out = {}
        n = check(L.osbl_kernel_stats(self.h, -1, None, 0, None, None))
        for i in range(n):
            name = ctypes.create_string_buffer(64)
            ms = ctypes.c_double()
            cnt = ctypes.c_int64()
            check(L.osbl_kernel_stats(self.h, i, name, 64, ctypes.byref(ms), ctypes.byref(cnt)))
            out[name.value.decode()] = (ms.value, cnt.value)
        return out

    def work_per_stage(self):
        b, f = ctypes.c_double(), ctypes.c_double()
        check(load().osbl_work_per_stage(self.h, ctypes.byref(b), ctypes.byref(f)))
        return b.value, f.value

    def kernel_work(self):
        """Per-launch algorithmic (bytes, flops) of every kernel family, by name."""
        L = load()
        out = {}
        n = check(L.osbl_kernel_stats(self.h, -1, None, 0, None, None))
        for i in range(n):
            name = ctypes.create_string_buffer(64)
            check(L.osbl_kernel_stats(self.h, i, name, 64, None, None))
            b, f = ctypes.c_double(), ctypes.c_double()
            check(L.osbl_kernel_work(self.h, i, ctypes.byref(b), ctypes.byref(f)))
            out[name.value.decode()] = (b.value, f.value)
        return out

    # --- adjoint (device pointers to packed interior arrays) ---
    def rhs_vjp_dev(self, q_ptr, mu_ptr, out_ptr):
        check(load().osbl_rhs_vjp_dev(self.h, ctypes.c_void_p(q_ptr), ctypes.c_void_p(mu_ptr), ctypes.c_void_p(out_ptr)))

    def step_stages_dev(self, stages_ptr):
        check(load().osbl_step_stages_dev(self.h, ctypes.c_void_p(stages_ptr)))

    def adjoint_step_dev(self, stages_ptr, lam_ptr):
        check(load().osbl_adjoint_step_dev(self.h, ctypes.c_void_p(stages_ptr), ctypes.c_void_p(lam_ptr)))

    def adjoint_step_dev_noalloc(self, stages_ptr, lam_ptr):
        """adjoint_step_dev with its work buffers in the library's fixed device scratch (no allocation)."""
        check(load().osbl_adjoint_step_dev_noalloc(self.h, ctypes.c_void_p(stages_ptr), ctypes.c_void_p(lam_ptr)))

    def num_stages(self):
        return check(load().osbl_num_stages(self.h))
