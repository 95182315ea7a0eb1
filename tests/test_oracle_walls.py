"""Pins of the oracle's non-periodic machinery: stencil closures, metric of a
stretched direction, Sutherland viscous terms, the boundary conditions and
the boundary-layer inflow profile.

Checks are against values the reference itself holds (the one-sided
coefficients printed in katzer_init.py:172, the wall temperatures of the two
flat-plate apps), against polynomial exactness, and against a manufactured
solution whose exact Navier-Stokes right-hand side is derived with SymPy, so
that a dropped term, a wrong sign or a wrong closure row changes the
convergence rate.
"""
import numpy as np
import pytest

from conftest import rel_l2
from oracle import ref
from oracle.ref import RefSolver
from opensbli_amd.boundary_layer import SimilarityProfile
from opensbli_amd.problems import flat_plate, stretched_coordinate


# ---------------------------------------------------------------------------
# closures
# ---------------------------------------------------------------------------
def test_carpenter_first_row_matches_reference_constants():
    # katzer_init.py:172 prints row 0 of the Carpenter closure (used for du/dy at the wall)
    printed = [-1.83333333333334, 3.00000000000002, -1.50000000000003, 0.333333333333356,
               -8.34657956545823e-15, 1.06910315192207e-15]
    w = ref.closure_weights("carpenter", 1)
    assert np.allclose(w[0], printed, rtol=0, atol=1e-13)


@pytest.mark.parametrize("scheme,order,exact_deg", [("carpenter", 1, 3), ("reduced_access", 1, 4),
                                                    ("carpenter", 2, 4), ("reduced_access", 2, 4)])
def test_closure_polynomial_exactness(scheme, order, exact_deg):
    w = ref.closure_weights(scheme, order)
    x = np.arange(6.0)
    for deg in range(exact_deg + 2):
        f = x ** deg
        if order == 1:
            ex = deg * x ** (deg - 1) if deg > 0 else np.zeros(6)
        else:
            ex = deg * (deg - 1) * x ** (deg - 2) if deg > 1 else np.zeros(6)
        err = max(abs(w[r] @ f - ex[r]) for r in range(w.shape[0]))
        if deg <= exact_deg:
            assert err < 1e-12 * max(1.0, 5.0 ** deg)
        else:
            assert err > 1e-3  # the stated order is the actual one


# ---------------------------------------------------------------------------
# metric of a stretched direction
# ---------------------------------------------------------------------------
def _block_2d(n, bc, coords=None, bscheme="carpenter", viscous="none", conv="weno_z", planes=None, **kw):
    d = [2.0 / (n[0] - 1), 2.0 / (n[1] - 1)]
    p = dict(ndim=2, n=list(n), delta=d, conv=conv, averaging="roe", viscous=viscous, rk="rk3", bc=bc,
             bscheme=[[bscheme] * 2] * 2 + [["carpenter"] * 2], stretch=[0, 1 if coords is not None else 0, 0],
             gama=1.4, Minf=0.5, Re=50.0, Pr=0.72, dt=1e-3, SuthT=110.4, RefT=288.0, Twall=1.2)
    if coords is not None:
        p["coords"] = {1: coords}
    if planes:
        p["planes"] = planes
    p.update(kw)
    return p


@pytest.mark.parametrize("bscheme", ["carpenter", "reduced_access"])
def test_metric_linear_map_is_exact(bscheme):
    n = (16, 24)
    a = 3.7
    xi = np.arange(n[1]) * 2.0 / (n[1] - 1)
    p = _block_2d(n, [["extrap0"] * 2, ["extrap0"] * 2], coords=a * xi + 0.3, bscheme=bscheme)
    dm, sd = RefSolver(p).metric(1)
    assert np.abs(dm - 1.0 / a).max() < 1e-14
    assert np.abs(sd).max() < 1e-12


@pytest.mark.parametrize("bscheme,rate", [("carpenter", 2.8), ("reduced_access", 3.6)])
def test_metric_sinh_map_converges(bscheme, rate):
    errs = []
    for n1 in (33, 65):
        L, by = 2.0, 2.0
        xi = np.arange(n1) * L / (n1 - 1)
        p = _block_2d((16, n1), [["extrap0"] * 2, ["extrap0"] * 2],
                      coords=stretched_coordinate(n1, L, by, "sinh"), bscheme=bscheme)
        dm, sd = RefSolver(p).metric(1)
        # x = L sinh(by xi/L)/sinh(by):  x' = by cosh(by xi/L)/sinh(by)
        xp = by * np.cosh(by * xi / L) / np.sinh(by)
        xpp = by * by / L * np.sinh(by * xi / L) / np.sinh(by)
        errs.append((np.abs(dm - 1.0 / xp).max(), np.abs(sd + xpp / xp ** 2).max()))
    r_dm = np.log2(errs[0][0] / errs[1][0])
    r_sd = np.log2(errs[0][1] / errs[1][1])
    assert r_dm > rate and r_sd > rate - 1.0


# ---------------------------------------------------------------------------
# manufactured solution: exact NS right-hand side with Sutherland viscosity
# ---------------------------------------------------------------------------
def _exact_rhs_functions(gama, Minf, Re, Pr, S):
    sp = pytest.importorskip("sympy")
    x, y = sp.symbols("x y")
    rho = 1 + sp.Rational(1, 10) * sp.sin(1.3 * x + 0.4) * sp.cos(0.7 * y + 0.2)
    u = sp.Rational(3, 10) + sp.Rational(1, 5) * sp.cos(0.9 * x) * sp.sin(1.1 * y + 0.3)
    v = sp.Rational(1, 10) * sp.sin(0.8 * x + 0.5) * sp.cos(1.2 * y)
    p = (1 / (gama * Minf ** 2)) * (1 + sp.Rational(1, 10) * sp.cos(x - 0.3) * sp.sin(0.6 * y + 0.1))
    E = p / (gama - 1) + rho * (u * u + v * v) / 2
    T = gama * Minf ** 2 * p / rho
    mu = T ** sp.Rational(3, 2) * (1 + S) / (T + S)
    X = [x, y]
    U = [u, v]
    div = sp.diff(u, x) + sp.diff(v, y)
    tau = [[mu / Re * (sp.diff(U[i], X[j]) + sp.diff(U[j], X[i]) - (sp.Rational(2, 3) * div if i == j else 0))
            for j in range(2)] for i in range(2)]
    q = [-mu / ((gama - 1) * Minf ** 2 * Pr * Re) * sp.diff(T, X[j]) for j in range(2)]
    R = [-(sp.diff(rho * u, x) + sp.diff(rho * v, y))]
    for i in range(2):
        R.append(-sum(sp.diff(rho * U[i] * U[j] + (p if i == j else 0), X[j]) for j in range(2))
                 + sum(sp.diff(tau[i][j], X[j]) for j in range(2)))
    R.append(-sum(sp.diff((E + p) * U[j], X[j]) for j in range(2)) - sum(sp.diff(q[j], X[j]) for j in range(2))
             + sum(sp.diff(U[i] * tau[i][j], X[j]) for i in range(2) for j in range(2)))
    state = [rho, rho * u, rho * v, E]
    f_state = sp.lambdify((x, y), state, "numpy")
    f_rhs = sp.lambdify((x, y), R, "numpy")
    return f_state, f_rhs


@pytest.mark.parametrize("bscheme,min_rate", [("carpenter", 2.5), ("reduced_access", 2.5)])
def test_manufactured_rhs_converges(bscheme, min_rate):
    """Closures are 3rd order at the faces (max-norm rate ~3), the interior
    stencils 4th order (rate ~4 away from the faces)."""
    gama, Minf, Re, Pr, S = 1.4, 0.5, 50.0, 0.72, 110.4 / 288.0
    f_state, f_rhs = _exact_rhs_functions(gama, Minf, Re, Pr, S)
    errs, errs_in = [], []
    for n in (32, 64):
        L = 2.0
        xs = np.arange(n) * L / (n - 1)
        ys = stretched_coordinate(n, L, 1.5, "sinh")
        Y, X = np.meshgrid(ys, xs, indexing="ij")
        q = np.stack([np.broadcast_to(np.asarray(a, dtype=float), X.shape) for a in f_state(X, Y)])
        planes = {(0, 0): q[:, :, 0][:, None, :], (0, 1): q[:, :, -1][:, None, :],
                  (1, 0): q[:, 0, :][:, None, :], (1, 1): q[:, -1, :][:, None, :]}
        p = _block_2d((n, n), [["dirichlet"] * 2, ["dirichlet"] * 2], coords=ys, bscheme=bscheme,
                      viscous="sutherland", conv="central_cons", planes=planes, gama=gama, Minf=Minf, Re=Re, Pr=Pr,
                      RefT=288.0, SuthT=110.4)
        p["delta"] = [L / (n - 1), L / (n - 1)]
        s = RefSolver(p)
        s.set_state(q)
        r = s.rhs()
        ex = np.stack([np.broadcast_to(np.asarray(a, dtype=float), X.shape) for a in f_rhs(X, Y)])
        scale = np.abs(ex).max(axis=(1, 2))
        errs.append(np.abs(r - ex).max(axis=(1, 2)) / scale)
        errs_in.append(np.abs(r - ex)[:, 4:-4, 4:-4].max(axis=(1, 2)) / scale)
    rates = np.log2(errs[0] / errs[1])
    rates_in = np.log2(errs_in[0] / errs_in[1])
    assert np.all(errs[1] < 5e-3), errs
    assert np.all(rates > min_rate), rates
    assert np.all(rates_in > 3.2), rates_in


# ---------------------------------------------------------------------------
# boundary conditions (the boundary rows after the BC kernels ran)
# ---------------------------------------------------------------------------
def _wall_state(n, seed=1, gama=1.4, Minf=0.5):
    rs = np.random.RandomState(seed)
    shape = (n[1], n[0])
    rho = 1.0 + 0.1 * rs.rand(*shape)
    u = 0.3 + 0.1 * rs.rand(*shape)
    v = 0.05 * rs.rand(*shape)
    p = 1.0 / (gama * Minf ** 2) * (1.0 + 0.05 * rs.rand(*shape))
    E = p / (gama - 1) + 0.5 * rho * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, E])


def _temperature(q, gama=1.4, Minf=0.5):
    ke = 0.5 * (q[1] ** 2 + q[2] ** 2) / q[0]
    return (q[3] - ke) * gama * (gama - 1) * Minf ** 2 / q[0]


def test_adiabatic_wall_zero_normal_gradient():
    n = (12, 16)
    p = _block_2d(n, [["extrap0"] * 2, ["adiabatic_wall", "extrap0"]])
    s = RefSolver(p)
    s.set_state(_wall_state(n))
    s.rhs()
    q = s.get_state()
    assert np.all(q[1:3, 0, :] == 0.0)
    T = _temperature(q)
    # 3rd-order one-sided dT/dy at the wall vanishes
    g = -11.0 / 6.0 * T[0] + 3.0 * T[1] - 1.5 * T[2] + T[3] / 3.0
    assert np.abs(g).max() < 1e-13 * np.abs(T).max()


def test_isothermal_wall_temperature():
    n = (12, 16)
    p = _block_2d(n, [["extrap0"] * 2, ["extrap0", "isothermal_wall"]])
    s = RefSolver(p)
    s.set_state(_wall_state(n))
    s.rhs()
    q = s.get_state()
    assert np.all(q[1:3, -1, :] == 0.0)
    assert np.abs(_temperature(q)[-1] - p["Twall"]).max() < 1e-14


def test_zero_gradient_and_extrapolation_rows():
    n = (12, 16)
    p = _block_2d(n, [["zero_gradient", "extrap0"], ["extrap1", "extrap1"]])
    s = RefSolver(p)
    q0 = _wall_state(n)
    s.set_state(q0)
    s.rhs()
    q = s.get_state()
    assert np.array_equal(q[:, :, 0], q[:, :, 1])       # zero gradient: boundary point takes its neighbour
    assert np.array_equal(q[:, :, -1], q[:, :, -2])     # order-0 extrapolation
    assert np.array_equal(q[:, 1:-1, 1:-1], q0[:, 1:-1, 1:-1])


@pytest.mark.parametrize("supersonic", [True, False])
def test_inlet_pressure_extrapolation(supersonic):
    n = (12, 16)
    gama, Minf = 1.4, 0.5
    q0 = _wall_state(n)
    if supersonic:  # |u| above the sound speed everywhere: a ~ 2, u ~ 3
        q0[1] = q0[0] * 3.0
        q0[3] = 1.0 / (gama * Minf ** 2) / (gama - 1) + 0.5 * (q0[1] ** 2 + q0[2] ** 2) / q0[0]
    plane = (q0[:, :, 0] * 1.01)[:, None, :]
    p = _block_2d(n, [["inlet_pextrap", "extrap0"], ["extrap0"] * 2], planes={(0, 0): plane})
    s = RefSolver(p)
    s.set_state(q0)
    s.rhs()
    q = s.get_state()
    rows = slice(1, -1)  # the y-boundary rows are rewritten by the y faces
    if supersonic:
        assert np.array_equal(q[:, rows, 0], plane[:, 0, rows])
    else:
        assert np.array_equal(q[:, rows, 0], q0[:, rows, 0])


def test_dirichlet_rows():
    n = (12, 16)
    q0 = _wall_state(n)
    top = (q0[:, -1, :] * 0.9)[:, None, :]
    p = _block_2d(n, [["extrap0"] * 2, ["extrap0", "dirichlet"]], planes={(1, 1): top})
    s = RefSolver(p)
    s.set_state(q0)
    s.rhs()
    q = s.get_state()
    assert np.array_equal(q[:, -1, 1:-1], top[:, 0, 1:-1])


# ---------------------------------------------------------------------------
# boundary-layer inflow
# ---------------------------------------------------------------------------
def test_similarity_wall_temperatures_match_reference_apps():
    # katzer_SBLI.py:147-148 (M 2, Re 950, Tinf 288): Twall '1.67619431'
    assert abs(SimilarityProfile(2.0, 950.0, Tinf=288.0).Twall - 1.67619431) < 2e-8
    # transitional_SBLI.py:184-186 (M 1.5, Re 750, Tinf 202.17): Twall '1.3809973268575328'
    # (the reference integrates with odeint at rtol 1e-6)
    assert abs(SimilarityProfile(1.5, 750.0, Tinf=202.17).Twall - 1.3809973268575328) < 5e-7


def test_similarity_blasius_limit_and_displacement_thickness():
    b = SimilarityProfile(0.01, 950.0)
    assert abs(b.fpp0 - 0.469600) < 2e-6          # Blasius f''(0) for f''' + f f'' = 0
    assert abs(b.Twall - 1.0) < 1e-4
    b = SimilarityProfile(2.0, 950.0)
    y = np.linspace(0.0, 30.0, 30001)
    rho, u, rv, T = b.evaluate(y)
    assert abs(np.trapezoid(1.0 - rho * u, y) - 1.0) < 1e-4
    # free-stream normal mass flux: the shock-generator app's pre-shock rhou1 '0.00565001630205'
    assert abs(rv[-1] - 0.00565001630205) < 5e-6


def test_flat_plate_oracle_runs():
    p, q0 = flat_plate(n=(40, 32, 8))
    s = RefSolver(p)
    s.set_state(q0)
    s.run(5)
    q = s.get_state()
    assert np.isfinite(q).all()
    assert np.all(q[1:4, :, 0, :] == 0.0)  # no-slip wall row


def _symmetry_problem(ny, top, conv):
    return dict(ndim=2, n=[24, ny], delta=[0.04, 0.03], conv=conv, averaging="roe", viscous="const", rk="ls3",
                bc=[["adiabatic_wall", "adiabatic_wall"], ["adiabatic_wall", top]],
                bscheme=[["reduced_access"] * 2] * 2, gama=1.4, Minf=1.0, Re=200.0, Pr=0.73, dt=2e-4, eps=1e-15,
                teno_ct=1e-6)


def _mirrored_state(nx, nf, yc):
    x = np.arange(nx) * 0.04
    y = np.arange(nf) * 0.03
    Y, X = np.meshgrid(y, x, indexing="ij")
    rho = 1 + 0.3 * np.exp(-((X - 0.5) ** 2 + (Y - yc) ** 2) / 0.05) + 0.1 * np.cos(3 * (Y - yc))
    u = 0.2 * np.sin(2 * X) * np.cos(2 * (Y - yc))
    v = 0.3 * np.sin(3 * X) * np.sin(2 * (Y - yc))  # odd about the plane
    p = 1 / 1.4 + 0.1 * np.cos(2 * X) * np.cos(Y - yc)
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])


@pytest.mark.parametrize("conv", ["weno_z", "teno5"])
def test_symmetry_plane_is_the_mirrored_domain(oracle, conv):
    """SymmetryBC (symmetry.py:24-48) on the top face of an n-point block
    evolves exactly like the mirror-symmetric 2n-1-point block with walls at
    both ends: its halos are the mirror images with the normal momentum
    negated, the boundary point is left to the scheme, and the central
    stencils are not closed there.  A BC that kept the normal momentum, or
    closed the stencils, fails this."""
    ny = 16
    nf = 2 * ny - 1
    q = _mirrored_state(24, nf, (ny - 1) * 0.03)
    full = oracle.RefSolver(_symmetry_problem(nf, "adiabatic_wall", conv))
    full.set_state(q)
    half = oracle.RefSolver(_symmetry_problem(ny, "symmetry", conv))
    half.set_state(np.ascontiguousarray(q[:, :ny]))
    full.run(20)
    half.run(20)
    err = rel_l2(half.get_state(), full.get_state()[:, :ny])
    assert np.all(err < 1e-14), err
    # the planted mistake: a top face that closes the stencils (a wall) differs
    wall = oracle.RefSolver(_symmetry_problem(ny, "zero_gradient", conv))
    wall.set_state(np.ascontiguousarray(q[:, :ny]))
    wall.run(20)
    assert np.max(rel_l2(wall.get_state(), full.get_state()[:, :ny])) > 1e-6


def test_body_force_adds_its_terms_pointwise(oracle):
    """The body force -KD(_i,_j)*c_j, -Dot(c_j, u_j) (apps/channel_flow) changes
    the right-hand side at every interior point by exactly (0, -c, -c.u)."""
    from opensbli_amd.problems import channel
    for n in ((16, 17), (12, 17, 8)):
        problem, q = channel(n)
        nd = len(n)
        base = dict(problem, body_force=[0.0] * nd)
        a = oracle.RefSolver(problem)
        a.set_state(q)
        b = oracle.RefSolver(base)
        b.set_state(q)
        dR = a.rhs() - b.rhs()
        c = problem["body_force"]
        u = [q[1 + m] / q[0] for m in range(nd)]
        assert np.abs(dR[0]).max() == 0.0
        for m in range(nd):
            assert np.allclose(dR[1 + m], -c[m], rtol=0, atol=1e-12)
        assert np.allclose(dR[nd + 1], -sum(c[m] * u[m] for m in range(nd)), rtol=0, atol=1e-12)
