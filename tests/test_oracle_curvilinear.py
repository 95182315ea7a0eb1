"""Pins of the oracle's curvilinear path: the general metric from grid
coordinates (metric.py:72-147), the Euler eigensystem about a metric
direction (euler_eigensystem.py:18-40,88-135), the fluxes detJ F_j of the
strong conservation form and the residual divided by detJ
(apps/euler_wave_curvilinear/euler_wave.py:20-25).

What fixes them, independently of the oracle's own formulas:
* L R = I and R diag(lambda) L = d(k.F)/dQ, the flux Jacobian along the
  (non-unit) direction k, by complex-step differentiation; k = e_d gives the
  Cartesian eigensystem bitwise;
* an affine mesh x = A xi has D = A^-1 and detJ = det A (the central
  derivative is exact on linear functions); the wavy mesh's metric converges
  to the analytic one at 4th order;
* on the identity mesh the curvilinear right-hand side is the Cartesian one;
  on a rotated uniform mesh (2-D) it is the Cartesian right-hand side of the
  state seen in the grid frame, rotated back (the 2-D characteristic fields
  are frame-invariant), which fails if D and its transpose are swapped
  anywhere;
* on the wavy mesh the density wave converges to the exact translation.
"""
import math

import numpy as np
import pytest

from conftest import rel_l2
from opensbli_amd.problems import euler_wave_curvilinear


def flux_k(q, k, nd, g):
    rho = q[0]
    u = [q[1 + m] / rho for m in range(nd)]
    ke = sum(ui * ui for ui in u)
    p = (g - 1.0) * (q[nd + 1] - 0.5 * rho * ke)
    U = sum(k[m] * u[m] for m in range(nd))
    return np.array([rho * U] + [q[1 + m] * U + p * k[m] for m in range(nd)] + [(q[nd + 1] + p) * U])


def jacobian_k(q, k, nd, g):
    h = 1e-30
    A = np.zeros((nd + 2, nd + 2))
    for c in range(nd + 2):
        qc = q.astype(complex)
        qc[c] += 1j * h
        A[:, c] = np.imag(flux_k(qc, k, nd, g)) / h
    return A


@pytest.mark.parametrize("nd", [2, 3])
def test_eigensystem_about_a_metric_direction(oracle, nd):
    g = 1.4
    rng = np.random.RandomState(70 + nd)
    for _ in range(8):
        rho = rng.uniform(0.5, 2.0)
        u = rng.uniform(-1.0, 1.0, nd)
        p = rng.uniform(0.5, 2.0)
        a = math.sqrt(g * p / rho)
        k = rng.uniform(-2.0, 2.0, nd)
        L, R, lam = oracle.eigensystem_k(g, rho, u, a, k)
        assert np.allclose(L @ R, np.eye(nd + 2), atol=1e-12)
        q = np.array([rho] + list(rho * u) + [p / (g - 1) + 0.5 * rho * np.dot(u, u)])
        A = jacobian_k(q, k, nd, g)
        assert np.allclose(R @ np.diag(lam) @ L, A, atol=1e-11 * (1 + np.abs(A).max()))
        # L and R depend on the unit normal only; the eigenvalues scale with |k|
        L2, R2, lam2 = oracle.eigensystem_k(g, rho, u, a, 3.0 * k)
        assert np.allclose(L2, L, rtol=1e-13, atol=1e-13) and np.allclose(R2, R, rtol=1e-13, atol=1e-13)
        assert np.allclose(lam2, 3.0 * lam, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("nd", [2, 3])
def test_unit_direction_is_the_cartesian_eigensystem(oracle, nd):
    from test_oracle_pins import eigensystem
    g = 1.4
    rho, u, a = 1.2, np.array([0.3, -0.4, 0.2][:nd]), 1.1
    for d in range(nd):
        k = np.zeros(nd)
        k[d] = 1.0
        L, R, lam = oracle.eigensystem_k(g, rho, u, a, k)
        Lc, Rc, lamc = eigensystem(oracle, nd, d, g, rho, u, a)
        assert np.array_equal(L, Lc) and np.array_equal(R, Rc) and np.array_equal(lam, lamc)


def _solver(oracle, problem, q=None):
    s = oracle.RefSolver(problem)
    if q is not None:
        s.set_state(q)
    return s


def test_affine_mesh_metric_is_exact(oracle):
    n, h = 12, 2
    d = 2.0 / n
    A = np.array([[1.3, 0.4], [-0.25, 0.9]])
    i = np.arange(-h, n + h) * d
    jj, ii = np.meshgrid(i, i, indexing="ij")
    coords = np.stack([A[0, 0] * ii + A[0, 1] * jj + 0.7, A[1, 0] * ii + A[1, 1] * jj - 0.3])
    problem, q = euler_wave_curvilinear(n, coords=coords)
    s = _solver(oracle, problem, q)
    Dinv = np.linalg.inv(A)
    for a in range(2):
        for b in range(2):
            assert np.allclose(s.metric_field(2 * a + b), Dinv[a, b], rtol=0, atol=1e-13)
    assert np.allclose(s.metric_field(4), np.linalg.det(A), rtol=0, atol=1e-13)


def test_affine_mesh_metric_3d(oracle):
    n, h = 8, 2
    d = 2.0 / n
    A = np.array([[1.1, 0.2, -0.1], [0.3, 0.8, 0.25], [-0.2, 0.15, 1.2]])
    i = np.arange(-h, n + h) * d
    kk, jj, ii = np.meshgrid(i, i, i, indexing="ij")
    xi = [ii, jj, kk]
    coords = np.stack([sum(A[a, b] * xi[b] for b in range(3)) for a in range(3)])
    problem = dict(ndim=3, n=[n] * 3, delta=[d] * 3, conv="weno_z", averaging="roe", viscous="none", rk="ls3",
                   bc=[["periodic", "periodic"]] * 3, gama=1.4, dt=1e-3, curvilinear=1, grid_coords=coords,
                   grid_halo=h)
    s = _solver(oracle, problem)
    Dinv = np.linalg.inv(A)
    for a in range(3):
        for b in range(3):
            assert np.allclose(s.metric_field(3 * a + b), Dinv[a, b], rtol=0, atol=1e-13)
    assert np.allclose(s.metric_field(9), np.linalg.det(A), rtol=0, atol=1e-13)


def test_wavy_mesh_metric_converges_at_fourth_order(oracle):
    errs = []
    for n in (32, 64):
        d = 2.0 / n
        problem, q = euler_wave_curvilinear(n)
        s = _solver(oracle, problem, q)
        sa = 0.04 * math.sin(2.0 * math.pi * 0.25)
        i = np.arange(n) * d
        y, x = np.meshgrid(i, i, indexing="ij")
        # J = d(x0, x1)/d(xi0, xi1) of the analytic map, then D = J^-1
        J00 = np.ones_like(x)
        J01 = sa * 3.0 * math.pi * np.cos(3.0 * math.pi * y)
        J10 = sa * 3.0 * math.pi * np.cos(3.0 * math.pi * x)
        J11 = np.ones_like(x)
        det = J00 * J11 - J01 * J10
        exact = [J11 / det, -J01 / det, -J10 / det, J00 / det, det]
        errs.append(max(np.abs(s.metric_field(m) - exact[m]).max() for m in range(5)))
    assert errs[1] < errs[0] / 12.0, errs


@pytest.mark.parametrize("conv", ["weno_z", "weno_js", "teno6"])
def test_identity_mesh_is_the_cartesian_rhs(oracle, conv):
    n, h = 16, 2
    d = 2.0 / n
    i = np.arange(-h, n + h) * d
    jj, ii = np.meshgrid(i, i, indexing="ij")
    problem, q = euler_wave_curvilinear(n, conv=conv, coords=np.stack([ii, jj]))
    rng = np.random.RandomState(5)
    q = q * (1.0 + 0.05 * rng.standard_normal(q.shape))
    cart = {k: v for k, v in problem.items() if k not in ("curvilinear", "grid_coords", "grid_halo")}
    Rc = _solver(oracle, cart, q).rhs()
    Rv = _solver(oracle, problem, q).rhs()
    assert np.all(rel_l2(Rv, Rc) < 1e-13), rel_l2(Rv, Rc)


def test_identity_mesh_3d_is_the_cartesian_rhs(oracle):
    from opensbli_amd.problems import taylor_green
    n, h = 12, 2
    problem, q = taylor_green(n, conv="weno_z", viscous="none")
    d = problem["delta"][0]
    i = np.arange(-h, n + h) * d
    kk, jj, ii = np.meshgrid(i, i, i, indexing="ij")
    curv = dict(problem, curvilinear=1, grid_coords=np.stack([ii, jj, kk]), grid_halo=h)
    # perturbed: on the exact TGV state (u2 = 0) WENO-Z's eps = 1e-16 turns
    # rounding-level smoothness indicators into O(1e-5) differences
    rng = np.random.RandomState(1)
    q = q * (1.0 + 0.05 * rng.standard_normal(q.shape))
    Rc = _solver(oracle, problem, q).rhs()
    Rv = _solver(oracle, curv, q).rhs()
    assert np.all(rel_l2(Rv, Rc) < 1e-13), rel_l2(Rv, Rc)


@pytest.mark.parametrize("theta", [0.5235987755982988, 1.1, -2.3])
def test_rotated_mesh_is_the_rotated_cartesian_rhs(oracle, theta):
    n, h = 16, 2
    d = 2.0 / n
    c, s_ = math.cos(theta), math.sin(theta)
    Rot = np.array([[c, -s_], [s_, c]])
    i = np.arange(-h, n + h) * d
    jj, ii = np.meshgrid(i, i, indexing="ij")
    coords = np.stack([Rot[0, 0] * ii + Rot[0, 1] * jj, Rot[1, 0] * ii + Rot[1, 1] * jj])
    problem, _ = euler_wave_curvilinear(n, coords=coords)
    cart = {k: v for k, v in problem.items() if k not in ("curvilinear", "grid_coords", "grid_halo")}
    # a non-trivial periodic state in the grid frame (velocities along the grid lines)
    g = 1.4
    x = np.arange(n) * d
    y, xx = np.meshgrid(x, x, indexing="ij")
    rho = 1.0 + 0.2 * np.sin(np.pi * xx) * np.cos(np.pi * y)
    ug = np.stack([0.4 + 0.3 * np.cos(np.pi * y), -0.2 + 0.25 * np.sin(np.pi * (xx + y))])
    p = 1.0 + 0.1 * np.cos(np.pi * (xx - y))
    E = p / (g - 1) + 0.5 * rho * (ug ** 2).sum(0)
    qg = np.stack([rho, rho * ug[0], rho * ug[1], E])
    up = np.einsum("ab,bij->aij", Rot, ug)  # physical-frame velocity
    qp = np.stack([rho, rho * up[0], rho * up[1], E])
    Rg = _solver(oracle, cart, qg).rhs()
    Rp = _solver(oracle, problem, qp).rhs()
    mom = np.einsum("ab,bij->aij", Rot, Rg[1:3])
    want = np.stack([Rg[0], mom[0], mom[1], Rg[3]])
    assert np.all(rel_l2(Rp, want) < 1e-12), rel_l2(Rp, want)
    # the transposed metric gives a different answer (the test can fail)
    coords_t = np.stack([Rot[0, 0] * ii + Rot[1, 0] * jj, Rot[0, 1] * ii + Rot[1, 1] * jj])
    problem_t, _ = euler_wave_curvilinear(n, coords=coords_t)
    Rt = _solver(oracle, problem_t, qp).rhs()
    assert np.max(rel_l2(Rt, want)) > 1e-3


def test_wavy_mesh_density_wave_converges(oracle):
    """Exact solution rho = 1 + 0.2 sin(pi (x0 + x1 - t (u0 + u1))) at the mesh points."""
    errs = []
    T = 0.2
    for n in (32, 64):
        d = 2.0 / n
        dt = 0.1 * d
        steps = int(round(T / dt))
        problem, q = euler_wave_curvilinear(n, dt=dt)
        s = _solver(oracle, problem, q)
        s.run(steps)
        rho = s.get_state()[0]
        c = problem["grid_coords"][:, 2:2 + n, 2:2 + n]
        exact = 1.0 + 0.2 * np.sin(np.pi * (c[0] + c[1] - steps * dt * 0.5))
        errs.append(np.abs(rho - exact).max())
    order = math.log2(errs[0] / errs[1])
    assert order > 3.5, (errs, order)
