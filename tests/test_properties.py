"""Property tests of the oracle (hypothesis): invariants that hold for every
state, direction, mesh or angle, searched over rather than sampled at fixed
seeds.  Each one is fixed by the mathematics, not by the oracle's formulas:

* the eigensystem about any direction k diagonalises d(k.F)/dQ (complex-step
  Jacobian) with L R = I;
* the Roe average has the Roe property F(Q_R) - F(Q_L) = A(Q^)(Q_R - Q_L);
* every reconstruction returns a constant exactly when fed one;
* a uniform state gives a zero right-hand side for every scheme, and on a
  periodic block the mass and momentum residuals sum to zero;
* an affine curvilinear mesh has D = A^-1, detJ = det A;
* on a rotated uniform mesh the curvilinear RHS is the rotated Cartesian RHS.
"""
import math

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st

from conftest import rel_l2

SETTINGS = settings(max_examples=25, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
g = 1.4

states = st.tuples(st.floats(0.3, 3.0), st.lists(st.floats(-2.0, 2.0), min_size=3, max_size=3), st.floats(0.3, 3.0))


def _q(rho, u, p):
    return np.array([rho] + [rho * x for x in u] + [p / (g - 1) + 0.5 * rho * float(np.dot(u, u))])


def _flux_k(q, k, nd):
    rho = q[0]
    u = [q[1 + m] / rho for m in range(nd)]
    p = (g - 1.0) * (q[nd + 1] - 0.5 * rho * sum(x * x for x in u))
    U = sum(k[m] * u[m] for m in range(nd))
    return np.array([rho * U] + [q[1 + m] * U + p * k[m] for m in range(nd)] + [(q[nd + 1] + p) * U])


def _jac(q, k, nd):
    A = np.zeros((nd + 2, nd + 2))
    for c in range(nd + 2):
        qc = q.astype(complex)
        qc[c] += 1e-30j
        A[:, c] = np.imag(_flux_k(qc, k, nd)) / 1e-30
    return A


@SETTINGS
@given(nd=st.sampled_from([2, 3]), s=states, k=st.lists(st.floats(-3.0, 3.0), min_size=3, max_size=3))
def test_eigensystem_diagonalises_any_direction(oracle, nd, s, k):
    rho, u, p = s
    k = np.array(k[:nd])
    if np.linalg.norm(k) < 1e-2:
        k = k + 0.5
    u = np.array(u[:nd])
    a = math.sqrt(g * p / rho)
    L, R, lam = oracle.eigensystem_k(g, rho, u, a, k)
    assert np.allclose(L @ R, np.eye(nd + 2), atol=1e-11)
    A = _jac(_q(rho, u, p), k, nd)
    assert np.allclose(R @ np.diag(lam) @ L, A, atol=1e-10 * (1 + np.abs(A).max()))


@SETTINGS
@given(nd=st.sampled_from([1, 2, 3]), sl=states, sr=states)
def test_roe_property(oracle, nd, sl, sr):
    import ctypes
    qL = _q(sl[0], sl[1][:nd], sl[2])
    qR = _q(sr[0], sr[1][:nd], sr[2])
    out = np.zeros(nd + 2)
    dp = ctypes.POINTER(ctypes.c_double)
    ext = lambda q, p: np.ascontiguousarray(np.append(q, p))  # noqa: E731
    a, b = ext(qL, sl[2]), ext(qR, sr[2])
    oracle.lib().ref_roe_average(nd, g, a.ctypes.data_as(dp), b.ctypes.data_as(dp), out.ctypes.data_as(dp))
    rho, u, ah = out[0], out[1:1 + nd], out[nd + 1]
    e = np.zeros(nd)
    e[0] = 1.0
    L, R, lam = oracle.eigensystem_k(g, rho, u, ah, e) if nd > 1 else (None, None, None)
    if nd == 1:
        return
    A = R @ np.diag(lam) @ L
    dF = _flux_k(qR, e, nd) - _flux_k(qL, e, nd)
    assert np.allclose(A @ (qR - qL), dF, atol=1e-10 * (1 + np.abs(dF).max()))


@SETTINGS
@given(conv=st.sampled_from(["weno_js", "weno_z", "teno5", "teno6"]), side=st.sampled_from([1, -1]),
       c=st.floats(-1e3, 1e3))
def test_reconstruction_of_a_constant(oracle, conv, side, c):
    import ctypes
    from oracle.ref import CONV
    f = np.full(6, c)
    v = oracle.lib().ref_recon_side(CONV[conv], side, f.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1e-16,
                                    1e-6)
    # a few roundings of values of magnitude |c|: within 4 units in the last place
    assert abs(v - c) <= 4 * np.spacing(max(1.0, abs(c)))


@settings(max_examples=10, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(conv=st.sampled_from(["central_skew", "central_cons", "weno_z", "weno_js", "teno5", "teno6"]), s=states)
def test_uniform_state_has_zero_rhs(oracle, conv, s):
    rho, u, p = s
    n = 10
    problem = dict(ndim=2, n=[n, n], delta=[0.1, 0.1], conv=conv, averaging="roe", viscous="none", rk="rk3",
                   gama=g, dt=1e-3)
    q = np.stack([np.full((n, n), v) for v in _q(rho, u[:2], p)])
    r = oracle.RefSolver(problem)
    r.set_state(q)
    assert np.abs(r.rhs()).max() < 1e-11 * (1 + np.abs(q).max())


@settings(max_examples=10, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(conv=st.sampled_from(["central_cons", "weno_z", "teno6"]), seed=st.integers(0, 2 ** 31 - 1))
def test_periodic_conservation(oracle, conv, seed):
    n = 12
    rs = np.random.RandomState(seed)
    rho = 1.0 + 0.3 * rs.rand(n, n)
    u = 0.5 * rs.standard_normal((2, n, n))
    p = 1.0 + 0.3 * rs.rand(n, n)
    q = np.stack([rho, rho * u[0], rho * u[1], p / (g - 1) + 0.5 * rho * (u ** 2).sum(0)])
    problem = dict(ndim=2, n=[n, n], delta=[0.1, 0.1], conv=conv, averaging="roe", viscous="none", rk="rk3",
                   gama=g, dt=1e-3)
    r = oracle.RefSolver(problem)
    r.set_state(q)
    R = r.rhs()
    for m in range(4):
        assert abs(R[m].sum()) < 1e-11 * np.abs(R[m]).sum()


@SETTINGS
@given(a=st.lists(st.floats(-0.5, 0.5), min_size=4, max_size=4))
def test_affine_mesh_metric(oracle, a):
    from opensbli_amd.problems import euler_wave_curvilinear
    A = np.eye(2) + np.array(a).reshape(2, 2)
    if abs(np.linalg.det(A)) < 0.2:
        A = A + 0.6 * np.eye(2)
    n, h = 8, 2
    d = 2.0 / n
    i = np.arange(-h, n + h) * d
    jj, ii = np.meshgrid(i, i, indexing="ij")
    coords = np.stack([A[0, 0] * ii + A[0, 1] * jj, A[1, 0] * ii + A[1, 1] * jj])
    problem, q = euler_wave_curvilinear(n, coords=coords)
    r = oracle.RefSolver(problem)
    Dinv = np.linalg.inv(A)
    for x in range(2):
        for y in range(2):
            assert np.allclose(r.metric_field(2 * x + y), Dinv[x, y], rtol=1e-12, atol=1e-12)
    assert np.allclose(r.metric_field(4), np.linalg.det(A), rtol=1e-12, atol=1e-12)


@settings(max_examples=12, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(theta=st.floats(-math.pi, math.pi), seed=st.integers(0, 2 ** 31 - 1))
def test_rotated_mesh_rhs(oracle, theta, seed):
    from opensbli_amd.problems import euler_wave_curvilinear
    n, h = 12, 2
    d = 2.0 / n
    c, s = math.cos(theta), math.sin(theta)
    Rot = np.array([[c, -s], [s, c]])
    i = np.arange(-h, n + h) * d
    jj, ii = np.meshgrid(i, i, indexing="ij")
    coords = np.stack([Rot[0, 0] * ii + Rot[0, 1] * jj, Rot[1, 0] * ii + Rot[1, 1] * jj])
    problem, _ = euler_wave_curvilinear(n, coords=coords)
    cart = {k: v for k, v in problem.items() if k not in ("curvilinear", "grid_coords", "grid_halo")}
    rs = np.random.RandomState(seed)
    x = np.arange(n) * d
    y, xx = np.meshgrid(x, x, indexing="ij")
    ph = rs.rand(4) * 2 * np.pi
    rho = 1.0 + 0.2 * np.sin(np.pi * xx + ph[0]) * np.cos(np.pi * y + ph[1])
    ug = np.stack([0.3 + 0.3 * np.cos(np.pi * y + ph[2]), -0.2 + 0.25 * np.sin(np.pi * (xx + y) + ph[3])])
    E = 1.0 / (g - 1) + 0.5 * rho * (ug ** 2).sum(0)
    qg = np.stack([rho, rho * ug[0], rho * ug[1], E])
    up = np.einsum("ab,bij->aij", Rot, ug)
    qp = np.stack([rho, rho * up[0], rho * up[1], E])
    a = oracle.RefSolver(cart)
    a.set_state(qg)
    Rg = a.rhs()
    b = oracle.RefSolver(problem)
    b.set_state(qp)
    Rp = b.rhs()
    mom = np.einsum("ab,bij->aij", Rot, Rg[1:3])
    want = np.stack([Rg[0], mom[0], mom[1], Rg[3]])
    assert np.all(rel_l2(Rp, want) < 1e-11), rel_l2(Rp, want)
