"""Pins of the oracle's post-step filters: the WENO non-linear filter
(WENO_filter.py:15-339, shock_capturing.py:82-96,401-416), selective frequency
damping (SFD.py:27-86) and the binomial filter (binomial_filter.py:15-123).

Each check is fixed by mathematics rather than by the oracle's own formula:
exact reconstruction of polynomials, the sensor's normalisation, the SFD map's
invariant and decay rate, the binomial filter's Fourier transfer function,
conservation of the flux form, and where the filter must leave the state
untouched.
"""
import math

import numpy as np
import pytest

from oracle import ref
from oracle.ref import RefSolver


def _cell_avg_power(p, j):
    """Average of x^p over the cell [j - 1/2, j + 1/2]."""
    return ((j + 0.5) ** (p + 1) - (j - 0.5) ** (p + 1)) / (p + 1)


# ---------------------------------------------------------------------------
# the filter flux of one characteristic field
# ---------------------------------------------------------------------------
def test_filter_flux_vanishes_on_quadratic_data(oracle):
    """On quadratic data the Jiang-Shu indicators are equal, the non-linear
    weights are the optimal ones (rj = 0) and both WENO5 and the sixth-order
    central flux reproduce the same half-node value: no filter flux."""
    rs = np.random.RandomState(1)
    for _ in range(20):
        a, b, c = rs.uniform(-1, 1, 3)
        lam = rs.uniform(0.1, 2.0)
        j = np.arange(-2, 4)
        cf = a + b * j + c * j * j
        cs = 0.3 * cf + rs.uniform(-1, 1) * j - 0.2
        cs = a * 0.5 + 0.1 * b * j + 0.7 * c * j * j
        fp = 0.5 * (cf + lam * cs)
        fm = 0.5 * (cf - lam * cs)
        v, rj = ref.weno_filter_flux(fp, fm, cf)
        assert rj < 1e-20
        assert abs(v) < 1e-14


def test_filter_sensor_is_one_on_a_step_seen_by_the_smallest_weight(oracle):
    """A step that leaves only the stencil of the smallest optimal weight
    (d = 1/10) smooth puts every weight there: sum |omega/d - 1|^4 = 9^4 + 2,
    the normalisation, so rj = 1 (for either split flux)."""
    step = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0])
    z = np.zeros(6)
    _, rj = ref.weno_filter_flux(step, z, step)
    assert abs(rj - 1.0) < 1e-9
    # the mirrored step seen by the left reconstruction of f-
    _, rj = ref.weno_filter_flux(z, 1.0 - step, 1.0 - step)
    assert abs(rj - 1.0) < 1e-9
    # partly one-sided data gives a sensor strictly between 0 and 1
    _, rj = ref.weno_filter_flux(np.array([0.0, 0.1, 0.3, 1.0, 1.2, 1.3]), z, step)
    assert 0.0 < rj < 1.0


@pytest.mark.parametrize("p", [0, 1, 2, 3, 4, 5])
def test_central_part_reconstructs_half_node_value_exactly(p, oracle):
    """With rj = 1 forced by the step in f+ (whose WENO value is 0) and
    f- = 0, the filter flux is minus the sixth-order central flux of cf.  That
    flux is the half-node value of h when cf are cell averages of h, exact for
    polynomials up to degree 5."""
    step = np.array([0.0, 0.0, 0.0, 1.0, 1.0, 1.0])
    cf = np.array([_cell_avg_power(p, j) for j in range(-2, 4)])
    v, rj = ref.weno_filter_flux(step, np.zeros(6), cf)
    # the WENO value of f+ is 0 up to the O(1e-12) weights the non-smooth stencils keep
    assert abs(rj - 1.0) < 1e-9
    np.testing.assert_allclose(v / rj, -0.5 ** p, rtol=0, atol=1e-10)


# ---------------------------------------------------------------------------
# end-to-end filters on a block
# ---------------------------------------------------------------------------
def _problem(n, conv="central_skew", bc=None, filters=None, dt=1e-3, **kw):
    nd = len(n)
    p = dict(ndim=nd, n=list(n), delta=[2 * math.pi / m for m in n], conv=conv, averaging="roe", viscous="none",
             rk="rk3", bc=bc or [["periodic"] * 2] * nd, gama=1.4, Minf=0.5, Re=100.0, Pr=0.72, dt=dt, eps=1e-16,
             sensor_eps=1e-12, filters=filters or {})
    p.update(kw)
    return p


def _uniform(n, rho=1.0, u=(0.0, 0.0), p=1.0 / (1.4 * 0.25)):
    shape = tuple(reversed(n))
    q = [np.full(shape, rho)]
    for m in range(len(n)):
        q.append(np.full(shape, rho * u[m]))
    q.append(np.full(shape, p / 0.4 + 0.5 * rho * sum(v * v for v in u[: len(n)])))
    return np.stack(q)


def test_sfd_invariant_and_decay(oracle):
    """On a state at rest the time step changes nothing, and the SFD map keeps
    v + (chi/omega) vfilt while v - vfilt decays by exp(-(chi + omega) dt) per
    step (SFD.py:76-80)."""
    n = (16, 12)
    chi, om, dt = 0.3, 1.7, 0.05
    p = _problem(n, filters={"sfd": {"chi": chi, "omega": om}}, dt=dt)
    q0 = _uniform(n)
    # a uniform vfilt different from the state keeps v uniform, so the RHS stays 0
    f0 = q0 * np.array([1.05, 1.0, 1.0, 0.97])[:, None, None] + np.array([0, 0.02, -0.01, 0])[:, None, None]
    s = RefSolver(p)
    s.set_state(q0)
    s.set_sfd_field(f0)
    steps = 7
    s.run(steps)
    v, f = s.get_state(), s.sfd_field()
    c, E = chi / om, math.exp(-(chi + om) * dt)
    np.testing.assert_allclose(v + c * f, q0 + c * f0, rtol=1e-13)
    np.testing.assert_allclose(v - f, (E ** steps) * (q0 - f0), rtol=1e-9, atol=1e-13)


def test_sfd_starts_from_the_state(oracle):
    """Without a restart field vfilt starts as the state: a steady state stays."""
    n = (12, 10)
    p = _problem(n, filters={"sfd": {}})
    q0 = _uniform(n)
    s = RefSolver(p)
    s.set_state(q0)
    s.run(3)
    np.testing.assert_allclose(s.get_state(), q0, rtol=1e-14, atol=1e-16)
    np.testing.assert_allclose(s.sfd_field(), q0, rtol=1e-14, atol=1e-16)


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10])
def test_binomial_transfer_function(order, oracle):
    """dt = 0 isolates the filter.  A density mode cos(k x) constant in y is
    multiplied per step by 1 - sigma + sigma (cos^N(k dx/2) + 1)/2: the x
    stencil's transfer function cos^N, the y stencil's 1, averaged."""
    n = (32, 8)
    sigma, k, amp = 0.2, 3, 0.05
    p = _problem(n, filters={"binomial": {"order": order, "sigma": sigma}}, dt=0.0)
    q0 = _uniform(n)
    x = 2 * math.pi * np.arange(n[0]) / n[0]
    q0[0] += amp * np.cos(k * x)[None, :]
    s = RefSolver(p)
    s.set_state(q0)
    steps = 3
    s.run(steps)
    dx = 2 * math.pi / n[0]
    g = 1.0 - sigma + sigma * (math.cos(k * dx / 2) ** order + 1.0) / 2.0
    q = s.get_state()
    np.testing.assert_allclose(q[0] - 1.0, amp * g ** steps * np.cos(k * x)[None, :] * np.ones((n[1], 1)),
                               atol=1e-14)
    np.testing.assert_allclose(q[1:], q0[1:], atol=1e-14)


def test_binomial_grid_condition(oracle):
    """Cells outside the grid condition keep their values."""
    n = (24, 16)
    mask = np.zeros((n[1], n[0]), dtype=np.uint8)
    mask[:, : n[0] // 2] = 1
    p = _problem(n, filters={"binomial": {"order": 4, "sigma": 0.5, "mask": mask}}, dt=0.0)
    rs = np.random.RandomState(5)
    q0 = _uniform(n) * (1.0 + 0.05 * rs.uniform(-1, 1, (4,) + (n[1], n[0])))
    s = RefSolver(p)
    s.set_state(q0)
    s.run(1)
    q = s.get_state()
    assert np.array_equal(q[:, mask == 0], q0[:, mask == 0])
    assert np.abs(q[:, mask == 1] - q0[:, mask == 1]).max() > 1e-3


def _shock_state(n, seed=0):
    """A periodic 2-D state with a density/pressure jump along x and smooth noise."""
    rs = np.random.RandomState(seed)
    y, x = np.meshgrid(np.arange(n[1]) / n[1], np.arange(n[0]) / n[0], indexing="ij")
    hi = (x > 0.3) & (x < 0.6)
    rho = np.where(hi, 1.5, 1.0) + 0.01 * np.sin(2 * np.pi * y)
    u = 0.2 + 0.05 * np.sin(2 * np.pi * y)
    v = 0.05 * np.cos(2 * np.pi * x)
    pr = np.where(hi, 2.5, 1.8) + 0.01 * rs.uniform(-1, 1, rho.shape)
    return np.stack([rho, rho * u, rho * v, pr / 0.4 + 0.5 * rho * (u * u + v * v)])


def test_weno_filter_conserves_on_periodic_block(oracle):
    """With the constant sensor the filter is a flux difference: on a periodic
    block the totals of the conservative variables are unchanged (to rounding),
    while the state differs from the unfiltered run."""
    n = (40, 16)
    q0 = _shock_state(n)
    base = _problem(n, conv="weno_js", dt=2e-3)
    filt = _problem(n, conv="weno_js", dt=2e-3, filters={"weno": {"sensor": "constant"}})
    a, b = RefSolver(base), RefSolver(filt)
    a.set_state(q0)
    b.set_state(q0)
    a.run(10)
    b.run(10)
    qa, qb = a.get_state(), b.get_state()
    tot0 = q0.sum(axis=(1, 2))
    np.testing.assert_allclose(qb.sum(axis=(1, 2)), tot0, rtol=1e-12, atol=1e-10)
    assert np.abs(qb - qa).max() > 1e-6


def test_ducros_sensor_scales_the_filter_per_point(oracle):
    """After one step the unfiltered, constant-sensor and Ducros-sensor runs
    share the post-step state, so the Ducros filter's change is kappa_max
    times the constant filter's (kappa_max in [0, 1], one value per point for
    every variable).  Away from compression kappa_max is ~0.  The Mach
    correction (a factor in [0, 1] on kappa) can only lower kappa_max."""
    n = (48, 16)
    q0 = _shock_state(n, seed=6)
    # a compressive velocity jump at the density jump gives the sensor something to see
    y, x = np.meshgrid(np.arange(n[1]) / n[1], np.arange(n[0]) / n[0], indexing="ij")
    u = np.where((x > 0.3) & (x < 0.6), 0.0, 0.4)
    q0[3] += 0.5 * q0[0] * (u * u - (q0[1] / q0[0]) ** 2)
    q0[1] = q0[0] * u
    runs = {}
    for name, filt in (("none", {}), ("const", {"weno": {"sensor": "constant"}}),
                       ("ducros", {"weno": {"sensor": "ducros"}}),
                       ("ducros_mach", {"weno": {"sensor": "ducros", "mach_correction": True}})):
        s = RefSolver(_problem(n, conv="weno_js", dt=1e-3, filters=filt))
        s.set_state(q0)
        s.run(1)
        runs[name] = s.get_state()
    dc = runs["const"] - runs["none"]
    big = np.all(np.abs(dc) > 1e-9, axis=0)
    assert big.sum() > 50
    ratios = {}
    for name in ("ducros", "ducros_mach"):
        ratio = (runs[name] - runs["none"])[:, big] / dc[:, big]
        assert ratio.min() > -1e-6 and ratio.max() < 1.0 + 1e-6
        np.testing.assert_allclose(ratio, np.broadcast_to(ratio[0], ratio.shape), atol=1e-6)
        ratios[name] = ratio[0]
    assert ratios["ducros"].max() > 0.1 and ratios["ducros"].min() < 1e-3
    assert np.all(ratios["ducros_mach"] <= ratios["ducros"] + 1e-6)
    assert ratios["ducros_mach"].max() < 0.9 * ratios["ducros"].max()


def test_weno_filter_off_near_walls_and_on_uniform_state(oracle):
    """One step with adiabatic walls at both y faces: rows within 6 points of a
    wall match the unfiltered step exactly; rows further in are filtered.  A
    uniform state is left untouched."""
    n = (32, 24)
    q0 = _shock_state(n, seed=2)
    q0[2] = 0.0
    bc = [["periodic"] * 2, ["adiabatic_wall"] * 2]
    kw = dict(conv="weno_js", bc=bc, dt=2e-3, bscheme=[["carpenter"] * 2] * 3)
    a = RefSolver(_problem(n, **kw))
    b = RefSolver(_problem(n, filters={"weno": {"sensor": "constant"}}, **kw))
    for s in (a, b):
        s.set_state(q0)
        s.run(1)
    qa, qb = a.get_state(), b.get_state()
    assert np.array_equal(qa[:, :6, :], qb[:, :6, :])
    assert np.array_equal(qa[:, -6:, :], qb[:, -6:, :])
    assert np.abs(qa[:, 6:-6, :] - qb[:, 6:-6, :]).max() > 1e-8

    u = RefSolver(_problem(n, conv="weno_js", filters={"weno": {"sensor": "constant"}}))
    qu = _uniform(n, u=(0.3, 0.1))
    u.set_state(qu)
    u.run(2)
    np.testing.assert_allclose(u.get_state(), qu, rtol=1e-13)


def test_weno_filter_dissipates_a_step(oracle):
    """An advected density step under the non-dissipative central scheme grows
    oscillations; the filter removes most of them: the total variation of
    density stays closer to the initial one."""
    n = (48, 8)
    y, x = np.meshgrid(np.arange(n[1]) / n[1], np.arange(n[0]) / n[0], indexing="ij")
    rho = np.where((x > 0.3) & (x < 0.6), 1.5, 1.0)
    q0 = np.stack([rho, 0.2 * rho, 0.0 * x, 1.8 / 0.4 + 0.5 * rho * 0.04])
    tv = lambda r: np.abs(np.diff(r, axis=1)).sum()
    out = {}
    for name, filt in (("central", {}), ("filtered", {"weno": {"sensor": "constant"}})):
        s = RefSolver(_problem(n, conv="central_skew", dt=5e-3, filters=filt))
        s.set_state(q0)
        s.run(10)
        out[name] = tv(s.get_state()[0]) - tv(q0[0])
    assert out["central"] > 0.5
    assert out["filtered"] < 0.5 * out["central"]
