"""Parity of the HIP filters with the oracle through the C-ABI: the WENO
non-linear filter (WENO_filter.py:15-339), selective frequency damping
(SFD.py:27-86) and the binomial filter (binomial_filter.py:15-123), each run
after every full time step.

Bar as in test_gpu_parity.py: relative L2 <= 1e-10 per conservative variable
(conftest.parity_errors), after 100 steps.
"""
import math
import sys

import numpy as np
import pytest

from conftest import check_parity

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(n, conv, filters, bc=None, viscous="none", dt=1e-3, **kw):
    nd = len(n)
    p = dict(ndim=nd, n=list(n), delta=[2 * math.pi / m for m in n], conv=conv, averaging="roe", viscous=viscous,
             rk="rk3", bc=bc or [["periodic"] * 2] * nd, gama=1.4, Minf=0.5, Re=200.0, Pr=0.72, dt=dt, eps=1e-16,
             sensor_eps=1e-12, filters=filters)
    p.update(kw)
    return p


def _state(n, seed=0):
    """A compressible state with a density/pressure/velocity jump along x and
    smooth variation elsewhere (interior arrays, x fastest)."""
    rs = np.random.RandomState(seed)
    nd = len(n)
    grids = np.meshgrid(*[np.arange(m) / m for m in reversed(n)], indexing="ij")
    x = grids[-1]
    y = grids[-2]
    hi = (x > 0.3) & (x < 0.6)
    rho = np.where(hi, 1.5, 1.0) + 0.05 * np.sin(2 * np.pi * y)
    vel = [np.where(hi, 0.0, 0.4) + 0.05 * np.cos(2 * np.pi * y), 0.1 * np.sin(2 * np.pi * x)]
    if nd == 3:
        z = grids[0]
        vel.append(0.05 * np.sin(2 * np.pi * (x + z)))
    p = np.where(hi, 3.0, 2.4) + 0.02 * rs.uniform(-1, 1, rho.shape)
    ke = sum(v * v for v in vel)
    return np.stack([rho] + [rho * v for v in vel] + [p / 0.4 + 0.5 * rho * ke])


def _compare(problem, q0, oracle, steps=100, sfd0=None):
    from opensbli_amd.native import Block
    g = Block(problem)
    g.set_state(q0)
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    if sfd0 is not None:
        for s in (g, r):
            s.set_sfd_field(sfd0)
    g.run(steps)
    r.run(steps)
    a, b = g.get_state(), r.get_state()
    assert np.isfinite(a).all()
    check_parity(a, b, "state")
    if "sfd" in problem["filters"]:
        check_parity(g.sfd_field(), r.sfd_field(), "vfilt")
    return a, b


@pytest.mark.parametrize("sensor,mach", [("ducros", False), ("ducros", True), ("constant", False),
                                         ("constant", True)])
def test_weno_filter_central_periodic_2d(sensor, mach, oracle):
    n = (64, 48)
    _say(f"case weno filter central skew 2-D sensor={sensor} mach={mach}")
    p = _problem(n, "central_skew", {"weno": {"sensor": sensor, "mach_correction": mach, "control": 1.0}})
    a, b = _compare(p, _state(n), oracle)
    # the filter acted: the run differs from one without it
    q = oracle.RefSolver(_problem(n, "central_skew", {}))
    q.set_state(_state(n))
    q.run(100)
    assert np.abs(q.get_state() - b).max() > 1e-6


def test_weno_filter_3d_with_viscous_terms(oracle):
    n = (32, 24, 16)
    _say("case weno filter central skew 3-D viscous ducros+mach")
    p = _problem(n, "central_skew", {"weno": {"sensor": "ducros", "mach_correction": True}}, viscous="const")
    _compare(p, _state(n, seed=1), oracle)


def test_weno_filter_walls_closures_and_metric(oracle):
    """The flat plate block (stretched y, Carpenter/ReducedAccess closures,
    walls) with the filter: wall zones, metric factor and closure rows of the
    sensor."""
    from opensbli_amd.problems import flat_plate
    _say("case weno filter on the flat plate 64x48")
    p, q0 = flat_plate(n=(64, 48), ndim=2)
    p["filters"] = {"weno": {"sensor": "ducros"}}
    p["sensor_eps"] = 1e-12
    _compare(p, q0, oracle)


def test_weno_filter_with_characteristic_scheme_and_walls(oracle):
    n = (48, 40)
    _say("case weno filter + weno_js, adiabatic walls in y")
    q0 = _state(n, seed=2)
    q0[2] = 0.0
    p = _problem(n, "weno_js", {"weno": {"sensor": "constant"}},
                 bc=[["periodic"] * 2, ["adiabatic_wall"] * 2], bscheme=[["carpenter"] * 2] * 3)
    _compare(p, q0, oracle)


@pytest.mark.parametrize("restart", [False, True])
def test_sfd(restart, oracle):
    n = (48, 32)
    _say(f"case sfd restart={restart}")
    q0 = _state(n, seed=3)
    p = _problem(n, "weno_z", {"sfd": {"chi": 0.5, "omega": 2.0}}, dt=2e-3)
    sfd0 = None
    if restart:
        rs = np.random.RandomState(7)
        sfd0 = q0 * (1.0 + 0.02 * rs.uniform(-1, 1, q0.shape))
    _compare(p, q0, oracle, sfd0=sfd0)


@pytest.mark.parametrize("order,conv,bc", [
    (2, "central_skew", "periodic"), (4, "central_skew", "periodic"), (6, "central_skew", "periodic"),
    (8, "weno_z", "periodic"), (6, "weno_js", "extrap0"), (4, "central_cons", "extrap1"),
])
def test_binomial(order, conv, bc, oracle):
    n = (40, 32)
    _say(f"case binomial order={order} conv={conv} bc={bc}")
    q0 = _state(n, seed=4)
    mask = np.ones((n[1], n[0]), dtype=np.uint8)
    mask[: n[1] // 3, :] = 0
    bcs = [[bc] * 2, ["periodic"] * 2]
    p = _problem(n, conv, {"binomial": {"order": order, "sigma": 0.1, "mask": mask if order == 4 else None}},
                 bc=bcs)
    _compare(p, q0, oracle)


def test_binomial_mask_cleared_between_runs(oracle):
    """Clearing the grid condition between two runs (set_filter_mask(None))
    takes effect in the next run, which replays the time loop from captured
    graphs: the state follows the oracle with the mask cleared at the same
    iteration, and differs from a run that keeps the mask."""
    from opensbli_amd.native import Block
    n = (40, 32)
    _say("case binomial order 4 with a grid condition, cleared after 4 of 8 iterations")
    q0 = _state(n, seed=6)
    mask = np.ones((n[1], n[0]), dtype=np.uint8)
    mask[: n[1] // 2, :] = 0
    p = _problem(n, "central_skew", {"binomial": {"order": 4, "sigma": 0.1, "mask": mask}})
    g = Block(p)
    g.set_state(q0)
    r = oracle.RefSolver(p)
    r.set_state(q0)
    kept = oracle.RefSolver(p)
    kept.set_state(q0)
    for s in (g, r, kept):
        s.run(4)
    g.set_filter_mask(None)
    r.set_filter_mask(None)
    for s in (g, r, kept):
        s.run(4)
    a, b = g.get_state(), r.get_state()
    check_parity(a, b, "mask cleared after 4 iterations")
    assert np.abs(kept.get_state() - b).max() > 1e-8


def test_all_filters_together_3d(oracle):
    n = (24, 20, 16)
    _say("case weno filter + sfd + binomial 3-D")
    p = _problem(n, "central_skew", {"weno": {"sensor": "ducros"}, "sfd": {}, "binomial": {"order": 4, "sigma": 0.05}},
                 viscous="const")
    _compare(p, _state(n, seed=5), oracle)


def test_filter_configuration_errors():
    from opensbli_amd.native import Block, OsblError
    n = (32, 16)
    with pytest.raises(OsblError, match="even"):
        Block(_problem(n, "central_skew", {"binomial": {"order": 3}}))
    with pytest.raises(OsblError, match="halo"):
        Block(_problem(n, "central_skew", {"binomial": {"order": 10}}))
    with pytest.raises(OsblError, match="halo"):
        Block(_problem(n, "central_skew", {"binomial": {"order": 6}}, bc=[["extrap0"] * 2, ["periodic"] * 2]))
    with pytest.raises(OsblError, match="2-D and 3-D"):
        Block(dict(_problem((32,), "weno_z", {"sfd": {}})))
    g = Block(_problem(n, "central_skew", {}))
    with pytest.raises(OsblError, match="no SFD"):
        g.sfd_field()
    with pytest.raises(OsblError, match="no binomial"):
        g.set_filter_mask(np.ones((16, 32), dtype=np.uint8))


def test_filters_script_on_gpu(oracle):
    """The TGV script with WENOFilter, SFD and a BinomialFilter with a grid
    condition, through HIPC.run, against the oracle on the same problem."""
    from opensbli_amd.apps.taylor_green import build
    _say("case TGV script 16^3 with the three filters, 20 steps")
    prog = build(16, "central", niter=20, filters=("weno", "sfd", "binomial"))
    res = prog.run()
    p = prog.problem()
    q0 = prog.initial_state(p)
    p = prog._attach_boundary_data(p, q0)
    r = oracle.RefSolver(p)
    r.set_state(q0)
    r.run(20)
    q = np.stack([res[k] for k in ("rho", "rhou0", "rhou1", "rhou2", "rhoE")])
    check_parity(q, r.get_state(), "TGV script with filters")
