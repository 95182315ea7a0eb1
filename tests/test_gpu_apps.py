"""The reference's Sod shock tube and Euler wave apps through the
problem-script facade on the GPU: against the oracle, and against their exact
solutions (the exact Sod density the reference ships; the translated wave,
apps/euler_wave/plot.py:93-103) at the reference's own sizes."""
import os
import sys

import numpy as np
import pytest

from conftest import check_parity

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _oracle_run(oracle, prog, steps):
    p = prog.problem()
    q0 = prog.initial_state(p)
    p = prog._attach_boundary_data(p, q0)
    r = oracle.RefSolver(p)
    r.set_state(q0)
    r.run(steps)
    return r.get_state()


def test_sod_app_against_oracle_and_exact(oracle):
    """apps/Sod_shock_tube as the reference runs it: 200 points, TENO5,
    Dirichlet ends, 1000 steps to t = 0.2."""
    from opensbli_amd.apps.sod_shock_tube import build
    _say("case Sod app 200 points, 1000 steps")
    prog = build(output=False)
    res = prog.run()
    q = np.stack([res[k] for k in ("rho", "rhou0", "rhoE")])
    check_parity(q, _oracle_run(oracle, prog, 1000), "Sod app vs oracle")
    g = np.loadtxt(os.path.join(os.path.dirname(__file__), "golden", "sod_exact_density_t0.2.txt"))
    x = np.arange(200) / 199.0
    l1 = np.abs(q[0] - np.interp(x, g[:, 0], g[:, 1])).mean()
    _say("L1 density error vs the exact solution", l1)
    assert l1 < 5e-3, l1


def test_euler_wave_app_against_oracle_and_exact(oracle):
    """apps/euler_wave: against the oracle at 64^2 (100 steps) and, as the
    reference runs it (400^2, 2500 steps to t = 2.5), against the translated
    wave."""
    from opensbli_amd.apps.euler_wave import build
    _say("case euler wave app 64^2 100 steps vs oracle; 400^2 2500 steps vs exact")
    prog = build(64, niter=100, output=False)
    res = prog.run()
    q = np.stack([res[k] for k in ("rho", "rhou0", "rhou1", "rhoE")])
    check_parity(q, _oracle_run(oracle, prog, 100), "Euler wave app vs oracle")
    res = build(output=False).run()
    n, t = 400, 2.5
    x = np.arange(n) * 2.0 / n
    y, xx = np.meshgrid(x, x, indexing="ij")
    exact = 1.0 + 0.2 * np.sin(np.pi * (xx + y - t * 0.5))
    e = res["rho"] - exact
    _say("L1 %.3e, Linf %.3e" % (np.abs(e).mean(), np.abs(e).max()))
    assert np.abs(e).max() < 1e-9  # measured 1.6e-10


def test_shu_osher_app_against_oracle(oracle):
    """apps/shu_osher at 400 points, dt 2e-3, 900 steps to t = 1.8 (the
    reference runs 3200 points): WENO5-Z with Dirichlet states from the
    initial condition at x = -5 and 5, against the oracle."""
    from opensbli_amd.apps.shu_osher import build
    _say("case Shu-Osher app 400 points, 900 steps")
    prog = build(400, dt="0.002", niter=900, output=False)
    res = prog.run()
    q = np.stack([res[k] for k in ("rho", "rhou0", "rhoE")])
    check_parity(q, _oracle_run(oracle, prog, 900), "Shu-Osher app vs oracle")


def test_simulation_monitor_and_iteration_print(capsys):
    """SimulationMonitor probes and print_iteration_ops on the Euler wave app:
    the printed iterations, times and probe values are those of the state
    after that many steps (read back from a second block run to the same
    step counts)."""
    from opensbli_amd import api, native
    from opensbli_amd.apps.euler_wave import build
    prog = build(32, niter=20, output=False)
    locs = [(3, 5), (10, 2)]
    prog.alg.simulation_monitor = api.SimulationMonitor(["rho", "rhoE"], locs, prog.block, print_frequency=7,
                                                        fp_precision=15)
    api.print_iteration_ops(every=5, NaN_check="rho")
    with capsys.disabled():
        _say("case monitor: Euler wave 32^2, 20 steps, probes every 7, iteration print every 5")
    prog.run()
    out = capsys.readouterr().out.splitlines()
    assert [ln for ln in out if ln.startswith("Iteration is")] == ["Iteration is %d" % k for k in (5, 10, 15, 20)]
    assert "Iteration, Time, rho(3, 5), rhoE(10, 2)" in out
    rows = [ln.split(", ") for ln in out if ln[:1].isdigit()]
    assert [int(r[0]) for r in rows] == [1, 7, 14]
    p = prog.problem()
    q0 = prog.initial_state(p)
    blk = native.Block(prog._attach_boundary_data(p, q0))
    try:
        blk.set_state(q0)
        done = 0
        for r in rows:
            k = int(r[0])
            blk.run(k - done)
            done = k
            q = blk.get_state()
            assert abs(float(r[1]) - k * p["dt"]) < 1e-14
            for v, (i, j), col in zip((0, 3), locs, r[2:]):
                assert abs(float(col) - q[v][j, i]) <= 1e-14 * max(1.0, abs(q[v][j, i])), (k, v)
                assert blk.point(v, (i, j)) == q[v][j, i]
    finally:
        blk.close()


def test_nan_count_and_check():
    """osbl_nan_count counts interior NaN values per variable; the facade's
    NaN check stops the run with FloatingPointError."""
    from opensbli_amd import api, native
    from opensbli_amd.apps.euler_wave import build
    _say("case NaN count: Euler wave 32^2 with planted NaN values")
    prog = build(32, niter=4, output=False)
    p = prog.problem()
    q0 = prog.initial_state(p)
    blk = native.Block(prog._attach_boundary_data(p, q0))
    try:
        blk.set_state(q0)
        assert blk.nan_count() == 0
        q = q0.copy()
        q[1][0, 0] = q[1][31, 31] = q[1][7, 19] = np.nan
        q[3][5, 5] = np.nan
        blk.set_state(q)
        assert (blk.nan_count(), blk.nan_count(0), blk.nan_count(1), blk.nan_count(3)) == (4, 0, 3, 1)
    finally:
        blk.close()
    bad = q0.copy()
    bad[0][4, 4] = np.nan
    prog.initial_state = lambda problem: bad
    api.print_iteration_ops(every=1, NaN_check="rho")
    with pytest.raises(FloatingPointError, match="NaN detected in rho"):
        prog.run()
