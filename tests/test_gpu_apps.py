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
