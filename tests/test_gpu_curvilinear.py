"""Curvilinear blocks on the GPU (through the C-ABI) against the oracle: the
metric from grid coordinates (metric.py:72-147), the residual of the strong
conservation form with the characteristic system about the metric normal
(apps/euler_wave_curvilinear/euler_wave.py), and the time loop.

Bar as in the other parity tests: relative L2 per conservative variable
<= 1e-10 (conftest.parity_errors).  The rotated-mesh case is also checked against the GPU's own Cartesian path (the
2-D characteristic fields are frame-invariant, so the two agree to rounding).
"""
import math
import sys

import numpy as np
import pytest

from conftest import check_parity, rel_l2

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _pair(problem, q0, oracle):
    from opensbli_amd.native import Block
    g = Block(problem)
    g.set_state(q0)
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    return g, r


def _problem3d(n, conv="weno_z", averaging="roe"):
    from opensbli_amd.problems import euler_wave_curvilinear_3d
    return euler_wave_curvilinear_3d(n, conv=conv, averaging=averaging)


def test_metric_matches_oracle():
    from opensbli_amd.native import Block
    from opensbli_amd.problems import euler_wave_curvilinear
    from oracle import ref
    _say("case curvilinear metric: wavy 2-D 48^2 and skewed wavy 3-D 16^3")
    for problem, q in (euler_wave_curvilinear(48), _problem3d(16)):
        nd = problem["ndim"]
        g = Block(problem)
        r = ref.RefSolver(problem)
        for idx in range(nd * nd + 1):
            a, b = g.metric_field(idx), r.metric_field(idx)
            assert np.allclose(a, b, rtol=1e-14, atol=1e-14), (nd, idx, np.abs(a - b).max())
        g.close()


@pytest.mark.parametrize("conv,avg", [("weno_z", "roe"), ("weno_js", "simple"), ("teno6", "roe"), ("teno5", "simple")])
def test_rhs_wavy_2d(conv, avg, oracle):
    from opensbli_amd.problems import euler_wave_curvilinear
    _say(f"case curvilinear rhs 2-D wavy 64^2 conv={conv} avg={avg}")
    problem, q0 = euler_wave_curvilinear(64, conv=conv, averaging=avg)
    g, r = _pair(problem, q0, oracle)
    check_parity(g.rhs(), r.rhs(), f"curvilinear 2-D rhs {conv} {avg}")


def test_euler_wave_curvilinear_100_steps(oracle):
    """The reference app's configuration (WENO5-Z, Roe, RK3 low storage, dt =
    5e-4, wavy mesh) at 64^2 for 100 steps."""
    from opensbli_amd.problems import euler_wave_curvilinear
    _say("case curvilinear euler wave 64^2, 100 steps")
    problem, q0 = euler_wave_curvilinear(64)
    g, r = _pair(problem, q0, oracle)
    g.run(100)
    r.run(100)
    check_parity(g.get_state(), r.get_state(), "curvilinear euler wave 100 steps")


def test_rotated_mesh_equals_the_cartesian_path(oracle):
    from opensbli_amd.native import Block
    from opensbli_amd.problems import euler_wave_curvilinear
    _say("case curvilinear rotated mesh 32^2 vs Cartesian, rhs and 20 steps")
    n, h, theta = 32, 2, 0.7
    d = 2.0 / n
    c, s = math.cos(theta), math.sin(theta)
    Rot = np.array([[c, -s], [s, c]])
    i = np.arange(-h, n + h) * d
    jj, ii = np.meshgrid(i, i, indexing="ij")
    coords = np.stack([Rot[0, 0] * ii + Rot[0, 1] * jj, Rot[1, 0] * ii + Rot[1, 1] * jj])
    problem, _ = euler_wave_curvilinear(n, coords=coords, dt=2e-3)
    cart = {k: v for k, v in problem.items() if k not in ("curvilinear", "grid_coords", "grid_halo")}
    x = np.arange(n) * d
    y, xx = np.meshgrid(x, x, indexing="ij")
    rho = 1.0 + 0.2 * np.sin(np.pi * xx) * np.cos(np.pi * y)
    ug = np.stack([0.4 + 0.3 * np.cos(np.pi * y), -0.2 + 0.25 * np.sin(np.pi * (xx + y))])
    E = (1.0 + 0.1 * np.cos(np.pi * (xx - y))) / 0.4 + 0.5 * rho * (ug ** 2).sum(0)
    qg = np.stack([rho, rho * ug[0], rho * ug[1], E])
    up = np.einsum("ab,bij->aij", Rot, ug)
    qp = np.stack([rho, rho * up[0], rho * up[1], E])

    def rotate(q):
        m = np.einsum("ab,bij->aij", Rot, q[1:3])
        return np.stack([q[0], m[0], m[1], q[3]])

    gc, gp = Block(cart), Block(problem)
    gc.set_state(qg)
    gp.set_state(qp)
    err = rel_l2(gp.rhs(), rotate(gc.rhs()))
    _say("rhs rel L2", err)
    assert np.all(err < 1e-12), err
    gc.run(20)
    gp.run(20)
    err = rel_l2(gp.get_state(), rotate(gc.get_state()))
    _say("20 steps rel L2", err)
    assert np.all(err < 1e-12), err
    gc.close()
    gp.close()


@pytest.mark.parametrize("conv", ["weno_z", "teno6"])
def test_skewed_wavy_3d(conv, oracle):
    _say(f"case curvilinear 3-D 16^3 conv={conv}: rhs and 20 steps")
    problem, q0 = _problem3d(16, conv=conv)
    g, r = _pair(problem, q0, oracle)
    check_parity(g.rhs(), r.rhs(), f"curvilinear 3-D rhs {conv}")
    g.run(20)
    r.run(20)
    check_parity(g.get_state(), r.get_state(), f"curvilinear 3-D {conv} 20 steps")


def test_run_needs_coordinates():
    from opensbli_amd.native import Block, OsblError
    from opensbli_amd.problems import euler_wave_curvilinear
    _say("case curvilinear block without coordinates refuses to run")
    problem, q0 = euler_wave_curvilinear(16)
    problem = {k: v for k, v in problem.items() if k != "grid_coords"}
    g = Block(problem)
    g.set_state(q0)
    with pytest.raises(OsblError, match="osbl_set_grid_coordinates"):
        g.run(1)
    g.close()
