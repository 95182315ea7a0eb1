"""Parity of the Ducros sensor with the adaptive TENO cut-off (HIP vs oracle)."""
import sys

import numpy as np
import pytest

from conftest import check_parity

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(problem, q0, oracle, steps):
    from opensbli_amd.native import Block
    g = Block(problem)
    g.set_state(q0)
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    if steps == 0:
        a, b = g.rhs(), r.rhs()
    else:
        g.run(steps)
        r.run(steps)
        a, b = g.get_state(), r.get_state()
    assert np.isfinite(a).all()
    check_parity(a, b, "rhs" if steps == 0 else f"{steps} steps")


@pytest.mark.parametrize("conv", ["teno5", "teno6"])
@pytest.mark.parametrize("steps", [0, 100])
def test_tgv_adaptive_teno(conv, steps, oracle):
    from opensbli_amd.problems import taylor_green
    _say(f"case tgv adaptive {conv} 24^3 steps={steps}")
    problem, q0 = taylor_green(24, conv=conv)
    problem.update(sensor=1, teno_a1=10.5, teno_a2=4.5, sensor_eps=1e-30)
    _compare(problem, q0, oracle, steps)


@pytest.mark.parametrize("nd", [2, 3])
def test_flat_plate_adaptive_teno5(nd, oracle):
    from opensbli_amd.problems import flat_plate
    n = (64, 48) if nd == 2 else (40, 32, 8)
    _say(f"case flat plate adaptive teno5 nd={nd}")
    p, q0 = flat_plate(n=n, ndim=nd, conv="teno5", sensor=True)
    _compare(p, q0, oracle, 0)
    _compare(p, q0, oracle, 100)
