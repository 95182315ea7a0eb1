"""Ensembles of uneven 1-D shock tubes: the persistent and the graph-staged
executions agree bitwise with each other and for any threads-per-instance
table (same per-segment code), agree with running every instance alone
through the single-block path to rounding (relative L2 1e-12: the two are
separately compiled kernels, so FMA contraction may differ), and agree with
the oracle (relative L2 <= 1e-10 per variable)."""
import sys

import numpy as np
import pytest

from conftest import TOL, check_parity, parity_errors, rel_l2

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _small(ninst=12, seed=7, steps=30):
    from opensbli_amd.ensemble import ensemble_parameters
    par = ensemble_parameters(ninst=ninst, seed=seed, nmin=40, nmax=700)
    par["nsteps"] = np.minimum(par["nsteps"], steps + np.arange(ninst) % 5)
    return par


def _single(par, i):
    from opensbli_amd.ensemble import instance_problem
    from opensbli_amd.native import Block
    problem, q = instance_problem(par, i)
    b = Block(problem)
    b.set_state(q)
    b.run(int(par["nsteps"][i]))
    out = b.get_state()
    b.close()
    return out


def test_ensemble_modes_vs_single_runs():
    from opensbli_amd.ensemble import Ensemble
    _say("case ensemble persistent/staged vs single runs")
    par = _small()
    e = Ensemble(par)
    e.save()
    e.run("persistent")
    pers = [e.get_state(i) for i in range(len(par["n"]))]
    e.restore()
    e.set_graph_steps(7)
    e.run("staged")
    for i in range(len(par["n"])):
        got = e.get_state(i)
        assert np.array_equal(got, pers[i]), (i, np.abs(got - pers[i]).max())
        ref = _single(par, i)
        assert np.all(rel_l2(got, ref) < 1e-12), (i, rel_l2(got, ref))


def test_ensemble_thread_tables_agree():
    from opensbli_amd.ensemble import Ensemble
    _say("case ensemble thread tables")
    par = _small(ninst=6, seed=11)
    outs = []
    for t in [64, 128, 256, 512]:
        e = Ensemble(par)
        e.set_threads([1 << 30], [t])
        e.run("persistent")
        outs.append([e.get_state(i) for i in range(6)])
        e.close()
    for o in outs[1:]:
        for a, b in zip(outs[0], o):
            assert np.array_equal(a, b)


def test_ensemble_vs_oracle(oracle):
    from opensbli_amd.ensemble import Ensemble, instance_problem
    _say("case ensemble vs oracle")
    par = _small(ninst=4, seed=5, steps=60)
    e = Ensemble(par)
    e.run("persistent")
    for i in range(4):
        problem, q = instance_problem(par, i)
        r = oracle.RefSolver(problem)
        r.set_state(q)
        r.run(int(par["nsteps"][i]))
        check_parity(e.get_state(i), r.get_state(), f"instance {i} n={int(par['n'][i])}")


def test_configured_ensemble_512(oracle):
    """The configured draw (512 tubes, N ~ U{200..4000}, seed 3, to t = 0.2,
    autotuned threads): the persistent launch and the graph-staged launches
    agree bitwise on every instance, and the smallest, the largest and eight
    strided instances agree with the oracle after their full step counts."""
    from opensbli_amd.ensemble import Ensemble, autotune, ensemble_parameters, instance_problem
    _say("case ensemble 512 seed 3 as configured: persistent vs staged, 10 instances vs oracle")
    par = ensemble_parameters(512, seed=3)
    max_n, threads = autotune(par)
    e = Ensemble(par)
    e.set_threads(max_n, threads)
    e.save()
    ms_p = e.run("persistent")
    pers = [e.get_state(i) for i in range(512)]
    e.restore()
    ms_s = e.run("staged")
    _say("persistent %.1f ms, staged %.1f ms" % (ms_p, ms_s))
    for i in range(512):
        got = e.get_state(i)
        assert np.array_equal(got, pers[i]), (i, np.abs(got - pers[i]).max())
    work = par["n"].astype(float) * par["nsteps"]
    pick = sorted({int(np.argmin(par["n"])), int(np.argmax(par["n"])), int(np.argmax(work))}
                  | set(range(5, 512, 64)))
    for i in pick:
        problem, q = instance_problem(par, i)
        r = oracle.RefSolver(problem)
        r.set_state(q)
        r.run(int(par["nsteps"][i]))
        err = parity_errors(pers[i], r.get_state())
        _say(i, int(par["n"][i]), int(par["nsteps"][i]), err)
        assert np.all(err <= TOL), (i, err)
