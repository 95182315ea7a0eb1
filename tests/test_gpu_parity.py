"""Parity of the HIP path (through the C-ABI) with the oracle on identical inputs.

Bar: relative L2 <= 1e-10 per conservative variable (FP64), after the step
counts of the benchmark configurations (100 steps) at sizes the oracle runs
in seconds (conftest.parity_errors: a momentum component that is small
against the momentum magnitude, such as the Taylor-Green vortex's rho*w which
starts at zero, is measured against that magnitude).  Every test prints its
per-variable errors.
"""
import sys

import numpy as np
import pytest

from conftest import TOL, check_parity, generic_state, kinetic_energy_imbalance, parity_errors, uniform_rho_p_state

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


@pytest.mark.parametrize("conv", ["central_skew", "central_cons", "weno_z", "weno_js", "teno6", "teno5"])
def test_rhs_generic3d(conv, oracle):
    """The residual (convective + viscous terms) at a state in which every
    variable varies at order one, so that no component of it is a
    cancellation: literal bar on every variable."""
    from opensbli_amd.problems import taylor_green
    _say(f"case rhs generic state conv={conv} n=24")
    problem, _ = taylor_green(24, conv=conv)
    problem["Minf"] = 1.0
    q0 = generic_state(24, 3, seed=1)
    g, r = _pair(problem, q0, oracle)
    check_parity(g.rhs(), r.rhs(), f"rhs generic {conv}")


@pytest.mark.parametrize("conv", ["central_skew", "central_cons", "weno_z", "weno_js", "teno6", "teno5"])
def test_rhs_tgv3d(conv, oracle):
    """The residual at the Taylor-Green vortex's initial state.  Its density
    and energy components are small differences of large flux derivatives
    (low Mach, divergence-free velocity), and with WENO5-Z the Z weights of
    the nearly constant characteristic fields are rounding-sensitive there:
    the oracle built with FMA contraction differs from itself by ~5e-10
    (rho) and ~8e-10 (rho*E).  Those two components of WENO5-Z are held to
    twice the oracle's own spread, measured here; every other component and
    scheme to the literal bar.  The bar that applies to the Taylor-Green
    workload, the state after 100 steps, is literal for every scheme
    (test_run_tgv3d_100_steps)."""
    from opensbli_amd.problems import taylor_green
    _say(f"case rhs tgv3d conv={conv} n=24")
    problem, q0 = taylor_green(24, conv=conv)
    g, r = _pair(problem, q0, oracle)
    rr = r.rhs()
    err = parity_errors(g.rhs(), rr)
    bar = np.full(len(err), TOL)
    if conv == "weno_z":
        f = oracle.RefSolver(problem, variant="fma")
        f.set_state(q0)
        spread = parity_errors(f.rhs(), rr)
        _say("oracle FMA-contraction spread per variable: " + " ".join("%.2e" % e for e in spread))
        bar[0] = max(TOL, 2.0 * spread[0])
        bar[-1] = max(TOL, 2.0 * spread[-1])
    _say(f"rhs tgv {conv} rel L2 per variable: " + " ".join("%.2e" % e for e in err) +
         " (bars " + " ".join("%.1e" % b for b in bar) + ")")
    assert np.all(err <= bar), (err, bar)


@pytest.mark.parametrize("conv,viscous,rk", [("central_skew", "const", "rk3"), ("central_skew", "const", "ls3"),
                                             ("central_cons", "const", "rk3"), ("central_skew", "none", "ls3"),
                                             ("central_cons", "none", "rk3")])
def test_tiled_central_130x130x49(conv, viscous, rk, oracle):
    """A block large enough for the tiled 3-D central kernels (the z-marching
    convective kernel, then the tiled viscous kernel with the RK update),
    with partial tiles in x and y and a one-plane z segment: the residual
    and 10 steps against the oracle at a generic state."""
    from opensbli_amd.problems import taylor_green
    n = (130, 130, 49)
    _say(f"case tiled central {conv} viscous={viscous} {rk} {n}")
    problem, _ = taylor_green(16, conv=conv, rk=rk, viscous=viscous)
    problem.update(n=list(n), delta=[2.0 * np.pi / m for m in n], Minf=1.0)
    q0 = generic_state(n, 3, seed=2)
    g, r = _pair(problem, q0, oracle)
    check_parity(g.rhs(), r.rhs(), f"tiled {conv} {viscous} rhs")
    g.set_state(q0)
    g.profile(True)
    g.run(10)
    stats = g.kernel_stats()
    g.profile(False)
    assert stats.get("central_convective", (0.0, 0))[1] > 0, stats  # the tiled path ran
    r.run(10)
    check_parity(g.get_state(), r.get_state(), f"tiled {conv} {viscous} {rk} 10 steps")


@pytest.mark.parametrize("nd", [2, 3])
def test_skew_form_kinetic_energy_identity(nd, oracle):
    """The skew-symmetric form's residual at uniform density and pressure
    produces no kinetic energy (sum u.R_rhou - |u|^2 R_rho / 2 = 0 to
    rounding on a periodic grid); the conservative form's does."""
    from opensbli_amd.native import Block
    from opensbli_amd.problems import taylor_green
    n = 32
    _say(f"case kinetic-energy identity nd={nd} n={n}")
    q = uniform_rho_p_state(n, nd, seed=nd)
    imb = {}
    for conv in ("central_skew", "central_cons"):
        problem, _ = taylor_green(n, conv=conv, viscous="none")
        if nd == 2:
            problem.update(ndim=2, n=[n, n], delta=problem["delta"][:2], bc=problem["bc"][:2])
        g = Block(problem)
        g.set_state(q)
        imb[conv] = kinetic_energy_imbalance(q, g.rhs())
        g.close()
    _say("relative kinetic-energy production", imb)
    assert imb["central_skew"] < 1e-12, imb
    assert imb["central_cons"] > 1e-5, imb


@pytest.mark.parametrize("conv", ["central_skew", "weno_z", "teno6", "weno_js", "teno5"])
def test_run_tgv3d_100_steps(conv, oracle):
    from opensbli_amd.problems import taylor_green
    n, steps = 24, 100
    _say(f"case run tgv3d conv={conv} n={n} steps={steps}")
    problem, q0 = taylor_green(n, conv=conv)
    g, r = _pair(problem, q0, oracle)
    g.run(steps)
    r.run(steps)
    check_parity(g.get_state(), r.get_state(), f"{conv} {steps} steps")


@pytest.mark.parametrize("conv", ["teno6", "central_skew", "weno_z"])
@pytest.mark.parametrize("rk", ["ls3", "ls3_ssp", "ls4"])
def test_run_tgv3d_lowstorage(rk, conv, oracle):
    from opensbli_amd.problems import taylor_green
    n, steps = 16, 20
    _say(f"case run tgv3d conv={conv} rk={rk} n={n} steps={steps}")
    problem, q0 = taylor_green(n, conv=conv, rk=rk)
    g, r = _pair(problem, q0, oracle)
    g.run(steps)
    r.run(steps)
    check_parity(g.get_state(), r.get_state(), f"{conv} {rk} {steps} steps")


@pytest.mark.parametrize("rk", ["rk3", "ls3"])
def test_c1_tgv64_central_100_steps(rk, oracle):
    """Configuration 1 at full size: 64^3, central skew, RK3 (the 3-register
    RungeKutta(3) and the low-storage RungeKuttaLS(3)), 100 steps."""
    from opensbli_amd.problems import taylor_green
    _say(f"case tgv 64^3 central skew {rk} 100 steps")
    problem, q0 = taylor_green(64, conv="central_skew", rk=rk)
    g, r = _pair(problem, q0, oracle)
    g.run(100)
    r.run(100)
    check_parity(g.get_state(), r.get_state(), f"tgv 64^3 central skew {rk}")


@pytest.mark.parametrize("conv", ["weno_z", "teno6", "central_skew", "central_cons"])
def test_2d_kh(conv, oracle):
    from opensbli_amd.problems import kelvin_helmholtz
    _say(f"case 2d kh conv={conv} n=64 steps=100")
    amps = np.random.RandomState(4).normal(0.0, 1e-2, 16)
    problem, q0 = kelvin_helmholtz(64, amps, conv=conv)
    g, r = _pair(problem, q0, oracle)
    check_parity(g.rhs(), r.rhs(), f"kh {conv} rhs")
    g.run(100)
    r.run(100)
    check_parity(g.get_state(), r.get_state(), f"kh {conv} 100 steps")


@pytest.mark.parametrize("conv", ["weno_z", "teno5", "teno6", "weno_js"])
@pytest.mark.parametrize("n", [200, 1037])
def test_1d_shock_tube(conv, n, oracle):
    from opensbli_amd.problems import shock_tube
    _say(f"case 1d sod conv={conv} n={n}")
    problem, q0 = shock_tube(n, conv=conv)
    g, r = _pair(problem, q0, oracle)
    nsteps = 100
    g.run(nsteps)
    r.run(nsteps)
    check_parity(g.get_state(), r.get_state(), f"sod {conv} n={n}")


def test_tgv_symmetry_256_weno_z():
    """Full size (256^3) WENO5-Z: properties that hold at any size.
    The TGV initial state and the scheme are invariant under x -> -x with
    (u, v, w) -> (-u, v, w) about x = pi, so rho stays mirror-symmetric; mass
    is conserved to rounding on the periodic grid."""
    from opensbli_amd.native import Block
    from opensbli_amd.problems import taylor_green
    _say("case tgv256 weno_z 2 steps symmetry/conservation")
    n = 256
    problem, q0 = taylor_green(n, conv="weno_z")
    g = Block(problem)
    g.set_state(q0)
    g.run(2)
    q = g.get_state()
    assert np.all(np.isfinite(q))
    m0, m1 = q0[0].sum(), q[0].sum()
    assert abs(m1 - m0) / abs(m0) < 1e-13
    # mirror about x = pi: index i -> (n - i) mod n
    idx = (-np.arange(n)) % n
    rho_m = q[0][:, :, idx]
    assert np.max(np.abs(rho_m - q[0])) < 1e-12 * np.max(np.abs(q[0]))
    ru_m = -q[1][:, :, idx]
    assert np.max(np.abs(ru_m - q[1])) < 1e-11 * np.max(np.abs(q[1]))


def test_api_script_tgv_weno_z(oracle):
    """The OpenSBLI-style script (HIPC target) runs on the GPU and matches the oracle."""
    from opensbli_amd.apps.taylor_green import build
    _say("case api tgv 16 weno_z 20 steps")
    prog = build(16, "weno_z", niter=20)
    out = prog.run()
    problem = prog.problem()
    q0 = prog.initial_state(problem)
    r = oracle.RefSolver(problem)
    r.set_state(q0)
    r.run(20)
    qr = r.get_state()
    q = np.stack([out[k] for k in ["rho", "rhou0", "rhou1", "rhou2", "rhoE"]])
    check_parity(q, qr, "api tgv weno_z")
