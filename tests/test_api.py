"""The OpenSBLI-style problem-definition surface (opensbli_amd.api) builds
the same configuration and initial state as the direct problem helpers, and
refuses what the kernels do not implement."""
import numpy as np
import pytest


@pytest.mark.parametrize("scheme", ["central", "weno_z", "weno_js", "teno5", "teno6"])
def test_tgv_problem_from_script(scheme):
    from opensbli_amd.apps.taylor_green import build
    from opensbli_amd.problems import taylor_green
    prog = build(16, scheme)
    p = prog.problem()
    conv = "central_skew" if scheme == "central" else scheme
    ref, q_ref = taylor_green(16, conv=conv)
    for k in ["ndim", "n", "conv", "viscous", "rk", "averaging", "bc"]:
        assert p[k] == ref[k], k
    for k in ["dt", "gama", "Minf", "Re", "Pr"]:
        assert abs(p[k] - ref[k]) < 1e-15 * abs(ref[k]), k
    assert np.allclose(p["delta"], ref["delta"], rtol=1e-15, atol=0)
    q = prog.initial_state(p)
    assert np.abs(q - q_ref).max() < 1e-13


def test_rejects_unsupported_schemes():
    from opensbli_amd import api
    with pytest.raises(NotImplementedError):
        api.Central(6)
    with pytest.raises(NotImplementedError):
        api.LLFWeno(7, formulation="Z")
    with pytest.raises(NotImplementedError):
        api.LLFTeno(8)
    with pytest.raises(NotImplementedError):
        api.RungeKutta(4)
    with pytest.raises(NotImplementedError):
        api.SimulationDataType.set_datatype(api.FloatC)


def test_rejects_unknown_equation():
    from opensbli_amd import api
    ee = api.EinsteinEquation()
    sim = api.SimulationEquations()
    sim.add_equations(ee.expand("Eq(Der(rho,t), - Conservative(rhou_j,x_j) + c_j)", 3, "x", [], []))
    sim.add_equations(ee.expand("Eq(Der(rhou_i,t), -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j))", 3, "x", [], []))
    sim.add_equations(ee.expand("Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j))", 3, "x", [], []))
    block = api.SimulationBlock(3)
    block.set_block_boundaries([api.PeriodicBC(d, s) for d in range(3) for s in (0, 1)])
    block.set_equations([sim])
    c = api.Central(4)
    r = api.RungeKutta(3)
    block.set_discretisation_schemes({c.name: c, r.name: r})
    block.discretise()
    with pytest.raises(NotImplementedError):
        api.HIPC(api.TraditionalAlgorithmRK(block))


def test_parameter_expressions():
    from opensbli_amd.apps.taylor_green import build
    prog = build(32, "central")
    assert abs(prog.values["Delta0block0"] - 2 * np.pi / 32) < 1e-16
    from opensbli_amd import api
    prog.set_values(["niter"], ["ceil(0.2/0.0002)"])
    assert prog.values["niter"] == 1000


def test_flat_plate_script_matches_problem():
    """The SBLI-style script (metric, Sutherland, walls, Katzer profile) gives
    the same block as problems.flat_plate without the noise."""
    from opensbli_amd.apps import flat_plate as fp
    from opensbli_amd.problems import flat_plate
    prog = fp.build(n=(40, 32, 8))
    p = prog.problem()
    q0 = prog.initial_state(p)
    p = prog._attach_boundary_data(p, q0)
    ref_p, ref_q = flat_plate(n=(40, 32, 8), noise=0.0)
    for k in ("ndim", "n", "conv", "averaging", "viscous", "rk", "bc", "stretch", "gama", "Minf", "Re", "Pr",
              "SuthT", "RefT"):
        assert p[k] == ref_p[k], k
    assert np.allclose(p["delta"], ref_p["delta"], rtol=1e-15)
    assert p["dt"] == ref_p["dt"]
    assert p["bscheme"][:2] == ref_p["bscheme"][:2]
    assert np.abs(p["coords"][1] - ref_p["coords"][1]).max() < 1e-12
    for f in ((0, 0), (1, 1)):
        assert np.abs(p["planes"][f] - ref_p["planes"][f]).max() < 1e-10
    assert np.abs(q0 - ref_q).max() < 1e-15


def test_facade_rejects_unsupported_boundaries():
    from opensbli_amd import api
    with pytest.raises(ValueError):
        api.InletPressureExtrapolateBC(0, 1)
    with pytest.raises(ValueError):
        api.ExtrapolationBC(0, 0, order=2)
    m = api.MetricsEquation()
    with pytest.raises(NotImplementedError):
        m.generate_transformations(2, "x", [(True, True), (True, False)], 2)
    with pytest.raises(ValueError):
        m.generate_transformations(2, "x", [(False, False)], 2)


def test_flat_plate_adaptive_teno_script():
    """The Katzer-style script with the Ducros sensor and adaptive TENO5."""
    from opensbli_amd.apps import flat_plate as fp
    from opensbli_amd.problems import flat_plate
    prog = fp.build(n=(40, 32), scheme="teno5")
    p = prog.problem()
    ref_p, _ = flat_plate(n=(40, 32), ndim=2, noise=0.0, conv="teno5", sensor=True)
    for k in ("conv", "sensor", "teno_a1", "teno_a2", "sensor_eps", "viscous", "stretch", "bc"):
        assert p[k] == ref_p[k], k


def test_filters_from_script():
    """WENOFilter, SFD and BinomialFilter on the TGV script: the problem
    carries the filters with the reference's default constants, substituted
    values override them, and the binomial grid condition becomes a mask."""
    from opensbli_amd.apps.taylor_green import build
    prog = build(16, "central", filters=("weno", "sfd", "binomial"))
    p = prog.problem()
    f = p["filters"]
    assert f["weno"] == {"sensor": "ducros", "mach_correction": False, "control": 1.0}
    assert p["sensor_eps"] == 1.0e-12
    assert f["sfd"] == {"chi": 0.1, "omega": 1.0 / 0.75}
    assert f["binomial"] == {"order": 4, "sigma": 0.05}
    from opensbli_amd.api import substitute_simulation_parameters
    substitute_simulation_parameters(["sigma_filt", "shock_filter_control", "epsilon"], ["0.2", "0.5", "1e-10"])
    p = prog.problem()
    assert p["filters"]["binomial"]["sigma"] == 0.2 and p["filters"]["weno"]["control"] == 0.5
    assert p["sensor_eps"] == 1e-10
    q0 = prog.initial_state(p)
    p = prog._attach_boundary_data(p, q0)
    mask = p["filters"]["binomial"]["mask"]
    assert mask.shape == (16, 16, 16) and mask.dtype == np.uint8
    assert mask[:, :, :8].all() and not mask[:, :, 8:].any()


def test_filter_constructors_follow_the_reference_checks():
    from opensbli_amd import api
    b2 = api.SimulationBlock(2)
    with pytest.raises(ValueError, match="even orders"):
        api.BinomialFilter(b2, 3)
    with pytest.raises(ValueError, match="halo points"):
        api.BinomialFilter(b2, 12)
    with pytest.raises(NotImplementedError, match="dissipation sensor"):
        api.WENOFilter(b2, 5, dissipation_sensor="Yee")
    w = api.WENOFilter(b2, 5, dissipation_sensor="Constant", Mach_correction=True)
    assert b2.shock_filter and w.sensor == "constant" and w.mach_correction


def test_curvilinear_app_matches_problem():
    """The curvilinear Euler wave script (apps/euler_wave_curvilinear) gives the
    problem and initial state of problems.euler_wave_curvilinear, with the
    coordinates evaluated over the interior and 2 points beyond it."""
    from opensbli_amd.apps.euler_wave_curvilinear import build
    from opensbli_amd.problems import euler_wave_curvilinear
    prog = build(32, niter=10)
    p = prog.problem()
    ref, q_ref = euler_wave_curvilinear(32, dt=0.0005)
    for k in ["ndim", "n", "conv", "averaging", "viscous", "rk", "bc", "curvilinear"]:
        assert p[k] == ref[k], k
    assert np.allclose(p["delta"], ref["delta"], rtol=1e-15, atol=0)
    q = prog.initial_state(p)
    p = prog._attach_boundary_data(p, q)
    assert p["grid_halo"] == 2
    assert np.array_equal(p["grid_coords"], ref["grid_coords"])
    assert np.array_equal(q, q_ref)


def test_curvilinear_rhs_through_the_facade_matches_oracle_problem(oracle):
    """The facade's problem runs in the oracle exactly like the helper's."""
    from opensbli_amd.apps.euler_wave_curvilinear import build
    from opensbli_amd.problems import euler_wave_curvilinear
    prog = build(16, niter=2)
    p = prog.problem()
    q = prog.initial_state(p)
    p = prog._attach_boundary_data(p, q)
    ref, q_ref = euler_wave_curvilinear(16, dt=0.0005)
    a = oracle.RefSolver(p)
    a.set_state(q)
    b = oracle.RefSolver(ref)
    b.set_state(q_ref)
    assert np.array_equal(a.rhs(), b.rhs())


def _curvilinear_block(metric_params, sim_eqs, with_metric=True):
    from opensbli_amd import api
    ee = api.EinsteinEquation()
    sim = api.SimulationEquations()
    for t in sim_eqs:
        sim.add_equations(ee.expand(t, 2, "xi", [], ["gama"]))
    cr = api.ConstituentRelations()
    for t in ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE-rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
              "Eq(a, (gama*p/rho)**0.5)", "Eq(U_i, D_i_j*u_j)"]:
        cr.add_equations(ee.expand(t, 2, "xi", [], ["gama"]))
    block = api.SimulationBlock(2)
    block.set_block_boundaries([api.PeriodicBC(d, s) for d in range(2) for s in (0, 1)])
    classes = [cr, sim]
    if with_metric:
        m = api.MetricsEquation()
        m.generate_transformations(2, "x", metric_params, 2)
        classes.append(m)
    block.set_equations(classes)
    w = api.LLFWeno(5, formulation="Z", averaging=api.RoeAverage([0, 1]))
    c = api.Central(4)
    r = api.RungeKuttaLS(3)
    block.set_discretisation_schemes({w.name: w, c.name: c, r.name: r})
    block.discretise()
    return api.HIPC(api.TraditionalAlgorithmRK(block))


CURV_EQS = ["Eq(Der(rho,t), - Conservative(detJ * rho*U_j,xi_j,**{'scheme':'Weno'})/detJ)",
            "Eq(Der(rhou_i,t) , -  Conservative(detJ * (rhou_i*U_j + p*D_j_i), xi_j , **{'scheme':'Weno'})/detJ)",
            "Eq(Der(rhoE,t), - Conservative(detJ * (p+rhoE)*U_j,xi_j, **{'scheme':'Weno'})/detJ)"]


def test_curvilinear_facade_checks():
    from opensbli_amd import api
    prog = _curvilinear_block([(True, True), (True, True)], CURV_EQS)
    assert prog.scheme["curvilinear"] == 1 and prog.scheme["conv"] == "weno_z"
    # metric in one direction only
    m = api.MetricsEquation()
    with pytest.raises(NotImplementedError, match="every direction"):
        m.generate_transformations(2, "x", [(True, True), (False, False)], 2)
    # strong-conservation equations without a curvilinear metric on the block
    with pytest.raises(ValueError, match="MetricsEquation"):
        _curvilinear_block(None, CURV_EQS, with_metric=False)
    # a wrong momentum flux (the transposed metric) is not recognised
    bad = list(CURV_EQS)
    bad[1] = bad[1].replace("p*D_j_i", "p*D_i_j")
    with pytest.raises(NotImplementedError, match="not recognised"):
        _curvilinear_block([(True, True), (True, True)], bad)


def test_apply_metrics_with_curvilinear_metric_is_refused():
    """Cartesian-form equations transformed by a curvilinear metric are not the
    stretched case; the facade refuses them rather than treating them as such."""
    from opensbli_amd import api
    ee = api.EinsteinEquation()
    sim = api.SimulationEquations()
    sc = "**{'scheme':'Weno'}"
    for t in ["Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc,
              "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s)  )" % sc,
              "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s))" % sc]:
        sim.add_equations(ee.expand(t, 2, "x", [], ["gama"]))
    m = api.MetricsEquation()
    m.generate_transformations(2, "x", [(True, True), (True, True)], 2)
    sim.apply_metrics(m)
    cr = api.ConstituentRelations()
    for t in ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE-rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
              "Eq(a, (gama*p/rho)**0.5)"]:
        cr.add_equations(ee.expand(t, 2, "x", [], ["gama"]))
    block = api.SimulationBlock(2)
    block.set_block_boundaries([api.PeriodicBC(d, s) for d in range(2) for s in (0, 1)])
    block.set_equations([cr, sim])
    w = api.LLFWeno(5, formulation="Z", averaging=api.RoeAverage([0, 1]))
    c = api.Central(4)
    r = api.RungeKuttaLS(3)
    block.set_discretisation_schemes({w.name: w, c.name: c, r.name: r})
    block.discretise()
    with pytest.raises(NotImplementedError, match="strong conservation form"):
        api.HIPC(api.TraditionalAlgorithmRK(block))


def test_channel_body_force_is_recognised():
    """The TENO channel's equations (apps/channel_flow/compressible_TCF_TENO/
    turbulent_channel.py:14-16): conservative TENO fluxes, viscous terms and
    the body force -KD(_i,_j)*c_j, -Dot(c_j, u_j) with c0, c1 from the
    parameters."""
    from opensbli_amd import api
    sc = "**{'scheme':'Teno'}"
    eqs = ["Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc,
           "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s) + Der(tau_i_j,x_j) - KD(_i,_j)*c_j )"
           % sc,
           "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s) - Der(q_j,x_j) + Der(u_i*tau_i_j ,x_j) - "
           "Dot(c_j, u_j))" % sc]
    subs = ["Eq(tau_i_j, (mu/Re)*(Der(u_i,x_j)+ Der(u_j,x_i) - (2/3)* KD(_i,_j)* Der(u_k,x_k)))",
            "Eq(q_j, (-mu/((gama-1)*Minf*Minf*Pr*Re))*Der(T,x_j))"]
    consts = ["Re", "Pr", "gama", "Minf", "c_j", "RefT", "SuthT"]
    ee = api.EinsteinEquation()
    sim = api.SimulationEquations()
    for t in eqs:
        sim.add_equations(ee.expand(t, 2, "x", subs, consts))
    cr = api.ConstituentRelations()
    for t in ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
              "Eq(a, (gama*p/rho)**0.5)", "Eq(T, p*gama*Minf*Minf/(rho))",
              "Eq(mu, (T**(1.5)*(1.0+SuthT/RefT)/(T+SuthT/RefT)))"]:
        cr.add_equations(ee.expand(t, 2, "x", subs, consts))
    block = api.SimulationBlock(2)
    block.set_block_boundaries([api.PeriodicBC(0, 0), api.PeriodicBC(0, 1), api.AdiabaticWallBC(1, 0),
                                api.AdiabaticWallBC(1, 1)])
    block.set_equations([cr, sim])
    t = api.LLFTeno(5, averaging=api.RoeAverage([0, 1]))
    c = api.StoreSome(4, "u0 u1 T")
    r = api.RungeKuttaLS(3)
    block.set_discretisation_schemes({t.name: t, c.name: c, r.name: r})
    block.discretise()
    prog = api.HIPC(api.TraditionalAlgorithmRK(block))
    assert prog.scheme["force"] == 1 and prog.scheme["viscous"] == "sutherland"
    prog.set_values(["Re", "Pr", "gama", "Minf", "dt", "block0np0", "block0np1", "Delta0block0", "Delta1block0",
                     "SuthT", "RefT"], ["190.0", "0.7", "1.4", "0.1", "1e-4", "16", "17", "0.1", "0.1", "110.4", "273.0"])
    with pytest.raises(ValueError, match="c0"):
        prog.problem()
    prog.set_values(["c0", "c1"], ["-1", "0"])
    assert prog.problem()["body_force"] == [-1.0, 0.0]


def _euler_block(eqs, rels):
    """A periodic 2-D block with the given Euler equations and constituent
    relations (WENO5-Z, RK3), discretised."""
    from opensbli_amd import api
    ee = api.EinsteinEquation()
    sim = api.SimulationEquations()
    cr = api.ConstituentRelations()
    for t in eqs:
        sim.add_equations(ee.expand(t, 2, "x", [], ["gama", "Minf"]))
    for t in rels:
        cr.add_equations(ee.expand(t, 2, "x", [], ["gama", "Minf"]))
    block = api.SimulationBlock(2)
    block.set_block_boundaries([api.PeriodicBC(d, s) for d in range(2) for s in (0, 1)])
    llf = api.LLFWeno(5, formulation="Z", averaging=api.RoeAverage([0, 1]))
    rk = api.RungeKutta(3)
    block.set_discretisation_schemes({llf.name: llf, rk.name: rk})
    block.set_equations([cr, sim])
    block.discretise()
    return block


def test_equations_are_recognised_up_to_algebra():
    """Reordered factors and terms, and a factor distributed over a sum, are
    the same equations; a different flux is refused."""
    from opensbli_amd import api
    sc = "**{'scheme':'Weno'}"
    ref_eqs = ["Eq(Der(rho,t), - Conservative(rhou_j,x_j,%s))" % sc,
               "Eq(Der(rhou_i,t) , -Conservative(rhou_i*u_j + KD(_i,_j)*p,x_j , %s)  )" % sc,
               "Eq(Der(rhoE,t), - Conservative((p+rhoE)*u_j,x_j, %s))" % sc]
    ref_rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (gama-1)*(rhoE - rho*(1/2)*(KD(_i,_j)*u_i*u_j)))",
                "Eq(a, (gama*p/rho)**0.5)"]
    alt_eqs = ["Eq(Der(rho,t), -Conservative(rhou_j, x_j, %s))" % sc,
               "Eq(Der(rhou_i,t), -Conservative(p*KD(_i,_j) + u_j*rhou_i, x_j, %s))" % sc,
               "Eq(Der(rhoE,t), -Conservative(u_j*rhoE + u_j*p, x_j, %s))" % sc]
    alt_rels = ["Eq(u_i, rhou_i/rho)", "Eq(p, (rhoE - (1/2)*rho*(KD(_i,_j)*u_j*u_i))*(gama-1))",
                "Eq(a, (p*gama/rho)**0.5)"]
    want = api._recognise(_euler_block(ref_eqs, ref_rels))
    assert want["conv"] == "weno_z"
    assert api._recognise(_euler_block(alt_eqs, alt_rels)) == want
    bad = list(ref_eqs)
    bad[2] = "Eq(Der(rhoE,t), - Conservative((p+2*rhoE)*u_j,x_j, %s))" % sc
    with pytest.raises(NotImplementedError, match="not recognised"):
        api._recognise(_euler_block(bad, ref_rels))
    bad_rels = list(ref_rels)
    bad_rels[1] = "Eq(p, (gama-1)*(rhoE - rho*(KD(_i,_j)*u_i*u_j)))"
    with pytest.raises(NotImplementedError, match="constituent relation for p"):
        api._recognise(_euler_block(ref_eqs, bad_rels))


def test_simulation_monitor_checks_and_format():
    """SimulationMonitor's checks (simulation_monitors.py:27-29) and its
    header and line formats."""
    from opensbli_amd import api
    block = api.SimulationBlock(2)
    with pytest.raises(ValueError, match="number of arrays"):
        api.SimulationMonitor(["rho"], [(1, 2), (3, 4)], block)
    with pytest.raises(NotImplementedError, match="conservative arrays"):
        api.SimulationMonitor(["p"], [(1, 2)], block)
    with pytest.raises(ValueError, match="indices"):
        api.SimulationMonitor(["rho"], [(1, 2, 3)], block)
    m = api.SimulationMonitor([api.DataObject("rho"), "rhou1"], [(1, 2), (3, 4)], block, print_frequency=10,
                              fp_precision=3)
    assert m.header() == "Iteration, Time, rho(1, 2), rhou1(3, 4)"
    assert m.line(10, 0.01, [1.0, -0.5]) == "10, 0.010, 1.000, -0.500"
