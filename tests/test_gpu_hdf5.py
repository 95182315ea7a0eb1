"""HDF5 output, restart and input of the problem scripts on the GPU: iohdf5
writes (io_hdf5.py:99-127, every save_every iterations and after the loop,
algorithm.py:443-465), iohdf5 reads (opsc.py:702-705) and output_hdf5
(helperfunctions.py:103-127), through the C-ABI entry osbl_get_field_halo.

Writing a file leaves the evolution unchanged (bitwise), a file holds the
state with the halo layers the boundary conditions fill, and a restart from
the file written after k iterations reproduces the uninterrupted run bitwise.
"""
import os
import sys

import numpy as np
import pytest

from conftest import check_parity

pytestmark = pytest.mark.gpu



def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _need_h5():
    from opensbli_amd import hdf5
    if not os.path.exists(hdf5.H5_LIB_PATH):
        pytest.skip("libosbl_h5.so is not built (no HDF5 library)")
    return hdf5


NAMES3 = ["rho", "rhou0", "rhou1", "rhou2", "rhoE"]


def test_tgv_output_files_halos_and_restart(tmp_path, monkeypatch):
    hdf5 = _need_h5()
    monkeypatch.chdir(tmp_path)
    from opensbli_amd.apps.taylor_green import build
    _say("case TGV 16^3 central, 20 steps, save_every 10, restart from step 10")
    res_io = build(16, "central", niter=20, save_every=10).run()
    files = sorted(f for f in os.listdir(".") if f.endswith(".h5"))
    assert files == ["opensbli_output.h5", "opensbli_output_000010.h5", "opensbli_output_000020.h5"], files
    res = build(16, "central", niter=20).run()
    for nm in NAMES3:
        assert np.array_equal(res_io[nm], res[nm]), nm
        a, info = hdf5.decl_dat_hdf5("opensbli_output.h5", "opensbliblock00", nm + "_B0")
        assert info["size"] == [16, 16, 16]
        h = -info["d_m"][0]
        assert info["d_m"] == [-h] * 3 and info["d_p"] == [h] * 3 and h >= 2
        # every halo layer of a periodic block is the periodic image
        assert np.array_equal(a, np.pad(res[nm], h, mode="wrap")), nm
        b, _ = hdf5.decl_dat_hdf5("opensbli_output_000020.h5", "opensbliblock00", nm + "_B0")
        assert np.array_equal(a, b)
    assert hdf5.list_names("opensbli_output.h5") == ["opensbliblock00"]
    # restart from the file of iteration 10: ten more iterations give the 20-iteration state
    res_rs = build(16, "central", niter=10, restart="opensbli_output_000010.h5").run()
    for nm in NAMES3:
        assert np.array_equal(res_rs[nm], res[nm]), nm


def test_kh_app_with_hdf5_noise_against_oracle(tmp_path, monkeypatch, oracle):
    """The KH app: noise written by output_hdf5, read by iohdf5 into the
    initial condition; WENO5-Z, 20 steps, against the oracle; its output file
    holds the state and the coordinates over the widened grid."""
    hdf5 = _need_h5()
    monkeypatch.chdir(tmp_path)
    from opensbli_amd.apps.kelvin_helmholtz import build
    n = 32
    _say("case KH app 32^2 WENO-Z 20 steps, noise from data.h5")
    prog = build(n, niter=20, save_every=10, seed=4)
    res = prog.run()
    p = prog.problem()
    q0 = prog.initial_state(p)
    q = np.stack([res[k] for k in ("rho", "rhou0", "rhou1", "rhoE")])
    r = oracle.RefSolver(p)
    r.set_state(q0)
    r.run(20)
    check_parity(q, r.get_state(), "20 steps")
    names = hdf5.list_names("opensbli_output.h5", "opensbliblock00")
    assert sorted(names) == sorted(["rho_B0", "rhou0_B0", "rhou1_B0", "rhoE_B0", "x0_B0", "x1_B0"])
    for m, nm in enumerate(("rho", "rhou0", "rhou1", "rhoE")):
        assert np.array_equal(hdf5.read_interior("opensbli_output.h5", "opensbliblock00", nm + "_B0"), q[m])
    x0, info = hdf5.decl_dat_hdf5("opensbli_output.h5", "opensbliblock00", "x0_B0")
    h = -info["d_m"][0]
    i = np.arange(-h, n + info["d_p"][0])
    assert np.allclose(x0, np.broadcast_to(-0.5 + i / n, x0.shape), rtol=0, atol=1e-15)
    assert os.path.exists("opensbli_output_000010.h5") and os.path.exists("opensbli_output_000020.h5")


def test_walls_output_leaves_the_run_unchanged(tmp_path, monkeypatch):
    """Flat plate 2-D (walls, closures, stretched y): output every 7 iterations
    over 20; the state equals the uninterrupted run's with its boundary
    conditions applied, which the output files hold."""
    hdf5 = _need_h5()
    monkeypatch.chdir(tmp_path)
    from opensbli_amd.api import DataObject, iohdf5
    from opensbli_amd.apps.flat_plate import build
    _say("case flat plate 2-D 64x48, 20 steps, save_every 7")
    prog = build(n=(64, 48), lengths=(400.0, 115.0), niter=20)
    h5 = iohdf5(save_every=7, **{"iotype": "Write"})
    h5.add_arrays([DataObject(nm) for nm in ("rho", "rhou0", "rhou1", "rhoE", "x0", "x1")])
    prog.block.setio([h5])
    res_io = prog.run()
    prog2 = build(n=(64, 48), lengths=(400.0, 115.0), niter=20)
    _, blk = prog2.run(return_block=True)
    try:
        for m, nm in enumerate(("rho", "rhou0", "rhou1", "rhoE")):
            a, dm, dp = blk.get_field_halo(m)
            inner = a[-dm[1]:a.shape[0] - dp[1], -dm[0]:a.shape[1] - dp[0]]
            assert np.array_equal(res_io[nm], inner), nm
            b, info = hdf5.decl_dat_hdf5("opensbli_output.h5", "opensbliblock00", nm + "_B0")
            assert info["d_m"] == dm and info["d_p"] == dp
            assert np.array_equal(a, b), nm
    finally:
        blk.close()
    files = sorted(f for f in os.listdir(".") if f.endswith(".h5"))
    assert files == ["opensbli_output.h5", "opensbli_output_000007.h5", "opensbli_output_000014.h5"], files
    x1 = hdf5.read_interior("opensbli_output.h5", "opensbliblock00", "x1_B0")
    assert np.all(np.diff(x1[:, 0]) > 0) and x1[0, 0] == 0.0


def test_curvilinear_app_output_and_parity(tmp_path, monkeypatch, oracle):
    """The curvilinear Euler wave app (apps/euler_wave_curvilinear): 20 steps
    against the oracle; its output file holds the metric D00..D11 with the
    periodic images in the halos and the coordinates over the widened grid."""
    hdf5 = _need_h5()
    monkeypatch.chdir(tmp_path)
    from opensbli_amd.apps.euler_wave_curvilinear import build
    from opensbli_amd.problems import euler_wave_curvilinear, wavy_grid
    n = 32
    _say("case curvilinear euler wave app 32^2, 20 steps, save_every 10")
    res = build(n, niter=20, save_every=10).run()
    q = np.stack([res[k] for k in ("rho", "rhou0", "rhou1", "rhoE")])
    p, q0 = euler_wave_curvilinear(n, dt=0.0005)
    r = oracle.RefSolver(p)
    r.set_state(q0)
    r.run(20)
    check_parity(q, r.get_state(), "20 steps")
    names = hdf5.list_names("opensbli_output.h5", "opensbliblock00")
    want = ["rho", "rhou0", "rhou1", "rhoE", "x0", "x1", "D00", "D01", "D10", "D11"]
    assert sorted(names) == sorted(nm + "_B0" for nm in want)
    for idx, nm in enumerate(["D00", "D01", "D10", "D11"]):
        a, info = hdf5.decl_dat_hdf5("opensbli_output.h5", "opensbliblock00", nm + "_B0")
        h = -info["d_m"][0]
        assert np.allclose(a, np.pad(r.metric_field(idx), h, mode="wrap"), rtol=1e-14, atol=1e-14), nm
    x0, info = hdf5.decl_dat_hdf5("opensbli_output.h5", "opensbliblock00", "x0_B0")
    h = -info["d_m"][0]
    assert np.allclose(x0, wavy_grid((n, n), (2.0 / n, 2.0 / n), halo=h)[0], rtol=0, atol=1e-15)
    assert os.path.exists("opensbli_output_000010.h5") and os.path.exists("opensbli_output_000020.h5")


def test_flat_plate_writes_metric_and_teno_cutoff(tmp_path, monkeypatch, oracle):
    """The flat plate with the adaptive TENO writes D11 and TENO as
    katzer_SBLI.py:135 asks: D11 of the stretched wall-normal direction
    mirrored across its faces (metric.py:359-379) and the adaptive cut-off of
    the last stage (teno.py:430-443), both against the oracle."""
    hdf5 = _need_h5()
    monkeypatch.chdir(tmp_path)
    from opensbli_amd.api import DataObject, _adaptive_ct, iohdf5
    from opensbli_amd.apps.flat_plate import build
    _say("case flat plate 2-D 64x48 adaptive TENO5, 10 steps, output D11 and TENO")
    prog = build(n=(64, 48), lengths=(400.0, 115.0), niter=10, scheme="teno5")
    h5 = iohdf5(**{"iotype": "Write"})
    h5.add_arrays([DataObject(nm) for nm in ("rho", "rhou0", "rhou1", "rhoE", "x0", "x1", "D11", "TENO")])
    prog.block.setio([h5])
    res = prog.run()
    p = prog.problem()
    q0 = prog.initial_state(p)
    p = prog._attach_boundary_data(p, q0)
    r = oracle.RefSolver(p)
    r.set_state(q0)
    r.run(10)
    q = np.stack([res[k] for k in ("rho", "rhou0", "rhou1", "rhoE")])
    check_parity(q, r.get_state(), "state")
    d11, info = hdf5.decl_dat_hdf5("opensbli_output.h5", "opensbliblock00", "D11_B0")
    dm, dp = info["d_m"], info["d_p"]
    line = r.metric(1)[0]
    hj = -dm[1]
    want = np.pad(line, (hj, dp[1]), mode="reflect")  # D[-t] = D[t], D[N-1+t] = D[N-1-t]
    assert d11.shape == (48 + hj + dp[1], 64 - dm[0] + dp[0])
    # the library forms D11 on the host in its own evaluation order: ulp-level differences
    assert np.allclose(d11, want[:, None], rtol=1e-13, atol=0), np.abs(d11 - want[:, None]).max()
    teno = hdf5.read_interior("opensbli_output.h5", "opensbliblock00", "TENO_B0")
    ct = _adaptive_ct(r.sensor(), p["teno_a1"], p["teno_a2"])
    same = np.mean(teno == ct)
    _say("TENO cut-off equal to the oracle's at", same, "of the points; range", teno.min(), teno.max())
    assert same > 0.999 and teno.min() > 0
