"""The counter summary groups kernels into the families osbl_kernel_stats
reports and scales the counters by a copy's readings."""
import csv

from opensbli_amd import profiling


def _csv(path, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for r in rows:
            w.writerow(dict(zip(["Kernel_Name", "Counter_Name", "Counter_Value"], r)))


def test_family_mapping():
    assert profiling.family("void osbl::k_char_sweep_xl<3, 3, 1, true>(double const*)") == "char_sweep_x"
    assert profiling.family("void osbl::k_char_sweep<3, 1, 3, 1, true, -1>(double const*)") == "char_sweep_y"
    assert profiling.family("void osbl::k_char_sweep<3, 2, 3, 1, true, 0>(double const*)") == "char_sweep_z"
    assert profiling.family("void osbl::k_visc_update3d<3>(double const*)") == "residual_rk_update"
    assert profiling.family("osbl::k_halo_periodic(double*, int)") == "boundary"
    assert profiling.family("__amd_rocclr_copyBuffer") is None


def test_summary_scales_by_copy(tmp_path):
    gib = 2.0 ** 30
    _csv(tmp_path / "cf.csv", [("__amd_rocclr_copyBuffer", "FETCH_SIZE", gib / 1024 / 2)])
    _csv(tmp_path / "cw.csv", [("__amd_rocclr_copyBuffer", "WRITE_SIZE", gib / 1024)])
    _csv(tmp_path / "pf.csv", [("void osbl::k_char_sweep<3, 1, 3, 1, true, -1>(x)", "FETCH_SIZE", 100.0),
                               ("void osbl::k_char_sweep<3, 1, 3, 1, true, -1>(x)", "FETCH_SIZE", 300.0)])
    _csv(tmp_path / "pw.csv", [("void osbl::k_char_sweep<3, 1, 3, 1, true, -1>(x)", "WRITE_SIZE", 50.0),
                               ("void osbl::k_char_sweep<3, 1, 3, 1, true, -1>(x)", "WRITE_SIZE", 50.0)])
    fams, cal = profiling.summarise(str(tmp_path / "pf.csv"), str(tmp_path / "pw.csv"), str(tmp_path / "cf.csv"),
                                    str(tmp_path / "cw.csv"), gib)
    assert abs(cal["fetch_factor"] - 2.0) < 1e-12 and abs(cal["write_factor"] - 1.0) < 1e-12
    y = fams["char_sweep_y"]
    assert y["dispatches"] == 2
    assert abs(y["fetch_bytes"] - 2.0 * 200.0 * 1024) < 1e-9
    assert abs(y["write_bytes"] - 50.0 * 1024) < 1e-9
