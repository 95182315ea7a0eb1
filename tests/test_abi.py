"""The C-ABI library loads and exports every entry point include/*.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="opensbli_amd.h"):
    """Functions declared by one header of include/ (every header: None)."""
    names = []
    pattern = "*.h" if header is None else header
    for h in glob.glob(os.path.join(ROOT, "include", pattern)):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(osbl_\w+)\s*\(", src, flags=re.M):
            names.append(m.group(1))
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "osbl_create" in names and "osbl_run" in names and len(names) >= 15


def test_library_exports_every_declared_symbol():
    from opensbli_amd import native
    lib = native.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.osbl_abi_version() == 1
    assert sorted(native.EXPORTS) == declared_functions()


def test_every_header_is_covered():
    """Each function of include/*.h belongs to one library's export list."""
    from opensbli_amd import hdf5, native
    assert sorted(native.EXPORTS + hdf5.EXPORTS) == declared_functions(None)


def test_hdf5_library_exports_every_declared_symbol():
    from opensbli_amd import build, hdf5
    if not os.path.exists(hdf5.H5_LIB_PATH) and not build.hdf5_available():
        pytest.skip("no HDF5 library in this image")
    lib = hdf5.load()
    missing = [n for n in declared_functions("opensbli_amd_h5.h") if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(hdf5.EXPORTS) == declared_functions("opensbli_amd_h5.h")
    assert hdf5.library_version() >= (1, 8, 0)


def test_config_struct_layout_matches_header(tmp_path):
    """osbl_config as ctypes lays it out equals the C compiler's layout of the header."""
    import subprocess
    from opensbli_amd.native import OsblConfig
    fields = [f[0] for f in OsblConfig._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("{f} %zu\\n", offsetof(osbl_config, {f}));\n' for f in fields)
    src.write_text('#include <stddef.h>\n#include <stdio.h>\n#include "opensbli_amd.h"\nint main(void) {\n'
                   + body + '  printf("sizeof %zu\\n", sizeof(osbl_config));\n  return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(l.split() for l in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.splitlines())
    for f in fields:
        assert int(out[f]) == getattr(OsblConfig, f).offset, f
    assert int(out["sizeof"]) == ctypes.sizeof(OsblConfig)


def test_create_without_device_fails_cleanly():
    """Without a GPU the library reports an error instead of falling back."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from opensbli_amd import native
    from opensbli_amd.problems import taylor_green
    problem, _ = taylor_green(16)
    with pytest.raises(native.OsblError):
        native.Block(problem)


def test_invalid_config_rejected():
    from opensbli_amd import native
    from opensbli_amd.problems import taylor_green
    problem, _ = taylor_green(16)
    problem["n"] = [4, 16, 16]
    with pytest.raises(native.OsblError, match="at least 8"):
        native.Block(problem)


@pytest.mark.parametrize("change,msg", [
    (dict(conv="central_cons"), "WENO/TENO"),
    (dict(bc=[["periodic", "periodic"], ["extrap0", "extrap0"]]), "periodic boundaries"),
    (dict(ndim=1, n=[16], delta=[0.125], bc=[["periodic", "periodic"]]), "2-D or 3-D"),
])
def test_curvilinear_config_rejected(change, msg):
    """Curvilinear blocks are periodic Euler blocks with a WENO/TENO scheme;
    other configurations are refused before any device call."""
    from opensbli_amd import native
    from opensbli_amd.problems import euler_wave_curvilinear
    problem, _ = euler_wave_curvilinear(16)
    problem = {k: v for k, v in problem.items() if k != "grid_coords"}
    problem.update(change)
    with pytest.raises(native.OsblError, match=msg):
        native.Block(problem)


def test_body_force_needs_the_general_kernels():
    """The body force is refused on blocks without closures, metric or
    Sutherland viscosity (checked before any device call)."""
    from opensbli_amd import native
    from opensbli_amd.problems import taylor_green
    problem, _ = taylor_green(16, conv="weno_z")
    problem["body_force"] = [-1.0, 0.0, 0.0]
    with pytest.raises(native.OsblError, match="body force"):
        native.Block(problem)
