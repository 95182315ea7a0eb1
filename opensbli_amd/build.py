"""Build recipe of libosbl_hip.so (hipcc, gfx950), in-tree."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libosbl_hip.so")
SOURCES = ["osbl_kernels.hip", "osbl_runtime.hip", "osbl_batch.hip", "osbl_adjoint.hip", "osbl_general.hip",
           "osbl_boundary.hip", "osbl_adjoint_char.hip", "osbl_filters.hip", "osbl_adjoint_visc.hip",
           "osbl_curvilinear.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# HDF5 C library of the image (headers and libhdf5.so); OSBL_HDF5_PREFIX overrides
HDF5_PREFIX = os.environ.get("OSBL_HDF5_PREFIX", "/opt/conda")
H5_LIB = os.path.join(HERE, "libosbl_h5.so")
ARCH = os.environ.get("OSBL_ARCH", "gfx950")
FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-result",
]


def build(verbose=False, extra=(), jobs=None):
    """Compile every HIP source (in parallel) and link libosbl_hip.so next to this file."""
    from concurrent.futures import ThreadPoolExecutor
    objs, cmds = [], []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmds.append([HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    jobs = jobs or min(len(cmds), max(1, (os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for cmd in cmds:
            if verbose:
                print(" ".join(cmd), flush=True)
        results = list(ex.map(lambda c: subprocess.run(c), cmds))
    bad = [c for c, r in zip(cmds, results) if r.returncode != 0]
    if bad:
        raise subprocess.CalledProcessError(1, bad[0])
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


def hdf5_available(prefix=None):
    prefix = prefix or HDF5_PREFIX
    return os.path.exists(os.path.join(prefix, "include", "hdf5.h")) and any(
        os.path.exists(os.path.join(prefix, "lib", n)) for n in ("libhdf5.so", "libhdf5.so.103"))


def build_h5(verbose=False):
    """Compile libosbl_h5.so (host C over the HDF5 library) next to this file.
    Returns its path, or None when the HDF5 library is not installed."""
    if not hdf5_available():
        if verbose:
            print("no HDF5 library under %s: libosbl_h5.so not built" % HDF5_PREFIX, flush=True)
        return None
    inc, lib = os.path.join(HDF5_PREFIX, "include"), os.path.join(HDF5_PREFIX, "lib")
    tmp = H5_LIB + ".tmp"
    cmd = [os.environ.get("CC", "gcc"), "-O2", "-std=c11", "-D_DEFAULT_SOURCE", "-Wall", "-Wextra", "-fPIC",
           "-shared", "-I" + inc, os.path.join(CSRC, "osbl_h5.c"), "-L" + lib, "-lhdf5",
           "-Wl,--enable-new-dtags,-rpath," + lib, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, H5_LIB)
    return H5_LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
    print(build_h5(verbose="-v" in sys.argv))
