"""HDF5 files of a simulation in the OPS layout, through libosbl_h5.so (the
C-ABI declared in include/opensbli_amd_h5.h).

A file holds one group per block ("opensbliblock00") and in it one dataset
per array, named "<array>_B<block number>" (rho_B0, rhou0_B0, x0_B0, ...),
with the array's halo layers and the OPS attributes (d_m, d_p, size, ...).
These are the files the reference's generated program writes with
ops_fetch_block_hdf5_file / ops_fetch_dat_hdf5_file (io_hdf5.py:99-127),
reads with ops_decl_dat_hdf5 (opsc.py:702-705), that its helper writes for
inputs (helperfunctions.py:80-127) and that its plot scripts read
(apps/Sod_shock_tube/plot.py:14-31).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
H5_LIB_PATH = os.path.join(HERE, "libosbl_h5.so")

# symbols the header declares (checked by the ABI test)
EXPORTS = ["osbl_h5_last_error", "osbl_h5_library_version", "osbl_h5_fetch_block", "osbl_h5_fetch_dat",
           "osbl_h5_dat_info", "osbl_h5_read_dat", "osbl_h5_list"]

E_ARG, E_IO, E_MISSING, E_FORMAT = -1, -2, -3, -4


class H5Error(OSError):
    pass


class H5MissingError(H5Error, KeyError):
    """The file, block group or dataset does not exist."""


_lib = None


def load():
    """Load libosbl_h5.so; raise if it has not been built (no HDF5 library)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(H5_LIB_PATH):
        raise H5Error(f"{H5_LIB_PATH} is missing: build it with opensbli_amd.build.build_h5() "
                      "(needs the HDF5 C library and headers)")
    L = ctypes.CDLL(H5_LIB_PATH)
    i32, i32p, cp = ctypes.c_int32, ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p
    dp = ctypes.POINTER(ctypes.c_double)
    L.osbl_h5_last_error.restype = cp
    L.osbl_h5_library_version.restype = ctypes.c_int
    L.osbl_h5_fetch_block.argtypes = [cp, cp, i32, i32, i32]
    L.osbl_h5_fetch_dat.argtypes = [cp, cp, i32, cp, i32, i32p, i32p, i32p, dp]
    L.osbl_h5_dat_info.argtypes = [cp, cp, cp, i32p, i32p, i32p, i32p]
    L.osbl_h5_read_dat.argtypes = [cp, cp, cp, dp, ctypes.c_int64]
    L.osbl_h5_list.argtypes = [cp, cp, cp, ctypes.c_int64]
    _lib = L
    return L


def _check(rc):
    if rc < 0:
        msg = load().osbl_h5_last_error().decode() or f"osbl_h5 error {rc}"
        raise (H5MissingError if rc == E_MISSING else H5Error)(msg)
    return rc


def _b(s):
    return None if s is None else os.fsencode(s) if isinstance(s, os.PathLike) else str(s).encode()


def library_version():
    v = _check(load().osbl_h5_library_version())
    return v // 10000, v // 100 % 100, v % 100


def dataset_name(array, block_number=0):
    """Dataset name of an array of a block: rho on block 0 -> rho_B0."""
    return "%s_B%d" % (array, block_number)


def fetch_block_hdf5_file(path, block_name, block_index, ndim, truncate=False):
    """Write the block's group and attributes (ops_fetch_block_hdf5_file)."""
    _check(load().osbl_h5_fetch_block(_b(path), _b(block_name), int(block_index), int(ndim), int(bool(truncate))))


def fetch_dat_hdf5_file(path, block_name, block_index, dat_name, data, d_m, d_p):
    """Write one array with its halos (ops_fetch_dat_hdf5_file).  data has its
    axes in reverse direction order (x last) and includes the halo layers;
    d_m[d] <= 0 and d_p[d] >= 0 are the halo depths of direction d (x first)."""
    a = np.ascontiguousarray(data, dtype=np.float64)
    nd = a.ndim
    if len(d_m) != nd or len(d_p) != nd:
        raise ValueError("d_m and d_p need one entry per dimension (%d)" % nd)
    size = [a.shape[nd - 1 - d] + int(d_m[d]) - int(d_p[d]) for d in range(nd)]
    arr = lambda v: (ctypes.c_int32 * 3)(*[int(x) for x in v])  # noqa: E731
    _check(load().osbl_h5_fetch_dat(_b(path), _b(block_name), int(block_index), _b(dat_name), nd, arr(size),
                                    arr(d_m), arr(d_p), a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


def dat_info(path, block_name, dat_name):
    """{'ndim', 'size', 'd_m', 'd_p'} of a dataset (per direction, x first)."""
    nd = ctypes.c_int32()
    s, m, p = (ctypes.c_int32 * 3)(), (ctypes.c_int32 * 3)(), (ctypes.c_int32 * 3)()
    _check(load().osbl_h5_dat_info(_b(path), _b(block_name), _b(dat_name), ctypes.byref(nd), s, m, p))
    n = nd.value
    return {"ndim": n, "size": list(s)[:n], "d_m": list(m)[:n], "d_p": list(p)[:n]}


def decl_dat_hdf5(path, block_name, dat_name):
    """Read a dataset with its halos (ops_decl_dat_hdf5): (array, info), the
    array's axes in reverse direction order (x last)."""
    info = dat_info(path, block_name, dat_name)
    nd = info["ndim"]
    shape = tuple(info["size"][d] - info["d_m"][d] + info["d_p"][d] for d in reversed(range(nd)))
    a = np.empty(shape)
    _check(load().osbl_h5_read_dat(_b(path), _b(block_name), _b(dat_name),
                                   a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size))
    return a, info


def read_interior(path, block_name, dat_name):
    """The interior points of a dataset (its halo layers removed), axes in
    reverse direction order -- what the reference's plot scripts read
    (apps/Sod_shock_tube/plot.py:19-31)."""
    a, info = decl_dat_hdf5(path, block_name, dat_name)
    nd = info["ndim"]
    sl = tuple(slice(-info["d_m"][d], -info["d_m"][d] + info["size"][d]) for d in reversed(range(nd)))
    return np.ascontiguousarray(a[sl])


def list_names(path, block_name=None):
    """Groups of the file, or datasets of a block group."""
    cap = 1 << 16
    while True:
        buf = ctypes.create_string_buffer(cap)
        rc = load().osbl_h5_list(_b(path), _b(block_name), buf, cap)
        if rc == E_ARG and cap < (1 << 24) and b"more than" in load().osbl_h5_last_error():
            cap *= 4
            continue
        _check(rc)
        return [n for n in buf.value.decode().split("\n") if n]
