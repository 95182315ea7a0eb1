"""Ensembles of independent 1-D Euler shock tubes of uneven size on one GPU.

The configuration: 512 tubes, N ~ U{200..4000}, rho_L ~ U[0.5,2], p_L ~
U[0.5,2], rho_R ~ U[0.1,0.5], p_R ~ U[0.05,0.3], u = 0, jump at x = 0.5,
transmissive (zeroth-order extrapolation) ends, WENO5-Z with Roe averaging,
RK3, constant dt = CFL * dx / S_max per tube with S_max the fastest wave of
the exact Riemann solution (OpenSBLI has only a constant dt, rk_LS.py:62-67),
integrated to t = 0.2 as the Sod app does (niter = ceil(0.2/dt),
Sod_shock_tube.py:115).
"""
import ctypes
import math
import time

import numpy as np

from . import native
from .problems import riemann_max_speed


def ensemble_parameters(ninst=512, seed=3, nmin=200, nmax=4000, cfl=0.5, t_end=0.2, gama=1.4):
    """Seeded draw of the ensemble (numpy RandomState: stream-stable)."""
    rs = np.random.RandomState(seed)
    n = rs.randint(nmin, nmax + 1, size=ninst)
    rho_l = rs.uniform(0.5, 2.0, ninst)
    p_l = rs.uniform(0.5, 2.0, ninst)
    rho_r = rs.uniform(0.1, 0.5, ninst)
    p_r = rs.uniform(0.05, 0.3, ninst)
    dx = 1.0 / (n - 1)
    smax = np.array([riemann_max_speed(rho_l[i], 0.0, p_l[i], rho_r[i], 0.0, p_r[i], gama) for i in range(ninst)])
    dt = cfl * dx / smax
    nsteps = np.ceil(t_end / dt).astype(np.int64)
    return dict(n=n.astype(np.int32), rho_l=rho_l, p_l=p_l, rho_r=rho_r, p_r=p_r, dx=dx, dt=dt, nsteps=nsteps,
                gama=gama, cfl=cfl, t_end=t_end, seed=seed)


def instance_problem(par, i, conv="weno_z", rk="rk3"):
    """Single-block problem dict and initial state of instance i."""
    n = int(par["n"][i])
    problem = dict(ndim=1, n=[n], delta=[float(par["dx"][i])], conv=conv, averaging="roe", viscous="none", rk=rk,
                   bc=[["extrap0", "extrap0"]], gama=par["gama"], dt=float(par["dt"][i]), eps=1e-16, teno_ct=1e-5)
    x = np.arange(n) * par["dx"][i]
    left = x < 0.5
    rho = np.where(left, par["rho_l"][i], par["rho_r"][i])
    p = np.where(left, par["p_l"][i], par["p_r"][i])
    q = np.stack([rho, 0.0 * rho, p / (par["gama"] - 1.0)])
    return problem, q


class Ensemble:
    """All instances of an ensemble on the GPU (osbl_batch_* of the C-ABI)."""

    def __init__(self, par, conv="weno_z", rk="rk3", nsteps=None):
        L = native.load()
        self.par = par
        self.ninst = len(par["n"])
        self.nsteps = np.asarray(par["nsteps"] if nsteps is None else nsteps, dtype=np.int64)
        base_problem, _ = instance_problem(par, 0, conv, rk)
        base = native.make_config(base_problem)
        self.h = ctypes.c_void_p()
        n = np.ascontiguousarray(par["n"], dtype=np.int32)
        dx = np.ascontiguousarray(par["dx"], dtype=np.float64)
        dt = np.ascontiguousarray(par["dt"], dtype=np.float64)
        st = np.ascontiguousarray(self.nsteps, dtype=np.int64)
        native.check(L.osbl_batch_create(ctypes.byref(base), self.ninst, n.ctypes.data, dx.ctypes.data,
                                         dt.ctypes.data, st.ctypes.data, ctypes.byref(self.h)))
        self.conv, self.rk = conv, rk
        for i in range(self.ninst):
            _, q = instance_problem(par, i, conv, rk)
            self.set_state(i, q)

    def close(self):
        if getattr(self, "h", None):
            native.load().osbl_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, i, q):
        a = np.ascontiguousarray(q, dtype=np.float64)
        native.check(native.load().osbl_batch_set_state(self.h, i, a.ctypes.data))

    def get_state(self, i):
        n = int(self.par["n"][i])
        a = np.empty((3, n))
        native.check(native.load().osbl_batch_get_state(self.h, i, a.ctypes.data))
        return a

    def save(self):
        native.check(native.load().osbl_batch_save(self.h))

    def restore(self):
        native.check(native.load().osbl_batch_restore(self.h))

    def run(self, mode="persistent"):
        ms = ctypes.c_double()
        native.check(native.load().osbl_batch_run(self.h, 0 if mode == "persistent" else 1, ctypes.byref(ms)))
        return ms.value

    def set_threads(self, max_n, threads):
        a = np.ascontiguousarray(max_n, dtype=np.int32)
        b = np.ascontiguousarray(threads, dtype=np.int32)
        native.check(native.load().osbl_batch_set_threads(self.h, len(a), a.ctypes.data, b.ctypes.data))

    def set_graph_steps(self, k):
        native.check(native.load().osbl_batch_set_graph_steps(self.h, int(k)))

    def cell_updates(self):
        """Cell updates of one full integration (cells x stages x steps)."""
        stages = 3
        return float(np.sum(self.par["n"].astype(np.float64) * self.nsteps * stages))


BUCKETS = [256, 512, 1024, 2048, 4096]
CANDIDATES = [64, 128, 256, 512]


def autotune(par, conv="weno_z", rk="rk3", probe_steps=40, buckets=BUCKETS, candidates=CANDIDATES, cache=None):
    """Pick the threads per workgroup of every size bucket by timing the
    persistent kernel on the bucket's largest instance for `probe_steps` steps.
    Returns (max_n, threads); results are memoised in `cache` (a dict) keyed by
    (scheme, rk, bucket)."""
    threads = []
    lo = 0
    for hi in buckets:
        key = (conv, rk, hi)
        if cache is not None and key in cache:
            threads.append(cache[key])
            lo = hi
            continue
        sel = np.where((par["n"] > lo) & (par["n"] <= hi))[0]
        if len(sel) == 0:
            threads.append(candidates[0])
            lo = hi
            continue
        i = int(sel[np.argmax(par["n"][sel])])
        sub = {k: (np.asarray(v)[[i]] if isinstance(v, np.ndarray) else v) for k, v in par.items()}
        sub["nsteps"] = np.array([probe_steps], dtype=np.int64)
        best, best_t = candidates[0], float("inf")
        for t in candidates:
            e = Ensemble(sub, conv, rk)
            e.set_threads([1 << 30], [t])
            e.run("persistent")  # warm-up
            ms = min(e.run("persistent") for _ in range(2))
            e.close()
            if ms < best_t:
                best, best_t = t, ms
        threads.append(best)
        if cache is not None:
            cache[key] = best
        lo = hi
    return list(buckets), threads


def plain_runs(par, conv="weno_z", rk="rk3", indices=None):
    """The per-instance baseline: every instance as its own single-block run
    (its own launches per stage), one after the other.  Returns seconds."""
    idx = range(len(par["n"])) if indices is None else indices
    total = 0.0
    for i in idx:
        problem, q = instance_problem(par, i, conv, rk)
        blk = native.Block(problem)
        blk.set_state(q)
        blk.sync()
        t0 = time.perf_counter()
        blk.run(int(par["nsteps"][i]))
        blk.sync()
        total += time.perf_counter() - t0
        blk.close()
    return total
