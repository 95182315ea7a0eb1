import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


from oracle.ref import SMALL_COMPONENT, TOL, parity_errors  # noqa: E402,F401  (shared with smoke())


def check_parity(q, qr, label="", tol=TOL):
    """Print the per-variable errors of q against qr and assert each is <= tol."""
    err = parity_errors(q, qr)
    print("%s rel L2 per variable: %s (bar %.0e)" % (label, " ".join("%.2e" % e for e in err), tol), flush=True)
    assert np.all(err <= tol), (label, err)
    return err


def rel_l2(a, b):
    """Relative L2 difference per conservative variable (leading axis)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = []
    for m in range(a.shape[0]):
        den = np.linalg.norm(b[m])
        num = np.linalg.norm(a[m] - b[m])
        out.append(num / den if den > 0 else num)
    return np.array(out)


@pytest.fixture(scope="session")
def oracle():
    from oracle import ref
    ref.build()
    return ref


def uniform_rho_p_state(n, nd, gama=1.4, Minf=0.1, seed=0, kmax=2):
    """Periodic state on [0, 2pi)^nd (interior arrays, x fastest) with uniform
    density and pressure and a smooth random velocity field with divergence
    and vorticity: every Fourier mode with |k_d| <= kmax, seeded amplitudes
    and phases (all modes present, so products of three of them interact)."""
    import itertools
    rs = np.random.RandomState(seed)
    x = np.arange(n) * (2.0 * np.pi / n)
    grids = np.meshgrid(*([x] * nd), indexing="ij")[::-1]     # grids[d] varies along direction d
    vel = []
    for _ in range(nd):
        v = np.zeros((n,) * nd)
        for k in itertools.product(range(-kmax, kmax + 1), repeat=nd):
            amp = rs.uniform(0.0, 0.3) / (1.0 + sum(c * c for c in k))
            v = v + amp * np.cos(sum(k[d] * grids[d] for d in range(nd)) + rs.uniform(0, 2 * np.pi))
        vel.append(v)
    rho = np.ones((n,) * nd)
    p = 1.0 / (gama * Minf * Minf)
    E = p / (gama - 1.0) + 0.5 * rho * sum(v * v for v in vel)
    return np.stack([rho] + [rho * v for v in vel] + [E])


def generic_state(n, nd, gama=1.4, seed=0):
    """Periodic state on [0, 2pi)^nd (interior arrays, x fastest) in which
    every variable varies at order one: density, pressure and each velocity
    component are sums of six seeded Fourier modes (|k_d| <= 2) about 1,
    1/gama and 0, Mach number of order 0.3.  No term of the residual is a
    small difference of large ones (unlike a low-Mach, divergence-free
    field such as the Taylor-Green vortex's initial state).  n: points per
    direction, one number or one per direction (x first)."""
    rs = np.random.RandomState(seed)
    ns = [n] * nd if np.isscalar(n) else list(n)
    axes = [np.arange(m) * (2.0 * np.pi / m) for m in ns]
    grids = np.meshgrid(*axes[::-1], indexing="ij")[::-1]  # grids[d] varies along direction d
    shape = tuple(ns[::-1])

    def field(mean, amp):
        f = np.full(shape, float(mean))
        for _ in range(6):
            k = rs.randint(-2, 3, nd)
            f = f + amp * rs.uniform(0.3, 1.0) * np.cos(sum(k[d] * grids[d] for d in range(nd)) +
                                                         rs.uniform(0.0, 2.0 * np.pi))
        return f

    rho = field(1.0, 0.08)
    p = field(1.0 / gama, 0.05)
    vel = [field(0.0, 0.25) for _ in range(nd)]
    E = p / (gama - 1.0) + 0.5 * rho * sum(v * v for v in vel)
    return np.stack([rho] + [rho * v for v in vel] + [E])


def kinetic_energy_imbalance(q, R):
    """Relative discrete kinetic-energy production of the residual R at state
    q: sum over cells of u.R_rhou - |u|^2 R_rho / 2, divided by the sum of the
    magnitudes of its terms.  For the skew-symmetric (Blaisdell) convective
    form at uniform density and pressure on a periodic grid the sum vanishes
    identically, because the central first derivative is antisymmetric there
    (sum a D b = -sum b D a); the conservative form has no such identity."""
    nd = q.shape[0] - 2
    rho = q[0]
    u = [q[1 + i] / rho for i in range(nd)]
    terms = [u[i] * R[1 + i] for i in range(nd)] + [-0.5 * sum(v * v for v in u) * R[0]]
    total = sum(t.sum() for t in terms)
    scale = sum(np.abs(t).sum() for t in terms)
    return abs(total) / scale
