"""Pins of the oracle's modified Ducros sensor and adaptive TENO cut-off
(shock_sensors.py:12-49, teno.py:430-443)."""
import math

import numpy as np
import pytest

from oracle.ref import RefSolver


def _block2d(n, q, conv="teno5", bc="periodic", **kw):
    p = dict(ndim=2, n=[n, n], delta=[1.0 / n, 1.0 / n], conv=conv, averaging="roe", viscous="none", rk="rk3",
             bc=[[bc, bc], [bc, bc]], gama=1.4, Minf=1.0, dt=1e-4, eps=1e-16, teno_ct=1e-6, sensor=1,
             teno_a1=10.5, teno_a2=4.5, sensor_eps=1e-30)
    p.update(kw)
    s = RefSolver(p)
    s.set_state(q)
    return p, s


def _linear_state(n, a, b, c, d):
    y, x = np.meshgrid(np.arange(n) / n, np.arange(n) / n, indexing="ij")
    rho = np.ones((n, n))
    u = a * x + b * y
    v = c * x + d * y
    p = np.full((n, n), 10.0)
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])


@pytest.mark.parametrize("coef", [(-0.3, 0.2, 0.5, -0.1), (0.004, -0.7, 0.2, 0.001), (-0.02, 0.0, 0.0, -0.01)])
def test_sensor_closed_form_on_linear_velocity(coef):
    """Central differences (and the closures) are exact on linear fields: at
    interior points theta follows the closed form in div and curl."""
    a, b, c, d = coef
    n = 16
    q = _linear_state(n, a, b, c, d)
    p, s = _block2d(n, q, bc="extrap0")
    s.rhs()
    th = s.sensor()
    div, w = a + d, c - b
    want = 0.5 * (1 - math.tanh(2.5 * (1 + 100 * div))) * div * div / (div * div + w * w + 1e-30)
    inner = th[4:-4, 4:-4]   # the extrapolation rows change the state at the faces
    assert np.abs(inner - want).max() < 1e-12 * max(1.0, abs(want))


def test_adaptive_cutoff_limits_match_fixed_cutoff():
    """A rotation (div = 0, curl = 2) gives theta = 0, hence C_T =
    10^-floor(a1) = 1e-10; a uniform compression (div = -2, no curl) gives
    theta = 1, hence C_T = 10^-floor(a1 - a2) = 1e-6.  Away from the faces the
    adaptive RHS then equals the RHS with that fixed cut-off exactly."""
    n = 24
    for coef, ct, th_want in (((0.0, -1.0, 1.0, 0.0), 1e-10, 0.0), ((-1.0, 0.0, 0.0, -1.0), 1e-6, 1.0)):
        q = _linear_state(n, *coef)
        # a steep density layer makes the choice of cut-off matter
        y, x = np.meshgrid(np.arange(n) / n, np.arange(n) / n, indexing="ij")
        q = q * (1.0 + 0.4 * np.tanh((x - 0.5) / 0.04) * np.tanh((y - 0.45) / 0.04))[None]
        _, sa = _block2d(n, q, bc="extrap0")
        ra = sa.rhs()
        th = sa.sensor()[6:-6, 6:-6]
        assert np.abs(th - th_want).max() < 1e-9
        _, sf = _block2d(n, q, bc="extrap0", teno_ct=ct, sensor=0)
        rf = sf.rhs()
        assert np.array_equal(ra[:, 8:-8, 8:-8], rf[:, 8:-8, 8:-8])
        other = 1e-6 if ct == 1e-10 else 1e-10
        _, so = _block2d(n, q, bc="extrap0", teno_ct=other, sensor=0)
        assert not np.array_equal(ra[:, 8:-8, 8:-8], so.rhs()[:, 8:-8, 8:-8])


def test_sensor_rejected_in_1d():
    from oracle import ref
    p = dict(ndim=1, n=[32], delta=[0.1], conv="teno5", dt=1e-3, bc=[["extrap0", "extrap0"]], sensor=1)
    cfg = ref.make_config(p)
    import ctypes
    h = ctypes.c_void_p()
    assert ref.lib().ref_create(ctypes.byref(cfg), ctypes.byref(h)) != 0
