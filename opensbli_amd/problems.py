"""Problem definitions and initial states of the benchmark configurations.

Each problem is a plain dict (the block's constants and scheme choices, in
the names substitute_simulation_parameters uses) and an initial state built
on the host, the counterpart of the GridBasedInitialisation kernel the
reference runs before the time loop (gridbasedinit.py:46-57).
"""
import math

import numpy as np


def taylor_green(n, conv="central_skew", rk="rk3", gama=1.4, Minf=0.1, Re=1600.0, Pr=0.71, dt=3.385e-3,
                 averaging="roe", eps=1e-16, teno_ct=1e-6, viscous="const"):
    """3-D Taylor-Green vortex on [0, 2pi)^3 (taylor_green_vortex.py:80-152).

    Returns (problem, q0) with q0 of shape (5, n, n, n) in (z, y, x) order.
    """
    d = 2.0 * math.pi / n
    problem = dict(ndim=3, n=[n, n, n], delta=[d, d, d], conv=conv, averaging=averaging, viscous=viscous, rk=rk,
                   bc=[["periodic", "periodic"]] * 3, gama=gama, Minf=Minf, Re=Re, Pr=Pr, dt=dt, eps=eps,
                   teno_ct=teno_ct)
    return problem, taylor_green_state(n, gama, Minf)


def taylor_green_state(n, gama=1.4, Minf=0.1):
    """Analytic TGV state (taylor_green_vortex.py:87-101): x_d = i_d * 2pi/n."""
    d = 2.0 * math.pi / n
    x = np.arange(n) * d
    z, y, xx = np.meshgrid(x, x, x, indexing="ij")
    u0 = np.sin(xx) * np.cos(y) * np.cos(z)
    u1 = -np.cos(xx) * np.sin(y) * np.cos(z)
    u2 = np.zeros_like(xx)
    p = 1.0 / (gama * Minf * Minf) + (1.0 / 16.0) * (np.cos(2.0 * xx) + np.cos(2.0 * y)) * (2.0 + np.cos(2.0 * z))
    r = gama * Minf * Minf * p
    rhoE = p / (gama - 1) + 0.5 * r * (u0 ** 2 + u1 ** 2 + u2 ** 2)
    return np.stack([r, r * u0, r * u1, r * u2, rhoE])


def shock_tube(n, rho_l=1.0, p_l=1.0, rho_r=0.125, p_r=0.1, gama=1.4, dt=None, conv="weno_z", rk="rk3",
               averaging="roe", cfl=0.5, eps=1e-16, teno_ct=1e-5):
    """1-D Euler shock tube on [0, 1] with transmissive (zeroth-order
    extrapolation) ends, x = i/(n-1), jump at x = 0.5 (Sod_shock_tube.py:54-60,115).

    dt defaults to cfl * dx / S, with S the largest wave speed of the exact
    Riemann solution of the initial jump.
    """
    dx = 1.0 / (n - 1)
    if dt is None:
        dt = cfl * dx / riemann_max_speed(rho_l, 0.0, p_l, rho_r, 0.0, p_r, gama)
    problem = dict(ndim=1, n=[n], delta=[dx], conv=conv, averaging=averaging, viscous="none", rk=rk,
                   bc=[["extrap0", "extrap0"]], gama=gama, dt=dt, eps=eps, teno_ct=teno_ct)
    x = np.arange(n) * dx
    left = x < 0.5
    rho = np.where(left, rho_l, rho_r)
    p = np.where(left, p_l, p_r)
    q = np.stack([rho, 0.0 * rho, p / (gama - 1.0)])
    return problem, q


def riemann_max_speed(rl, ul, pl, rr, ur, pr, gama=1.4):
    """Largest signal speed of the exact Riemann solution (Toro, ch. 4):
    Newton iteration for the star pressure, then the fastest shock/rarefaction head."""
    al, ar = math.sqrt(gama * pl / rl), math.sqrt(gama * pr / rr)

    def f(p, rk, pk, ak):
        if p > pk:
            A = 2.0 / ((gama + 1.0) * rk)
            B = (gama - 1.0) / (gama + 1.0) * pk
            s = math.sqrt(A / (p + B))
            return (p - pk) * s, s * (1.0 - 0.5 * (p - pk) / (B + p))
        pr_ = p / pk
        return (2.0 * ak / (gama - 1.0)) * (pr_ ** ((gama - 1.0) / (2.0 * gama)) - 1.0), \
            (1.0 / (rk * ak)) * pr_ ** (-(gama + 1.0) / (2.0 * gama))

    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = f(p, rl, pl, al)
        fr, dr = f(p, rr, pr, ar)
        dp = (fl + fr + (ur - ul)) / (dl + dr)
        p_new = max(1e-10, p - dp)
        if abs(p_new - p) < 1e-14 * (p + p_new):
            p = p_new
            break
        p = p_new
    fl, _ = f(p, rl, pl, al)
    fr, _ = f(p, rr, pr, ar)
    us = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    if p > pl:
        sl = ul - al * math.sqrt((gama + 1.0) / (2.0 * gama) * p / pl + (gama - 1.0) / (2.0 * gama))
    else:
        sl = ul - al
    if p > pr:
        sr = ur + ar * math.sqrt((gama + 1.0) / (2.0 * gama) * p / pr + (gama - 1.0) / (2.0 * gama))
    else:
        sr = ur + ar
    return max(abs(sl), abs(sr), abs(us))


def kelvin_helmholtz(n, amplitudes, sigma=0.05, gama=1.4, dt=None, conv="central_skew", rk="rk3"):
    """2-D periodic Kelvin-Helmholtz shear layer on [0,1)^2: rho = 2, u = 0.5
    for 0.25 <= y < 0.75, else rho = 1, u = -0.5; p = 2.5 (kelvin_helmholtz.py:65-68),
    with v' = sum_m A_m sin(2 pi m x)(exp(-(y-.25)^2/(2 s^2)) + exp(-(y-.75)^2/(2 s^2))).
    """
    d = 1.0 / n
    if dt is None:
        dt = 0.2 * d / (0.5 + math.sqrt(gama * 2.5 / 1.0))
    problem = dict(ndim=2, n=[n, n], delta=[d, d], conv=conv, averaging="roe", viscous="none", rk=rk,
                   bc=[["periodic", "periodic"]] * 2, gama=gama, dt=dt)
    return problem, kelvin_helmholtz_state(n, amplitudes, sigma, gama)


def kelvin_helmholtz_state(n, amplitudes, sigma=0.05, gama=1.4):
    d = 1.0 / n
    x = np.arange(n) * d
    y, xx = np.meshgrid(x, x, indexing="ij")
    inner = (y >= 0.25) & (y < 0.75)
    rho = np.where(inner, 2.0, 1.0)
    u = np.where(inner, 0.5, -0.5)
    env = np.exp(-(y - 0.25) ** 2 / (2 * sigma ** 2)) + np.exp(-(y - 0.75) ** 2 / (2 * sigma ** 2))
    v = np.zeros_like(xx)
    for m, a in enumerate(np.asarray(amplitudes, dtype=np.float64), start=1):
        v = v + a * np.sin(2.0 * math.pi * m * xx) * env
    p = 2.5
    return np.stack([rho, rho * u, rho * v, p / (gama - 1.0) + 0.5 * rho * (u * u + v * v)])
