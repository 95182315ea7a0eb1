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


def stretched_coordinate(n, length, factor, kind="sinh"):
    """Wall-clustered coordinate x(j), j = 0..n-1, on [0, length].

    "sinh": length*sinh(by*xi/length)/sinh(by), xi = j*length/(n-1), the map of
    the reference's flat-plate apps (katzer_SBLI.py:131, transitional_SBLI.py:168);
    "tanh": length*(1 + tanh(by*(eta - 1))/tanh(by)), eta = j/(n-1), a one-sided
    form of the channel apps' map (turbulent_channel.py:157).
    """
    eta = np.arange(n, dtype=np.float64) / (n - 1)
    if kind == "sinh":
        return length * np.sinh(factor * eta) / np.sinh(factor)
    if kind == "tanh":
        return length * (1.0 + np.tanh(factor * (eta - 1.0)) / np.tanh(factor))
    raise ValueError("unknown stretching %r" % kind)


def stable_dt(y, rho, u, T, deltas, Minf, Re, Pr=0.72, gama=1.4, SuthT=110.4, Tinf=288.0, cfl=0.8, vnn=0.3):
    """Explicit time step of a wall-normal profile (rho, u, T at the points y)
    on a grid with uniform spacings `deltas` in the other directions:
    min(cfl / max(sum_d (|u_d| + a) / dx_d), vnn / max((gama/Pr) nu sum_d 1/dx_d^2)),
    a = sqrt(T)/Minf, nu = mu(T)/(rho Re) with Sutherland's mu.  cfl = 0.8 and
    vnn = 0.3 keep RK3's stability region (imaginary-axis reach sqrt(3),
    negative-real reach 2.51) clear of the 4th-order operators' largest
    eigenvalues (advection 1.37/dx, diffusion 16/3 nu/dx^2)."""
    dy = np.diff(y)
    dyc = np.minimum(np.r_[dy[0], dy], np.r_[dy, dy[-1]])        # smaller neighbouring spacing
    a = np.sqrt(T) / Minf
    s = SuthT / Tinf
    mu = T ** 1.5 * (1.0 + s) / (T + s)
    nu = mu / (rho * Re)
    inv2 = 1.0 / dyc ** 2 + sum(1.0 / d ** 2 for d in deltas)
    lam_a = (np.abs(u) + a) / deltas[0] + a / dyc + sum(a / d for d in deltas[1:])
    lam_v = gama / Pr * nu * inv2
    return float(min(cfl / lam_a.max(), vnn / lam_v.max()))


def flat_plate_dt(n=(480, 256, 128), lengths=(400.0, 115.0, 35.0), Minf=2.0, Re=950.0, Pr=0.72, gama=1.4,
                  Tinf=288.0, SuthT=110.4, stretch_factor=5.0, stretch="tanh", wall="adiabatic_wall", Twall=None):
    """The time step flat_plate takes by default for these arguments (stable_dt
    of the similarity profile on the grid)."""
    from .boundary_layer import SimilarityProfile
    nd = len(n)
    delta = [lengths[0] / (n[0] - 1)] + ([lengths[2] / n[2]] if nd == 3 else [])
    bl = SimilarityProfile(Minf, Re, Pr=Pr, gama=gama, Tw=-1.0 if wall == "adiabatic_wall" else Twall,
                           Tinf=Tinf, suth=SuthT)
    y = stretched_coordinate(n[1], lengths[1], stretch_factor, stretch)
    rho, u, _, T = bl.evaluate(y)
    return stable_dt(y, rho, u, T, delta, Minf, Re, Pr, gama, SuthT, Tinf)


def flat_plate(n=(480, 256, 128), lengths=(400.0, 115.0, 35.0), Minf=2.0, Re=950.0, Pr=0.72, gama=1.4,
               Tinf=288.0, SuthT=110.4, stretch_factor=5.0, stretch="tanh", dt=None, conv="weno_z", rk="ls3",
               noise=1e-4, seed=5, bscheme="reduced_access", wall="adiabatic_wall", Twall=None, ndim=None,
               sensor=False, teno_a1=10.5, teno_a2=4.5, sensor_eps=1.0e-30):
    """Laminar flat-plate boundary layer, wall-normal stretched grid.

    The wall-normal map is `stretch` with `stretch_factor` (stretched_coordinate):
    by default tanh with factor 5, clustered at the wall; "sinh" is the map of
    the reference's Katzer app.  dt = None takes stable_dt of the similarity
    profile on this grid (the tanh map's first cell is 75 times thinner than
    the sinh map's, so its explicit step is set by the viscous terms there).

    x0 streamwise (inlet with pressure extrapolation, zeroth-order outlet),
    x1 wall-normal (stretched; no-slip wall below, free-stream Dirichlet
    above), x2 spanwise periodic (3-D).  The inflow is the compressible
    similarity profile at Re_delta* (boundary_layer.SimilarityProfile) and the
    interior velocity carries N(0, noise) perturbations from
    numpy.random.RandomState(seed).  Viscosity follows Sutherland's law
    (katzer_SBLI.py:26), the convective flux is the characteristic LLF
    scheme `conv` with Roe averaging (sensor=True: TENO with the Ducros sensor
    and the adaptive cut-off, katzer_SBLI.py:55-62), the viscous terms are
    Central(4) with the metric of the stretched direction
    (katzer_SBLI.py:37-70).

    Returns (problem, q0); the problem carries the wall-normal coordinates
    ("coords") and the planes of the inlet and of the Dirichlet top.
    """
    from .boundary_layer import SimilarityProfile
    nd = len(n) if ndim is None else ndim
    n = list(n[:nd])
    lengths = list(lengths[:nd])
    delta = [lengths[0] / (n[0] - 1), lengths[1] / (n[1] - 1)] + ([lengths[2] / n[2]] if nd == 3 else [])
    bl = SimilarityProfile(Minf, Re, Pr=Pr, gama=gama, Tw=-1.0 if wall == "adiabatic_wall" else Twall,
                           Tinf=Tinf, suth=SuthT)
    y = stretched_coordinate(n[1], lengths[1], stretch_factor, stretch)
    rho, u, rv, T = bl.evaluate(y)
    if dt is None:
        dt = stable_dt(y, rho, u, T, [delta[0]] + delta[2:], Minf, Re, Pr, gama, SuthT, Tinf)
    dt = float(dt)
    # profile as (nvar, n1) columns
    cons = [rho, rho * u, rv] + ([np.zeros_like(rho)] if nd == 3 else [])
    ke = 0.5 * (cons[1] ** 2 + cons[2] ** 2) / rho
    rhoE = rho * T / (gama * (gama - 1.0) * Minf ** 2) + ke
    prof = np.stack(cons + [rhoE])                                      # (nvar, n1)
    shape = tuple(reversed(n))                                          # (z,) y, x
    bshape = (1,) * (nd - 2) + (n[1], 1)
    q0 = np.stack([np.broadcast_to(prof[m].reshape(bshape), shape) for m in range(nd + 2)]).copy()
    if noise:
        rs = np.random.RandomState(seed)
        vel = [q0[1 + m] / q0[0] for m in range(nd)]
        kin0 = 0.5 * sum(v * v for v in vel) * q0[0]
        for m in range(nd):
            vel[m] = vel[m] + rs.normal(0.0, noise, shape)
        kin1 = 0.5 * sum(v * v for v in vel) * q0[0]
        for m in range(nd):
            q0[1 + m] = q0[0] * vel[m]
        q0[nd + 1] = q0[nd + 1] - kin0 + kin1     # keep the temperature profile
        # the wall rows stay at rest
        q0[1:nd + 1, ..., 0, :] = 0.0
    # inlet plane: the clean profile over (x1, x2) -> [nvar][n_e2][n_e1] with e1 = 1, e2 = 2
    n2 = n[2] if nd == 3 else 1
    inlet = np.repeat(prof[:, None, :], n2, axis=1)                    # (nvar, n2, n1)
    # top plane (normal 1): free-stream values of the profile's last row over (x0, x2)
    top = np.repeat(np.repeat(prof[:, -1][:, None, None], n2, axis=1), n[0], axis=2)  # (nvar, n2, n0)
    bc = [["inlet_pextrap", "extrap0"], [wall, "dirichlet"]] + ([["periodic", "periodic"]] if nd == 3 else [])
    bs = [[bscheme, bscheme], [bscheme, bscheme]] + ([["carpenter", "carpenter"]] if nd == 3 else [])
    problem = dict(ndim=nd, n=n, delta=delta, conv=conv, averaging="roe", viscous="sutherland", rk=rk, bc=bc,
                   bscheme=bs, stretch=[0, 1, 0][:nd] + [0] * (3 - nd), gama=gama, Minf=Minf, Re=Re, Pr=Pr, dt=dt,
                   eps=1e-16, teno_ct=1e-6, SuthT=SuthT, RefT=Tinf, Twall=bl.Twall if Twall is None else Twall,
                   coords={1: y}, planes={(0, 0): inlet, (1, 1): top}, stretch_kind=stretch,
                   stretch_factor=float(stretch_factor))
    if sensor:
        problem.update(sensor=1, teno_a1=teno_a1, teno_a2=teno_a2, sensor_eps=sensor_eps)
    return problem, q0


def wavy_grid(n, delta, amp=0.04, waves=6.0, omega=0.25, length=2.0, halo=2):
    """Coordinates of the wavy curvilinear mesh of the curvilinear Euler wave
    (apps/euler_wave_curvilinear/euler_wave.py:97-107):
        x0 = i dx + A sin(2 pi omega) sin(S pi j dy / L)
        x1 = j dy + A sin(2 pi omega) sin(S pi i dx / L)
    over the interior extended by `halo` points on every side (the initial
    condition covers the halos, gridbasedinit.py:46-57).  Returns an array of
    shape (2, n1 + 2 halo, n0 + 2 halo) (x fastest)."""
    n0, n1 = n
    dx, dy = delta
    i = np.arange(-halo, n0 + halo) * 1.0
    j = np.arange(-halo, n1 + halo) * 1.0
    jj, ii = np.meshgrid(j, i, indexing="ij")
    s = amp * math.sin(2.0 * math.pi * omega)
    x0 = ii * dx + s * np.sin(waves * math.pi * jj * dy / length)
    x1 = jj * dy + s * np.sin(waves * math.pi * ii * dx / length)
    return np.stack([x0, x1])


def euler_wave_curvilinear(n, conv="weno_z", averaging="roe", rk="ls3", gama=1.4, dt=0.0005, amp=0.04,
                           coords=None, halo=2):
    """2-D Euler density wave on a curvilinear mesh of [0, 2)^2, periodic
    (apps/euler_wave_curvilinear/euler_wave.py): rho = 1 + 0.2 sin(pi (x0 + x1))
    at the mesh points, u0 = 1, u1 = -0.5, p = 1; LLF WENO5-Z with the Roe
    average and RK3 low storage, Delta = 2/n.  `coords` overrides the mesh
    (shape (2, n + 2 halo, n + 2 halo)); the default is the wavy mesh with
    amplitude `amp`.  Returns (problem, q0), q0 of shape (4, n, n)."""
    d = 2.0 / n
    if coords is None:
        coords = wavy_grid((n, n), (d, d), amp=amp, halo=halo)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    if coords.shape != (2, n + 2 * halo, n + 2 * halo):
        raise ValueError("coords must have shape (2, n + 2 halo, n + 2 halo)")
    problem = dict(ndim=2, n=[n, n], delta=[d, d], conv=conv, averaging=averaging, viscous="none", rk=rk,
                   bc=[["periodic", "periodic"]] * 2, gama=gama, dt=dt, curvilinear=1, grid_coords=coords,
                   grid_halo=halo)
    x0 = coords[0, halo:halo + n, halo:halo + n]
    x1 = coords[1, halo:halo + n, halo:halo + n]
    rho = 1.0 + 0.2 * np.sin(np.pi * (x0 + x1))
    u0, u1, p = 1.0, -0.5, 1.0
    q = np.stack([rho, rho * u0, rho * u1, p / (gama - 1) + 0.5 * rho * (u0 * u0 + u1 * u1)])
    return problem, q


def skewed_wavy_grid3d(n, halo=2):
    """A 3-D curvilinear mesh of the cube [0, 2)^3 in index space: a skewing
    affine map plus waves across each direction (periodic metric).  Returns
    (coords of shape (3, n + 2 halo, n + 2 halo, n + 2 halo), Delta)."""
    d = 2.0 / n
    i = np.arange(-halo, n + halo) * d
    kk, jj, ii = np.meshgrid(i, i, i, indexing="ij")
    A = np.array([[1.0, 0.15, -0.1], [0.1, 0.9, 0.2], [-0.05, 0.1, 1.1]])
    xi = [ii, jj, kk]
    x = [sum(A[a, b] * xi[b] for b in range(3)) for a in range(3)]
    x[0] = x[0] + 0.03 * np.sin(np.pi * jj) * np.cos(np.pi * kk)
    x[1] = x[1] + 0.03 * np.sin(np.pi * kk + 0.3) * np.sin(np.pi * ii)
    x[2] = x[2] + 0.02 * np.cos(np.pi * ii) * np.sin(np.pi * jj)
    return np.stack(x), d


def euler_wave_curvilinear_3d(n, conv="weno_z", averaging="roe", rk="ls3", gama=1.4, dt=2e-3, halo=2):
    """3-D Euler density wave on the skewed wavy mesh (periodic): rho = 1 +
    0.2 sin(pi (x0 + x1 + x2)), u = (0.6, -0.3, 0.25), p = 1 + 0.1 cos(pi (x0 - x2)).
    Returns (problem, q0), q0 of shape (5, n, n, n)."""
    coords, d = skewed_wavy_grid3d(n, halo)
    problem = dict(ndim=3, n=[n] * 3, delta=[d] * 3, conv=conv, averaging=averaging, viscous="none", rk=rk,
                   bc=[["periodic", "periodic"]] * 3, gama=gama, dt=dt, curvilinear=1, grid_coords=coords,
                   grid_halo=halo)
    c = coords[:, halo:halo + n, halo:halo + n, halo:halo + n]
    rho = 1.0 + 0.2 * np.sin(np.pi * (c[0] + c[1] + c[2]))
    u = [0.6, -0.3, 0.25]
    p = 1.0 + 0.1 * np.cos(np.pi * (c[0] - c[2]))
    q = np.stack([rho] + [rho * ui for ui in u] + [p / (gama - 1) + 0.5 * rho * sum(ui * ui for ui in u)])
    return problem, q


def channel(n=(24, 33, 12), conv="teno5", c=(-1.0, 0.0, 0.0), Re=190.71, Minf=0.0955, Pr=0.7, dt=2e-4, by=1.7,
            seed=3):
    """Compressible channel driven by a constant body force c_j: conservative
    Navier-Stokes with a characteristic scheme, Sutherland viscosity,
    periodic x0 (and x2), isothermal walls at x1 = -1 and 1 on a
    wall-clustered grid (tanh map from both walls), T_wall = 1.  The body
    force terms are those of the reference's channel apps
    (apps/channel_flow/compressible_TCF_TENO/turbulent_channel.py:15-16);
    those apps themselves use the power law mu = T**0.7 (TCF_TENO, TCF_Central)
    or the Feiereisen split form (turbulent_3D), which this build does not
    recognise, so none of them runs here: this block exercises the body-force
    terms with Sutherland viscosity instead.  Initial state: a laminar-like
    profile u0 = 1.5 (1 - x1^2) plus seeded noise of amplitude 0.02, rho = 1,
    T = 1.  Returns (problem, q0)."""
    nd = len(n)
    d = [4.0 * math.pi / n[0], 2.0 / (n[1] - 1)] + ([4.0 * math.pi / 3.0 / n[2]] if nd == 3 else [])
    eta = np.arange(n[1]) / (n[1] - 1.0)
    y = -np.tanh(by * (1.0 - 2.0 * eta)) / np.tanh(by)
    bc = [["periodic", "periodic"], ["isothermal_wall", "isothermal_wall"]] + ([["periodic", "periodic"]] if nd == 3
                                                                              else [])
    problem = dict(ndim=nd, n=list(n), delta=d, conv=conv, averaging="roe", viscous="sutherland", rk="ls3", bc=bc,
                   bscheme=[["carpenter", "carpenter"]] * nd, gama=1.4, Minf=Minf, Re=Re, Pr=Pr, dt=dt, eps=1e-15,
                   teno_ct=1e-7, SuthT=110.4, RefT=273.0, Twall=1.0, stretch=[0, 1, 0][:nd] + [0] * (3 - nd),
                   coords={1: y}, body_force=list(c[:nd]))
    shape = tuple(reversed(n))
    rs = np.random.RandomState(seed)
    Y = y.reshape((1,) * (nd - 2) + (n[1], 1)) * np.ones(shape)
    u0 = 1.5 * (1.0 - Y ** 2) + 0.02 * rs.standard_normal(shape) * (1.0 - Y ** 2)
    u1 = 0.02 * rs.standard_normal(shape) * (1.0 - Y ** 2)
    vel = [u0, u1] + ([0.02 * rs.standard_normal(shape) * (1.0 - Y ** 2)] if nd == 3 else [])
    rho = np.ones(shape)
    p = 1.0 / (1.4 * Minf ** 2)
    q = np.stack([rho] + [rho * v for v in vel] + [p / 0.4 + 0.5 * rho * sum(v * v for v in vel)])
    return problem, q
