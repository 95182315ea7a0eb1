"""Plain PyTorch FP64 statement of the central skew-symmetric scheme, the
LLF characteristic WENO5-Z residual and the constant-viscosity terms with
RungeKutta(3) on periodic grids (test infrastructure): the autograd
reference for the hand-written adjoint.

The same discretisation as the reference's TGV-style Skew form
(parsing.py:75-111, scheme.py:187-252): 4th-order central first derivatives
with periodic wrap, 3-register RK3 (rk_sbli.py:58-60,117-133); the WENO5-Z
and viscous parts are described where they start.
"""
import torch


def d1(f, axis, inv):
    """4th-order central first derivative along a tensor axis (periodic)."""
    return inv * ((1.0 / 12.0) * (torch.roll(f, 2, axis) - torch.roll(f, -2, axis))
                  + (2.0 / 3.0) * (torch.roll(f, -1, axis) - torch.roll(f, 1, axis)))


def rhs(q, delta, gama):
    nd = q.dim() - 1
    nv = nd + 2
    rho = q[0]
    u = [q[1 + d] / rho for d in range(nd)]
    p = (gama - 1.0) * (q[nv - 1] - 0.5 * rho * sum(ui * ui for ui in u))
    r = [torch.zeros_like(rho) for _ in range(nv)]
    for j in range(nd):
        axis = nd - 1 - j  # tensor axis of coordinate direction j (x fastest)
        inv = 1.0 / delta[j]
        uj = u[j]
        duj = d1(uj, axis, inv)
        for m in range(nv):
            r[m] = r[m] - 0.5 * (d1(q[m] * uj, axis, inv) + uj * d1(q[m], axis, inv) + q[m] * duj)
        r[1 + j] = r[1 + j] - d1(p, axis, inv)
        r[nv - 1] = r[nv - 1] - d1(p * uj, axis, inv)
    return torch.stack(r)


def rk3_step(q, dt, delta, gama):
    rknew = [2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0]
    rkold = [1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0]
    qold = q
    cur = q
    for s in range(3):
        r = rhs(cur, delta, gama)
        cur = qold + dt * rknew[s] * r
        qold = qold + dt * rkold[s] * r
    return cur


def integrate(q, nsteps, dt, delta, gama):
    for _ in range(nsteps):
        q = rk3_step(q, dt, delta, gama)
    return q


# ---------------------------------------------------------------------------
# LLF characteristic WENO5-Z (Roe average) on a periodic 2-D grid: the
# reference's shock-capturing residual (shock_capturing.py:357-448,
# weno.py:283-338, averaging.py:70-114, euler_eigensystem.py:57-135) written
# with whole-array torch operations, for autograd.
# ---------------------------------------------------------------------------
def _shift(f, k, axis):
    """f at index i+k along the tensor axis (periodic)."""
    return torch.roll(f, -k, axis)


def _weno_z(fm2, fm1, f0, f1, f2):
    """WENO-Z (p = 2, eps = 1e-14) reconstruction at i+1/2 from f[i-2..i+2]."""
    b0 = (13.0 / 12.0) * (f0 - 2 * f1 + f2) ** 2 + 0.25 * (3 * f0 - 4 * f1 + f2) ** 2
    b1 = (13.0 / 12.0) * (fm1 - 2 * f0 + f1) ** 2 + 0.25 * (fm1 - f1) ** 2
    b2 = (13.0 / 12.0) * (fm2 - 2 * fm1 + f0) ** 2 + 0.25 * (fm2 - 4 * fm1 + 3 * f0) ** 2
    tau = torch.abs(b0 - b2)
    a0 = 0.3 * (1 + (tau / (1e-14 + b0)) ** 2)
    a1 = 0.6 * (1 + (tau / (1e-14 + b1)) ** 2)
    a2 = 0.1 * (1 + (tau / (1e-14 + b2)) ** 2)
    s = a0 + a1 + a2
    r0 = (2 * f0 + 5 * f1 - f2) / 6.0
    r1 = (-fm1 + 5 * f0 + 2 * f1) / 6.0
    r2 = (2 * fm2 - 7 * fm1 + 11 * f0) / 6.0
    return (a0 * r0 + a1 * r1 + a2 * r2) / s


def rhs_weno_z(q, delta, gama):
    """Residual of the 2-D Euler equations with the LLF characteristic WENO5-Z
    flux and Roe averaging (q: (4, ny, nx), x along the last axis)."""
    gm1 = gama - 1.0
    rho = q[0]
    u = [q[1] / rho, q[2] / rho]
    p = gm1 * (q[3] - 0.5 * rho * (u[0] ** 2 + u[1] ** 2))
    a = torch.sqrt(gama * p / rho)
    res = torch.zeros_like(q)
    for d in range(2):
        ax = 1 - d
        kt = [1.0 if d == 0 else 0.0, 1.0 if d == 1 else 0.0]
        sh = lambda f, k: _shift(f, k, ax)  # noqa: E731
        # Roe average between i and i+1
        sL, sR = torch.sqrt(rho), torch.sqrt(sh(rho, 1))
        inv = 1.0 / (sL + sR)
        ra = torch.sqrt(rho * sh(rho, 1))
        ua = [(sL * u[m] + sR * sh(u[m], 1)) * inv for m in range(2)]
        H = (q[3] + p) / rho
        Ha = (sL * H + sR * sh(H, 1)) * inv
        aa = torch.sqrt(gm1 * (Ha - 0.5 * (ua[0] ** 2 + ua[1] ** 2)))
        ia, ir = 1.0 / aa, 1.0 / ra
        alpha = ra / (aa * 2.0 ** 0.5)
        bta = ir * ia / 2.0 ** 0.5
        th = kt[0] * ua[0] + kt[1] * ua[1]
        phi = gm1 * (ua[0] ** 2 + ua[1] ** 2) / 2.0
        u0, u1 = ua
        R = [[1.0 + 0 * ra, 0 * ra, alpha, alpha],
             [u0, kt[1] * ra, alpha * (u0 + kt[0] * aa), alpha * (u0 - kt[0] * aa)],
             [u1, -kt[0] * ra, alpha * (u1 + kt[1] * aa), alpha * (u1 - kt[1] * aa)],
             [phi / gm1, ra * (kt[1] * u0 - kt[0] * u1), alpha * ((phi + aa * aa) / gm1 + aa * th),
              alpha * ((phi + aa * aa) / gm1 - aa * th)]]
        L = [[1 - phi * ia * ia, gm1 * u0 * ia * ia, gm1 * u1 * ia * ia, -gm1 * ia * ia],
             [-(kt[1] * u0 - kt[0] * u1) * ir, kt[1] * ir + 0 * ra, -kt[0] * ir + 0 * ra, 0 * ra],
             [bta * (phi - aa * th), bta * (kt[0] * aa - gm1 * u0), bta * (kt[1] * aa - gm1 * u1), bta * gm1],
             [bta * (phi + aa * th), -bta * (kt[0] * aa + gm1 * u0), -bta * (kt[1] * aa + gm1 * u1), bta * gm1]]
        ud = u[d]
        F = [q[1 + d], q[1] * ud + kt[0] * p, q[2] * ud + kt[1] * p, (q[3] + p) * ud]
        lam_pt = [torch.abs(ud), torch.abs(ud), torch.abs(ud + a), torch.abs(ud - a)]
        flux = [0.0] * 4
        for k in range(4):
            # lambda_max over the six points i-2..i+3, first maximum kept
            lam = sh(lam_pt[k], -2)
            for s in range(-1, 4):
                nxt = sh(lam_pt[k], s)
                lam = torch.where(nxt > lam, nxt, lam)
            fp, fm = [], []
            for s in range(-2, 4):
                cf = sum(L[k][m] * sh(F[m], s) for m in range(4))
                cs = sum(L[k][m] * sh(q[m], s) for m in range(4))
                fp.append(0.5 * (cf + lam * cs))
                fm.append(0.5 * (cf - lam * cs))
            rec = _weno_z(*fp[:5]) + _weno_z(fm[5], fm[4], fm[3], fm[2], fm[1])
            for m in range(4):
                flux[m] = flux[m] + R[m][k] * rec
        for m in range(4):
            res[m] = res[m] - (flux[m] - sh(flux[m], -1)) / delta[d]
    return res


def rk3_step_weno(q, dt, delta, gama):
    rknew = [2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0]
    rkold = [1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0]
    qold, cur = q, q
    for s in range(3):
        r = rhs_weno_z(cur, delta, gama)
        cur = qold + dt * rknew[s] * r
        qold = qold + dt * rkold[s] * r
    return cur


def integrate_weno(q, nsteps, dt, delta, gama):
    for _ in range(nsteps):
        q = rk3_step_weno(q, dt, delta, gama)
    return q


# ---------------------------------------------------------------------------
# Viscous terms with constant viscosity (taylor_green_vortex.py:14-19): the
# stress tensor and heat flux differentiated by the product rule as
# Der.differentiate does (parsing.py:141-186), second derivatives with the
# 4th-order D2 weights, mixed derivatives as D_b(D_a u) for a < b
# (scheme.py:81-85,360-387); any dimension, periodic.
# ---------------------------------------------------------------------------
def d2(f, axis, inv):
    """4th-order central second derivative along a tensor axis (periodic)."""
    return inv * inv * ((-1.0 / 12.0) * (torch.roll(f, 2, axis) + torch.roll(f, -2, axis))
                        + (4.0 / 3.0) * (torch.roll(f, 1, axis) + torch.roll(f, -1, axis)) - 2.5 * f)


def viscous(q, delta, gama, Minf, Re, Pr):
    nd = q.dim() - 1
    nv = nd + 2
    rho = q[0]
    u = [q[1 + d] / rho for d in range(nd)]
    p = (gama - 1.0) * (q[nv - 1] - 0.5 * rho * sum(ui * ui for ui in u))
    T = p * gama * Minf * Minf / rho
    inv_re = 1.0 / Re
    cq = 1.0 / ((gama - 1.0) * Minf * Minf * Pr * Re)
    ax = [nd - 1 - j for j in range(nd)]
    inv = [1.0 / delta[j] for j in range(nd)]
    du = [[d1(u[m], ax[j], inv[j]) for j in range(nd)] for m in range(nd)]
    dd = [[[None] * nd for _ in range(nd)] for _ in range(nd)]
    for m in range(nd):
        for a in range(nd):
            dd[m][a][a] = d2(u[m], ax[a], inv[a])
            for b in range(a + 1, nd):
                v = d1(d1(u[m], ax[a], inv[a]), ax[b], inv[b])
                dd[m][a][b] = v
                dd[m][b][a] = v
    divu = sum(du[k][k] for k in range(nd))
    energy = sum(cq * d2(T, ax[j], inv[j]) for j in range(nd))
    out = [torch.zeros_like(rho)]
    for m in range(nd):
        sk = sum(dd[k][m][k] for k in range(nd))
        visc = torch.zeros_like(rho)
        for j in range(nd):
            tt = dd[m][j][j] + dd[j][m][j] - ((2.0 / 3.0) * sk if j == m else 0.0)
            dtau = inv_re * tt
            tau = inv_re * (du[m][j] + du[j][m] - ((2.0 / 3.0) * divu if j == m else 0.0))
            visc = visc + dtau
            energy = energy + du[m][j] * tau + u[m] * dtau
        out.append(visc)
    out.append(energy)
    return torch.stack(out)


def rhs_ns(q, delta, gama, Minf, Re, Pr, conv="central_skew"):
    """Central skew (any dimension) or WENO5-Z (2-D) convective residual plus the viscous terms."""
    r = rhs(q, delta, gama) if conv == "central_skew" else rhs_weno_z(q, delta, gama)
    return r + viscous(q, delta, gama, Minf, Re, Pr)


def integrate_ns(q, nsteps, dt, delta, gama, Minf, Re, Pr, conv="central_skew"):
    rknew = [2.0 / 3.0, 5.0 / 12.0, 3.0 / 5.0]
    rkold = [1.0 / 4.0, 3.0 / 20.0, 3.0 / 5.0]
    for _ in range(nsteps):
        qold, cur = q, q
        for s in range(3):
            r = rhs_ns(cur, delta, gama, Minf, Re, Pr, conv)
            cur = qold + dt * rknew[s] * r
            qold = qold + dt * rkold[s] * r
        q = cur
    return q
