// osbl_general.hip -- the RHS of blocks with non-periodic faces, stretched
// directions or a temperature-dependent viscosity:
//   * central first/second derivatives with the one-sided closures of the
//     faces' boundary schemes (Carpenter_scheme.py:6-102,
//     reduced_access_scheme.py:6-83, applied as in opensblifunctions.py:523-534),
//   * the metric of directions stretched along themselves (metric.py:72-212),
//   * viscous terms with mu(T) by Sutherland's law as a constituent relation
//     (katzer_SBLI.py:12-26), expanded by the product rule of
//     Der.differentiate (parsing.py:141-186),
//   * the central convective forms with the same closures,
// fused with the RK update.  The characteristic fluxes come from the sweeps of
// osbl_kernels.hip, which apply the metric to their flux differences.
#include "osbl_kernels.h"

namespace osbl {

// Derivative in the computational coordinate (without 1/Delta) at the point
// x, grid index c of N along direction d (stride st).  ORDER 1 or 2.  f(y)
// evaluates the differentiated expression at storage offset y.  A closure row
// replaces the stencil only at its own grid index (side-0 rows tested first,
// as in the reference's Piecewise); halo points keep the central stencil.
template <int ORDER, class F>
__device__ __forceinline__ double dxi(const Bnd& b, int d, int c, int N, int x, int st, F&& f) {
  const int lo = b.mod[d][0], hi = b.mod[d][1];
  if constexpr (ORDER == 1) {
    if (lo) {
      const int nr = lo == 1 ? 4 : 2, np = lo == 1 ? 6 : 5;
      if (c >= 0 && c < nr) {
        const double* w = b.cw + (lo == 1 ? CW_CARP : CW_RA) + 6 * c;
        double acc = 0.0;
        for (int k = 0; k < np; ++k) acc += w[k] * f(x + (k - c) * st);
        return acc;
      }
    }
    if (hi) {
      const int nr = hi == 1 ? 4 : 2, np = hi == 1 ? 6 : 5;
      const int r = N - 1 - c;
      if (r >= 0 && r < nr) {
        const double* w = b.cw + (hi == 1 ? CW_CARP : CW_RA) + 6 * r;
        double acc = 0.0;
        for (int k = 0; k < np; ++k) acc += w[k] * f(x + (r - k) * st);
        return -acc;
      }
    }
    return (1.0 / 12.0) * (f(x - 2 * st) - f(x + 2 * st)) + (2.0 / 3.0) * (f(x + st) - f(x - st));
  } else {
    if (lo && c >= 0 && c < 2) {
      const double* w = b.cw + CW_D2 + 6 * c;
      double acc = 0.0;
      for (int k = 0; k < 5; ++k) acc += w[k] * f(x + (k - c) * st);
      return acc;
    }
    if (hi && N - 1 - c >= 0 && N - 1 - c < 2) {
      const int r = N - 1 - c;
      const double* w = b.cw + CW_D2 + 6 * r;
      double acc = 0.0;
      for (int k = 0; k < 5; ++k) acc += w[k] * f(x + (r - k) * st);
      return acc;
    }
    return (-1.0 / 12.0) * (f(x - 2 * st) + f(x + 2 * st)) + (4.0 / 3.0) * (f(x - st) + f(x + st)) - 2.5 * f(x);
  }
}

// Viscous terms (stress tensor and heat flux of the reference's NS scripts):
//   D_j tau_mj = (mu/Re)(D_jj u_m + D_j D_m u_j - 2/3 d_mj D_m D_k u_k)
//              + (1/Re) D_j mu (D_j u_m + D_m u_j - 2/3 d_mj D_k u_k)
//   -D_j q_j   = cq (mu D_jj T + D_j mu D_j T)
//   D_j(u_m tau_mj) = D_j u_m tau_mj + u_m D_j tau_mj
// mu = 1 and D mu = 0 for constant viscosity.  Derivatives are taken in the
// computational coordinates and transformed on stretched directions:
// D_j -> G_j D, D_jj -> G_j^2 D2 + G_j S_j D, D_a D_b -> G_a G_b D_b(D_a).
// P holds u_0..u_{ND-1}, p, T, mu.
template <int ND>
__device__ __forceinline__ void viscous_assemble(double mu, const double um_[3], const double (&DU)[ND][ND],
                                                 const double (&DD)[ND][ND][ND], const double (&D2T)[ND],
                                                 const double (&DT)[ND], const double (&DMU)[ND], const Phys& ph,
                                                 bool suth, double vr[ND + 2]);

template <int ND>
__device__ __forceinline__ void viscous_gen(const double* __restrict__ P, long long fs, int x, const int st[3],
                                            const int ci[3], const Grid& g, const Bnd& b, const Phys& ph, bool suth,
                                            double vr[ND + 2]) {
  const double* U[3] = {P, P + fs, P + 2 * fs};
  const double* T = P + (ND + 1) * fs;
  const double* MU = P + (ND + 2) * fs;
  bool metric = false;
  double G[ND], S[ND];
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    metric |= b.dm[d] != nullptr;
    G[d] = b.dm[d] ? b.dm[d][ci[d]] : 1.0;
    S[d] = b.sd[d] ? b.sd[d][ci[d]] : 0.0;
  }
  double du[ND][ND];       // computational first derivatives (with 1/Delta)
  double DU[ND][ND];       // transformed
  double DD[ND][ND][ND];   // transformed second / mixed derivatives, DD[m][a][b] = D_a D_b u_m
#pragma unroll
  for (int m = 0; m < ND; ++m)
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      const double* f = U[m];
      auto fv = [&](int y) { return f[y]; };
      du[m][j] = ph.inv_delta[j] * dxi<1>(b, j, ci[j], g.n[j], x, st[j], fv);
      const double d2 = ph.inv_delta[j] * ph.inv_delta[j] * dxi<2>(b, j, ci[j], g.n[j], x, st[j], fv);
      DU[m][j] = metric ? G[j] * du[m][j] : du[m][j];
      DD[m][j][j] = metric ? G[j] * G[j] * d2 + G[j] * S[j] * du[m][j] : d2;
    }
  // mixed D_b(D_a u_m), a < b, for m in {a, b}
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int bb = a + 1; bb < ND; ++bb)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int m = q == 0 ? a : bb;
        const double* f = U[m];
        auto inner = [&](int y) {
          return ph.inv_delta[a] * dxi<1>(b, a, ci[a], g.n[a], y, st[a], [&](int z) { return f[z]; });
        };
        double v = ph.inv_delta[bb] * dxi<1>(b, bb, ci[bb], g.n[bb], x, st[bb], inner);
        if (metric) v = G[a] * G[bb] * v;
        DD[m][a][bb] = v;
        DD[m][bb][a] = v;
      }
  double D2T[ND], DT[ND], DMU[ND];
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    auto tv = [&](int y) { return T[y]; };
    const double d2 = ph.inv_delta[j] * ph.inv_delta[j] * dxi<2>(b, j, ci[j], g.n[j], x, st[j], tv);
    DT[j] = 0.0;
    DMU[j] = 0.0;
    if (suth || metric) {
      const double d1 = ph.inv_delta[j] * dxi<1>(b, j, ci[j], g.n[j], x, st[j], tv);
      DT[j] = metric ? G[j] * d1 : d1;
      D2T[j] = metric ? G[j] * G[j] * d2 + G[j] * S[j] * d1 : d2;
    } else {
      D2T[j] = d2;
    }
    if (suth) {
      const double dm1 = ph.inv_delta[j] * dxi<1>(b, j, ci[j], g.n[j], x, st[j], [&](int y) { return MU[y]; });
      DMU[j] = metric ? G[j] * dm1 : dm1;
    }
  }
  const double um[3] = {U[0][x], ND > 1 ? U[1][x] : 0.0, ND > 2 ? U[2][x] : 0.0};
  viscous_assemble<ND>(suth ? MU[x] : 1.0, um, DU, DD, D2T, DT, DMU, ph, suth, vr);
}

// The viscous residual from the transformed derivatives at a point (the
// second half of viscous_gen, shared with the tiled kernel).
template <int ND>
__device__ __forceinline__ void viscous_assemble(double mu, const double um_[3], const double (&DU)[ND][ND],
                                                 const double (&DD)[ND][ND][ND], const double (&D2T)[ND],
                                                 const double (&DT)[ND], const double (&DMU)[ND], const Phys& ph,
                                                 bool suth, double vr[ND + 2]) {
  double divu = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) divu += DU[k][k];
  double energy = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    if (suth) energy += ph.cq * (mu * D2T[j] + DMU[j] * DT[j]);
    else energy += ph.cq * D2T[j];
  }
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double um = um_[m];
    double visc = 0.0;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      double tt = DD[m][j][j] + DD[j][m][j];
      if (j == m) {
        double sk = 0.0;
#pragma unroll
        for (int k = 0; k < ND; ++k) sk += DD[k][m][k];
        tt -= (2.0 / 3.0) * sk;
      }
      const double sr = DU[m][j] + DU[j][m] - (j == m ? (2.0 / 3.0) * divu : 0.0);
      double dtau, tau;
      if (suth) {
        dtau = ph.inv_re * mu * tt + ph.inv_re * DMU[j] * sr;
        tau = ph.inv_re * mu * sr;
      } else {
        dtau = ph.inv_re * tt;
        tau = ph.inv_re * sr;
      }
      visc += dtau;
      energy += DU[m][j] * tau + um * dtau;
    }
    vr[1 + m] += visc;
  }
  vr[ND + 1] += energy;
}

// Central convective terms with the closures and metric: the skew form
// (parsing.py:75-111, scheme.py:187-252) or the conservative flux form
// (parsing.py:189-216).  P holds u_0..u_{ND-1}, p.
template <int ND, int CONV>
__device__ __forceinline__ void central_gen(const double* __restrict__ Q, const double* __restrict__ P, long long fs,
                                            int x, const int st[3], const int ci[3], const Grid& g, const Bnd& b,
                                            const Phys& ph, double r[ND + 2]) {
  constexpr int NV = ND + 2;
  const double* U[3] = {P, P + fs, P + 2 * fs};
  const double* Pp = P + ND * fs;
#pragma unroll
  for (int m = 0; m < NV; ++m) r[m] = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    const double inv = ph.inv_delta[j];
    const bool met = b.dm[j] != nullptr;
    const double G = met ? b.dm[j][ci[j]] : 1.0;
    auto D = [&](auto&& f) {
      const double v = inv * dxi<1>(b, j, ci[j], g.n[j], x, st[j], f);
      return met ? G * v : v;
    };
    const double* uj = U[j];
    if constexpr (CONV == OSBL_CONV_CENTRAL_SKEW) {
      const double duj = D([&](int y) { return uj[y]; });
      const double dp = D([&](int y) { return Pp[y]; });
      const double dpu = D([&](int y) { return Pp[y] * uj[y]; });
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double* qm = Q + m * fs;
        const double dru = D([&](int y) { return qm[y] * uj[y]; });
        const double dq = D([&](int y) { return qm[y]; });
        double t = -0.5 * (dru + uj[x] * dq + qm[x] * duj);
        if (m == 1 + j) t -= dp;
        if (m == NV - 1) t -= dpu;
        r[m] += t;
      }
    } else {
      r[0] -= D([&](int y) { return Q[(1 + j) * fs + y]; });
#pragma unroll
      for (int m = 0; m < ND; ++m)
        r[1 + m] -= D([&](int y) { return Q[(1 + m) * fs + y] * uj[y] + (m == j ? Pp[y] : 0.0); });
      r[NV - 1] -= D([&](int y) { return (Pp[y] + Q[(NV - 1) * fs + y]) * uj[y]; });
    }
  }
}

// Residual + RK update of one interior cell (modes as k_update: 0 3-register
// RK3, 1 low-storage, 2 residual only into R).
// the constant body force: - KD(_i,_j)*c_j and - Dot(c_j, u_j)
// (apps/channel_flow/compressible_TCF_TENO/turbulent_channel.py:15-16)
template <int ND>
__device__ __forceinline__ void add_body_force(const double* __restrict__ Q, long long fs, int x, const Phys& ph,
                                               double* r) {
  double cu = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    r[1 + m] -= ph.force[m];
    cu += ph.force[m] * (Q[(1 + m) * fs + x] / Q[x]);
  }
  r[ND + 1] -= cu;
}

// R += the body force over the interior (the tiled general schedule)
__global__ void __launch_bounds__(256) k_body_force3d(const double* __restrict__ Q, double* __restrict__ R, Grid g,
                                                      Phys ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int x = g.idx(i, blockIdx.y, blockIdx.z);
  double r[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) r[m] = R[m * g.fs + x];
  add_body_force<3>(Q, g.fs, x, ph, r);
#pragma unroll
  for (int m = 0; m < 5; ++m) R[m * g.fs + x] = r[m];
}

template <int ND, int CONV, bool VISC>
__global__ void __launch_bounds__(128) k_update_gen(const double* __restrict__ Q, const double* __restrict__ P,
                                                    double* __restrict__ R, double* __restrict__ Qo,
                                                    double* __restrict__ Qn, Grid g, Phys ph, Bnd b, double cA,
                                                    double cB, int first, int mode, int suth) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j = blockIdx.y, k = blockIdx.z;
  const int x = g.idx(i, j, k);
  const int st[3] = {1, g.pitch, g.plane};
  const int ci[3] = {i, j, k};
  double r[NV];
  if constexpr (CONV == OSBL_CONV_CENTRAL_SKEW || CONV == OSBL_CONV_CENTRAL_CONS) {
    central_gen<ND, CONV>(Q, P, g.fs, x, st, ci, g, b, ph, r);
  } else {
#pragma unroll
    for (int m = 0; m < NV; ++m) r[m] = R[m * g.fs + x];
  }
  if constexpr (VISC) viscous_gen<ND>(P, g.fs, x, st, ci, g, b, ph, suth != 0, r);
  if (ph.has_force) add_body_force<ND>(Q, g.fs, x, ph, r);
  if (mode == 2) {
#pragma unroll
    for (int m = 0; m < NV; ++m) R[m * g.fs + x] = r[m];
  } else if (mode == 0) {
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const double qo = first ? Q[m * g.fs + x] : Qo[m * g.fs + x];
      Qn[m * g.fs + x] = qo + ph.dt * cA * r[m];
      Qo[m * g.fs + x] = qo + ph.dt * cB * r[m];
    }
  } else {
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const double tmp = (cA == 0.0 ? 0.0 : cA * Qo[m * g.fs + x]) + ph.dt * r[m];
      Qo[m * g.fs + x] = tmp;
      Qn[m * g.fs + x] = Q[m * g.fs + x] + cB * tmp;
    }
  }
}

// ---------------------------------------------------------------------------
// 3-D viscous terms of the general path, tiled 2.5-D like k_visc_update3d: a
// 64 x 4 block of (x, y) columns marches along z; u_0..u_2, T (and mu) of each
// z plane are staged from P into one of three rotating LDS slots (2-point
// x/y halo); the z stencils read a per-thread queue of the planes z-2..z+2
// holding u, T, mu and the inner derivatives D_x u_0, D_x u_2, D_y u_1,
// D_y u_2 of the mixed terms.  The viscous terms are added to R (which holds
// the x sweep's flux differences).  Points inside the closure rows of
// an x or y face are left to k_visc_shell; z must have neither closures nor a
// metric (checked by the caller).  The derivative and assembly formulas are
// those of viscous_gen.
// ---------------------------------------------------------------------------
// u_0..u_{ND-1}, p, T (and mu(T) by Sutherland's law) of the point at x --
// the constituent relations k_prims / k_prims_mu store, formed in place
template <int ND, bool MU>
__device__ __forceinline__ void prims_at(const double* __restrict__ Q, long long fs, int x, const Phys& ph,
                                         double (&u)[ND], double& p, double& T, double& mu) {
  const double rho = Q[x];
  const double ir = recip_pos(rho);
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    u[m] = Q[(1 + m) * fs + x] * ir;
    ke += u[m] * u[m];
  }
  p = ph.gm1 * (Q[(ND + 1) * fs + x] - 0.5 * rho * ke);
  T = p * ph.gama * ph.Minf * ph.Minf * ir;
  // T^(3/2) as T sqrt(T) (instead of pow); the reciprocals, quotient and
  // root of these positive operands by the refined hardware estimates
  if constexpr (MU) mu = T * sqrt_pos(T) * quot_pos(1.0 + ph.suth, T + ph.suth);
}

// the same constituent relations from a point's conservative variables in registers
template <int ND, bool MU>
__device__ __forceinline__ void prims_of(const double (&q)[ND + 2], const Phys& ph, double (&u)[ND], double& p,
                                         double& T, double& mu) {
  const double rho = q[0];
  const double ir = recip_pos(rho);
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    u[m] = q[1 + m] * ir;
    ke += u[m] * u[m];
  }
  p = ph.gm1 * (q[ND + 1] - 0.5 * rho * ke);
  T = p * ph.gama * ph.Minf * ph.Minf * ir;
  if constexpr (MU) mu = T * sqrt_pos(T) * quot_pos(1.0 + ph.suth, T + ph.suth);
}

constexpr int GTX = 64, GTY = 8, GPX = GTX + 4, GPY = GTY + 4, GPL = GPX * GPY;

// closure rows of the first (Carpenter 4, ReducedAccess 2) and second (2) derivatives
__host__ __device__ __forceinline__ int closure_rows(int mod) { return mod == 0 ? 0 : (mod == 1 ? 4 : 2); }

// MS: the stretched directions (bit 0: x, bit 1: y); the metric factors of
// the others are 1 and 0, and their products are left out.
template <bool SUTH, int MS>
__global__ void __launch_bounds__(GTX * GTY, 1) k_visc_gen3d(const double* __restrict__ Q, double* __restrict__ R,
                                                            Grid g, Phys ph, Bnd b, int sz) {
  constexpr int NF = SUTH ? 5 : 4;  // u0, u1, u2, T (, mu)
  constexpr int NQ = NF + 4;        // + D_x u0, D_x u2, D_y u1, D_y u2
  constexpr int GNT = GTX * GTY, GPP = (GPL + GNT - 1) / GNT;  // staged points per thread and plane
  __shared__ double sp[4][NF][GPL];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * GTX + tx;
  const int x0 = blockIdx.x * GTX, y0 = blockIdx.y * GTY;
  const int z0 = blockIdx.z * sz;
  const int z1 = min(z0 + sz, g.n[2]);
  const int x = x0 + tx, y = y0 + ty;
  const bool inside = x < g.n[0] && y < g.n[1];
  const double ix = ph.inv_delta[0], iy = ph.inv_delta[1], iz = ph.inv_delta[2];
  const bool slow = x < closure_rows(b.mod[0][0]) || g.n[0] - 1 - x < closure_rows(b.mod[0][1]) ||
                    y < closure_rows(b.mod[1][0]) || g.n[1] - 1 - y < closure_rows(b.mod[1][1]);
  constexpr bool metric = MS != 0;
  auto str = [](int j) { return ((MS >> j) & 1) != 0; };
  const double G0 = ((MS & 1) && b.dm[0] && inside) ? b.dm[0][x] : 1.0;
  const double S0 = ((MS & 1) && b.sd[0] && inside) ? b.sd[0][x] : 0.0;
  const double G1 = ((MS & 2) && b.dm[1] && inside) ? b.dm[1][y] : 1.0;
  const double S1 = ((MS & 2) && b.sd[1] && inside) ? b.sd[1][y] : 0.0;
  // G_a G_b v, the factors of unstretched directions (1) left out
  auto gg = [&](int a, int bb, double v, const double (&G)[3]) {
    if (str(a) && str(bb)) return G[a] * G[bb] * v;
    if (str(a)) return G[a] * v;
    if (str(bb)) return G[bb] * v;
    return v;
  };

  // this thread's staged points (plane offset at z = 0) and the raw Q of the
  // plane in flight; the constituent relations of every staged point are
  // formed from Q when the plane goes into its LDS slot
  int soff[GPP];
  bool sval[GPP];
#pragma unroll
  for (int k = 0; k < GPP; ++k) {
    const int pt = tid + k * GNT;
    const int gx = x0 - 2 + pt % GPX, gy = y0 - 2 + pt / GPX;
    sval[k] = pt < GPL && gx < g.n[0] + g.h[0] && gy < g.n[1] + g.h[1];
    soff[k] = sval[k] ? g.idx(gx, gy, 0) : 0;
  }
  double pq[GPP][5];
  auto fetch = [&](int zp) {
    const int zo = zp * g.plane;
#pragma unroll
    for (int k = 0; k < GPP; ++k)
      if (sval[k]) {
#pragma unroll
        for (int m = 0; m < 5; ++m) pq[k][m] = Q[m * g.fs + soff[k] + zo];
      }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int k = 0; k < GPP; ++k)
      if (sval[k]) {
        double u[3], pr, T, mu = 0.0;
        prims_of<3, SUTH>(pq[k], ph, u, pr, T, mu);
        const int pt = tid + k * GNT;
        sp[slot][0][pt] = u[0];
        sp[slot][1][pt] = u[1];
        sp[slot][2][pt] = u[2];
        sp[slot][3][pt] = T;
        if constexpr (SUTH) sp[slot][4][pt] = mu;
      }
  };
  auto slot_of = [](int zp) { return (zp + 8) & 3; };
  const int c = (ty + 2) * GPX + tx + 2;
  auto d1 = [](double fm2, double fm1, double fp1, double fp2) {
    return (1.0 / 12.0) * (fm2 - fp2) + (2.0 / 3.0) * (fp1 - fm1);
  };
  auto d2 = [](double fm2, double fm1, double f0, double fp1, double fp2) {
    return (-1.0 / 12.0) * (fm2 + fp2) + (4.0 / 3.0) * (fm1 + fp1) - 2.5 * f0;
  };
  auto dx1 = [&](const double* f, int o) { return d1(f[o - 2], f[o - 1], f[o + 1], f[o + 2]); };
  auto dy1 = [&](const double* f, int o) { return d1(f[o - 2 * GPX], f[o - GPX], f[o + GPX], f[o + 2 * GPX]); };
  auto dx2 = [&](const double* f, int o) { return d2(f[o - 2], f[o - 1], f[o], f[o + 1], f[o + 2]); };
  auto dy2 = [&](const double* f, int o) {
    return d2(f[o - 2 * GPX], f[o - GPX], f[o], f[o + GPX], f[o + 2 * GPX]);
  };
  auto column = [&](int slot, double* cq) {
#pragma unroll
    for (int e = 0; e < NF; ++e) cq[e] = sp[slot][e][c];
    cq[NF + 0] = ix * dx1(sp[slot][0], c);
    cq[NF + 1] = ix * dx1(sp[slot][2], c);
    cq[NF + 2] = iy * dy1(sp[slot][1], c);
    cq[NF + 3] = iy * dy1(sp[slot][2], c);
  };

  // software pipeline: planes z0-2 .. z0+2 are staged first; while plane z
  // is evaluated, the raw Q of plane z+3 is in flight, and plane z+3 then
  // takes the slot plane z-1 left, one barrier per plane
  double zq[5][NQ];
  for (int k = 0; k < 4; ++k) {
    fetch(z0 - 2 + k);
    store(slot_of(z0 - 2 + k));
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) column(slot_of(z0 - 2 + k), zq[k]);
  fetch(z0 + 2);
  __syncthreads();
  store(slot_of(z0 + 2));
  __syncthreads();
  column(slot_of(z0 + 2), zq[4]);
  for (int z = z0; z < z1; ++z) {
    const bool more = z + 1 < z1;  // plane z+3 feeds the next plane's queue entry
    if (more) fetch(z + 3);
    if (inside && !slow) {
    const int o = g.idx(x, y, z);
    double r[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    {
      const int sl = slot_of(z);
      const double* U[3] = {sp[sl][0], sp[sl][1], sp[sl][2]};
      const double* T = sp[sl][3];
      const double G[3] = {G0, G1, 1.0}, S[3] = {S0, S1, 0.0};
      auto qz1 = [&](int e) { return iz * d1(zq[0][e], zq[1][e], zq[3][e], zq[4][e]); };
      auto qz2 = [&](int e) { return iz * iz * d2(zq[0][e], zq[1][e], zq[2][e], zq[3][e], zq[4][e]); };
      double DU[3][3], DD[3][3][3];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        double du[3], dd[3];
        du[0] = ix * dx1(U[m], c);
        dd[0] = ix * ix * dx2(U[m], c);
        du[1] = iy * dy1(U[m], c);
        dd[1] = iy * iy * dy2(U[m], c);
        du[2] = qz1(m);
        dd[2] = qz2(m);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          DU[m][j] = str(j) ? G[j] * du[j] : du[j];
          DD[m][j][j] = str(j) ? G[j] * G[j] * dd[j] + G[j] * S[j] * du[j] : dd[j];
        }
      }
      // mixed D_y(D_x u_m), m = 0, 1
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const double* f = U[m];
        double v = iy * d1(ix * dx1(f, c - 2 * GPX), ix * dx1(f, c - GPX), ix * dx1(f, c + GPX),
                           ix * dx1(f, c + 2 * GPX));
        v = gg(0, 1, v, G);
        DD[m][0][1] = v;
        DD[m][1][0] = v;
      }
      // mixed D_z(D_x u_m), m = 0, 2 and D_z(D_y u_m), m = 1, 2 from the queue
      {
        double v = gg(0, 2, qz1(NF + 0), G);
        DD[0][0][2] = DD[0][2][0] = v;
        v = gg(0, 2, qz1(NF + 1), G);
        DD[2][0][2] = DD[2][2][0] = v;
        v = gg(1, 2, qz1(NF + 2), G);
        DD[1][1][2] = DD[1][2][1] = v;
        v = gg(1, 2, qz1(NF + 3), G);
        DD[2][1][2] = DD[2][2][1] = v;
      }
      double D2T[3], DT[3], DMU[3];
      {
        double t1[3], t2[3];
        t1[0] = ix * dx1(T, c);
        t2[0] = ix * ix * dx2(T, c);
        t1[1] = iy * dy1(T, c);
        t2[1] = iy * iy * dy2(T, c);
        t1[2] = qz1(3);
        t2[2] = qz2(3);
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          DT[j] = 0.0;
          DMU[j] = 0.0;
          if constexpr (SUTH || metric) {
            DT[j] = str(j) ? G[j] * t1[j] : t1[j];
            D2T[j] = str(j) ? G[j] * G[j] * t2[j] + G[j] * S[j] * t1[j] : t2[j];
          } else {
            D2T[j] = t2[j];
          }
        }
        if constexpr (SUTH) {
          const double* MU = sp[sl][4];
          const double m1[3] = {ix * dx1(MU, c), iy * dy1(MU, c), qz1(4)};
#pragma unroll
          for (int j = 0; j < 3; ++j) DMU[j] = str(j) ? G[j] * m1[j] : m1[j];
        }
      }
      const double um[3] = {zq[2][0], zq[2][1], zq[2][2]};
      viscous_assemble<3>(SUTH ? zq[2][4] : 1.0, um, DU, DD, D2T, DT, DMU, ph, SUTH, r);
    }
    // the terms add to R's momentum and energy components (r[0] stays 0)
#pragma unroll
    for (int m = 1; m < 5; ++m) R[m * g.fs + o] += r[m];
    }
    if (more) {
      store(slot_of(z + 3));
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int e = 0; e < NQ; ++e) zq[k][e] = zq[k + 1][e];
      column(slot_of(z + 3), zq[4]);
    }
  }
}

// The viscous terms (added to R) at the points k_visc_gen3d leaves out:
// the closure rows of the x or y faces.  part 0: rows y in ys[] (all x);
// part 1: columns x in xs[] at the other rows.
struct ShellRows {
  int ys[8], ny, xs[8], nx;
};

__global__ void __launch_bounds__(128) k_visc_shell(const double* __restrict__ P, double* __restrict__ R, Grid g,
                                                    Phys ph, Bnd b, int suth, ShellRows sr, int part) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.z;
  int i, j;
  if (part == 0) {
    if (t >= g.n[0]) return;
    i = t;
    j = sr.ys[blockIdx.y];
  } else {
    // the columns of one y are adjacent threads (a face's columns share cache lines)
    if (t >= g.n[1] * sr.nx) return;
    j = t / sr.nx;
    i = sr.xs[t % sr.nx];
    for (int q = 0; q < sr.ny; ++q)
      if (sr.ys[q] == j) return;
  }
  const int x = g.idx(i, j, k);
  const int st[3] = {1, g.pitch, g.plane};
  const int ci[3] = {i, j, k};
  double r[5];
#pragma unroll
  for (int m = 0; m < 5; ++m) r[m] = R[m * g.fs + x];
  viscous_gen<3>(P, g.fs, x, st, ci, g, b, ph, suth != 0, r);
#pragma unroll
  for (int m = 0; m < 5; ++m) R[m * g.fs + x] = r[m];
}

void launch_visc_gen3d(const double* Q, const double* P, double* R, const Grid& g, const Phys& ph, const Bnd& b,
                       int suth, int sz, hipStream_t s) {
  dim3 grid((g.n[0] + GTX - 1) / GTX, (g.n[1] + GTY - 1) / GTY, (g.n[2] + sz - 1) / sz);
  dim3 block(GTX, GTY);
  const int ms = ((b.dm[0] || b.sd[0]) ? 1 : 0) | ((b.dm[1] || b.sd[1]) ? 2 : 0);
#define OSBL_VISC_GEN3D(SU, MS) k_visc_gen3d<SU, MS><<<grid, block, 0, s>>>(Q, R, g, ph, b, sz)
  switch (ms + (suth ? 4 : 0)) {
    case 0: OSBL_VISC_GEN3D(false, 0); break;
    case 1: OSBL_VISC_GEN3D(false, 1); break;
    case 2: OSBL_VISC_GEN3D(false, 2); break;
    case 3: OSBL_VISC_GEN3D(false, 3); break;
    case 4: OSBL_VISC_GEN3D(true, 0); break;
    case 5: OSBL_VISC_GEN3D(true, 1); break;
    case 6: OSBL_VISC_GEN3D(true, 2); break;
    default: OSBL_VISC_GEN3D(true, 3); break;
  }
#undef OSBL_VISC_GEN3D
  // the closure rows and columns (distinct indices; a face's rows may overlap
  // the other face's on tiny blocks, hence the de-duplication)
  ShellRows sr{};
  auto add = [](int* list, int& n, int v) {
    for (int q = 0; q < n; ++q)
      if (list[q] == v) return;
    list[n++] = v;
  };
  for (int r = 0; r < closure_rows(b.mod[1][0]); ++r) add(sr.ys, sr.ny, r);
  for (int r = 0; r < closure_rows(b.mod[1][1]); ++r) add(sr.ys, sr.ny, g.n[1] - 1 - r);
  for (int r = 0; r < closure_rows(b.mod[0][0]); ++r) add(sr.xs, sr.nx, r);
  for (int r = 0; r < closure_rows(b.mod[0][1]); ++r) add(sr.xs, sr.nx, g.n[0] - 1 - r);
  if (sr.ny) k_visc_shell<<<dim3((g.n[0] + 127) / 128, sr.ny, g.n[2]), 128, 0, s>>>(P, R, g, ph, b, suth, sr, 0);
  if (sr.nx) k_visc_shell<<<dim3((g.n[1] * sr.nx + 127) / 128, 1, g.n[2]), 128, 0, s>>>(P, R, g, ph, b, suth, sr, 1);
}

// Constituent relations with the Sutherland viscosity over the padded domain:
// u_m, p, T, mu (katzer_SBLI.py:21-26).
template <int ND>
__global__ void __launch_bounds__(256) k_prims_mu(const double* __restrict__ Q, double* __restrict__ P, Grid g,
                                                  Phys ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x - g.h[0];
  const int j = blockIdx.y - g.h[1];
  const int k = blockIdx.z - g.h[2];
  if (i >= g.n[0] + g.h[0]) return;
  const int x = g.idx(i, j, k);
  double u[ND], p, T, mu;
  prims_at<ND, true>(Q, g.fs, x, ph, u, p, T, mu);
#pragma unroll
  for (int m = 0; m < ND; ++m) P[m * g.fs + x] = u[m];
  P[ND * g.fs + x] = p;
  P[(ND + 1) * g.fs + x] = T;
  P[(ND + 2) * g.fs + x] = mu;
}

// The same over the slabs within W points of the x and y faces only (every
// z): what k_visc_shell reads when the tiled kernel forms the constituent
// relations itself.  Closure rows (at most 4) plus the reach of their
// one-sided and mixed stencils (at most 6) stay inside W = 10.
constexpr int PRIM_SLAB = 10;
template <int ND, bool MU>
__global__ void __launch_bounds__(256) k_prims_slab(const double* __restrict__ Q, double* __restrict__ P, Grid g,
                                                    Phys ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x - g.h[0];
  const int j = blockIdx.y - g.h[1];
  const int k = blockIdx.z - g.h[2];
  if (i >= g.n[0] + g.h[0]) return;
  if (!(i < PRIM_SLAB || i >= g.n[0] - PRIM_SLAB || j < PRIM_SLAB || j >= g.n[1] - PRIM_SLAB)) return;
  const int x = g.idx(i, j, k);
  double u[ND], p, T, mu = 0.0;
  prims_at<ND, MU>(Q, g.fs, x, ph, u, p, T, mu);
#pragma unroll
  for (int m = 0; m < ND; ++m) P[m * g.fs + x] = u[m];
  P[ND * g.fs + x] = p;
  P[(ND + 1) * g.fs + x] = T;
  if constexpr (MU) P[(ND + 2) * g.fs + x] = mu;
}

// Modified Ducros sensor (shock_sensors.py:12-49) as a constituent relation:
// theta = 1/2 (1 - tanh(2.5 (1 + 100 div u))) div^2 / (div^2 + |curl u|^2 + epsilon)
// over the interior and one halo point on each side (the sweeps read it at
// every interface's left point, interfaces -1/2 .. N-1/2), with the block's
// central derivatives (closures, metric; halo points of a stretched
// direction take the mirrored metric of metric.py:359-379).
template <int ND>
__global__ void __launch_bounds__(256) k_sensor(const double* __restrict__ Q, double* __restrict__ TH, Grid g,
                                                Phys ph, Bnd b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x - 1;
  const int j = ND >= 2 ? (int)blockIdx.y - 1 : 0;
  const int k = ND >= 3 ? (int)blockIdx.z - 1 : 0;
  if (i > g.n[0]) return;
  const int x = g.idx(i, j, k);
  const int st[3] = {1, g.pitch, g.plane};
  const int ci[3] = {i, j, k};
  double du[3][3] = {{0.0}};
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    double G = 1.0;
    if (b.dm[d]) {
      int c = ci[d];
      const int N = g.n[d];
      if (c < 0) c = -c;
      if (c > N - 1) c = 2 * (N - 1) - c;
      G = b.dm[d][c];
    }
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      const double* qm = Q + (1 + m) * g.fs;
      const double v = ph.inv_delta[d] * dxi<1>(b, d, ci[d], g.n[d], x, st[d], [&](int y) { return qm[y] / Q[y]; });
      du[m][d] = G * v;
    }
  }
  double div = 0.0, w2;
#pragma unroll
  for (int d = 0; d < ND; ++d) div += du[d][d];
  if constexpr (ND == 2) {
    const double w = du[1][0] - du[0][1];
    w2 = w * w;
  } else {
    const double wa = du[2][1] - du[1][2], wb = du[0][2] - du[2][0], wc = du[1][0] - du[0][1];
    w2 = wa * wa + wb * wb + wc * wc;
  }
  const double filt = 0.5 * (1.0 - tanh(2.5 * (1.0 + 100.0 * div)));
  TH[x] = filt * (div * div) / (div * div + w2 + ph.sensor_eps);
}

void launch_sensor(int nd, const double* Q, double* theta, const Grid& g, const Phys& ph, const Bnd& b,
                   hipStream_t s) {
  dim3 grid((g.n[0] + 2 + 255) / 256, nd >= 2 ? g.n[1] + 2 : 1, nd >= 3 ? g.n[2] + 2 : 1);
  if (nd == 3) k_sensor<3><<<grid, 256, 0, s>>>(Q, theta, g, ph, b);
  else k_sensor<2><<<grid, 256, 0, s>>>(Q, theta, g, ph, b);
}

template <int ND, int CONV, bool VISC>
static void launch_gen_t(const double* Q, const double* P, double* R, double* Qo, double* Qn, const Grid& g,
                         const Phys& ph, const Bnd& b, double cA, double cB, int first, int mode, int suth,
                         hipStream_t s) {
  dim3 grid((g.n[0] + 127) / 128, g.n[1], g.n[2]);
  k_update_gen<ND, CONV, VISC><<<grid, 128, 0, s>>>(Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth);
}

template <int ND>
static void launch_gen_nd(int conv, bool visc, const double* Q, const double* P, double* R, double* Qo, double* Qn,
                          const Grid& g, const Phys& ph, const Bnd& b, double cA, double cB, int first, int mode,
                          int suth, hipStream_t s) {
  if (conv == OSBL_CONV_CENTRAL_SKEW) {
    if (visc) launch_gen_t<ND, OSBL_CONV_CENTRAL_SKEW, true>(Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
    else launch_gen_t<ND, OSBL_CONV_CENTRAL_SKEW, false>(Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
  } else if (conv == OSBL_CONV_CENTRAL_CONS) {
    if (visc) launch_gen_t<ND, OSBL_CONV_CENTRAL_CONS, true>(Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
    else launch_gen_t<ND, OSBL_CONV_CENTRAL_CONS, false>(Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
  } else {
    if (visc) launch_gen_t<ND, OSBL_CONV_WENO_Z, true>(Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
    else launch_gen_t<ND, OSBL_CONV_WENO_Z, false>(Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
  }
}

void launch_update_gen(int nd, int conv, bool visc, int mode, const double* Q, const double* P, double* R,
                       double* Qo, double* Qn, const Grid& g, const Phys& ph, const Bnd& b, double cA, double cB,
                       int first, int suth, hipStream_t s) {
  if (nd == 3) launch_gen_nd<3>(conv, visc, Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
  else if (nd == 2) launch_gen_nd<2>(conv, visc, Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
  else launch_gen_nd<1>(conv, visc, Q, P, R, Qo, Qn, g, ph, b, cA, cB, first, mode, suth, s);
}

void launch_body_force3d(const double* Q, double* R, const Grid& g, const Phys& ph, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  k_body_force3d<<<grid, 256, 0, s>>>(Q, R, g, ph);
}

void launch_prims_slab(const double* Q, double* P, const Grid& g, const Phys& ph, int suth, hipStream_t s) {
  const int PX = g.n[0] + 2 * g.h[0];
  dim3 grid((PX + 255) / 256, g.n[1] + 2 * g.h[1], g.n[2] + 2 * g.h[2]);
  if (suth) k_prims_slab<3, true><<<grid, 256, 0, s>>>(Q, P, g, ph);
  else k_prims_slab<3, false><<<grid, 256, 0, s>>>(Q, P, g, ph);
}

void launch_prims_mu(int nd, const double* Q, double* P, const Grid& g, const Phys& ph, hipStream_t s) {
  const int PX = g.n[0] + 2 * g.h[0];
  dim3 grid((PX + 255) / 256, g.n[1] + 2 * g.h[1], g.n[2] + 2 * g.h[2]);
  if (nd == 3) k_prims_mu<3><<<grid, 256, 0, s>>>(Q, P, g, ph);
  else if (nd == 2) k_prims_mu<2><<<grid, 256, 0, s>>>(Q, P, g, ph);
  else k_prims_mu<1><<<grid, 256, 0, s>>>(Q, P, g, ph);
}

}  // namespace osbl
