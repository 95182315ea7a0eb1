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
  const double mu = suth ? MU[x] : 1.0;
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
    const double um = U[m][x];
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
  const double rho = Q[x];
  const double ir = 1.0 / rho;
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double u = Q[(1 + m) * g.fs + x] * ir;
    P[m * g.fs + x] = u;
    ke += u * u;
  }
  const double p = ph.gm1 * (Q[(ND + 1) * g.fs + x] - 0.5 * rho * ke);
  const double T = p * ph.gama * ph.Minf * ph.Minf * ir;
  P[ND * g.fs + x] = p;
  P[(ND + 1) * g.fs + x] = T;
  P[(ND + 2) * g.fs + x] = pow(T, 1.5) * (1.0 + ph.suth) / (T + ph.suth);
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

void launch_prims_mu(int nd, const double* Q, double* P, const Grid& g, const Phys& ph, hipStream_t s) {
  const int PX = g.n[0] + 2 * g.h[0];
  dim3 grid((PX + 255) / 256, g.n[1] + 2 * g.h[1], g.n[2] + 2 * g.h[2]);
  if (nd == 3) k_prims_mu<3><<<grid, 256, 0, s>>>(Q, P, g, ph);
  else if (nd == 2) k_prims_mu<2><<<grid, 256, 0, s>>>(Q, P, g, ph);
  else k_prims_mu<1><<<grid, 256, 0, s>>>(Q, P, g, ph);
}

}  // namespace osbl
