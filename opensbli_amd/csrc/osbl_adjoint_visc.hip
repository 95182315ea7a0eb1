// osbl_adjoint_visc.hip -- discrete adjoint of the constant-viscosity terms
// (taylor_green_vortex.py:14-19; the product-rule expansion of
// parsing.py:141-186 with the second and mixed derivatives of
// scheme.py:81-85,360-387) on periodic blocks.
//
// Forward, per cell (D1_j, D2_j the 4th-order central first and second
// derivatives, M_aa = D2_a, M_ab = M_ba = D1_b D1_a for a < b):
//   du_mj = D1_j u_m,  tau_mj = (du_mj + du_jm - 2/3 delta_mj div u) / Re
//   V_{1+m} = (sum_j M_jj u_m + M_mj u_j - 2/3 delta_mj sum_k M_km u_k) / Re
//   V_E     = cq sum_j D2_j T + sum_mj du_mj tau_mj + sum_m u_m V_{1+m}
// On a periodic grid D1^T = -D1 and D2, M_ab are symmetric, so for weights
// mu the vector-Jacobian product is again a stencil:
//   w_m  = mu_{1+m} + mu_E u_m                    (the momentum terms and u_m V_{1+m})
//   gu_k = (sum_j M_jj w_k + 1/3 sum_m M_mk w_m) / Re - sum_b D1_b(2 mu_E tau_kb) + mu_E V_{1+k}
//   gT   = cq sum_j D2_j mu_E
// followed by the chain rule through u_m = rhou_m / rho and
// T = gama Minf^2 p / rho, p = (gama-1)(rhoE - rho |u|^2 / 2).
// Two passes over the padded arrays: the first forms w, 2 mu_E tau and
// mu_E V at every cell and on two halo layers (from periodic images, so no
// halo fill of its own), the second applies the transposed stencils and adds
// the result to `out`.
#include "osbl_kernels.h"

namespace osbl {

namespace {

__device__ __forceinline__ double vd1(const double* __restrict__ f, int x, int st) {
  return (1.0 / 12.0) * (f[x - 2 * st] - f[x + 2 * st]) + (2.0 / 3.0) * (f[x + st] - f[x - st]);
}
__device__ __forceinline__ double vd2(const double* __restrict__ f, int x, int st) {
  return (-1.0 / 12.0) * (f[x - 2 * st] + f[x + 2 * st]) + (4.0 / 3.0) * (f[x - st] + f[x + st]) - 2.5 * f[x];
}
// M_ab f at x: D2_a (a == b) or D1_b(D1_a f) (a != b; the two orders agree on a periodic grid)
__device__ __forceinline__ double vmix(const double* __restrict__ f, int x, int a, int b, const int st[3],
                                       const double inv[3]) {
  if (a == b) return inv[a] * inv[a] * vd2(f, x, st[a]);
  const int sa = st[a], sb = st[b];
  const double im2 = vd1(f, x - 2 * sb, sa), im1 = vd1(f, x - sb, sa);
  const double ip1 = vd1(f, x + sb, sa), ip2 = vd1(f, x + 2 * sb, sa);
  return inv[a] * inv[b] * ((1.0 / 12.0) * (im2 - ip2) + (2.0 / 3.0) * (ip1 - im1));
}

// pass 1: W[m] = w_m, W[ND + ND a + b] = 2 mu_E tau_ab, W[ND + ND^2 + m] = mu_E V_{1+m}
template <int ND>
__global__ void __launch_bounds__(128) k_visc_vjp_pre(const double* __restrict__ P, const double* __restrict__ M,
                                                      double* __restrict__ W, Grid g, Phys ph) {
  constexpr int NV = ND + 2, E = NV - 1;
  // the interior and two halo layers (the second pass's stencil reach), from
  // P and M over the padded grid (halo >= 4): no halo fill of W is needed
  const int w2[3] = {2, ND >= 2 ? 2 : 0, ND >= 3 ? 2 : 0};
  const int i = blockIdx.x * blockDim.x + threadIdx.x - w2[0];
  if (i >= g.n[0] + w2[0]) return;
  const int x = g.idx(i, (int)blockIdx.y - w2[1], (int)blockIdx.z - w2[2]);
  const int st[3] = {1, g.pitch, g.plane};
  const long long fs = g.fs;
  const double* U[3] = {P, P + fs, P + 2 * fs};
  double du[ND][ND];
#pragma unroll
  for (int m = 0; m < ND; ++m)
#pragma unroll
    for (int j = 0; j < ND; ++j) du[m][j] = ph.inv_delta[j] * vd1(U[m], x, st[j]);
  double divu = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) divu += du[k][k];
  const double muE = M[E * fs + x];
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int b = 0; b < ND; ++b) {
      const double tau = ph.inv_re * (du[a][b] + du[b][a] - (a == b ? (2.0 / 3.0) * divu : 0.0));
      W[(ND + ND * a + b) * fs + x] = 2.0 * muE * tau;
    }
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    W[m * fs + x] = M[(1 + m) * fs + x] + muE * U[m][x];
    // V_{1+m} = (sum_j M_jj u_m + 1/3 sum_j M_mj u_j) / Re
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < ND; ++j) v += vmix(U[m], x, j, j, st, ph.inv_delta) + (1.0 / 3.0) * vmix(U[j], x, m, j, st, ph.inv_delta);
    W[(ND + ND * ND + m) * fs + x] = muE * ph.inv_re * v;
  }
}

// pass 2: the transposed stencils and the chain rule to the conservative
// variables, added to out (packed, x fastest)
template <int ND>
__global__ void __launch_bounds__(128) k_visc_vjp(const double* __restrict__ q, const double* __restrict__ P,
                                                  const double* __restrict__ M, const double* __restrict__ W,
                                                  double* __restrict__ out, Grid g, Phys ph) {
  constexpr int NV = ND + 2, E = NV - 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j0 = blockIdx.y, k0 = blockIdx.z;
  const int x = g.idx(i, j0, k0);
  const int st[3] = {1, g.pitch, g.plane};
  const long long fs = g.fs;
  double gu[ND];
#pragma unroll
  for (int k = 0; k < ND; ++k) {
    double v = 0.0;
#pragma unroll
    for (int j = 0; j < ND; ++j) v += vmix(W + k * fs, x, j, j, st, ph.inv_delta);
#pragma unroll
    for (int m = 0; m < ND; ++m) v += (1.0 / 3.0) * vmix(W + m * fs, x, m, k, st, ph.inv_delta);
    double s = ph.inv_re * v + W[(ND + ND * ND + k) * fs + x];
#pragma unroll
    for (int b = 0; b < ND; ++b) s -= ph.inv_delta[b] * vd1(W + (ND + ND * k + b) * fs, x, st[b]);
    gu[k] = s;
  }
  double gT = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) gT += ph.inv_delta[j] * ph.inv_delta[j] * vd2(M + E * fs, x, st[j]);
  gT *= ph.cq;
  // chain rule at this cell
  const long long cells = (long long)g.n[0] * g.n[1] * g.n[2];
  const long long o = i + (long long)g.n[0] * (j0 + (long long)g.n[1] * k0);
  const double rho = q[o], ir = 1.0 / rho;
  const double T = P[(ND + 1) * fs + x];
  double uu = 0.0, g0 = 0.0;
  double u[ND];
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    u[m] = P[m * fs + x];
    uu += u[m] * u[m];
  }
  const double gp = gT * ph.gama * ph.Minf * ph.Minf * ir;
  g0 += -gT * T * ir + gp * ph.gm1 * 0.5 * uu;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    g0 += -gu[m] * u[m] * ir;
    out[(1 + m) * cells + o] += gu[m] * ir - gp * ph.gm1 * u[m];
  }
  out[o] += g0;
  out[E * cells + o] += gp * ph.gm1;
}

// packed index of the periodic image of padded point (i, j, k)
__device__ __forceinline__ long long wrap_index(const Grid& g, int i, int j, int k) {
  const int a = (i % g.n[0] + g.n[0]) % g.n[0];
  const int b = (j % g.n[1] + g.n[1]) % g.n[1];
  const int c = (k % g.n[2] + g.n[2]) % g.n[2];
  return a + (long long)g.n[0] * (b + (long long)g.n[1] * c);
}

// the constituent relations u_i, p, T of the packed state q (k_prims's
// formulas) and M = alpha A + beta B (B optional) over the interior and every
// halo layer, each halo point from its periodic image in the packed arrays
template <int ND>
__global__ void __launch_bounds__(256) k_adj_unpack(const double* __restrict__ q, const double* __restrict__ A,
                                                    const double* __restrict__ B, double alpha, double beta,
                                                    double* __restrict__ P, double* __restrict__ M, Grid g, Phys ph) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x - g.h[0];
  if (i >= g.n[0] + g.h[0]) return;
  const int j = (int)blockIdx.y - g.h[1], k = (int)blockIdx.z - g.h[2];
  const long long cells = (long long)g.n[0] * g.n[1] * g.n[2];
  const long long o = wrap_index(g, i, j, k);
  const int x = g.idx(i, j, k);
  double v[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    v[m] = q[m * cells + o];
    double w = alpha * A[m * cells + o];
    if (B) w += beta * B[m * cells + o];
    M[m * g.fs + x] = w;
  }
  const double rho = v[0];
  const double ir = 1.0 / rho;
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double u = v[1 + m] * ir;
    P[m * g.fs + x] = u;
    ke += u * u;
  }
  const double p = ph.gm1 * (v[NV - 1] - 0.5 * rho * ke);
  P[ND * g.fs + x] = p;
  P[(ND + 1) * g.fs + x] = p * ph.gama * ph.Minf * ph.Minf * ir;
}

// the padded state and its constituent relations u_i, p, T (k_prims's
// formulas) over the interior and every halo layer, each halo point from its
// periodic image in the packed state
template <int ND>
__global__ void __launch_bounds__(256) k_unpack_prims(const double* __restrict__ q, double* __restrict__ Q,
                                                      double* __restrict__ P, Grid g, Phys ph) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x - g.h[0];
  if (i >= g.n[0] + g.h[0]) return;
  const int j = (int)blockIdx.y - g.h[1], k = (int)blockIdx.z - g.h[2];
  const long long cells = (long long)g.n[0] * g.n[1] * g.n[2];
  const long long o = wrap_index(g, i, j, k);
  const int x = g.idx(i, j, k);
  double v[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    v[m] = q[m * cells + o];
    Q[m * g.fs + x] = v[m];
  }
  const double rho = v[0];
  const double ir = 1.0 / rho;
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double u = v[1 + m] * ir;
    P[m * g.fs + x] = u;
    ke += u * u;
  }
  const double p = ph.gm1 * (v[NV - 1] - 0.5 * rho * ke);
  P[ND * g.fs + x] = p;
  P[(ND + 1) * g.fs + x] = p * ph.gama * ph.Minf * ph.Minf * ir;
}

}  // namespace

// fields of the scratch launch_visc_vjp needs (2 ND + ND^2 padded fields)
int visc_vjp_fields(int nd) { return 2 * nd + nd * nd; }

// the padded state (every halo layer filled with periodic images) and its
// constituent relations from a packed state, in one pass
void launch_unpack_prims(int nd, const double* q, double* Qpad, double* P, const Grid& g, const Phys& ph,
                         hipStream_t s) {
  dim3 gu((g.n[0] + 2 * g.h[0] + 255) / 256, g.n[1] + 2 * g.h[1], g.n[2] + 2 * g.h[2]);
  if (nd == 3) k_unpack_prims<3><<<gu, 256, 0, s>>>(q, Qpad, P, g, ph);
  else if (nd == 2) k_unpack_prims<2><<<gu, 256, 0, s>>>(q, Qpad, P, g, ph);
  else k_unpack_prims<1><<<gu, 256, 0, s>>>(q, Qpad, P, g, ph);
}

// out (packed) += (dV/dQ at q)^T (alpha A + beta B) for the viscous terms V,
// q the packed state.  P (nd + 2 padded fields), Mpad (nv) and W
// (visc_vjp_fields(nd)) are scratch; the grid's halo is >= 4 (the first pass
// reads P and M two layers beyond the two halo layers of W it forms).
void launch_visc_vjp(int nd, const double* q, const double* A, const double* B, double alpha, double beta, double* P,
                     double* Mpad, double* W, double* out, const Grid& g, const Phys& ph, hipStream_t s) {
  dim3 gu((g.n[0] + 2 * g.h[0] + 255) / 256, g.n[1] + 2 * g.h[1], g.n[2] + 2 * g.h[2]);
  if (nd == 3) k_adj_unpack<3><<<gu, 256, 0, s>>>(q, A, B, alpha, beta, P, Mpad, g, ph);
  else if (nd == 2) k_adj_unpack<2><<<gu, 256, 0, s>>>(q, A, B, alpha, beta, P, Mpad, g, ph);
  else k_adj_unpack<1><<<gu, 256, 0, s>>>(q, A, B, alpha, beta, P, Mpad, g, ph);
  // w and 2 mu_E tau over the interior and two halo layers
  dim3 gp((g.n[0] + 4 + 127) / 128, g.n[1] + (nd >= 2 ? 4 : 0), g.n[2] + (nd >= 3 ? 4 : 0));
  if (nd == 3) k_visc_vjp_pre<3><<<gp, 128, 0, s>>>(P, Mpad, W, g, ph);
  else if (nd == 2) k_visc_vjp_pre<2><<<gp, 128, 0, s>>>(P, Mpad, W, g, ph);
  else k_visc_vjp_pre<1><<<gp, 128, 0, s>>>(P, Mpad, W, g, ph);
  dim3 grid((g.n[0] + 127) / 128, g.n[1], g.n[2]);
  if (nd == 3) k_visc_vjp<3><<<grid, 128, 0, s>>>(q, P, Mpad, W, out, g, ph);
  else if (nd == 2) k_visc_vjp<2><<<grid, 128, 0, s>>>(q, P, Mpad, W, out, g, ph);
  else k_visc_vjp<1><<<grid, 128, 0, s>>>(q, P, Mpad, W, out, g, ph);
}

}  // namespace osbl
