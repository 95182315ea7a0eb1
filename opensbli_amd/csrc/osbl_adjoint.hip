// osbl_adjoint.hip -- discrete adjoint of the central (skew-symmetric) Euler
// RHS and of the 3-register RK3 step, on periodic blocks.
//
// Forward residual (parsing.py:75-111, scheme.py:187-252), per direction j:
//   R_m -= 1/2 [D_j(q_m u_j) + u_j D_j q_m + q_m D_j u_j]
//   R_{1+j} -= D_j p,   R_E -= D_j(p u_j)
// with u_j = q_{1+j}/rho, p = (gama-1)(q_E - rho |u|^2/2).  The central D_j
// has antisymmetric weights, so on a periodic grid D_j^T = -D_j and the
// vector-Jacobian product lambda = (dR/dQ)^T mu is again a stencil:
//   g_q[m] += 1/2 u_j D_j mu_m + 1/2 D_j(mu_m u_j) - 1/2 mu_m D_j u_j
//   g_u[j] += 1/2 sum_m q_m D_j mu_m - 1/2 sum_m mu_m D_j q_m
//             + 1/2 D_j(sum_m mu_m q_m) + p D_j mu_E
//   g_p    += D_j mu_{1+j} + u_j D_j mu_E
// followed by the chain rule through u_j and p.
#include "osbl_kernels.h"

namespace osbl {

template <int ND>
__global__ void __launch_bounds__(128) k_rhs_vjp(const double* __restrict__ Q, const double* __restrict__ P,
                                                 const double* __restrict__ M, double* __restrict__ out, Grid g,
                                                 Phys ph) {
  constexpr int NV = ND + 2, E = NV - 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j0 = blockIdx.y, k0 = blockIdx.z;
  const int x = g.idx(i, j0, k0);
  const int st[3] = {1, g.pitch, g.plane};
  const long long fs = g.fs;
  double q[NV], mu[NV], u[ND];
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    q[m] = Q[m * fs + x];
    mu[m] = M[m * fs + x];
  }
#pragma unroll
  for (int d = 0; d < ND; ++d) u[d] = P[d * fs + x];
  const double p = P[ND * fs + x];
  double gq[NV], gu[ND], gp = 0.0;
#pragma unroll
  for (int m = 0; m < NV; ++m) gq[m] = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    const int sj = st[j];
    const double inv = ph.inv_delta[j];
    const int offs[4] = {x - 2 * sj, x - sj, x + sj, x + 2 * sj};
    constexpr double W1A = 1.0 / 12.0, W1B = 2.0 / 3.0;
    const double wts[4] = {W1A, -W1B, W1B, -W1A};
    double dmu[NV], dmuu[NV], dq[NV], dmuq = 0.0, duj = 0.0;
#pragma unroll
    for (int m = 0; m < NV; ++m) { dmu[m] = 0.0; dmuu[m] = 0.0; dq[m] = 0.0; }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const int y = offs[s];
      const double w = wts[s];
      const double uj = P[j * fs + y];
      double sq = 0.0;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double my = M[m * fs + y], qy = Q[m * fs + y];
        dmu[m] += w * my;
        dmuu[m] += w * (my * uj);
        dq[m] += w * qy;
        sq += my * qy;
      }
      dmuq += w * sq;
      duj += w * uj;
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) { dmu[m] *= inv; dmuu[m] *= inv; dq[m] *= inv; }
    dmuq *= inv;
    duj *= inv;
    double gu_j = 0.5 * dmuq + p * dmu[E];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      gq[m] += 0.5 * u[j] * dmu[m] + 0.5 * dmuu[m] - 0.5 * mu[m] * duj;
      gu_j += 0.5 * q[m] * dmu[m] - 0.5 * mu[m] * dq[m];
    }
    gu[j] = gu_j;
    gp += dmu[1 + j] + u[j] * dmu[E];
  }
  const double rho = q[0], ir = 1.0 / rho;
  double uu = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) uu += u[d] * u[d];
  const long long cells = (long long)g.n[0] * g.n[1] * g.n[2];
  const long long c = (long long)i + (long long)g.n[0] * (j0 + (long long)g.n[1] * k0);
  double l0 = gq[0] + gp * ph.gm1 * 0.5 * uu;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    l0 -= gu[d] * u[d] * ir;
    out[(1 + d) * cells + c] += gq[1 + d] + gu[d] * ir - gp * ph.gm1 * u[d];
  }
  out[c] += l0;
  out[E * cells + c] += gq[E] + gp * ph.gm1;
}

// out = a*x + b*y (packed arrays of n doubles); y may be null
__global__ void __launch_bounds__(256) k_lincomb(double* __restrict__ out, double a, const double* __restrict__ xv,
                                                  double b, const double* __restrict__ yv, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double v = a * xv[t];
  if (yv) v += b * yv[t];
  out[t] = v;
}

// ---------------------------------------------------------------------------
// One reverse RK stage in one launch, on packed interior arrays with periodic
// wrap (no halo copies): the adjoint weights are mu = alpha*A + beta*B,
// formed on the fly at every stencil point, V = (dR/dQ at Qs)^T mu, and
//   out1 = c*(A + B) + V      (c = 0 or 1)
//   out2 = A + B              (when out2 is given)
// ---------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(128) k_adj_stage(const double* __restrict__ Qs, const double* __restrict__ A,
                                                   const double* __restrict__ B, double alpha, double beta,
                                                   double c, double* __restrict__ out1, double* __restrict__ out2,
                                                   int n0, int n1, int n2, Phys ph) {
  constexpr int NV = ND + 2, E = NV - 1;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n0) return;
  const int j0 = blockIdx.y, k0 = blockIdx.z;
  const long long cells = (long long)n0 * n1 * n2;
  const int nn[3] = {n0, n1, n2};
  const int co[3] = {i, j0, k0};
  const long long sd[3] = {1, n0, (long long)n0 * n1};
  const long long x = i + (long long)n0 * (j0 + (long long)n1 * k0);
  auto muv = [&](int m, long long y) {
    double v = alpha * A[m * cells + y];
    if (B) v += beta * B[m * cells + y];
    return v;
  };
  auto prim = [&](long long y, double* u, double& p, double* q) {
#pragma unroll
    for (int m = 0; m < NV; ++m) q[m] = Qs[m * cells + y];
    const double ir = 1.0 / q[0];
    double ke = 0.0;
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      u[d] = q[1 + d] * ir;
      ke += u[d] * u[d];
    }
    p = ph.gm1 * (q[E] - 0.5 * q[0] * ke);
  };
  double q[NV], u[ND], p, mu[NV];
  prim(x, u, p, q);
#pragma unroll
  for (int m = 0; m < NV; ++m) mu[m] = muv(m, x);
  double gq[NV], gu[ND], gp = 0.0;
#pragma unroll
  for (int m = 0; m < NV; ++m) gq[m] = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    const double inv = ph.inv_delta[j];
    constexpr double W1A = 1.0 / 12.0, W1B = 2.0 / 3.0;
    const int offs[4] = {-2, -1, 1, 2};
    const double wts[4] = {W1A, -W1B, W1B, -W1A};
    double dmu[NV], dmuu[NV], dq[NV], dmuq = 0.0, duj = 0.0;
#pragma unroll
    for (int m = 0; m < NV; ++m) { dmu[m] = 0.0; dmuu[m] = 0.0; dq[m] = 0.0; }
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      int cj = co[j] + offs[s];
      cj = cj < 0 ? cj + nn[j] : (cj >= nn[j] ? cj - nn[j] : cj);
      const long long y = x + (long long)(cj - co[j]) * sd[j];
      const double w = wts[s];
      double qy[NV], uy[ND], py;
      prim(y, uy, py, qy);
      const double uj = uy[j];
      double sq = 0.0;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double my = muv(m, y);
        dmu[m] += w * my;
        dmuu[m] += w * (my * uj);
        dq[m] += w * qy[m];
        sq += my * qy[m];
      }
      dmuq += w * sq;
      duj += w * uj;
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) { dmu[m] *= inv; dmuu[m] *= inv; dq[m] *= inv; }
    dmuq *= inv;
    duj *= inv;
    double gu_j = 0.5 * dmuq + p * dmu[E];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      gq[m] += 0.5 * u[j] * dmu[m] + 0.5 * dmuu[m] - 0.5 * mu[m] * duj;
      gu_j += 0.5 * q[m] * dmu[m] - 0.5 * mu[m] * dq[m];
    }
    gu[j] = gu_j;
    gp += dmu[1 + j] + u[j] * dmu[E];
  }
  const double ir = 1.0 / q[0];
  double uu = 0.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) uu += u[d] * u[d];
  double v[NV];
  v[0] = gq[0] + gp * ph.gm1 * 0.5 * uu;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    v[0] -= gu[d] * u[d] * ir;
    v[1 + d] = gq[1 + d] + gu[d] * ir - gp * ph.gm1 * u[d];
  }
  v[E] = gq[E] + gp * ph.gm1;
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    const double ab = A[m * cells + x] + (B ? B[m * cells + x] : 0.0);
    out1[m * cells + x] = c * ab + v[m];
    if (out2) out2[m * cells + x] = ab;
  }
}

// 2-D variant of k_adj_stage with the stencil inputs staged through LDS: a
// 64 x 4 block of cells loads the (64+4) x (4+4) tile of q, u and the adjoint
// weights mu = alpha A + beta B once (periodic wrap), and every cell's four
// neighbours per direction are then read from LDS.  Same arithmetic as
// k_adj_stage.
constexpr int ATX = 64, ATY = 4, APX = ATX + 4, APY = ATY + 4, ANP = APX * APY;

__global__ void __launch_bounds__(256) k_adj_stage2d(const double* __restrict__ Qs, const double* __restrict__ A,
                                                     const double* __restrict__ B, double alpha, double beta,
                                                     double c, double* __restrict__ out1, double* __restrict__ out2,
                                                     int n0, int n1, Phys ph) {
  constexpr int NV = 4, E = 3;
  __shared__ double s_mu[NV][ANP];
  __shared__ double s_q[NV][ANP];
  __shared__ double s_u[2][ANP];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * ATX + tx;
  const int x0 = blockIdx.x * ATX, y0 = blockIdx.y * ATY;
  const long long cells = (long long)n0 * n1;
  for (int p = tid; p < ANP; p += ATX * ATY) {
    const int xx = p % APX, yy = p / APX;
    int gx = (x0 - 2 + xx) % n0;
    int gy = (y0 - 2 + yy) % n1;
    gx = gx < 0 ? gx + n0 : gx;
    gy = gy < 0 ? gy + n1 : gy;
    const long long y = gx + (long long)n0 * gy;
    double q[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      q[m] = Qs[m * cells + y];
      s_q[m][p] = q[m];
      double v = alpha * A[m * cells + y];
      if (B) v += beta * B[m * cells + y];
      s_mu[m][p] = v;
    }
    const double ir = 1.0 / q[0];
    s_u[0][p] = q[1] * ir;
    s_u[1][p] = q[2] * ir;
  }
  __syncthreads();
  const int i = x0 + tx, j0 = y0 + ty;
  if (i >= n0 || j0 >= n1) return;
  const int cc = (ty + 2) * APX + tx + 2;
  const long long x = i + (long long)n0 * j0;
  double q[NV], u[2], mu[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    q[m] = s_q[m][cc];
    mu[m] = s_mu[m][cc];
  }
  u[0] = s_u[0][cc];
  u[1] = s_u[1][cc];
  const double p = ph.gm1 * (q[E] - 0.5 * q[0] * (u[0] * u[0] + u[1] * u[1]));
  double gq[NV], gu[2], gp = 0.0;
#pragma unroll
  for (int m = 0; m < NV; ++m) gq[m] = 0.0;
#pragma unroll
  for (int jd = 0; jd < 2; ++jd) {
    const double inv = ph.inv_delta[jd];
    constexpr double W1A = 1.0 / 12.0, W1B = 2.0 / 3.0;
    const int st = jd == 0 ? 1 : APX;
    const int offs[4] = {-2 * st, -st, st, 2 * st};
    const double wts[4] = {W1A, -W1B, W1B, -W1A};
    double dmu[NV], dmuu[NV], dq[NV], dmuq = 0.0, duj = 0.0;
#pragma unroll
    for (int m = 0; m < NV; ++m) { dmu[m] = 0.0; dmuu[m] = 0.0; dq[m] = 0.0; }
#pragma unroll
    for (int sI = 0; sI < 4; ++sI) {
      const int y = cc + offs[sI];
      const double w = wts[sI];
      const double uj = s_u[jd][y];
      double sq = 0.0;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double my = s_mu[m][y];
        const double qy = s_q[m][y];
        dmu[m] += w * my;
        dmuu[m] += w * (my * uj);
        dq[m] += w * qy;
        sq += my * qy;
      }
      dmuq += w * sq;
      duj += w * uj;
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) { dmu[m] *= inv; dmuu[m] *= inv; dq[m] *= inv; }
    dmuq *= inv;
    duj *= inv;
    double gu_j = 0.5 * dmuq + p * dmu[E];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      gq[m] += 0.5 * u[jd] * dmu[m] + 0.5 * dmuu[m] - 0.5 * mu[m] * duj;
      gu_j += 0.5 * q[m] * dmu[m] - 0.5 * mu[m] * dq[m];
    }
    gu[jd] = gu_j;
    gp += dmu[1 + jd] + u[jd] * dmu[E];
  }
  const double ir = 1.0 / q[0];
  const double uu = u[0] * u[0] + u[1] * u[1];
  double v[NV];
  v[0] = gq[0] + gp * ph.gm1 * 0.5 * uu;
#pragma unroll
  for (int d = 0; d < 2; ++d) {
    v[0] -= gu[d] * u[d] * ir;
    v[1 + d] = gq[1 + d] + gu[d] * ir - gp * ph.gm1 * u[d];
  }
  v[E] = gq[E] + gp * ph.gm1;
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    const double ab = A[m * cells + x] + (B ? B[m * cells + x] : 0.0);
    out1[m * cells + x] = c * ab + v[m];
    if (out2) out2[m * cells + x] = ab;
  }
}

void launch_adj_stage(int nd, const double* Qs, const double* A, const double* B, double alpha, double beta, double c,
                      double* out1, double* out2, const Grid& g, const Phys& ph, hipStream_t s) {
  if (nd == 2) {
    dim3 grid((g.n[0] + ATX - 1) / ATX, (g.n[1] + ATY - 1) / ATY);
    k_adj_stage2d<<<grid, dim3(ATX, ATY), 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, g.n[0], g.n[1], ph);
    return;
  }
  dim3 grid((g.n[0] + 127) / 128, g.n[1], g.n[2]);
  if (nd == 3) k_adj_stage<3><<<grid, 128, 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, g.n[0], g.n[1], g.n[2], ph);
  else k_adj_stage<1><<<grid, 128, 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, g.n[0], g.n[1], g.n[2], ph);
}

// pack every variable of a bundle into contiguous interior blocks in one launch
__global__ void __launch_bounds__(256) k_pack_all(const double* __restrict__ F, double* __restrict__ out, Grid g, int nv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j = blockIdx.y, k = blockIdx.z;
  const long long cells = (long long)g.n[0] * g.n[1] * g.n[2];
  const long long c = (long long)i + (long long)g.n[0] * (j + (long long)g.n[1] * k);
  const int x = g.idx(i, j, k);
  for (int m = 0; m < nv; ++m) out[m * cells + c] = F[m * g.fs + x];
}

void launch_pack_all(const double* F, double* out, const Grid& g, int nv, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  k_pack_all<<<grid, 256, 0, s>>>(F, out, g, nv);
}

void launch_rhs_vjp(int nd, const double* Q, const double* P, const double* M, double* out, const Grid& g,
                    const Phys& ph, hipStream_t s) {
  dim3 grid((g.n[0] + 127) / 128, g.n[1], g.n[2]);
  if (nd == 3) k_rhs_vjp<3><<<grid, 128, 0, s>>>(Q, P, M, out, g, ph);
  else if (nd == 2) k_rhs_vjp<2><<<grid, 128, 0, s>>>(Q, P, M, out, g, ph);
  else k_rhs_vjp<1><<<grid, 128, 0, s>>>(Q, P, M, out, g, ph);
}

void launch_lincomb(double* out, double a, const double* x, double b, const double* y, long long n, hipStream_t s) {
  k_lincomb<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(out, a, x, b, y, n);
}

}  // namespace osbl
