// osbl_curvilinear.hip -- the shock-capturing right-hand side on a general
// curvilinear grid: the Euler equations in strong conservation form
//   d(Q)/dt = -(1/detJ) sum_j d(detJ F_j)/d xi_j,
//   F_j = (rho U_j, rhou_i U_j + p D_ji, (p + rhoE) U_j),  U_j = D_jm u_m
// (apps/euler_wave_curvilinear/euler_wave.py:17-47), with the metric
// D = d xi/d x and detJ = det(d x/d xi) of MetricsEquation (metric.py:72-147).
//
// The LLF characteristic reconstruction (shock_capturing.py:357-448) runs
// about the metric direction k = (D_d0, .., D_d(nd-1)) of each interface:
// k is the simple average of the two neighbours' metric (averaging.py:49-57,
// 104-113), L and R are the Euler eigenvectors about the unit normal k/|k|
// (euler_eigensystem.py:18-40, 88-135), and the local wave speeds use every
// stencil point's own metric, |U|, |U +- a|k|| (shock_capturing.py:512-536).
// The characteristic fluxes are L (detJ F_d) and the characteristic state L Q.
//
// Layout: the padded fields of the block (Grid); the metric holds nd*nd + 1
// padded fields, D_ab at a*nd + b and detJ last.  Per direction: a point
// kernel forms detJ F_d, U and a|k| once per point; one thread per interface
// then projects the six stencil points onto the characteristic fields and
// reconstructs, writing the interface fluxes to a scratch bundle; a
// difference kernel accumulates -(F[i] - F[i-1])/Delta_d and the last
// direction divides by detJ.
#include "osbl_kernels.h"

namespace osbl {

namespace {

// L (rows) and R (columns) of the Euler flux Jacobian about the unit normal
// kt, at the averaged state v; th = kt . u.  Field order (U, U(, U), U+a, U-a).
template <int ND>
__device__ __forceinline__ void curv_lrow(const Avg<ND>& v, const double kt[ND], double th, const Phys& ph, int f,
                                          double l[ND + 2]) {
  constexpr int NV = ND + 2;
  const double g1 = ph.gm1;
  const double ia2 = v.ia * v.ia;
  if (f == ND) {  // U + a
#pragma unroll
    for (int c = 0; c < ND; ++c) l[1 + c] = v.bta * (kt[c] * v.a - g1 * v.u[c]);
    l[0] = v.bta * (v.phi - v.a * th);
    l[NV - 1] = v.bta * g1;
    return;
  }
  if (f == ND + 1) {  // U - a
#pragma unroll
    for (int c = 0; c < ND; ++c) l[1 + c] = -v.bta * (kt[c] * v.a + g1 * v.u[c]);
    l[0] = v.bta * (v.phi + v.a * th);
    l[NV - 1] = v.bta * g1;
    return;
  }
  if constexpr (ND == 2) {
    if (f == 0) {  // entropy wave
      l[0] = 1.0 - v.phi * ia2;
      l[1] = g1 * v.u[0] * ia2;
      l[2] = g1 * v.u[1] * ia2;
      l[3] = -g1 * ia2;
    } else {  // shear wave: the velocity across the normal
      l[0] = -(kt[1] * v.u[0] - kt[0] * v.u[1]) * v.ir;
      l[1] = kt[1] * v.ir;
      l[2] = -kt[0] * v.ir;
      l[3] = 0.0;
    }
  } else {
    // row f: kt_f times the entropy row plus the f-th component of the
    // rotation (kt x u, and e_c x kt for the momentum columns) over rho
    const double cr[3] = {kt[1] * v.u[2] - kt[2] * v.u[1], kt[2] * v.u[0] - kt[0] * v.u[2],
                          kt[0] * v.u[1] - kt[1] * v.u[0]};
    l[0] = kt[f] * (1.0 - v.phi * ia2) + cr[f] * v.ir;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int j = 3 - f - c;  // eps(f, c, j) kt_j; zero for c == f
      const double e = (c == f) ? 0.0 : ((((c - f + 3) % 3) == 1) ? kt[j] : -kt[j]);
      l[1 + c] = kt[f] * g1 * v.u[c] * ia2 + e * v.ir;
    }
    l[4] = -kt[f] * g1 * ia2;
  }
}

template <int ND>
__device__ __forceinline__ void curv_rcol(const Avg<ND>& v, const double kt[ND], double th, const Phys& ph, int f,
                                          double r[ND + 2]) {
  constexpr int NV = ND + 2;
  const double ig1 = ph.igm1;
  if (f >= ND) {  // acoustic waves
    const double sg = (f == ND) ? 1.0 : -1.0;
    r[0] = v.alpha;
#pragma unroll
    for (int c = 0; c < ND; ++c) r[1 + c] = v.alpha * (v.u[c] + sg * kt[c] * v.a);
    r[NV - 1] = v.alpha * ((v.phi + v.a * v.a) * ig1 + sg * th * v.a);
    return;
  }
  if constexpr (ND == 2) {
    if (f == 0) {
      r[0] = 1.0;
      r[1] = v.u[0];
      r[2] = v.u[1];
      r[3] = v.phi * ig1;
    } else {
      r[0] = 0.0;
      r[1] = kt[1] * v.rho;
      r[2] = -kt[0] * v.rho;
      r[3] = v.rho * (kt[1] * v.u[0] - kt[0] * v.u[1]);
    }
  } else {
    const double cr[3] = {kt[1] * v.u[2] - kt[2] * v.u[1], kt[2] * v.u[0] - kt[0] * v.u[2],
                          kt[0] * v.u[1] - kt[1] * v.u[0]};
    r[0] = kt[f];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int j = 3 - f - c;  // eps(f, c, j) kt_j
      const double e = (c == f) ? 0.0 : ((((c - f + 3) % 3) == 1) ? kt[j] : -kt[j]);
      r[1 + c] = kt[f] * v.u[c] + v.rho * e;
    }
    r[4] = kt[f] * v.phi * ig1 - v.rho * cr[f];
  }
}

// thread t of a 1-D launch over the box e[0] x e[1] x e[2] (x fastest);
// false past its end
__device__ __forceinline__ bool box_index(const int e[3], int c[3]) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)e[0] * e[1] * e[2]) return false;
  c[0] = (int)(t % e[0]);
  const long long r = t / e[0];
  c[1] = (int)(r % e[1]);
  c[2] = (int)(r / e[1]);
  return true;
}

// metric of point y along direction d, the contravariant velocity U = k.u and |k|
template <int ND>
__device__ __forceinline__ void point_metric(const double* __restrict__ M, long long fs, int d, int y,
                                             const Pt<ND>& s, double k[ND], double& U, double& km) {
  U = 0.0;
  double kk = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    k[m] = M[(d * ND + m) * fs + y];
    U += k[m] * s.u[m];
    kk += k[m] * k[m];
  }
  km = sqrt(kk);
}

// per point along direction d (d index in [-3, n_d + 3), the others
// interior): the fluxes detJ F_d (nv fields), U = k.u and a|k| with the
// point's own metric k = (D_d0, ..)
template <int ND>
__global__ void __launch_bounds__(256) k_curv_point(const double* __restrict__ Q, const double* __restrict__ M,
                                                    double* __restrict__ PF, Grid g, Phys ph, int d) {
  constexpr int NV = ND + 2;
  const int e[3] = {g.n[0] + (d == 0 ? 6 : 0), g.n[1] + (d == 1 ? 6 : 0), g.n[2] + (d == 2 ? 6 : 0)};
  int c[3];
  if (!box_index(e, c)) return;
  c[d] -= 3;
  const int y = g.idx(c[0], c[1], c[2]);
  const long long fs = g.fs;
  Pt<ND> s;
  load_pt<ND>(s, Q, fs, y, ph);
  double k[ND], U, km;
  point_metric<ND>(M, fs, d, y, s, k, U, km);
  const double dj = M[(ND * ND) * fs + y];
  PF[y] = dj * s.q[0] * U;
#pragma unroll
  for (int m = 0; m < ND; ++m) PF[(1 + m) * fs + y] = dj * (s.q[1 + m] * U + s.p * k[m]);
  PF[(NV - 1) * fs + y] = dj * (s.p + s.q[NV - 1]) * U;
  PF[NV * fs + y] = U;
  PF[(NV + 1) * fs + y] = s.a * km;
}

// one interface i+1/2 along d per thread, i in [-1, n_d); the other indices interior
template <int ND, int CONV, int AVG>
__global__ void __launch_bounds__(128) k_curv_flux(const double* __restrict__ Q, const double* __restrict__ M,
                                                   const double* __restrict__ PF, double* __restrict__ F, Grid g,
                                                   Phys ph, int d) {
  constexpr int NV = ND + 2;
  const int e[3] = {g.n[0] + (d == 0 ? 1 : 0), g.n[1] + (d == 1 ? 1 : 0), g.n[2] + (d == 2 ? 1 : 0)};
  int c[3];
  if (!box_index(e, c)) return;
  c[d] -= 1;
  const int x = g.idx(c[0], c[1], c[2]);
  const int st = g.stride(d);
  const long long fs = g.fs;
  Pt<ND> pl, pr;
  load_pt<ND>(pl, Q, fs, x, ph);
  load_pt<ND>(pr, Q, fs, x + st, ph);
  Avg<ND> v;
  interface_state<ND, AVG>(v, pl, pr, ph);
  double kt[ND];
  double kk = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    kt[m] = 0.5 * (M[(d * ND + m) * fs + x] + M[(d * ND + m) * fs + x + st]);
    kk += kt[m] * kt[m];
  }
  const double ikm = 1.0 / sqrt(kk);
  double th = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    kt[m] *= ikm;
    th += kt[m] * v.u[m];
  }
  double L[NV][NV];
#pragma unroll
  for (int f = 0; f < NV; ++f) curv_lrow<ND>(v, kt, th, ph, f, L[f]);
  // characteristic fluxes and states on the six stencil points, and the
  // local wave speeds (maximum over the points)
  double cf[NV][6], cs[NV][6], lmax[NV];
#pragma unroll
  for (int f = 0; f < NV; ++f) lmax[f] = 0.0;
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    const int y = x + (p - 2) * st;
    double Fv[NV], qv[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      Fv[m] = PF[m * fs + y];
      qv[m] = Q[m * fs + y];
    }
    const double U = PF[NV * fs + y], ak = PF[(NV + 1) * fs + y];
#pragma unroll
    for (int f = 0; f < ND; ++f) lmax[f] = fmax(lmax[f], fabs(U));
    lmax[ND] = fmax(lmax[ND], fabs(U + ak));
    lmax[ND + 1] = fmax(lmax[ND + 1], fabs(U - ak));
#pragma unroll
    for (int f = 0; f < NV; ++f) {
      double a = 0.0, b = 0.0;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        a += L[f][m] * Fv[m];
        b += L[f][m] * qv[m];
      }
      cf[f][p] = a;
      cs[f][p] = b;
    }
  }
  // per characteristic field: split fluxes, reconstruction, and its
  // contribution R[:, f] rec_f to the interface flux
  double out[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) out[m] = 0.0;
#pragma unroll
  for (int f = 0; f < NV; ++f) {
    double fp[6], fm[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      fp[p] = 0.5 * (cf[f][p] + lmax[f] * cs[f][p]);
      fm[p] = 0.5 * (cf[f][p] - lmax[f] * cs[f][p]);
    }
    const double rec = recon_pair<CONV>(fp, fm, ph.eps, ph.teno_ct);
    double r[NV];
    curv_rcol<ND>(v, kt, th, ph, f, r);
#pragma unroll
    for (int m = 0; m < NV; ++m) out[m] += r[m] * rec;
  }
#pragma unroll
  for (int m = 0; m < NV; ++m) F[m * fs + x] = out[m];
}

// R = (first ? 0 : R) - (F[i] - F[i-1])/Delta_d over the interior; the last
// direction divides the sum by detJ
template <int ND>
__global__ void __launch_bounds__(256) k_curv_diff(const double* __restrict__ F, const double* __restrict__ M,
                                                   double* __restrict__ R, Grid g, double delta, int d, int first,
                                                   int last) {
  int c[3];
  if (!box_index(g.n, c)) return;
  const int x = g.idx(c[0], c[1], c[2]);
  const int st = g.stride(d);
  const long long fs = g.fs;
  const double dj = last ? M[(ND * ND) * fs + x] : 1.0;
#pragma unroll
  for (int m = 0; m < ND + 2; ++m) {
    const double v = (F[m * fs + x] - F[m * fs + x - st]) / delta;
    double r = first ? -v : R[m * fs + x] - v;
    if (last) r = r / dj;
    R[m * fs + x] = r;
  }
}

// J_ab = CD(x_a, xi_b) from the coordinates X (nd arrays over the interior
// extended by 2 points on every side, x fastest, extents e), detJ = det J and
// D = adj(J)/detJ over the interior
template <int ND>
__global__ void __launch_bounds__(256) k_curv_metric(const double* __restrict__ X, double* __restrict__ M, Grid g,
                                                     Phys ph, int e0, int e1, int e2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j = blockIdx.y, k = blockIdx.z;
  const long long per = (long long)e0 * e1 * e2;
  const long long base = (i + 2) + (long long)e0 * ((j + (ND > 1 ? 2 : 0)) + (long long)e1 * (k + (ND > 2 ? 2 : 0)));
  const long long sx[3] = {1, e0, (long long)e0 * e1};
  double J[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int b = 0; b < ND; ++b) {
      const double* f = X + a * per + base;
      const long long s = sx[b];
      J[a][b] = ph.inv_delta[b] *
                ((1.0 / 12.0) * f[-2 * s] + (-2.0 / 3.0) * f[-s] + (2.0 / 3.0) * f[s] + (-1.0 / 12.0) * f[2 * s]);
    }
  double adj[3][3], det;
  if constexpr (ND == 2) {
    det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    adj[0][0] = J[1][1];
    adj[0][1] = -J[0][1];
    adj[1][0] = -J[1][0];
    adj[1][1] = J[0][0];
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        const int r0 = (b + 1) % 3, r1 = (b + 2) % 3, c0 = (a + 1) % 3, c1 = (a + 2) % 3;
        adj[a][b] = J[r0][c0] * J[r1][c1] - J[r0][c1] * J[r1][c0];  // cofactor C_ba
      }
    det = J[0][0] * adj[0][0] + J[0][1] * adj[1][0] + J[0][2] * adj[2][0];
  }
  const int x = g.idx(i, j, k);
  const long long fs = g.fs;
  M[(ND * ND) * fs + x] = det;
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int b = 0; b < ND; ++b) M[(a * ND + b) * fs + x] = adj[a][b] / det;
}

template <int ND, int CONV>
void launch_flux_c(int avg, const double* Q, const double* M, const double* PF, double* F, const Grid& g,
                   const Phys& ph, int d, hipStream_t s) {
  const long long total = (long long)(g.n[0] + (d == 0)) * (g.n[1] + (d == 1)) * (g.n[2] + (d == 2));
  const unsigned grid = (unsigned)((total + 127) / 128);
  if (avg == OSBL_AVG_ROE) k_curv_flux<ND, CONV, OSBL_AVG_ROE><<<grid, 128, 0, s>>>(Q, M, PF, F, g, ph, d);
  else k_curv_flux<ND, CONV, OSBL_AVG_SIMPLE><<<grid, 128, 0, s>>>(Q, M, PF, F, g, ph, d);
}

template <int ND>
void launch_flux_nd(int conv, int avg, const double* Q, const double* M, double* PF, double* F, const Grid& g,
                    const Phys& ph, int d, hipStream_t s) {
  const long long pts = (long long)(g.n[0] + (d == 0 ? 6 : 0)) * (g.n[1] + (d == 1 ? 6 : 0)) *
                        (g.n[2] + (d == 2 ? 6 : 0));
  k_curv_point<ND><<<(unsigned)((pts + 255) / 256), 256, 0, s>>>(Q, M, PF, g, ph, d);
  switch (conv) {
    case OSBL_CONV_WENO_JS: launch_flux_c<ND, OSBL_CONV_WENO_JS>(avg, Q, M, PF, F, g, ph, d, s); break;
    case OSBL_CONV_WENO_Z: launch_flux_c<ND, OSBL_CONV_WENO_Z>(avg, Q, M, PF, F, g, ph, d, s); break;
    case OSBL_CONV_TENO5: launch_flux_c<ND, OSBL_CONV_TENO5>(avg, Q, M, PF, F, g, ph, d, s); break;
    default: launch_flux_c<ND, OSBL_CONV_TENO6>(avg, Q, M, PF, F, g, ph, d, s); break;
  }
}

}  // namespace

// -(1/detJ) sum_d (F_d[i] - F_d[i-1])/Delta_d of direction d into R (assigned
// at d = 0, divided by detJ at the last direction); PF (nd + 4 padded fields)
// and F (nd + 2) are scratch bundles.  Q needs filled halos (4 points), M the metric with
// its periodic halos.
void launch_curv_direction(int nd, int d, int conv, int avg, const double* Q, const double* M, double* PF, double* F,
                           double* R, const Grid& g, const Phys& ph, double delta, hipStream_t s) {
  if (nd == 3) launch_flux_nd<3>(conv, avg, Q, M, PF, F, g, ph, d, s);
  else launch_flux_nd<2>(conv, avg, Q, M, PF, F, g, ph, d, s);
  const unsigned grid = (unsigned)(((long long)g.n[0] * g.n[1] * g.n[2] + 255) / 256);
  const int first = d == 0, last = d == nd - 1;
  if (nd == 3) k_curv_diff<3><<<grid, 256, 0, s>>>(F, M, R, g, delta, d, first, last);
  else k_curv_diff<2><<<grid, 256, 0, s>>>(F, M, R, g, delta, d, first, last);
}

// metric over the interior from the coordinates X (extents e[3]: n[d] + 4 on
// the block's directions, 1 elsewhere), then its periodic halos
void launch_curv_metric(int nd, const double* X, double* M, const Grid& g, const Phys& ph, const int e[3],
                        hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  if (nd == 3) k_curv_metric<3><<<grid, 256, 0, s>>>(X, M, g, ph, e[0], e[1], e[2]);
  else k_curv_metric<2><<<grid, 256, 0, s>>>(X, M, g, ph, e[0], e[1], e[2]);
  launch_halo_periodic(M, nd * nd + 1, g, s);
}

}  // namespace osbl
