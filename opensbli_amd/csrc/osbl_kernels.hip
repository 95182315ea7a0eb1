// osbl_kernels.hip -- gfx950 kernels of one RK stage of the OpenSBLI RHS:
//   * boundary fills (periodic wrap, zeroth-order extrapolation),
//   * constituent relations for the central/viscous stencils,
//   * one characteristic LLF sweep per direction (WENO5-JS/Z, TENO5/6),
//   * the fused central (skew or conservative) + viscous + RK update.
#include "osbl_kernels.h"

namespace osbl {

// ---------------------------------------------------------------------------
// Periodic boundary (periodic.py:42-56): every halo point of the shell takes
// the value of its periodic image, corners and edges included, in one pass.
// The shell is enumerated as z-slabs (full x,y extent), then y-slabs of the
// interior z range, then x-slabs of the interior y,z range.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_halo_periodic(double* __restrict__ Q, int nv, Grid g, long long total) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int hx = g.h[0], hy = g.h[1], hz = g.h[2];
  const int PX = g.n[0] + 2 * hx, PY = g.n[1] + 2 * hy;
  const long long nzs = 2LL * hz * PY * PX;
  const long long nys = 2LL * hy * g.n[2] * PX;
  int i, j, k;
  if (t < nzs) {
    const int ii = (int)(t % PX);
    const long long r = t / PX;
    const int jj = (int)(r % PY);
    const int kk = (int)(r / PY);
    i = ii - hx;
    j = jj - hy;
    k = kk < hz ? kk - hz : g.n[2] + (kk - hz);
  } else if (t < nzs + nys) {
    const long long u = t - nzs;
    const int ii = (int)(u % PX);
    const long long r = u / PX;
    const int jj = (int)(r % (2 * hy));
    k = (int)(r / (2 * hy));
    i = ii - hx;
    j = jj < hy ? jj - hy : g.n[1] + (jj - hy);
  } else {
    const long long u = t - nzs - nys;
    const int ii = (int)(u % (2 * hx));
    const long long r = u / (2 * hx);
    j = (int)(r % g.n[1]);
    k = (int)(r / g.n[1]);
    i = ii < hx ? ii - hx : g.n[0] + (ii - hx);
  }
  auto wrap = [](int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); };
  const int dst = g.idx(i, j, k);
  const int src = g.idx(wrap(i, g.n[0]), wrap(j, g.n[1]), wrap(k, g.n[2]));
  for (int m = 0; m < nv; ++m) Q[m * g.fs + dst] = Q[m * g.fs + src];
}

// Zeroth-order extrapolation (extrapolation.py:29-58): on the boundary plane
// of direction d and side s (the other directions' halos included), the
// boundary point and its halo points take the value of the first point inside.
__global__ void __launch_bounds__(256) k_extrap0(double* __restrict__ Q, int nv, Grid g, int d, int side) {
  const int e1 = (d == 0) ? 1 : 0, e2 = (d == 2) ? 1 : 2;
  const int n1 = g.n[e1] + 2 * g.h[e1], n2 = g.n[e2] + 2 * g.h[e2];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n1 * n2) return;
  int c[3];
  c[e1] = t % n1 - g.h[e1];
  c[e2] = t / n1 - g.h[e2];
  const int N = g.n[d];
  const int b = side == 0 ? 0 : N - 1;
  const int from = side == 0 ? 1 : N - 2;
  const int dir = side == 0 ? -1 : 1;
  c[d] = from;
  const int src = g.idx(c[0], c[1], c[2]);
  for (int m = 0; m < nv; ++m) {
    const double v = Q[m * g.fs + src];
    for (int hh = 0; hh <= g.h[d]; ++hh) {
      c[d] = b + dir * hh;
      Q[m * g.fs + g.idx(c[0], c[1], c[2])] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Constituent relations over the padded domain: u_m, p, T
// (taylor_green_vortex.py:29-31).  Read by the central and viscous stencils.
// ---------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(256) k_prims(const double* __restrict__ Q, double* __restrict__ P, Grid g, Phys ph) {
  const int PX = g.n[0] + 2 * g.h[0];
  const int PY = g.n[1] + 2 * g.h[1];
  const int PZ = g.n[2] + 2 * g.h[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x - g.h[0];
  const int j = blockIdx.y - g.h[1];
  const int k = blockIdx.z - g.h[2];
  if (i >= g.n[0] + g.h[0]) return;
  (void)PX; (void)PY; (void)PZ;
  const int x = g.idx(i, j, k);
  const double rho = Q[x];
  const double ir = recip_pos(rho);
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double u = Q[(1 + m) * g.fs + x] * ir;
    P[m * g.fs + x] = u;
    ke += u * u;
  }
  const double p = ph.gm1 * (Q[(ND + 1) * g.fs + x] - 0.5 * rho * ke);
  P[ND * g.fs + x] = p;
  P[(ND + 1) * g.fs + x] = p * ph.gama * ph.Minf * ph.Minf * ir;
}

// ---------------------------------------------------------------------------
// Characteristic sweep along direction D.  Each thread owns a segment of
// `seg` cells of one grid line and slides a 6-point window along it: every
// point's constituent relations are evaluated once, every interface flux
// once, and the residual of a cell is the difference of its two interface
// fluxes (shock_capturing.py:21-34).  Threads of a wave are adjacent in x,
// except for D = 0, where they are adjacent lines.
//
// RKM >= 0: the sweep is the stage's last contribution to the residual, so a
// cell's complete residual R + (this direction's difference) goes straight
// into the RK update (RKM = 0: 3-register stage, 1: low-storage stage, as in
// k_update) and R is not written back.
// ---------------------------------------------------------------------------
struct RkArgs {
  double* Qo;
  double* Qn;
  double cA, cB;
  int first;
};

template <int ND, int D, int CONV, int AVG, bool ACC, int RKM = -1>
__global__ void __launch_bounds__(256) k_char_sweep(const double* __restrict__ Q, double* __restrict__ R, Grid g,
                                                     Phys ph, int seg, long long total,
                                                     const double* __restrict__ met,
                                                     const double* __restrict__ theta, RkArgs rk) {
  constexpr int NV = ND + 2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  int c[3];
  int s;
  if constexpr (D == 0) {
    c[1] = (int)(t % g.n[1]);
    const long long r = t / g.n[1];
    c[2] = (int)(r % g.n[2]);
    s = (int)(r / g.n[2]);
  } else {
    constexpr int O = (D == 1) ? 2 : 1;
    c[0] = (int)(t % g.n[0]);
    const long long r = t / g.n[0];
    c[O] = (int)(r % g.n[O]);
    s = (int)(r / g.n[O]);
  }
  const int c0 = s * seg;
  const int c1 = min(c0 + seg, g.n[D]);
  c[D] = c0 - 3;
  const int st = g.stride(D);
  int off = g.idx(c[0], c[1], c[2]);
  const double inv = ph.inv_delta[D];

  Pt<ND> w[6];
#pragma unroll
  for (int p = 0; p < 5; ++p) load_pt<ND>(w[p], Q, g.fs, off + p * st, ph);
  // the next point's raw values are fetched one iteration ahead
  double nxt[NV];
  fetch_q<ND>(nxt, Q, g.fs, off + 5 * st);
  double fprev[NV];
  // Roe factors of the interface's two points, carried along the window
  [[maybe_unused]] RoeF rf{};
  if constexpr (AVG == OSBL_AVG_ROE) roe_factors<ND>(w[2], rf.sL, rf.hL);
  for (int f = c0 - 1; f < c1; ++f) {
#pragma unroll
    for (int m = 0; m < NV; ++m) w[5].q[m] = nxt[m];
    finish_pt<ND>(w[5], ph);
    if (f + 1 < c1) fetch_q<ND>(nxt, Q, g.fs, off + 6 * st);
    double fh[NV];
    // the residual (and RK register) of this iteration's cell are fetched
    // before the flux, so that the loads overlap its arithmetic instead of
    // stalling the update after it -- except for TENO6, whose sweep would
    // then need more registers than two waves per SIMD leave it
    constexpr bool EARLY = CONV != OSBL_CONV_TENO6;
    [[maybe_unused]] double rin[NV], qoin[NV];
    if constexpr ((RKM >= 0 || ACC) && EARLY) {
      if (f >= c0) {
        const int x = off + 2 * st;
#pragma unroll
        for (int m = 0; m < NV; ++m) {
          rin[m] = ld_once(&R[m * g.fs + x], ph.nt);
          if constexpr (RKM >= 0) qoin[m] = (RKM == 0 && rk.first) ? 0.0 : ld_once(&rk.Qo[m * g.fs + x], ph.nt);
        }
      }
    }
    // interface between window points 2 and 3: its left point is at off + 2 st
    const double ct = theta ? adaptive_ct(theta[off + 2 * st], ph) : ph.teno_ct;
    if constexpr (AVG == OSBL_AVG_ROE) roe_factors<ND>(w[3], rf.sR, rf.hR);
    char_flux<ND, D, CONV, AVG, true>(w, ph, fh, ct, rf);
    if constexpr (AVG == OSBL_AVG_ROE) {
      rf.sL = rf.sR;
      rf.hL = rf.hR;
    }
    if (f >= c0) {
      const int x = off + 2 * st;
      // a stretched direction scales the flux difference by D_dd (opensblifunctions.py:607-616)
      const double gm = met ? met[f] : 1.0;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double v = met ? -(gm * ((fh[m] - fprev[m]) * inv)) : -(fh[m] - fprev[m]) * inv;
        if constexpr (RKM == 0) {
          const double r = (EARLY ? rin[m] : R[m * g.fs + x]) + v;
          const double qo = rk.first ? w[2].q[m] : (EARLY ? qoin[m] : rk.Qo[m * g.fs + x]);
          st_once(qo + ph.dt * rk.cA * r, &rk.Qn[m * g.fs + x], ph.nt);
          st_once(qo + ph.dt * rk.cB * r, &rk.Qo[m * g.fs + x], ph.nt);
        } else if constexpr (RKM == 1) {
          const double r = (EARLY ? rin[m] : R[m * g.fs + x]) + v;
          const double tmp = (rk.cA == 0.0 ? 0.0 : rk.cA * (EARLY ? qoin[m] : rk.Qo[m * g.fs + x])) + ph.dt * r;
          st_once(tmp, &rk.Qo[m * g.fs + x], ph.nt);
          st_once(w[2].q[m] + rk.cB * tmp, &rk.Qn[m * g.fs + x], ph.nt);
        } else if constexpr (ACC) {
          st_once((EARLY ? rin[m] : R[m * g.fs + x]) + v, &R[m * g.fs + x], ph.nt);
        } else {
          st_once(v, &R[m * g.fs + x], ph.nt);
        }
      }
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) fprev[m] = fh[m];
#pragma unroll
    for (int p = 0; p < 5; ++p) w[p] = w[p + 1];
    off += st;
  }
}

// ---------------------------------------------------------------------------
// Characteristic sweep along x with the lanes of a wave along x.  A wave owns
// one grid line (one (y, z) row) and walks it in chunks of 64 cells: the
// chunk's 69 points (x0-2 .. x0+66) are loaded with coalesced reads, their
// constituent relations computed once and kept in LDS, then lane l evaluates
// the flux at interface x0+l+1/2 from LDS and the residual of
// cell x0+l is its difference with lane l-1's flux (a shuffle; lane 0 takes
// the previous chunk's last flux).  Stores are coalesced 512-byte rows.  On a
// periodic line the flux left of cell 0 is the line's last interface flux, so
// cell 0 is finished after the walk; otherwise it is evaluated first.
// ---------------------------------------------------------------------------
constexpr int XW = 4;             // waves (lines) per block
constexpr int XP = 64 + 5;        // points staged per chunk

// (at least three waves per SIMD: TENO6 needs ~178 VGPRs otherwise and
// would run two; held to 168 it spills four and is 6 % faster)
template <int ND, int CONV, int AVG, bool ACC>
__global__ void __launch_bounds__(64 * XW, 3) k_char_sweep_xl(const double* __restrict__ Q, double* __restrict__ R,
                                                          Grid g, Phys ph, int nrows, int periodic,
                                                          const double* __restrict__ met,
                                                          const double* __restrict__ theta) {
  constexpr int NV = ND + 2;
  constexpr bool ROE = AVG == OSBL_AVG_ROE;
  constexpr int NF = NV + ND + 2 + (ROE ? 2 : 0);  // q, u, p, a (, sqrt(rho), sqrt(rho) H)
  __shared__ double sh[XW][NF][XP];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int row = blockIdx.x * XW + wv;
  if (row >= nrows) return;  // whole waves only: no block-wide barrier below
  double (*S)[XP] = sh[wv];
  const int base = g.idx(0, row % g.n[1], row / g.n[1]);
  const int N = g.n[0];
  const double inv = ph.inv_delta[0];
  auto stage_point = [&](int slot, int x) {
    Pt<ND> p;
    load_pt<ND>(p, Q, g.fs, base + x, ph);
#pragma unroll
    for (int m = 0; m < NV; ++m) S[m][slot] = p.q[m];
#pragma unroll
    for (int m = 0; m < ND; ++m) S[NV + m][slot] = p.u[m];
    S[NV + ND][slot] = p.p;
    S[NV + ND + 1][slot] = p.a;
    if constexpr (ROE) roe_factors<ND>(p, S[NV + ND + 2][slot], S[NV + ND + 3][slot]);
  };
  // flux at interface x0 + lane + 1/2 (points x0+lane-2 .. x0+lane+3 = slots lane .. lane+5)
  const int xmax = N + g.h[0] - 1;  // last stored point of the line
  auto flux = [&](double fh[NV], int xi) {
    const LdsWindow<ND> w{&S[0][lane], XP};
    const double ct = theta ? adaptive_ct(theta[base + min(xi, xmax)], ph) : ph.teno_ct;
    [[maybe_unused]] RoeF rf{};
    if constexpr (ROE) {
      rf.sL = S[NV + ND + 2][lane + 2];
      rf.hL = S[NV + ND + 3][lane + 2];
      rf.sR = S[NV + ND + 2][lane + 3];
      rf.hR = S[NV + ND + 3][lane + 3];
    }
    char_flux_w<ND, 0, CONV, AVG, true>(w, ph, fh, ct, rf);
  };
  auto stage_chunk = [&](int x0) {
    // slots 0..68 hold points x0-2 .. x0+66 (clamped to the stored line: the
    // clamped slots only feed lanes whose cells lie beyond the line)
    __builtin_amdgcn_wave_barrier();
    stage_point(lane, min(x0 - 2 + lane, xmax));
    if (lane < XP - 64) stage_point(64 + lane, min(x0 + 62 + lane, xmax));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  auto emit = [&](int x, const double* r) {
    const double gm = met ? met[x] : 1.0;
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const double v = met ? -(gm * (r[m] * inv)) : -r[m] * inv;
      if constexpr (ACC) st_once(ld_once(&R[m * g.fs + base + x], ph.nt) + v, &R[m * g.fs + base + x], ph.nt);
      else st_once(v, &R[m * g.fs + base + x], ph.nt);
    }
  };
  if constexpr (CONV == OSBL_CONV_WENO_Z || CONV == OSBL_CONV_WENO_JS) {
    // Pipelined walk (the WENO sweeps have the registers for it): slot s
    // holds point x0 - 7 + s, lane l evaluates the flux at interface
    // I = x0 - 5 + l (points I-2 .. I+3 = slots l .. l+5) and finishes cell
    // I.  Slots 0..4 of a chunk are slots 64..68 of the one before, so each
    // chunk stages one new point per lane, whose raw values are fetched while
    // the previous chunk evaluates its fluxes.  Periodic lines start at I = 0
    // (the flux left of cell 0 is the line's last one, used after the walk),
    // others at I = -1.
    auto clampx = [&](int x) { return max(-g.h[0], min(x, xmax)); };
    auto stage_q = [&](int slot, const double (&q)[NV]) {
      Pt<ND> p;
#pragma unroll
      for (int m = 0; m < NV; ++m) p.q[m] = q[m];
      finish_pt<ND>(p, ph);
#pragma unroll
      for (int m = 0; m < NV; ++m) S[m][slot] = p.q[m];
#pragma unroll
      for (int m = 0; m < ND; ++m) S[NV + m][slot] = p.u[m];
      S[NV + ND][slot] = p.p;
      S[NV + ND + 1][slot] = p.a;
      if constexpr (ROE) roe_factors<ND>(p, S[NV + ND + 2][slot], S[NV + ND + 3][slot]);
    };
    auto wave_sync = [&] {
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto flux_at = [&](double fh[NV]) {
      const LdsWindow<ND> w{&S[0][lane], XP};
      [[maybe_unused]] RoeF rf{};
      if constexpr (ROE) {
        rf.sL = S[NV + ND + 2][lane + 2];
        rf.hL = S[NV + ND + 3][lane + 2];
        rf.sR = S[NV + ND + 2][lane + 3];
        rf.hR = S[NV + ND + 3][lane + 3];
      }
      char_flux_w<ND, 0, CONV, AVG, true>(w, ph, fh, ph.teno_ct, rf);
    };
    const int x0s = periodic ? 5 : 4;
    double nq[NV];
    if (lane < 5) {
      double q0[NV];
      fetch_q<ND>(q0, Q, g.fs, base + clampx(x0s - 7 + lane));
      stage_q(lane, q0);
    }
    fetch_q<ND>(nq, Q, g.fs, base + clampx(x0s - 2 + lane));
    double prev[NV], first[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      prev[m] = 0.0;
      first[m] = 0.0;
    }
    for (int x0 = x0s; x0 - 5 <= N - 1; x0 += 64) {
      stage_q(5 + lane, nq);
      wave_sync();
      // the next chunk's point, in flight while this chunk's fluxes are evaluated
      fetch_q<ND>(nq, Q, g.fs, base + clampx(x0 + 62 + lane));
      double fh[NV];
      const int I = x0 - 5 + lane;
      flux_at(fh);
      double left[NV];
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double up = __shfl_up(fh[m], 1);
        left[m] = lane == 0 ? prev[m] : up;
      }
      const int last = min(63, N - 1 - (x0 - 5));  // lane holding the line's last interface of the chunk
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        if (x0 == x0s) first[m] = __shfl(fh[m], 0);
        prev[m] = __shfl(fh[m], last);
      }
      if (I >= 0 && I < N && !(periodic && I == 0)) {
        double r[NV];
#pragma unroll
        for (int m = 0; m < NV; ++m) r[m] = fh[m] - left[m];
        emit(I, r);
      }
      // slots 64..68 become slots 0..4 of the next chunk
      wave_sync();
      if (lane < 5) {
#pragma unroll
        for (int e = 0; e < NF; ++e) S[e][lane] = S[e][64 + lane];
      }
      wave_sync();
    }
    if (periodic && lane == 0) {
      // the line's last interface flux is the flux left of cell 0
      double r[NV];
#pragma unroll
      for (int m = 0; m < NV; ++m) r[m] = first[m] - prev[m];
      emit(0, r);
    }
  } else {
    double prev[NV];  // flux left of the chunk's first cell (valid in every lane)
    double first[NV];
    if (!periodic) {
      // flux at interface -1/2 (points -3..2): lane 0 of a chunk staged at x0 = -1
      stage_chunk(-1);
      double fh[NV];
      flux(fh, -1 + lane);
  #pragma unroll
      for (int m = 0; m < NV; ++m) prev[m] = __shfl(fh[m], 0);
    }
    for (int x0 = 0; x0 < N; x0 += 64) {
      stage_chunk(x0);
      double fh[NV];
      flux(fh, x0 + lane);
      const int x = x0 + lane;
      double left[NV];
  #pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double up = __shfl_up(fh[m], 1);
        left[m] = lane == 0 ? prev[m] : up;
      }
      const int last = min(63, N - 1 - x0);  // lane holding the chunk's last valid interface
  #pragma unroll
      for (int m = 0; m < NV; ++m) {
        if (x0 == 0) first[m] = fh[m];
        prev[m] = __shfl(fh[m], last);
      }
      if (x < N && !(periodic && x == 0)) {
        double r[NV];
  #pragma unroll
        for (int m = 0; m < NV; ++m) r[m] = fh[m] - left[m];
        emit(x, r);
      }
    }
    if (periodic && lane == 0) {
      // the line's last interface flux is the flux left of cell 0
      double r[NV];
  #pragma unroll
      for (int m = 0; m < NV; ++m) r[m] = first[m] - prev[m];
      emit(0, r);
    }
  }
}

// ---------------------------------------------------------------------------
// Central fourth-order stencils (scheme.py:81-85): D1 weights
// [1/12, -2/3, 0, 2/3, -1/12], D2 weights [-1/12, 4/3, -5/2, 4/3, -1/12].
// ---------------------------------------------------------------------------
__device__ __forceinline__ double d1(const double* __restrict__ f, int x, int st) {
  return (1.0 / 12.0) * (f[x - 2 * st] - f[x + 2 * st]) + (2.0 / 3.0) * (f[x + st] - f[x - st]);
}
__device__ __forceinline__ double d2(const double* __restrict__ f, int x, int st) {
  return (-1.0 / 12.0) * (f[x - 2 * st] + f[x + 2 * st]) + (4.0 / 3.0) * (f[x - st] + f[x + st]) - 2.5 * f[x];
}

// Viscous terms of the constant-viscosity stress tensor and heat flux
// (taylor_green_vortex.py:14-19), expanded by the product rule as the
// reference's Der.differentiate does (parsing.py:141-186): second derivatives
// with the D2 stencil, mixed derivatives as D_b(D_a u) with a < b
// (opensblifunctions.py:341-353, scheme.py:360-387).  P holds u_0..u_{ND-1}, p, T.
template <int ND>
__device__ __forceinline__ void viscous_terms(const double* __restrict__ P, long long fs, int x, const int st[3],
                                              const Phys& ph, double vr[ND + 2]) {
  const double* U[3] = {P, P + fs, P + 2 * fs};
  const double* T = P + (ND + 1) * fs;
  double du[ND][ND], dd[ND][ND][ND];  // du[m][j] = D_j u_m; dd[m][a][b] = D_a D_b u_m (symmetric use)
#pragma unroll
  for (int m = 0; m < ND; ++m)
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      du[m][j] = ph.inv_delta[j] * d1(U[m], x, st[j]);
      dd[m][j][j] = ph.inv_delta[j] * ph.inv_delta[j] * d2(U[m], x, st[j]);
    }
  // mixed derivatives needed: u_a and u_b for every pair a < b
#pragma unroll
  for (int a = 0; a < ND; ++a)
#pragma unroll
    for (int b = a + 1; b < ND; ++b)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const int m = q == 0 ? a : b;
        const double* f = U[m];
        const int sa = st[a], sb = st[b];
        const double im2 = ph.inv_delta[a] * d1(f, x - 2 * sb, sa);
        const double im1 = ph.inv_delta[a] * d1(f, x - sb, sa);
        const double ip1 = ph.inv_delta[a] * d1(f, x + sb, sa);
        const double ip2 = ph.inv_delta[a] * d1(f, x + 2 * sb, sa);
        const double v = ph.inv_delta[b] * ((1.0 / 12.0) * (im2 - ip2) + (2.0 / 3.0) * (ip1 - im1));
        dd[m][a][b] = v;
        dd[m][b][a] = v;
      }
  double divu = 0.0;
#pragma unroll
  for (int k = 0; k < ND; ++k) divu += du[k][k];
  double energy = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) energy += ph.cq * ph.inv_delta[j] * ph.inv_delta[j] * d2(T, x, st[j]);
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double um = U[m][x];
    double sk = 0.0;
#pragma unroll
    for (int k = 0; k < ND; ++k) sk += dd[k][m][k];
    double visc = 0.0;
#pragma unroll
    for (int j = 0; j < ND; ++j) {
      double tt = dd[m][j][j] + dd[j][m][j];
      if (j == m) tt -= (2.0 / 3.0) * sk;
      const double dtau = ph.inv_re * tt;
      const double tau = ph.inv_re * (du[m][j] + du[j][m] - (j == m ? (2.0 / 3.0) * divu : 0.0));
      visc += dtau;
      energy += du[m][j] * tau + um * dtau;
    }
    vr[1 + m] += visc;
  }
  vr[ND + 1] += energy;
}

// Central convective terms.  Skew/Blaisdell form (parsing.py:75-111,
// scheme.py:187-252) or the conservative flux form (parsing.py:189-216).
template <int ND, int CONV>
__device__ __forceinline__ void central_convective(const double* __restrict__ Q, const double* __restrict__ P,
                                                   long long fs, int x, const int st[3], const Phys& ph,
                                                   double r[ND + 2]) {
  constexpr int NV = ND + 2;
  const double* U[3] = {P, P + fs, P + 2 * fs};
  const double* Pp = P + ND * fs;
#pragma unroll
  for (int m = 0; m < NV; ++m) r[m] = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    const int sj = st[j];
    const double inv = ph.inv_delta[j];
    constexpr double W1A = 1.0 / 12.0, W1B = 2.0 / 3.0;
    // points -2, -1, +1, +2 with weights  W1A, -W1B, W1B, -W1A
    const int offs[4] = {x - 2 * sj, x - sj, x + sj, x + 2 * sj};
    const double wts[4] = {W1A, -W1B, W1B, -W1A};
    if constexpr (CONV == OSBL_CONV_CENTRAL_SKEW) {
      double d_ru[NV], d_q[NV], d_uj = 0.0, d_p = 0.0, d_puj = 0.0;
#pragma unroll
      for (int m = 0; m < NV; ++m) { d_ru[m] = 0.0; d_q[m] = 0.0; }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int y = offs[q];
        const double w = wts[q];
        const double uj = U[j][y];
        const double pp = Pp[y];
#pragma unroll
        for (int m = 0; m < NV; ++m) {
          const double qm = Q[m * fs + y];
          d_ru[m] += w * (qm * uj);
          d_q[m] += w * qm;
        }
        d_uj += w * uj;
        d_p += w * pp;
        d_puj += w * (pp * uj);
      }
      const double uj = U[j][x];
      const double duj = inv * d_uj;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double qm = Q[m * fs + x];
        double t = -0.5 * (inv * d_ru[m] + uj * (inv * d_q[m]) + qm * duj);
        if (m == 1 + j) t -= inv * d_p;
        if (m == NV - 1) t -= inv * d_puj;
        r[m] += t;
      }
    } else {
      double dF[NV];
#pragma unroll
      for (int m = 0; m < NV; ++m) dF[m] = 0.0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int y = offs[q];
        const double w = wts[q];
        const double uj = U[j][y];
        const double pp = Pp[y];
        dF[0] += w * Q[(1 + j) * fs + y];
#pragma unroll
        for (int m = 0; m < ND; ++m) dF[1 + m] += w * (Q[(1 + m) * fs + y] * uj + (m == j ? pp : 0.0));
        dF[NV - 1] += w * ((pp + Q[(NV - 1) * fs + y]) * uj);
      }
#pragma unroll
      for (int m = 0; m < NV; ++m) r[m] -= inv * dF[m];
    }
  }
}

// ---------------------------------------------------------------------------
// Residual assembly + RK update for one interior cell.
//   MODE 0: 3-register RK3 stage (rk_sbli.py:117-124): Qn = Qo + dt*cA*R,
//           Qo += dt*cB*R; at the first stage Qo is read from Q (the save of
//           rk_sbli.py:126-133 fused into the stage).
//   MODE 1: low-storage stage (rk_LS.py:155-166): Qo = cA*Qo + dt*R,
//           Qn = Q + cB*Qo (Qo not read when cA == 0).
//   MODE 2: residual only (written to R).
//   MODE 3 (k_visc_update3d only): the viscous terms alone, assigned to R.
// For characteristic schemes the convective residual is already in R.
// ---------------------------------------------------------------------------
template <int ND, int CONV, bool VISC, int MODE>
__global__ void __launch_bounds__(256) k_update(const double* __restrict__ Q, const double* __restrict__ P,
                                                double* __restrict__ R, double* __restrict__ Qo,
                                                double* __restrict__ Qn, Grid g, Phys ph, double cA, double cB,
                                                int first) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j = blockIdx.y, k = blockIdx.z;
  const int x = g.idx(i, j, k);
  const int st[3] = {1, g.pitch, g.plane};
  double r[NV];
  if constexpr (CONV == OSBL_CONV_CENTRAL_SKEW || CONV == OSBL_CONV_CENTRAL_CONS) {
    central_convective<ND, CONV>(Q, P, g.fs, x, st, ph, r);
  } else {
#pragma unroll
    for (int m = 0; m < NV; ++m) r[m] = R[m * g.fs + x];
  }
  if constexpr (VISC) viscous_terms<ND>(P, g.fs, x, st, ph, r);
  if constexpr (MODE == 2) {
#pragma unroll
    for (int m = 0; m < NV; ++m) R[m * g.fs + x] = r[m];
  } else if constexpr (MODE == 0) {
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
// 3-D viscous terms + RK update for the characteristic schemes, tiled 2.5-D.
// A VTX x VTY block of (x, y) columns marches along z.  Each z plane of the
// constituent relations u_0..u_2, T (computed here from Q, with a 2-point
// x/y halo) is staged in one of four rotating LDS slots; the x/y stencils
// read LDS, the z stencils read a per-thread register queue of the five
// planes z-2..z+2 holding u, T and the inner derivatives D_x u_0, D_x u_2,
// D_y u_1, D_y u_2 of the mixed terms D_z(D_x u), D_z(D_y u).  The formulas
// are those of viscous_terms (scheme.py:299-387).
//
// Software pipeline: while plane z is evaluated, the raw Q of plane z+3 and
// the cell's residual / RK registers are already in flight (loaded at the
// top of the iteration into registers); plane z+3 then goes into the slot
// plane z-1 vacated, and one barrier per plane separates those writes from
// the next plane's reads.
// ---------------------------------------------------------------------------
constexpr int VTX = 64, VTY = 8, VPX = VTX + 4, VPY = VTY + 4, VPL = VPX * VPY;
constexpr int VNT = VTX * VTY;                 // threads per block
constexpr int VPP = (VPL + VNT - 1) / VNT;     // staged points per thread and plane

template <int MODE>
__global__ void __launch_bounds__(VNT, 2) k_visc_update3d(const double* __restrict__ Q, double* __restrict__ R,
                                                       double* __restrict__ Qo, double* __restrict__ Qn, Grid g,
                                                       Phys ph, double cA, double cB, int first, int sz) {
  __shared__ double sp[4][4][VPL];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * VTX + tx;
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int x0 = bx * VTX, y0 = by * VTY;
  const int z0 = bz * sz;
  const int z1 = min(z0 + sz, g.n[2]);
  const int x = x0 + tx, y = y0 + ty;
  const bool inside = x < g.n[0] && y < g.n[1];
  const double ix = ph.inv_delta[0], iy = ph.inv_delta[1], iz = ph.inv_delta[2];

  // this thread's staged points: slot index and in-plane offset (plane z = 0)
  int sidx[VPP], soff[VPP];
  bool sval[VPP];
#pragma unroll
  for (int k = 0; k < VPP; ++k) {
    const int p = tid + k * VNT;
    const int xx = p % VPX, yy = p / VPX;
    const int gx = x0 - 2 + xx, gy = y0 - 2 + yy;
    sidx[k] = p;
    sval[k] = p < VPL && gx < g.n[0] + g.h[0] && gy < g.n[1] + g.h[1];
    soff[k] = sval[k] ? g.idx(gx, gy, 0) : 0;
  }
  double pq[VPP][5];  // raw Q of the plane in flight
  auto fetch = [&](int zp) {
    const int zo = zp * g.plane;
#pragma unroll
    for (int k = 0; k < VPP; ++k)
      if (sval[k]) {
#pragma unroll
        for (int m = 0; m < 5; ++m) pq[k][m] = Q[m * g.fs + soff[k] + zo];
      }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int k = 0; k < VPP; ++k)
      if (sval[k]) {
        const double rho = pq[k][0];
        const double ir = recip_pos(rho);
        const double u0 = pq[k][1] * ir, u1 = pq[k][2] * ir, u2 = pq[k][3] * ir;
        const double pr = ph.gm1 * (pq[k][4] - 0.5 * rho * (u0 * u0 + u1 * u1 + u2 * u2));
        const int p = sidx[k];
        sp[slot][0][p] = u0;
        sp[slot][1][p] = u1;
        sp[slot][2][p] = u2;
        sp[slot][3][p] = pr * ph.gama * ph.Minf * ph.Minf * ir;
      }
  };
  const int c = (ty + 2) * VPX + tx + 2;  // this thread's point in a slot
  auto dx1 = [&](const double* f, int o) {
    return (1.0 / 12.0) * (f[o - 2] - f[o + 2]) + (2.0 / 3.0) * (f[o + 1] - f[o - 1]);
  };
  auto dy1 = [&](const double* f, int o) {
    return (1.0 / 12.0) * (f[o - 2 * VPX] - f[o + 2 * VPX]) + (2.0 / 3.0) * (f[o + VPX] - f[o - VPX]);
  };
  auto dx2 = [&](const double* f, int o) {
    return (-1.0 / 12.0) * (f[o - 2] + f[o + 2]) + (4.0 / 3.0) * (f[o - 1] + f[o + 1]) - 2.5 * f[o];
  };
  auto dy2 = [&](const double* f, int o) {
    return (-1.0 / 12.0) * (f[o - 2 * VPX] + f[o + 2 * VPX]) + (4.0 / 3.0) * (f[o - VPX] + f[o + VPX]) - 2.5 * f[o];
  };
  // column quantities of a plane: u0,u1,u2,T, Dx u0, Dx u2, Dy u1, Dy u2
  auto column = [&](int slot, double* cq) {
    const double* u0 = sp[slot][0];
    const double* u1 = sp[slot][1];
    const double* u2 = sp[slot][2];
    cq[0] = u0[c];
    cq[1] = u1[c];
    cq[2] = u2[c];
    cq[3] = sp[slot][3][c];
    cq[4] = ix * dx1(u0, c);
    cq[5] = ix * dx1(u2, c);
    cq[6] = iy * dy1(u1, c);
    cq[7] = iy * dy1(u2, c);
  };
  auto slot_of = [](int zp) { return (zp + 8) & 3; };

  double zq[5][8];
  // prologue: planes z0-2 .. z0+1 into the four slots, their columns into
  // queue entries 0..3, then plane z0+2 into the slot of plane z0-2
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
    const int o = g.idx(x, y, z);
    // the cell's residual and RK registers, in flight during the arithmetic
    double rin[5], qa[5], qb[5];
    if (inside) {
#pragma unroll
      for (int m = 0; m < 5; ++m) {
        if constexpr (MODE == 2) rin[m] = ld_once(&R[m * g.fs + o], ph.nt);
        if constexpr (MODE == 0 || MODE == 1) rin[m] = ld_once(&R[m * g.fs + o], ph.nt);
        if constexpr (MODE == 0) qa[m] = first ? Q[m * g.fs + o] : ld_once(&Qo[m * g.fs + o], ph.nt);
        if constexpr (MODE == 1) {
          qa[m] = cA == 0.0 ? 0.0 : ld_once(&Qo[m * g.fs + o], ph.nt);
          qb[m] = Q[m * g.fs + o];
        }
      }
      const int sl = slot_of(z);
      const double* U[3] = {sp[sl][0], sp[sl][1], sp[sl][2]};
      const double* T = sp[sl][3];
      double du[3][3], dd[3][3][3];
      // x and y first / second derivatives from LDS
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        du[m][0] = ix * dx1(U[m], c);
        du[m][1] = iy * dy1(U[m], c);
        dd[m][0][0] = ix * ix * dx2(U[m], c);
        dd[m][1][1] = iy * iy * dy2(U[m], c);
      }
      // z derivatives from the queue (planes z-2..z+2 at indices 0..4)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        du[m][2] = iz * ((1.0 / 12.0) * (zq[0][m] - zq[4][m]) + (2.0 / 3.0) * (zq[3][m] - zq[1][m]));
        dd[m][2][2] =
            iz * iz * ((-1.0 / 12.0) * (zq[0][m] + zq[4][m]) + (4.0 / 3.0) * (zq[1][m] + zq[3][m]) - 2.5 * zq[2][m]);
      }
      // mixed D_y(D_x u_m), m = 0, 1, from LDS
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        const double* f = U[m];
        const double a2 = ix * dx1(f, c - 2 * VPX), a1 = ix * dx1(f, c - VPX);
        const double b1 = ix * dx1(f, c + VPX), b2 = ix * dx1(f, c + 2 * VPX);
        const double v = iy * ((1.0 / 12.0) * (a2 - b2) + (2.0 / 3.0) * (b1 - a1));
        dd[m][0][1] = v;
        dd[m][1][0] = v;
      }
      // mixed D_z(D_x u_m), m = 0 (queue 4), 2 (queue 5); D_z(D_y u_m), m = 1 (6), 2 (7)
      auto dzq = [&](int e) {
        return iz * ((1.0 / 12.0) * (zq[0][e] - zq[4][e]) + (2.0 / 3.0) * (zq[3][e] - zq[1][e]));
      };
      dd[0][0][2] = dd[0][2][0] = dzq(4);
      dd[2][0][2] = dd[2][2][0] = dzq(5);
      dd[1][1][2] = dd[1][2][1] = dzq(6);
      dd[2][1][2] = dd[2][2][1] = dzq(7);
      const double divu = du[0][0] + du[1][1] + du[2][2];
      double r[5];
#pragma unroll
      for (int m = 0; m < 5; ++m) r[m] = 0.0;
      double energy = ph.cq * (ix * ix * dx2(T, c) + iy * iy * dy2(T, c) +
                               iz * iz * ((-1.0 / 12.0) * (zq[0][3] + zq[4][3]) +
                                          (4.0 / 3.0) * (zq[1][3] + zq[3][3]) - 2.5 * zq[2][3]));
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double um = zq[2][m];
        double sk = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) sk += dd[k][m][k];
        double visc = 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          double tt = dd[m][j][j] + dd[j][m][j];
          if (j == m) tt -= (2.0 / 3.0) * sk;
          const double dtau = ph.inv_re * tt;
          const double tau = ph.inv_re * (du[m][j] + du[j][m] - (j == m ? (2.0 / 3.0) * divu : 0.0));
          visc += dtau;
          energy += du[m][j] * tau + um * dtau;
        }
        r[1 + m] = visc;
      }
      r[4] = energy;
      if constexpr (MODE != 3) {
#pragma unroll
        for (int m = 0; m < 5; ++m) r[m] = rin[m] + r[m];
      }
      if constexpr (MODE == 2 || MODE == 3) {
#pragma unroll
        for (int m = 0; m < 5; ++m) st_once(r[m], &R[m * g.fs + o], ph.nt);
      } else if constexpr (MODE == 0) {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          st_once(qa[m] + ph.dt * cA * r[m], &Qn[m * g.fs + o], ph.nt);
          st_once(qa[m] + ph.dt * cB * r[m], &Qo[m * g.fs + o], ph.nt);
        }
      } else {
#pragma unroll
        for (int m = 0; m < 5; ++m) {
          const double tmp = (cA == 0.0 ? 0.0 : cA * qa[m]) + ph.dt * r[m];
          st_once(tmp, &Qo[m * g.fs + o], ph.nt);
          st_once(qb[m] + cB * tmp, &Qn[m * g.fs + o], ph.nt);
        }
      }
    }
    if (more) {
      // plane z+3 replaces plane z-1 (last read before the previous barrier)
      store(slot_of(z + 3));
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int e = 0; e < 8; ++e) zq[k][e] = zq[k + 1][e];
      column(slot_of(z + 3), zq[4]);
    }
  }
}

void launch_visc_update3d(int mode, const double* Q, double* R, double* Qo, double* Qn, const Grid& g, const Phys& ph,
                          double cA, double cB, int first, int sz, hipStream_t s) {
  dim3 grid((g.n[0] + VTX - 1) / VTX, (g.n[1] + VTY - 1) / VTY, (g.n[2] + sz - 1) / sz);
  dim3 block(VTX, VTY);
  if (mode == 0) k_visc_update3d<0><<<grid, block, 0, s>>>(Q, R, Qo, Qn, g, ph, cA, cB, first, sz);
  else if (mode == 1) k_visc_update3d<1><<<grid, block, 0, s>>>(Q, R, Qo, Qn, g, ph, cA, cB, first, sz);
  else if (mode == 2) k_visc_update3d<2><<<grid, block, 0, s>>>(Q, R, Qo, Qn, g, ph, cA, cB, first, sz);
  else k_visc_update3d<3><<<grid, block, 0, s>>>(Q, R, Qo, Qn, g, ph, cA, cB, first, sz);
}

// ---------------------------------------------------------------------------
// 3-D central convective terms, tiled 2.5-D.  Skew-symmetric (Blaisdell) form
// (parsing.py:75-111, scheme.py:187-252) or the conservative flux form
// (parsing.py:189-216), with the formulas of central_convective.  A CTX x CTY
// block of (x, y) columns marches along z.  The plane being evaluated is
// staged in LDS (Q, u_x, u_y, p with a 2-point x/y halo, the x/y stencils);
// each thread keeps a register queue of its own column's Q, u_z, p over the
// five planes z-2..z+2 (the z stencils).  Software pipeline: the raw Q of the
// next plane's staged points and of the column point three planes ahead are
// loaded at the top of an iteration and consumed after its arithmetic; two
// LDS slots alternate, one barrier per plane.
//   MODE 0/1: RK update (no viscous terms); MODE 2: the terms are assigned to
//   R (the viscous kernel then adds its terms and applies the update).
// ---------------------------------------------------------------------------
constexpr int CTX = 64, CTY = 4, CPX = CTX + 4, CPY = CTY + 4, CPL = CPX * CPY;
constexpr int CNT = CTX * CTY;                 // threads per block
constexpr int CPP = (CPL + CNT - 1) / CNT;     // staged points per thread and plane
constexpr int CNF = 8;                         // staged fields: Q0..Q4, u_x, u_y, p

template <int CONV, int MODE>
__global__ void __launch_bounds__(CNT, 2) k_conv3d(const double* __restrict__ Q, double* __restrict__ R,
                                                   double* __restrict__ Qo, double* __restrict__ Qn, Grid g, Phys ph,
                                                   double cA, double cB, int first, int sz) {
  constexpr int NV = 5;
  __shared__ double sp[2][CNF][CPL];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * CTX + tx;
  int bx, by, bz;
  xcd_tile(bx, by, bz);
  const int x0 = bx * CTX, y0 = by * CTY;
  const int z0 = bz * sz;
  const int z1 = min(z0 + sz, g.n[2]);
  const int x = x0 + tx, y = y0 + ty;
  const bool inside = x < g.n[0] && y < g.n[1];

  int sidx[CPP], soff[CPP];
  bool sval[CPP];
#pragma unroll
  for (int k = 0; k < CPP; ++k) {
    const int p = tid + k * CNT;
    const int xx = p % CPX, yy = p / CPX;
    const int gx = x0 - 2 + xx, gy = y0 - 2 + yy;
    sidx[k] = p;
    sval[k] = p < CPL && gx < g.n[0] + g.h[0] && gy < g.n[1] + g.h[1];
    soff[k] = sval[k] ? g.idx(gx, gy, 0) : 0;
  }
  double pq[CPP][NV];  // raw Q of the next plane's staged points
  auto fetch = [&](int zp) {
    const int zo = zp * g.plane;
#pragma unroll
    for (int k = 0; k < CPP; ++k)
      if (sval[k]) {
#pragma unroll
        for (int m = 0; m < NV; ++m) pq[k][m] = Q[m * g.fs + soff[k] + zo];
      }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int k = 0; k < CPP; ++k)
      if (sval[k]) {
        const double rho = pq[k][0];
        const double ir = recip_pos(rho);
        const double u0 = pq[k][1] * ir, u1 = pq[k][2] * ir, u2 = pq[k][3] * ir;
        const double pr = ph.gm1 * (pq[k][4] - 0.5 * rho * (u0 * u0 + u1 * u1 + u2 * u2));
        const int p = sidx[k];
#pragma unroll
        for (int m = 0; m < NV; ++m) sp[slot][m][p] = pq[k][m];
        sp[slot][5][p] = u0;
        sp[slot][6][p] = u1;
        sp[slot][7][p] = pr;
      }
  };
  // this thread's column: raw Q of one plane, and its queue entries Q0..Q4, u_z, p
  const int col = inside ? g.idx(x, y, 0) : 0;
  double oq[NV];
  auto fetch_col = [&](int zp) {
    if (inside) {
      const int o = col + zp * g.plane;
#pragma unroll
      for (int m = 0; m < NV; ++m) oq[m] = Q[m * g.fs + o];
    }
  };
  auto col_entry = [&](double* e) {
    const double rho = oq[0];
    const double ir = recip_pos(rho);
    const double u0 = oq[1] * ir, u1 = oq[2] * ir, u2 = oq[3] * ir;
#pragma unroll
    for (int m = 0; m < NV; ++m) e[m] = oq[m];
    e[5] = u2;
    e[6] = ph.gm1 * (oq[4] - 0.5 * rho * (u0 * u0 + u1 * u1 + u2 * u2));
  };
  const int c = (ty + 2) * CPX + tx + 2;  // this thread's point in a slot

  double zq[5][7];  // planes z-2..z+2 of this column
  // prologue: the queue for z0 (planes z0-2 .. z0+2), plane z0 staged
  for (int k = 0; k < 5; ++k) {
    fetch_col(z0 - 2 + k);
    if (inside) col_entry(zq[k]);
  }
  fetch(z0);
  store(z0 & 1);
  if (z0 + 1 < z1) {  // in flight during the first plane
    fetch(z0 + 1);
    fetch_col(z0 + 3);
  }
  __syncthreads();

  constexpr double W1A = 1.0 / 12.0, W1B = 2.0 / 3.0;
  for (int z = z0; z < z1; ++z) {
    const bool more = z + 1 < z1;
    const int o = g.idx(x, y, z);
    if (inside) {
      // RK registers of the cell, in flight during the arithmetic
      [[maybe_unused]] double qa[NV], qb[NV];
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        if constexpr (MODE == 0) qa[m] = first ? 0.0 : ld_once(&Qo[m * g.fs + o], ph.nt);
        if constexpr (MODE == 1) qa[m] = cA == 0.0 ? 0.0 : ld_once(&Qo[m * g.fs + o], ph.nt);
      }
      const int sl = z & 1;
      double r[NV];
#pragma unroll
      for (int m = 0; m < NV; ++m) r[m] = 0.0;
      // x and y from LDS (stride 1 and CPX), then z from the queue
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double inv = ph.inv_delta[j];
        const double wts[4] = {W1A, -W1B, W1B, -W1A};
        auto qv = [&](int q, int m) -> double {  // Q_m at stencil point q (offsets -2,-1,+1,+2)
          if (j == 2) return zq[q < 2 ? q : q + 1][m];
          const int st = j == 0 ? 1 : CPX;
          const int off = (q < 2 ? q - 2 : q - 1) * st;
          return sp[sl][m][c + off];
        };
        auto uv = [&](int q) -> double {  // u_j at stencil point q
          if (j == 2) return zq[q < 2 ? q : q + 1][5];
          const int st = j == 0 ? 1 : CPX;
          const int off = (q < 2 ? q - 2 : q - 1) * st;
          return sp[sl][5 + j][c + off];
        };
        auto pv = [&](int q) -> double {
          if (j == 2) return zq[q < 2 ? q : q + 1][6];
          const int st = j == 0 ? 1 : CPX;
          const int off = (q < 2 ? q - 2 : q - 1) * st;
          return sp[sl][7][c + off];
        };
        if constexpr (CONV == OSBL_CONV_CENTRAL_SKEW) {
          double d_ru[NV], d_q[NV], d_uj = 0.0, d_p = 0.0, d_puj = 0.0;
#pragma unroll
          for (int m = 0; m < NV; ++m) { d_ru[m] = 0.0; d_q[m] = 0.0; }
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double w = wts[q];
            const double uj = uv(q);
            const double pp = pv(q);
#pragma unroll
            for (int m = 0; m < NV; ++m) {
              const double qm = qv(q, m);
              d_ru[m] += w * (qm * uj);
              d_q[m] += w * qm;
            }
            d_uj += w * uj;
            d_p += w * pp;
            d_puj += w * (pp * uj);
          }
          const double uj = j == 2 ? zq[2][5] : sp[sl][5 + j][c];
          const double duj = inv * d_uj;
#pragma unroll
          for (int m = 0; m < NV; ++m) {
            const double qm = zq[2][m];
            double t = -0.5 * (inv * d_ru[m] + uj * (inv * d_q[m]) + qm * duj);
            if (m == 1 + j) t -= inv * d_p;
            if (m == NV - 1) t -= inv * d_puj;
            r[m] += t;
          }
        } else {
          double dF[NV];
#pragma unroll
          for (int m = 0; m < NV; ++m) dF[m] = 0.0;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const double w = wts[q];
            const double uj = uv(q);
            const double pp = pv(q);
            dF[0] += w * qv(q, 1 + j);
#pragma unroll
            for (int m = 0; m < 3; ++m) dF[1 + m] += w * (qv(q, 1 + m) * uj + (m == j ? pp : 0.0));
            dF[NV - 1] += w * ((pp + qv(q, NV - 1)) * uj);
          }
#pragma unroll
          for (int m = 0; m < NV; ++m) r[m] -= inv * dF[m];
        }
      }
      if constexpr (MODE == 2) {
#pragma unroll
        for (int m = 0; m < NV; ++m) st_once(r[m], &R[m * g.fs + o], ph.nt);
      } else if constexpr (MODE == 0) {
#pragma unroll
        for (int m = 0; m < NV; ++m) {
          const double qo = first ? zq[2][m] : qa[m];
          st_once(qo + ph.dt * cA * r[m], &Qn[m * g.fs + o], ph.nt);
          st_once(qo + ph.dt * cB * r[m], &Qo[m * g.fs + o], ph.nt);
        }
      } else {
#pragma unroll
        for (int m = 0; m < NV; ++m) {
          const double tmp = (cA == 0.0 ? 0.0 : cA * qa[m]) + ph.dt * r[m];
          st_once(tmp, &Qo[m * g.fs + o], ph.nt);
          st_once(zq[2][m] + cB * tmp, &Qn[m * g.fs + o], ph.nt);
        }
      }
    }
    if (more) {
      // plane z+1 goes into the slot plane z-1 left (last read before the
      // previous barrier); the queue moves on by one plane
      store((z + 1) & 1);
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int e = 0; e < 7; ++e) zq[k][e] = zq[k + 1][e];
      if (inside) col_entry(zq[4]);
      if (z + 2 < z1) {  // planes up to z1 + 1 are read
        fetch(z + 2);
        fetch_col(z + 4);
      }
      __syncthreads();
    }
  }
}

// z segment of a 2.5-D tiled kernel over tx x ty columns: `start` planes,
// halved (down to 2) while the launch would have fewer than `want` blocks.
// The viscous kernels (one 512-thread block per CU, bound by their 104-130 KB
// of LDS) take 64-plane segments from 512 blocks on: each segment stages four
// planes beyond its own, and at 256^3 two rounds of 64 planes beat four of 32
// (0.684 -> 0.673 ms for the central viscous + RK kernel).
static int tile_segment(const Grid& g, int tx, int ty, long long want, int start) {
  const long long tiles = (long long)((g.n[0] + tx - 1) / tx) * ((g.n[1] + ty - 1) / ty);
  int sz = start;
  while (sz > 2 && tiles * ((g.n[2] + sz - 1) / sz) < want) sz /= 2;
  return sz;
}
int visc3d_segment(const Grid& g) { return tile_segment(g, VTX, VTY, 512, 64); }
int conv3d_segment(const Grid& g) { return tile_segment(g, CTX, CTY, 1024, 32); }
long long conv3d_blocks(const Grid& g, int sz) {
  return (long long)((g.n[0] + CTX - 1) / CTX) * ((g.n[1] + CTY - 1) / CTY) * ((g.n[2] + sz - 1) / sz);
}

void launch_conv3d(int conv, int mode, const double* Q, double* R, double* Qo, double* Qn, const Grid& g,
                   const Phys& ph, double cA, double cB, int first, int sz, hipStream_t s) {
  dim3 grid((g.n[0] + CTX - 1) / CTX, (g.n[1] + CTY - 1) / CTY, (g.n[2] + sz - 1) / sz);
  dim3 block(CTX, CTY);
#define OSBL_CONV3D(C, M) k_conv3d<C, M><<<grid, block, 0, s>>>(Q, R, Qo, Qn, g, ph, cA, cB, first, sz)
  if (conv == OSBL_CONV_CENTRAL_SKEW) {
    if (mode == 0) OSBL_CONV3D(OSBL_CONV_CENTRAL_SKEW, 0);
    else if (mode == 1) OSBL_CONV3D(OSBL_CONV_CENTRAL_SKEW, 1);
    else OSBL_CONV3D(OSBL_CONV_CENTRAL_SKEW, 2);
  } else {
    if (mode == 0) OSBL_CONV3D(OSBL_CONV_CENTRAL_CONS, 0);
    else if (mode == 1) OSBL_CONV3D(OSBL_CONV_CENTRAL_CONS, 1);
    else OSBL_CONV3D(OSBL_CONV_CENTRAL_CONS, 2);
  }
#undef OSBL_CONV3D
}

// interior <-> contiguous copies
__global__ void __launch_bounds__(256) k_pack(const double* __restrict__ F, double* __restrict__ out, Grid g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j = blockIdx.y, k = blockIdx.z;
  out[(long long)i + (long long)g.n[0] * (j + (long long)g.n[1] * k)] = F[g.idx(i, j, k)];
}
__global__ void __launch_bounds__(256) k_unpack(const double* __restrict__ in, double* __restrict__ F, Grid g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int j = blockIdx.y, k = blockIdx.z;
  F[g.idx(i, j, k)] = in[(long long)i + (long long)g.n[0] * (j + (long long)g.n[1] * k)];
}

// NaN values among the interior values of fields v0 .. v1-1 (OPS's
// ops_NaNcheck): counted per wave with a ballot, a wave adds to the counter
// only when it found one, so a finite state costs no atomics
__global__ void __launch_bounds__(256) k_count_nan(const double* __restrict__ F, Grid g, int v0, int v1,
                                                  unsigned long long* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  unsigned long long found = 0;
  if (i < g.n[0]) {
    const int x = g.idx(i, j, k);
    for (int v = v0; v < v1; ++v) found += __builtin_isnan(F[v * g.fs + x]) ? 1 : 0;
  }
  unsigned long long wave = 0;
  for (int v = 0; v < v1 - v0; ++v) wave += __popcll(__ballot(found > (unsigned long long)v));
  if ((threadIdx.x & 63) == 0 && wave) atomicAdd(count, wave);
}
__global__ void k_clear_count(unsigned long long* count) { *count = 0; }

// ---------------------------------------------------------------------------
// launch helpers (instantiation tables)
// ---------------------------------------------------------------------------
template <int ND, int D, int CONV, int AVG>
static void launch_sweep_t(const double* Q, double* R, const Grid& g, const Phys& ph, int seg, bool acc,
                           const double* met, const double* theta, hipStream_t s) {
  if constexpr (D == 0 && ND >= 2) {
    const int nrows = g.n[1] * g.n[2];
    const unsigned blocks = (unsigned)((nrows + XW - 1) / XW);
    if (acc)
      k_char_sweep_xl<ND, CONV, AVG, true><<<blocks, 64 * XW, 0, s>>>(Q, R, g, ph, nrows, g.periodic_x, met, theta);
    else
      k_char_sweep_xl<ND, CONV, AVG, false><<<blocks, 64 * XW, 0, s>>>(Q, R, g, ph, nrows, g.periodic_x, met, theta);
    return;
  }
  long long lines = 1;
  for (int e = 0; e < 3; ++e)
    if (e != D) lines *= g.n[e];
  seg = fit_segment(lines, g.n[D], seg, 4);
  const long long nseg = (g.n[D] + seg - 1) / seg;
  const long long total = lines * nseg;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  const RkArgs none{nullptr, nullptr, 0.0, 0.0, 0};
  if (acc) k_char_sweep<ND, D, CONV, AVG, true><<<blocks, 256, 0, s>>>(Q, R, g, ph, seg, total, met, theta, none);
  else k_char_sweep<ND, D, CONV, AVG, false><<<blocks, 256, 0, s>>>(Q, R, g, ph, seg, total, met, theta, none);
}

// the last (z) sweep of a 3-D stage fused with the RK update
template <int CONV, int AVG>
static void launch_sweep_rk_t(int mode, const double* Q, double* R, const Grid& g, const Phys& ph, int seg,
                              const double* theta, const RkArgs& rk, hipStream_t s) {
  const long long lines = (long long)g.n[0] * g.n[1];
  seg = fit_segment(lines, g.n[2], seg, 4);
  const long long nseg = (g.n[2] + seg - 1) / seg;
  const long long total = lines * nseg;
  const unsigned blocks = (unsigned)((total + 255) / 256);
  if (mode == 0)
    k_char_sweep<3, 2, CONV, AVG, true, 0><<<blocks, 256, 0, s>>>(Q, R, g, ph, seg, total, nullptr, theta, rk);
  else
    k_char_sweep<3, 2, CONV, AVG, true, 1><<<blocks, 256, 0, s>>>(Q, R, g, ph, seg, total, nullptr, theta, rk);
}

template <int CONV>
static void launch_sweep_rk_avg(int avg, int mode, const double* Q, double* R, const Grid& g, const Phys& ph, int seg,
                                const double* theta, const RkArgs& rk, hipStream_t s) {
  if (avg == OSBL_AVG_ROE) launch_sweep_rk_t<CONV, OSBL_AVG_ROE>(mode, Q, R, g, ph, seg, theta, rk, s);
  else launch_sweep_rk_t<CONV, OSBL_AVG_SIMPLE>(mode, Q, R, g, ph, seg, theta, rk, s);
}

void launch_sweep_z_rk(int conv, int avg, int mode, const double* Q, double* R, double* Qo, double* Qn, const Grid& g,
                       const Phys& ph, int seg, const double* theta, double cA, double cB, int first, hipStream_t s) {
  const RkArgs rk{Qo, Qn, cA, cB, first};
  switch (conv) {
    case OSBL_CONV_WENO_JS: launch_sweep_rk_avg<OSBL_CONV_WENO_JS>(avg, mode, Q, R, g, ph, seg, theta, rk, s); break;
    case OSBL_CONV_WENO_Z: launch_sweep_rk_avg<OSBL_CONV_WENO_Z>(avg, mode, Q, R, g, ph, seg, theta, rk, s); break;
    case OSBL_CONV_TENO5: launch_sweep_rk_avg<OSBL_CONV_TENO5>(avg, mode, Q, R, g, ph, seg, theta, rk, s); break;
    default: launch_sweep_rk_avg<OSBL_CONV_TENO6>(avg, mode, Q, R, g, ph, seg, theta, rk, s); break;
  }
}

template <int ND, int D, int CONV>
static void launch_sweep_avg(int avg, const double* Q, double* R, const Grid& g, const Phys& ph, int seg, bool acc,
                             const double* met, const double* theta, hipStream_t s) {
  if (avg == OSBL_AVG_ROE) launch_sweep_t<ND, D, CONV, OSBL_AVG_ROE>(Q, R, g, ph, seg, acc, met, theta, s);
  else launch_sweep_t<ND, D, CONV, OSBL_AVG_SIMPLE>(Q, R, g, ph, seg, acc, met, theta, s);
}

template <int ND, int D>
static void launch_sweep_conv(int conv, int avg, const double* Q, double* R, const Grid& g, const Phys& ph, int seg,
                              const double* met, const double* theta, bool acc, hipStream_t s) {
  switch (conv) {
    case OSBL_CONV_WENO_JS: launch_sweep_avg<ND, D, OSBL_CONV_WENO_JS>(avg, Q, R, g, ph, seg, acc, met, theta, s); break;
    case OSBL_CONV_WENO_Z: launch_sweep_avg<ND, D, OSBL_CONV_WENO_Z>(avg, Q, R, g, ph, seg, acc, met, theta, s); break;
    case OSBL_CONV_TENO5: launch_sweep_avg<ND, D, OSBL_CONV_TENO5>(avg, Q, R, g, ph, seg, acc, met, theta, s); break;
    case CONV_WFILTER:
      // the WENO filter reconstructs with SimpleAverage (WENO_filter.py:324)
      if constexpr (ND >= 2) launch_sweep_t<ND, D, CONV_WFILTER, OSBL_AVG_SIMPLE>(Q, R, g, ph, seg, acc, met, theta, s);
      break;
    default: launch_sweep_avg<ND, D, OSBL_CONV_TENO6>(avg, Q, R, g, ph, seg, acc, met, theta, s); break;
  }
}

void launch_sweep(int nd, int d, int conv, int avg, const double* Q, double* R, const Grid& g, const Phys& ph, int seg,
                  const double* met, const double* theta, bool acc, hipStream_t s) {
  if (nd == 3) {
    if (d == 0) launch_sweep_conv<3, 0>(conv, avg, Q, R, g, ph, seg, met, theta, acc, s);
    else if (d == 1) launch_sweep_conv<3, 1>(conv, avg, Q, R, g, ph, seg, met, theta, acc, s);
    else launch_sweep_conv<3, 2>(conv, avg, Q, R, g, ph, seg, met, theta, acc, s);
  } else if (nd == 2) {
    if (d == 0) launch_sweep_conv<2, 0>(conv, avg, Q, R, g, ph, seg, met, theta, acc, s);
    else launch_sweep_conv<2, 1>(conv, avg, Q, R, g, ph, seg, met, theta, acc, s);
  } else {
    launch_sweep_conv<1, 0>(conv, avg, Q, R, g, ph, seg, met, theta, acc, s);
  }
}

template <int ND, int CONV, bool VISC>
static void launch_update_m(int mode, const double* Q, const double* P, double* R, double* Qo, double* Qn,
                            const Grid& g, const Phys& ph, double cA, double cB, int first, hipStream_t s) {
  dim3 grid((g.n[0] + 127) / 128, g.n[1], g.n[2]);
  if (mode == 0) k_update<ND, CONV, VISC, 0><<<grid, 128, 0, s>>>(Q, P, R, Qo, Qn, g, ph, cA, cB, first);
  else if (mode == 1) k_update<ND, CONV, VISC, 1><<<grid, 128, 0, s>>>(Q, P, R, Qo, Qn, g, ph, cA, cB, first);
  else k_update<ND, CONV, VISC, 2><<<grid, 128, 0, s>>>(Q, P, R, Qo, Qn, g, ph, cA, cB, first);
}

template <int ND>
static void launch_update_nd(int conv, bool visc, int mode, const double* Q, const double* P, double* R, double* Qo,
                             double* Qn, const Grid& g, const Phys& ph, double cA, double cB, int first,
                             hipStream_t s) {
  const bool central = conv == OSBL_CONV_CENTRAL_SKEW || conv == OSBL_CONV_CENTRAL_CONS;
  if (central) {
    if (conv == OSBL_CONV_CENTRAL_SKEW) {
      if (visc) launch_update_m<ND, OSBL_CONV_CENTRAL_SKEW, true>(mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
      else launch_update_m<ND, OSBL_CONV_CENTRAL_SKEW, false>(mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
    } else {
      if (visc) launch_update_m<ND, OSBL_CONV_CENTRAL_CONS, true>(mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
      else launch_update_m<ND, OSBL_CONV_CENTRAL_CONS, false>(mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
    }
  } else {
    if (visc) launch_update_m<ND, OSBL_CONV_WENO_Z, true>(mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
    else launch_update_m<ND, OSBL_CONV_WENO_Z, false>(mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
  }
}

void launch_update(int nd, int conv, bool visc, int mode, const double* Q, const double* P, double* R, double* Qo,
                   double* Qn, const Grid& g, const Phys& ph, double cA, double cB, int first, hipStream_t s) {
  if (nd == 3) launch_update_nd<3>(conv, visc, mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
  else if (nd == 2) launch_update_nd<2>(conv, visc, mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
  else launch_update_nd<1>(conv, visc, mode, Q, P, R, Qo, Qn, g, ph, cA, cB, first, s);
}

void launch_prims(int nd, const double* Q, double* P, const Grid& g, const Phys& ph, hipStream_t s) {
  const int PX = g.n[0] + 2 * g.h[0];
  dim3 grid((PX + 255) / 256, g.n[1] + 2 * g.h[1], g.n[2] + 2 * g.h[2]);
  if (nd == 3) k_prims<3><<<grid, 256, 0, s>>>(Q, P, g, ph);
  else if (nd == 2) k_prims<2><<<grid, 256, 0, s>>>(Q, P, g, ph);
  else k_prims<1><<<grid, 256, 0, s>>>(Q, P, g, ph);
}

long long halo_points(const Grid& g) {
  const long long PX = g.n[0] + 2 * g.h[0], PY = g.n[1] + 2 * g.h[1];
  return 2LL * g.h[2] * PY * PX + 2LL * g.h[1] * g.n[2] * PX + 2LL * g.h[0] * g.n[1] * (long long)g.n[2];
}

void launch_halo_periodic(double* Q, int nv, const Grid& g, hipStream_t s) {
  const long long total = halo_points(g);
  if (total == 0) return;
  k_halo_periodic<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(Q, nv, g, total);
}

void launch_extrap0(double* Q, int nv, const Grid& g, int d, int side, hipStream_t s) {
  const int e1 = (d == 0) ? 1 : 0, e2 = (d == 2) ? 1 : 2;
  const int n = (g.n[e1] + 2 * g.h[e1]) * (g.n[e2] + 2 * g.h[e2]);
  k_extrap0<<<(n + 255) / 256, 256, 0, s>>>(Q, nv, g, d, side);
}

void launch_count_nan(const double* F, int v0, int v1, unsigned long long* count, const Grid& g, hipStream_t s) {
  k_clear_count<<<1, 1, 0, s>>>(count);
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  k_count_nan<<<grid, 256, 0, s>>>(F, g, v0, v1, count);
}

void launch_pack(const double* F, double* out, const Grid& g, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  k_pack<<<grid, 256, 0, s>>>(F, out, g);
}
void launch_unpack(const double* in, double* F, const Grid& g, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  k_unpack<<<grid, 256, 0, s>>>(in, F, g);
}

}  // namespace osbl
