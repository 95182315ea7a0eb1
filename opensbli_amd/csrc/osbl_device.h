// osbl_device.h -- device-side building blocks of the OpenSBLI RHS on gfx950:
// grid indexing, constituent relations, Roe averaging, the Euler
// eigensystems, WENO5-JS/Z and TENO5/6 reconstructions.
//
// All arithmetic is FP64.  Formulas follow the reference's symbolic
// definitions; where an expression is rearranged to save divisions the
// rearrangement is algebraically exact and noted at the function.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#include "../../include/opensbli_amd.h"

namespace osbl {

// ---------------------------------------------------------------------------
// Storage: one array per field, x fastest.  The interior x range starts at a
// 128-byte boundary (offset OX doubles into each row), y and z carry `h`
// halo points.  Absent dimensions have n = 1 and no halo.
// ---------------------------------------------------------------------------
constexpr int OX = 16;

struct Grid {
  int nd;
  int n[3];
  int h[3];
  int pitch;  // doubles per row
  int rows;   // rows per plane
  int planes;
  int plane;  // doubles per plane
  long long fs;  // doubles per field (padded)
  int periodic_x;  // x faces periodic
  __host__ __device__ __forceinline__ int idx(int i, int j, int k) const {
    return (i + OX) + pitch * ((j + h[1]) + rows * (k + h[2]));
  }
  __host__ __device__ __forceinline__ int stride(int d) const { return d == 0 ? 1 : (d == 1 ? pitch : plane); }
};

// Segment length of a sliding-window sweep: `seg` cells per thread, halved
// (down to `lo`) while the launch would have fewer than 2^17 threads, so small
// blocks (2-D, 1-D lines) still give the GPU a few thousand waves.
__host__ __forceinline__ int fit_segment(long long lines, int n, int seg, int lo) {
  while (seg > lo && lines * ((n + seg - 1) / seg) < (1LL << 17)) seg /= 2;
  return seg;
}

// Workgroups are dispatched to the 8 XCDs round-robin by linear id.  The
// tile of a workgroup: linear ids are renumbered so that each XCD gets one
// contiguous range of tiles (x fastest, then y, then z) and tiles sharing
// halo rows sit behind one L2.  A bijection when the tile count is a
// multiple of 8, the identity otherwise.
__device__ __forceinline__ void xcd_tile(int& bx, int& by, int& bz) {
  const int nx = gridDim.x, ny = gridDim.y, total = nx * ny * gridDim.z;
  const int lin = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const int logical = (total % 8 == 0) ? (lin % 8) * (total / 8) + lin / 8 : lin;
  bx = logical % nx;
  by = (logical / nx) % ny;
  bz = logical / (nx * ny);
}

struct Phys {
  double gama, gm1, Minf, Re, Pr, dt, eps, teno_ct;
  double inv_delta[3];
  double inv_re, cq;  // 1/Re and 1/((gama-1) Minf^2 Pr Re)
  double suth;        // SuthT/RefT of the Sutherland law
  double teno_a1, teno_a2, sensor_eps;  // adaptive TENO cut-off and Ducros sensor constants
  double igm1;        // 1/(gama-1)
  double force[3];    // c_j of the constant body force (zero without one)
  int has_force;
  int nt;             // non-temporal hint on the residual and RK registers
};

// 10^n for an integer n by repeated products (exact for |n| <= 22)
__host__ __device__ __forceinline__ double pow10i(int n) {
  double r = 1.0;
  const int m = n < 0 ? -n : n;
  for (int i = 0; i < m; ++i) r *= 10.0;
  return n < 0 ? 1.0 / r : r;
}

// adaptive TENO cut-off from the sensor at an interface's left point
// (teno.py:430-443): C_T = 1/10^floor(a1 - a2 (1 - (1-theta)^4 (1 + 4 theta))),
// each operation rounded separately so that floor() sees the same value as a
// build without contraction
__device__ __forceinline__ double adaptive_ct(double m, const Phys& ph) {
  const double t = __dsub_rn(1.0, m);
  const double t2 = __dmul_rn(t, t);
  const double t4 = __dmul_rn(t2, t2);
  const double q = __dmul_rn(t4, __dadd_rn(1.0, __dmul_rn(4.0, m)));
  const double beta = __dsub_rn(ph.teno_a1, __dmul_rn(ph.teno_a2, __dsub_rn(1.0, q)));
  return 1.0 / pow10i((int)floor(beta));
}

// Near-boundary stencils and metric of a block with non-periodic faces
// (bc_core.py:15-16, opensblifunctions.py:523-534, metric.py:72-212).
struct Bnd {
  int mod[3][2];        // 0: central stencil; 1 + OSBL_BS_* where the face's BC replaces it
  const double* dm[3];  // D_dd of a direction stretched along itself (n[d] values), or nullptr
  const double* sd[3];  // SD_ddd = CD(D_dd, xi_d)
  const double* cw;     // closure rows (6 weights each): Carpenter first derivative rows 0-3 at 0,
                        // ReducedAccess first derivative rows 0-1 at 24, second derivative rows 0-1 at 36
};
constexpr int CW_CARP = 0, CW_RA = 24, CW_D2 = 36, CW_SIZE = 48;

// ---------------------------------------------------------------------------
// compile-time loop
// ---------------------------------------------------------------------------
template <int B, int E, class F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// ---------------------------------------------------------------------------
// Point state: conservative variables and the constituent relations
// u_i = rhou_i/rho, p = (gama-1)(rhoE - rho/2 u.u), a = (gama p/rho)^0.5
// (taylor_green_vortex.py:29-31, Sod_shock_tube.py:26-28).
// ---------------------------------------------------------------------------
template <int ND>
struct Pt {
  double q[ND + 2];
  double u[ND];
  double p, a;
};

// Loads and stores of the arrays a stage kernel touches once per cell (the
// residual and the RK registers).  On blocks whose fields exceed the
// last-level cache (Phys::nt) they carry the non-temporal hint, so that they
// stream through L2 without evicting the Q planes the stencils of
// neighbouring tiles and rows still re-read from it; on smaller blocks the
// whole working set stays cached between kernels and plain accesses keep it
// there.
__device__ __forceinline__ double ld_once(const double* p, int nt) {
  return nt ? __builtin_nontemporal_load(p) : *p;
}
__device__ __forceinline__ void st_once(double v, double* p, int nt) {
  if (nt) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// a/b (a of either sign) and 1/b for positive normal b far from the
// overflow range:
// the hardware reciprocal estimate refined by a Newton-Raphson step, the
// quotient then corrected once by its residual (Markstein's method); within
// an ulp of the correctly rounded result, in about half the instructions of
// the IEEE division sequence, which must also handle scaling and special
// operands.
__device__ __forceinline__ double quot_pos(double a, double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = __fma_rn(__fma_rn(-b, r, 1.0), r, r);
  const double q = a * r;
  return __fma_rn(__fma_rn(-b, q, a), r, q);
}
// sqrt(x) for positive normal x: the hardware reciprocal-square-root
// estimate refined by a Newton-Raphson step, then y = x r corrected once by
// its residual, y + r (x - y^2)/2; within an ulp of the correctly rounded root.
__device__ __forceinline__ double sqrt_pos(double x) {
  double r = __builtin_amdgcn_rsq(x);
  r = r * __fma_rn(-0.5 * x * r, r, 1.5);
  const double y = x * r;
  return __fma_rn(0.5 * r, __fma_rn(-y, y, x), y);
}
__device__ __forceinline__ double recip_pos(double b) {
  double r = __builtin_amdgcn_rcp(b);
  r = __fma_rn(__fma_rn(-b, r, 1.0), r, r);
  return __fma_rn(__fma_rn(-b, r, 1.0), r, r);
}

// Every operation is rounded on its own (no contraction into FMAs), so a
// point's values are bit-identical wherever a kernel evaluates them -- in a
// sweep's prologue or in its loop -- and a direction along which the state is
// uniform keeps exactly zero flux differences.
template <int ND>
__device__ __forceinline__ void finish_pt(Pt<ND>& s, const Phys& ph) {
  const double ir = recip_pos(s.q[0]);
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    s.u[m] = __dmul_rn(s.q[1 + m], ir);
    ke = __dadd_rn(ke, __dmul_rn(s.u[m], s.u[m]));
  }
  s.p = __dmul_rn(ph.gm1, __dsub_rn(s.q[ND + 1], __dmul_rn(__dmul_rn(0.5, s.q[0]), ke)));
  s.a = sqrt_pos(__dmul_rn(__dmul_rn(ph.gama, s.p), ir));
}

// raw conservative variables of one point
template <int ND>
__device__ __forceinline__ void fetch_q(double (&q)[ND + 2], const double* __restrict__ Q, long long fs, int off) {
#pragma unroll
  for (int m = 0; m < ND + 2; ++m) q[m] = Q[m * fs + off];
}

template <int ND>
__device__ __forceinline__ void load_pt(Pt<ND>& s, const double* __restrict__ Q, long long fs, int off,
                                         const Phys& ph) {
  fetch_q<ND>(s.q, Q, fs, off);
  finish_pt<ND>(s, ph);
}

// ---------------------------------------------------------------------------
// Interface state between points i and i+1 (the averaged symbols of the
// eigensystem).  Roe: averaging.py:70-114.  Simple: averaging.py:38-59.
// ---------------------------------------------------------------------------
template <int ND, class T = double>
struct AvgT {
  T rho, a, u[ND];
  T ia, ir;           // 1/a, 1/rho (inv_AVG_a, inv_AVG_rho, shock_capturing.py:319-331)
  T phi;              // (gama-1)|u|^2/2
  T alpha, bta;       // rho/(a sqrt2), 1/(rho a sqrt2)
  T Hh;               // a^2/(gama-1) + |u|^2/2  (1-D eigensystem)
};
template <int ND>
using Avg = AvgT<ND, double>;

// the derived symbols of an averaged state (rho, a, u)
template <int ND, class T>
__device__ __forceinline__ void finish_avg(AvgT<ND, T>& v, const Phys& ph) {
  T uu = v.u[0] * v.u[0];
#pragma unroll
  for (int m = 1; m < ND; ++m) uu = uu + v.u[m] * v.u[m];
  if constexpr (std::is_same_v<T, double>) {
    // 1/a and 1/rho from one division
    const double t = recip_pos(v.a * v.rho);
    v.ia = v.rho * t;
    v.ir = v.a * t;
  } else {
    // dual numbers (the adjoint's tangents): two plain reciprocals are cheaper
    v.ia = 1.0 / v.a;
    v.ir = 1.0 / v.rho;
  }
  v.phi = ph.gm1 * 0.5 * uu;
  constexpr double RS2 = 0.70710678118654752440;  // 1/sqrt(2.0)
  v.alpha = v.rho * v.ia * RS2;
  v.bta = v.ir * v.ia * RS2;
  v.Hh = v.a * v.a * ph.igm1 + 0.5 * uu;
}

// the per-point factors of the Roe average: sqrt(rho) and
// sqrt(rho) H = (rhoE + p)/sqrt(rho).  Sweeps that visit every point once
// compute them once per point and pass them in (RoeF); the result is the same.
struct RoeF {
  double sL, hL, sR, hR;
};
template <int ND>
__device__ __forceinline__ void roe_factors(const Pt<ND>& P, double& sq, double& h) {
  sq = sqrt_pos(P.q[0]);
  h = (P.q[ND + 1] + P.p) * quot_pos(sq, P.q[0]);
}

template <int ND>
__device__ __forceinline__ void roe_state(Avg<ND>& v, const Pt<ND>& L, const Pt<ND>& R, const RoeF& f,
                                          const Phys& ph) {
  const double inv = recip_pos(f.sR + f.sL);
  v.rho = f.sL * f.sR;  // sqrt(rho_L rho_R)
  double ke = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    v.u[m] = (f.sL * L.u[m] + f.sR * R.u[m]) * inv;
    ke += v.u[m] * v.u[m];
  }
  const double hh = (f.hL + f.hR) * inv;
  v.a = sqrt_pos(ph.gm1 * (hh - 0.5 * ke));
}

template <int ND, int AVG>
__device__ __forceinline__ void interface_state(Avg<ND>& v, const Pt<ND>& L, const Pt<ND>& R, const Phys& ph) {
  if constexpr (AVG == OSBL_AVG_ROE) {
    RoeF f;
    roe_factors<ND>(L, f.sL, f.hL);
    roe_factors<ND>(R, f.sR, f.hR);
    roe_state<ND>(v, L, R, f, ph);
  } else {
    v.rho = 0.5 * (L.q[0] + R.q[0]);
#pragma unroll
    for (int m = 0; m < ND; ++m) v.u[m] = 0.5 * (L.u[m] + R.u[m]);
    v.a = 0.5 * (L.a + R.a);
  }
  finish_avg<ND>(v, ph);
}

// ---------------------------------------------------------------------------
// Left/right eigenvectors of the Euler flux Jacobian for a Cartesian
// direction D (unit normal k = e_D), euler_eigensystem.py:57-135.  Field
// order follows the reference: 3-D (U,U,U,U+a,U-a), 2-D (U,U,U+a,U-a),
// 1-D (u-a,u,u+a).  In 3-D the rows k != D are the two shear waves whose
// non-zero entries sit at component c = 3-D-k with sign eps(D,k,c); this is
// the reference's k_t algebra with k_t = e_D substituted.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int lc3(int D, int k) { return 3 - D - k; }
__host__ __device__ constexpr int eps3(int a, int b, int c) {
  return (a == 0 && b == 1 && c == 2) || (a == 1 && b == 2 && c == 0) || (a == 2 && b == 0 && c == 1) ? 1 : -1;
}

// bit m set when L[k][m] is structurally non-zero
template <int ND, int D, int K>
__host__ __device__ constexpr unsigned lmask() {
  if constexpr (ND == 3) {
    if constexpr (K < 3 && K != D) return 1u | (1u << (1 + lc3(D, K)));
    else return 0x1Fu;
  } else if constexpr (ND == 2) {
    if constexpr (K == 1) return 1u | (1u << (1 + (1 - D)));
    else return 0xFu;
  } else {
    return 0x7u;
  }
}
template <int ND, int D, int K>
__host__ __device__ constexpr unsigned rmask() {
  if constexpr (ND == 3) {
    if constexpr (K < 3 && K != D) return (1u << (1 + lc3(D, K))) | (1u << 4);
    else return 0x1Fu;
  } else if constexpr (ND == 2) {
    if constexpr (K == 1) return (1u << (1 + (1 - D))) | (1u << 3);
    else return 0xFu;
  } else {
    return 0x7u;
  }
}

// row K of the left eigenvector matrix (T: double, or a dual number type for
// the derivatives with respect to the averaged state)
template <int ND, int D, int K, class T = double>
__device__ __forceinline__ void lrow(const AvgT<ND, T>& v, const Phys& ph, T l[ND + 2]) {
  const double gm1 = ph.gm1;
  if constexpr (ND == 3) {
    const T ia2 = v.ia * v.ia;
    if constexpr (K == D) {
      l[0] = 1.0 - v.phi * ia2;
      l[1] = gm1 * v.u[0] * ia2;
      l[2] = gm1 * v.u[1] * ia2;
      l[3] = gm1 * v.u[2] * ia2;
      l[4] = -gm1 * ia2;
    } else if constexpr (K < 3) {
      constexpr int c = lc3(D, K);
      constexpr double s = (double)eps3(D, K, c);
      l[0] = -s * v.u[c] * v.ir;
      l[1 + c] = s * v.ir;
    } else {
      constexpr double sg = (K == 3) ? 1.0 : -1.0;
      const T th = v.u[D];
      l[0] = v.bta * (v.phi - sg * th * v.a);
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        if (m == D) l[1 + m] = -v.bta * (gm1 * v.u[m] - sg * v.a);
        else l[1 + m] = -v.bta * (gm1 * v.u[m]);
      }
      l[4] = v.bta * gm1;
    }
  } else if constexpr (ND == 2) {
    const T ia2 = v.ia * v.ia;
    if constexpr (K == 0) {
      l[0] = 1.0 - v.phi * ia2;
      l[1] = gm1 * v.u[0] * ia2;
      l[2] = gm1 * v.u[1] * ia2;
      l[3] = -gm1 * ia2;
    } else if constexpr (K == 1) {
      constexpr int c = 1 - D;
      constexpr double s = (D == 0) ? -1.0 : 1.0;
      l[0] = -s * v.u[c] * v.ir;
      l[1 + c] = s * v.ir;
    } else {
      constexpr double sg = (K == 2) ? 1.0 : -1.0;
      const T th = v.u[D];
      l[0] = v.bta * (v.phi - sg * v.a * th);
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        if (m == D) l[1 + m] = -v.bta * (gm1 * v.u[m] - sg * v.a);
        else l[1 + m] = -v.bta * (gm1 * v.u[m]);
      }
      l[3] = v.bta * gm1;
    }
  } else {
    const T u = v.u[0], a = v.a, Hh = v.Hh;
    const T den = 2.0 * Hh - u * u;
    const T iden = 1.0 / den;
    if constexpr (K == 0) {
      l[0] = u * (2.0 * Hh + a * u - u * u) * v.ia * 0.5 * iden;
      l[1] = (-Hh - a * u + 0.5 * u * u) * v.ia * iden;
      l[2] = iden;
    } else if constexpr (K == 1) {
      l[0] = 2.0 * (Hh - u * u) * iden;
      l[1] = 2.0 * u * iden;
      l[2] = -2.0 * iden;
    } else {
      l[0] = u * (-2.0 * Hh + a * u + u * u) * v.ia * 0.5 * iden;
      l[1] = (Hh - a * u - 0.5 * u * u) * v.ia * iden;
      l[2] = iden;
    }
  }
}

// column K of the right eigenvector matrix
template <int ND, int D, int K, class T = double>
__device__ __forceinline__ void rcol(const AvgT<ND, T>& v, const Phys& ph, T r[ND + 2]) {
  const double igm1 = ph.igm1;
  if constexpr (ND == 3) {
    if constexpr (K == D) {
      r[0] = 1.0;
      r[1] = v.u[0];
      r[2] = v.u[1];
      r[3] = v.u[2];
      r[4] = v.phi * igm1;
    } else if constexpr (K < 3) {
      constexpr int c = lc3(D, K);
      constexpr double s = (double)eps3(D, K, c);
      r[1 + c] = s * v.rho;
      r[4] = s * v.rho * v.u[c];
    } else {
      constexpr double sg = (K == 3) ? 1.0 : -1.0;
      r[0] = v.alpha;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        if (m == D) r[1 + m] = v.alpha * (v.u[m] + sg * v.a);
        else r[1 + m] = v.alpha * v.u[m];
      }
      r[4] = v.alpha * ((v.phi + v.a * v.a) * igm1 + sg * v.u[D] * v.a);
    }
  } else if constexpr (ND == 2) {
    if constexpr (K == 0) {
      r[0] = 1.0;
      r[1] = v.u[0];
      r[2] = v.u[1];
      r[3] = v.phi * igm1;
    } else if constexpr (K == 1) {
      constexpr int c = 1 - D;
      constexpr double s = (D == 0) ? -1.0 : 1.0;
      r[1 + c] = s * v.rho;
      r[3] = s * v.rho * v.u[c];
    } else {
      constexpr double sg = (K == 2) ? 1.0 : -1.0;
      r[0] = v.alpha;
#pragma unroll
      for (int m = 0; m < 2; ++m) {
        if (m == D) r[1 + m] = v.alpha * (v.u[m] + sg * v.a);
        else r[1 + m] = v.alpha * v.u[m];
      }
      r[3] = v.alpha * ((v.phi + v.a * v.a) * igm1 + sg * v.a * v.u[D]);
    }
  } else {
    const T u = v.u[0], a = v.a, Hh = v.Hh;
    if constexpr (K == 0) {
      r[0] = 1.0; r[1] = u - a; r[2] = Hh - u * a;
    } else if constexpr (K == 1) {
      r[0] = 1.0; r[1] = u; r[2] = 0.5 * u * u;
    } else {
      r[0] = 1.0; r[1] = u + a; r[2] = Hh + u * a;
    }
  }
}

// |lambda_K| at a stencil point from its own u_D and a (shock_capturing.py:512-536)
template <int ND, int D, int K>
__device__ __forceinline__ double abs_lambda_ua(double ud, double a) {
  if constexpr (ND == 1) {
    if constexpr (K == 0) return fabs(ud - a);
    else if constexpr (K == 1) return fabs(ud);
    else return fabs(ud + a);
  } else {
    constexpr int NV = ND + 2;
    if constexpr (K == NV - 2) return fabs(ud + a);
    else if constexpr (K == NV - 1) return fabs(ud - a);
    else return fabs(ud);
  }
}
template <int ND, int D, int K>
__device__ __forceinline__ double abs_lambda(const Pt<ND>& s) {
  return abs_lambda_ua<ND, D, K>(s.u[D], s.a);
}

// ---------------------------------------------------------------------------
// Reconstructions at i+1/2.  Arguments are the 6 points f[-2..3] in the
// orientation of the "right" reconstruction (weno.py side +1, applied to
// f+).  The left reconstruction (side -1, applied to f-) is the right one
// applied to the mirrored points f[3-p] -- except for TENO6, whose
// right-side fourth indicator has the unsquared 781/720 term
// (teno.py:166-167), hence the QUIRK flag.
// ---------------------------------------------------------------------------

// WENO5-Z (weno.py:283-338) with eps = 1e-14, p = 2.  The weights are
// alpha_r = d_r (1 + tau^2/b_r^2), b_r = eps + beta_r; multiplying every
// alpha_r by b0^2 b1^2 b2^2 leaves omega_r = alpha_r / sum(alpha) unchanged
// and needs one division instead of four.
// numerator and denominator of the WENO5-Z value (value = num/den)
__device__ __forceinline__ void weno5z_nd(double fm2, double fm1, double f0, double f1, double f2, double& num,
                                          double& den) {
  const double t0 = f0 - 2.0 * f1 + f2, s0 = 3.0 * f0 - 4.0 * f1 + f2;
  const double t1 = fm1 - 2.0 * f0 + f1, s1 = fm1 - f1;
  const double t2 = fm2 - 2.0 * fm1 + f0, s2 = fm2 - 4.0 * fm1 + 3.0 * f0;
  constexpr double C13 = 13.0 / 12.0;
  const double b0 = C13 * t0 * t0 + 0.25 * s0 * s0;
  const double b1 = C13 * t1 * t1 + 0.25 * s1 * s1;
  const double b2 = C13 * t2 * t2 + 0.25 * s2 * s2;
  const double tau = fabs(b0 - b2);
  const double tt = tau * tau;
  const double e0 = 1.0e-14 + b0, e1 = 1.0e-14 + b1, e2 = 1.0e-14 + b2;
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double A0 = 0.3 * (q0 + tt) * (q1 * q2);
  const double A1 = 0.6 * (q1 + tt) * (q0 * q2);
  const double A2 = 0.1 * (q2 + tt) * (q0 * q1);
  // the candidates' common factor 1/6 moves into the denominator
  const double r0 = 2.0 * f0 + 5.0 * f1 - f2;
  const double r1 = -fm1 + 5.0 * f0 + 2.0 * f1;
  const double r2 = 2.0 * fm2 - 7.0 * fm1 + 11.0 * f0;
  num = A0 * r0 + A1 * r1 + A2 * r2;
  den = 6.0 * (A0 + A1 + A2);
}

__device__ __forceinline__ double weno5z(double fm2, double fm1, double f0, double f1, double f2) {
  double n, d;
  weno5z_nd(fm2, fm1, f0, f1, f2, n, d);
  return n / d;
}

// n1/d1 + n2/d2 with one division: (n1 d2 + n2 d1)/(d1 d2).  The weights'
// denominators grow like beta^6 (at least ~1e-85 from eps), so d1 d2 is a
// normal number unless |f| exceeds ~1e12; then the two quotients are formed
// separately.
__device__ __forceinline__ double sum_quotients(double n1, double d1, double n2, double d2) {
  const double dd = d1 * d2;
  if (dd < 1.0e300) return quot_pos(n1 * d2 + n2 * d1, dd);
  return n1 / d1 + n2 / d2;
}

// WENO5-JS (weno.py:341-369) with eps = 1e-6: alpha_r = d_r/(eps+beta_r)^2,
// rewritten as d_r * prod_{s!=r} (eps+beta_s)^2 with one division.
__device__ __forceinline__ void weno5js_nd(double fm2, double fm1, double f0, double f1, double f2, double& num,
                                           double& den) {
  const double t0 = f0 - 2.0 * f1 + f2, s0 = 3.0 * f0 - 4.0 * f1 + f2;
  const double t1 = fm1 - 2.0 * f0 + f1, s1 = fm1 - f1;
  const double t2 = fm2 - 2.0 * fm1 + f0, s2 = fm2 - 4.0 * fm1 + 3.0 * f0;
  constexpr double C13 = 13.0 / 12.0;
  const double e0 = 1.0e-6 + (C13 * t0 * t0 + 0.25 * s0 * s0);
  const double e1 = 1.0e-6 + (C13 * t1 * t1 + 0.25 * s1 * s1);
  const double e2 = 1.0e-6 + (C13 * t2 * t2 + 0.25 * s2 * s2);
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double A0 = 0.3 * (q1 * q2), A1 = 0.6 * (q0 * q2), A2 = 0.1 * (q0 * q1);
  // the candidates' common factor 1/6 moves into the denominator
  const double r0 = 2.0 * f0 + 5.0 * f1 - f2;
  const double r1 = -fm1 + 5.0 * f0 + 2.0 * f1;
  const double r2 = 2.0 * fm2 - 7.0 * fm1 + 11.0 * f0;
  num = A0 * r0 + A1 * r1 + A2 * r2;
  den = 6.0 * (A0 + A1 + A2);
}

__device__ __forceinline__ double weno5js(double fm2, double fm1, double f0, double f1, double f2) {
  double n, d;
  weno5js_nd(fm2, fm1, f0, f1, f2, n, d);
  return n / d;
}

// WENO5-JS as above, also returning the shock sensor of its non-linear
// weights (weno.py:415-431, normalised, sensor_theta = 4):
// rj = sum_r |omega_r/d_r - 1|^4 / (|1/min d - 1|^4 + 2) = (...)/6563.
__device__ __forceinline__ double weno5js_sensor(double fm2, double fm1, double f0, double f1, double f2, double& rj) {
  const double t0 = f0 - 2.0 * f1 + f2, s0 = 3.0 * f0 - 4.0 * f1 + f2;
  const double t1 = fm1 - 2.0 * f0 + f1, s1 = fm1 - f1;
  const double t2 = fm2 - 2.0 * fm1 + f0, s2 = fm2 - 4.0 * fm1 + 3.0 * f0;
  constexpr double C13 = 13.0 / 12.0;
  const double e0 = 1.0e-6 + (C13 * t0 * t0 + 0.25 * s0 * s0);
  const double e1 = 1.0e-6 + (C13 * t1 * t1 + 0.25 * s1 * s1);
  const double e2 = 1.0e-6 + (C13 * t2 * t2 + 0.25 * s2 * s2);
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double A0 = 0.3 * (q1 * q2), A1 = 0.6 * (q0 * q2), A2 = 0.1 * (q0 * q1);
  const double ia = 1.0 / (A0 + A1 + A2);
  const double w0 = A0 * ia, w1 = A1 * ia, w2 = A2 * ia;
  const double z0 = w0 / 0.3 - 1.0, z1 = w1 / 0.6 - 1.0, z2 = w2 / 0.1 - 1.0;
  const double y0 = z0 * z0, y1 = z1 * z1, y2 = z2 * z2;
  rj = (y0 * y0 + y1 * y1 + y2 * y2) / 6563.0;
  constexpr double S6 = 1.0 / 6.0;
  const double r0 = (2.0 * f0 + 5.0 * f1 - f2) * S6;
  const double r1 = (-fm1 + 5.0 * f0 + 2.0 * f1) * S6;
  const double r2 = (2.0 * fm2 - 7.0 * fm1 + 11.0 * f0) * S6;
  return w0 * r0 + w1 * r1 + w2 * r2;
}

// Flux of the WENO non-linear filter for one characteristic field
// (shock_capturing.py:82-96, 401-416, 438-443; WENO_filter.py:324-326): the
// WENO-JS reconstructions of f+ and f- summed, minus the sixth-order central
// flux (1, -8, 37, 37, -8, 1)/60 of the characteristic fluxes on points
// -2..3, times rj = max(sensor(f+ side), sensor(f- side)).
__device__ __forceinline__ double weno_filter_flux(const double fp[6], const double fm[6], const double cf[6]) {
  double rr, rl;
  const double rec = weno5js_sensor(fp[0], fp[1], fp[2], fp[3], fp[4], rr) +
                     weno5js_sensor(fm[5], fm[4], fm[3], fm[2], fm[1], rl);
  const double central = (cf[0] - 8.0 * cf[1] + 37.0 * cf[2] + 37.0 * cf[3] - 8.0 * cf[4] + cf[5]) / 60.0;
  return fmax(rr, rl) * (rec - central);
}

// scheme tag of the WENO filter's characteristic sweeps (not a convective scheme)
constexpr int CONV_WFILTER = 16;

__device__ __forceinline__ double pow6(double x) {
  const double x2 = x * x;
  return x2 * x2 * x2;
}

// TENO5 / TENO6 (teno.py:57-234, 407-465):
// alpha_r = (1 + tau/(eps + beta_r))^6, delta_r = 0 if alpha_r*(1/sum alpha) < C_T,
// omega_r = d_r delta_r / sum(d delta).
// Numerator and denominator of the value (value = num/den).  The four (three)
// quotients tau/(eps + beta_r) share one division: with e_r = eps + beta_r,
// 1/e_0 = e_1 e_2 e_3 / (e_0 e_1 e_2 e_3) and so on, unless that product
// leaves the normal range (then each quotient is formed on its own).  The
// cut-off compares alpha_r * (1/sum alpha) with C_T as written.
template <bool SIX, bool QUIRK>
__device__ __forceinline__ void teno_nd(double fm2, double fm1, double f0, double f1, double f2, double f3,
                                        double eps, double ct, double& num, double& den) {
  constexpr double C13 = 13.0 / 12.0;
  double q;
  q = fm1 - f1;
  double b0 = 0.25 * q * q;
  q = fm1 - 2.0 * f0 + f1;
  b0 += C13 * q * q;
  q = 3.0 * f0 - 4.0 * f1 + f2;
  double b1 = 0.25 * q * q;
  q = f0 - 2.0 * f1 + f2;
  b1 += C13 * q * q;
  q = fm2 - 4.0 * fm1 + 3.0 * f0;
  double b2 = 0.25 * q * q;
  q = fm2 - 2.0 * fm1 + f0;
  b2 += C13 * q * q;
  constexpr double S6 = 1.0 / 6.0;
  const double r0 = (-fm1 + 5.0 * f0 + 2.0 * f1) * S6;
  const double r1 = (2.0 * f0 + 5.0 * f1 - f2) * S6;
  const double r2 = (2.0 * fm2 - 7.0 * fm1 + 11.0 * f0) * S6;
  const double e0 = eps + b0, e1 = eps + b1, e2 = eps + b2;
  if constexpr (SIX) {
    q = -11.0 * f0 + 18.0 * f1 - 9.0 * f2 + 2.0 * f3;
    double b3 = (1.0 / 36.0) * q * q;
    q = 2.0 * f0 - 5.0 * f1 + 4.0 * f2 - f3;
    b3 += C13 * q * q;
    q = -f0 + 3.0 * f1 - 3.0 * f2 + f3;
    if constexpr (QUIRK) b3 += (781.0 / 720.0) * q;
    else b3 += (781.0 / 720.0) * q * q;
    const double r3 = (3.0 * f0 + 13.0 * f1 - 5.0 * f2 + f3) * (1.0 / 12.0);
    const double tau = fabs(b3 - (b0 + b2 - 4.0 * b1) * S6);
    const double e3 = eps + b3;
    const double p01 = e0 * e1, p23 = e2 * e3;
    const double pr = p01 * p23;
    double a0, a1, a2, a3;
    if (fabs(pr) > 1.0e-300 && fabs(pr) < 1.0e300) {
      const double tp = quot_pos(tau, pr);
      a0 = pow6(1.0 + (e1 * p23) * tp);
      a1 = pow6(1.0 + (e0 * p23) * tp);
      a2 = pow6(1.0 + (e3 * p01) * tp);
      a3 = pow6(1.0 + (e2 * p01) * tp);
    } else {
      a0 = pow6(1.0 + tau / e0);
      a1 = pow6(1.0 + tau / e1);
      a2 = pow6(1.0 + tau / e2);
      a3 = pow6(1.0 + tau / e3);
    }
    const double ias = recip_pos(a0 + a1 + a2 + a3);
    const double d0 = (a0 * ias < ct) ? 0.0 : 1.0;
    const double d1 = (a1 * ias < ct) ? 0.0 : 1.0;
    const double d2 = (a2 * ias < ct) ? 0.0 : 1.0;
    const double d3 = (a3 * ias < ct) ? 0.0 : 1.0;
    constexpr double D0 = 231.0 / 500.0, D1 = 3.0 / 10.0, D2 = 27.0 / 500.0, D3 = 23.0 / 125.0;
    const double w0 = D0 * d0, w1 = D1 * d1, w2 = D2 * d2, w3 = D3 * d3;
    num = w0 * r0 + w1 * r1 + w2 * r2 + w3 * r3;
    den = w0 + w1 + w2 + w3;
  } else {
    const double tau = fabs(b0 - b2);
    const double p01 = e0 * e1;
    const double pr = p01 * e2;
    double a0, a1, a2;
    if (fabs(pr) > 1.0e-300 && fabs(pr) < 1.0e300) {
      const double tp = quot_pos(tau, pr);
      a0 = pow6(1.0 + (e1 * e2) * tp);
      a1 = pow6(1.0 + (e0 * e2) * tp);
      a2 = pow6(1.0 + p01 * tp);
    } else {
      a0 = pow6(1.0 + tau / e0);
      a1 = pow6(1.0 + tau / e1);
      a2 = pow6(1.0 + tau / e2);
    }
    const double ias = recip_pos(a0 + a1 + a2);
    const double d0 = (a0 * ias < ct) ? 0.0 : 1.0;
    const double d1 = (a1 * ias < ct) ? 0.0 : 1.0;
    const double d2 = (a2 * ias < ct) ? 0.0 : 1.0;
    constexpr double D0 = 11.0 / 20.0, D1 = 4.0 / 10.0, D2 = 1.0 / 20.0;
    const double w0 = D0 * d0, w1 = D1 * d1, w2 = D2 * d2;
    num = w0 * r0 + w1 * r1 + w2 * r2;
    den = w0 + w1 + w2;
  }
}

// The same with one division per quotient (kernels that keep their stencil
// in registers: the shared-division form needs more registers than they
// have); the quotients are all of positive operands (eps + beta > 0, sums of
// weights >= 1), formed by quot_pos / recip_pos.
template <bool SIX, bool QUIRK>
__device__ __forceinline__ double teno(double fm2, double fm1, double f0, double f1, double f2, double f3,
                                       double eps, double ct) {
  constexpr double C13 = 13.0 / 12.0;
  double q;
  q = fm1 - f1;
  double b0 = 0.25 * q * q;
  q = fm1 - 2.0 * f0 + f1;
  b0 += C13 * q * q;
  q = 3.0 * f0 - 4.0 * f1 + f2;
  double b1 = 0.25 * q * q;
  q = f0 - 2.0 * f1 + f2;
  b1 += C13 * q * q;
  q = fm2 - 4.0 * fm1 + 3.0 * f0;
  double b2 = 0.25 * q * q;
  q = fm2 - 2.0 * fm1 + f0;
  b2 += C13 * q * q;
  constexpr double S6 = 1.0 / 6.0;
  const double r0 = (-fm1 + 5.0 * f0 + 2.0 * f1) * S6;
  const double r1 = (2.0 * f0 + 5.0 * f1 - f2) * S6;
  const double r2 = (2.0 * fm2 - 7.0 * fm1 + 11.0 * f0) * S6;
  if constexpr (SIX) {
    q = -11.0 * f0 + 18.0 * f1 - 9.0 * f2 + 2.0 * f3;
    double b3 = (1.0 / 36.0) * q * q;
    q = 2.0 * f0 - 5.0 * f1 + 4.0 * f2 - f3;
    b3 += C13 * q * q;
    q = -f0 + 3.0 * f1 - 3.0 * f2 + f3;
    if constexpr (QUIRK) b3 += (781.0 / 720.0) * q;
    else b3 += (781.0 / 720.0) * q * q;
    const double r3 = (3.0 * f0 + 13.0 * f1 - 5.0 * f2 + f3) * (1.0 / 12.0);
    const double tau = fabs(b3 - (b0 + b2 - 4.0 * b1) * S6);
    const double a0 = pow6(1.0 + quot_pos(tau, eps + b0));
    const double a1 = pow6(1.0 + quot_pos(tau, eps + b1));
    const double a2 = pow6(1.0 + quot_pos(tau, eps + b2));
    const double a3 = pow6(1.0 + quot_pos(tau, eps + b3));
    const double ias = recip_pos(a0 + a1 + a2 + a3);
    const double d0 = (a0 * ias < ct) ? 0.0 : 1.0;
    const double d1 = (a1 * ias < ct) ? 0.0 : 1.0;
    const double d2 = (a2 * ias < ct) ? 0.0 : 1.0;
    const double d3 = (a3 * ias < ct) ? 0.0 : 1.0;
    constexpr double D0 = 231.0 / 500.0, D1 = 3.0 / 10.0, D2 = 27.0 / 500.0, D3 = 23.0 / 125.0;
    const double w0 = D0 * d0, w1 = D1 * d1, w2 = D2 * d2, w3 = D3 * d3;
    const double iws = recip_pos(w0 + w1 + w2 + w3);
    return (w0 * r0 + w1 * r1 + w2 * r2 + w3 * r3) * iws;
  } else {
    const double tau = fabs(b0 - b2);
    const double a0 = pow6(1.0 + quot_pos(tau, eps + b0));
    const double a1 = pow6(1.0 + quot_pos(tau, eps + b1));
    const double a2 = pow6(1.0 + quot_pos(tau, eps + b2));
    const double ias = recip_pos(a0 + a1 + a2);
    const double d0 = (a0 * ias < ct) ? 0.0 : 1.0;
    const double d1 = (a1 * ias < ct) ? 0.0 : 1.0;
    const double d2 = (a2 * ias < ct) ? 0.0 : 1.0;
    constexpr double D0 = 11.0 / 20.0, D1 = 4.0 / 10.0, D2 = 1.0 / 20.0;
    const double w0 = D0 * d0, w1 = D1 * d1, w2 = D2 * d2;
    const double iws = recip_pos(w0 + w1 + w2);
    return (w0 * r0 + w1 * r1 + w2 * r2) * iws;
  }
}

// Sum of the right reconstruction of f+ and the left reconstruction of f-
// (shock_capturing.py:257-265) for one characteristic field.
template <int CONV, bool SHARED = true>
__device__ __forceinline__ double recon_pair(const double fp[6], const double fm[6], double eps, double ct) {
  if constexpr (CONV == OSBL_CONV_WENO_Z) {
    double n1, d1, n2, d2;
    weno5z_nd(fp[0], fp[1], fp[2], fp[3], fp[4], n1, d1);
    weno5z_nd(fm[5], fm[4], fm[3], fm[2], fm[1], n2, d2);
    return sum_quotients(n1, d1, n2, d2);
  } else if constexpr (CONV == OSBL_CONV_WENO_JS) {
    double n1, d1, n2, d2;
    weno5js_nd(fp[0], fp[1], fp[2], fp[3], fp[4], n1, d1);
    weno5js_nd(fm[5], fm[4], fm[3], fm[2], fm[1], n2, d2);
    return sum_quotients(n1, d1, n2, d2);
  } else if constexpr (!SHARED && CONV == OSBL_CONV_TENO6) {
    return teno<true, true>(fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], eps, ct) +
           teno<true, false>(fm[5], fm[4], fm[3], fm[2], fm[1], fm[0], eps, ct);
  } else if constexpr (!SHARED) {
    return teno<false, false>(fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], eps, ct) +
           teno<false, false>(fm[5], fm[4], fm[3], fm[2], fm[1], fm[0], eps, ct);
  } else if constexpr (CONV == OSBL_CONV_TENO6) {
    double n1, d1, n2, d2;
    teno_nd<true, true>(fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], eps, ct, n1, d1);
    teno_nd<true, false>(fm[5], fm[4], fm[3], fm[2], fm[1], fm[0], eps, ct, n2, d2);
    return sum_quotients(n1, d1, n2, d2);
  } else {
    double n1, d1, n2, d2;
    teno_nd<false, false>(fp[0], fp[1], fp[2], fp[3], fp[4], fp[5], eps, ct, n1, d1);
    teno_nd<false, false>(fm[5], fm[4], fm[3], fm[2], fm[1], fm[0], eps, ct, n2, d2);
    return sum_quotients(n1, d1, n2, d2);
  }
}

// ---------------------------------------------------------------------------
// The six stencil points of an interface, as char_flux reads them: from a
// sliding register window (RegWindow), or from the LDS slots of a staged
// chunk (LdsWindow, fields q, u, p, a in rows of `stride` doubles).  With the
// LDS window every characteristic field re-reads the points it needs instead
// of holding all six in registers across the fields.
// ---------------------------------------------------------------------------
template <int ND>
struct RegWindow {
  static constexpr bool STREAM = false;
  const Pt<ND>* w;
  __device__ __forceinline__ double q(int p, int m) const { return w[p].q[m]; }
  __device__ __forceinline__ double u(int p, int m) const { return w[p].u[m]; }
  __device__ __forceinline__ double pr(int p) const { return w[p].p; }
  __device__ __forceinline__ double a(int p) const { return w[p].a; }
  __device__ __forceinline__ void point(int p, Pt<ND>& out) const { out = w[p]; }
};

template <int ND>
struct LdsWindow {
  static constexpr bool STREAM = true;
  const double* s;  // field 0 of point 0
  int stride;       // doubles between fields
  __device__ __forceinline__ double f(int p, int e) const { return s[e * stride + p]; }
  __device__ __forceinline__ double q(int p, int m) const { return f(p, m); }
  __device__ __forceinline__ double u(int p, int m) const { return f(p, ND + 2 + m); }
  __device__ __forceinline__ double pr(int p) const { return f(p, 2 * ND + 2); }
  __device__ __forceinline__ double a(int p) const { return f(p, 2 * ND + 3); }
  __device__ __forceinline__ void point(int p, Pt<ND>& out) const {
#pragma unroll
    for (int m = 0; m < ND + 2; ++m) out.q[m] = q(p, m);
#pragma unroll
    for (int m = 0; m < ND; ++m) out.u[m] = u(p, m);
    out.p = pr(p);
    out.a = a(p);
  }
};

// ---------------------------------------------------------------------------
// LLF characteristic flux at the interface between window points 2 and 3
// (shock_capturing.py:357-448).  CF_k = L_k . F is formed from CS_k = L_k . Q
// through F = u_D Q + p (e_{1+D} + u_D e_last), which holds exactly for the
// Euler flux vector.
// ---------------------------------------------------------------------------
template <int ND, int D, int CONV, int AVG, bool CACHED = false, class W>
__device__ __forceinline__ void char_flux_w(const W& w, const Phys& ph, double fh[ND + 2], double ct,
                                            const RoeF& rf = RoeF{}) {
  constexpr int NV = ND + 2;
  Avg<ND> v;
  {
    Pt<ND> L, R;
    w.point(2, L);
    w.point(3, R);
    if constexpr (CACHED && AVG == OSBL_AVG_ROE) {
      roe_state<ND>(v, L, R, rf, ph);
      finish_avg<ND>(v, ph);
    } else {
      interface_state<ND, AVG>(v, L, R, ph);
    }
  }
#pragma unroll
  for (int m = 0; m < NV; ++m) fh[m] = 0.0;
  sfor<0, NV>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    constexpr unsigned LM = lmask<ND, D, K>();
    constexpr unsigned RM = rmask<ND, D, K>();
    double l[NV];
    lrow<ND, D, K>(v, ph, l);
    // the split fluxes 1/2 (CF +- lam CS) are formed from the half row L_k/2:
    // halving is exact, so CS/2 and CF/2 come out exactly and the two
    // products by 1/2 per point disappear
#pragma unroll
    for (int m = 0; m < NV; ++m) l[m] *= 0.5;
    double lam = 0.0;
#pragma unroll
    for (int p = 0; p < 6; ++p) lam = fmax(lam, abs_lambda_ua<ND, D, K>(w.u(p, D), w.a(p)));
    // coefficient of p in L_k . F - u_D L_k . Q
    double fp[6], fm[6];
    [[maybe_unused]] double cfa[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      double hcs = 0.0;
      sfor<0, NV>([&](auto mc) {
        constexpr int M = decltype(mc)::value;
        if constexpr ((LM >> M) & 1u) hcs += l[M] * w.q(p, M);
      });
      const double ud = w.u(p, D);
      double hlp = 0.0;
      if constexpr ((LM >> (1 + D)) & 1u) hlp = l[1 + D];
      if constexpr ((LM >> (NV - 1)) & 1u) hlp += ud * l[NV - 1];
      const double hcf = ud * hcs + w.pr(p) * hlp;
      fp[p] = __fma_rn(lam, hcs, hcf);
      fm[p] = __fma_rn(-lam, hcs, hcf);
      if constexpr (CONV == CONV_WFILTER) cfa[p] = 2.0 * hcf;
    }
    double rc;
    if constexpr (CONV == CONV_WFILTER) rc = weno_filter_flux(fp, fm, cfa);
    else rc = recon_pair<CONV, W::STREAM>(fp, fm, ph.eps, ct);
    double r[NV];
    rcol<ND, D, K>(v, ph, r);
    sfor<0, NV>([&](auto mc) {
      constexpr int M = decltype(mc)::value;
      if constexpr ((RM >> M) & 1u) fh[M] += r[M] * rc;
    });
    // LDS window: the next field reloads its points (nothing stays live across fields)
    if constexpr (W::STREAM) asm volatile("" ::: "memory");
  });
}

template <int ND, int D, int CONV, int AVG, bool CACHED = false>
__device__ __forceinline__ void char_flux(const Pt<ND> (&w)[6], const Phys& ph, double fh[ND + 2], double ct,
                                          const RoeF& rf = RoeF{}) {
  char_flux_w<ND, D, CONV, AVG, CACHED>(RegWindow<ND>{w}, ph, fh, ct, rf);
}

}  // namespace osbl
