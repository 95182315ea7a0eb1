// osbl_adjoint_char.hip -- discrete adjoint of the LLF characteristic WENO5-Z
// residual (shock_capturing.py:357-448, weno.py:283-338, averaging.py:38-114,
// euler_eigensystem.py:57-135) on periodic blocks.
//
// R_i = -sum_d (F_{i+1/2} - F_{i-1/2}) / Delta_d with F_{i+1/2} a function of
// the six states Q_{i-2..i+3}.  Its vector-Jacobian product with weights mu is
// sum over interfaces of J_p^T nu, nu = (mu_{i+1} - mu_i)/Delta_d.  A thread
// slides the forward 6-point window along a grid line; per interface it
// re-evaluates the forward flux of every characteristic field and runs it
// backwards:
//   * the reconstructions through a hand-written reverse of the one-division
//     WENO-Z weights (tau^2 is smooth; no branch),
//   * lambda_max = max_p |lambda(p)| to its first maximising point (the
//     subgradient the forward's max picks),
//   * the characteristic projections CS = L.Q, CF = u_D CS + p (L_{1+D} + u_D L_E),
//   * the eigenvector rows and columns and the Roe (or simple) average, by
//     hand-written reverses (dual numbers over the averaged state for the
//     1-D eigensystem),
//   * the constituent relations u_D, p, a of every stencil point.
// The lanes of a wave take 64 consecutive interfaces of a grid line: the
// chunk's points sit in LDS with their constituent relations, every lane adds
// its interface's contributions to its own LDS column, and each cell gathers
// the six contributions that reach it from its neighbours' columns; no
// partial sums leave the wave.
#include <algorithm>

#include "osbl_kernels.h"

namespace osbl {

namespace {

// forward-mode dual number with N derivative slots
template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ __forceinline__ Dual& operator=(double c) {
    v = c;
#pragma unroll
    for (int i = 0; i < N; ++i) d[i] = 0.0;
    return *this;
  }
};
template <int N>
__device__ __forceinline__ Dual<N> mkc(double c) {
  Dual<N> r;
  r.v = c;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = 0.0;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> mkvar(double c, int slot) {
  Dual<N> r = mkc<N>(c);
  r.d[slot] = 1.0;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a) {
  Dual<N> r;
  r.v = -a.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -a.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
  Dual<N> r;
  const double ib = 1.0 / b.v;
  r.v = a.v * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v += b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(double b, const Dual<N>& a) { return a + b; }
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, double b) {
  Dual<N> r = a;
  r.v -= b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(double b, const Dual<N>& a) {
  Dual<N> r = -a;
  r.v += b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, double b) {
  Dual<N> r;
  r.v = a.v * b;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double b, const Dual<N>& a) { return a * b; }
template <int N>
__device__ __forceinline__ Dual<N> operator/(double a, const Dual<N>& b) {
  Dual<N> r;
  const double ib = 1.0 / b.v;
  r.v = a * ib;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = -r.v * b.d[i] * ib;
  return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, double b) {
  return a * (1.0 / b);
}
template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = sqrt(a.v);
  const double h = 0.5 / r.v;
#pragma unroll
  for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * h;
  return r;
}

__device__ __forceinline__ double sgn(double x) { return (double)((x > 0.0) - (x < 0.0)); }

// Gradient of weno5z(x0..x4) (the one-division WENO-Z of osbl_device.h) with
// respect to its five arguments, scaled by gbar; returns the reconstruction
// itself (the same weights, the candidates' 1/6 applied to each candidate).
__device__ __forceinline__ double weno5z_grad(const double x[5], double gbar, double g[5]) {
  const double fm2 = x[0], fm1 = x[1], f0 = x[2], f1 = x[3], f2 = x[4];
  const double t0 = f0 - 2.0 * f1 + f2, s0 = 3.0 * f0 - 4.0 * f1 + f2;
  const double t1 = fm1 - 2.0 * f0 + f1, s1 = fm1 - f1;
  const double t2 = fm2 - 2.0 * fm1 + f0, s2 = fm2 - 4.0 * fm1 + 3.0 * f0;
  constexpr double C13 = 13.0 / 12.0;
  const double b0 = C13 * t0 * t0 + 0.25 * s0 * s0;
  const double b1 = C13 * t1 * t1 + 0.25 * s1 * s1;
  const double b2 = C13 * t2 * t2 + 0.25 * s2 * s2;
  const double dtau = b0 - b2;  // tau^2 = (b0 - b2)^2
  const double tt = dtau * dtau;
  const double e0 = 1.0e-14 + b0, e1 = 1.0e-14 + b1, e2 = 1.0e-14 + b2;
  const double q0 = e0 * e0, q1 = e1 * e1, q2 = e2 * e2;
  const double A0 = 0.3 * (q0 + tt) * (q1 * q2);
  const double A1 = 0.6 * (q1 + tt) * (q0 * q2);
  const double A2 = 0.1 * (q2 + tt) * (q0 * q1);
  constexpr double S6 = 1.0 / 6.0;
  const double r0 = (2.0 * f0 + 5.0 * f1 - f2) * S6;
  const double r1 = (-fm1 + 5.0 * f0 + 2.0 * f1) * S6;
  const double r2 = (2.0 * fm2 - 7.0 * fm1 + 11.0 * f0) * S6;
  const double iS = recip_pos(A0 + A1 + A2);
  const double W = (A0 * r0 + A1 * r1 + A2 * r2) * iS;
  // reverse
  const double rb0 = A0 * iS * gbar, rb1 = A1 * iS * gbar, rb2 = A2 * iS * gbar;
  const double Ab0 = (r0 - W) * iS * gbar, Ab1 = (r1 - W) * iS * gbar, Ab2 = (r2 - W) * iS * gbar;
  const double q12 = q1 * q2, q02 = q0 * q2, q01 = q0 * q1;
  const double qb0 = Ab0 * 0.3 * q12 + Ab1 * 0.6 * (q1 + tt) * q2 + Ab2 * 0.1 * (q2 + tt) * q1;
  const double qb1 = Ab0 * 0.3 * (q0 + tt) * q2 + Ab1 * 0.6 * q02 + Ab2 * 0.1 * (q2 + tt) * q0;
  const double qb2 = Ab0 * 0.3 * (q0 + tt) * q1 + Ab1 * 0.6 * (q1 + tt) * q0 + Ab2 * 0.1 * q01;
  const double ttb = Ab0 * 0.3 * q12 + Ab1 * 0.6 * q02 + Ab2 * 0.1 * q01;
  double bb0 = 2.0 * e0 * qb0, bb1 = 2.0 * e1 * qb1, bb2 = 2.0 * e2 * qb2;
  const double db = 2.0 * dtau * ttb;
  bb0 += db;
  bb2 -= db;
  const double tb0 = 2.0 * C13 * t0 * bb0, sb0 = 0.5 * s0 * bb0;
  const double tb1 = 2.0 * C13 * t1 * bb1, sb1 = 0.5 * s1 * bb1;
  const double tb2 = 2.0 * C13 * t2 * bb2, sb2 = 0.5 * s2 * bb2;
  // fm2
  g[0] = tb2 + sb2 + rb2 * (2.0 * S6);
  // fm1
  g[1] = tb1 + sb1 - 2.0 * tb2 - 4.0 * sb2 - rb1 * S6 - rb2 * (7.0 * S6);
  // f0
  g[2] = tb0 + 3.0 * sb0 - 2.0 * tb1 + tb2 + 3.0 * sb2 + rb0 * (2.0 * S6) + rb1 * (5.0 * S6) + rb2 * (11.0 * S6);
  // f1
  g[3] = -2.0 * tb0 - 4.0 * sb0 + tb1 - sb1 + rb0 * (5.0 * S6) + rb1 * (2.0 * S6);
  // f2
  g[4] = tb0 + sb0 - rb0 * S6;
  return W;
}

// Adjoints of the averaged state's symbols (AvgT's members).
template <int ND>
struct AvgBar {
  double rho = 0.0, a = 0.0, u[ND] = {}, ia = 0.0, ir = 0.0, phi = 0.0, alpha = 0.0, bta = 0.0, Hh = 0.0;
};

// Reverse of lrow<ND, D, K> (ND >= 2): b += (d l / d v)^T nl.
template <int ND, int D, int K>
__device__ __forceinline__ void lrow_rev(const Avg<ND>& v, const Phys& ph, const double nl[ND + 2], AvgBar<ND>& b) {
  constexpr int NV = ND + 2;
  const double gm1 = ph.gm1;
  constexpr bool dir_row = (ND == 3) ? (K == D) : (K == 0);
  constexpr bool shear = (ND == 3) ? (K < 3 && K != D) : (K == 1);
  if constexpr (dir_row) {
    // l0 = 1 - phi ia^2, l_{1+m} = gm1 u_m ia^2, l_E = -gm1 ia^2
    const double ia2 = v.ia * v.ia;
    double ia2b = -v.phi * nl[0] - gm1 * nl[NV - 1];
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      ia2b += gm1 * v.u[m] * nl[1 + m];
      b.u[m] += gm1 * ia2 * nl[1 + m];
    }
    b.phi += -ia2 * nl[0];
    b.ia += 2.0 * v.ia * ia2b;
  } else if constexpr (shear) {
    // l0 = -s u_c ir, l_{1+c} = s ir
    constexpr int c = (ND == 3) ? lc3(D, K) : 1 - D;
    constexpr double sgn_ = (ND == 3) ? (double)eps3(D, K, lc3(D, K)) : (D == 0 ? -1.0 : 1.0);
    b.u[c] += -sgn_ * v.ir * nl[0];
    b.ir += -sgn_ * v.u[c] * nl[0] + sgn_ * nl[1 + c];
  } else {
    // l0 = bta (phi - sg u_D a), l_{1+m} = -bta (gm1 u_m - [m == D] sg a), l_E = bta gm1
    constexpr double sg = (K == NV - 2) ? 1.0 : -1.0;
    double btab = nl[0] * (v.phi - sg * v.u[D] * v.a) + gm1 * nl[NV - 1];
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      btab -= nl[1 + m] * (gm1 * v.u[m] - (m == D ? sg * v.a : 0.0));
      b.u[m] += -v.bta * gm1 * nl[1 + m];
    }
    b.bta += btab;
    b.phi += v.bta * nl[0];
    b.u[D] += -v.bta * sg * v.a * nl[0];
    b.a += -v.bta * sg * v.u[D] * nl[0] + v.bta * sg * nl[1 + D];
  }
}

// Reverse of rcol<ND, D, K> (ND >= 2): b += (d r / d v)^T nr.
template <int ND, int D, int K>
__device__ __forceinline__ void rcol_rev(const Avg<ND>& v, const Phys& ph, const double nr[ND + 2], AvgBar<ND>& b) {
  constexpr int NV = ND + 2;
  const double igm1 = ph.igm1;
  constexpr bool dir_row = (ND == 3) ? (K == D) : (K == 0);
  constexpr bool shear = (ND == 3) ? (K < 3 && K != D) : (K == 1);
  if constexpr (dir_row) {
    // r0 = 1, r_{1+m} = u_m, r_E = phi / (gama - 1)
#pragma unroll
    for (int m = 0; m < ND; ++m) b.u[m] += nr[1 + m];
    b.phi += igm1 * nr[NV - 1];
  } else if constexpr (shear) {
    // r_{1+c} = s rho, r_E = s rho u_c
    constexpr int c = (ND == 3) ? lc3(D, K) : 1 - D;
    constexpr double sgn_ = (ND == 3) ? (double)eps3(D, K, lc3(D, K)) : (D == 0 ? -1.0 : 1.0);
    b.rho += sgn_ * nr[1 + c] + sgn_ * v.u[c] * nr[NV - 1];
    b.u[c] += sgn_ * v.rho * nr[NV - 1];
  } else {
    // r0 = alpha, r_{1+m} = alpha (u_m + [m == D] sg a), r_E = alpha ((phi + a^2)/(gama-1) + sg u_D a)
    constexpr double sg = (K == NV - 2) ? 1.0 : -1.0;
    const double rE = (v.phi + v.a * v.a) * igm1 + sg * v.u[D] * v.a;
    double alb = nr[0] + rE * nr[NV - 1];
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      alb += (v.u[m] + (m == D ? sg * v.a : 0.0)) * nr[1 + m];
      b.u[m] += v.alpha * nr[1 + m];
    }
    b.alpha += alb;
    b.a += v.alpha * sg * nr[1 + D] + v.alpha * (2.0 * v.a * igm1 + sg * v.u[D]) * nr[NV - 1];
    b.phi += v.alpha * igm1 * nr[NV - 1];
    b.u[D] += v.alpha * sg * v.a * nr[NV - 1];
  }
}

// Reverse of finish_avg: folds the adjoints of ia, ir, phi, alpha, bta, Hh
// into those of rho, a, u.
template <int ND>
__device__ __forceinline__ void finish_avg_rev(const Avg<ND>& v, const Phys& ph, AvgBar<ND>& b) {
  constexpr double RS2 = 0.70710678118654752440;
  b.rho += v.ia * RS2 * b.alpha;
  b.ia += v.rho * RS2 * b.alpha + v.ir * RS2 * b.bta;
  b.ir += v.ia * RS2 * b.bta;
  b.a += 2.0 * v.a * ph.igm1 * b.Hh;
#pragma unroll
  for (int m = 0; m < ND; ++m) b.u[m] += (ph.gm1 * b.phi + b.Hh) * v.u[m];
  b.a += -v.ia * v.ia * b.ia;
  b.rho += -v.ir * v.ir * b.ir;
}

// Reverse of interface_state's average: the adjoints (rho, a, u) of the
// averaged state added to those of the two points' conservative variables.
template <int ND, int AVG>
__device__ __forceinline__ void average_rev(const Pt<ND>& L, const Pt<ND>& R, const Phys& ph, const AvgBar<ND>& b,
                                            double (&gL)[ND + 2], double (&gR)[ND + 2]) {
  constexpr int NV = ND + 2;
  double pLb, pRb, rLb, rRb, uLb[ND], uRb[ND];
  if constexpr (AVG == OSBL_AVG_ROE) {
    const double sL = sqrt_pos(L.q[0]), sR = sqrt_pos(R.q[0]);
    const double isL = recip_pos(sL), isR = recip_pos(sR);
    const double inv = recip_pos(sL + sR);
    const double HL = (L.q[NV - 1] + L.p) * isL, HR = (R.q[NV - 1] + R.p) * isR;
    double ke = 0.0, u[ND];
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      u[m] = (sL * L.u[m] + sR * R.u[m]) * inv;
      ke += u[m] * u[m];
    }
    const double hh = (HL + HR) * inv;
    const double a = sqrt_pos(ph.gm1 * (hh - 0.5 * ke));
    // a = sqrt(gm1 (hh - ke/2))
    const double tb = quot_pos(b.a * ph.gm1, 2.0 * a);
    double invb = tb * (HL + HR);
    const double Hb = tb * inv;
    double sLb = b.rho * sR, sRb = b.rho * sL;
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      const double ub = b.u[m] - tb * u[m];
      sLb += ub * L.u[m] * inv;
      sRb += ub * R.u[m] * inv;
      uLb[m] = ub * sL * inv;
      uRb[m] = ub * sR * inv;
      invb += ub * (sL * L.u[m] + sR * R.u[m]);
    }
    const double sb = -inv * inv * invb;
    sLb += sb - Hb * HL * isL;
    sRb += sb - Hb * HR * isR;
    // H = (E + p) / s
    pLb = Hb * isL;
    pRb = Hb * isR;
    gL[NV - 1] += Hb * isL;
    gR[NV - 1] += Hb * isR;
    rLb = 0.5 * sLb * isL;
    rRb = 0.5 * sRb * isR;
  } else {
    // rho = (rho_L + rho_R)/2, u = (u_L + u_R)/2, a = (a_L + a_R)/2, a = sqrt(gama p / rho)
    rLb = 0.5 * b.rho;
    rRb = 0.5 * b.rho;
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      uLb[m] = 0.5 * b.u[m];
      uRb[m] = 0.5 * b.u[m];
    }
    const double aLb = 0.5 * b.a, aRb = 0.5 * b.a;
    pLb = quot_pos(aLb * ph.gama, 2.0 * L.a * L.q[0]);
    pRb = quot_pos(aRb * ph.gama, 2.0 * R.a * R.q[0]);
    rLb += quot_pos(-aLb * L.a, 2.0 * L.q[0]);
    rRb += quot_pos(-aRb * R.a, 2.0 * R.q[0]);
  }
  // p = gm1 (E - rho |u|^2 / 2), u_m = (rho u_m) / rho
  double keL = 0.0, keR = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    keL += L.u[m] * L.u[m];
    keR += R.u[m] * R.u[m];
  }
  gL[NV - 1] += ph.gm1 * pLb;
  gR[NV - 1] += ph.gm1 * pRb;
  rLb += -0.5 * ph.gm1 * keL * pLb;
  rRb += -0.5 * ph.gm1 * keR * pRb;
  const double irL = recip_pos(L.q[0]), irR = recip_pos(R.q[0]);
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double uLt = uLb[m] - ph.gm1 * L.q[0] * L.u[m] * pLb;
    const double uRt = uRb[m] - ph.gm1 * R.q[0] * R.u[m] * pRb;
    gL[1 + m] += uLt * irL;
    gR[1 + m] += uRt * irR;
    rLb += -uLt * L.u[m] * irL;
    rRb += -uRt * R.u[m] * irR;
  }
  gL[0] += rLb;
  gR[0] += rRb;
}

// |lambda_K| at a point and its derivative with respect to (u_D, a)
template <int ND, int K>
__device__ __forceinline__ double lam_point(double ud, double a, double& dud, double& da) {
  double l;
  if constexpr (ND == 1) {
    if constexpr (K == 0) { l = ud - a; dud = 1.0; da = -1.0; }
    else if constexpr (K == 1) { l = ud; dud = 1.0; da = 0.0; }
    else { l = ud + a; dud = 1.0; da = 1.0; }
  } else {
    constexpr int NV = ND + 2;
    if constexpr (K == NV - 2) { l = ud + a; dud = 1.0; da = 1.0; }
    else if constexpr (K == NV - 1) { l = ud - a; dud = 1.0; da = -1.0; }
    else { l = ud; dud = 1.0; da = 0.0; }
  }
  const double s = sgn(l);
  dud *= s;
  da *= s;
  return fabs(l);
}

// Accumulator of the VJP's contributions g[p][m] (window point p, variable m)
// in a lane's own LDS column: an atomic add that returns nothing, so no
// register waits for the old value (every lane owns its column: no contention)
template <int ND>
struct LdsAcc {
  double* s;   // slot (0, 0) of this lane
  int stride;  // doubles between slots
  __device__ __forceinline__ void add(int p, int m, double v) const { atomicAdd(&s[(p * (ND + 2) + m) * stride], v); }
};

// g[p][m] += (d F_{i+1/2} / d Q_p)^T nu for the window w = Q_{i-2..i+3}
// (points 0..5, read through the LdsWindow accessors)
template <int ND, int D, int AVG, class W, class G>
__device__ __forceinline__ void char_flux_vjp(const W& w, const double nu[ND + 2], const Phys& ph, const G& g) {
  constexpr int NV = ND + 2, NS = ND + 2;  // averaged symbols: rho, a, u_0..u_{ND-1}
  Avg<ND> v;
  {
    Pt<ND> L, R;
    w.point(2, L);
    w.point(3, R);
    interface_state<ND, AVG>(v, L, R, ph);
  }
  // adjoints of the averaged state: by the hand-written reverse of the
  // eigenvector entries (ND >= 2) or by dual numbers (the 1-D eigensystem)
  AvgBar<ND> b;
  [[maybe_unused]] AvgT<ND, Dual<NS>> vd;
  [[maybe_unused]] double nvs[NS];
  if constexpr (ND == 1) {
    vd.rho = mkvar<NS>(v.rho, 0);
    vd.a = mkvar<NS>(v.a, 1);
#pragma unroll
    for (int m = 0; m < ND; ++m) vd.u[m] = mkvar<NS>(v.u[m], 2 + m);
    finish_avg<ND>(vd, ph);
#pragma unroll
    for (int q = 0; q < NS; ++q) nvs[q] = 0.0;
  }
  double nud[6], npp[6], na[6];
#pragma unroll
  for (int p = 0; p < 6; ++p) { nud[p] = 0.0; npp[p] = 0.0; na[p] = 0.0; }

  auto field = [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    constexpr unsigned LM = lmask<ND, D, K>();
    constexpr unsigned RM = rmask<ND, D, K>();
    double l[NV], r[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) { l[m] = 0.0; r[m] = 0.0; }
    [[maybe_unused]] Dual<NS> ld[NV], rd[NV];
    if constexpr (ND == 1) {
#pragma unroll
      for (int m = 0; m < NV; ++m) { ld[m] = mkc<NS>(0.0); rd[m] = mkc<NS>(0.0); }
      lrow<ND, D, K>(vd, ph, ld);
      rcol<ND, D, K>(vd, ph, rd);
#pragma unroll
      for (int m = 0; m < NV; ++m) { l[m] = ld[m].v; r[m] = rd[m].v; }
    } else {
      lrow<ND, D, K>(v, ph, l);
      rcol<ND, D, K>(v, ph, r);
    }
    // lambda_max and its first maximising point
    double lam = 0.0, dlu = 0.0, dla = 0.0;
    int pst = -1;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      double du_, da_;
      const double lp = lam_point<ND, K>(w.u(p, D), w.a(p), du_, da_);
      if (lp > lam) {
        lam = lp;
        pst = p;
        dlu = du_;
        dla = da_;
      }
    }
    double cs[6], lpv[6], fp[6], fm[6];
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      double c = 0.0;
#pragma unroll
      for (int m = 0; m < NV; ++m)
        if ((LM >> m) & 1u) c += l[m] * w.q(p, m);
      double lpp = 0.0;
      if constexpr ((LM >> (1 + D)) & 1u) lpp = l[1 + D];
      if constexpr ((LM >> (NV - 1)) & 1u) lpp += w.u(p, D) * l[NV - 1];
      const double cf = w.u(p, D) * c + w.pr(p) * lpp;
      cs[p] = c;
      lpv[p] = lpp;
      fp[p] = 0.5 * (cf + lam * c);
      fm[p] = 0.5 * (cf - lam * c);
    }
    // F += r rc: rc's adjoint, then the two reconstructions (values and
    // gradients from one evaluation of their weights each)
    double nrc = 0.0;
#pragma unroll
    for (int m = 0; m < NV; ++m)
      if ((RM >> m) & 1u) nrc += r[m] * nu[m];
    double gp[5], gm[5];
    const double fmr[5] = {fm[5], fm[4], fm[3], fm[2], fm[1]};
    const double rc = weno5z_grad(fp, nrc, gp) + weno5z_grad(fmr, nrc, gm);
    double nr[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) nr[m] = ((RM >> m) & 1u) ? rc * nu[m] : 0.0;
    if constexpr (ND == 1) {
#pragma unroll
      for (int m = 0; m < NV; ++m)
        if ((RM >> m) & 1u) {
#pragma unroll
          for (int q = 0; q < NS; ++q) nvs[q] += rd[m].d[q] * nr[m];
        }
    } else {
      rcol_rev<ND, D, K>(v, ph, nr, b);
    }
    double nl[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) nl[m] = 0.0;
    double nlam = 0.0;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      const double a_ = p < 5 ? gp[p] : 0.0;
      const double b_ = p > 0 ? gm[5 - p] : 0.0;
      const double ncf = 0.5 * (a_ + b_);
      double ncs = 0.5 * lam * (a_ - b_);
      nlam += 0.5 * cs[p] * (a_ - b_);
      // cf = u_D cs + p lp, lp = l_{1+D} + u_D l_E
      ncs += w.u(p, D) * ncf;
      nud[p] += cs[p] * ncf;
      npp[p] += lpv[p] * ncf;
      if constexpr ((LM >> (1 + D)) & 1u) nl[1 + D] += w.pr(p) * ncf;
      if constexpr ((LM >> (NV - 1)) & 1u) {
        nl[NV - 1] += w.pr(p) * w.u(p, D) * ncf;
        nud[p] += w.pr(p) * l[NV - 1] * ncf;
      }
      // cs = sum_m l_m q_m
#pragma unroll
      for (int m = 0; m < NV; ++m)
        if ((LM >> m) & 1u) {
          g.add(p, m, l[m] * ncs);
          nl[m] += w.q(p, m) * ncs;
        }
    }
    // lambda_max
#pragma unroll
    for (int p = 0; p < 6; ++p)
      if (p == pst) {
        nud[p] += dlu * nlam;
        na[p] += dla * nlam;
      }
    if constexpr (ND == 1) {
#pragma unroll
      for (int m = 0; m < NV; ++m)
        if ((LM >> m) & 1u) {
#pragma unroll
          for (int q = 0; q < NS; ++q) nvs[q] += ld[m].d[q] * nl[m];
        }
    } else {
      lrow_rev<ND, D, K>(v, ph, nl, b);
    }
    // LDS window: the next field reloads its points (its work does not
    // overlap this one's, which keeps the live registers to one field's)
    if constexpr (W::STREAM) asm volatile("" ::: "memory");
  };
  // one characteristic field at a time (a loop the compiler keeps rolled,
  // so no field's work is scheduled into another's registers)
#pragma nounroll
  for (int k = 0; k < NV; ++k) {
    switch (k) {
      case 0: field(std::integral_constant<int, 0>{}); break;
      case 1: field(std::integral_constant<int, 1>{}); break;
      case 2: field(std::integral_constant<int, 2>{}); break;
      case 3:
        if constexpr (NV > 3) field(std::integral_constant<int, (NV > 3 ? 3 : 0)>{});
        break;
      default:
        if constexpr (NV > 4) field(std::integral_constant<int, (NV > 4 ? 4 : 0)>{});
        break;
    }
  }

  // averaged state -> the two states
  if constexpr (ND == 1) {
    b.rho = nvs[0];
    b.a = nvs[1];
    b.u[0] = nvs[2];
  } else {
    finish_avg_rev<ND>(v, ph, b);
  }
  {
    Pt<ND> L, R;
    w.point(2, L);
    w.point(3, R);
    double gl[NV], gr[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) { gl[m] = 0.0; gr[m] = 0.0; }
    average_rev<ND, AVG>(L, R, ph, b, gl, gr);
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      g.add(2, m, gl[m]);
      g.add(3, m, gr[m]);
    }
  }
  // u_D, p, a of every point -> its conservative variables
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    const double rho = w.q(p, 0), ir = recip_pos(rho);
    const double a = w.a(p);
    double uu = 0.0;
#pragma unroll
    for (int m = 0; m < ND; ++m) uu += w.u(p, m) * w.u(p, m);
    // a = sqrt(gama p / rho)
    const double np_ = npp[p] + quot_pos(na[p] * ph.gama * 0.5, a * rho);
    double g0 = -na[p] * 0.5 * a * ir;
    // p = (gama-1)(E - |rho u|^2 / (2 rho))
    g0 += np_ * ph.gm1 * 0.5 * uu;
#pragma unroll
    for (int m = 0; m < ND; ++m) g.add(p, 1 + m, -np_ * ph.gm1 * w.u(p, m));
    g.add(p, NV - 1, np_ * ph.gm1);
    // u_D = rhou_D / rho
    g0 += -nud[p] * w.u(p, D) * ir;
    g.add(p, 1 + D, nud[p] * ir);
    g.add(p, 0, g0);
  }
}

}  // namespace

// origin and stride of grid line `line` along D (lines numbered with the
// lower remaining direction fastest)
template <int D>
__device__ __forceinline__ void vjp_line(long long line, int n0, int n1, long long& base, long long& sd) {
  int co[3];
  if constexpr (D == 0) {
    co[1] = (int)(line % n1);
    co[2] = (int)(line / n1);
  } else {
    constexpr int O = (D == 1) ? 2 : 1;
    co[0] = (int)(line % n0);
    co[O] = (int)(line / n0);
  }
  co[D] = 0;
  base = co[0] + (long long)n0 * (co[1] + (long long)n1 * co[2]);
  sd = D == 0 ? 1 : (D == 1 ? (long long)n0 : (long long)n0 * n1);
}

// One direction of the characteristic residual's VJP with the lanes of a wave
// along the grid line: a wave owns a chunk of 64 consecutive interfaces
// I = a .. a+63 of one line.  The chunk's 69 points (a-2 .. a+66, periodic
// images) are staged in LDS with their constituent relations, lane l
// evaluates the VJP of interface a+l from LDS, and the contributions of the
// six interfaces j-3 .. j+2 that reach cell j = a+l are gathered by
// cross-lane shuffles; lanes 3..61 hold complete cells, so consecutive
// chunks of a line start 59 cells apart (five interfaces are evaluated
// twice) and no partial sums leave the wave.
//   ACC = false: out1 = c (A + B) + V (and out2 = A + B when given);
//   ACC = true:  out1 += V.
constexpr int VJW = 2;         // waves per block
constexpr int VJP = 64 + 5;    // points staged per chunk
constexpr int VJSTEP = 59;     // cells completed per chunk

// The work of wave `wid` of one direction's VJP (chunk wid / lines of line
// wid % lines); S and G are the wave's LDS staging window and contribution
// columns.  OUT 0: out1 = c (A + B) + V (and out2 = A + B when given);
// OUT 1: out1 += V; OUT 2: out1 = V.
template <int ND, int D, int AVG, int OUT>
__device__ __forceinline__ void vjp_chunk(long long wid, int lane, double (*S)[VJP], double (*G)[64],
                                          const double* __restrict__ Qs, const double* __restrict__ A,
                                          const double* __restrict__ B, double alpha, double beta, double c,
                                          double* __restrict__ out1, double* __restrict__ out2, int n0, int n1,
                                          int n2, const Phys& ph) {
  constexpr int NV = ND + 2;
  const int nn[3] = {n0, n1, n2};
  const long long cells = (long long)n0 * n1 * n2;
  const int N = nn[D];
  long long lines = 1;
  for (int e = 0; e < 3; ++e)
    if (e != D) lines *= nn[e];
  const long long line = wid % lines;
  const int chunk = (int)(wid / lines);
  long long base, sd;
  vjp_line<D>(line, n0, n1, base, sd);
  auto at = [&](int i) { return base + (long long)(((i % N) + N) % N) * sd; };
  const int a0 = chunk * VJSTEP - 3;  // interface of lane 0; lane 3 completes cell chunk * VJSTEP
  auto stage = [&](int slot, int i) {
    Pt<ND> pt;
    const long long y = at(i);
#pragma unroll
    for (int m = 0; m < NV; ++m) pt.q[m] = Qs[m * cells + y];
    finish_pt<ND>(pt, ph);
#pragma unroll
    for (int m = 0; m < NV; ++m) S[m][slot] = pt.q[m];
#pragma unroll
    for (int m = 0; m < ND; ++m) S[NV + m][slot] = pt.u[m];
    S[NV + ND][slot] = pt.p;
    S[NV + ND + 1][slot] = pt.a;
  };
  // slot s holds point a0 - 2 + s
  stage(lane, a0 - 2 + lane);
  if (lane < VJP - 64) stage(64 + lane, a0 + 62 + lane);
  // the interface's weights nu = (mu_{I+1} - mu_I) / Delta_D, mu = alpha A + beta B
  const int I = a0 + lane;
  double nu[NV];
  {
    const long long y0 = at(I), y1 = at(I + 1);
    const double inv = ph.inv_delta[D];
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      double m0 = alpha * A[m * cells + y0], m1 = alpha * A[m * cells + y1];
      if (B) {
        m0 += beta * B[m * cells + y0];
        m1 += beta * B[m * cells + y1];
      }
      nu[m] = inv * (m1 - m0);
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int r = 0; r < 6 * NV; ++r) G[r][lane] = 0.0;
  char_flux_vjp<ND, D, AVG>(LdsWindow<ND>{&S[0][lane], VJP}, nu, ph, LdsAcc<ND>{&G[0][lane], 64});
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // cell j = a0 + lane receives slot k of interface j + 2 - k (lane + 2 - k)
  const int j = a0 + lane;
  if (lane < 3 || lane > 61 || j >= N) return;
  double v[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) v[m] = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int m = 0; m < NV; ++m) v[m] += G[k * NV + m][lane + 2 - k];
  const long long x = base + (long long)j * sd;
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    if constexpr (OUT == 1) {
      out1[m * cells + x] += v[m];
    } else if constexpr (OUT == 2) {
      out1[m * cells + x] = v[m];
    } else {
      const double ab = A[m * cells + x] + (B ? B[m * cells + x] : 0.0);
      out1[m * cells + x] = c * ab + v[m];
      if (out2) out2[m * cells + x] = ab;
    }
  }
}

template <int ND, int D, int AVG, bool ACC>
__global__ void __launch_bounds__(64 * VJW) k_char_vjp_lanes(const double* __restrict__ Qs,
                                                             const double* __restrict__ A,
                                                             const double* __restrict__ B, double alpha, double beta,
                                                             double c, double* __restrict__ out1,
                                                             double* __restrict__ out2, int n0, int n1, int n2,
                                                             Phys ph) {
  constexpr int NV = ND + 2;
  constexpr int NF = NV + ND + 2;  // q, u, p, a
  __shared__ double sh[VJW][NF][VJP];
  __shared__ double ga[VJW][6 * NV][64];  // each lane's contributions, slot (p, m) at row p * NV + m
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int nn[3] = {n0, n1, n2};
  long long lines = 1;
  for (int e = 0; e < 3; ++e)
    if (e != D) lines *= nn[e];
  const long long waves = lines * ((nn[D] + VJSTEP - 1) / VJSTEP);
  // adjacent waves take adjacent lines (for y and z: adjacent x, sharing cache
  // lines); blocks go to the eight XCDs in turn, so consecutive blocks of one
  // XCD take consecutive waves and the shared lines are fetched into one L2
  const unsigned nb = gridDim.x, bl = blockIdx.x;
  const unsigned lb = (nb % 8 == 0) ? (bl % 8) * (nb / 8) + bl / 8 : bl;
  const long long wid = (long long)lb * VJW + wv;
  if (wid >= waves) return;  // whole waves only: no block-wide barrier in vjp_chunk
  vjp_chunk<ND, D, AVG, ACC ? 1 : 0>(wid, lane, sh[wv], ga[wv], Qs, A, B, alpha, beta, c, out1, out2, n0, n1, n2,
                                     ph);
}

// Both directions of a 2-D VJP in one launch, so that the y waves fill the
// SIMDs the x waves leave idle in their last round (one wave per SIMD):
// x waves as k_char_vjp_lanes<2, 0, AVG, false> (out1, out2), y waves write
// their V to vy, which k_add_packed then adds to out1.  With both block
// counts multiples of 8, each XCD takes an eighth of the x blocks and then an
// eighth of the y blocks, in order; otherwise blocks run in launch order.
template <int AVG>
__global__ void __launch_bounds__(64 * VJW) k_char_vjp_xy(const double* __restrict__ Qs, const double* __restrict__ A,
                                                          const double* __restrict__ B, double alpha, double beta,
                                                          double c, double* __restrict__ out1,
                                                          double* __restrict__ out2, double* __restrict__ vy, int n0,
                                                          int n1, int bx, int by, Phys ph) {
  constexpr int NV = 4, NF = 8;
  __shared__ double sh[VJW][NF][VJP];
  __shared__ double ga[VJW][6 * NV][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const unsigned bl = blockIdx.x;
  int dir;
  long long lb;
  if (bx % 8 == 0 && by % 8 == 0) {
    const int px = bx / 8, py = by / 8, slot = (int)(bl / 8), xcd = (int)(bl % 8);
    dir = slot < px ? 0 : 1;
    lb = dir == 0 ? (long long)xcd * px + slot : (long long)xcd * py + (slot - px);
  } else {
    dir = (int)bl < bx ? 0 : 1;
    lb = dir == 0 ? (long long)bl : (long long)bl - bx;
  }
  const long long wid = lb * VJW + wv;
  if (dir == 0) {
    if (wid >= (long long)n1 * ((n0 + VJSTEP - 1) / VJSTEP)) return;
    vjp_chunk<2, 0, AVG, 0>(wid, lane, sh[wv], ga[wv], Qs, A, B, alpha, beta, c, out1, out2, n0, n1, 1, ph);
  } else {
    if (wid >= (long long)n0 * ((n1 + VJSTEP - 1) / VJSTEP)) return;
    vjp_chunk<2, 1, AVG, 2>(wid, lane, sh[wv], ga[wv], Qs, A, B, alpha, beta, 0.0, vy, nullptr, n0, n1, 1, ph);
  }
}

__global__ void __launch_bounds__(256) k_add_packed(double* __restrict__ out, const double* __restrict__ v, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    out[i] += v[i];
}

template <int ND, int D, int AVG>
static void launch_cv_d(const double* Qs, const double* A, const double* B, double alpha, double beta, double c,
                        double* out1, double* out2, const Grid& g, const Phys& ph, bool acc, hipStream_t s) {
  long long lines = 1;
  for (int e = 0; e < 3; ++e)
    if (e != D) lines *= g.n[e];
  const long long waves = lines * ((g.n[D] + VJSTEP - 1) / VJSTEP);
  const unsigned blocks = (unsigned)((waves + VJW - 1) / VJW);
  if (acc)
    k_char_vjp_lanes<ND, D, AVG, true><<<blocks, 64 * VJW, 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, g.n[0],
                                                                     g.n[1], g.n[2], ph);
  else
    k_char_vjp_lanes<ND, D, AVG, false><<<blocks, 64 * VJW, 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, g.n[0],
                                                                      g.n[1], g.n[2], ph);
}

template <int ND, int AVG>
static void launch_cv_nd(const double* Qs, const double* A, const double* B, double alpha, double beta, double c,
                         double* out1, double* out2, const Grid& g, const Phys& ph, bool acc, hipStream_t s,
                         double* vy) {
  if constexpr (ND == 2) {
    if (!acc && vy) {
      const long long wx = (long long)g.n[1] * ((g.n[0] + VJSTEP - 1) / VJSTEP);
      const long long wy = (long long)g.n[0] * ((g.n[1] + VJSTEP - 1) / VJSTEP);
      const int bx = (int)((wx + VJW - 1) / VJW), by = (int)((wy + VJW - 1) / VJW);
      k_char_vjp_xy<AVG><<<bx + by, 64 * VJW, 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, vy, g.n[0], g.n[1], bx,
                                                       by, ph);
      const long long n = (long long)g.n[0] * g.n[1] * 4;
      k_add_packed<<<(unsigned)std::min<long long>((n + 255) / 256, 4096), 256, 0, s>>>(out1, vy, n);
      return;
    }
  }
  launch_cv_d<ND, 0, AVG>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, acc, s);
  if constexpr (ND >= 2) launch_cv_d<ND, 1, AVG>(Qs, A, B, alpha, beta, 0.0, out1, nullptr, g, ph, true, s);
  if constexpr (ND >= 3) launch_cv_d<ND, 2, AVG>(Qs, A, B, alpha, beta, 0.0, out1, nullptr, g, ph, true, s);
}

// out1 = c (A + B) + (dR/dQ at Qs)^T (alpha A + beta B) for the WENO5-Z
// characteristic residual, out2 = A + B (optional); with acc, out1 += the
// product instead.  Packed periodic arrays; out1 aliases neither A nor B.
// vy (optional, 2-D without acc: a packed array of the state's size) lets both
// directions run in one launch, the y part staged in vy; the sums are the
// same, in the same order, as with the two launches.
void launch_adj_stage_char(int nd, int avg, const double* Qs, const double* A, const double* B, double alpha,
                           double beta, double c, double* out1, double* out2, const Grid& g, const Phys& ph, bool acc,
                           hipStream_t s, double* vy) {
  if (avg == OSBL_AVG_ROE) {
    if (nd == 3) launch_cv_nd<3, OSBL_AVG_ROE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, acc, s, vy);
    else if (nd == 2) launch_cv_nd<2, OSBL_AVG_ROE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, acc, s, vy);
    else launch_cv_nd<1, OSBL_AVG_ROE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, acc, s, vy);
  } else {
    if (nd == 3) launch_cv_nd<3, OSBL_AVG_SIMPLE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, acc, s, vy);
    else if (nd == 2) launch_cv_nd<2, OSBL_AVG_SIMPLE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, acc, s, vy);
    else launch_cv_nd<1, OSBL_AVG_SIMPLE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, acc, s, vy);
  }
}

}  // namespace osbl
