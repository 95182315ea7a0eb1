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
//   * the eigenvector rows and columns, and the Roe (or simple) average, with
//     forward-mode dual numbers over the averaged state and the two states it
//     is built from,
//   * the constituent relations u_D, p, a of every stencil point.
// Contributions to the six cells of the window accumulate in registers; a
// cell is written once no later interface of the line touches it.
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
// respect to its five arguments, scaled by gbar.
__device__ __forceinline__ void weno5z_grad(const double x[5], double gbar, double g[5]) {
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
  const double iS = 1.0 / (A0 + A1 + A2);
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
}

// The averaged state (rho, a, u_0..u_{ND-1}) as dual numbers over the
// conservative variables of the two states (slots 0..NV-1 left, NV..2NV-1
// right), by the same formulas as interface_state.
template <int ND, int AVG>
__device__ __forceinline__ void average_dual(const double qL[ND + 2], const double qR[ND + 2], const Phys& ph,
                                             Dual<2 * (ND + 2)>& rho, Dual<2 * (ND + 2)>& a,
                                             Dual<2 * (ND + 2)> (&u)[ND]) {
  constexpr int NV = ND + 2, N = 2 * NV;
  Dual<N> L[NV], R[NV];
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    L[m] = mkvar<N>(qL[m], m);
    R[m] = mkvar<N>(qR[m], NV + m);
  }
  Dual<N> uL[ND], uR[ND];
  const Dual<N> irL = 1.0 / L[0], irR = 1.0 / R[0];
  Dual<N> keL = mkc<N>(0.0), keR = mkc<N>(0.0);
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    uL[m] = L[1 + m] * irL;
    uR[m] = R[1 + m] * irR;
    keL = keL + uL[m] * uL[m];
    keR = keR + uR[m] * uR[m];
  }
  const Dual<N> pL = ph.gm1 * (L[NV - 1] - 0.5 * L[0] * keL);
  const Dual<N> pR = ph.gm1 * (R[NV - 1] - 0.5 * R[0] * keR);
  if constexpr (AVG == OSBL_AVG_ROE) {
    const Dual<N> sL = dsqrt(L[0]), sR = dsqrt(R[0]);
    const Dual<N> inv = 1.0 / (sR + sL);
    rho = sL * sR;
    Dual<N> ke = mkc<N>(0.0);
#pragma unroll
    for (int m = 0; m < ND; ++m) {
      u[m] = (sL * uL[m] + sR * uR[m]) * inv;
      ke = ke + u[m] * u[m];
    }
    const Dual<N> hh = ((L[NV - 1] + pL) / sL + (R[NV - 1] + pR) / sR) * inv;
    a = dsqrt(ph.gm1 * (hh - 0.5 * ke));
  } else {
    rho = 0.5 * (L[0] + R[0]);
#pragma unroll
    for (int m = 0; m < ND; ++m) u[m] = 0.5 * (uL[m] + uR[m]);
    const Dual<N> aL = dsqrt(ph.gama * pL * irL), aR = dsqrt(ph.gama * pR * irR);
    a = 0.5 * (aL + aR);
  }
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

// g[p][m] += (d F_{i+1/2} / d Q_p)^T nu for the window w[0..5] = Q_{i-2..i+3}
template <int ND, int D, int AVG>
__device__ __forceinline__ void char_flux_vjp(const Pt<ND> (&w)[6], const double nu[ND + 2], const Phys& ph,
                                              double (&g)[6][ND + 2]) {
  constexpr int NV = ND + 2, NS = ND + 2;  // averaged symbols: rho, a, u_0..u_{ND-1}
  Avg<ND> v;
  interface_state<ND, AVG>(v, w[2], w[3], ph);
  AvgT<ND, Dual<NS>> vd;
  vd.rho = mkvar<NS>(v.rho, 0);
  vd.a = mkvar<NS>(v.a, 1);
#pragma unroll
  for (int m = 0; m < ND; ++m) vd.u[m] = mkvar<NS>(v.u[m], 2 + m);
  finish_avg<ND>(vd, ph);
  double nvs[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) nvs[s] = 0.0;
  double nud[6], npp[6], na[6];
#pragma unroll
  for (int p = 0; p < 6; ++p) { nud[p] = 0.0; npp[p] = 0.0; na[p] = 0.0; }

  sfor<0, NV>([&](auto kc) {
    constexpr int K = decltype(kc)::value;
    constexpr unsigned LM = lmask<ND, D, K>();
    constexpr unsigned RM = rmask<ND, D, K>();
    Dual<NS> ld[NV], rd[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) { ld[m] = mkc<NS>(0.0); rd[m] = mkc<NS>(0.0); }
    lrow<ND, D, K>(vd, ph, ld);
    rcol<ND, D, K>(vd, ph, rd);
    double l[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) l[m] = ld[m].v;
    // lambda_max and its first maximising point
    double lam = 0.0, dlu = 0.0, dla = 0.0;
    int pst = -1;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
      double du_, da_;
      const double lp = lam_point<ND, K>(w[p].u[D], w[p].a, du_, da_);
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
        if ((LM >> m) & 1u) c += l[m] * w[p].q[m];
      double lpp = 0.0;
      if constexpr ((LM >> (1 + D)) & 1u) lpp = l[1 + D];
      if constexpr ((LM >> (NV - 1)) & 1u) lpp += w[p].u[D] * l[NV - 1];
      const double cf = w[p].u[D] * c + w[p].p * lpp;
      cs[p] = c;
      lpv[p] = lpp;
      fp[p] = 0.5 * (cf + lam * c);
      fm[p] = 0.5 * (cf - lam * c);
    }
    const double rc = weno5z(fp[0], fp[1], fp[2], fp[3], fp[4]) + weno5z(fm[5], fm[4], fm[3], fm[2], fm[1]);
    // F += r rc
    double nrc = 0.0;
#pragma unroll
    for (int m = 0; m < NV; ++m)
      if ((RM >> m) & 1u) {
        nrc += rd[m].v * nu[m];
#pragma unroll
        for (int s = 0; s < NS; ++s) nvs[s] += rd[m].d[s] * rc * nu[m];
      }
    // reconstructions
    double gp[5], gm[5];
    weno5z_grad(fp, nrc, gp);
    const double fmr[5] = {fm[5], fm[4], fm[3], fm[2], fm[1]};
    weno5z_grad(fmr, nrc, gm);
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
      ncs += w[p].u[D] * ncf;
      nud[p] += cs[p] * ncf;
      npp[p] += lpv[p] * ncf;
      if constexpr ((LM >> (1 + D)) & 1u) nl[1 + D] += w[p].p * ncf;
      if constexpr ((LM >> (NV - 1)) & 1u) {
        nl[NV - 1] += w[p].p * w[p].u[D] * ncf;
        nud[p] += w[p].p * l[NV - 1] * ncf;
      }
      // cs = sum_m l_m q_m
#pragma unroll
      for (int m = 0; m < NV; ++m)
        if ((LM >> m) & 1u) {
          g[p][m] += l[m] * ncs;
          nl[m] += w[p].q[m] * ncs;
        }
    }
    // lambda_max
#pragma unroll
    for (int p = 0; p < 6; ++p)
      if (p == pst) {
        nud[p] += dlu * nlam;
        na[p] += dla * nlam;
      }
#pragma unroll
    for (int m = 0; m < NV; ++m)
      if ((LM >> m) & 1u) {
#pragma unroll
        for (int s = 0; s < NS; ++s) nvs[s] += ld[m].d[s] * nl[m];
      }
  });

  // averaged state -> the two states
  {
    constexpr int N2 = 2 * NV;
    Dual<N2> rho, a, u[ND];
    average_dual<ND, AVG>(w[2].q, w[3].q, ph, rho, a, u);
#pragma unroll
    for (int k = 0; k < N2; ++k) {
      double acc = nvs[0] * rho.d[k] + nvs[1] * a.d[k];
#pragma unroll
      for (int m = 0; m < ND; ++m) acc += nvs[2 + m] * u[m].d[k];
      if (k < NV) g[2][k] += acc;
      else g[3][k - NV] += acc;
    }
  }
  // u_D, p, a of every point -> its conservative variables
#pragma unroll
  for (int p = 0; p < 6; ++p) {
    const double rho = w[p].q[0], ir = 1.0 / rho;
    const double a = w[p].a, pr = w[p].p;
    double uu = 0.0;
#pragma unroll
    for (int m = 0; m < ND; ++m) uu += w[p].u[m] * w[p].u[m];
    // a = sqrt(gama p / rho)
    const double np_ = npp[p] + na[p] * ph.gama * 0.5 / (a * rho);
    double g0 = -na[p] * 0.5 * a * ir;
    // p = (gama-1)(E - |rho u|^2 / (2 rho))
    g0 += np_ * ph.gm1 * 0.5 * uu;
#pragma unroll
    for (int m = 0; m < ND; ++m) g[p][1 + m] += -np_ * ph.gm1 * w[p].u[m];
    g[p][NV - 1] += np_ * ph.gm1;
    // u_D = rhou_D / rho
    g0 += -nud[p] * w[p].u[D] * ir;
    g[p][1 + D] += nud[p] * ir;
    g[p][0] += g0;
    (void)pr;
  }
}

}  // namespace

// One direction of the characteristic residual's VJP on packed periodic
// arrays (n0 x n1 x n2, x fastest).  mu = alpha A + beta B (B optional) is
// formed on the fly.  ACC = false: out1 = c (A + B) + V (and out2 = A + B when
// given); ACC = true: out1 += V.  Each thread owns `seg` cells of one line.
template <int ND, int D, int AVG, bool ACC>
__global__ void __launch_bounds__(128) k_char_vjp(const double* __restrict__ Qs, const double* __restrict__ A,
                                                  const double* __restrict__ B, double alpha, double beta, double c,
                                                  double* __restrict__ out1, double* __restrict__ out2, int n0,
                                                  int n1, int n2, Phys ph, int seg) {
  constexpr int NV = ND + 2;
  const int nn[3] = {n0, n1, n2};
  const long long cells = (long long)n0 * n1 * n2;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int N = nn[D];
  const long long nseg = (N + seg - 1) / seg;
  long long lines = 1;
  for (int e = 0; e < 3; ++e)
    if (e != D) lines *= nn[e];
  if (t >= lines * nseg) return;
  int co[3];
  int s;
  if constexpr (D == 0) {
    co[1] = (int)(t % n1);
    const long long r = t / n1;
    co[2] = (int)(r % n2);
    s = (int)(r / n2);
  } else {
    constexpr int O = (D == 1) ? 2 : 1;
    co[0] = (int)(t % n0);
    const long long r = t / n0;
    co[O] = (int)(r % nn[O]);
    s = (int)(r / nn[O]);
  }
  co[D] = 0;
  const long long base = co[0] + (long long)n0 * (co[1] + (long long)n1 * co[2]);
  const long long sd = D == 0 ? 1 : (D == 1 ? (long long)n0 : (long long)n0 * n1);
  const int c0 = s * seg, c1 = min(c0 + seg, N);
  auto at = [&](int i) { return base + (long long)(((i % N) + N) % N) * sd; };
  auto load = [&](Pt<ND>& pt, double* mu, int i) {
    const long long y = at(i);
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      pt.q[m] = Qs[m * cells + y];
      double v = alpha * A[m * cells + y];
      if (B) v += beta * B[m * cells + y];
      mu[m] = v;
    }
    finish_pt<ND>(pt, ph);
  };
  const double inv = ph.inv_delta[D];
  Pt<ND> w[6];
  double mw[6][NV];
  double acc[6][NV];
#pragma unroll
  for (int p = 0; p < 6; ++p)
#pragma unroll
    for (int m = 0; m < NV; ++m) acc[p][m] = 0.0;
  // window for interface i holds points i-2..i+3; the first interface is c0-3
#pragma unroll
  for (int p = 0; p < 5; ++p) load(w[p], mw[p], c0 - 5 + p);
  for (int i = c0 - 3; i < c1 + 2; ++i) {
    load(w[5], mw[5], i + 3);
    double nu[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) nu[m] = inv * (mw[3][m] - mw[2][m]);
    char_flux_vjp<ND, D, AVG>(w, nu, ph, acc);
    const int cell = i - 2;
    if (cell >= c0) {
      const long long x = at(cell);
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        if constexpr (ACC) {
          out1[m * cells + x] += acc[0][m];
        } else {
          const double ab = A[m * cells + x] + (B ? B[m * cells + x] : 0.0);
          out1[m * cells + x] = c * ab + acc[0][m];
          if (out2) out2[m * cells + x] = ab;
        }
      }
    }
#pragma unroll
    for (int p = 0; p < 5; ++p) {
      w[p] = w[p + 1];
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        mw[p][m] = mw[p + 1][m];
        acc[p][m] = acc[p + 1][m];
      }
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) acc[5][m] = 0.0;
  }
}

template <int ND, int D, int AVG>
static void launch_cv_d(const double* Qs, const double* A, const double* B, double alpha, double beta, double c,
                        double* out1, double* out2, const Grid& g, const Phys& ph, int seg, bool acc,
                        hipStream_t s) {
  long long lines = 1;
  for (int e = 0; e < 3; ++e)
    if (e != D) lines *= g.n[e];
  seg = fit_segment(lines, g.n[D], seg, 4);
  const long long total = lines * ((g.n[D] + seg - 1) / seg);
  const unsigned blocks = (unsigned)((total + 127) / 128);
  if (acc)
    k_char_vjp<ND, D, AVG, true><<<blocks, 128, 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, g.n[0], g.n[1],
                                                         g.n[2], ph, seg);
  else
    k_char_vjp<ND, D, AVG, false><<<blocks, 128, 0, s>>>(Qs, A, B, alpha, beta, c, out1, out2, g.n[0], g.n[1],
                                                          g.n[2], ph, seg);
}

template <int ND, int AVG>
static void launch_cv_nd(const double* Qs, const double* A, const double* B, double alpha, double beta, double c,
                         double* out1, double* out2, const Grid& g, const Phys& ph, int seg, bool acc,
                         hipStream_t s) {
  launch_cv_d<ND, 0, AVG>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, seg, acc, s);
  if constexpr (ND >= 2) launch_cv_d<ND, 1, AVG>(Qs, A, B, alpha, beta, 0.0, out1, nullptr, g, ph, seg, true, s);
  if constexpr (ND >= 3) launch_cv_d<ND, 2, AVG>(Qs, A, B, alpha, beta, 0.0, out1, nullptr, g, ph, seg, true, s);
}

// out1 = c (A + B) + (dR/dQ at Qs)^T (alpha A + beta B) for the WENO5-Z
// characteristic residual, out2 = A + B (optional); with acc, out1 += the
// product instead.  Packed periodic arrays.
void launch_adj_stage_char(int nd, int avg, const double* Qs, const double* A, const double* B, double alpha,
                           double beta, double c, double* out1, double* out2, const Grid& g, const Phys& ph, int seg,
                           bool acc, hipStream_t s) {
  if (avg == OSBL_AVG_ROE) {
    if (nd == 3) launch_cv_nd<3, OSBL_AVG_ROE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, seg, acc, s);
    else if (nd == 2) launch_cv_nd<2, OSBL_AVG_ROE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, seg, acc, s);
    else launch_cv_nd<1, OSBL_AVG_ROE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, seg, acc, s);
  } else {
    if (nd == 3) launch_cv_nd<3, OSBL_AVG_SIMPLE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, seg, acc, s);
    else if (nd == 2) launch_cv_nd<2, OSBL_AVG_SIMPLE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, seg, acc, s);
    else launch_cv_nd<1, OSBL_AVG_SIMPLE>(Qs, A, B, alpha, beta, c, out1, out2, g, ph, seg, acc, s);
  }
}

}  // namespace osbl
