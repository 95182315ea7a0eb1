// osbl_boundary.hip -- boundary kernels of non-periodic faces and the
// periodic exchange of one direction.  Each kernel runs over the boundary
// plane of its face with the other directions' halos included
// (bc_core.py:163-171); one thread owns one line normal to the face.
#include "osbl_kernels.h"

namespace osbl {

namespace {

// temperature from the conservative variables at x (ns_physics.py:105-106)
template <int ND>
__device__ __forceinline__ double t_cons(const double* __restrict__ Q, long long fs, int x, const Phys& ph) {
  double m2 = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) m2 += Q[(1 + m) * fs + x] * Q[(1 + m) * fs + x];
  return (Q[(ND + 1) * fs + x] - 0.5 * m2 / Q[x]) * (ph.gama * ph.gm1 * ph.Minf * ph.Minf) / Q[x];
}

}  // namespace

// One face: type OSBL_BC_* (not periodic, not zeroth-order extrapolation),
// nh halo points filled, `st` the storage stride pointing inward, `plane` the
// face's values [nv][n_e2][n_e1] (Dirichlet), Q2 the other state buffer
// (inlet: its rhoE halos are kept equal to those of Q).
template <int ND>
__global__ void __launch_bounds__(256) k_bc_face(double* __restrict__ Q, double* __restrict__ Q2, Grid g, Phys ph,
                                                 int d, int side, int type, int nh, double twall,
                                                 const double* __restrict__ plane) {
  constexpr int NV = ND + 2;
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int n1 = g.n[e1] + 2 * g.h[e1], n2 = g.n[e2] + 2 * g.h[e2];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n1 * n2) return;
  int c[3];
  const int ia = t % n1 - g.h[e1], ib = t / n1 - g.h[e2];
  c[e1] = ia;
  c[e2] = ib;
  c[d] = side == 0 ? 0 : g.n[d] - 1;
  const int x = g.idx(c[0], c[1], c[2]);
  const int st = (side == 0 ? 1 : -1) * g.stride(d);
  const long long fs = g.fs;
  const double gm = ph.gama, M2 = ph.Minf * ph.Minf;
  switch (type) {
    case OSBL_BC_EXTRAP1:  // halo i = 2 q[i-1] - q[i-2], outward (extrapolation.py:42-54)
      for (int v = 0; v < NV; ++v)
        for (int hh = 1; hh <= nh; ++hh)
          Q[v * fs + x - hh * st] = 2.0 * Q[v * fs + x - (hh - 1) * st] - Q[v * fs + x - (hh - 2) * st];
      break;
    case OSBL_BC_SYMMETRY:  // symmetry.py:24-48: mirror images, normal momentum negated; the boundary point stays
      for (int hh = 1; hh <= nh; ++hh)
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const double s = Q[v * fs + x + hh * st];
          Q[v * fs + x - hh * st] = (v == 1 + d) ? -s : s;
        }
      break;
    case OSBL_BC_ZERO_GRADIENT:  // q[0] = q[1]; q[-i] = q[i] (zero_gradient_outlet.py:16-23)
      for (int v = 0; v < NV; ++v) {
        Q[v * fs + x] = Q[v * fs + x + st];
        for (int hh = 1; hh <= nh; ++hh) Q[v * fs + x - hh * st] = Q[v * fs + x + hh * st];
      }
      break;
    case OSBL_BC_ADIABATIC_WALL: {  // adiabatic_wall.py:28-69
#pragma unroll
      for (int m = 0; m < ND; ++m) Q[(1 + m) * fs + x] = 0.0;
      const double T1 = t_cons<ND>(Q, fs, x + st, ph), T2 = t_cons<ND>(Q, fs, x + 2 * st, ph),
                   T3 = t_cons<ND>(Q, fs, x + 3 * st, ph);
      // Carpenter's one-sided dT/dn = 0 at the wall
      const double Tw = (6.0 / 11.0) * (3.0 * T1 + (1.0 / 3.0) * T3 - (3.0 / 2.0) * T2);
      const double pw = Q[x] * Tw / (gm * M2);
      Q[(ND + 1) * fs + x] = pw / ph.gm1;
      for (int hh = 1; hh <= nh; ++hh) {
        Q[x - hh * st] = Q[x + hh * st];
#pragma unroll
        for (int m = 0; m < ND; ++m) Q[(1 + m) * fs + x - hh * st] = -Q[(1 + m) * fs + x + hh * st];
        Q[(ND + 1) * fs + x - hh * st] = Q[(ND + 1) * fs + x + hh * st];
      }
      break;
    }
    case OSBL_BC_ISOTHERMAL_WALL: {  // isothermal_wall.py:33-88
#pragma unroll
      for (int m = 0; m < ND; ++m) Q[(1 + m) * fs + x] = 0.0;
      Q[(ND + 1) * fs + x] = Q[x] * twall / (gm * ph.gm1 * M2);
      // wall pressure from the updated wall state (momentum is zero there)
      const double pw = ph.gm1 * (Q[(ND + 1) * fs + x] - 0.5 * 0.0 / Q[x]);
      const double Ta = t_cons<ND>(Q, fs, x + st, ph);
      for (int hh = 1; hh <= nh; ++hh) {
        const int xi = x + hh * st, xo = x - hh * st;
        double uh[ND];
        double ke = 0.0;
#pragma unroll
        for (int m = 0; m < ND; ++m) {
          uh[m] = Q[(1 + m) * fs + xi] / Q[xi];
          ke += uh[m] * uh[m];
        }
        const double Th = (hh + 1) * twall - hh * Ta;
        const double rh = pw * gm * M2 / Th;
#pragma unroll
        for (int m = 0; m < ND; ++m) Q[(1 + m) * fs + xo] = -rh * uh[m];
        Q[(ND + 1) * fs + xo] = pw / ph.gm1 + 0.5 * rh * ke;
        Q[xo] = rh;
      }
      break;
    }
    case OSBL_BC_DIRICHLET: {  // dirichlet.py:28-41
      const int n1i = g.n[e1], n2i = g.n[e2];
      const int pa = ia < 0 ? 0 : (ia >= n1i ? n1i - 1 : ia);
      const int pb = ib < 0 ? 0 : (ib >= n2i ? n2i - 1 : ib);
      for (int v = 0; v < NV; ++v) {
        const double val = plane[((long long)v * n2i + pb) * n1i + pa];
        for (int hh = 0; hh <= nh; ++hh) Q[v * fs + x - hh * st] = val;
      }
      break;
    }
    case OSBL_BC_INLET_PEXTRAP: {  // inlet_pressure_extrapolate.py:32-66, side 0
      const double rb = Q[x];
      double ke = 0.0, un = 0.0;
#pragma unroll
      for (int m = 0; m < ND; ++m) {
        const double ub = fabs(Q[(1 + m) * fs + x] / rb);
        ke += ub * ub;
        if (m == d) un = ub;
      }
      const double pb = ph.gm1 * (Q[(ND + 1) * fs + x] - 0.5 * rb * ke);
      const double ab = sqrt(gm * pb / rb);
      if (un >= ab) {
#pragma unroll
        for (int v = 0; v < NV; ++v) Q[v * fs + x] = Q[v * fs + x - st];
      } else {
        const double e0 = Q[(ND + 1) * fs + x];
        for (int hh = 1; hh <= nh; ++hh) {
          Q[(ND + 1) * fs + x - hh * st] = e0;
          Q2[(ND + 1) * fs + x - hh * st] = e0;
        }
      }
      break;
    }
    default:
      break;
  }
}

// Periodic exchange of one direction over the full extents of the others
// (periodic.py:42-56): halo points [-h, 0) and [N, N+h) take their images.
__global__ void __launch_bounds__(256) k_periodic_dir(double* __restrict__ Q, int nv, Grid g, int d) {
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int n1 = g.n[e1] + 2 * g.h[e1], n2 = g.n[e2] + 2 * g.h[e2];
  const int h = g.h[d], N = g.n[d];
  const long long total = (long long)n1 * n2 * 2 * h;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int hh = (int)(t % (2 * h));
  const long long r = t / (2 * h);
  int c[3], s[3];
  c[e1] = s[e1] = (int)(r % n1) - g.h[e1];
  c[e2] = s[e2] = (int)(r / n1) - g.h[e2];
  c[d] = hh < h ? hh - h : N + (hh - h);
  s[d] = hh < h ? c[d] + N : c[d] - N;
  const int dst = g.idx(c[0], c[1], c[2]), src = g.idx(s[0], s[1], s[2]);
  for (int m = 0; m < nv; ++m) Q[m * g.fs + dst] = Q[m * g.fs + src];
}

// Every halo point of face (d, side) takes the plane's values (the inflow
// state an inlet face keeps in its halos).
__global__ void __launch_bounds__(256) k_fill_face_halo(double* __restrict__ Q, int nv, Grid g, int d, int side,
                                                        const double* __restrict__ plane) {
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int n1 = g.n[e1] + 2 * g.h[e1], n2 = g.n[e2] + 2 * g.h[e2];
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n1 * n2) return;
  int c[3];
  const int ia = t % n1 - g.h[e1], ib = t / n1 - g.h[e2];
  c[e1] = ia;
  c[e2] = ib;
  c[d] = side == 0 ? 0 : g.n[d] - 1;
  const int x = g.idx(c[0], c[1], c[2]);
  const int st = (side == 0 ? 1 : -1) * g.stride(d);
  const int n1i = g.n[e1], n2i = g.n[e2];
  const int pa = ia < 0 ? 0 : (ia >= n1i ? n1i - 1 : ia);
  const int pb = ib < 0 ? 0 : (ib >= n2i ? n2i - 1 : ib);
  for (int v = 0; v < nv; ++v) {
    const double val = plane[((long long)v * n2i + pb) * n1i + pa];
    for (int hh = 1; hh <= g.h[d]; ++hh) Q[v * g.fs + x - hh * st] = val;
  }
}

void launch_bc_face(int nd, double* Q, double* Q2, const Grid& g, const Phys& ph, int d, int side, int type, int nh,
                    double twall, const double* plane, hipStream_t s) {
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int n = (g.n[e1] + 2 * g.h[e1]) * (g.n[e2] + 2 * g.h[e2]);
  if (nh > g.h[d]) nh = g.h[d];
  if (nd == 3) k_bc_face<3><<<(n + 255) / 256, 256, 0, s>>>(Q, Q2, g, ph, d, side, type, nh, twall, plane);
  else if (nd == 2) k_bc_face<2><<<(n + 255) / 256, 256, 0, s>>>(Q, Q2, g, ph, d, side, type, nh, twall, plane);
  else k_bc_face<1><<<(n + 255) / 256, 256, 0, s>>>(Q, Q2, g, ph, d, side, type, nh, twall, plane);
}

void launch_periodic_dir(double* Q, int nv, const Grid& g, int d, hipStream_t s) {
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const long long total = (long long)(g.n[e1] + 2 * g.h[e1]) * (g.n[e2] + 2 * g.h[e2]) * 2 * g.h[d];
  k_periodic_dir<<<(unsigned)((total + 255) / 256), 256, 0, s>>>(Q, nv, g, d);
}

void launch_fill_face_halo(double* Q, int nv, const Grid& g, int d, int side, const double* plane, hipStream_t s) {
  const int e1 = d == 0 ? 1 : 0, e2 = d == 2 ? 1 : 2;
  const int n = (g.n[e1] + 2 * g.h[e1]) * (g.n[e2] + 2 * g.h[e2]);
  k_fill_face_halo<<<(n + 255) / 256, 256, 0, s>>>(Q, nv, g, d, side, plane);
}

}  // namespace osbl
