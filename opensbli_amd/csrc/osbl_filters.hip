// osbl_filters.hip -- the filters OpenSBLI places after every full time step
// (UserDefinedEquations with InTheSimulation(frequency=False),
// algorithm.py:412-470): the WENO non-linear filter, selective frequency
// damping and the binomial filter.
//
// The WENO filter's characteristic fluxes come from the characteristic
// sweeps (scheme tag CONV_WFILTER); this file holds its pointwise
// application and the two other filters.  All kernels cover the interior
// only; the halos are refreshed by the boundary conditions that follow.

#include "osbl_kernels.h"

namespace osbl {

namespace {

// Yee's low-Mach correction from the conservative values at x
// (WENO_filter.py:178-196): min(M^2/2 sqrt(4 + (1-M^2)^2)/(1+M^2), 1).
template <int ND>
__device__ __forceinline__ double mach_sensor(const double* __restrict__ Q, long long fs, int x, const Phys& ph) {
  const double inv_rho = 1.0 / Q[x];
  double m2 = 0.0, uu = 0.0;
#pragma unroll
  for (int m = 0; m < ND; ++m) {
    const double rm = Q[(1 + m) * fs + x];
    const double u = rm * inv_rho;
    m2 += rm * rm;
    uu += u * u;
  }
  const double p = ph.gm1 * (Q[(ND + 1) * fs + x] - 0.5 * m2 * inv_rho);
  const double inv_a = 1.0 / sqrt(ph.gama * p * inv_rho);
  const double M = sqrt(uu) * inv_a, M2 = M * M;
  return fmin(0.5 * M2 * sqrt(4.0 + (1.0 - M2) * (1.0 - M2)) / (1.0 + M2), 1.0);
}

// kappa = Mach sensor x Ducros sensor at every interior point, in place
// (WENO_filter.py:211-216); pointwise, so no thread reads another's output.
template <int ND>
__global__ void __launch_bounds__(256) k_kappa_mach(const double* __restrict__ Q, double* __restrict__ theta, Grid g,
                                                    Phys ph) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int x = g.idx(i, blockIdx.y, blockIdx.z);
  theta[x] = mach_sensor<ND>(Q, g.fs, x, ph) * theta[x];
}

// Application of the WENO filter (WENO_filter.py:265-318):
// Q += Wall * shock_filter_control * kappa_max * dt * R, with R the filter
// residual of the sweeps.  With the Ducros sensor, kappa (already scaled by
// the Mach sensor when corrected) is read from theta; the reference's sensor
// array is never written in its halos (zeros), so kappa_max, the maximum over
// the points +-2 along each direction, only sees interior points.  With the
// constant sensor kappa_max is 1 or the Mach sensor of the point.  Wall = 0
// within 6 points of a wall face (walls: bit 2d+side).  Each thread reads and
// writes Q at its own point only.
template <int ND, bool DUCROS, bool MACH>
__global__ void __launch_bounds__(256) k_weno_filter_apply(double* __restrict__ Q, const double* __restrict__ R,
                                                           const double* __restrict__ theta, Grid g, Phys ph,
                                                           double control, int walls) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.n[0]) return;
  const int cc[3] = {i, j, k};
  const int x = g.idx(i, j, k);
  double kmax;
  if constexpr (DUCROS) {
    kmax = theta[x];
#pragma unroll
    for (int d = 0; d < ND; ++d) {
      const int st = g.stride(d);
#pragma unroll
      for (int o = -2; o <= 2; ++o) {
        const int c = cc[d] + o;
        if (c < 0 || c >= g.n[d]) continue;
        kmax = fmax(kmax, theta[x + o * st]);
      }
    }
  } else {
    kmax = MACH ? mach_sensor<ND>(Q, g.fs, x, ph) : 1.0;
  }
  double w = 1.0;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    if (((walls >> (2 * d)) & 1) && cc[d] <= 5) w = 0.0;
    if (((walls >> (2 * d + 1)) & 1) && cc[d] >= g.n[d] - 6) w = 0.0;
  }
  const double f = w * control * kmax * ph.dt;
#pragma unroll
  for (int m = 0; m < NV; ++m) Q[m * g.fs + x] = Q[m * g.fs + x] + f * R[m * g.fs + x];
}

// Selective frequency damping (SFD.py:64-86) at every interior point:
// vfilt' = (v (1 - E) + vfilt (c + E))/(c + 1), v' = (v (E c + 1) + vfilt c (1 - E))/(c + 1).
template <int ND>
__global__ void __launch_bounds__(256) k_sfd(double* __restrict__ Q, double* __restrict__ F, Grid g, double c,
                                             double E) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.n[0]) return;
  const int x = g.idx(i, blockIdx.y, blockIdx.z);
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    const double v = Q[m * g.fs + x], old = F[m * g.fs + x];
    F[m * g.fs + x] = (v * (1.0 - E) + old * (c + E)) / (c + 1.0);
    Q[m * g.fs + x] = (v * (E * c + 1.0) + old * c * (1.0 - E)) / (c + 1.0);
  }
}

struct BinomW {
  double w[11];
};

// Binomial filter of even order N (binomial_filter.py:36-110), from the
// pre-filter state Q into Qn: per direction sum_t w_t q[x + (t - N/2) e_d],
// averaged over the directions, blended as (1 - sigma) q + sigma q_f where the
// grid condition holds (mask of interior cells, or every cell).
template <int ND>
__global__ void __launch_bounds__(256) k_binomial(const double* __restrict__ Q, double* __restrict__ Qn, Grid g,
                                                  BinomW bw, int N, double sigma,
                                                  const unsigned char* __restrict__ mask) {
  constexpr int NV = ND + 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y, k = blockIdx.z;
  if (i >= g.n[0]) return;
  const int x = g.idx(i, j, k);
  bool on = true;
  if (mask) on = mask[(long long)i + (long long)g.n[0] * (j + (long long)g.n[1] * k)] != 0;
  const int h = N / 2;
#pragma unroll
  for (int m = 0; m < NV; ++m) {
    const double* q = Q + m * g.fs;
    const double v = q[x];
    double out = v;
    if (on) {
      double sum = 0.0;
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const int st = g.stride(d);
        double fd = 0.0;
        for (int t = 0; t <= N; ++t) fd += bw.w[t] * q[x + (t - h) * st];
        sum += fd;
      }
      const double qf = sum / (double)ND;
      out = (1.0 - sigma) * v + sigma * qf;
    }
    Qn[m * g.fs + x] = out;
  }
}

}  // namespace

void launch_weno_filter_apply(int nd, double* Q, const double* R, double* theta, const Grid& g, const Phys& ph,
                              bool ducros, bool mach, double control, int walls, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  auto go = [&](auto ndc) {
    constexpr int ND = decltype(ndc)::value;
    if (ducros && mach) k_kappa_mach<ND><<<grid, 256, 0, s>>>(Q, theta, g, ph);
    if (ducros) k_weno_filter_apply<ND, true, false><<<grid, 256, 0, s>>>(Q, R, theta, g, ph, control, walls);
    else if (mach) k_weno_filter_apply<ND, false, true><<<grid, 256, 0, s>>>(Q, R, theta, g, ph, control, walls);
    else k_weno_filter_apply<ND, false, false><<<grid, 256, 0, s>>>(Q, R, theta, g, ph, control, walls);
  };
  if (nd == 3) go(std::integral_constant<int, 3>{});
  else go(std::integral_constant<int, 2>{});
}

void launch_sfd(int nd, double* Q, double* F, const Grid& g, double c, double E, hipStream_t s) {
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  if (nd == 3) k_sfd<3><<<grid, 256, 0, s>>>(Q, F, g, c, E);
  else k_sfd<2><<<grid, 256, 0, s>>>(Q, F, g, c, E);
}

void launch_binomial(int nd, const double* Q, double* Qn, const Grid& g, const double* w, int N, double sigma,
                     const unsigned char* mask, hipStream_t s) {
  BinomW bw{};
  for (int t = 0; t <= N && t < 11; ++t) bw.w[t] = w[t];
  dim3 grid((g.n[0] + 255) / 256, g.n[1], g.n[2]);
  if (nd == 3) k_binomial<3><<<grid, 256, 0, s>>>(Q, Qn, g, bw, N, sigma, mask);
  else k_binomial<2><<<grid, 256, 0, s>>>(Q, Qn, g, bw, N, sigma, mask);
}

}  // namespace osbl
