// osbl_batch.hip -- ensembles of independent 1-D Euler problems of uneven
// size (WENO/TENO characteristic LLF fluxes, zeroth-order extrapolation ends,
// 3-register or low-storage RK), integrated on one GPU.
//
// Storage: each variable is one array holding every instance's points back
// to back, each instance with H halo points on both sides.
//
// Two executions, same arithmetic as the single-block path (char_flux<1,...>,
// the same residual and RK formulas), so results are bitwise identical to
// running every instance alone:
//  * persistent: one launch; workgroup b integrates instance order[b]
//    through all its steps and stages, with block barriers between phases.
//    Threads per workgroup come from a per-size table the autotuner fills.
//  * staged: one launch per RK stage covering every instance still running;
//    the stages of a chunk of steps are captured once into a hipGraph (two,
//    when a chunk ends on the other ping-pong buffer) and replayed for every
//    chunk, the step number coming from a device counter.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "osbl_kernels.h"

using namespace osbl;

namespace {

constexpr int BH = 4;  // halo points per side

struct Inst {
  int n;           // points
  int off;         // index of point 0 in every variable array
  int nsteps;
  int threads;     // persistent mode: threads used by this instance's workgroup
  double dt;
  double inv_dx;
};

// one RK stage of the sweep + update for the cells [c0, c1) of one instance
template <int CONV, int AVG, int MODE>
__device__ __forceinline__ void stage_segment(const double* __restrict__ Q, double* __restrict__ Qo,
                                              double* __restrict__ Qn, long long fs, const Inst& in, int c0, int c1,
                                              const Phys& ph, double cA, double cB, bool first) {
  constexpr int ND = 1, NV = 3;
  if (c0 >= c1) return;
  int off = in.off + c0 - 3;
  Pt<ND> w[6];
#pragma unroll
  for (int p = 0; p < 5; ++p) load_pt<ND>(w[p], Q, fs, off + p, ph);
  // the next point's raw values are fetched one iteration ahead
  double nxt[NV];
  fetch_q<ND>(nxt, Q, fs, off + 5);
  double fprev[NV];
  for (int f = c0 - 1; f < c1; ++f) {
#pragma unroll
    for (int m = 0; m < NV; ++m) w[5].q[m] = nxt[m];
    finish_pt<ND>(w[5], ph);
    if (f + 1 < c1) fetch_q<ND>(nxt, Q, fs, off + 6);
    double fh[NV];
    // the RK register of this iteration's cell is fetched before the flux so
    // that the load overlaps its arithmetic; the cell's state is window point 2
    double qoin[NV];
#pragma unroll
    for (int m = 0; m < NV; ++m) qoin[m] = 0.0;
    if (f >= c0 && !(MODE == 0 && first)) {
#pragma unroll
      for (int m = 0; m < NV; ++m) qoin[m] = Qo[m * fs + off + 2];
    }
    char_flux<ND, 0, CONV, AVG>(w, ph, fh, ph.teno_ct);
    if (f >= c0) {
      const int x = off + 2;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const double r = -(fh[m] - fprev[m]) * in.inv_dx;
        if constexpr (MODE == 0) {
          const double qo = first ? w[2].q[m] : qoin[m];
          Qn[m * fs + x] = qo + in.dt * cA * r;
          Qo[m * fs + x] = qo + in.dt * cB * r;
        } else {
          const double tmp = (cA == 0.0 ? 0.0 : cA * qoin[m]) + in.dt * r;
          Qo[m * fs + x] = tmp;
          Qn[m * fs + x] = w[2].q[m] + cB * tmp;
        }
      }
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) fprev[m] = fh[m];
#pragma unroll
    for (int p = 0; p < 5; ++p) w[p] = w[p + 1];
    ++off;
  }
}

// zeroth-order extrapolation of both ends (extrapolation.py:29-58) by lanes 0, 1
__device__ __forceinline__ void extrap_ends(double* __restrict__ Q, long long fs, const Inst& in, int lane) {
  if (lane > 1) return;
  const int b = lane == 0 ? 0 : in.n - 1;
  const int from = lane == 0 ? 1 : in.n - 2;
  const int dir = lane == 0 ? -1 : 1;
  for (int m = 0; m < 3; ++m) {
    const double v = Q[m * fs + in.off + from];
    for (int h = 0; h <= BH; ++h) Q[m * fs + in.off + b + dir * h] = v;
  }
}

struct Coeffs {
  double A[5], B[5];
  int ns;
};

template <int CONV, int AVG, int MODE>
__global__ void __launch_bounds__(512) k_batch_persistent(const Inst* __restrict__ insts, const int* __restrict__ order,
                                                            double* __restrict__ QA, double* __restrict__ QB,
                                                            double* __restrict__ Qo, long long fs, Phys ph, Coeffs rk) {
  const Inst in = insts[order[blockIdx.x]];
  const int T = in.threads;
  const int t = threadIdx.x;
  double* Q = QA;
  double* Qn = QB;
  const int seg = (in.n + T - 1) / T;
  const int c0 = min(t * seg, in.n), c1 = min(c0 + seg, in.n);
  const bool act = t < T;
  for (int step = 0; step < in.nsteps; ++step) {
    for (int s = 0; s < rk.ns; ++s) {
      if (act) extrap_ends(Q, fs, in, t);
      __syncthreads();
      if (act) stage_segment<CONV, AVG, MODE>(Q, Qo, Qn, fs, in, c0, c1, ph, rk.A[s], rk.B[s], s == 0);
      __syncthreads();
      double* tmp = Q;
      Q = Qn;
      Qn = tmp;
    }
  }
  // leave the final state of every instance in QA with filled halos
  if ((in.nsteps * rk.ns) & 1) {
    for (int i = t - BH; i < in.n + BH; i += blockDim.x)
      for (int m = 0; m < 3; ++m) QA[m * fs + in.off + i] = Q[m * fs + in.off + i];
    __syncthreads();
  }
  if (t < 2) extrap_ends(QA, fs, in, t);
}

// Staged mode.  The step a launch belongs to is *step0 + rel: step0 is a
// device counter the chunk's last kernel advances, so one captured chunk of
// steps serves every chunk of the run.
//
// one thread per (instance, segment); inactive instances skip
template <int CONV, int AVG, int MODE>
__global__ void __launch_bounds__(256) k_batch_stage(const Inst* __restrict__ insts, const int* __restrict__ seg_inst,
                                                      const int* __restrict__ seg_c0, int nseg, const double* __restrict__ Q,
                                                      double* __restrict__ Qo, double* __restrict__ Qn, long long fs,
                                                      Phys ph, double cA, double cB, int first,
                                                      const int* __restrict__ step0, int rel, int seglen) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nseg) return;
  const Inst in = insts[seg_inst[t]];
  const int step = *step0 + rel;
  if (step >= in.nsteps) return;
  const int c0 = seg_c0[t];
  const int c1 = min(c0 + seglen, in.n);
  stage_segment<CONV, AVG, MODE>(Q, Qo, Qn, fs, in, c0, c1, ph, cA, cB, first != 0);
}

__global__ void __launch_bounds__(256) k_batch_bc(const Inst* __restrict__ insts, int ninst, double* __restrict__ Q,
                                                   long long fs, const int* __restrict__ step0, int rel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 2 * ninst) return;
  const Inst in = insts[t >> 1];
  if (*step0 + rel > in.nsteps) return;
  extrap_ends(Q, fs, in, t & 1);
}

// inactive instances keep their state: copy Q -> Qn so the ping-pong stays consistent
__global__ void __launch_bounds__(256) k_batch_carry(const Inst* __restrict__ insts, const int* __restrict__ pt_inst,
                                                      const int* __restrict__ pt_i, int npts, const double* __restrict__ Q,
                                                      double* __restrict__ Qn, long long fs, const int* __restrict__ step0,
                                                      int rel) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npts) return;
  const Inst in = insts[pt_inst[t]];
  if (*step0 + rel < in.nsteps) return;
  const int x = in.off + pt_i[t];
  for (int m = 0; m < 3; ++m) Qn[m * fs + x] = Q[m * fs + x];
}

// the staged mode's step counter: set before a run, advanced after each chunk
__global__ void k_batch_step_set(int* __restrict__ step0, int value) {
  if (threadIdx.x == 0) step0[0] = value;
}
__global__ void k_batch_step_add(int* __restrict__ step0, int add) {
  if (threadIdx.x == 0) step0[0] += add;
}

int bfail(int code, const std::string& msg) { return osbl::set_error(code, msg.c_str()); }
#define BCHK(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return bfail(OSBL_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

struct osbl_batch {
  osbl_config base{};
  int ninst = 0;
  long long fs = 0;
  Phys ph{};
  Coeffs rk{};
  int mode_rk = 0;
  std::vector<Inst> h_inst;
  std::vector<int> order;  // instances by decreasing work
  Inst* d_inst = nullptr;
  int* d_order = nullptr;
  double* QA = nullptr;
  double* QB = nullptr;
  double* Qo = nullptr;
  double* stage_buf = nullptr;
  hipStream_t stream = nullptr;
  // staged mode
  int seglen = 16;
  int* d_seg_inst = nullptr;
  int* d_seg_c0 = nullptr;
  int nseg = 0;
  int* d_pt_inst = nullptr;
  int* d_pt_i = nullptr;
  int npts = 0;
  int graph_steps = 8;
  int* d_step0 = nullptr;
  // captured chunk of graph_steps steps, by the buffer it starts from (0: QA, 1: QB)
  hipGraphExec_t chunk[2] = {nullptr, nullptr};
  // per-size thread table: threads for n <= bucket_n[k]
  std::vector<int> bucket_n{256, 512, 1024, 2048, 4096, 1 << 30};
  std::vector<int> bucket_t{64, 64, 128, 256, 256, 512};
  double last_ms = 0.0;
};

namespace {

int pick_threads(const osbl_batch* b, int n) {
  for (size_t k = 0; k < b->bucket_n.size(); ++k)
    if (n <= b->bucket_n[k]) return b->bucket_t[k];
  return b->bucket_t.back();
}

int upload_instances(osbl_batch* b) {
  for (auto& in : b->h_inst) in.threads = pick_threads(b, in.n);
  BCHK(hipMemcpy(b->d_inst, b->h_inst.data(), sizeof(Inst) * b->ninst, hipMemcpyHostToDevice));
  return OSBL_OK;
}

template <int CONV, int AVG>
void launch_persistent_t(osbl_batch* b) {
  int maxT = 64;
  for (auto& in : b->h_inst) maxT = std::max(maxT, in.threads);
  if (b->mode_rk == 0)
    k_batch_persistent<CONV, AVG, 0><<<b->ninst, maxT, 0, b->stream>>>(b->d_inst, b->d_order, b->QA, b->QB, b->Qo,
                                                                       b->fs, b->ph, b->rk);
  else
    k_batch_persistent<CONV, AVG, 1><<<b->ninst, maxT, 0, b->stream>>>(b->d_inst, b->d_order, b->QA, b->QB, b->Qo,
                                                                       b->fs, b->ph, b->rk);
}

template <int CONV, int AVG>
void launch_stage_t(osbl_batch* b, const double* Q, double* Qn, int s, int rel) {
  const unsigned blocks = (unsigned)((b->nseg + 255) / 256);
  if (b->mode_rk == 0)
    k_batch_stage<CONV, AVG, 0><<<blocks, 256, 0, b->stream>>>(b->d_inst, b->d_seg_inst, b->d_seg_c0, b->nseg, Q, b->Qo,
                                                               Qn, b->fs, b->ph, b->rk.A[s], b->rk.B[s], s == 0,
                                                               b->d_step0, rel, b->seglen);
  else
    k_batch_stage<CONV, AVG, 1><<<blocks, 256, 0, b->stream>>>(b->d_inst, b->d_seg_inst, b->d_seg_c0, b->nseg, Q, b->Qo,
                                                               Qn, b->fs, b->ph, b->rk.A[s], b->rk.B[s], s == 0,
                                                               b->d_step0, rel, b->seglen);
}

template <class F>
void dispatch(int conv, int avg, F&& f) {
  auto go = [&](auto cc) {
    constexpr int C = decltype(cc)::value;
    if (avg == OSBL_AVG_ROE) f(std::integral_constant<int, C>{}, std::integral_constant<int, OSBL_AVG_ROE>{});
    else f(std::integral_constant<int, C>{}, std::integral_constant<int, OSBL_AVG_SIMPLE>{});
  };
  switch (conv) {
    case OSBL_CONV_WENO_JS: go(std::integral_constant<int, OSBL_CONV_WENO_JS>{}); break;
    case OSBL_CONV_WENO_Z: go(std::integral_constant<int, OSBL_CONV_WENO_Z>{}); break;
    case OSBL_CONV_TENO5: go(std::integral_constant<int, OSBL_CONV_TENO5>{}); break;
    default: go(std::integral_constant<int, OSBL_CONV_TENO6>{}); break;
  }
}

int run_persistent(osbl_batch* b) {
  dispatch(b->base.conv, b->base.averaging, [&](auto c, auto a) {
    launch_persistent_t<decltype(c)::value, decltype(a)::value>(b);
  });
  BCHK(hipGetLastError());
  return OSBL_OK;
}

void drop_chunks(osbl_batch* b) {
  for (auto& g : b->chunk)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// capture one chunk of graph_steps steps starting from buffer `start`
int capture_chunk(osbl_batch* b, int start) {
  const unsigned bcb = (unsigned)((2 * b->ninst + 255) / 256);
  const unsigned cpb = (unsigned)((b->npts + 255) / 256);
  double* Q = start ? b->QB : b->QA;
  double* Qn = start ? b->QA : b->QB;
  hipGraph_t graph;
  BCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  for (int rel = 0; rel < b->graph_steps; ++rel) {
    for (int s = 0; s < b->rk.ns; ++s) {
      k_batch_bc<<<bcb, 256, 0, b->stream>>>(b->d_inst, b->ninst, Q, b->fs, b->d_step0, rel);
      dispatch(b->base.conv, b->base.averaging, [&](auto c, auto a) {
        launch_stage_t<decltype(c)::value, decltype(a)::value>(b, Q, Qn, s, rel);
      });
      k_batch_carry<<<cpb, 256, 0, b->stream>>>(b->d_inst, b->d_pt_inst, b->d_pt_i, b->npts, Q, Qn, b->fs,
                                                b->d_step0, rel);
      std::swap(Q, Qn);
    }
  }
  k_batch_step_add<<<1, 64, 0, b->stream>>>(b->d_step0, b->graph_steps);
  BCHK(hipStreamEndCapture(b->stream, &graph));
  hipError_t e = hipGraphInstantiate(&b->chunk[start], graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    b->chunk[start] = nullptr;
    return bfail(OSBL_E_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  return OSBL_OK;
}

int run_staged(osbl_batch* b) {
  int maxsteps = 0;
  for (auto& in : b->h_inst) maxsteps = std::max(maxsteps, in.nsteps);
  const unsigned bcb = (unsigned)((2 * b->ninst + 255) / 256);
  // every chunk replays the same graph (at most one per starting buffer);
  // steps past an instance's nsteps leave it unchanged, so the last chunk may
  // run beyond maxsteps
  const int swaps = (b->rk.ns * b->graph_steps) & 1;
  int cur = 0;
  k_batch_step_set<<<1, 64, 0, b->stream>>>(b->d_step0, 0);
  BCHK(hipGetLastError());
  for (int step0 = 0; step0 < maxsteps; step0 += b->graph_steps) {
    if (!b->chunk[cur])
      if (int rc = capture_chunk(b, cur)) return rc;
    BCHK(hipGraphLaunch(b->chunk[cur], b->stream));
    cur ^= swaps;
  }
  if (cur) BCHK(hipMemcpyAsync(b->QA, b->QB, sizeof(double) * 3 * b->fs, hipMemcpyDeviceToDevice, b->stream));
  k_batch_step_set<<<1, 64, 0, b->stream>>>(b->d_step0, 0);
  k_batch_bc<<<bcb, 256, 0, b->stream>>>(b->d_inst, b->ninst, b->QA, b->fs, b->d_step0, 0);
  BCHK(hipGetLastError());
  return OSBL_OK;
}

}  // namespace

extern "C" {

int osbl_batch_create(const osbl_config* base, int ninst, const int32_t* n, const double* delta, const double* dt,
                      const int64_t* nsteps, osbl_batch** out) {
  if (!base || !n || !delta || !dt || !nsteps || !out || ninst <= 0) return bfail(OSBL_E_ARG, "invalid argument");
  *out = nullptr;
  if (base->ndim != 1) return bfail(OSBL_E_ARG, "ensembles are 1-D");
  if (base->conv < OSBL_CONV_WENO_JS || base->conv > OSBL_CONV_TENO6)
    return bfail(OSBL_E_ARG, "ensembles use a characteristic scheme (WENO/TENO)");
  if (base->rk < OSBL_RK3 || base->rk > OSBL_RK4_LS) return bfail(OSBL_E_ARG, "unknown RK scheme");
  for (int s = 0; s < 2; ++s)
    if (base->bc[0][s] != OSBL_BC_EXTRAP0) return bfail(OSBL_E_ARG, "ensembles use extrapolation ends");
  auto* b = new osbl_batch();
  b->base = *base;
  b->ninst = ninst;
  b->ph.gama = base->gama;
  b->ph.gm1 = base->gama - 1.0;
  b->ph.igm1 = 1.0 / b->ph.gm1;
  b->ph.eps = base->eps;
  b->ph.teno_ct = base->teno_ct;
  b->mode_rk = base->rk == OSBL_RK3 ? 0 : 1;
  {
    double A[5], B[5];
    // same coefficients as the single-block runtime
    int ns;
    switch (base->rk) {
      case OSBL_RK3: ns = 3; A[0] = 2.0 / 3.0; A[1] = 5.0 / 12.0; A[2] = 3.0 / 5.0; B[0] = 1.0 / 4.0; B[1] = 3.0 / 20.0; B[2] = 3.0 / 5.0; break;
      case OSBL_RK3_LS: ns = 3; A[0] = 0.0; A[1] = -5.0 / 9.0; A[2] = -153.0 / 128.0; B[0] = 1.0 / 3.0; B[1] = 15.0 / 16.0; B[2] = 8.0 / 15.0; break;
      default: delete b; return bfail(OSBL_E_ARG, "ensembles support RungeKutta(3) and RungeKuttaLS(3)");
    }
    for (int i = 0; i < ns; ++i) { b->rk.A[i] = A[i]; b->rk.B[i] = B[i]; }
    b->rk.ns = ns;
  }
  long long off = BH;
  b->h_inst.resize(ninst);
  for (int i = 0; i < ninst; ++i) {
    if (n[i] < 8) { delete b; return bfail(OSBL_E_ARG, "instance with fewer than 8 points"); }
    if (!(delta[i] > 0.0) || !(dt[i] > 0.0) || nsteps[i] < 0) { delete b; return bfail(OSBL_E_ARG, "bad instance parameters"); }
    Inst& in = b->h_inst[i];
    in.n = n[i];
    in.off = (int)off;
    in.nsteps = (int)nsteps[i];
    in.dt = dt[i];
    in.inv_dx = 1.0 / delta[i];
    off += n[i] + 2 * BH;
    off = (off + 7) / 8 * 8 + BH;  // 64-byte aligned interiors
  }
  b->fs = (off + 63) / 64 * 64;
  if (b->fs * 3 >= (1LL << 31)) { delete b; return bfail(OSBL_E_ARG, "ensemble too large"); }
  b->order.resize(ninst);
  for (int i = 0; i < ninst; ++i) b->order[i] = i;
  std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) {
    return (double)b->h_inst[a].n * b->h_inst[a].nsteps > (double)b->h_inst[c].n * b->h_inst[c].nsteps;
  });
  // staged mode segment lists
  std::vector<int> si, sc, pi, pp;
  for (int i = 0; i < ninst; ++i) {
    for (int c = 0; c < n[i]; c += b->seglen) { si.push_back(i); sc.push_back(c); }
    for (int p = -BH; p < n[i] + BH; ++p) { pi.push_back(i); pp.push_back(p); }
  }
  b->nseg = (int)si.size();
  b->npts = (int)pi.size();
  auto fail_free = [&](int rc) { osbl_batch_destroy(b); return rc; };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) return fail_free(bfail(OSBL_E_HIP, "stream"));
  const size_t bytes = sizeof(double) * 3 * b->fs;
  if (hipMalloc((void**)&b->QA, bytes) != hipSuccess || hipMalloc((void**)&b->QB, bytes) != hipSuccess ||
      hipMalloc((void**)&b->Qo, bytes) != hipSuccess || hipMalloc((void**)&b->stage_buf, bytes) != hipSuccess ||
      hipMalloc((void**)&b->d_inst, sizeof(Inst) * ninst) != hipSuccess ||
      hipMalloc((void**)&b->d_order, sizeof(int) * ninst) != hipSuccess ||
      hipMalloc((void**)&b->d_seg_inst, sizeof(int) * b->nseg) != hipSuccess ||
      hipMalloc((void**)&b->d_seg_c0, sizeof(int) * b->nseg) != hipSuccess ||
      hipMalloc((void**)&b->d_pt_inst, sizeof(int) * b->npts) != hipSuccess ||
      hipMalloc((void**)&b->d_pt_i, sizeof(int) * b->npts) != hipSuccess ||
      hipMalloc((void**)&b->d_step0, sizeof(int)) != hipSuccess)
    return fail_free(bfail(OSBL_E_NOMEM, "hipMalloc failed"));
  if (hipMemset(b->QA, 0, bytes) != hipSuccess || hipMemset(b->QB, 0, bytes) != hipSuccess ||
      hipMemset(b->Qo, 0, bytes) != hipSuccess ||
      hipMemcpy(b->d_order, b->order.data(), sizeof(int) * ninst, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_seg_inst, si.data(), sizeof(int) * b->nseg, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_seg_c0, sc.data(), sizeof(int) * b->nseg, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_pt_inst, pi.data(), sizeof(int) * b->npts, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_pt_i, pp.data(), sizeof(int) * b->npts, hipMemcpyHostToDevice) != hipSuccess)
    return fail_free(bfail(OSBL_E_HIP, "initialisation copy failed"));
  int rc = upload_instances(b);
  if (rc) return fail_free(rc);
  *out = b;
  return OSBL_OK;
}

void osbl_batch_destroy(osbl_batch* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_chunks(b);
  for (void* p : {(void*)b->QA, (void*)b->QB, (void*)b->Qo, (void*)b->stage_buf, (void*)b->d_inst, (void*)b->d_order,
                  (void*)b->d_seg_inst, (void*)b->d_seg_c0, (void*)b->d_pt_inst, (void*)b->d_pt_i,
                  (void*)b->d_step0})
    if (p) (void)hipFree(p);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int osbl_batch_set_state(osbl_batch* b, int inst, const double* host) {
  if (!b || !host || inst < 0 || inst >= b->ninst) return bfail(OSBL_E_ARG, "invalid argument");
  const Inst& in = b->h_inst[inst];
  for (int m = 0; m < 3; ++m)
    BCHK(hipMemcpy(b->QA + m * b->fs + in.off, host + (size_t)m * in.n, sizeof(double) * in.n, hipMemcpyHostToDevice));
  return OSBL_OK;
}

int osbl_batch_get_state(osbl_batch* b, int inst, double* host) {
  if (!b || !host || inst < 0 || inst >= b->ninst) return bfail(OSBL_E_ARG, "invalid argument");
  BCHK(hipStreamSynchronize(b->stream));
  const Inst& in = b->h_inst[inst];
  for (int m = 0; m < 3; ++m)
    BCHK(hipMemcpy(host + (size_t)m * in.n, b->QA + m * b->fs + in.off, sizeof(double) * in.n, hipMemcpyDeviceToHost));
  return OSBL_OK;
}

int osbl_batch_save(osbl_batch* b) {
  if (!b) return bfail(OSBL_E_ARG, "null");
  BCHK(hipMemcpyAsync(b->stage_buf, b->QA, sizeof(double) * 3 * b->fs, hipMemcpyDeviceToDevice, b->stream));
  return OSBL_OK;
}

int osbl_batch_restore(osbl_batch* b) {
  if (!b) return bfail(OSBL_E_ARG, "null");
  BCHK(hipMemcpyAsync(b->QA, b->stage_buf, sizeof(double) * 3 * b->fs, hipMemcpyDeviceToDevice, b->stream));
  return OSBL_OK;
}

int osbl_batch_run(osbl_batch* b, int mode, double* elapsed_ms) {
  if (!b) return bfail(OSBL_E_ARG, "null");
  hipEvent_t e0, e1;
  BCHK(hipEventCreate(&e0));
  BCHK(hipEventCreate(&e1));
  BCHK(hipEventRecord(e0, b->stream));
  int rc = mode == 0 ? run_persistent(b) : run_staged(b);
  if (rc) return rc;
  BCHK(hipEventRecord(e1, b->stream));
  BCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  BCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->last_ms = ms;
  if (elapsed_ms) *elapsed_ms = ms;
  return OSBL_OK;
}

int osbl_batch_set_threads(osbl_batch* b, int nbuckets, const int32_t* max_n, const int32_t* threads) {
  if (!b || nbuckets <= 0 || !max_n || !threads) return bfail(OSBL_E_ARG, "invalid argument");
  for (int k = 0; k < nbuckets; ++k)
    if (threads[k] < 64 || threads[k] > 512 || threads[k] % 64) return bfail(OSBL_E_ARG, "threads must be 64..512, multiple of 64");
  b->bucket_n.assign(max_n, max_n + nbuckets);
  b->bucket_t.assign(threads, threads + nbuckets);
  b->bucket_n.back() = 1 << 30;
  return upload_instances(b);
}

int osbl_batch_set_graph_steps(osbl_batch* b, int steps) {
  if (!b || steps < 1) return bfail(OSBL_E_ARG, "invalid argument");
  if (steps != b->graph_steps) {
    BCHK(hipStreamSynchronize(b->stream));
    drop_chunks(b);
  }
  b->graph_steps = steps;
  return OSBL_OK;
}

}  // extern "C"
