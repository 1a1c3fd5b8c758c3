// Optimality-criteria design update on the device, with no host round trip.
//
// The reference bisects lambda in [1e-9, 1e9] (TopOpt/_cuda.py:242-280) with a
// host synchronisation per step.  Here one pass evaluates the volume at every
// midpoint of the next kOcDepth bisection levels at once (the 2^kOcDepth - 1
// nodes of the subtree below the current bracket, formed exactly as the
// sequential 0.5 * (l1 + l2) would form them), and a one-block kernel then
// walks the subtree with the reference's decisions and stopping rule.  A fixed
// number of passes is enqueued, enough for the longest bisection the bracket
// and the 1e-4 stopping rule allow (73 steps); passes after the stop return at
// once.  The new design is evaluated once, at the last midpoint tried.  Every
// kernel is stream-ordered: the whole update can be captured in a graph.
#include <algorithm>
#include <mutex>

#include "common.hpp"
#include "blas.hpp"

namespace et {

// Bisection state at the start of the work buffer (callers may read `steps`
// there once the stream has reached the update).
struct OcState {
  double l1, l2, lmid, sabs;
  int done, steps;
};
static_assert(sizeof(OcState) == 40, "OcState layout is part of the C ABI");

constexpr int kOcDepth = 3;                       // bisection levels per pass
constexpr int kOcCand = (1 << kOcDepth) - 1;      // midpoints evaluated per pass
constexpr int kOcPasses = 26;                     // 78 >= 73 steps: every bisection ends
constexpr int64_t kOcStateBytes = 256;            // state slot (keeps ocp 256-byte aligned)

__host__ __device__ inline bool oc_continue(double l1, double l2) { return (l2 - l1) / (l2 + l1) > 1e-4; }

// midpoints of the subtree below [l1, l2] in heap order (node i: left child
// 2i+1 = [lo, m_i], right child 2i+2 = [m_i, hi])
__device__ inline void oc_midpoints(double l1, double l2, double (&m)[kOcCand]) {
  double lo[kOcCand], hi[kOcCand];
  lo[0] = l1;
  hi[0] = l2;
#pragma unroll
  for (int i = 0; i < kOcCand; ++i) {
    m[i] = 0.5 * (lo[i] + hi[i]);
    if (2 * i + 2 < kOcCand) {
      lo[2 * i + 1] = lo[i];
      hi[2 * i + 1] = m[i];
      lo[2 * i + 2] = m[i];
      hi[2 * i + 2] = hi[i];
    }
  }
}

__device__ __forceinline__ double oc_value(double x, double ocp_raw, bool all_zero, double move,
                                           double lmid) {
  const double xu = fmin(fmax(x + move, 0.0), 1.0);
  const double xl = fmin(fmax(x - move, 0.0), 1.0);
  const double p = all_zero ? 1e-3 : fmax(1e-12, ocp_raw);
  return fmax(0.0, fmax(xl, fmin(1.0, fmin(xu, p / lmid))));
}

// ocp = x * sqrt(max(-dc/A, 0)) with NaN -> 0 (SingleMaterial.__ocP__,
// MaterialModels.py:83-90), and per-block sums of |ocp|.  A: the element measure,
// Ae[t] per element where Ae is given (general meshes, mesh.As of shape (n, 1)).
__global__ __launch_bounds__(256) void oc_prep_kernel(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ dc, double A,
                                                      const double* __restrict__ Ae,
                                                      double* __restrict__ ocp,
                                                      double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double q = -dc[t] / (Ae != nullptr ? Ae[t] : A);
    q = q > 0.0 ? q : 0.0;
    double v = x[t] * sqrt(q);
    if (v != v) v = 0.0;  // nan_to_num(nan=0)
    ocp[t] = v;
    s += fabs(v);
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void oc_init_kernel(int nb, const double* __restrict__ partial, OcState* st) {
  double s = 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) s += partial[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) {
    st->sabs = s;
    st->l1 = 1e-9;
    st->l2 = 1e9;
    st->lmid = 0.5 * (st->l1 + st->l2);
    st->done = !oc_continue(st->l1, st->l2);
    st->steps = 0;
  }
}

// per-block volume sums sum(x_new(m_c) * A) at every midpoint m_c of the pass
__global__ __launch_bounds__(256) void oc_pass_kernel(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ ocp, double A,
                                                      const double* __restrict__ Ae,
                                                      double move, const OcState* __restrict__ st,
                                                      double* __restrict__ partial, int nb) {
  if (st->done) return;
  double m[kOcCand];
  oc_midpoints(st->l1, st->l2, m);
  const bool all_zero = st->sabs == 0.0;
  double s[kOcCand];
#pragma unroll
  for (int c = 0; c < kOcCand; ++c) s[c] = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double xv = x[t], pv = ocp[t], av = Ae != nullptr ? Ae[t] : A;
#pragma unroll
    for (int c = 0; c < kOcCand; ++c) s[c] += oc_value(xv, pv, all_zero, move, m[c]) * av;
  }
#pragma unroll
  for (int c = 0; c < kOcCand; ++c) {
    const double v = block_sum256(s[c]);
    if (threadIdx.x == 0) partial[c * nb + blockIdx.x] = v;
  }
}

// sums each midpoint's volume (fixed order) and takes up to kOcDepth of the
// reference's steps: l2 = lmid if the volume fraction is within vf, else l1 = lmid
__global__ void oc_decide_kernel(const double* __restrict__ partial, int nb, double vol, double vf,
                                 OcState* st) {
  if (st->done) return;
  __shared__ double tot[kOcCand];
  for (int c = 0; c < kOcCand; ++c) {
    double s = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) s += partial[c * nb + t];
    s = block_sum256(s);
    if (threadIdx.x == 0) tot[c] = s;
  }
  if (threadIdx.x == 0) {
    double m[kOcCand];
    double l1 = st->l1, l2 = st->l2, lmid = st->lmid;
    oc_midpoints(l1, l2, m);
    int node = 0, steps = st->steps, done = 0;
    for (int d = 0; d < kOcDepth; ++d) {
      lmid = m[node];
      const bool feasible = tot[node] / vol <= vf;
      if (feasible)
        l2 = lmid;
      else
        l1 = lmid;
      ++steps;
      if (!oc_continue(l1, l2)) {
        done = 1;
        break;
      }
      node = feasible ? 2 * node + 1 : 2 * node + 2;
    }
    st->l1 = l1;
    st->l2 = l2;
    st->lmid = lmid;
    st->steps = steps;
    st->done = done;
  }
}

__global__ __launch_bounds__(256) void oc_final_kernel(int64_t n, const double* __restrict__ x,
                                                       const double* __restrict__ ocp, double move,
                                                       const OcState* __restrict__ st,
                                                       double* __restrict__ xn) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  xn[t] = oc_value(x[t], ocp[t], st->sabs == 0.0, move, st->lmid);
}

int64_t oc_work_size(int64_t n) {
  return kOcStateBytes + n * (int64_t)sizeof(double) + kOcCand * kRedBlocks * (int64_t)sizeof(double);
}

int oc_update(int64_t n, const double* x, const double* dc, double A, double vol, double vf,
              double move, double* xn, void* work, int* steps, hipStream_t s, const double* Ae) {
  ET_REQUIRE(n > 0 && x && dc && xn && work, "oc_update: bad arguments");
  ET_REQUIRE((Ae != nullptr || A > 0) && vol > 0, "oc_update: element measure and volume must be positive");
  ET_REQUIRE(xn != x && xn != dc, "oc_update: x_new must not alias x or dc");
  ET_REQUIRE((reinterpret_cast<uintptr_t>(work) & 7) == 0, "oc_update: work must be 8-byte aligned");
  OcState* st = static_cast<OcState*>(work);
  double* ocp = reinterpret_cast<double*>(static_cast<char*>(work) + kOcStateBytes);
  double* partial = ocp + n;
  const int nb = (int)std::min<int64_t>(kRedBlocks, (n + 255) / 256);
  oc_prep_kernel<<<nb, 256, 0, s>>>(n, x, dc, A, Ae, ocp, partial);
  oc_init_kernel<<<1, 256, 0, s>>>(nb, partial, st);
  for (int pass = 0; pass < kOcPasses; ++pass) {
    oc_pass_kernel<<<nb, 256, 0, s>>>(n, x, ocp, A, Ae, move, st, partial, nb);
    oc_decide_kernel<<<1, 256, 0, s>>>(partial, nb, vol, vf, st);
  }
  oc_final_kernel<<<grid_for(n), kBlock, 0, s>>>(n, x, ocp, move, st, xn);
  ET_CHECK_LAUNCH();
  if (steps != nullptr) {
    OcState h{};
    ET_CHECK_HIP(hipMemcpyAsync(&h, st, sizeof(OcState), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    *steps = h.steps;
    if (!h.done) {
      set_error("oc_update: bisection did not terminate");
      return ET_ENOCONV;
    }
  }
  return ET_OK;
}

// ------------------------------------------------ allocation-free variant ---
//
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops): the same kernels on a work buffer in a file-scope
// device array.  Each GPU holds its own copy of the array for the life of the
// process; its device address is looked up once per device (the host-side
// name of a __device__ array is only a stand-in).  Every call on one GPU
// shares that array: calls that overlap on one GPU (different streams, or
// host threads) race.  Nothing in it needs clearing between calls: the state
// is initialised by the first kernel of every update, ocp and the partial
// sums are written before they are read.

constexpr int64_t kOcNoAllocMaxN = int64_t(1) << 20;  // elements
constexpr int64_t kOcScratchBytes =
    kOcStateBytes + kOcNoAllocMaxN * (int64_t)sizeof(double) + kOcCand * kRedBlocks * (int64_t)sizeof(double);

static __device__ __align__(256) unsigned char scratch[kOcScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

int64_t oc_noalloc_max_n() { return kOcNoAllocMaxN; }

int oc_update_noalloc(int64_t n, const double* x, const double* dc, double A, double vol, double vf,
                      double move, double* xn, int* steps, hipStream_t s) {
  ET_REQUIRE(n > 0 && n <= kOcNoAllocMaxN,
             "oc_update_noalloc: n must be in [1, 1048576] (size of the static work array)");
  int dev = -1;
  ET_CHECK_HIP(hipGetDevice(&dev));
  ET_REQUIRE(dev >= 0 && dev < kMaxDevices, "oc_update_noalloc: device index out of range");
  void* work = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
      void* p = nullptr;
      ET_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      ET_REQUIRE(p != nullptr, "oc_update_noalloc: no device address for the work array");
      g_scratch_ptr[dev] = p;
    }
    work = g_scratch_ptr[dev];
  }
  return oc_update(n, x, dc, A, vol, vf, move, xn, work, steps, s, nullptr);
}

// sum (a - b)^2, used for the RMS design change of the SIMP loop
__global__ __launch_bounds__(256) void sqdist_partial_kernel(int64_t n, const double* __restrict__ a,
                                                             const double* __restrict__ b,
                                                             double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const double d = a[t] - b[t];
    s += d * d;
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

int sqdist(int64_t n, const double* a, const double* b, double* partial, double* out,
           hipStream_t s) {
  sqdist_partial_kernel<<<kRedBlocks, 256, 0, s>>>(n, a, b, partial);
  ET_CHECK_LAUNCH();
  return reduce_finalize(partial, out, s);
}

}  // namespace et
