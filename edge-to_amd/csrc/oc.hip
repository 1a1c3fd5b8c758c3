// Optimality-criteria design update on the device.
//
// The reference bisects lambda in [1e-9, 1e9] with a host round trip per
// step.  Here each step is two kernels (a partial volume sum at the current
// midpoint, then a one-block kernel that finalises the sum, moves l1/l2 and
// raises a `done` flag); steps are enqueued in chunks and every kernel is a
// no-op once `done` is set, so the host synchronises once per chunk instead
// of once per step.  The midpoint sequence and the decisions are those of
// the reference; the new design is evaluated once, at the last midpoint.
#include "common.hpp"
#include "blas.hpp"

namespace et {

struct OcScalars {
  double l1, l2, lmid, sabs;
  int done, steps;
};

__global__ __launch_bounds__(256) void oc_prep_kernel(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ dc, double A,
                                                      double* __restrict__ ocp,
                                                      double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double q = -dc[t] / A;
    q = q > 0.0 ? q : 0.0;
    double v = x[t] * sqrt(q);
    if (v != v) v = 0.0;  // nan_to_num(nan=0)
    ocp[t] = v;
    s += fabs(v);
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void oc_init_kernel(const double* __restrict__ partial, OcScalars* sc) {
  double s = 0.0;
  for (int t = threadIdx.x; t < kRedBlocks; t += blockDim.x) s += partial[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) {
    sc->sabs = s;
    sc->l1 = 1e-9;
    sc->l2 = 1e9;
    sc->lmid = 0.5 * (sc->l1 + sc->l2);
    sc->done = !((sc->l2 - sc->l1) / (sc->l2 + sc->l1) > 1e-4);
    sc->steps = 0;
  }
}

__device__ __forceinline__ double oc_value(double x, double ocp_raw, bool all_zero, double move,
                                           double lmid) {
  const double xu = fmin(fmax(x + move, 0.0), 1.0);
  const double xl = fmin(fmax(x - move, 0.0), 1.0);
  const double p = all_zero ? 1e-3 : fmax(1e-12, ocp_raw);
  return fmax(0.0, fmax(xl, fmin(1.0, fmin(xu, p / lmid))));
}

__global__ __launch_bounds__(256) void oc_step_kernel(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ ocp, double A,
                                                      double move, const OcScalars* sc,
                                                      double* __restrict__ partial) {
  if (sc->done) return;
  const double lmid = 0.5 * (sc->l1 + sc->l2);
  const bool all_zero = sc->sabs == 0.0;
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    s += oc_value(x[t], ocp[t], all_zero, move, lmid) * A;
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void oc_decide_kernel(const double* __restrict__ partial, double vol, double vf,
                                 OcScalars* sc) {
  if (sc->done) return;
  double s = 0.0;
  for (int t = threadIdx.x; t < kRedBlocks; t += blockDim.x) s += partial[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) {
    const double lmid = 0.5 * (sc->l1 + sc->l2);
    sc->lmid = lmid;
    if (s / vol <= vf)
      sc->l2 = lmid;
    else
      sc->l1 = lmid;
    sc->steps += 1;
    sc->done = !((sc->l2 - sc->l1) / (sc->l2 + sc->l1) > 1e-4);
  }
}

__global__ __launch_bounds__(256) void oc_final_kernel(int64_t n, const double* __restrict__ x,
                                                       const double* __restrict__ ocp, double move,
                                                       const OcScalars* sc,
                                                       double* __restrict__ xn) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  xn[t] = oc_value(x[t], ocp[t], sc->sabs == 0.0, move, sc->lmid);
}

int64_t oc_work_size(int64_t n) {
  return n * (int64_t)sizeof(double) + kRedBlocks * (int64_t)sizeof(double) + 256;
}

int oc_update(int64_t n, const double* x, const double* dc, double A, double vol, double vf,
              double move, double* xn, void* work, int* steps, hipStream_t s) {
  ET_REQUIRE(n > 0 && x && dc && xn && work, "oc_update: bad arguments");
  ET_REQUIRE(A > 0 && vol > 0, "oc_update: element measure and volume must be positive");
  double* ocp = static_cast<double*>(work);
  double* partial = ocp + n;
  OcScalars* sc = reinterpret_cast<OcScalars*>(partial + kRedBlocks);
  // pinned read-back slot, allocated once per host thread and kept for the process
  static thread_local OcScalars* h = nullptr;
  if (h == nullptr) ET_CHECK_HIP(hipHostMalloc(&h, sizeof(OcScalars)));
  int status = ET_OK;
  do {
    oc_prep_kernel<<<kRedBlocks, 256, 0, s>>>(n, x, dc, A, ocp, partial);
    oc_init_kernel<<<1, 256, 0, s>>>(partial, sc);
    constexpr int kChunk = 12;
    for (int round = 0; round < 40; ++round) {  // bisection needs < 120 steps
      for (int c = 0; c < kChunk; ++c) {
        oc_step_kernel<<<kRedBlocks, 256, 0, s>>>(n, x, ocp, A, move, sc, partial);
        oc_decide_kernel<<<1, 256, 0, s>>>(partial, vol, vf, sc);
      }
      if (hipMemcpyAsync(h, sc, sizeof(OcScalars), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("oc_update: HIP error during bisection");
        status = ET_EHIP;
        break;
      }
      if (h->done) break;
    }
    if (status != ET_OK) break;
    if (!h->done) {
      set_error("oc_update: bisection did not terminate");
      status = ET_ENOCONV;
      break;
    }
    oc_final_kernel<<<grid_for(n), kBlock, 0, s>>>(n, x, ocp, move, sc, xn);
    if (hipGetLastError() != hipSuccess) {
      set_error("oc_update: launch failed");
      status = ET_EHIP;
      break;
    }
    if (steps) *steps = h->steps;
  } while (0);
  return status;
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
