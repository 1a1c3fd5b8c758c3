// Deterministic FP64 reductions: a fixed grid of kRedBlocks blocks
// grid-strides over the vector and writes one partial per block; a single
// block then sums the partials in a fixed order.  The result is bitwise
// reproducible run to run and independent of timing.
#include "common.hpp"
#include "blas.hpp"

namespace et {

__global__ __launch_bounds__(256) void dot_partial_kernel(int64_t n, const double* __restrict__ a,
                                                          const double* __restrict__ b,
                                                          double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    s += a[t] * b[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void finalize_kernel(int m, const double* __restrict__ partial,
                                                       double* __restrict__ out) {
  double s = 0.0;
  for (int t = threadIdx.x; t < m; t += blockDim.x) s += partial[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) *out = s;
}

int reduce_finalize(const double* partial, double* out, hipStream_t s) {
  finalize_kernel<<<1, 256, 0, s>>>(kRedBlocks, partial, out);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int reduce_finalize_n(const double* partial, int m, double* out, hipStream_t s) {
  finalize_kernel<<<1, 256, 0, s>>>(m, partial, out);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int dot_partials(int64_t n, const double* a, const double* b, double* partial, hipStream_t s) {
  dot_partial_kernel<<<kRedBlocks, 256, 0, s>>>(n, a, b, partial);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int dot(int64_t n, const double* a, const double* b, double* partial, double* out, hipStream_t s) {
  dot_partial_kernel<<<kRedBlocks, 256, 0, s>>>(n, a, b, partial);
  ET_CHECK_LAUNCH();
  return reduce_finalize(partial, out, s);
}

}  // namespace et
