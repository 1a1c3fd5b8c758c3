// Linear-hat density filter on structured element grids: normaliser,
// forward (normalised weighted average over in-bounds neighbours) and the
// exact transpose.  One thread per element; the offset/weight table (at most
// a few hundred entries) is staged once per block in LDS and read as a
// broadcast.  Summation follows the offset order, so results are
// deterministic.
#include "common.hpp"
#include "stiffness.hpp"

namespace et {

constexpr int kMaxOff = 1024;

template <int DIM>
__device__ __forceinline__ void elem_coords(const Grid& g, int64_t e, int& i, int& j, int& k) {
  if (DIM == 3) {
    i = (int)(e / ((int64_t)g.ey * g.ez));
    const int r = (int)(e % ((int64_t)g.ey * g.ez));
    j = r / g.ez;
    k = r % g.ez;
  } else {
    i = (int)(e / g.ey);
    j = (int)(e % g.ey);
    k = 0;
  }
}

// MODE 0: normaliser, 1: forward, 2: transpose
template <int DIM, int MODE>
__global__ __launch_bounds__(256) void filter_kernel(const Grid g, int n_off,
                                                     const int32_t* __restrict__ off,
                                                     const double* __restrict__ w,
                                                     const double* __restrict__ norm,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
  __shared__ int s_off[kMaxOff * DIM];
  __shared__ double s_w[kMaxOff];
  for (int t = threadIdx.x; t < n_off; t += blockDim.x) {
    for (int d = 0; d < DIM; ++d) s_off[t * DIM + d] = off[t * DIM + d];
    s_w[t] = w[t];
  }
  __syncthreads();
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_elem()) return;
  int i, j, k;
  elem_coords<DIM>(g, e, i, j, k);
  double sv = 0.0, sw = 0.0;
  for (int t = 0; t < n_off; ++t) {
    const int a = i + s_off[t * DIM], b = j + s_off[t * DIM + 1];
    const int c = DIM == 3 ? k + s_off[t * DIM + 2] : 0;
    bool ok = a >= 0 && a < g.ex && b >= 0 && b < g.ey;
    if (DIM == 3) ok = ok && c >= 0 && c < g.ez;
    if (!ok) continue;
    const int64_t q = DIM == 3 ? ((int64_t)a * g.ey + b) * g.ez + c : (int64_t)a * g.ey + b;
    if (MODE == 0) {
      sw += s_w[t];
    } else if (MODE == 1) {
      sv += s_w[t] * x[q];
      sw += s_w[t];
    } else {
      sv += (s_w[t] / norm[q]) * x[q];
    }
  }
  if (MODE == 0)
    y[e] = sw;
  else if (MODE == 1)
    y[e] = sw > 0.0 ? sv / sw : 0.0;
  else
    y[e] = sv;
}

template <int MODE>
static int launch_filter(const Grid& g, int n_off, const int32_t* off, const double* w,
                         const double* norm, const double* x, double* y, hipStream_t s) {
  ET_REQUIRE(n_off > 0 && n_off <= kMaxOff, "filter offset count out of range (1..1024)");
  const unsigned grid = grid_for(g.n_elem());
  if (g.dim == 3)
    filter_kernel<3, MODE><<<grid, kBlock, 0, s>>>(g, n_off, off, w, norm, x, y);
  else
    filter_kernel<2, MODE><<<grid, kBlock, 0, s>>>(g, n_off, off, w, norm, x, y);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int filter_normalizer(const Grid& g, int n_off, const int32_t* off, const double* w, double* norm,
                      hipStream_t s) {
  return launch_filter<0>(g, n_off, off, w, nullptr, nullptr, norm, s);
}
int filter_fwd(const Grid& g, int n_off, const int32_t* off, const double* w, const double* x,
               double* y, hipStream_t s) {
  return launch_filter<1>(g, n_off, off, w, nullptr, x, y, s);
}
int filter_adj(const Grid& g, int n_off, const int32_t* off, const double* w, const double* norm,
               const double* v, double* y, hipStream_t s) {
  return launch_filter<2>(g, n_off, off, w, norm, v, y, s);
}

}  // namespace et
