// Shared helpers for the EDGE-TO MI355X kernels: status handling, launch
// geometry, structured-grid index helpers and deterministic reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/edgeto_amd.h"

namespace et {

// Thread-local last-error text returned by et_last_error().
void set_error(const std::string& msg);
const char* last_error();

// On failure the runtime's last-error slot is cleared as well, so a failed
// allocation is not reported again by the next, unrelated launch check.
#define ET_CHECK_HIP(expr)                                                           \
  do {                                                                               \
    hipError_t _e = (expr);                                                          \
    if (_e != hipSuccess) {                                                          \
      (void)hipGetLastError();                                                       \
      ::et::set_error(std::string(#expr " failed: ") + hipGetErrorString(_e));       \
      return ET_EHIP;                                                                \
    }                                                                                \
  } while (0)

#define ET_CHECK_LAUNCH() ET_CHECK_HIP(hipGetLastError())

#define ET_REQUIRE(cond, msg)                                                        \
  do {                                                                               \
    if (!(cond)) {                                                                   \
      ::et::set_error(msg);                                                          \
      return ET_EINVAL;                                                              \
    }                                                                                \
  } while (0)

#define ET_TRY(expr)                                                                 \
  do {                                                                               \
    int _s = (expr);                                                                 \
    if (_s != ET_OK) return _s;                                                      \
  } while (0)

constexpr int kBlock = 256;

inline unsigned grid_for(int64_t n, int block = kBlock) {
  return static_cast<unsigned>((n + block - 1) / block);
}

// Structured grid geometry. nel = elements per axis (x, y[, z]); node counts
// are nel+1. Node numbering: 2-D j*nx + i; 3-D j*nx*nz + i*nz + k.
// Element numbering: 2-D i*nely + j; 3-D i*nely*nelz + j*nelz + k.
struct Grid {
  int dim;
  int ex, ey, ez;  // elements per axis (ez = 1 in 2-D)
  __host__ __device__ int nx() const { return ex + 1; }
  __host__ __device__ int ny() const { return ey + 1; }
  __host__ __device__ int nz() const { return dim == 3 ? ez + 1 : 1; }
  __host__ __device__ int64_t n_nodes() const { return (int64_t)nx() * ny() * nz(); }
  __host__ __device__ int64_t n_elem() const {
    return (int64_t)ex * ey * (dim == 3 ? ez : 1);
  }
  __host__ __device__ Grid coarse() const {
    Grid g = *this;
    g.ex = ex / 2;
    g.ey = ey / 2;
    g.ez = dim == 3 ? ez / 2 : 1;
    return g;
  }
};

inline int make_grid(int dim, const int* nel, Grid* g) {
  ET_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  ET_REQUIRE(nel != nullptr, "nel is null");
  g->dim = dim;
  g->ex = nel[0];
  g->ey = nel[1];
  g->ez = dim == 3 ? nel[2] : 1;
  ET_REQUIRE(g->ex > 0 && g->ey > 0 && g->ez > 0, "element counts must be positive");
  ET_REQUIRE(g->n_nodes() * (dim) < (int64_t)INT32_MAX, "grid too large for 32-bit DOF indices");
  return ET_OK;
}

// Local node offsets of the H8 element (bottom face CCW, then top face CCW):
// node a sits at (i + HX[a], j + HY[a], k + HZ[a]) relative to the element origin.
__host__ __device__ constexpr int hx(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0; }
__host__ __device__ constexpr int hy(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0; }
__host__ __device__ constexpr int hz(int a) { return a >= 4 ? 1 : 0; }
// Q4 local node offsets: (0,0), (1,0), (1,1), (0,1).
__host__ __device__ constexpr int qx(int a) { return (a == 1 || a == 2) ? 1 : 0; }
__host__ __device__ constexpr int qy(int a) { return (a == 2 || a == 3) ? 1 : 0; }

// Wavefront (64-lane) sum via DPP-backed shuffles.
template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum for blockDim.x == 256 (4 waves); result valid in thread 0.
template <typename T>
__device__ inline T block_sum256(T v) {
  __shared__ T part[4];
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) r = (part[0] + part[1]) + (part[2] + part[3]);
  __syncthreads();
  return r;
}

// block_sum256 with the total returned to every thread of the block
template <typename T>
__device__ inline T block_sum256_all(T v) {
  __shared__ T tot;
  v = block_sum256(v);
  if (threadIdx.x == 0) tot = v;
  __syncthreads();
  return tot;
}

// Global accesses at a 32-bit element index from a kernel-argument base: the
// byte offset is formed in 32 bits and zero-extended, so the access takes the
// scalar-base + vector-offset form with no 64-bit address arithmetic per access
// (the host checks that every byte offset fits in 32 bits).
// Element i + k: the constant k goes into the instruction's offset field.
template <typename T>
__device__ inline T ld32(const T* __restrict__ p, int i, int k = 0) {
  const char* a = reinterpret_cast<const char*>(p) + (size_t)((unsigned)i * (unsigned)sizeof(T));
  return *reinterpret_cast<const T*>(a + (size_t)k * sizeof(T));
}
template <typename T>
__device__ inline void st32(T* __restrict__ p, int i, int k, T v) {
  char* a = reinterpret_cast<char*>(p) + (size_t)((unsigned)i * (unsigned)sizeof(T));
  *reinterpret_cast<T*>(a + (size_t)k * sizeof(T)) = v;
}

}  // namespace et
