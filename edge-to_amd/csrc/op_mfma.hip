// Fine-grid H8 operator on the matrix cores (FP32 in/out, FP16 MFMA with a
// hi/lo split, FP32 accumulation): the V-cycle's fine-level sweeps in
// mixed precision.
//
// Element-centric contraction.  For a group of 16 elements the 32 x 16 block
//     V^T = K0p * U^T
// is formed with v_mfma_f32_16x16x32_f16, where U^T holds each element's 8
// corner displacements padded to (x, y, z, 0) (K = 32) and K0p is K0 padded
// the same way (M = 32, two 16-row tiles).  Each operand is split into an
// FP16 high part and an FP16 remainder; hi*hi + hi*lo + lo*hi recovers ~22
// bits.  Every element's 24 values are scaled by a power of two that maps
// their maximum into [0.5, 1) before the split, so no FP16 underflow or
// overflow occurs whatever the vector's magnitude; the scale, E_e and the K0
// scale are applied to the FP32 result.
//
// Tiling.  A workgroup (4 wavefronts) owns 7 node columns in x (all z) and
// marches over JC node planes in y.  Per element plane it stages in LDS the
// FP32 displacements of the two node planes the plane touches (9 node
// columns, constrained DOFs zeroed) and the E values of its 8 element
// columns.  Element contributions are added into two node-plane
// accumulators in LDS (the plane below and the plane above the element
// plane).  The scatter is colour-ordered: in phase p every wavefront takes
// element column 2w + p, so concurrently processed columns never share a
// node, and within a wavefront the elements of one MFMA group are distinct
// along z; no atomics are used and the summation order is fixed.  Once an
// element plane is done the node plane below it is complete and goes
// through the mode epilogue (y = Ax, b - Ax, Jacobi, Jacobi + add).
#include <cmath>

#include "common.hpp"
#include "op_tiled.hpp"
#include "stiffness.hpp"

namespace et {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float4_t __attribute__((ext_vector_type(4)));

constexpr int kMfBX = 7;             // output node columns per workgroup
constexpr int kMfCols = kMfBX + 1;   // element columns per workgroup
constexpr int kMfUCols = kMfBX + 2;  // node columns staged

struct MfmaArgs {
  Grid g;
  int64_t e_sx, e_sy;
  int jc;
  float kscale;  // K0 was divided by this power of two before the split
};

__host__ inline size_t mfma_lds_bytes(const Grid& g) {
  const size_t nz = g.nz();
  return 2 * kMfUCols * nz * 16 + 2 * kMfBX * nz * 16 + (size_t)kMfCols * g.ez * 4;
}

__device__ inline void split_half(float v, _Float16& hi, _Float16& lo) {
  hi = (_Float16)v;
  lo = (_Float16)(v - (float)hi);
}

template <int MODE>
__global__ __launch_bounds__(256) void op0_mfma_kernel(const MfmaArgs A, const float* __restrict__ K0,
                                                       const float* __restrict__ E,
                                                       const uint8_t* __restrict__ cons,
                                                       const float* __restrict__ x,
                                                       const float* __restrict__ b,
                                                       const float* __restrict__ Dinv,
                                                       const double* __restrict__ omega,
                                                       double wscale, const float* __restrict__ xadd,
                                                       float* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const Grid& g = A.g;
  const int nx = g.nx(), ny = g.ny(), nz = g.nz();
  float4_t* su = reinterpret_cast<float4_t*>(lds_raw);         // [2][kMfUCols][nz]
  float4_t* sa = su + 2 * kMfUCols * nz;                         // [2][kMfBX][nz]
  float* se = reinterpret_cast<float*>(sa + 2 * kMfBX * nz);     // [kMfCols][nelz]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.x * kMfBX;
  const int j0 = blockIdx.y * A.jc;
  const int j1 = min(j0 + A.jc, ny);
  const int64_t plane = (int64_t)nx * nz;

  // ---- constant A operands: K0 padded to 32x32, two 16-row tiles, hi/lo ----
  half8_t ka_hi[2], ka_lo[2];
  {
    const int row16 = lane & 15, kq = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      const int row = 16 * mt + row16;
      const int a = row >> 2, r = row & 3;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int k = 8 * kq + v;
        const int bb = k >> 2, c = k & 3;
        const float kv = (r < 3 && c < 3) ? K0[(3 * a + r) * 24 + 3 * bb + c] / A.kscale : 0.0f;
        _Float16 hi, lo;
        split_half(kv, hi, lo);
        ka_hi[mt][v] = hi;
        ka_lo[mt][v] = lo;
      }
    }
  }

  // ---- zero both accumulator planes ----
  for (int t = tid; t < 2 * kMfBX * nz; t += blockDim.x) sa[t] = float4_t{0.f, 0.f, 0.f, 0.f};

  auto load_u = [&](int jp) {  // node plane jp -> slot jp & 1 (columns i0-1 .. i0+kMfBX)
    float4_t* dst = su + (size_t)(jp & 1) * kMfUCols * nz;
    const bool in = jp >= 0 && jp < ny;
    const int64_t q0 = (int64_t)(i0 - 1) * nz;
    const int nvals = kMfUCols * nz * 3;
    for (int t = tid; t < nvals; t += blockDim.x) {
      const int node = t / 3, c = t - 3 * node;
      const int64_t q = q0 + node;
      float v = 0.f;
      if (in && q >= 0 && q < plane) {
        const int64_t d = ((int64_t)jp * plane + q) * 3 + c;
        v = (cons != nullptr && cons[d]) ? 0.f : x[d];
      }
      reinterpret_cast<float*>(dst + node)[c] = v;
    }
    for (int t = tid; t < kMfUCols * nz; t += blockDim.x) reinterpret_cast<float*>(dst + t)[3] = 0.f;
  };
  auto load_e = [&](int ey) {
    for (int t = tid; t < kMfCols * g.ez; t += blockDim.x) {
      const int col = t / g.ez, ez = t - col * g.ez;
      const int ex = i0 - 1 + col;
      float v = 0.f;
      if (ex >= 0 && ex < g.ex) v = E[(int64_t)ex * A.e_sx + (int64_t)ey * A.e_sy + ez];
      se[t] = v;
    }
  };

  if (j0 - 1 >= 0) load_u(j0 - 1);
  const int ngroups = (g.ez + 15) >> 4;
  const int ql = lane >> 4, el = lane & 15;
  for (int ey = j0 - 1; ey < j1; ++ey) {
    const bool has_elems = ey >= 0 && ey < g.ey;
    load_u(ey + 1);  // zero-filled beyond the grid
    if (has_elems) load_e(ey);
    __syncthreads();
    if (has_elems) {
      const float4_t* ulo = su + (size_t)(ey & 1) * kMfUCols * nz;        // node plane ey
      const float4_t* uhi = su + (size_t)((ey + 1) & 1) * kMfUCols * nz;  // node plane ey+1
      float4_t* alo = sa + (size_t)(ey & 1) * kMfBX * nz;
      float4_t* ahi = sa + (size_t)((ey + 1) & 1) * kMfBX * nz;
#pragma unroll
      for (int phase = 0; phase < 2; ++phase) {
        const int col = 2 * wave + phase;  // element column within the tile
        for (int gi = 0; gi < ngroups; ++gi) {
          const int ez = 16 * gi + el;
          const bool ev = ez < g.ez;
          const int ezc = ev ? ez : g.ez - 1;
          // B operand: corners 2*ql, 2*ql+1 of element (col, ey, ez), padded xyz0
          float u[8];
#pragma unroll
          for (int h = 0; h < 2; ++h) {
            const int cb = 2 * ql + h;
            const int ucol = col + hx(cb);  // staged column index (column i0-1 is 0)
            const float4_t* pl = hy(cb) ? uhi : ulo;
            const float4_t val = pl[ucol * nz + ezc + hz(cb)];
            u[4 * h + 0] = ev ? val[0] : 0.f;
            u[4 * h + 1] = ev ? val[1] : 0.f;
            u[4 * h + 2] = ev ? val[2] : 0.f;
            u[4 * h + 3] = 0.f;
          }
          // per-element power-of-two scale (max over the 4 lanes holding the element)
          float m = 0.f;
#pragma unroll
          for (int v = 0; v < 8; ++v) m = fmaxf(m, fabsf(u[v]));
          m = fmaxf(m, __shfl_xor(m, 16, 64));
          m = fmaxf(m, __shfl_xor(m, 32, 64));
          int ex2 = 0;
          (void)frexpf(m, &ex2);
          const float inv = m > 0.f ? ldexpf(1.0f, -ex2) : 0.f;
          const float scl = m > 0.f ? ldexpf(1.0f, ex2) : 0.f;
          half8_t bh, bl;
#pragma unroll
          for (int v = 0; v < 8; ++v) {
            _Float16 hi, lo;
            split_half(u[v] * inv, hi, lo);
            bh[v] = hi;
            bl[v] = lo;
          }
          const float Ee = ev ? se[col * g.ez + ez] : 0.f;
          const float f = Ee * scl * A.kscale;
#pragma unroll
          for (int mt = 0; mt < 2; ++mt) {
            float4_t d = {0.f, 0.f, 0.f, 0.f};
            d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ka_hi[mt], bh, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ka_hi[mt], bl, d, 0, 0, 0);
            d = __builtin_amdgcn_mfma_f32_16x16x32_f16(ka_lo[mt], bh, d, 0, 0, 0);
            // lane holds corner a = 4*mt + ql of element el: rows 4*ql + v -> components v
            const int a = 4 * mt + ql;
            const int ncol = col + hx(a) - 1;  // output node column (tile-relative)
            if (ev && ncol >= 0 && ncol < kMfBX) {
              float4_t* dst = (hy(a) ? ahi : alo) + ncol * nz + ez + hz(a);
              float4_t cur = *dst;
              cur[0] += f * d[0];
              cur[1] += f * d[1];
              cur[2] += f * d[2];
              *dst = cur;
            }
          }
        }
        __syncthreads();
      }
    }
    // node plane ey is complete: epilogue (plane j0-1 belongs to the tile below:
    // its accumulator is only cleared)
    {
      float4_t* acc = sa + (size_t)(ey & 1) * kMfBX * nz;
      const float w = MODE >= TM_JAC ? (float)((omega != nullptr ? *omega : 1.0) * wscale) : 0.f;
      const bool out = ey >= j0;
      for (int t = tid; t < kMfBX * nz; t += blockDim.x) {
        const int ncol = t / nz, k = t - ncol * nz;
        const int i = i0 + ncol;
        const float4_t av = acc[t];
        acc[t] = float4_t{0.f, 0.f, 0.f, 0.f};
        if (!out || i >= nx) continue;
        const int64_t n = ((int64_t)ey * nx + i) * nz + k;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          const int64_t d = n * 3 + r;
          const float xv = x[d];
          const float ax = (cons != nullptr && cons[d]) ? xv : av[r];
          float v;
          if (MODE == TM_Y)
            v = ax;
          else if (MODE == TM_RES)
            v = b[d] - ax;
          else
            v = xv + w * Dinv[d] * (b[d] - ax);
          if (MODE == TM_JAC_ADD) v += xadd[d];
          y[d] = v;
        }
      }
    }
    __syncthreads();
  }
}

static int g_mfma_jc = 8;
static int g_mixed_mfma = 0;
void set_mixed_mfma(int on) { g_mixed_mfma = on ? 1 : 0; }
int get_mixed_mfma() { return g_mixed_mfma; }

float k0_scale(const double* K0_host, int n) {
  double m = 0.0;
  for (int t = 0; t < n; ++t) m = std::fmax(m, std::fabs(K0_host[t]));
  int e = 0;
  (void)std::frexp(m, &e);
  return m > 0.0 ? (float)std::ldexp(1.0, e) : 1.0f;
}
void set_mfma_jc(int jc) { g_mfma_jc = jc > 0 ? jc : 8; }
int get_mfma_jc() { return g_mfma_jc; }

bool mfma_supported(const Grid& g) { return g.dim == 3 && mfma_lds_bytes(g) <= 160 * 1024; }

template <int MODE>
static void launch_mfma(dim3 grid, size_t lds, hipStream_t s, const MfmaArgs& A, const float* K0,
                        const float* E, const uint8_t* cons, const float* x, const float* b,
                        const float* Dinv, const double* omega, double wscale, const float* xadd,
                        float* y) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&op0_mfma_kernel<MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  op0_mfma_kernel<MODE><<<grid, 256, lds, s>>>(A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y);
}

int op0_mfma(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, float kscale, const float* K0,
             const float* E, const uint8_t* cons, const float* x, const float* b, const float* Dinv,
             const double* omega, double wscale, const float* xadd, float* y, hipStream_t s) {
  ET_REQUIRE(mfma_supported(g), "MFMA operator: 3-D grids with nz <= ~290 only");
  MfmaArgs A{g, e_sx, e_sy, g_mfma_jc, kscale};
  const size_t lds = mfma_lds_bytes(g);
  dim3 grid((unsigned)((g.nx() + kMfBX - 1) / kMfBX), (unsigned)((g.ny() + A.jc - 1) / A.jc));
  switch (mode) {
    case TM_Y: launch_mfma<TM_Y>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_RES: launch_mfma<TM_RES>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_JAC: launch_mfma<TM_JAC>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    default: launch_mfma<TM_JAC_ADD>(grid, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

}  // namespace et
