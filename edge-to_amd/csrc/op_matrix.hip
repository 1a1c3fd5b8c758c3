// Fine-grid H8 operator on the matrix cores at full working precision:
// v_mfma_f32_16x16x4f32 for FP32 vectors, v_mfma_f64_16x16x4f64 for FP64.
//
// Element-centric contraction.  A group of 16 elements consecutive along z
// (one per MFMA column j) is multiplied by K0:
//     D (32 x 16) = K0p (32 x 24) * U (24 x 16)
// in two 16-row tiles of six K=4 steps.  Output rows are padded per corner to
// (x, y, z, 0), so row 4g+v of tile t is component v of corner (hx = g&1,
// hy = t, hz = g>>1); the MFMA result layout puts corner g of both tiles in
// lane group g = lane/16.  The K ordering is chosen so that lane group g
// supplies the two corners (hx = g&1, hy = g>>1, hz = 0/1): its six operand
// values are two consecutive z nodes of one staged node column.
//
// Assembly without atomics, in a fixed order:
//  * E_e scales column j (lane j of every group);
//  * z: the hz = 1 rows are moved onto the hz = 0 rows (permlane32 swap) and
//    shifted one lane up the row (DPP row_ror:1); lane 0 takes the value that
//    lane 15 of the previous group produced, carried in a register;
//  * y: the hy = 1 tile of an element layer is carried in registers to the
//    next layer, where it is added to the hy = 0 tile;
//  * x: the two per-column partial results are written to LDS and the node
//    phase adds the contribution of element column i (hx = 0) and of column
//    i-1 (hx = 1).
//
// Tiling.  A workgroup owns x tiles of 8 element columns (2 per wavefront)
// that complete 7 node columns, over all z, and marches in y.  The grid is
// persistent: the (x tile, node plane) units are split evenly over the
// launched workgroups, each taking a contiguous range; a range starts with
// one extra element layer whose top half seeds the y carry.  LDS holds two
// staged node planes (SoA, constrained DOFs zeroed), the partial results of
// one layer and the E values of one layer; the next plane and layer are
// loaded into registers while the matrix cores work.
#include "common.hpp"
#include "op_tiled.hpp"
#include "stiffness.hpp"

namespace et {

constexpr int kMxCols = 8;                // element columns per x tile (2 per wavefront)
constexpr int kMxOut = kMxCols - 1;       // node columns completed per x tile
constexpr int kMxNodeCols = kMxCols + 1;  // node columns staged per plane
constexpr int kMxZG = 8;                  // z groups of 16 elements (nelz <= 128)
constexpr int kMxStage = 5;               // staged nodes per thread per plane (9 * 129 <= 1280)
constexpr int kMxEStage = 4;              // staged E values per thread per layer (8 * 128 <= 1024)

typedef float mxf4 __attribute__((ext_vector_type(4)));
typedef double mxd4 __attribute__((ext_vector_type(4)));
template <typename T>
struct MxAcc;
template <>
struct MxAcc<float> {
  typedef mxf4 type;
};
template <>
struct MxAcc<double> {
  typedef mxd4 type;
};

__device__ inline mxf4 mx_mfma(float a, float b, mxf4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ inline mxd4 mx_mfma(double a, double b, mxd4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// lanes 0-31 receive lanes 32-63, then each row of 16 lanes is rotated by one
// (lane j takes lane j-1, lane 0 takes lane 15)
__device__ inline unsigned mx_shift32(unsigned v) {
  const auto sw = __builtin_amdgcn_permlane32_swap(v, v, false, false);
  return __builtin_amdgcn_update_dpp(0u, sw[1], 0x121, 0xf, 0xf, false);
}
__device__ inline float mx_shift(float v) {
  return __builtin_bit_cast(float, mx_shift32(__builtin_bit_cast(unsigned, v)));
}
__device__ inline double mx_shift(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = mx_shift32((unsigned)u), hi = mx_shift32((unsigned)(u >> 32));
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

__host__ inline size_t mx_lds_values(const Grid& g) {
  const size_t nz = g.nz(), span = (size_t)kMxNodeCols * nz;
  return 6 * span + 6 * (size_t)kMxCols * nz + (size_t)kMxCols * g.ez;
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void op0_mx_kernel(const MxArgs A, const T* __restrict__ K0,
                                                     const T* __restrict__ E,
                                                     const uint8_t* __restrict__ cons,
                                                     const T* __restrict__ x,
                                                     const T* __restrict__ b,
                                                     const T* __restrict__ Dinv,
                                                     const double* __restrict__ omega,
                                                     double wscale, const T* __restrict__ xadd,
                                                     T* __restrict__ y) {
  typedef typename MxAcc<T>::type acc_t;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const Grid& g = A.g;
  const int nx = g.nx(), ny = g.ny(), nz = g.nz();
  const int nelx = g.ex, nely = g.ey, nelz = g.ez;
  const int span = kMxNodeCols * nz;
  T* su = reinterpret_cast<T*>(lds_raw);  // [2 slots][3 comps][span]
  T* sw = su + 6 * span;                  // [3 comps][2 hx][kMxCols][nz]
  T* se = sw + 6 * kMxCols * nz;          // [kMxCols][nelz]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j16 = lane & 15, lg = lane >> 4;
  const int gx = lg & 1, gh = lg >> 1;
  const int64_t plane = (int64_t)nx * nz;

  // A operands: row (lane&15) of tile t, K column 4kk + lg.  Result row i of a
  // tile lands in lane group i/4, register i%4 for the FP32 instruction and in
  // lane group i%4, register i/4 for the FP64 one; the row order is chosen so
  // that lane group g receives corner g, register v its component v.
  T a0[6], a1[6];
  {
    const int r = lane & 15;
    const int oc = sizeof(T) == 8 ? (r & 3) : (r >> 2);
    const int comp = sizeof(T) == 8 ? (r >> 2) : (r & 3);
    const int ohx = oc & 1, ohz = oc >> 1;
#pragma unroll
    for (int kk = 0; kk < 6; ++kk) {
      const int ci = h8_corner(lg & 1, lg >> 1, kk / 3), ic = kk % 3;
      a0[kk] = comp < 3 ? K0[(3 * h8_corner(ohx, 0, ohz) + comp) * 24 + 3 * ci + ic] : T(0);
      a1[kk] = comp < 3 ? K0[(3 * h8_corner(ohx, 1, ohz) + comp) * 24 + 3 * ci + ic] : T(0);
    }
  }
  const T w = MODE >= TM_JAC ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);

  // staged-node and E slots of this thread: (local column, z) of s = tid + 256 m
  int snl[kMxStage], sk[kMxStage], ecl[kMxEStage], eez[kMxEStage];
#pragma unroll
  for (int m = 0; m < kMxStage; ++m) {
    const int s = tid + 256 * m;
    snl[m] = s / nz;
    sk[m] = s - snl[m] * nz;
  }
#pragma unroll
  for (int m = 0; m < kMxEStage; ++m) {
    const int t = tid + 256 * m;
    ecl[m] = t / nelz;
    eez[m] = t - ecl[m] * nelz;
  }

  const int64_t u0 = (int64_t)blockIdx.x * A.units / gridDim.x;
  const int64_t u1 = (int64_t)(blockIdx.x + 1) * A.units / gridDim.x;
  int64_t un = u0;
  while (un < u1) {
    const int xt = (int)(un / ny);
    const int ja = (int)(un - (int64_t)xt * ny);
    const int jb = (int)min((int64_t)ny, (int64_t)ja + (u1 - un));
    un += jb - ja;
    const int c0 = xt * kMxOut - 1;  // element column of local column 0 = node column of staged column 0

    T pu[kMxStage][3];
    T pe[kMxEStage];
    auto fetch_u = [&](int jp) {
      const bool pin = jp >= 0 && jp < ny;
      const int64_t pb = (int64_t)(pin ? jp : 0) * plane * 3 + (int64_t)c0 * nz * 3;
#pragma unroll
      for (int m = 0; m < kMxStage; ++m) {
        const int s = tid + 256 * m;
        const int n = c0 + snl[m];
        const bool in = pin && s < span && n >= 0 && n < nx;
#pragma unroll
        for (int r = 0; r < 3; ++r) {
          T v = T(0);
          if (in) {
            const int64_t d = pb + (int64_t)s * 3 + r;
            v = x[d];
            if (cons != nullptr && cons[d]) v = T(0);
          }
          pu[m][r] = v;
        }
      }
    };
    auto store_u = [&](int jp) {
      T* dst = su + (size_t)(jp & 1) * 3 * span;
#pragma unroll
      for (int m = 0; m < kMxStage; ++m) {
        const int s = tid + 256 * m;
        if (s < span) {
#pragma unroll
          for (int r = 0; r < 3; ++r) dst[r * span + s] = pu[m][r];
        }
      }
    };
    auto fetch_e = [&](int ey) {
      const bool lin = ey >= 0 && ey < nely;
#pragma unroll
      for (int m = 0; m < kMxEStage; ++m) {
        const int c = c0 + ecl[m];
        const bool in = lin && ecl[m] < kMxCols && c >= 0 && c < nelx;
        pe[m] = in ? E[(int64_t)c * A.e_sx + (int64_t)ey * A.e_sy + eez[m]] : T(0);
      }
    };
    auto store_e = [&]() {
#pragma unroll
      for (int m = 0; m < kMxEStage; ++m) {
        const int t = tid + 256 * m;
        if (t < kMxCols * nelz) se[t] = pe[m];
      }
    };

    // prologue: node planes ja-1 and ja, element layer ja-1
    fetch_u(ja - 1);
    store_u(ja - 1);
    fetch_u(ja);
    store_u(ja);
    fetch_e(ja - 1);
    store_e();
    __syncthreads();

    T yc[2][kMxZG][3];  // y carry: hy = 1 half of the previous layer
    T yt[2][3];         // y carry of the node k = nelz when nelz is a multiple of 16
#pragma unroll
    for (int cc = 0; cc < 2; ++cc) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        yt[cc][r] = T(0);
#pragma unroll
        for (int zg = 0; zg < kMxZG; ++zg) yc[cc][zg][r] = T(0);
      }
    }

    for (int ey = ja - 1; ey < jb; ++ey) {
      const bool out = ey >= ja;
      const bool more = ey + 2 <= jb;
      if (more) {
        fetch_u(ey + 2);
        fetch_e(ey + 1);
      }

      // matrix-core phase: element layer ey, two element columns per wavefront
#pragma unroll
      for (int cc = 0; cc < 2; ++cc) {
        const int cl = 2 * wave + cc;
        const T* ub = su + (size_t)((ey + gh) & 1) * 3 * span + (cl + gx) * nz;
        T zc0[3] = {T(0), T(0), T(0)}, zc1[3] = {T(0), T(0), T(0)};
#pragma unroll
        for (int zg = 0; zg < kMxZG; ++zg) {
          if (zg < A.nzg) {
            const int ez = 16 * zg + j16;
            const int ezc = min(ez, nelz - 1);
            T bv[6];
#pragma unroll
            for (int kk = 0; kk < 6; ++kk) bv[kk] = ub[(kk % 3) * span + ezc + kk / 3];
            acc_t d0 = {T(0), T(0), T(0), T(0)}, d1 = {T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int kk = 0; kk < 6; ++kk) {
              d0 = mx_mfma(a0[kk], bv[kk], d0);
              d1 = mx_mfma(a1[kk], bv[kk], d1);
            }
            const T e = ez < nelz ? se[cl * nelz + ez] : T(0);
            const int k = 16 * zg + j16;
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              const T p0 = d0[r] * e, p1 = d1[r] * e;
              const T s0 = mx_shift(p0), s1 = mx_shift(p1);
              const T z0 = p0 + (j16 == 0 ? zc0[r] : s0);
              const T z1 = p1 + (j16 == 0 ? zc1[r] : s1);
              zc0[r] = s0;
              zc1[r] = s1;
              const T yv = z0 + yc[cc][zg][r];
              yc[cc][zg][r] = z1;
              if (out && lane < 32 && k < nz) sw[((r * 2 + gx) * kMxCols + cl) * nz + k] = yv;
            }
          }
        }
        if (nelz == 16 * A.nzg) {
          // node k = nelz: only the hz = 1 half of element nelz-1, now in lane 0
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            const T yv = zc0[r] + yt[cc][r];
            yt[cc][r] = zc1[r];
            if (out && lane < 32 && j16 == 0) sw[((r * 2 + gx) * kMxCols + cl) * nz + nelz] = yv;
          }
        }
      }
      __syncthreads();

      // node phase: node plane ey, then stage the prefetched plane and layer.
      // All global operands of the plane are requested before any is used.
      if (out) {
        const int64_t jb3 = (int64_t)ey * plane * 3 + (int64_t)c0 * nz * 3;
        const T* own = su + (size_t)(ey & 1) * 3 * span;
        bool act[kMxStage];
        int64_t dd[kMxStage];
        T bq[kMxStage][3], dq[kMxStage][3];
        uint8_t cq[kMxStage][3];
#pragma unroll
        for (int m = 0; m < kMxStage; ++m) {
          const int s = tid + 256 * m;
          act[m] = s < span && snl[m] >= 1 && snl[m] < kMxCols && c0 + snl[m] < nx;
          dd[m] = act[m] ? jb3 + (int64_t)s * 3 : 0;
#pragma unroll
          for (int r = 0; r < 3; ++r) {
            cq[m][r] = cons != nullptr ? cons[dd[m] + r] : (uint8_t)0;
            bq[m][r] = MODE != TM_Y ? b[dd[m] + r] : T(0);
            dq[m][r] = MODE >= TM_JAC ? Dinv[dd[m] + r] : T(0);
          }
        }
#pragma unroll
        for (int m = 0; m < kMxStage; ++m) {
          if (act[m]) {
            const int s = tid + 256 * m, nl = snl[m], k = sk[m];
#pragma unroll
            for (int r = 0; r < 3; ++r) {
              const int64_t d = dd[m] + r;
              const T ax0 = sw[((r * 2) * kMxCols + nl) * nz + k] + sw[((r * 2 + 1) * kMxCols + nl - 1) * nz + k];
              const bool cr = cq[m][r] != 0;
              T xv = own[r * span + s];  // zero on constrained DOFs
              if (cr) xv = x[d];
              const T ax = cr ? xv : ax0;
              T v;
              if (MODE == TM_Y)
                v = ax;
              else if (MODE == TM_RES)
                v = bq[m][r] - ax;
              else
                v = xv + w * dq[m][r] * (bq[m][r] - ax);
              if (MODE == TM_JAC_ADD) v += xadd[d];
              y[d] = v;
            }
          }
        }
      }
      if (more) {
        store_u(ey + 2);
        store_e();
      }
      __syncthreads();
    }
  }
}

static int g_mx_wgs_per_cu[2] = {2, 1};  // [float, double]
static int g_mx_enable[2] = {1, 0};

void set_mx(int is_double, int on) { g_mx_enable[is_double ? 1 : 0] = on ? 1 : 0; }
int get_mx(int is_double) { return g_mx_enable[is_double ? 1 : 0]; }
void set_mx_wgs(int is_double, int n) { g_mx_wgs_per_cu[is_double ? 1 : 0] = n > 0 ? n : 1; }
int get_mx_wgs(int is_double) { return g_mx_wgs_per_cu[is_double ? 1 : 0]; }

constexpr size_t kMxMaxLds = 160 * 1024;

bool mx_supported(const Grid& g, size_t elem_bytes) {
  return g.dim == 3 && g.ez <= 16 * kMxZG && kMxNodeCols * g.nz() <= 256 * kMxStage &&
         kMxCols * g.ez <= 256 * kMxEStage && mx_lds_values(g) * elem_bytes <= kMxMaxLds;
}

static int device_cus() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
      cus = 256;
  }
  return cus;
}

template <typename T, int MODE>
static void launch_mx(int n_wg, size_t lds, hipStream_t s, const MxArgs& A, const T* K0, const T* E,
                      const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
                      double wscale, const T* xadd, T* y) {
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&op0_mx_kernel<T, MODE>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMxMaxLds);
    attr_set = true;
  }
  op0_mx_kernel<T, MODE><<<n_wg, 256, lds, s>>>(A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y);
}

template <typename T>
int op0_mx(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const T* K0, const T* E,
           const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
           double wscale, const T* xadd, T* y, hipStream_t s) {
  ET_REQUIRE(mx_supported(g, sizeof(T)), "matrix-core operator: 3-D grids with nelz <= 128 only");
  MxArgs A;
  A.g = g;
  A.e_sx = e_sx;
  A.e_sy = e_sy;
  A.n_xt = (g.nx() + kMxOut - 1) / kMxOut;
  A.nzg = (g.ez + 15) / 16;
  A.units = (int64_t)A.n_xt * g.ny();
  const size_t lds = mx_lds_values(g) * sizeof(T);
  const int64_t slots = (int64_t)device_cus() * g_mx_wgs_per_cu[sizeof(T) == 8 ? 1 : 0];
  const int n_wg = (int)(A.units < slots ? A.units : slots);
  switch (mode) {
    case TM_Y: launch_mx<T, TM_Y>(n_wg, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_RES: launch_mx<T, TM_RES>(n_wg, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    case TM_JAC: launch_mx<T, TM_JAC>(n_wg, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
    default: launch_mx<T, TM_JAC_ADD>(n_wg, lds, s, A, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y); break;
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

template int op0_mx<double>(int, const Grid&, int64_t, int64_t, const double*, const double*,
                            const uint8_t*, const double*, const double*, const double*,
                            const double*, double, const double*, double*, hipStream_t);
template int op0_mx<float>(int, const Grid&, int64_t, int64_t, const float*, const float*,
                           const uint8_t*, const float*, const float*, const float*,
                           const double*, double, const float*, float*, hipStream_t);

}  // namespace et

namespace et {

bool fine_supported(const Grid& g, size_t elem_bytes) {
  return (get_wht() && wht_supported(g, elem_bytes)) || (get_mx(elem_bytes == 8) && mx_supported(g, elem_bytes)) ||
         tiled_supported(g, elem_bytes);
}

template <typename T>
int op0_fine(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const T* K0, const WhtTable& tab, const T* E,
             const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
             double wscale, const T* xadd, T* y, hipStream_t s) {
  if (get_wht() && wht_supported(g, sizeof(T)))
    return op0_wht<T>(mode, g, e_sx, e_sy, tab, E, cons, x, b, Dinv, omega, wscale, xadd, y, s);
  ET_REQUIRE(mode >= 0 && mode <= 3, "fine operator: mode not available without the transformed-basis operator");
  if (get_mx(sizeof(T) == 8) && mx_supported(g, sizeof(T)))
    return op0_mx<T>(mode, g, e_sx, e_sy, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y, s);
  return op0_tiled<T>(mode, g, e_sx, e_sy, K0, E, cons, x, b, Dinv, omega, wscale, xadd, y, s);
}

bool fine_jac0_supported(const Grid& g, size_t elem_bytes) { return get_wht() && wht_supported(g, elem_bytes); }

template int op0_fine<double>(int, const Grid&, int64_t, int64_t, const double*, const WhtTable&, const double*,
                              const uint8_t*, const double*, const double*, const double*,
                              const double*, double, const double*, double*, hipStream_t);
template int op0_fine<float>(int, const Grid&, int64_t, int64_t, const float*, const WhtTable&, const float*,
                             const uint8_t*, const float*, const float*, const float*,
                             const double*, double, const float*, float*, hipStream_t);

}  // namespace et
