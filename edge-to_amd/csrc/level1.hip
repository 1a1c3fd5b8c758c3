// Matrix-free operator of the first coarse multigrid level.
//
// With even fine element counts every fine element is one of the 8 children
// t of a coarse element C, and the Galerkin operator A_1 = R A_hat R^T
// (R = D^-1 P^T, D the total weights tw) splits into
//     A_1 = D^-1 ( sum_C Q_C^T K_C Q_C ) D^-1 + dA,   K_C = sum_t E_t M_t,
// where M_t = W_t^T K0 W_t and dA collects everything constrained fine DOFs
// change (masked children and the identity rows); dA is non-zero only at the
// coarse nodes of coarse elements with a constrained child.
//
// Element product in the corner Walsh-Hadamard basis (8-point transform H per
// displacement component): reflecting child 0 onto child t is diagonal there,
// so  H K_C H / 64 = Mh o Eh  with a fixed 24 x 24 matrix Mh (279 structural
// non-zeros, level1_terms.hpp) and Eh[(M,c),(N,d)] = (H E_C)[M^N^e_c^e_d],
// H E_C being the transform of the 8 children's densities (computed once per
// solve).  Per coarse element: one forward transform of the tw-scaled corner
// values, 279 + 150 FMAs, one inverse transform.
//
// Passes per application, no atomics, fixed summation order:
//   1. one thread per coarse element computes its 8 corner results; the top
//      (bz = 1) ones move one lane up to the element above, so each element
//      stores the 4 node sums of its bottom face (12 values, plane by plane);
//   2. one thread per coarse node sums the bottom-face sums of its (up to 4)
//      element columns, scales by 1/tw and applies the mode epilogue (y = Ax,
//      b - Ax, Jacobi, Jacobi + add); nodes with a correction row store the
//      raw sum instead;
//   3. one thread per (corrected node, component) adds its 27 correction
//      blocks and applies the epilogue.
#include <cmath>
#include <vector>

#include "common.hpp"
#include "galerkin_elem.hpp"
#include "level1.hpp"
#include "level1_terms.hpp"

namespace et {

static int g_level1_mf = 1;
void set_level1_mf(int on) { g_level1_mf = on ? 1 : 0; }
int get_level1_mf() { return g_level1_mf; }

bool level1_supported(const Grid& g0, int n_level) {
  return g0.dim == 3 && n_level >= 2 && galerkin_elem_supported(g0);
}

// corner-bit order (beta = bx + 2 by + 4 bz) -> H8 local corner index
__host__ __device__ constexpr int l1_corner(int beta) {
  return (beta & 4 ? 4 : 0) + ((beta & 2) ? ((beta & 1) ? 2 : 3) : ((beta & 1) ? 1 : 0));
}

__device__ __forceinline__ double l1_inv_tw(const Grid& g, int i, int j, int k) {
  double t = (i == 0 || i == g.nx() - 1) ? 1.5 : 2.0;
  t *= (j == 0 || j == g.ny() - 1) ? 1.5 : 2.0;
  t *= (k == 0 || k == g.nz() - 1) ? 1.5 : 2.0;
  return 1.0 / t;
}

// the same weight in the working precision, as a product of per-axis factors
template <typename T>
__device__ __forceinline__ T l1_inv_tw_t(const Grid& g, int i, int j, int k) {
  const T bnd = T(2.0 / 3.0), mid = T(0.5);
  T t = (i == 0 || i == g.nx() - 1) ? bnd : mid;
  t *= (j == 0 || j == g.ny() - 1) ? bnd : mid;
  t *= (k == 0 || k == g.nz() - 1) ? bnd : mid;
  return t;
}

// in-place 8-point Walsh-Hadamard transform, natural order
template <typename T>
__device__ __forceinline__ void l1_wht8(T (&v)[8]) {
#pragma unroll
  for (int bit = 1; bit < 8; bit <<= 1)
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (!(i & bit)) {
        const T a = v[i], b = v[i | bit];
        v[i] = a + b;
        v[i | bit] = a - b;
      }
}

// ------------------------------------------------------------ setup ----

template <typename T>
__global__ void l1_eh_kernel(const Grid g0, const Grid g1, const double* __restrict__ E,
                             const uint8_t* __restrict__ eflag, T* __restrict__ Eh,
                             uint8_t* __restrict__ special) {
  const int C = blockIdx.x * blockDim.x + threadIdx.x;
  if (C >= (int)g1.n_elem()) return;
  const int K = C % g1.ez, IJ = C / g1.ez, J = IJ % g1.ey, I = IJ / g1.ey;
  double v[8];
  uint8_t sp = 0;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const int64_t e = ((int64_t)(2 * I + (t & 1)) * g0.ey + 2 * J + ((t >> 1) & 1)) * g0.ez + 2 * K + (t >> 2);
    v[t] = E[e];
    sp |= eflag[e];
  }
  l1_wht8(v);
#pragma unroll
  for (int m = 0; m < 8; ++m) Eh[(int64_t)C * 8 + m] = (T)v[m];
  special[C] = sp ? 1 : 0;
}

__global__ void l1_affected_kernel(const Grid g1, const uint8_t* __restrict__ special,
                                   int* __restrict__ aidx, int* __restrict__ alist, int* __restrict__ count) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int)g1.n_nodes()) return;
  const int k = n % g1.nz(), ij = n / g1.nz(), i = ij % g1.nx(), j = ij / g1.nx();
  bool hit = false;
  for (int beta = 0; beta < 8; ++beta) {
    const int I = i - (beta & 1), J = j - ((beta >> 1) & 1), K = k - (beta >> 2);
    if (I < 0 || I >= g1.ex || J < 0 || J >= g1.ey || K < 0 || K >= g1.ez) continue;
    if (special[(I * g1.ey + J) * g1.ez + K]) hit = true;
  }
  int idx = -1;
  if (hit) {
    idx = atomicAdd(count, 1);
    alist[idx] = n;
  }
  aidx[n] = idx;
}

// dA for listed node c and offset o: the true Galerkin block S1 minus the
// block of the unconstrained element formula (child tables only, no identity).
__global__ void l1_delta_kernel(const Grid g0, const Grid g1, const double* __restrict__ Mt,
                                const double* __restrict__ E, const double* __restrict__ S1,
                                const int* __restrict__ alist, int n_aff, double* __restrict__ dA) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n_aff * 27) return;
  const int o = tid / n_aff, a = tid - o * n_aff;
  const int c = alist[a];
  const int64_t Nc = g1.n_nodes();
  const int ck = c % g1.nz(), cij = c / g1.nz(), ci = cij % g1.nx(), cj = cij / g1.nx();
  const int oi = o / 9 - 1, oj = (o / 3) % 3 - 1, ok = o % 3 - 1;
  const int ti = ci + oi, tj = cj + oj, tk = ck + ok;
  double blk[3][3] = {};
  const bool tin = ti >= 0 && ti < g1.nx() && tj >= 0 && tj < g1.ny() && tk >= 0 && tk < g1.nz();
  if (tin) {
    const int I0 = oi == 0 ? ci - 1 : (oi > 0 ? ci : ci - 1), I1 = oi == 0 ? ci : I0;
    const int J0 = oj == 0 ? cj - 1 : (oj > 0 ? cj : cj - 1), J1 = oj == 0 ? cj : J0;
    const int K0_ = ok == 0 ? ck - 1 : (ok > 0 ? ck : ck - 1), K1_ = ok == 0 ? ck : K0_;
    for (int I = I0; I <= I1; ++I) {
      if (I < 0 || I >= g1.ex) continue;
      for (int J = J0; J <= J1; ++J) {
        if (J < 0 || J >= g1.ey) continue;
        for (int K = K0_; K <= K1_; ++K) {
          if (K < 0 || K >= g1.ez) continue;
          const int ca = l1_corner((ci - I) + 2 * (cj - J) + 4 * (ck - K));
          const int cb = l1_corner((ti - I) + 2 * (tj - J) + 4 * (tk - K));
          for (int t = 0; t < 8; ++t) {
            const int tb = (t == 1 || t == 2 || t == 5 || t == 6 ? 1 : 0) + 2 * (t == 2 || t == 3 || t == 6 || t == 7 ? 1 : 0) +
                           4 * (t >= 4 ? 1 : 0);
            const int64_t e = ((int64_t)(2 * I + (tb & 1)) * g0.ey + 2 * J + ((tb >> 1) & 1)) * g0.ez + 2 * K + (tb >> 2);
            const double Ee = E[e];
            const double* m = Mt + (((size_t)t * 8 + ca) * 8 + cb) * 9;
            for (int r = 0; r < 3; ++r)
              for (int s = 0; s < 3; ++s) blk[r][s] += Ee * m[r * 3 + s];
          }
        }
      }
    }
    const double inv = l1_inv_tw(g1, ci, cj, ck) * l1_inv_tw(g1, ti, tj, tk);
    for (int r = 0; r < 3; ++r)
      for (int s = 0; s < 3; ++s) blk[r][s] *= inv;
  }
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) {
      const double full = tin ? S1[((int64_t)(o * 3 + r) * 3 + s) * Nc + c] : 0.0;
      dA[((int64_t)(o * 3 + r) * 3 + s) * n_aff + a] = full - blk[r][s];
    }
}

template <typename TI, typename TO>
__global__ void l1_convert_kernel(int64_t n, const TI* __restrict__ a, TO* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (TO)a[t];
}

// ------------------------------------------------------------ apply ----

// lane j takes lane j-1's value (lane 0 takes 0)
__device__ __forceinline__ float l1_shr(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0u, __builtin_bit_cast(unsigned, v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ double l1_shr(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_update_dpp(0u, (unsigned)u, 0x138, 0xf, 0xf, false);
  const unsigned hi = __builtin_amdgcn_update_dpp(0u, (unsigned)(u >> 32), 0x138, 0xf, 0xf, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Element pass.  Thread = coarse element C (lane = C mod 64).  The bz = 1
// corner results of element C-1 move one lane up and are added to the bz = 0
// results of C when both lie in one column and one wavefront, so lo holds per
// element the 4 node sums (bx, by) x 3 of its bottom face; results that cannot
// move (lane 63, the top element of a column) are stored in top instead.
// Plane (bx + 2 by) * 3 + c of lo / top, element index C.
template <typename T>
__global__ __launch_bounds__(256) void l1_elem_kernel(const Grid g1, const T* __restrict__ cf,
                                                      const T* __restrict__ Eh, const T* __restrict__ x,
                                                      T* __restrict__ lo, T* __restrict__ top) {
  const int ne = (int)g1.n_elem();
  const int C0 = blockIdx.x * blockDim.x + threadIdx.x;
  const bool valid = C0 < ne;
  const int C = valid ? C0 : ne - 1;  // every lane takes part in the lane shift
  const int lane = threadIdx.x & 63;
  const int K = C % g1.ez, IJ = C / g1.ez, J = IJ % g1.ey, I = IJ / g1.ey;
  const int nx = g1.nx(), nz = g1.nz();
  // the element's transformed densities are requested before the corner values
  // are waited on (they are needed only after the forward transforms)
  T eh[8];
#pragma unroll
  for (int m = 0; m < 8; ++m) eh[m] = Eh[(int64_t)C * 8 + m];
  T xh[3][8];
#pragma unroll
  for (int beta = 0; beta < 8; ++beta) {
    const int i = I + (beta & 1), j = J + ((beta >> 1) & 1), k = K + (beta >> 2);
    const int n = (j * nx + i) * nz + k;
    const T sw = l1_inv_tw_t<T>(g1, i, j, k);
#pragma unroll
    for (int c = 0; c < 3; ++c) xh[c][beta] = x[n * 3 + c] * sw;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) l1_wht8(xh[c]);
  T yh[3][8];
  // the coefficients are wave-uniform: scalar loads into SGPRs (staging them
  // in LDS lets the compiler hoist every read and quadruples the VGPRs)
  level1_product(xh, eh, [&](int k) -> T { return cf[k]; }, yh);
#pragma unroll
  for (int c = 0; c < 3; ++c) l1_wht8(yh[c]);
  const bool from_prev = lane > 0 && K > 0;          // lane-1 holds C-1 of this column
  const bool keep_top = lane == 63 || K == g1.ez - 1;  // C's bz = 1 results stay here
#pragma unroll
  for (int f = 0; f < 4; ++f)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T up = yh[c][f + 4];
      const T moved = l1_shr(up);
      if (valid) {
        lo[(int64_t)(f * 3 + c) * ne + C] = yh[c][f] + (from_prev ? moved : T(0));
        if (keep_top) top[(int64_t)(f * 3 + c) * ne + C] = up;
      }
    }
}

// Node pass: sum of the (up to 4) element columns' bottom-face sums, plus the
// unmoved top results of the element below where there are any; scale by
// 1/tw; epilogue.  Nodes with a correction row store the raw sum (finished by
// l1_fix_kernel).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void l1_node_kernel(const Grid g1, const T* __restrict__ lo,
                                                      const T* __restrict__ top,
                                                      const int* __restrict__ aidx, int n_aff,
                                                      const T* __restrict__ dA, const T* __restrict__ x,
                                                      const T* __restrict__ b, const T* __restrict__ Dinv,
                                                      const double* __restrict__ omega, double wscale,
                                                      const T* __restrict__ xadd, T* __restrict__ y) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= (int)g1.n_nodes()) return;
  const int nx = g1.nx(), nz = g1.nz();
  const int k = n % nz, ij = n / nz, i = ij % nx, j = ij / nx;
  const int ne = (int)g1.n_elem();
  T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
  for (int f = 0; f < 4; ++f) {
    const int I = i - (f & 1), J = j - (f >> 1);
    const bool col = I >= 0 && I < g1.ex && J >= 0 && J < g1.ey;
    const int Cb = (I * g1.ey + J) * g1.ez;  // first element of the column
    // element k of the column: its bottom face (with C-1's moved results)
    const bool in = col && k < g1.ez;
    const int C = in ? Cb + k : 0;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      const T v = lo[(int64_t)(f * 3 + r) * ne + C];
      acc[r] += in ? v : T(0);
    }
    // element k-1 whose top results were not moved (rare: lane 63 or column top)
    const int Cm = Cb + k - 1;
    if (col && k > 0 && ((Cm & 63) == 63 || k == g1.ez)) {
#pragma unroll
      for (int r = 0; r < 3; ++r) acc[r] += top[(int64_t)(f * 3 + r) * ne + Cm];
    }
  }
  const T s = l1_inv_tw_t<T>(g1, i, j, k);
#pragma unroll
  for (int r = 0; r < 3; ++r) acc[r] *= s;
  if (aidx[n] >= 0) {
    // corrected node: the raw product goes to y, l1_fix_kernel finishes it
#pragma unroll
    for (int r = 0; r < 3; ++r) y[(int64_t)n * 3 + r] = acc[r];
    return;
  }
  const T w = MODE >= 2 ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    const int64_t d = (int64_t)n * 3 + r;
    T v;
    if (MODE == 0)
      v = acc[r];
    else if (MODE == 1)
      v = b[d] - acc[r];
    else
      v = x[d] + w * Dinv[d] * (b[d] - acc[r]);
    if (MODE == 3) v += xadd[d];
    y[d] = v;
  }
}

// Corrected nodes: y = epilogue(raw + dA x), one thread per (listed node, component).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void l1_fix_kernel(const Grid g1, const int* __restrict__ alist, int n_aff,
                                                     const T* __restrict__ dA, const T* __restrict__ x,
                                                     const T* __restrict__ b, const T* __restrict__ Dinv,
                                                     const double* __restrict__ omega, double wscale,
                                                     const T* __restrict__ xadd, T* __restrict__ y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_aff * 3) return;
  const int r = t / n_aff, a = t - r * n_aff;
  const int n = alist[a];
  const int nx = g1.nx(), ny = g1.ny(), nz = g1.nz();
  const int k = n % nz, ij = n / nz, i = ij % nx, j = ij / nx;
  const int64_t d = (int64_t)n * 3 + r;
  T acc = y[d];
  // all 81 products are independent: out-of-grid neighbours read the node
  // itself and are weighted by zero, so every load can be in flight at once
#pragma unroll
  for (int o = 0; o < 27; ++o) {
    const int mi = i + o / 9 - 1, mj = j + (o / 3) % 3 - 1, mk = k + o % 3 - 1;
    const bool in = mi >= 0 && mi < nx && mj >= 0 && mj < ny && mk >= 0 && mk < nz;
    const int m = in ? (mj * nx + mi) * nz + mk : n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T coef = dA[(int64_t)((o * 3 + r) * 3 + c) * n_aff + a];
      acc += (in ? coef : T(0)) * x[(int64_t)m * 3 + c];
    }
  }
  const T w = MODE >= 2 ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);
  T v;
  if (MODE == 0)
    v = acc;
  else if (MODE == 1)
    v = b[d] - acc;
  else
    v = x[d] + w * Dinv[d] * (b[d] - acc);
  if (MODE == 3) v += xadd[d];
  y[d] = v;
}

// ------------------------------------------------------------ host ----

namespace {
inline int l1_at(int B, int c, int D, int d) { return ((B * 3 + c) * 8 + D) * 3 + d; }

// Mh = H M_0 H / 64, M_0 = W_0^T K0 W_0 in corner-bit order, [M][c][N][d]
std::vector<double> level1_mh(const double* K0_host) {
  double Kb[8][3][8][3];
  for (int b = 0; b < 8; ++b)
    for (int b2 = 0; b2 < 8; ++b2)
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) Kb[b][c][b2][d] = K0_host[(3 * l1_corner(b) + c) * 24 + 3 * l1_corner(b2) + d];
  double W[8][8];  // child 0: [child corner b][coarse corner B]
  for (int b = 0; b < 8; ++b)
    for (int B = 0; B < 8; ++B) {
      double w = 1.0;
      for (int d = 0; d < 3; ++d) {
        const int f = (b >> d) & 1, cpos = 2 * ((B >> d) & 1);
        w *= std::fmax(0.0, 1.0 - std::fabs((double)(f - cpos)) / 2.0);
      }
      W[b][B] = w;
    }
  std::vector<double> M0(8 * 3 * 8 * 3, 0.0), Mh(8 * 3 * 8 * 3, 0.0);
  for (int B = 0; B < 8; ++B)
    for (int c = 0; c < 3; ++c)
      for (int D = 0; D < 8; ++D)
        for (int d = 0; d < 3; ++d) {
          double v = 0.0;
          for (int b = 0; b < 8; ++b)
            for (int b2 = 0; b2 < 8; ++b2) v += W[b][B] * Kb[b][c][b2][d] * W[b2][D];
          M0[l1_at(B, c, D, d)] = v;
        }
  auto h = [](int m, int b) { return (__builtin_popcount(m & b) & 1) ? -1.0 : 1.0; };
  for (int M = 0; M < 8; ++M)
    for (int c = 0; c < 3; ++c)
      for (int N = 0; N < 8; ++N)
        for (int d = 0; d < 3; ++d) {
          double v = 0.0;
          for (int B = 0; B < 8; ++B)
            for (int D = 0; D < 8; ++D) v += h(M, B) * M0[l1_at(B, c, D, d)] * h(N, D);
          Mh[l1_at(M, c, N, d)] = v / 64.0;
        }
  return Mh;
}
}  // namespace

bool level1_terms_cover(const double* K0_host, double rtol) {
  const std::vector<double> Mh = level1_mh(K0_host);
  std::vector<char> kept(Mh.size(), 0);
  for (int t = 0; t < kL1Terms; ++t) kept[l1_at(kL1Term[t][1], kL1Term[t][0], kL1Term[t][3], kL1Term[t][2])] = 1;
  double mx = 0.0;
  for (double v : Mh) mx = std::fmax(mx, std::fabs(v));
  for (size_t i = 0; i < Mh.size(); ++i)
    if (!kept[i] && std::fabs(Mh[i]) > rtol * mx) return false;
  return true;
}

int Level1MF::init(const Grid& fine, const double* K0_host, bool mixed_, hipStream_t s) {
  g0 = fine;
  g1 = fine.coarse();
  mixed = mixed_;
  const std::vector<double> Mh = level1_mh(K0_host);
  auto at = l1_at;
  std::vector<double> cd(kL1Terms);
  std::vector<float> cfl(kL1Terms);
  for (int t = 0; t < kL1Terms; ++t) {
    cd[t] = Mh[at(kL1Term[t][1], kL1Term[t][0], kL1Term[t][3], kL1Term[t][2])];
    cfl[t] = (float)cd[t];
  }
  const size_t ts = mixed ? sizeof(float) : sizeof(double);
  ET_CHECK_HIP(hipMalloc(&cf, kL1Terms * ts));
  ET_CHECK_HIP(hipMemcpy(cf, mixed ? (const void*)cfl.data() : (const void*)cd.data(), kL1Terms * ts,
                         hipMemcpyHostToDevice));
  const int64_t ne1 = g1.n_elem(), nn1 = g1.n_nodes();
  ET_CHECK_HIP(hipMalloc(&Eh, ne1 * 8 * ts));
  ET_CHECK_HIP(hipMalloc(&scratch, ne1 * 24 * ts));
  ET_CHECK_HIP(hipMalloc(&special, ne1));
  ET_CHECK_HIP(hipMalloc(&aidx, nn1 * sizeof(int)));
  ET_CHECK_HIP(hipMalloc(&alist, nn1 * sizeof(int)));
  ET_CHECK_HIP(hipMalloc(&acount, sizeof(int)));
  ET_CHECK_HIP(hipHostMalloc(&acount_h, sizeof(int)));
  (void)s;
  return ET_OK;
}

int Level1MF::setup(const double* E, const double* Mt, const uint8_t* eflag, const double* S1, hipStream_t s) {
  const int64_t ne1 = g1.n_elem(), nn1 = g1.n_nodes();
  if (mixed)
    l1_eh_kernel<float><<<grid_for(ne1), kBlock, 0, s>>>(g0, g1, E, eflag, (float*)Eh, special);
  else
    l1_eh_kernel<double><<<grid_for(ne1), kBlock, 0, s>>>(g0, g1, E, eflag, (double*)Eh, special);
  ET_CHECK_LAUNCH();
  ET_CHECK_HIP(hipMemsetAsync(acount, 0, sizeof(int), s));
  l1_affected_kernel<<<grid_for(nn1), kBlock, 0, s>>>(g1, special, aidx, alist, acount);
  ET_CHECK_LAUNCH();
  ET_CHECK_HIP(hipMemcpyAsync(acount_h, acount, sizeof(int), hipMemcpyDeviceToHost, s));
  ET_CHECK_HIP(hipStreamSynchronize(s));
  n_aff = *acount_h;
  if (n_aff == 0) return ET_OK;
  if (n_aff > dA_cap) {
    if (dA64) (void)hipFree(dA64);
    if (dA) (void)hipFree(dA);
    dA64 = nullptr;
    dA = nullptr;
    dA_cap = n_aff;
    ET_CHECK_HIP(hipMalloc(&dA64, (size_t)dA_cap * 243 * sizeof(double)));
    if (mixed) ET_CHECK_HIP(hipMalloc(&dA, (size_t)dA_cap * 243 * sizeof(float)));
  }
  l1_delta_kernel<<<grid_for((int64_t)n_aff * 27), kBlock, 0, s>>>(g0, g1, Mt, E, S1, alist, n_aff, dA64);
  ET_CHECK_LAUNCH();
  if (mixed) {
    const int64_t cnt = (int64_t)n_aff * 243;
    l1_convert_kernel<double, float><<<grid_for(cnt), kBlock, 0, s>>>(cnt, dA64, (float*)dA);
    ET_CHECK_LAUNCH();
  }
  return ET_OK;
}

template <typename T>
int Level1MF::apply(int mode, const T* x, const T* b, const T* Dinv, const double* omega, double wscale,
                    const T* xadd, T* y, hipStream_t s) const {
  ET_REQUIRE((sizeof(T) == sizeof(float)) == mixed, "level-1 operator: precision mismatch");
  const T* dAt = mixed ? (const T*)dA : (const T*)dA64;
  T* lo = (T*)scratch;
  T* top = lo + 12 * g1.n_elem();
  l1_elem_kernel<T><<<grid_for(g1.n_elem()), 256, 0, s>>>(g1, (const T*)cf, (const T*)Eh, x, lo, top);
  const unsigned gn = grid_for(g1.n_nodes());
  const unsigned gf = grid_for((int64_t)n_aff * 3);
#define ET_L1(M)                                                                                     \
  do {                                                                                               \
    l1_node_kernel<T, M><<<gn, 256, 0, s>>>(g1, lo, top, aidx, n_aff, dAt, x, b, Dinv, omega, wscale, \
                                            xadd, y);                                                \
    if (n_aff > 0)                                                                                   \
      l1_fix_kernel<T, M><<<gf, 256, 0, s>>>(g1, alist, n_aff, dAt, x, b, Dinv, omega, wscale, xadd, y); \
  } while (0)
  switch (mode) {
    case 0: ET_L1(0); break;
    case 1: ET_L1(1); break;
    case 2: ET_L1(2); break;
    default: ET_L1(3); break;
  }
#undef ET_L1
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int Level1MF::apply<float>(int, const float*, const float*, const float*, const double*, double,
                                    const float*, float*, hipStream_t) const;
template int Level1MF::apply<double>(int, const double*, const double*, const double*, const double*, double,
                                     const double*, double*, hipStream_t) const;

void Level1MF::release() {
  for (void* p : {cf, Eh, scratch, (void*)special, (void*)aidx, (void*)alist, (void*)acount, (void*)dA64, dA})
    if (p) (void)hipFree(p);
  if (acount_h) (void)hipHostFree(acount_h);
  cf = Eh = scratch = dA = nullptr;
  special = nullptr;
  aidx = alist = acount = acount_h = nullptr;
  dA64 = nullptr;
  n_aff = dA_cap = 0;
}

}  // namespace et
