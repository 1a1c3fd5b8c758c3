// LDS-tiled, node-centric H8 operator for the fine grid (device code).
//
// A workgroup owns a strip of 256 consecutive nodes of one x-z node plane
// (plane index p = i*nz + k, contiguous in memory) and marches over JC
// consecutive y planes.  LDS holds a ring of four u planes (the strip plus a
// halo of nz+1 nodes on each side, SoA per component, constrained entries
// zeroed) and a ring of three E planes (the element columns the strip
// touches).  The loop is software-pipelined: while plane j is computed, the
// global loads of u plane j+2, E plane j+1 and of this plane's epilogue
// operands are in flight in registers; they are written to LDS after the
// compute, so one barrier per plane suffices and memory latency overlaps the
// FMAs.  Each thread forms
//     y_n = sum_a E_a * sum_o K0[a, a+o] u_{n+o}
// visiting the 27 neighbours o once (81 LDS reads) and keeping one partial
// 3-vector per adjacent element a.  Every output is written once: no
// atomics, fixed summation order.
//
// Element matrix: a structured-grid element is an axis-aligned box, so K0 is
// invariant under the reflections of the box.  Reflecting by the corner bits
// of a maps corner a to 0 and flips the sign of the reflected displacement
// components:  K0[a, b] = R_a K0[0, a^b] R_a  (corner bits XORed, R_a the
// diagonal sign matrix).  Only the first block row of K0 (8 blocks of 3x3,
// 72 values) is therefore needed; it is read with scalar loads and the signs
// are compile-time constants folded into the FMAs.  For a fixed neighbour
// offset o every pair (a, a+o) uses the same base block, the one whose XOR
// pattern marks the non-zero components of o.
#pragma once
#include "common.hpp"

namespace et {

constexpr int kStrip = 256;
constexpr int kUSlots = 3;  // u nodes staged per thread per plane (u_span <= 768)
constexpr int kESlots = 4;  // E values staged per thread per plane (e_span <= 1024)

// Base-block entry (r, block m, c) = K0[r][3m + c]: wave-uniform, compile-time
// offsets, so it is fetched with scalar loads into SGPRs.
#define ET_TILED_K(r, m, c) K0[(r) * 24 + 3 * (m) + (c)]

// Transformed-basis operator only: TM_YDOT, y = A x plus per-workgroup partial
// sums of x.y; TM_JAC0, the Jacobi sweep from the zero guess fused with its
// first step, y = x0 + w D^-1 (b - A x0) with x0 = w D^-1 b (x is not read).
enum TiledMode { TM_Y = 0, TM_RES = 1, TM_JAC = 2, TM_JAC_ADD = 3, TM_YDOT = 4, TM_JAC0 = 5, TM_JAC_ADD_DOT = 6 };

struct TiledArgs {
  Grid g;
  int64_t e_sx, e_sy;  // element index = ex*e_sx + ey*e_sy + ez
  int jc;              // y planes per workgroup
  int halo;            // nz + 1
  int u_span;          // kStrip + 2*halo
  int e_span;          // element slots per element plane
};

__host__ inline TiledArgs tiled_args(const Grid& g, int64_t e_sx, int64_t e_sy, int jc) {
  TiledArgs a;
  a.g = g;
  a.e_sx = e_sx;
  a.e_sy = e_sy;
  a.jc = jc;
  a.halo = g.nz() + 1;
  a.u_span = kStrip + 2 * a.halo;
  // x columns spanned by a strip (+1 on the low side for the elements behind it)
  const int cols = kStrip / g.nz() + 3;
  a.e_span = cols * g.ez;
  return a;
}

__host__ inline bool tiled_shape_ok(const TiledArgs& a) {
  return a.g.dim == 3 && a.u_span <= kUSlots * kStrip && a.e_span <= kESlots * kStrip;
}

template <typename T>
__host__ inline size_t tiled_lds_bytes(const TiledArgs& a) {
  return (size_t)4 * 3 * a.u_span * sizeof(T) + (size_t)3 * a.e_span * sizeof(T);
}

// Arguments of the element-centric fine-grid operators (matrix-core and
// transformed-basis kernels): persistent grid over (x tile, node plane) units.
struct MxArgs {
  Grid g;
  int64_t e_sx, e_sy;  // element index = ex*e_sx + ey*e_sy + ez
  int n_xt;            // x tiles
  int nzg;             // z groups of 16 elements (matrix-core kernel)
  int64_t units;       // n_xt * ny
};

// offset bit of H8 corner a along axis d (0 = x, 1 = y, 2 = z)
__host__ __device__ constexpr int h8_bit(int a, int d) { return d == 0 ? hx(a) : (d == 1 ? hy(a) : hz(a)); }

// corner index of the H8 node with offset bits (bx, by, bz)
__host__ __device__ constexpr int h8_corner(int bx, int by, int bz) {
  return (bz ? 4 : 0) + (by ? (bx ? 2 : 3) : (bx ? 1 : 0));
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void op0_tiled_kernel(const TiledArgs A, const T* __restrict__ K0,
                                                        const T* __restrict__ E,
                                                        const uint8_t* __restrict__ cons,
                                                        const T* __restrict__ x,
                                                        const T* __restrict__ b,
                                                        const T* __restrict__ Dinv,
                                                        const double* __restrict__ omega,
                                                        double wscale, const T* __restrict__ xadd,
                                                        T* __restrict__ y) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* su = reinterpret_cast<T*>(lds_raw);  // [4 planes][3 comps][u_span]
  T* se = su + 12 * A.u_span;             // [3 planes][e_span]
  const Grid& g = A.g;
  const int nx = g.nx(), ny = g.ny(), nz = g.nz();
  const int plane = nx * nz;
  const int p0 = blockIdx.x * kStrip;
  const int j0 = blockIdx.y * A.jc;
  const int j1 = min(j0 + A.jc, ny);
  const int tid = threadIdx.x;
  const int p = p0 + tid;
  const bool active = p < plane;
  const int i = p / nz, k = p - (p / nz) * nz;
  const int ex0 = p0 / nz - 1;    // first element column staged
  const int ubase = p0 - A.halo;  // plane index of LDS slot 0

  // per-thread fixed staging slots
  int uq[kUSlots];  // plane index of the u node staged by this thread
  bool uin[kUSlots];
#pragma unroll
  for (int m = 0; m < kUSlots; ++m) {
    const int s = tid + m * kStrip;
    const int q = ubase + s;
    uq[m] = q;
    uin[m] = s < A.u_span && q >= 0 && q < plane;
  }
  int eoff[kESlots];  // ex*e_sx + ez (fits 32 bits: < n_elem)
  bool ein[kESlots];
#pragma unroll
  for (int m = 0; m < kESlots; ++m) {
    const int t = tid + m * kStrip;
    const int col = t / g.ez, ez = t - col * g.ez;
    const int ex = ex0 + col;
    ein[m] = t < A.e_span && ex >= 0 && ex < g.ex;
    eoff[m] = ein[m] ? (int)((int64_t)ex * A.e_sx + ez) : 0;
  }

  T pu[kUSlots][3];
  T pe[kESlots];
  auto fetch_u = [&](int jp) {
    const bool in = jp >= 0 && jp < ny;
    // plane base pointers are wave-uniform; per-thread offsets stay 32-bit
    const int64_t pb = (int64_t)(in ? jp : 0) * plane * 3;
    const T* xp = x + pb;
    const uint8_t* cp = cons != nullptr ? cons + pb : nullptr;
#pragma unroll
    for (int m = 0; m < kUSlots; ++m) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T v = T(0);
        if (in && uin[m]) {
          const int d = uq[m] * 3 + c;
          v = xp[d];
          if (cp != nullptr && cp[d]) v = T(0);
        }
        pu[m][c] = v;
      }
    }
  };
  auto store_u = [&](int jp) {
    T* dst = su + (size_t)((jp + 4) & 3) * 3 * A.u_span;
#pragma unroll
    for (int m = 0; m < kUSlots; ++m) {
      const int s = tid + m * kStrip;
      if (s < A.u_span) {
#pragma unroll
        for (int c = 0; c < 3; ++c) dst[c * A.u_span + s] = pu[m][c];
      }
    }
  };
  auto fetch_e = [&](int ey) {
    const bool in = ey >= 0 && ey < g.ey;
    const T* ep = E + (int64_t)(in ? ey : 0) * A.e_sy;
#pragma unroll
    for (int m = 0; m < kESlots; ++m) pe[m] = (in && ein[m]) ? ep[eoff[m]] : T(0);
  };
  auto store_e = [&](int ey) {
    T* dst = se + (size_t)((ey + 3) % 3) * A.e_span;
#pragma unroll
    for (int m = 0; m < kESlots; ++m) {
      const int t = tid + m * kStrip;
      if (t < A.e_span) dst[t] = pe[m];
    }
  };

  // prologue: u planes j0-1, j0, j0+1 and E planes j0-1, j0
  fetch_u(j0 - 1);
  store_u(j0 - 1);
  fetch_u(j0);
  store_u(j0);
  fetch_u(j0 + 1);
  store_u(j0 + 1);
  fetch_e(j0 - 1);
  store_e(j0 - 1);
  fetch_e(j0);
  store_e(j0);
  __syncthreads();

  const T w = MODE >= TM_JAC ? (T)((omega != nullptr ? *omega : 1.0) * wscale) : T(0);
  for (int j = j0; j < j1; ++j) {
    const bool more = j + 1 < j1;
    if (more) {
      fetch_u(j + 2);
      fetch_e(j + 1);
    }
    // epilogue operands of this plane, issued before the FMAs
    const int64_t jb = (int64_t)j * plane * 3;  // wave-uniform plane offset
    const int pd = p * 3;                        // per-thread offset in the plane
    T bv[3] = {T(0), T(0), T(0)}, dv[3] = {T(0), T(0), T(0)};
    uint8_t cv[3] = {0, 0, 0};
    if (active) {
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        if (cons != nullptr) cv[r] = cons[jb + pd + r];
        if (MODE != TM_Y) bv[r] = b[jb + pd + r];
        if (MODE >= TM_JAC) dv[r] = Dinv[jb + pd + r];
      }
    }

    if (active) {
      T s[8][3];
#pragma unroll
      for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int r = 0; r < 3; ++r) s[a][r] = T(0);
      const int lp = tid + A.halo;  // LDS slot of this node
      // Offsets are visited grouped by their base block (XOR pattern mb): one
      // 3x3 block is live at a time, each neighbour is read once.
#pragma unroll
      for (int mb = 0; mb < 8; ++mb) {
        const int ax = mb & 1, ay = (mb >> 1) & 1, az = (mb >> 2) & 1;
        T kv[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
          for (int c = 0; c < 3; ++c) kv[r][c] = ET_TILED_K(r, h8_corner(ax, ay, az), c);
#pragma unroll
        for (int sc = 0; sc < 8; ++sc) {
          const int sx = sc & 1, sy = (sc >> 1) & 1, sz = (sc >> 2) & 1;
          if ((!ax && sx) || (!ay && sy) || (!az && sz)) continue;
          const int di = ax ? (sx ? 1 : -1) : 0;
          const int dj = ay ? (sy ? 1 : -1) : 0;
          const int dk = az ? (sz ? 1 : -1) : 0;
          const T* pl = su + (size_t)((j + dj + 4) & 3) * 3 * A.u_span;
          const int q = lp + di * nz + dk;
          // At k = 0 / k = nz-1 the slot q +- 1 wraps into the neighbouring x column;
          // such values only feed elements outside the grid, whose E is zero.
          const T u[3] = {pl[q], pl[A.u_span + q], pl[2 * A.u_span + q]};
#pragma unroll
          for (int a = 0; a < 8; ++a) {
            const int bx = hx(a) + di, by = hy(a) + dj, bz = hz(a) + dk;
            if (bx < 0 || bx > 1 || by < 0 || by > 1 || bz < 0 || bz > 1) continue;
            // (R_a K R_a)[r][c] = sign(a, r) sign(a, c) K[r][c]: a compile-time sign
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                const bool neg = (h8_bit(a, r) ^ h8_bit(a, c)) != 0;
                s[a][r] = neg ? fma(-kv[r][c], u[c], s[a][r]) : fma(kv[r][c], u[c], s[a][r]);
              }
          }
        }
      }
      T acc[3] = {T(0), T(0), T(0)};
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const int ex = i - hx(a), ey = j - hy(a), ez = k - hz(a);
        const bool zin = ez >= 0 && ez < g.ez;
        const T Ea = se[(size_t)((ey + 3) % 3) * A.e_span + (ex - ex0) * g.ez + (zin ? ez : 0)];
#pragma unroll
        for (int r = 0; r < 3; ++r) acc[r] += (zin ? Ea : T(0)) * s[a][r];
      }
      const T* own = su + (size_t)((j + 4) & 3) * 3 * A.u_span;
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        const int64_t d = jb + pd + r;
        T xv = own[r * A.u_span + lp];  // zero on constrained DOFs
        if (cv[r]) xv = x[d];
        const T ax = cv[r] ? xv : acc[r];
        T v;
        if (MODE == TM_Y)
          v = ax;
        else if (MODE == TM_RES)
          v = bv[r] - ax;
        else
          v = xv + w * dv[r] * (bv[r] - ax);
        if (MODE == TM_JAC_ADD) v += xadd[d];
        y[d] = v;
      }
    }
    if (more) {
      store_u(j + 2);
      store_e(j + 1);
    }
    __syncthreads();
  }
}

}  // namespace et
