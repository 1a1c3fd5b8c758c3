// Linear-hat density filter on structured element grids: normaliser, forward
// (normalised weighted average over in-bounds neighbours), exact transpose, and
// the SIMP adjoint fused with the transpose.
//
// Forward and transpose are LDS-halo stencils.  A workgroup owns a tile of
// TY x TZ element columns (3-D TY x 64 with TY 4 or 8, a wavefront per row along the
// fastest axis; 2-D 1 x 256) and marches along x through a chunk of planes (4 to 32, as
// many as keep about two workgroups per compute unit busy).  Each step stages one
// new element plane of the tile plus its y/z halo into a ring of 8 plane slots
// in LDS (zero outside the grid) and produces one output plane from the ring,
// so every input value is read from memory once per tile (plus the halo rows)
// and every output is written once, with no atomics.  Sums follow the offset
// table's order, so results are deterministic; the forward pass is bitwise the
// reference's (out-of-grid taps add an exact 0.0).  The forward's denominator
// depends only on the output's distances to the six faces clipped to the radius,
// so each workgroup tabulates it once per boundary class (at most 3^6 classes).
//
//   forward    y_i = (sum_k w_k x_{i+o_k}) / (sum_k [i+o_k in grid] w_k)
//              (apply_filter_{2D,3D}_kernel, core/_templated_cuda.cu:1526-1583)
//   transpose  y_j = sum_k w_k (v / N)_{j+o_k}, the offsets being symmetric; v / N is
//              formed once per staged value, not once per tap
//              (apply_filter_{2D,3D}_transpose_kernel, core/_templated_cuda.cu:1632-1679)
//   adjoint    y_j = sum_k w_k s_{j+o_k} / N_{j+o_k} with the SIMP sensitivity
//              s_e = p clip(xt_e, void, 1)^(p-1) (-u_e^T K0 u_e) computed while staging
//              (process_dk + SingleMaterial.__backward__ + the transpose of
//              CuTopOpt.solve_and_process, TopOpt/_cuda.py:293-300), in one kernel.
//
// In 3-D the element energy is evaluated in the corner Walsh-Hadamard basis,
// u^T K0 u = sum over the kept couplings of uh[M][c] G[M][c][d] uh[M^e_c^e_d][d]
// with uh = H u_e and G = H K0 H / 64 (about 140 flops instead of 1152).  The
// host picks that form only for element matrices with the brick's reflection
// structure (k0_structure); any other K0 takes the full product.
//
// 2-D grids run as 3-D grids of one element in y (element i * nely + j sits at
// x = i, z = j).  Each workgroup reads the offset radii from the table; taps
// wider than the ring (kMaxR per axis) are gathered from memory directly (the
// adjoint then recomputes the neighbour's sensitivity per tap).
#include "common.hpp"
#include "op_tiled.hpp"
#include "stiffness.hpp"

namespace et {

constexpr int kMaxOff = 1024;
constexpr int kMaxR = 2;     // halo radius per axis held in the stencil kernels' ring (rmin <= 3)
constexpr int kMaxTaps = 512;  // taps held in LDS by the stencil kernels (rmin up to about 4.9)
constexpr int kMaxCX = 32;   // x planes per workgroup, at most

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

// One thread per element (normaliser; fallback for wide stencils).
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

// ------------------------------------------------------- stencil geometry ---

// element grid seen as (X, Y, Z) with index (x * Y + y) * Z + z
struct SGrid {
  int X, Y, Z;
};

__host__ inline SGrid sgrid(const Grid& g) {
  if (g.dim == 3) return SGrid{g.ex, g.ey, g.ez};
  return SGrid{g.ex, 1, g.ey};  // 2-D: (i, j) -> x = i, z = j
}

struct AdjArgs {
  const double* K0;
  const double* U;
  const double* xt;
  double penal, vmin;
  const double* norm;
};

// corner-bit order beta = bx + 2 by + 4 bz of the H8 corners
__device__ __forceinline__ int fcorner(int beta) { return h8_corner(beta & 1, (beta >> 1) & 1, beta >> 2); }

// kept couplings of G[M][c][d] (M = 1..7), as in the transformed-basis operator
__host__ __device__ constexpr bool fk_kept(int M, int c, int d) {
  if (M == 0) return false;
  if (c == d) return true;
  if (M == 7) return false;
  const int pair = (1 << c) | (1 << d);
  const int bits = __builtin_popcount(M);
  return bits == 1 ? (pair & M) != 0 : pair != M;
}

// -u_e^T K0 u_e of element (x, y, z) of a 3-D grid: transformed basis or full product
template <bool WHT>
__device__ __forceinline__ double elem_energy3(const Grid& g, int ex, int ey, int ez, const double* __restrict__ U,
                                               const double* K0, const double* sG) {
  const int nx = g.nx(), nz = g.nz();
  // node index of corner-bit beta = bx + 2 by + 4 bz
  const int64_t m0 = ((int64_t)ey * nx + ex) * nz + ez;
  const int64_t sx = nz, sy = (int64_t)nx * nz;
  double e = 0.0;
  if (WHT) {
    // one displacement component at a time: its 8 corner values, then their
    // 8-point Walsh-Hadamard transform (the memory clobber keeps the three
    // components' loads from being scheduled together: fewer live registers)
    double uh[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double a[8];
#pragma unroll
      for (int b = 0; b < 8; ++b) a[b] = U[(m0 + (b & 1) * sx + ((b >> 1) & 1) * sy + (b >> 2)) * 3 + c];
#pragma unroll
      for (int h = 1; h < 8; h <<= 1)
#pragma unroll
        for (int b = 0; b < 8; ++b)
          if (!(b & h)) {
            const double p = a[b], q = a[b | h];
            a[b] = p + q;
            a[b | h] = p - q;
          }
#pragma unroll
      for (int M = 0; M < 8; ++M) uh[c][M] = a[M];
      asm volatile("" ::: "memory");
    }
#pragma unroll
    for (int M = 1; M < 8; ++M) {
      // the coefficients of one M at a time (a clobber per M bounds the loads in flight)
      asm volatile("" ::: "memory");
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (fk_kept(M, c, d)) e = fma(uh[c][M] * sG[(M - 1) * 9 + c * 3 + d], uh[d][M ^ (1 << c) ^ (1 << d)], e);
    }
  } else {
    // full product in the element's own DOF order
    double u[8][3];  // corner-bit order
#pragma unroll
    for (int beta = 0; beta < 8; ++beta)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        u[beta][c] = U[(m0 + (beta & 1) * sx + ((beta >> 1) & 1) * sy + (beta >> 2)) * 3 + c];
#pragma unroll 1
    for (int p = 0; p < 24; ++p) {
      const int ap = p / 3, cp = p % 3;
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < 24; ++q) {
        const int aq = q / 3, cq = q % 3;
        v += K0[p * 24 + q] * u[hx(aq) + 2 * hy(aq) + 4 * hz(aq)][cq];
      }
      // u[corner of p][component of p], selected without dynamic register indexing
      double up = 0.0;
#pragma unroll
      for (int b = 0; b < 8; ++b)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          up = (b == hx(ap) + 2 * hy(ap) + 4 * hz(ap) && c == cp) ? u[b][c] : up;
      e += v * up;
    }
  }
  return -e;
}

__device__ inline double elem_energy2(const Grid& g, int ex, int ey, const double* __restrict__ U,
                                      const double* K0) {
  const int nx = g.nx();
  double ue[8];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const int64_t m = (int64_t)(ey + qy(b)) * nx + ex + qx(b);
    ue[2 * b] = U[m * 2];
    ue[2 * b + 1] = U[m * 2 + 1];
  }
  double e = 0.0;
#pragma unroll
  for (int p = 0; p < 8; ++p) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < 8; ++q) v += K0[p * 8 + q] * ue[q];
    e += v * ue[p];
  }
  return -e;
}

// out of line: pow's many registers stay out of the stencil loops
__device__ __attribute__((noinline)) double simp_pow(double x, double e) { return pow(x, e); }

// SIMP factor p x^(p-1); for the usual p = 3 (GP false, chosen on the host) x^2 = x * x,
// which is pow(x, 2) correctly rounded, and pow's registers stay out of the kernel
template <bool GP>
__device__ __forceinline__ double simp_factor(double x, double penal) {
  return GP ? penal * simp_pow(x, penal - 1.0) : penal * (x * x);
}

// value staged for element e at (xs, gy, gz) of the stencil grid:
// MODE 1 x_e, 2 v_e / N_e, 3 s_e / N_e (SIMP adjoint; EN 1 transformed-basis energy, 2 full)
template <int MODE, int DIM, int EN>
__device__ __forceinline__ double stencil_value(const Grid& g, const SGrid& sg, int xs, int gy, int gz,
                                                const double* __restrict__ x, const AdjArgs& adj,
                                                const double* sK, const double* sG) {
  const int64_t e = ((int64_t)xs * sg.Y + gy) * sg.Z + gz;
  if (MODE == 1) return x[e];
  if (MODE == 2) return x[e] / adj.norm[e];
  const double xe = fmin(fmax(adj.xt[e], adj.vmin), 1.0);
  const double en = DIM == 3 ? elem_energy3<(EN & 3) == 1>(g, xs, gy, gz, adj.U, sK, sG)
                             : elem_energy2(g, xs, gz, adj.U, sK);
  return simp_factor<(EN & 4) != 0>(xe, adj.penal) * en / adj.norm[e];
}

// Tap sum over the full (2R+1)^DIM cube around the output, in the offset table's
// (dx, dy, dz) ascending order, with the weight of each cube position (0 where the
// table has no tap, and such positions are skipped): the loop is unrolled, every
// LDS address is an immediate offset from one base per x plane.
template <int R, int DIM, int PZ, int PL, int RING>
__device__ __forceinline__ double cube_sum(const double* ring, const double* s_wc, int xo, int me) {
  constexpr int W = 2 * R + 1, WY = DIM == 3 ? W : 1, RY = DIM == 3 ? R : 0;
  double sv = 0.0;
#pragma unroll
  for (int ix = 0; ix < W; ++ix) {
    const double* p = ring + ((xo + ix - R) & (RING - 1)) * PL + me;
#pragma unroll
    for (int iy = 0; iy < WY; ++iy)
#pragma unroll
      for (int iz = 0; iz < W; ++iz) {
        const double wk = s_wc[(ix * WY + iy) * W + iz];
        const double v = p[(iy - RY) * PZ + iz - R];
        sv = wk != 0.0 ? fma(wk, v, sv) : sv;
      }
  }
  return sv;
}

// MODE 1: forward, 2: transpose (values v / N), 3: SIMP adjoint (values s / N).
// Tile: TY x TZ outputs per plane (3-D TY x 64, 2-D 1 x 256), one per thread.  Ring of
// 8 element planes (x mod 8): 2 kMaxR + 1 in use plus the one being filled.  EN (MODE 3, 3-D): 1 = the
// element energy in the transformed basis (element matrices with the brick's reflection
// structure, chosen on the host), 2 = the full product; EN + 4: a penalty other than 3.
template <int MODE, int DIM, int EN, int TY>
__global__ __launch_bounds__(DIM == 3 ? 64 * TY : 256)
__attribute__((amdgpu_waves_per_eu(MODE == 3 ? 2 : 4))) void stencil_kernel(const Grid g, const SGrid sg, int cx,
                                                                          int n_off,
                                                      const int32_t* __restrict__ off,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ x,
                                                      const AdjArgs adj,
                                                      double* __restrict__ y) {
  static_assert(DIM == 3 || TY == 1, "2-D tiles are one row");
  constexpr int TZ = DIM == 3 ? 64 : 256, NT = TY * TZ;
  constexpr int RYM = DIM == 3 ? kMaxR : 0;  // y halo capacity (2-D grids have one y row)
  constexpr int PY = TY + 2 * RYM, PZ = TZ + 2 * kMaxR, PL = PY * PZ;
  constexpr int RING = 8;
  constexpr int NK = MODE == 3 ? (DIM == 3 ? 576 : 64) : 1;
  __shared__ double ring[RING * PL];
  __shared__ int s_tap[kMaxTaps];  // dx + 8 in bits 16+, in-plane LDS offset dy * PZ + dz + 2^15 below
  __shared__ double s_w[kMaxTaps];
  __shared__ double sK[NK];
  __shared__ double sG[MODE == 3 ? 63 : 1];
  __shared__ int s_r[3];
  // MODE 1: weight sum of every boundary class.  An output's in-grid taps depend only on its
  // distances to the six faces clipped to the radius, (lx, hx, ly, hy, lz, hz) in [0, r]^6.
  constexpr int NB = MODE == 1 ? (DIM == 3 ? 729 : 81) : 1;
  __shared__ double s_bw[NB];
  __shared__ int s_d[kMaxTaps];  // dx, dy, dz + 8, 5 bits each
  constexpr int NC = DIM == 3 ? 125 : 25;  // (2 kMaxR + 1)^DIM cube positions
  __shared__ double s_wc[NC];
  __shared__ int s_cube;
  const int tid = threadIdx.x;
  if (tid < 3) s_r[tid] = 0;
  for (int t = tid; t < NC; t += NT) s_wc[t] = 0.0;
  if (MODE == 3)
    for (int t = tid; t < NK; t += NT) sK[t] = adj.K0[t];
  __syncthreads();
  for (int t = tid; t < n_off; t += NT) {
    const int dx = off[t * DIM];
    const int dy = DIM == 3 ? off[t * 3 + 1] : 0;
    const int dz = DIM == 3 ? off[t * 3 + 2] : off[t * 2 + 1];
    s_tap[t] = ((dx + 8) << 16) | (dy * PZ + dz + (1 << 15));
    s_w[t] = w[t];
    s_d[t] = ((dx + 8) << 10) | ((dy + 8) << 5) | (dz + 8);
    atomicMax(&s_r[0], abs(dx));
    atomicMax(&s_r[1], abs(dy));
    atomicMax(&s_r[2], abs(dz));
  }
  if (MODE == 3 && DIM == 3 && (EN & 3) == 1 && tid < 63) {
    // kept couplings of G = H K0 H / 64
    const int M = 1 + tid / 9, c = (tid / 3) % 3, d = tid % 3;
    const int Mp = M ^ (1 << c) ^ (1 << d);
    double sum = 0.0;
    for (int bb = 0; bb < 8; ++bb)
      for (int bp = 0; bp < 8; ++bp) {
        const double k = sK[(3 * fcorner(bb) + c) * 24 + 3 * fcorner(bp) + d];
        sum += ((__builtin_popcount(M & bb) ^ __builtin_popcount(Mp & bp)) & 1) ? -k : k;
      }
    sG[tid] = fk_kept(M, c, d) ? sum / 64.0 : 0.0;
  }
  __syncthreads();
  const int rx = s_r[0], ry = DIM == 3 ? s_r[1] : 0, rz = s_r[2];
  const bool wide = rx > kMaxR || ry > RYM || rz > kMaxR;
  // boundary classes: index ((((lx * (rx+1) + hx) * (ry+1) + ly) * (ry+1) + hy) * (rz+1) + lz) * (rz+1) + hz;
  // each sum adds the in-grid taps in tap order, as the reference's loop does
  const int bx = rx + 1, by = ry + 1, bz = rz + 1;
  // cube form of the tap sum: taps strictly ascending in (dx, dy, dz), radius <= kMaxR
  const int rc = max(rx, max(ry, rz));
  if (!wide) {
    const int wc = 2 * rc + 1, wcy = DIM == 3 ? wc : 1, rcy = DIM == 3 ? rc : 0;
    for (int t = tid; t < n_off; t += NT) {
      const int d = s_d[t];
      const int dx = (d >> 10) - 8, dy = ((d >> 5) & 31) - 8, dz = (d & 31) - 8;
      s_wc[((dx + rc) * wcy + dy + rcy) * wc + dz + rc] = s_w[t];
    }
    if (tid == 0) {
      int ok = 1;
      for (int t = 1; t < n_off; ++t) ok &= s_d[t] > s_d[t - 1];
      s_cube = ok;
    }
    __syncthreads();
  }
  const bool cube = !wide && s_cube != 0;
  if (MODE == 1 && !wide) {
    const int nbc = bx * bx * by * by * bz * bz;
    for (int c = tid; c < nbc; c += NT) {
      int r = c;
      const int hz = r % bz; r /= bz;
      const int lz = r % bz; r /= bz;
      const int hy = r % by; r /= by;
      const int ly = r % by; r /= by;
      const int hx = r % bx, lx = r / bx;
      double ws = 0.0;
      for (int t = 0; t < n_off; ++t) {
        const int d = s_d[t];
        const int dx = (d >> 10) - 8, dy = ((d >> 5) & 31) - 8, dz = (d & 31) - 8;
        ws += (dx >= -lx && dx <= hx && dy >= -ly && dy <= hy && dz >= -lz && dz <= hz) ? s_w[t] : 0.0;
      }
      s_bw[c] = ws;
    }
    __syncthreads();
  }

  const int y0 = blockIdx.x * TY, z0 = blockIdx.y * TZ;
  const int x0 = blockIdx.z * cx, x1 = min(x0 + cx, sg.X);
  const int ty = tid / TZ, tz = tid % TZ;
  const int oy = y0 + ty, oz = z0 + tz;
  const bool mine = oy < sg.Y && oz < sg.Z;
  if (wide) {
    // taps beyond the ring: gathers from memory (the adjoint recomputes each tap's value)
    if (!mine) return;
#pragma unroll 1
    for (int xo = x0; xo < x1; ++xo) {
      double sv = 0.0, sw = 0.0;
#pragma unroll 1
      for (int t = 0; t < n_off; ++t) {
        const int dx = off[t * DIM];
        const int dy = DIM == 3 ? off[t * 3 + 1] : 0;
        const int dz = DIM == 3 ? off[t * 3 + 2] : off[t * 2 + 1];
        const int a = xo + dx, b = oy + dy, c = oz + dz;
        if (a < 0 || a >= sg.X || b < 0 || b >= sg.Y || c < 0 || c >= sg.Z) continue;
        if (MODE == 3) asm volatile("" ::: "memory");
        sv = fma(s_w[t], stencil_value<MODE, DIM, EN>(g, sg, a, b, c, x, adj, sK, sG), sv);
        sw += s_w[t];
      }
      const int64_t e = ((int64_t)xo * sg.Y + oy) * sg.Z + oz;
      y[e] = MODE == 1 ? (sw > 0.0 ? sv / sw : 0.0) : sv;
    }
    return;
  }
  const int py = TY + 2 * ry, pz = TZ + 2 * rz, pn = py * pz;
  // A plane is staged in two halves: fetch() forms this thread's values of plane xs in
  // registers (zero outside the grid), put() stores them to ring slot xs mod 8.  The loop
  // fetches plane xo + rx + 1 before it computes plane xo, so the loads are in flight during
  // the taps; the ring's spare slot (8 > 2 rx + 1) means the store needs no extra barrier.
  constexpr int KP = (PL + NT - 1) / NT;
  // MODES 1 and 2: plane xs is fetched into registers PD steps before it is stored to its
  // ring slot, so PD planes' loads are in flight while the taps run.  MODE 3 computes each
  // staged value (an element energy) and stores it at once, one at a time: holding a
  // plane's share across the taps would take most of the register file.  A ring slot is
  // written one step before it is read and is not in use then (8 slots, at most
  // 2 rx + 1 = 5 read per step).
  constexpr bool kDirect = MODE == 3;
  constexpr int PD = kDirect ? 1 : 4;
  double pre[PD][KP];
  // per staged slot q: its element's offset within a plane (-1: outside the grid in y or z)
  // and its LDS offset (-1: past the tile); both are the same for every plane
  int qe[KP], ql[KP];
#pragma unroll
  for (int q = 0; q < KP; ++q) {
    const int t = tid + NT * q;
    const int ly = t / pz, lz = t - ly * pz;
    const int gy = y0 - ry + ly, gz = z0 - rz + lz;
    const bool in = t < pn;
    ql[q] = in ? (ly + RYM - ry) * PZ + lz + kMaxR - rz : -1;
    qe[q] = in && gy >= 0 && gy < sg.Y && gz >= 0 && gz < sg.Z ? gy * sg.Z + gz : -1;
  }
  // the loads are unconditional (slots outside the grid read element 0 and are zeroed
  // afterwards), so the loop has no branches around them and a plane's loads stay in
  // flight across the steps until its store
  // MODE 2 also keeps N of each slot (prn), the quotient v / N is formed at the store
  double prn[MODE == 2 ? PD : 1][KP];
  auto fetch = [&](double (&r)[KP], double (&rn)[KP], int xs) {
    // planes outside the grid read the nearest one (every lane loads, no branch)
    const int64_t base = (int64_t)min(max(xs, 0), sg.X - 1) * sg.Y * sg.Z;
#pragma unroll
    for (int q = 0; q < KP; ++q) {
      const int64_t e = base + max(qe[q], 0);
      r[q] = x[e];
      if (MODE == 2) rn[q] = adj.norm[e];
    }
  };
  auto put = [&](const double (&r)[KP], const double (&rn)[KP], int xs) {
    double* pl = ring + (xs & (RING - 1)) * PL;
    const bool xin = xs >= 0 && xs < sg.X;
#pragma unroll
    for (int q = 0; q < KP; ++q) {
      const double v = MODE == 2 ? r[q] / rn[q] : r[q];
      if (ql[q] >= 0) pl[ql[q]] = xin && qe[q] >= 0 ? v : 0.0;
    }
  };
  auto fill = [&](int xs) {
    double* pl = ring + (xs & (RING - 1)) * PL;
#pragma unroll 1
    for (int q = 0; q < KP; ++q) {
      const int t = tid + NT * q;
      if (t < pn) {
        const int ly = t / pz, lz = t - (t / pz) * pz;
        const int gy = y0 - ry + ly, gz = z0 - rz + lz;
        double v = 0.0;
        asm volatile("" ::: "memory");  // K0 / G re-read from LDS per element
        if (xs >= 0 && xs < sg.X && gy >= 0 && gy < sg.Y && gz >= 0 && gz < sg.Z)
          v = stencil_value<MODE, DIM, EN>(g, sg, xs, gy, gz, x, adj, sK, sG);
        pl[(ly + RYM - ry) * PZ + lz + kMaxR - rz] = v;
      }
    }
  };
  // this output's boundary class in y and z (x is added per plane)
  const int byz = ((min(oy, ry) * by + min(sg.Y - 1 - oy, ry)) * bz + min(oz, rz)) * bz + min(sg.Z - 1 - oz, rz);
  const int me = (ty + RYM) * PZ + tz + kMaxR - (1 << 15);
  auto taps = [&](int xo) {
    if (!mine) return;
    double sv = 0.0;
    // the cube weights are re-read from LDS per plane (uniform broadcasts), not held in
    // registers across the march
    asm volatile("" ::: "memory");
    if (cube && rc == 1) {
      sv = cube_sum<1, DIM, PZ, PL, RING>(ring, s_wc, xo, me + (1 << 15));
    } else if (cube && rc == 2) {
      sv = cube_sum<2, DIM, PZ, PL, RING>(ring, s_wc, xo, me + (1 << 15));
    } else {
      for (int t = 0; t < n_off; ++t) {
        const int tp = s_tap[t];
        const int slot = (xo + (tp >> 16) - 8) & (RING - 1);
        sv = fma(s_w[t], ring[slot * PL + (tp & 0xffff) + me], sv);
      }
    }
    double out = sv;
    if (MODE == 1) {
      const double sw = s_bw[(min(xo, rx) * bx + min(sg.X - 1 - xo, rx)) * (by * by * bz * bz) + byz];
      out = sw > 0.0 ? sv / sw : 0.0;
    }
    y[((int64_t)xo * sg.Y + oy) * sg.Z + oz] = out;
  };
  if (kDirect) {
    for (int xs = x0 - rx; xs <= x0 + rx; ++xs) fill(xs);
    for (int xo = x0; xo < x1; ++xo) {
      __syncthreads();
      if (xo + 1 < x1) fill(xo + rx + 1);
      taps(xo);
    }
  } else {
    // planes x0 - rx .. x0 + rx - 1 go to the ring now; x0 + rx + k is in flight in pre[k]
    for (int xs = x0 - rx; xs < x0 + rx; ++xs) {
      fetch(pre[0], prn[0], xs);
      put(pre[0], prn[0], xs);
    }
#pragma unroll
    for (int k = 0; k < PD; ++k)
      if (x0 + k < x1) fetch(pre[k], prn[MODE == 2 ? k : 0], x0 + rx + k);
    // step xo stores plane xo + rx from pre[k] and refills pre[k] with plane xo + rx + PD.
    // Whole groups of PD steps run without conditions (planes past the grid are fetched
    // clamped and never stored), so every path through the loop issues the same loads and
    // each store waits only for its own plane's loads.
    int xb = x0;
    for (; xb + PD <= x1; xb += PD) {
#pragma unroll
      for (int k = 0; k < PD; ++k) {
        const int xo = xb + k;
        put(pre[k], prn[MODE == 2 ? k : 0], xo + rx);
        __syncthreads();
        fetch(pre[k], prn[MODE == 2 ? k : 0], xo + rx + PD);
        taps(xo);
      }
    }
#pragma unroll
    for (int k = 0; k < PD; ++k) {
      const int xo = xb + k;
      if (xo < x1) {
        put(pre[k], prn[MODE == 2 ? k : 0], xo + rx);
        __syncthreads();
        taps(xo);
      }
    }
  }
}

// ------------------------------------------------------------------- host ---

template <int MODE>
static int launch_gather(const Grid& g, int n_off, const int32_t* off, const double* w, const double* norm,
                         const double* x, double* y, hipStream_t s) {
  const unsigned grid = grid_for(g.n_elem());
  if (g.dim == 3)
    filter_kernel<3, MODE><<<grid, kBlock, 0, s>>>(g, n_off, off, w, norm, x, y);
  else
    filter_kernel<2, MODE><<<grid, kBlock, 0, s>>>(g, n_off, off, w, norm, x, y);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

// rows per 3-D tile, 4 or 8
static int g_stencil_ty = 8;
void set_stencil_rows(int ty) { g_stencil_ty = ty; }
int get_stencil_rows() { return g_stencil_ty; }

template <int MODE, int EN>
static int launch_stencil(const Grid& g, int n_off, const int32_t* off, const double* w, const double* x,
                          const AdjArgs& adj, double* y, hipStream_t s) {
  const SGrid sg = sgrid(g);
  const int ty = g.dim == 3 ? g_stencil_ty : 1, tz = g.dim == 3 ? 64 : 256;
  const int64_t tiles = (int64_t)((sg.Y + ty - 1) / ty) * ((sg.Z + tz - 1) / tz);
  // the longest x chunk that still gives about 512 workgroups (two per compute unit)
  int cx = kMaxCX;
  while (cx > 4 && tiles * ((sg.X + cx - 1) / cx) < 512) cx /= 2;
  const dim3 grid((sg.Y + ty - 1) / ty, (sg.Z + tz - 1) / tz, (sg.X + cx - 1) / cx);
  if (g.dim == 2)
    stencil_kernel<MODE, 2, EN, 1><<<grid, 256, 0, s>>>(g, sg, cx, n_off, off, w, x, adj, y);
  else if (ty == 4)
    stencil_kernel<MODE, 3, EN, 4><<<grid, 256, 0, s>>>(g, sg, cx, n_off, off, w, x, adj, y);
  else
    stencil_kernel<MODE, 3, EN, 8><<<grid, 512, 0, s>>>(g, sg, cx, n_off, off, w, x, adj, y);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int filter_normalizer(const Grid& g, int n_off, const int32_t* off, const double* w, double* norm,
                      hipStream_t s) {
  ET_REQUIRE(n_off > 0 && n_off <= kMaxOff, "filter offset count out of range (1..1024)");
  return launch_gather<0>(g, n_off, off, w, nullptr, nullptr, norm, s);
}

int filter_fwd(const Grid& g, int n_off, const int32_t* off, const double* w, const double* x,
               double* y, hipStream_t s) {
  ET_REQUIRE(n_off > 0 && n_off <= kMaxOff, "filter offset count out of range (1..1024)");
  if (n_off > kMaxTaps) return launch_gather<1>(g, n_off, off, w, nullptr, x, y, s);
  return launch_stencil<1, 0>(g, n_off, off, w, x, AdjArgs{}, y, s);
}

int filter_adj(const Grid& g, int n_off, const int32_t* off, const double* w, const double* norm,
               const double* v, double* y, hipStream_t s) {
  ET_REQUIRE(n_off > 0 && n_off <= kMaxOff, "filter offset count out of range (1..1024)");
  if (n_off > kMaxTaps) return launch_gather<2>(g, n_off, off, w, norm, v, y, s);
  AdjArgs a{};
  a.norm = norm;
  return launch_stencil<2, 0>(g, n_off, off, w, v, a, y, s);
}

int simp_adjoint(const Grid& g, const double* K0, const double* U, const double* xt, double penal, double vmin,
                 int n_off, const int32_t* off, const double* w, const double* norm, int k0_flags, double* y,
                 hipStream_t s) {
  ET_REQUIRE(n_off > 0 && n_off <= kMaxTaps, "simp adjoint: filter offset count out of range (1..512)");
  const AdjArgs a{K0, U, xt, penal, vmin, norm};
  const bool wht = g.dim == 3 && (k0_flags & K0S_WHT);
  if (penal == 3.0)
    return wht ? launch_stencil<3, 1>(g, n_off, off, w, nullptr, a, y, s)
               : launch_stencil<3, 2>(g, n_off, off, w, nullptr, a, y, s);
  return wht ? launch_stencil<3, 5>(g, n_off, off, w, nullptr, a, y, s)
             : launch_stencil<3, 6>(g, n_off, off, w, nullptr, a, y, s);
}

}  // namespace et
