// Linear-hat density filter on structured element grids: normaliser, forward
// (normalised weighted average over in-bounds neighbours), exact transpose, and
// the SIMP adjoint fused with the transpose.
//
// Forward and transpose are LDS-halo stencils.  A workgroup owns a tile of
// TY x TZ element columns (3-D 4 x 64, a wavefront per row along the fastest
// axis; 2-D 1 x 256) and marches along x through a chunk of kCX planes.  Each step stages one
// new element plane of the tile plus its y/z halo into a ring of 2 rx + 1 planes
// in LDS (zero outside the grid) and produces one output plane from the ring,
// so every input value is read from memory once per tile (plus the halo rows)
// and every output is written once, with no atomics.  Sums follow the offset
// table's order, so results are deterministic; the forward pass is bitwise the
// reference's (out-of-grid taps add an exact 0.0).
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
// with uh = H u_e and G = H K0 H / 64 (about 140 flops instead of 1152).  Each
// workgroup forms the kept G entries from K0 and checks that they rebuild K0
// exactly (to 1e-12 of its largest entry); if they do not (an element matrix
// without the brick's reflection structure) it evaluates the full product.
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
constexpr int kMaxR = 3;     // halo radius per axis held in the stencil kernels' ring
constexpr int kCX = 16;      // x planes per workgroup

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
                                               const double* __restrict__ K0, const double* sG) {
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
    for (int M = 1; M < 8; ++M)
#pragma unroll
      for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int d = 0; d < 3; ++d)
          if (fk_kept(M, c, d)) e = fma(uh[c][M] * sG[(M - 1) * 9 + c * 3 + d], uh[d][M ^ (1 << c) ^ (1 << d)], e);
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
                                      const double* __restrict__ K0) {
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

// value staged for element e at (xs, gy, gz) of the stencil grid:
// MODE 1 x_e, 2 v_e / N_e, 3 s_e / N_e (SIMP adjoint)
// (EN: element energy in the transformed basis, 1, or as the full product, 2)
template <int MODE, int DIM, int EN>
__device__ __forceinline__ double stencil_value(const Grid& g, const SGrid& sg, int xs, int gy, int gz,
                                                const double* __restrict__ x, const AdjArgs& adj,
                                                const double* sG) {
  const int64_t e = ((int64_t)xs * sg.Y + gy) * sg.Z + gz;
  if (MODE == 1) return x[e];
  if (MODE == 2) return x[e] / adj.norm[e];
  const double xe = fmin(fmax(adj.xt[e], adj.vmin), 1.0);
  const double en = DIM == 3 ? elem_energy3<EN == 1>(g, xs, gy, gz, adj.U, adj.K0, sG)
                             : elem_energy2(g, xs, gz, adj.U, adj.K0);
  return adj.penal * pow(xe, adj.penal - 1.0) * en / adj.norm[e];
}

// the same, out of line: the wide-stencil path calls it per tap
template <int MODE, int DIM, int EN>
__device__ __attribute__((noinline)) double stencil_value_call(const Grid& g, const SGrid& sg, int xs, int gy,
                                                                int gz, const double* __restrict__ x,
                                                                const AdjArgs& adj, const double* sG) {
  return stencil_value<MODE, DIM, EN>(g, sg, xs, gy, gz, x, adj, sG);
}

// MODE 1: forward, 2: transpose (values v / N), 3: SIMP adjoint (values s / N).
// Tile: TY x TZ outputs per plane (3-D 4 x 64, 2-D 1 x 256), one per thread.
// The 3-D adjoint is launched twice, EN = 1 (transformed-basis energy) and EN = 2
// (full product): each workgroup checks K0's structure and the launch that does
// not apply returns after that check, so each variant keeps only its own registers.
template <int MODE, int DIM, int EN>
__global__ __launch_bounds__(256) void stencil_kernel(const Grid g, const SGrid sg, int n_off,
                                                      const int32_t* __restrict__ off,
                                                      const double* __restrict__ w,
                                                      const double* __restrict__ x,
                                                      const AdjArgs adj,
                                                      double* __restrict__ y) {
  constexpr int TY = DIM == 3 ? 4 : 1, TZ = 256 / TY;
  constexpr int RYM = DIM == 3 ? kMaxR : 0;  // y halo capacity (2-D grids have one y row)
  constexpr int PY = TY + 2 * RYM, PZ = TZ + 2 * kMaxR, PL = PY * PZ;
  constexpr int RING = 2 * kMaxR + 1;
  __shared__ double ring[RING * PL];
  __shared__ int s_off[kMaxOff];  // packed tap (dx, dy, dz), each + 128
  __shared__ double s_w[kMaxOff];
  __shared__ double sG[63];
  __shared__ int s_r[3], s_wht;
  const int tid = threadIdx.x;
  if (tid < 3) s_r[tid] = 0;
  if (tid == 0) s_wht = 1;
  __syncthreads();
  for (int t = tid; t < n_off; t += 256) {
    const int dx = off[t * DIM];
    const int dy = DIM == 3 ? off[t * 3 + 1] : 0;
    const int dz = DIM == 3 ? off[t * 3 + 2] : off[t * 2 + 1];
    s_off[t] = ((dx + 128) << 16) | ((dy + 128) << 8) | (dz + 128);
    s_w[t] = w[t];
    atomicMax(&s_r[0], abs(dx));
    atomicMax(&s_r[1], abs(dy));
    atomicMax(&s_r[2], abs(dz));
  }
  if (MODE == 3 && DIM == 3) {
    // kept couplings of G = H K0 H / 64
    if (tid < 63) {
      const int M = 1 + tid / 9, c = (tid / 3) % 3, d = tid % 3;
      const int Mp = M ^ (1 << c) ^ (1 << d);
      double s = 0.0;
#pragma unroll 1
      for (int bb = 0; bb < 8; ++bb)
#pragma unroll 2
        for (int bp = 0; bp < 8; ++bp) {
          const double k = adj.K0[(3 * fcorner(bb) + c) * 24 + 3 * fcorner(bp) + d];
          s += ((__builtin_popcount(M & bb) ^ __builtin_popcount(Mp & bp)) & 1) ? -k : k;
        }
      sG[tid] = fk_kept(M, c, d) ? s / 64.0 : 0.0;
    }
    __syncthreads();
    // they must rebuild K0: K0[(b,c),(b',d)] = sum_M H[b][M] G[M][c][d] H[b'][M^e_c^e_d]
    double kmax = 0.0;
#pragma unroll 4
    for (int t = 0; t < 576; ++t) kmax = fmax(kmax, fabs(adj.K0[t]));
#pragma unroll 1
    for (int t = tid; t < 576; t += 256) {
      const int b = t / 72, c = (t / 24) % 3, b2 = (t / 3) % 8, d = t % 3;
      double s = 0.0;
      for (int M = 1; M < 8; ++M) {
        const int Mp = M ^ (1 << c) ^ (1 << d);
        const double h = ((__builtin_popcount(M & b) ^ __builtin_popcount(Mp & b2)) & 1) ? -1.0 : 1.0;
        s += h * sG[(M - 1) * 9 + c * 3 + d];
      }
      const double k = adj.K0[(3 * fcorner(b) + c) * 24 + 3 * fcorner(b2) + d];
      if (!(fabs(s - k) <= 1e-12 * kmax)) s_wht = 0;
    }
  }
  __syncthreads();
  if (MODE == 3 && DIM == 3 && (EN == 1) != (s_wht != 0)) return;
  const int rx = s_r[0], ry = DIM == 3 ? s_r[1] : 0, rz = s_r[2];
  const bool wide = rx > kMaxR || ry > RYM || rz > kMaxR;

  const int y0 = blockIdx.x * TY, z0 = blockIdx.y * TZ;
  const int x0 = blockIdx.z * kCX, x1 = min(x0 + kCX, sg.X);
  const int ty = tid / TZ, tz = tid % TZ;
  const int oy = y0 + ty, oz = z0 + tz;
  const bool mine = oy < sg.Y && oz < sg.Z;
  if (wide) {
    // taps beyond the ring: direct gathers (adjoint values recomputed per tap)
    if (!mine) return;
    for (int xo = x0; xo < x1; ++xo) {
      double sv = 0.0, sw = 0.0;
      for (int t = 0; t < n_off; ++t) {
        const int pk = s_off[t];
        const int a = xo + (pk >> 16) - 128, b = oy + ((pk >> 8) & 0xff) - 128, c = oz + (pk & 0xff) - 128;
        if (a < 0 || a >= sg.X || b < 0 || b >= sg.Y || c < 0 || c >= sg.Z) continue;
        sv = fma(s_w[t], stencil_value_call<MODE, DIM, EN>(g, sg, a, b, c, x, adj, sG), sv);
        sw += s_w[t];
      }
      const int64_t e = ((int64_t)xo * sg.Y + oy) * sg.Z + oz;
      y[e] = MODE == 1 ? (sw > 0.0 ? sv / sw : 0.0) : sv;
    }
    return;
  }
  const int py = TY + 2 * ry, pz = TZ + 2 * rz, pn = py * pz;
  // stage element plane xs (zero outside the grid) into ring slot xs mod RING
  auto stage = [&](int xs) {
    double* pl = ring + (((xs % RING) + RING) % RING) * PL;
#pragma unroll 1
    for (int t = tid; t < pn; t += 256) {
      const int ly = t / pz, lz = t % pz;
      const int gy = y0 - ry + ly, gz = z0 - rz + lz;
      double v = 0.0;
      // the G coefficients are re-read from LDS per element rather than held in
      // registers across the loop (63 doubles would take half the register file)
      if (MODE == 3) asm volatile("" ::: "memory");
      if (xs >= 0 && xs < sg.X && gy >= 0 && gy < sg.Y && gz >= 0 && gz < sg.Z)
        v = stencil_value<MODE, DIM, EN>(g, sg, xs, gy, gz, x, adj, sG);
      pl[(ly + RYM - ry) * PZ + lz + kMaxR - rz] = v;
    }
  };
  for (int xs = x0 - rx; xs < x0 + rx; ++xs) stage(xs);
  for (int xo = x0; xo < x1; ++xo) {
    stage(xo + rx);
    __syncthreads();
    if (mine) {
      double sv = 0.0, sw = 0.0;
      for (int t = 0; t < n_off; ++t) {
        const int pk = s_off[t];
        const int dx = (pk >> 16) - 128, dy = ((pk >> 8) & 0xff) - 128, dz = (pk & 0xff) - 128;
        const int slot = (((xo + dx) % RING) + RING) % RING;
        sv = fma(s_w[t], ring[slot * PL + (ty + RYM + dy) * PZ + tz + kMaxR + dz], sv);
        if (MODE == 1) {
          const int a = xo + dx, b = oy + dy, c = oz + dz;
          sw += (a >= 0 && a < sg.X && b >= 0 && b < sg.Y && c >= 0 && c < sg.Z) ? s_w[t] : 0.0;
        }
      }
      const int64_t e = ((int64_t)xo * sg.Y + oy) * sg.Z + oz;
      y[e] = MODE == 1 ? (sw > 0.0 ? sv / sw : 0.0) : sv;
    }
    __syncthreads();  // every read of this step precedes the next restaging
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

template <int MODE, int EN>
static int launch_stencil(const Grid& g, int n_off, const int32_t* off, const double* w, const double* x,
                          const AdjArgs& adj, double* y, hipStream_t s) {
  const SGrid sg = sgrid(g);
  const int ty = g.dim == 3 ? 4 : 1, tz = 256 / ty;
  const dim3 grid((sg.Y + ty - 1) / ty, (sg.Z + tz - 1) / tz, (sg.X + kCX - 1) / kCX);
  if (g.dim == 3)
    stencil_kernel<MODE, 3, EN><<<grid, 256, 0, s>>>(g, sg, n_off, off, w, x, adj, y);
  else
    stencil_kernel<MODE, 2, EN><<<grid, 256, 0, s>>>(g, sg, n_off, off, w, x, adj, y);
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
  return launch_stencil<1, 0>(g, n_off, off, w, x, AdjArgs{}, y, s);
}

int filter_adj(const Grid& g, int n_off, const int32_t* off, const double* w, const double* norm,
               const double* v, double* y, hipStream_t s) {
  ET_REQUIRE(n_off > 0 && n_off <= kMaxOff, "filter offset count out of range (1..1024)");
  AdjArgs a{};
  a.norm = norm;
  return launch_stencil<2, 0>(g, n_off, off, w, v, a, y, s);
}

int simp_adjoint(const Grid& g, const double* K0, const double* U, const double* xt, double penal, double vmin,
                 int n_off, const int32_t* off, const double* w, const double* norm, double* y, hipStream_t s) {
  ET_REQUIRE(n_off > 0 && n_off <= kMaxOff, "filter offset count out of range (1..1024)");
  const AdjArgs a{K0, U, xt, penal, vmin, norm};
  if (g.dim == 3) ET_TRY((launch_stencil<3, 1>(g, n_off, off, w, nullptr, a, y, s)));
  return launch_stencil<3, 2>(g, n_off, off, w, nullptr, a, y, s);
}

}  // namespace et
