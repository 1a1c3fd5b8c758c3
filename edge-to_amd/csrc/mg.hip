// Geometric multigrid building blocks on structured grids:
//   * restriction (normalised full weighting) and prolongation,
//   * Galerkin coarse operators S_c = R S_f R^T on dense block stencils,
//   * block-stencil application.
// Stencil layout: plane ((o*dof + r)*dof + c) holds the (r, c) entry of the
// block coupling node n to neighbour n + offset(o), one value per node, so a
// wavefront reads each plane with unit stride.
#include "common.hpp"
#include "stiffness.hpp"

namespace et {

__device__ __forceinline__ double pw(int d) {  // 2^-|d| for d in {-1, 0, 1}
  return d == 0 ? 1.0 : 0.5;
}

template <int DIM>
__device__ __forceinline__ void node_coords(const Grid& g, int64_t n, int& i, int& j, int& k) {
  const int nz = g.nz(), nx = g.nx();
  k = DIM == 3 ? (int)(n % nz) : 0;
  const int64_t t = DIM == 3 ? n / nz : n;
  i = (int)(t % nx);
  j = (int)(t / nx);
}

template <int DIM>
__device__ __forceinline__ int64_t node_index(const Grid& g, int i, int j, int k) {
  return DIM == 3 ? ((int64_t)j * g.nx() + i) * g.nz() + k : (int64_t)j * g.nx() + i;
}

// total weight tw(c): product over axes of 1.5 on the boundary, 2 inside
template <int DIM>
__device__ __forceinline__ double tw_of(const Grid& gc, int i, int j, int k) {
  double t = (i == 0 || i == gc.nx() - 1) ? 1.5 : 2.0;
  t *= (j == 0 || j == gc.ny() - 1) ? 1.5 : 2.0;
  if (DIM == 3) t *= (k == 0 || k == gc.nz() - 1) ? 1.5 : 2.0;
  return t;
}

// 1/tw(c) as a product of per-axis factors (2/3 on the boundary, 1/2 inside),
// in the working precision: no division in the transfer kernels
template <typename T, int DIM>
__device__ __forceinline__ T inv_tw_of(const Grid& gc, int i, int j, int k) {
  const T bnd = T(2.0 / 3.0), mid = T(0.5);
  T t = (i == 0 || i == gc.nx() - 1) ? bnd : mid;
  t *= (j == 0 || j == gc.ny() - 1) ? bnd : mid;
  if (DIM == 3) t *= (k == 0 || k == gc.nz() - 1) ? bnd : mid;
  return t;
}

// ------------------------------------------------------------ restriction ---
// One thread per coarse node, all DOF components at once; 32-bit node
// arithmetic (node counts are below 2^31 by construction of the Grid).

template <int DIM>
__device__ __forceinline__ void node_coords32(int n, int nx, int nz, int& i, int& j, int& k) {
  if (DIM == 3) {
    k = n % nz;
    n /= nz;
  } else {
    k = 0;
  }
  i = n % nx;
  j = n / nx;
}

// Branch-free: out-of-range neighbours read a clamped (in-range) node with
// weight 0, so all loads of a thread are in flight together.
// With x0 given, also x0 = w D^-1 vc, w = omega * wscale (the coarse level's first
// Jacobi sweep from the zero guess, rounded as scale_kernel rounds it).
template <typename T, int DIM, int DOF>
__global__ __launch_bounds__(256) void restrict_kernel(const Grid gf, const Grid gc,
                                                       const T* __restrict__ vf,
                                                       T* __restrict__ vc,
                                                       const T* __restrict__ Dinv,
                                                       const double* __restrict__ omega, double wscale,
                                                       T* __restrict__ x0) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int)gc.n_nodes()) return;
  const int fnx = gf.nx(), fny = gf.ny(), fnz = gf.nz();
  int ci, cj, ck;
  node_coords32<DIM>(c, gc.nx(), gc.nz(), ci, cj, ck);
  T acc[DOF];
#pragma unroll
  for (int d = 0; d < DOF; ++d) acc[d] = T(0);
#pragma unroll
  for (int dj = -1; dj <= 1; ++dj) {
    const int fj0 = 2 * cj + dj;
    const bool jin = fj0 >= 0 && fj0 < fny;
    const int fj = min(max(fj0, 0), fny - 1);
#pragma unroll
    for (int di = -1; di <= 1; ++di) {
      const int fi0 = 2 * ci + di;
      const bool iin = fi0 >= 0 && fi0 < fnx;
      const int fi = min(max(fi0, 0), fnx - 1);
#pragma unroll
      for (int dk = (DIM == 3 ? -1 : 0); dk <= (DIM == 3 ? 1 : 0); ++dk) {
        const int fk0 = 2 * ck + dk;
        const bool kin = DIM != 3 || (fk0 >= 0 && fk0 < fnz);
        const int fk = DIM == 3 ? min(max(fk0, 0), fnz - 1) : 0;
        const T wgt = (jin && iin && kin) ? (T)(pw(di) * pw(dj) * pw(dk)) : T(0);
        const int f = DIM == 3 ? (fj * fnx + fi) * fnz + fk : fj * fnx + fi;
#pragma unroll
        for (int d = 0; d < DOF; ++d) acc[d] += wgt * vf[f * DOF + d];
      }
    }
  }
  const T inv = inv_tw_of<T, DIM>(gc, ci, cj, ck);
#pragma unroll
  for (int d = 0; d < DOF; ++d) vc[c * DOF + d] = acc[d] * inv;
  if (x0 != nullptr) {
    const T w = (T)(*omega * wscale);
#pragma unroll
    for (int d = 0; d < DOF; ++d) x0[c * DOF + d] = w * Dinv[c * DOF + d] * (acc[d] * inv);
  }
}

template <typename T>
int restrict_op(const Grid& gf, int dof, const T* vf, T* vc, hipStream_t s, const T* Dinv,
                const double* omega, double wscale, T* x0) {
  ET_REQUIRE(dof >= 1 && dof <= 3, "restriction: dof must be 1, 2 or 3");
  ET_REQUIRE(x0 == nullptr || (Dinv != nullptr && omega != nullptr), "restriction: x0 needs Dinv and omega");
  const Grid gc = gf.coarse();
  const unsigned grid = grid_for(gc.n_nodes());
#define ET_R(DIM, DOF) \
  restrict_kernel<T, DIM, DOF><<<grid, kBlock, 0, s>>>(gf, gc, vf, vc, Dinv, omega, wscale, x0)
  if (gf.dim == 3) {
    if (dof == 3) ET_R(3, 3); else if (dof == 2) ET_R(3, 2); else ET_R(3, 1);
  } else {
    if (dof == 3) ET_R(2, 3); else if (dof == 2) ET_R(2, 2); else ET_R(2, 1);
  }
#undef ET_R
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int restrict_op<double>(const Grid&, int, const double*, double*, hipStream_t, const double*,
                                  const double*, double, double*);
template int restrict_op<float>(const Grid&, int, const float*, float*, hipStream_t, const float*,
                                 const double*, double, float*);

// ----------------------------------------------------------- prolongation ---
// One thread per fine node: average of the (1, 2, 4 or 8) in-range coarse
// nodes of its cell, each divided by its total weight tw.

// Prolongation P = P_lin D_tw^-1, one thread per coarse node (ci, cj, ck): it
// writes the fine nodes (2ci + a, 2cj + b, 2ck + c), a, b, c in {0, 1}, whose
// parents all lie in {ci, ci+1} x {cj, cj+1} x {ck, ck+1}, so the 8 scaled
// parent values are loaded once per thread (not once per fine node).  A fine
// value is the sum of its parents' scaled values, in j, i, k order, times
// 1 / (number of parents); parents past the upper boundary are read clamped
// and not used.
template <typename T, int DIM, int DOF>
__global__ __launch_bounds__(256) void prolong_kernel(const Grid gf, const Grid gc,
                                                      const T* __restrict__ vc,
                                                      T* __restrict__ vf) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (int)gc.n_nodes()) return;
  const int cnx = gc.nx(), cny = gc.ny(), cnz = gc.nz();
  const int fnx = gf.nx(), fny = gf.ny(), fnz = gf.nz();
  int ci, cj, ck;
  node_coords32<DIM>(c, cnx, cnz, ci, cj, ck);
  constexpr int KZ = DIM == 3 ? 2 : 1;
  T sv[2][2][KZ][DOF];
#pragma unroll
  for (int dj = 0; dj < 2; ++dj)
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
      for (int dk = 0; dk < KZ; ++dk) {
        const int i = min(ci + di, cnx - 1), j = min(cj + dj, cny - 1), k = DIM == 3 ? min(ck + dk, cnz - 1) : 0;
        const T inv = inv_tw_of<T, DIM>(gc, i, j, k);
        const int n = DIM == 3 ? (j * cnx + i) * cnz + k : j * cnx + i;
#pragma unroll
        for (int d = 0; d < DOF; ++d) sv[dj][di][dk][d] = vc[n * DOF + d] * inv;
      }
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int fj = 2 * cj + b;
    const bool two_j = b == 1 && cj + 1 < cny;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int fi = 2 * ci + a;
      const bool two_i = a == 1 && ci + 1 < cnx;
#pragma unroll
      for (int cc = 0; cc < KZ; ++cc) {
        const int fk = 2 * ck + cc;
        const bool two_k = cc == 1 && ck + 1 < cnz;
        if (fj >= fny || fi >= fnx || (DIM == 3 && fk >= fnz)) continue;
        T acc[DOF];
#pragma unroll
        for (int d = 0; d < DOF; ++d) acc[d] = T(0);
#pragma unroll
        for (int pj = 0; pj < 2; ++pj)
#pragma unroll
          for (int pi = 0; pi < 2; ++pi)
#pragma unroll
            for (int pk = 0; pk < KZ; ++pk) {
              if ((pj == 1 && !two_j) || (pi == 1 && !two_i) || (pk == 1 && !two_k)) continue;
#pragma unroll
              for (int d = 0; d < DOF; ++d) acc[d] += sv[pj][pi][pk][d];
            }
        const int count = (two_i ? 2 : 1) * (two_j ? 2 : 1) * (two_k ? 2 : 1);
        const T scl = T(1) / (T)count;
        const int f = DIM == 3 ? (fj * fnx + fi) * fnz + fk : fj * fnx + fi;
#pragma unroll
        for (int d = 0; d < DOF; ++d) vf[f * DOF + d] = acc[d] * scl;
      }
    }
  }
}

template <typename T>
int prolong_op(const Grid& gf, int dof, const T* vc, T* vf, hipStream_t s) {
  ET_REQUIRE(dof >= 1 && dof <= 3, "prolongation: dof must be 1, 2 or 3");
  const Grid gc = gf.coarse();
  const unsigned grid = grid_for(gc.n_nodes());
#define ET_P(DIM, DOF) prolong_kernel<T, DIM, DOF><<<grid, kBlock, 0, s>>>(gf, gc, vc, vf)
  if (gf.dim == 3) {
    if (dof == 3) ET_P(3, 3); else if (dof == 2) ET_P(3, 2); else ET_P(3, 1);
  } else {
    if (dof == 3) ET_P(2, 3); else if (dof == 2) ET_P(2, 2); else ET_P(2, 1);
  }
#undef ET_P
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int prolong_op<double>(const Grid&, int, const double*, double*, hipStream_t);
template int prolong_op<float>(const Grid&, int, const float*, float*, hipStream_t);

// --------------------------------------------------------------- Galerkin ---

// One thread per (coarse node c, neighbour offset o): the dof x dof block of
// R S_f R^T coupling c and t = c + o.
template <int DIM, int DOF>
__global__ __launch_bounds__(256) void galerkin_kernel(const Grid gf, const Grid gc,
                                                       const double* __restrict__ Sf,
                                                       double* __restrict__ Sc) {
  constexpr int NO = DIM == 3 ? 27 : 9;
  const int64_t Nc = gc.n_nodes(), Nf = gf.n_nodes();
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= Nc * NO) return;
  const int o = (int)(tid / Nc);
  const int64_t c = tid % Nc;
  const int oi = DIM == 3 ? o / 9 - 1 : o / 3 - 1;
  const int oj = DIM == 3 ? (o / 3) % 3 - 1 : o % 3 - 1;
  const int ok = DIM == 3 ? o % 3 - 1 : 0;
  int ci, cj, ck;
  node_coords<DIM>(gc, c, ci, cj, ck);
  const int ti = ci + oi, tj = cj + oj, tk = ck + ok;
  double blk[DOF][DOF];
#pragma unroll
  for (int r = 0; r < DOF; ++r)
#pragma unroll
    for (int q = 0; q < DOF; ++q) blk[r][q] = 0.0;
  const bool tin = ti >= 0 && ti < gc.nx() && tj >= 0 && tj < gc.ny() &&
                   (DIM == 2 || (tk >= 0 && tk < gc.nz()));
  if (tin) {
    for (int di = -1; di <= 1; ++di) {
      const int fi = 2 * ci + di;
      if (fi < 0 || fi >= gf.nx()) continue;
      for (int dj = -1; dj <= 1; ++dj) {
        const int fj = 2 * cj + dj;
        if (fj < 0 || fj >= gf.ny()) continue;
        for (int dk = (DIM == 3 ? -1 : 0); dk <= (DIM == 3 ? 1 : 0); ++dk) {
          const int fk = 2 * ck + dk;
          if (DIM == 3 && (fk < 0 || fk >= gf.nz())) continue;
          const double wf = pw(di) * pw(dj) * pw(dk);
          const int64_t f = node_index<DIM>(gf, fi, fj, fk);
          // fine neighbours g = 2t + dg that are stencil neighbours of f
          for (int gi = -1; gi <= 1; ++gi) {
            const int si = 2 * ti + gi - fi;
            if (si < -1 || si > 1 || 2 * ti + gi < 0 || 2 * ti + gi >= gf.nx()) continue;
            for (int gj = -1; gj <= 1; ++gj) {
              const int sj = 2 * tj + gj - fj;
              if (sj < -1 || sj > 1 || 2 * tj + gj < 0 || 2 * tj + gj >= gf.ny()) continue;
              for (int gk = (DIM == 3 ? -1 : 0); gk <= (DIM == 3 ? 1 : 0); ++gk) {
                const int sk = DIM == 3 ? 2 * tk + gk - fk : 0;
                if (DIM == 3 && (sk < -1 || sk > 1 || 2 * tk + gk < 0 || 2 * tk + gk >= gf.nz()))
                  continue;
                const double wgt = wf * pw(gi) * pw(gj) * pw(gk);
                const int so = DIM == 3 ? (si + 1) * 9 + (sj + 1) * 3 + (sk + 1)
                                        : (si + 1) * 3 + (sj + 1);
#pragma unroll
                for (int r = 0; r < DOF; ++r)
#pragma unroll
                  for (int q = 0; q < DOF; ++q)
                    blk[r][q] += wgt * Sf[((int64_t)(so * DOF + r) * DOF + q) * Nf + f];
              }
            }
          }
        }
      }
    }
    const double inv = 1.0 / (tw_of<DIM>(gc, ci, cj, ck) * tw_of<DIM>(gc, ti, tj, tk));
#pragma unroll
    for (int r = 0; r < DOF; ++r)
#pragma unroll
      for (int q = 0; q < DOF; ++q) blk[r][q] *= inv;
  }
#pragma unroll
  for (int r = 0; r < DOF; ++r)
#pragma unroll
    for (int q = 0; q < DOF; ++q) Sc[((int64_t)(o * DOF + r) * DOF + q) * Nc + c] = blk[r][q];
}

// 3-D, dof 3: the same product with the fine pairs (f, g) enumerated per axis.
// For coarse offset o_a on an axis, f = 2c + d and g = 2t + e (d, e in -1..1)
// couple through the fine stencil only when s = 2 o_a + e - d is in -1..1:
// 7 (d, e) pairs for o_a = 0, 3 for o_a = +-1.  Workgroups share o (grid y and
// the template), so the pair loops are uniform; the z pairs are unrolled and
// out-of-grid nodes read a clamped node with weight zero, so each (x, y) pair
// issues all of its loads at once.
__device__ __forceinline__ int gal_npairs(int oa) { return oa == 0 ? 7 : 3; }
__device__ __forceinline__ void gal_pair(int oa, int a, int& d, int& e) {
  // oa = -1: (-1,0) (-1,1) (0,1); oa = 0: |e - d| <= 1; oa = 1: (0,-1) (1,-1) (1,0)
  if (oa < 0) {
    d = a == 2 ? 0 : -1;
    e = a == 0 ? 0 : 1;
  } else if (oa > 0) {
    d = a == 0 ? 0 : 1;
    e = a == 2 ? 0 : -1;
  } else {
    constexpr int dd[7] = {-1, -1, 0, 0, 0, 1, 1}, ee[7] = {-1, 0, -1, 0, 1, 0, 1};
    d = dd[a];
    e = ee[a];
  }
}

template <int OK>
// two waves per SIMD: lets the scheduler keep all loads of a pair group in flight
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void galerkin3_kernel(const Grid gf, const Grid gc, const double* __restrict__ Sf,
                                                        double* __restrict__ Sc) {
  const int64_t Nc = gc.n_nodes(), Nf = gf.n_nodes();
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= Nc) return;
  const int oi = (int)blockIdx.y / 3 - 1, oj = (int)blockIdx.y % 3 - 1;
  const int o = (oi + 1) * 9 + (oj + 1) * 3 + (OK + 1);
  int ci, cj, ck;
  node_coords<3>(gc, c, ci, cj, ck);
  const int ti = ci + oi, tj = cj + oj, tk = ck + OK;
  const int fnx = gf.nx(), fny = gf.ny(), fnz = gf.nz();
  double blk[3][3] = {};
  const bool tin = ti >= 0 && ti < gc.nx() && tj >= 0 && tj < gc.ny() && tk >= 0 && tk < gc.nz();
  if (tin) {
    const int nxp = gal_npairs(oi), nyp = gal_npairs(oj);
    constexpr int nzp = OK == 0 ? 7 : 3;
    for (int a = 0; a < nxp; ++a) {
      int di, ei;
      gal_pair(oi, a, di, ei);
      const int fi = 2 * ci + di, gi = 2 * ti + ei;
      const bool xin = fi >= 0 && fi < fnx && gi >= 0 && gi < fnx;
      const int si = 2 * oi + ei - di;
      for (int b2 = 0; b2 < nyp; ++b2) {
        int dj, ej;
        gal_pair(oj, b2, dj, ej);
        const int fj = 2 * cj + dj, gj = 2 * tj + ej;
        const bool yin = fj >= 0 && fj < fny && gj >= 0 && gj < fny;
        const int sj = 2 * oj + ej - dj;
        const double wxy = (xin && yin) ? pw(di) * pw(ei) * pw(dj) * pw(ej) : 0.0;
        const int fic = min(max(fi, 0), fnx - 1), fjc = min(max(fj, 0), fny - 1);
#pragma unroll
        for (int q3 = 0; q3 < nzp; ++q3) {
          int dk, ek;
          gal_pair(OK, q3, dk, ek);
          const int fk = 2 * ck + dk, gk = 2 * tk + ek;
          const bool zin = fk >= 0 && fk < fnz && gk >= 0 && gk < fnz;
          const double w = zin ? wxy * pw(dk) * pw(ek) : 0.0;
          const int sk = 2 * OK + ek - dk;
          const int64_t f = ((int64_t)fjc * fnx + fic) * fnz + min(max(fk, 0), fnz - 1);
          const int so = (si + 1) * 9 + (sj + 1) * 3 + (sk + 1);
#pragma unroll
          for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int q = 0; q < 3; ++q) blk[r][q] += w * Sf[((int64_t)(so * 3 + r) * 3 + q) * Nf + f];
        }
      }
    }
    const double inv = 1.0 / (tw_of<3>(gc, ci, cj, ck) * tw_of<3>(gc, ti, tj, tk));
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) blk[r][q] *= inv;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) Sc[((int64_t)(o * 3 + r) * 3 + q) * Nc + c] = blk[r][q];
}

// One axis of the 3-D Galerkin product in separable form.  The input is a
// block stencil (27 offsets of 3 x 3 blocks, [(o * 3 + r) * 3 + q] * N + node)
// on a node grid that is fine along axis AX (and coarse along the axes already
// reduced); the output has axis AX coarse.  Output node c, offset o: the sum
// over the (d, e) pairs of o's component along AX (gal_pair) of
// pw(d) pw(e) in[c with c_AX -> 2 c_AX + d][o with o_AX -> 2 o_AX + e - d]
// over pairs whose fine nodes 2 c_AX + d and 2 (c_AX + o_AX) + e lie in the
// grid, times 1 / (tw(c_AX) tw(c_AX + o_AX)) (tw is a product of per-axis
// factors); zero where c_AX + o_AX is outside the coarse grid.  The passes z,
// y, x give galerkin3_kernel's sums with each fine stencil entry read once from
// memory instead of once per coarse pair that needs it.  Thread per (output
// node, offsets on the other two axes): 9 input blocks, 27 sums.
__device__ __forceinline__ double tw_axis(int c, int cn) { return (c == 0 || c == cn - 1) ? 1.5 : 2.0; }

template <int AX>
__global__ __launch_bounds__(256) void galerkin_axis_kernel(int inx, int iny, int inz, int onx, int ony, int onz,
                                                            const double* __restrict__ in,
                                                            double* __restrict__ out) {
  const int64_t No = (int64_t)onx * ony * onz, Ni = (int64_t)inx * iny * inz;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= No) return;
  // grid y: the offsets on the two other axes; this thread forms all three
  // offsets o_AX = -1, 0, 1, so each input block it needs is loaded once
  const int ob = (int)blockIdx.y;
  int oo[3];
  {
    const int u0 = ob / 3 - 1, u1 = ob % 3 - 1;  // the other axes in x, y, z order
    int k = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) oo[a] = a == AX ? 0 : (k++ == 0 ? u0 : u1);
  }
  const int64_t u = t / onz;
  const int cc[3] = {(int)(u % onx), (int)(u / onx), (int)(t % onz)};
  const int ca = cc[AX];
  const int cn = AX == 0 ? onx : (AX == 1 ? ony : onz);
  const int fn = AX == 0 ? inx : (AX == 1 ? iny : inz);
  double acc[3][9];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[a][q] = 0.0;
  // fine positions f = 2 c + d and fine offsets s along AX; output offset oa
  // takes the pair (d, e = s - 2 oa + d) when e is in -1..1.  Loop order d, s
  // sums each output's pairs in gal_pair's order.  One fine position at a time
  // (its three blocks, 27 loads in flight) keeps the registers for 4 waves per SIMD.
#pragma unroll 1
  for (int d = -1; d <= 1; ++d) {
    const int f = 2 * ca + d;
    const bool fin = f >= 0 && f < fn;
    int nn[3] = {cc[0], cc[1], cc[2]};
    nn[AX] = min(max(f, 0), fn - 1);
    const int64_t ni = ((int64_t)nn[1] * inx + nn[0]) * inz + nn[2];
#pragma unroll
    for (int sa = -1; sa <= 1; ++sa) {
      int ss[3] = {oo[0], oo[1], oo[2]};
      ss[AX] = sa;
      const int so = (ss[0] + 1) * 9 + (ss[1] + 1) * 3 + (ss[2] + 1);
      double v[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) v[q] = in[(int64_t)(so * 9 + q) * Ni + ni];
#pragma unroll
      for (int oa = -1; oa <= 1; ++oa) {
        const int e = sa - 2 * oa + d;
        if (e < -1 || e > 1) continue;  // uniform: d, sa, oa are the same in every lane
        const int g = 2 * (ca + oa) + e;
        const double w = (fin && g >= 0 && g < fn) ? pw(d) * pw(e) : 0.0;
#pragma unroll
        for (int q = 0; q < 9; ++q) acc[oa + 1][q] += w * v[q];
      }
    }
  }
#pragma unroll
  for (int oa = -1; oa <= 1; ++oa) {
    const bool tin = ca + oa >= 0 && ca + oa < cn;
    const double inv = tin ? 1.0 / (tw_axis(ca, cn) * tw_axis(ca + oa, cn)) : 0.0;
    int os[3] = {oo[0], oo[1], oo[2]};
    os[AX] = oa;
    const int o = (os[0] + 1) * 9 + (os[1] + 1) * 3 + (os[2] + 1);
#pragma unroll
    for (int q = 0; q < 9; ++q) out[(int64_t)(o * 9 + q) * No + t] = tin ? acc[oa + 1][q] * inv : 0.0;
  }
}

int64_t galerkin_work_doubles(const Grid& gf) {
  if (gf.dim != 3) return 0;
  const Grid gc = gf.coarse();
  const int64_t fx = gf.nx(), fy = gf.ny(), cy = gc.ny(), cz = gc.nz();
  return 27 * 9 * (fx * fy * cz + fx * cy * cz);
}

int galerkin(const Grid& gf, int dof, const double* Sf, double* Sc, hipStream_t s, double* work) {
  const Grid gc = gf.coarse();
  const int NO = gf.dim == 3 ? 27 : 9;
  const unsigned grid = grid_for(gc.n_nodes() * NO);
  if (gf.dim == 3 && work != nullptr) {
    ET_REQUIRE(dof == 3, "3-D Galerkin expects dof 3");
    const int fx = gf.nx(), fy = gf.ny(), fz = gf.nz(), cx = gc.nx(), cy = gc.ny(), cz = gc.nz();
    double* T1 = work;                                  // (fx, fy, cz)
    double* T2 = work + (int64_t)27 * 9 * fx * fy * cz;  // (fx, cy, cz)
    galerkin_axis_kernel<2><<<dim3(grid_for((int64_t)fx * fy * cz), 9), kBlock, 0, s>>>(fx, fy, fz, fx, fy, cz,
                                                                                        Sf, T1);
    galerkin_axis_kernel<1><<<dim3(grid_for((int64_t)fx * cy * cz), 9), kBlock, 0, s>>>(fx, fy, cz, fx, cy, cz,
                                                                                        T1, T2);
    galerkin_axis_kernel<0><<<dim3(grid_for((int64_t)cx * cy * cz), 9), kBlock, 0, s>>>(fx, cy, cz, cx, cy, cz,
                                                                                        T2, Sc);
    (void)grid;
  } else if (gf.dim == 3) {
    ET_REQUIRE(dof == 3, "3-D Galerkin expects dof 3");
    const dim3 g3(grid_for(gc.n_nodes()), 9);
    galerkin3_kernel<-1><<<g3, kBlock, 0, s>>>(gf, gc, Sf, Sc);
    galerkin3_kernel<0><<<g3, kBlock, 0, s>>>(gf, gc, Sf, Sc);
    galerkin3_kernel<1><<<g3, kBlock, 0, s>>>(gf, gc, Sf, Sc);
    (void)grid;
  } else {
    ET_REQUIRE(dof == 2, "2-D Galerkin expects dof 2");
    galerkin_kernel<2, 2><<<grid, kBlock, 0, s>>>(gf, gc, Sf, Sc);
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

// ------------------------------------------------------- stencil apply ---

template <typename T, int DIM, int DOF>
__global__ __launch_bounds__(256) void stencil_apply_kernel(const Grid g, const T* __restrict__ S,
                                                            const T* __restrict__ x,
                                                            T* __restrict__ y) {
  constexpr int NO = DIM == 3 ? 27 : 9;
  const int64_t N = g.n_nodes();
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  int i, j, k;
  node_coords<DIM>(g, n, i, j, k);
  T acc[DOF];
#pragma unroll
  for (int r = 0; r < DOF; ++r) acc[r] = 0;
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    const int oi = DIM == 3 ? o / 9 - 1 : o / 3 - 1;
    const int oj = DIM == 3 ? (o / 3) % 3 - 1 : o % 3 - 1;
    const int ok = DIM == 3 ? o % 3 - 1 : 0;
    const int mi = i + oi, mj = j + oj, mk = k + ok;
    if (mi < 0 || mi >= g.nx() || mj < 0 || mj >= g.ny()) continue;
    if (DIM == 3 && (mk < 0 || mk >= g.nz())) continue;
    const int64_t m = node_index<DIM>(g, mi, mj, mk);
#pragma unroll
    for (int q = 0; q < DOF; ++q) {
      const T xv = x[m * DOF + q];
#pragma unroll
      for (int r = 0; r < DOF; ++r) acc[r] += S[((int64_t)(o * DOF + r) * DOF + q) * N + n] * xv;
    }
  }
#pragma unroll
  for (int r = 0; r < DOF; ++r) y[n * DOF + r] = acc[r];
}

template <typename T>
int stencil_apply(const Grid& g, int dof, const T* S, const T* x, T* y, hipStream_t s) {
  const unsigned grid = grid_for(g.n_nodes());
  if (g.dim == 3)
    stencil_apply_kernel<T, 3, 3><<<grid, kBlock, 0, s>>>(g, S, x, y);
  else
    stencil_apply_kernel<T, 2, 2><<<grid, kBlock, 0, s>>>(g, S, x, y);
  ET_CHECK_LAUNCH();
  return ET_OK;
}
template int stencil_apply<double>(const Grid&, int, const double*, const double*, double*,
                                   hipStream_t);
template int stencil_apply<float>(const Grid&, int, const float*, const float*, float*,
                                  hipStream_t);

}  // namespace et
