// First coarse Galerkin operator A_1 = R A_hat R^T built directly from the
// element densities, without forming a fine-grid stencil.
//
// R = D^-1 P^T, where P is the (unnormalised) multilinear interpolation and D
// holds the boundary-dependent total weights tw(c).  On grids with even
// element counts every fine element e lies inside exactly one coarse element
// C and is one of its 2^dim children t; P restricted to e's nodes only
// involves C's corners, through a fixed matrix W_t.  Hence
//   P^T A_hat P = sum_C Q_C^T ( sum_t E_{C,t} W_t^T K~_{C,t} W_t ) Q_C + P^T I_c P,
// with K~ = K0 masked on constrained DOFs.  For the (common) children without
// constrained DOFs W_t^T K0 W_t is a constant table M_t (built on the host);
// children touching constrained DOFs are contracted on the fly.
//
// One thread per (coarse node c, neighbour offset o) forms the dof x dof
// block coupling c and c+o, summing over the coarse elements that contain
// both, then scales it by 1 / (tw(c) tw(c+o)).  Output layout is the block
// stencil used by the solver (plane (o*dof + r)*dof + s, one value per node).
#include <vector>

#include "common.hpp"
#include "galerkin_elem.hpp"

namespace et {

// 1-D interpolation weight of a fine node at half-step s in {0,1,2} of a
// coarse element, for coarse corner bit ab.
__host__ __device__ inline double w1(int s, int ab) {
  return ab == 0 ? (s == 0 ? 1.0 : (s == 1 ? 0.5 : 0.0)) : (s == 0 ? 0.0 : (s == 1 ? 0.5 : 1.0));
}

template <int DIM>
__host__ __device__ inline int corner_of(int bx, int by, int bz) {
  if (DIM == 3) return (bz ? 4 : 0) + (by ? (bx ? 2 : 3) : (bx ? 1 : 0));
  return by ? (bx ? 2 : 3) : (bx ? 1 : 0);
}
template <int DIM>
__host__ __device__ inline void bits_of(int a, int& bx, int& by, int& bz) {
  bx = (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : 0;
  by = (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : 0;
  bz = DIM == 3 && a >= 4 ? 1 : 0;
}

// W_t[p][a]: weight of coarse corner a at node p of child t
template <int DIM>
__host__ __device__ inline double wt(int t, int p, int a) {
  int tx, ty, tz, px, py, pz, ax, ay, az;
  bits_of<DIM>(t, tx, ty, tz);
  bits_of<DIM>(p, px, py, pz);
  bits_of<DIM>(a, ax, ay, az);
  double w = w1(tx + px, ax) * w1(ty + py, ay);
  if (DIM == 3) w *= w1(tz + pz, az);
  return w;
}

std::vector<double> galerkin_child_tables(int dim, const double* K0) {
  const int NE = dim == 3 ? 8 : 4, ND = NE * dim;
  std::vector<double> M((size_t)NE * NE * NE * dim * dim, 0.0);
  for (int t = 0; t < NE; ++t)
    for (int a = 0; a < NE; ++a)
      for (int b = 0; b < NE; ++b)
        for (int r = 0; r < dim; ++r)
          for (int s = 0; s < dim; ++s) {
            double v = 0.0;
            for (int p = 0; p < NE; ++p) {
              const double wa = dim == 3 ? wt<3>(t, p, a) : wt<2>(t, p, a);
              if (wa == 0.0) continue;
              for (int q = 0; q < NE; ++q) {
                const double wb = dim == 3 ? wt<3>(t, q, b) : wt<2>(t, q, b);
                if (wb == 0.0) continue;
                v += wa * wb * K0[(p * dim + r) * ND + q * dim + s];
              }
            }
            M[((((size_t)t * NE + a) * NE + b) * dim + r) * dim + s] = v;
          }
  return M;
}

// OR of the constraint bytes of each element's DOFs
template <int DIM>
__global__ void elem_cons_flag_kernel(const Grid g, const uint8_t* __restrict__ cons,
                                      uint8_t* __restrict__ flag) {
  constexpr int NE = DIM == 3 ? 8 : 4;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= g.n_elem()) return;
  int ex, ey, ez = 0;
  if (DIM == 3) {
    ex = (int)(e / ((int64_t)g.ey * g.ez));
    const int r = (int)(e % ((int64_t)g.ey * g.ez));
    ey = r / g.ez;
    ez = r % g.ez;
  } else {
    ex = (int)(e / g.ey);
    ey = (int)(e % g.ey);
  }
  uint8_t f = 0;
  for (int p = 0; p < NE; ++p) {
    int px, py, pz;
    bits_of<DIM>(p, px, py, pz);
    const int64_t m = DIM == 3 ? ((int64_t)(ey + py) * g.nx() + ex + px) * g.nz() + ez + pz
                               : (int64_t)(ey + py) * g.nx() + ex + px;
    for (int d = 0; d < DIM; ++d) f |= cons[m * DIM + d];
  }
  flag[e] = f ? 1 : 0;
}

template <int DIM>
__device__ inline double tw_coarse(const Grid& gc, int i, int j, int k) {
  double t = (i == 0 || i == gc.nx() - 1) ? 1.5 : 2.0;
  t *= (j == 0 || j == gc.ny() - 1) ? 1.5 : 2.0;
  if (DIM == 3) t *= (k == 0 || k == gc.nz() - 1) ? 1.5 : 2.0;
  return t;
}

template <int DIM>
__global__ __launch_bounds__(256) void galerkin_elem_kernel(
    const Grid gf, const Grid gc, const double* __restrict__ K0, const double* __restrict__ Mt,
    const double* __restrict__ E, const uint8_t* __restrict__ cons,
    const uint8_t* __restrict__ eflag, double* __restrict__ Sc) {
  constexpr int NE = DIM == 3 ? 8 : 4, ND = NE * DIM, NO = DIM == 3 ? 27 : 9;
  const int64_t Nc = gc.n_nodes();
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= Nc * NO) return;
  const int o = (int)(tid / Nc);
  const int64_t c = tid % Nc;
  const int oi = DIM == 3 ? o / 9 - 1 : o / 3 - 1;
  const int oj = DIM == 3 ? (o / 3) % 3 - 1 : o % 3 - 1;
  const int ok = DIM == 3 ? o % 3 - 1 : 0;
  int ci, cj, ck;
  {
    const int nzc = gc.nz(), nxc = gc.nx();
    ck = DIM == 3 ? (int)(c % nzc) : 0;
    const int64_t t = DIM == 3 ? c / nzc : c;
    ci = (int)(t % nxc);
    cj = (int)(t / nxc);
  }
  const int ti = ci + oi, tj = cj + oj, tk = ck + ok;
  double blk[DIM][DIM];
#pragma unroll
  for (int r = 0; r < DIM; ++r)
#pragma unroll
    for (int s = 0; s < DIM; ++s) blk[r][s] = 0.0;
  const bool tin = ti >= 0 && ti < gc.nx() && tj >= 0 && tj < gc.ny() &&
                   (DIM == 2 || (tk >= 0 && tk < gc.nz()));
  if (tin) {
    // coarse elements containing c and t
    const int I0 = oi == 0 ? ci - 1 : (oi > 0 ? ci : ci - 1), I1 = oi == 0 ? ci : I0;
    const int J0 = oj == 0 ? cj - 1 : (oj > 0 ? cj : cj - 1), J1 = oj == 0 ? cj : J0;
    const int K0_ = DIM == 3 ? (ok == 0 ? ck - 1 : (ok > 0 ? ck : ck - 1)) : 0;
    const int K1_ = DIM == 3 ? (ok == 0 ? ck : K0_) : 0;
    for (int I = I0; I <= I1; ++I) {
      if (I < 0 || I >= gc.ex) continue;
      for (int J = J0; J <= J1; ++J) {
        if (J < 0 || J >= gc.ey) continue;
        for (int K = K0_; K <= K1_; ++K) {
          if (DIM == 3 && (K < 0 || K >= gc.ez)) continue;
          const int a = corner_of<DIM>(ci - I, cj - J, DIM == 3 ? ck - K : 0);
          const int b = corner_of<DIM>(ti - I, tj - J, DIM == 3 ? tk - K : 0);
          for (int t = 0; t < NE; ++t) {
            int tx, ty, tz;
            bits_of<DIM>(t, tx, ty, tz);
            const int fx = 2 * I + tx, fy = 2 * J + ty, fz = 2 * K + tz;
            const int64_t e = DIM == 3 ? ((int64_t)fx * gf.ey + fy) * gf.ez + fz
                                       : (int64_t)fx * gf.ey + fy;
            const double Ee = E[e];
            if (!eflag[e]) {
              const double* m = Mt + (((size_t)t * NE + a) * NE + b) * DIM * DIM;
#pragma unroll
              for (int r = 0; r < DIM; ++r)
#pragma unroll
                for (int s = 0; s < DIM; ++s) blk[r][s] += Ee * m[r * DIM + s];
            } else {
              // child touching constrained DOFs: W^T (M K0 M) W on the fly
              for (int p = 0; p < NE; ++p) {
                const double wa = wt<DIM>(t, p, a);
                if (wa == 0.0) continue;
                int px, py, pz;
                bits_of<DIM>(p, px, py, pz);
                const int64_t np = DIM == 3 ? ((int64_t)(fy + py) * gf.nx() + fx + px) * gf.nz() + fz + pz
                                            : (int64_t)(fy + py) * gf.nx() + fx + px;
                for (int q = 0; q < NE; ++q) {
                  const double wb = wt<DIM>(t, q, b);
                  if (wb == 0.0) continue;
                  int qx, qy, qz;
                  bits_of<DIM>(q, qx, qy, qz);
                  const int64_t nq = DIM == 3 ? ((int64_t)(fy + qy) * gf.nx() + fx + qx) * gf.nz() + fz + qz
                                              : (int64_t)(fy + qy) * gf.nx() + fx + qx;
                  for (int r = 0; r < DIM; ++r) {
                    if (cons[np * DIM + r]) continue;
                    for (int s = 0; s < DIM; ++s) {
                      if (cons[nq * DIM + s]) continue;
                      blk[r][s] += Ee * wa * wb * K0[(p * DIM + r) * ND + q * DIM + s];
                    }
                  }
                }
              }
            }
          }
        }
      }
    }
    // identity rows of constrained fine DOFs: sum_f P[f,c] P[f,t] over constrained (f, d)
    for (int fx = max(2 * ci - 1, 2 * ti - 1); fx <= min(2 * ci + 1, 2 * ti + 1); ++fx) {
      if (fx < 0 || fx >= gf.nx()) continue;
      for (int fy = max(2 * cj - 1, 2 * tj - 1); fy <= min(2 * cj + 1, 2 * tj + 1); ++fy) {
        if (fy < 0 || fy >= gf.ny()) continue;
        const int fz0 = DIM == 3 ? max(2 * ck - 1, 2 * tk - 1) : 0;
        const int fz1 = DIM == 3 ? min(2 * ck + 1, 2 * tk + 1) : 0;
        for (int fz = fz0; fz <= fz1; ++fz) {
          if (DIM == 3 && (fz < 0 || fz >= gf.nz())) continue;
          const int64_t f = DIM == 3 ? ((int64_t)fy * gf.nx() + fx) * gf.nz() + fz : (int64_t)fy * gf.nx() + fx;
          double w = (fx == 2 * ci ? 1.0 : 0.5) * (fx == 2 * ti ? 1.0 : 0.5) *
                     (fy == 2 * cj ? 1.0 : 0.5) * (fy == 2 * tj ? 1.0 : 0.5);
          if (DIM == 3) w *= (fz == 2 * ck ? 1.0 : 0.5) * (fz == 2 * tk ? 1.0 : 0.5);
#pragma unroll
          for (int d = 0; d < DIM; ++d)
            if (cons[f * DIM + d]) blk[d][d] += w;
        }
      }
    }
    const double inv = 1.0 / (tw_coarse<DIM>(gc, ci, cj, ck) * tw_coarse<DIM>(gc, ti, tj, tk));
#pragma unroll
    for (int r = 0; r < DIM; ++r)
#pragma unroll
      for (int s = 0; s < DIM; ++s) blk[r][s] *= inv;
  }
#pragma unroll
  for (int r = 0; r < DIM; ++r)
#pragma unroll
    for (int s = 0; s < DIM; ++s) Sc[((int64_t)(o * DIM + r) * DIM + s) * Nc + c] = blk[r][s];
}

bool galerkin_elem_supported(const Grid& g) {
  return g.ex % 2 == 0 && g.ey % 2 == 0 && (g.dim == 2 || g.ez % 2 == 0);
}

int galerkin_from_elements(const Grid& gf, const double* K0, const double* Mt, const double* E,
                           const uint8_t* cons, uint8_t* eflag, double* Sc, hipStream_t s) {
  ET_REQUIRE(galerkin_elem_supported(gf), "element Galerkin needs even element counts");
  const Grid gc = gf.coarse();
  const int NO = gf.dim == 3 ? 27 : 9;
  if (gf.dim == 3) {
    elem_cons_flag_kernel<3><<<grid_for(gf.n_elem()), kBlock, 0, s>>>(gf, cons, eflag);
    galerkin_elem_kernel<3><<<grid_for(gc.n_nodes() * NO), kBlock, 0, s>>>(gf, gc, K0, Mt, E, cons, eflag, Sc);
  } else {
    elem_cons_flag_kernel<2><<<grid_for(gf.n_elem()), kBlock, 0, s>>>(gf, cons, eflag);
    galerkin_elem_kernel<2><<<grid_for(gc.n_nodes() * NO), kBlock, 0, s>>>(gf, gc, K0, Mt, E, cons, eflag, Sc);
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

}  // namespace et
