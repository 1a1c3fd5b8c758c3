// Per-element post-processing of a solved displacement field: strain, stress,
// strain energy and von Mises stress of every element of a structured grid.
//
// Replaces FEA_Integrals_node_basis_parallel_cuda_kernel
// (core/_templated_cuda.cu:1681-1719) and the von Mises expressions of
// TopOpt.FEA_integrals (TopOpt/_cuda.py:121-124), fused into one pass: one
// thread per element gathers u_e (closed-form node indices, no connectivity
// array), forms B u_e, D (B u_e) and 1/2 E_e u_e^T K0 u_e, and the von Mises
// stress from the stress it just formed.  B, D and K0 are indexed with
// compile-time offsets only, so every entry is a wave-uniform (scalar) load.
// Sums run in the reference's order (k ascending) for each output.
#include "common.hpp"
#include "stiffness.hpp"

namespace et {

template <int DIM>
__global__ __launch_bounds__(256) void fea_kernel(const Grid g, const double* __restrict__ K0,
                                                  const double* __restrict__ D,
                                                  const double* __restrict__ B,
                                                  const double* __restrict__ E,
                                                  const double* __restrict__ U,
                                                  double* __restrict__ strain,
                                                  double* __restrict__ stress,
                                                  double* __restrict__ energy,
                                                  double* __restrict__ vm) {
  constexpr int NE = DIM == 3 ? 8 : 4;
  constexpr int ND = NE * DIM;
  constexpr int NB = DIM == 3 ? 6 : 3;
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
  const int nx = g.nx(), nz = g.nz();
  double ue[ND];
#pragma unroll
  for (int b = 0; b < NE; ++b) {
    const int bx = ex + (DIM == 3 ? hx(b) : qx(b));
    const int by = ey + (DIM == 3 ? hy(b) : qy(b));
    const int bz = ez + (DIM == 3 ? hz(b) : 0);
    const int64_t m = DIM == 3 ? ((int64_t)by * nx + bx) * nz + bz : (int64_t)by * nx + bx;
#pragma unroll
    for (int c = 0; c < DIM; ++c) ue[DIM * b + c] = U[m * DIM + c];
  }
  double eps[NB], sig[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < ND; ++k) v += B[j * ND + k] * ue[k];
    eps[j] = v;
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) v += D[j * NB + k] * eps[k];
    sig[j] = v;
  }
  const double w = E[e];
  double en = 0.0;
#pragma unroll
  for (int j = 0; j < ND; ++j) {
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < ND; ++k) v += w * K0[j * ND + k] * ue[k];
    en += 0.5 * ue[j] * v;
  }
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    strain[e * NB + j] = eps[j];
    stress[e * NB + j] = sig[j];
  }
  energy[e] = en;
  if (vm != nullptr) {
    double q;
    if (DIM == 2) {
      q = sig[0] * sig[0] + sig[1] * sig[1] - sig[0] * sig[1] + 3.0 * sig[2] * sig[2];
    } else {
      const double a = sig[0] - sig[1], b = sig[1] - sig[2], c = sig[2] - sig[0];
      q = 0.5 * (a * a + b * b + c * c + 6.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5]));
    }
    vm[e] = sqrt(q);
  }
}

int fea_integrals(const Grid& g, const double* K0, const double* D, const double* B, const double* E,
                  const double* U, double* strain, double* stress, double* energy, double* vm,
                  hipStream_t s) {
  const unsigned grid = grid_for(g.n_elem());
  if (g.dim == 3)
    fea_kernel<3><<<grid, kBlock, 0, s>>>(g, K0, D, B, E, U, strain, stress, energy, vm);
  else
    fea_kernel<2><<<grid, kBlock, 0, s>>>(g, K0, D, B, E, U, strain, stress, energy, vm);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

}  // namespace et
