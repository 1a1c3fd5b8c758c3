// Multi-material SIMP models and their OC update.
//
// PenalizedMultiMaterial / PenalizedMultiMaterialRP (MaterialModels.py:100-369):
// densities r (n_el x m, row-major), moduli E_mat (m):
//   E = clip(sum_i E_i r_i^p prod_{j != i} w_j, void, inf),   w_j = 1 - r_j^p
//       (RP: w_j = (1 - r_j)^p),  r = clip(rho, void, 1);  plain: clip(rho . E_mat, void, 1)
//   dE/dr_k = E_k p r_k^(p-1) prod_{j != k} w_j - sum_{i != k} E_i r_i^p dw_k prod_{j != i,k} w_j
//       with dw_k = p r_k^(p-1) (RP: p (1 - r_k)^(p-1)), the expression :178-206 evaluates.
// One thread per element computes all m outputs from registers.
//
// OC with one multiplier per material (TopOpt/_cuda.py:242-280 with n_sets = m): the m
// brackets are bisected in lockstep until every one has closed; each pass evaluates the
// volume of every material at the 7 midpoints of its next 3 bisection levels (as the
// single-material update does), and a one-block kernel takes the steps.  Per material
// the constraint is sum_e r_ek A / vol <= mass_k; with rho_mat it is the total
// sum_e sum_k r_ek A rho_mat_k / vol <= mass, one decision for every material (:208-221).
#include <algorithm>
#include <cmath>

#include "blas.hpp"
#include "common.hpp"
#include "stiffness.hpp"

namespace et {

constexpr int kMmMax = 4;  // materials per element handled on the device

template <int M>
__global__ __launch_bounds__(256) void multimat_kernel(int64_t n, const double* __restrict__ rho, const double* Emat,
                                                       double penal, double vmin, int rp, int mode,
                                                       double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  double r[M], w[M], rp_[M], dw[M];
  if (mode == 1) {  // plain
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) s += rho[e * M + i] * Emat[i];
    out[e] = fmin(fmax(s, vmin), 1.0);
    return;
  }
#pragma unroll
  for (int i = 0; i < M; ++i) {
    r[i] = fmin(fmax(rho[e * M + i], vmin), 1.0);
    rp_[i] = pow(r[i], penal);
    w[i] = rp ? pow(1.0 - r[i], penal) : 1.0 - rp_[i];
    dw[i] = rp ? penal * pow(1.0 - r[i], penal - 1.0) : penal * pow(r[i], penal - 1.0);
  }
  if (mode == 0) {
    double E = 0.0;
#pragma unroll
    for (int i = 0; i < M; ++i) {
      double t = rp_[i];
#pragma unroll
      for (int j = 0; j < M; ++j)
        if (j != i) t *= w[j];
      E += Emat[i] * t;
    }
    out[e] = fmax(E, vmin);
    return;
  }
#pragma unroll
  for (int k = 0; k < M; ++k) {
    double t = Emat[k] * penal * pow(r[k], penal - 1.0);
#pragma unroll
    for (int j = 0; j < M; ++j)
      if (j != k) t *= w[j];
#pragma unroll
    for (int i = 0; i < M; ++i) {
      if (i == k) continue;
      double s = Emat[i] * rp_[i] * dw[k];
#pragma unroll
      for (int j = 0; j < M; ++j)
        if (j != i && j != k) s *= w[j];
      t -= s;
    }
    out[e * M + k] = t;
  }
}

int multimat(int64_t n, int m, const double* rho, const double* Emat, double penal, double vmin, int rp, int mode,
             double* out, hipStream_t s) {
  ET_REQUIRE(m >= 1 && m <= kMmMax, "multi-material: 1 to 4 materials on the device");
  const unsigned g = grid_for(n);
  switch (m) {
    case 1: multimat_kernel<1><<<g, kBlock, 0, s>>>(n, rho, Emat, penal, vmin, rp, mode, out); break;
    case 2: multimat_kernel<2><<<g, kBlock, 0, s>>>(n, rho, Emat, penal, vmin, rp, mode, out); break;
    case 3: multimat_kernel<3><<<g, kBlock, 0, s>>>(n, rho, Emat, penal, vmin, rp, mode, out); break;
    default: multimat_kernel<4><<<g, kBlock, 0, s>>>(n, rho, Emat, penal, vmin, rp, mode, out); break;
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

// ------------------------------------------------------------------- OC ---

constexpr int kMDepth = 3, kMCand = 7, kMPasses = 26;

struct OcMState {
  double l1[kMmMax], l2[kMmMax], lmid[kMmMax], mass[kMmMax], rhom[kMmMax];
  double sabs;
  int done, steps, m, use_rhom;
};

__device__ inline bool mo_open(double l1, double l2) { return (l2 - l1) / (l2 + l1) > 1e-4; }

__device__ inline void mo_midpoints(double l1, double l2, double (&mid)[kMCand]) {
  double lo[kMCand], hi[kMCand];
  lo[0] = l1;
  hi[0] = l2;
#pragma unroll
  for (int i = 0; i < kMCand; ++i) {
    mid[i] = 0.5 * (lo[i] + hi[i]);
    if (2 * i + 2 < kMCand) {
      lo[2 * i + 1] = lo[i];
      hi[2 * i + 1] = mid[i];
      lo[2 * i + 2] = mid[i];
      hi[2 * i + 2] = hi[i];
    }
  }
}

__device__ __forceinline__ double mo_value(double x, double p_raw, bool all_zero, double move, double lmid) {
  const double xu = fmin(fmax(x + move, 0.0), 1.0);
  const double xl = fmin(fmax(x - move, 0.0), 1.0);
  const double p = all_zero ? 1e-3 : fmax(1e-12, p_raw);
  return fmax(0.0, fmax(xl, fmin(1.0, fmin(xu, p / lmid))));
}

__global__ __launch_bounds__(256) void mo_prep_kernel(int64_t nm, const double* __restrict__ x,
                                                      const double* __restrict__ dc, double A,
                                                      double* __restrict__ ocp, double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nm; t += (int64_t)gridDim.x * blockDim.x) {
    double q = -dc[t] / A;
    double v = x[t] * sqrt(q);  // sqrt of a negative value: NaN -> 0 (nan_to_num)
    if (v != v) v = 0.0;
    ocp[t] = v;
    s += fabs(v);
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void mo_init_kernel(int nb, const double* __restrict__ partial, OcMState* st) {
  double s = 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) s += partial[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) {
    st->sabs = s;
    int open = 0;
    for (int k = 0; k < st->m; ++k) {
      st->l1[k] = 1e-9;
      st->l2[k] = 1e9;
      st->lmid[k] = 0.5 * (st->l1[k] + st->l2[k]);
      open |= mo_open(st->l1[k], st->l2[k]);
    }
    st->done = !open;
    st->steps = 0;
  }
}

template <int M>
__global__ __launch_bounds__(256) void mo_pass_kernel(int64_t n, const double* __restrict__ x,
                                                      const double* __restrict__ ocp, double A, double move,
                                                      const OcMState* __restrict__ st, double* __restrict__ partial,
                                                      int nb) {
  if (st->done) return;
  double mid[M][kMCand];
#pragma unroll
  for (int k = 0; k < M; ++k) mo_midpoints(st->l1[k], st->l2[k], mid[k]);
  const bool all_zero = st->sabs == 0.0;
  double s[M][kMCand];
#pragma unroll
  for (int k = 0; k < M; ++k)
#pragma unroll
    for (int c = 0; c < kMCand; ++c) s[k][c] = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < M; ++k) {
      const double xv = x[e * M + k], pv = ocp[e * M + k];
#pragma unroll
      for (int c = 0; c < kMCand; ++c) s[k][c] += mo_value(xv, pv, all_zero, move, mid[k][c]) * A;
    }
  }
#pragma unroll
  for (int k = 0; k < M; ++k)
#pragma unroll
    for (int c = 0; c < kMCand; ++c) {
      const double v = block_sum256(s[k][c]);
      if (threadIdx.x == 0) partial[(k * kMCand + c) * nb + blockIdx.x] = v;
    }
}

__global__ void mo_decide_kernel(const double* __restrict__ partial, int nb, double vol, OcMState* st) {
  if (st->done) return;
  __shared__ double tot[kMmMax * kMCand];
  const int m = st->m;
  for (int q = 0; q < m * kMCand; ++q) {
    double s = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) s += partial[q * nb + t];
    s = block_sum256(s);
    if (threadIdx.x == 0) tot[q] = s;
  }
  if (threadIdx.x != 0) return;
  double mid[kMmMax][kMCand];
  for (int k = 0; k < m; ++k) mo_midpoints(st->l1[k], st->l2[k], mid[k]);
  int node[kMmMax] = {0, 0, 0, 0};
  for (int d = 0; d < kMDepth; ++d) {
    bool feas[kMmMax];
    if (st->use_rhom) {
      double V = 0.0;
      for (int k = 0; k < m; ++k) V += tot[k * kMCand + node[k]] * st->rhom[k];
      for (int k = 0; k < m; ++k) feas[k] = V / vol <= st->mass[0];
    } else {
      for (int k = 0; k < m; ++k) feas[k] = tot[k * kMCand + node[k]] / vol - st->mass[k] <= 0.0;
    }
    int open = 0;
    for (int k = 0; k < m; ++k) {
      const double lm = mid[k][node[k]];
      st->lmid[k] = lm;
      if (feas[k])
        st->l2[k] = lm;
      else
        st->l1[k] = lm;
      open |= mo_open(st->l1[k], st->l2[k]);
      node[k] = feas[k] ? 2 * node[k] + 1 : 2 * node[k] + 2;
    }
    st->steps += 1;
    if (!open) {
      st->done = 1;
      break;
    }
  }
}

template <int M>
__global__ __launch_bounds__(256) void mo_final_kernel(int64_t n, const double* __restrict__ x,
                                                       const double* __restrict__ ocp, double move,
                                                       const OcMState* __restrict__ st, double* __restrict__ xn) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
#pragma unroll
  for (int k = 0; k < M; ++k)
    xn[e * M + k] = mo_value(x[e * M + k], ocp[e * M + k], st->sabs == 0.0, move, st->lmid[k]);
}

int64_t oc_multi_work_size(int64_t n, int m) {
  return 256 + n * m * (int64_t)sizeof(double) + (int64_t)kMmMax * kMCand * kRedBlocks * (int64_t)sizeof(double);
}

template <int M>
static void mo_launch_passes(int64_t n, const double* x, const double* ocp, double A, double move, OcMState* st,
                             double* partial, int nb, double vol, hipStream_t s) {
  for (int pass = 0; pass < kMPasses; ++pass) {
    mo_pass_kernel<M><<<nb, 256, 0, s>>>(n, x, ocp, A, move, st, partial, nb);
    mo_decide_kernel<<<1, 256, 0, s>>>(partial, nb, vol, st);
  }
}

int oc_update_multi(int64_t n, int m, const double* x, const double* dc, double A, double vol, double move,
                    const double* mass, const double* rho_mat, double* xn, void* work, int* steps, hipStream_t s) {
  ET_REQUIRE(n > 0 && m >= 1 && m <= kMmMax && x && dc && xn && work && mass, "oc_update_multi: bad arguments");
  ET_REQUIRE(A > 0 && vol > 0, "oc_update_multi: element measure and volume must be positive");
  ET_REQUIRE((reinterpret_cast<uintptr_t>(work) & 7) == 0, "oc_update_multi: work must be 8-byte aligned");
  OcMState h{};
  h.m = m;
  h.use_rhom = rho_mat != nullptr;
  for (int k = 0; k < m; ++k) {
    h.mass[k] = rho_mat ? mass[0] : mass[k];
    h.rhom[k] = rho_mat ? rho_mat[k] : 0.0;
  }
  OcMState* st = static_cast<OcMState*>(work);
  double* ocp = reinterpret_cast<double*>(static_cast<char*>(work) + 256);
  double* partial = ocp + n * m;
  const int nb = (int)std::min<int64_t>(kRedBlocks, (n * m + 255) / 256);
  // the constant part of the state (m, masses, material densities) goes with the stream
  ET_CHECK_HIP(hipMemcpyAsync(st, &h, sizeof(OcMState), hipMemcpyHostToDevice, s));
  mo_prep_kernel<<<nb, 256, 0, s>>>(n * m, x, dc, A, ocp, partial);
  mo_init_kernel<<<1, 256, 0, s>>>(nb, partial, st);
  switch (m) {
    case 1: mo_launch_passes<1>(n, x, ocp, A, move, st, partial, nb, vol, s); break;
    case 2: mo_launch_passes<2>(n, x, ocp, A, move, st, partial, nb, vol, s); break;
    case 3: mo_launch_passes<3>(n, x, ocp, A, move, st, partial, nb, vol, s); break;
    default: mo_launch_passes<4>(n, x, ocp, A, move, st, partial, nb, vol, s); break;
  }
  const unsigned g = grid_for(n);
  switch (m) {
    case 1: mo_final_kernel<1><<<g, kBlock, 0, s>>>(n, x, ocp, move, st, xn); break;
    case 2: mo_final_kernel<2><<<g, kBlock, 0, s>>>(n, x, ocp, move, st, xn); break;
    case 3: mo_final_kernel<3><<<g, kBlock, 0, s>>>(n, x, ocp, move, st, xn); break;
    default: mo_final_kernel<4><<<g, kBlock, 0, s>>>(n, x, ocp, move, st, xn); break;
  }
  ET_CHECK_LAUNCH();
  // the host copy of the state must outlive the transfer
  ET_CHECK_HIP(hipMemcpyAsync(&h, st, sizeof(OcMState), hipMemcpyDeviceToHost, s));
  ET_CHECK_HIP(hipStreamSynchronize(s));
  if (steps) *steps = h.steps;
  if (!h.done) {
    set_error("oc_update_multi: bisection did not terminate");
    return ET_ENOCONV;
  }
  return ET_OK;
}

}  // namespace et
