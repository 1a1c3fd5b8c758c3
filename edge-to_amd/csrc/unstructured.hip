// General (unstructured) meshes: element matrices stored per element, any mix of
// element types (T3 / Q4 in 2-D, T4 / H8 in 3-D), DOFs interleaved node * dof + d.
//
// Layout (et_umesh): element e owns the node list elem_nodes[elem_ptr[e] ..
// elem_ptr[e+1]) and the row-major (size * dof)^2 matrix at K + K_ptr[e]; the
// node -> element star is the stable argsort of the flat node list, star_ptr /
// star / el_of (the reference's node_ids / sorter / el_ids,
// kernels/_cuda.py:320-331), built here on the device.
//
// Every operator is a gather with a fixed summation order, no atomics:
//   matvec    one thread per DOF row: for each element of the node's star (in star
//             order), the element row K_e[(a, k), :] . u_e over unconstrained columns,
//             times E_e (mat_vec_node_basis_parallel_{full,flat}_wcon,
//             core/_templated_cuda.cu:669-780); constrained rows return u;
//   assembly  the same rows as a CSR matrix (construct()), whose product reads 12 bytes
//             per stored entry instead of each element matrix per node of the element;
//   diagonal  the same rows' diagonal entries (get_diagonal_node_basis_{full,flat},
//             :1116-1213), 1 on constrained rows;
//   energy    one thread per element, -u_e^T K_e u_e (process_dk_{full,flat},
//             :196-231), optionally times the SIMP factor p clip(x, void, 1)^(p-1);
//   FEA       strain B_e u_e, stress D_e strain, energy 1/2 E_e u_e^T K_e u_e and
//             von Mises (FEA_Integrals_node_basis_parallel_{full,flat},
//             :1722-1800 and TopOpt/_cuda.py:123-127).
// The density filter of general meshes is a CSR matrix (core/_filter.py:126-182);
// its product runs 8 lanes per row with a fixed-shape lane reduction.
#include <algorithm>
#include <vector>

#include "blas.hpp"
#include "common.hpp"
#include "stiffness.hpp"

namespace et {

// ------------------------------------------------------------------ star ---

// node ids outside [0, n_nodes) are counted in *bad instead
__global__ void ustar_count_kernel(int64_t n_flat, int64_t n_nodes, const int* __restrict__ nodes,
                                   int* __restrict__ count, int* __restrict__ bad) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_flat) return;
  const int nd = nodes[f];
  if (nd < 0 || nd >= n_nodes)
    atomicAdd(bad, 1);
  else
    atomicAdd(&count[nd], 1);
}

// exclusive scan of count[0..n) into out[0..n]: per-block sums, one block scans the
// block sums, then each block adds its offset (2048 values per block)
constexpr int kScanPer = 2048;

__global__ __launch_bounds__(256) void scan_block_sum_kernel(int64_t n, const int* __restrict__ in,
                                                            int64_t* __restrict__ bsum) {
  const int64_t base = (int64_t)blockIdx.x * kScanPer;
  int64_t s = 0;
  for (int t = threadIdx.x; t < kScanPer; t += 256)
    if (base + t < n) s += in[base + t];
  s = block_sum256(s);
  if (threadIdx.x == 0) bsum[blockIdx.x] = s;
}

__global__ void scan_offsets_kernel(int nb, int64_t* __restrict__ bsum) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int64_t run = 0;
    for (int b = 0; b < nb; ++b) {
      const int64_t v = bsum[b];
      bsum[b] = run;
      run += v;
    }
  }
}

__global__ __launch_bounds__(256) void scan_apply_kernel(int64_t n, const int* __restrict__ in,
                                                        const int64_t* __restrict__ boff, int* __restrict__ out) {
  // one block per 2048 values; thread t scans 8 consecutive values, the block scans the thread sums
  __shared__ int64_t part[256];
  const int64_t base = (int64_t)blockIdx.x * kScanPer + threadIdx.x * 8;
  int64_t v[8], s = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    v[q] = base + q < n ? in[base + q] : 0;
    s += v[q];
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    const int64_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
    __syncthreads();
    part[threadIdx.x] += add;
    __syncthreads();
  }
  int64_t run = boff[blockIdx.x] + part[threadIdx.x] - s;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    if (base + q < n) out[base + q] = (int)run;
    run += v[q];
  }
}

// out[n] = out[n-1] + in[n-1] (the total)
__global__ void scan_total_kernel(int64_t n, const int* __restrict__ in, int* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = out[n - 1] + in[n - 1];
}

__global__ void ustar_fill_kernel(int64_t n_flat, const int* __restrict__ nodes, const int* __restrict__ star_ptr,
                                  int* __restrict__ cursor, int* __restrict__ star) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_flat) return;
  const int nd = nodes[f];
  star[star_ptr[nd] + atomicAdd(&cursor[nd], 1)] = (int)f;
}

// the positions of each node's segment in ascending order (a stable argsort)
__global__ void ustar_sort_kernel(int64_t n_nodes, const int* __restrict__ star_ptr, int* __restrict__ star) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const int a = star_ptr[i], b = star_ptr[i + 1];
  for (int p = a + 1; p < b; ++p) {
    const int v = star[p];
    int q = p - 1;
    while (q >= a && star[q] > v) {
      star[q + 1] = star[q];
      --q;
    }
    star[q + 1] = v;
  }
}

__global__ void uel_of_kernel(int64_t n_el, const int* __restrict__ elem_ptr, int* __restrict__ el_of) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_el) return;
  for (int f = elem_ptr[e]; f < elem_ptr[e + 1]; ++f) el_of[f] = (int)e;
}

int umesh_star(int64_t n_nodes, int64_t n_el, const int* elem_ptr, const int* elem_nodes, int* star_ptr, int* star,
               int* el_of, hipStream_t s) {
  ET_REQUIRE(n_nodes > 0 && n_el > 0 && elem_ptr && elem_nodes && star_ptr && star && el_of,
             "umesh_star: bad arguments");
  int flat_end = 0;
  ET_CHECK_HIP(hipMemcpyAsync(&flat_end, elem_ptr + n_el, sizeof(int), hipMemcpyDeviceToHost, s));
  ET_CHECK_HIP(hipStreamSynchronize(s));
  const int64_t n_flat = flat_end;
  ET_REQUIRE(n_flat > 0, "umesh_star: empty element lists");
  const int nb = (int)((n_nodes + kScanPer - 1) / kScanPer);
  // count[n_nodes] holds the out-of-range count
  int* count = nullptr;
  int64_t* bsum = nullptr;
  ET_CHECK_HIP(hipMalloc(&count, (size_t)(n_nodes + 1) * sizeof(int)));
  if (hipMalloc(&bsum, (size_t)nb * sizeof(int64_t)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(count);
    set_error("umesh_star: device allocation failed");
    return ET_EHIP;
  }
  auto run = [&]() -> int {
    ET_CHECK_HIP(hipMemsetAsync(count, 0, (size_t)(n_nodes + 1) * sizeof(int), s));
    ustar_count_kernel<<<grid_for(n_flat), kBlock, 0, s>>>(n_flat, n_nodes, elem_nodes, count, count + n_nodes);
    ET_CHECK_LAUNCH();
    int bad = 0;
    ET_CHECK_HIP(hipMemcpyAsync(&bad, count + n_nodes, sizeof(int), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    ET_REQUIRE(bad == 0, "umesh_star: element node ids outside [0, n_nodes)");
    scan_block_sum_kernel<<<nb, 256, 0, s>>>(n_nodes, count, bsum);
    scan_offsets_kernel<<<1, 64, 0, s>>>(nb, bsum);
    scan_apply_kernel<<<nb, 256, 0, s>>>(n_nodes, count, bsum, star_ptr);
    scan_total_kernel<<<1, 64, 0, s>>>(n_nodes, count, star_ptr);
    ET_CHECK_LAUNCH();
    ET_CHECK_HIP(hipMemsetAsync(count, 0, (size_t)n_nodes * sizeof(int), s));
    ustar_fill_kernel<<<grid_for(n_flat), kBlock, 0, s>>>(n_flat, elem_nodes, star_ptr, count, star);
    ustar_sort_kernel<<<grid_for(n_nodes), kBlock, 0, s>>>(n_nodes, star_ptr, star);
    uel_of_kernel<<<grid_for(n_el), kBlock, 0, s>>>(n_el, elem_ptr, el_of);
    ET_CHECK_LAUNCH();
    ET_CHECK_HIP(hipStreamSynchronize(s));
    return ET_OK;
  };
  const int st = run();
  (void)hipFree(count);
  (void)hipFree(bsum);
  return st;
}

// -------------------------------------------------------------- operator ---

// MODE 0: y = A u; 1: y = b - A u.  One thread per DOF row, the star's elements in star
// order (the reference's per-row order: out += (K row . u) * E_e per element).  Measured
// against a two-pass form (element rows into a work array, then per-row sums of the
// star): 188 us against 223 us for a 165 888-element tetrahedral mesh.
template <int DOF, int MODE>
__global__ __launch_bounds__(256) void umatvec_kernel(const et_umesh m, const double* __restrict__ E,
                                                      const uint8_t* __restrict__ cons,
                                                      const double* __restrict__ u, const double* __restrict__ b,
                                                      double* __restrict__ y) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= m.n_nodes * DOF) return;
  const int64_t i = row / DOF;
  const int k = (int)(row - i * DOF);
  double out;
  if (cons != nullptr && cons[row]) {
    out = u[row];
  } else {
    out = 0.0;
    for (int p = m.star_ptr[i]; p < m.star_ptr[i + 1]; ++p) {
      const int f = m.star[p];
      const int e = m.el_of[f];
      const int st = m.elem_ptr[e], size = m.elem_ptr[e + 1] - st;
      const int nc = size * DOF;
      const double* Kr = m.K + m.K_ptr[e] + (int64_t)((f - st) * DOF + k) * nc;
      double v = 0.0;
      for (int a = 0; a < size; ++a) {
        const int64_t c0 = (int64_t)m.elem_nodes[st + a] * DOF;
#pragma unroll
        for (int d = 0; d < DOF; ++d) {
          const double kv = Kr[a * DOF + d];
          if (cons == nullptr || !cons[c0 + d]) v += kv * u[c0 + d];
        }
      }
      out += v * E[e];
    }
  }
  y[row] = MODE == 0 ? out : b[row] - out;
}

template <int DOF>
__global__ __launch_bounds__(256) void udiag_kernel(const et_umesh m, const double* __restrict__ E,
                                                    const uint8_t* __restrict__ cons, double* __restrict__ diag) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= m.n_nodes * DOF) return;
  const int64_t i = row / DOF;
  const int k = (int)(row - i * DOF);
  if (cons != nullptr && cons[row]) {
    diag[row] = 1.0;
    return;
  }
  double out = 0.0;
  for (int p = m.star_ptr[i]; p < m.star_ptr[i + 1]; ++p) {
    const int f = m.star[p];
    const int e = m.el_of[f];
    const int st = m.elem_ptr[e], size = m.elem_ptr[e + 1] - st;
    const int nc = size * DOF;
    const double* Kr = m.K + m.K_ptr[e] + (int64_t)((f - st) * DOF + k) * nc;
    // every column of the row that is this DOF (an element may list a node twice)
    for (int a = 0; a < size; ++a)
      if (m.elem_nodes[st + a] == i) out += Kr[a * DOF + k] * E[e];
  }
  diag[row] = out;
}

// MODE 0: out = -u_e^T K_e u_e; 1: times p clip(xt, void, 1)^(p-1)
template <int DOF, int MODE>
__global__ __launch_bounds__(256) void uenergy_kernel(const et_umesh m, const double* __restrict__ U,
                                                      const double* __restrict__ xt, double penal, double vmin,
                                                      double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.n_el) return;
  const int st = m.elem_ptr[e], size = m.elem_ptr[e + 1] - st;
  const int nc = size * DOF;
  const double* Ke = m.K + m.K_ptr[e];
  double acc = 0.0;
  for (int j = 0; j < nc; ++j) {
    double v = 0.0;
    for (int a = 0; a < size; ++a) {
      const int64_t c0 = (int64_t)m.elem_nodes[st + a] * DOF;
#pragma unroll
      for (int d = 0; d < DOF; ++d) v += Ke[(int64_t)j * nc + a * DOF + d] * U[c0 + d];
    }
    acc -= v * U[(int64_t)m.elem_nodes[st + j / DOF] * DOF + j % DOF];
  }
  if (MODE == 1) {
    const double x = fmin(fmax(xt[e], vmin), 1.0);
    acc *= penal == 3.0 ? 3.0 * (x * x) : penal * pow(x, penal - 1.0);
  }
  out[e] = acc;
}

template <int DOF>
__global__ __launch_bounds__(256) void ufea_kernel(const et_umesh m, int nb, const double* __restrict__ B,
                                                   const int64_t* __restrict__ B_ptr, const double* __restrict__ D,
                                                   const int64_t* __restrict__ D_ptr, const double* __restrict__ E,
                                                   const double* __restrict__ U, double* __restrict__ strain,
                                                   double* __restrict__ stress, double* __restrict__ energy,
                                                   double* __restrict__ vm) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.n_el) return;
  const int st = m.elem_ptr[e], size = m.elem_ptr[e + 1] - st;
  const int nc = size * DOF;
  const double* Be = B + B_ptr[e];
  const double* De = D + D_ptr[e];
  const double w = E[e];
  double eps[6], sig[6];
  for (int j = 0; j < nb; ++j) {
    double s = 0.0;
    for (int a = 0; a < size; ++a) {
      const int64_t c0 = (int64_t)m.elem_nodes[st + a] * DOF;
#pragma unroll
      for (int d = 0; d < DOF; ++d) s += Be[(int64_t)j * nc + a * DOF + d] * U[c0 + d];
    }
    eps[j] = s;
    strain[e * nb + j] = s;
  }
  for (int j = 0; j < nb; ++j) {
    double s = 0.0;
    for (int q = 0; q < nb; ++q) s += De[j * nb + q] * eps[q];
    sig[j] = s;
    stress[e * nb + j] = s;
  }
  const double* Ke = m.K + m.K_ptr[e];
  double en = 0.0;
  for (int j = 0; j < nc; ++j) {
    double v = 0.0;
    for (int a = 0; a < size; ++a) {
      const int64_t c0 = (int64_t)m.elem_nodes[st + a] * DOF;
#pragma unroll
      for (int d = 0; d < DOF; ++d) v += w * Ke[(int64_t)j * nc + a * DOF + d] * U[c0 + d];
    }
    en += 0.5 * U[(int64_t)m.elem_nodes[st + j / DOF] * DOF + j % DOF] * v;
  }
  energy[e] = en;
  if (vm != nullptr) {
    double q;
    if (nb == 3) {
      q = sig[0] * sig[0] + sig[1] * sig[1] - sig[0] * sig[1] + 3.0 * sig[2] * sig[2];
    } else {
      const double a = sig[0] - sig[1], b2 = sig[1] - sig[2], c = sig[2] - sig[0];
      q = 0.5 * (a * a + b2 * b2 + c * c + 6.0 * (sig[3] * sig[3] + sig[4] * sig[4] + sig[5] * sig[5]));
    }
    vm[e] = sqrt(q);
  }
}

// ------------------------------------------------------------------- CSR ---

// y = A x (MODE 0) or y = b - A x (MODE 1), 8 lanes per row: lane q sums entries q,
// q + 8, ... of the row, then the lanes meet in a fixed tree (4, 2, 1)
template <int MODE>
__global__ __launch_bounds__(256) void csr_spmv_kernel(int64_t n_rows, const int* __restrict__ row_ptr,
                                                       const int* __restrict__ cols,
                                                       const double* __restrict__ vals,
                                                       const double* __restrict__ x, const double* __restrict__ b,
                                                       double* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t row = t >> 3;
  const int q = (int)(t & 7);
  double s = 0.0;
  if (row < n_rows) {
    const int a = row_ptr[row], e = row_ptr[row + 1];
    for (int p = a + q; p < e; p += 8) s += vals[p] * x[cols[p]];
  }
  s += __shfl_xor(s, 4, 8);
  s += __shfl_xor(s, 2, 8);
  s += __shfl_xor(s, 1, 8);
  if (row < n_rows && q == 0) y[row] = MODE == 0 ? s : b[row] - s;
}

// ------------------------------------------------------------- assembly ---
//
// The assembled operator A_hat = M K(E) M + I_c as a CSR matrix with one row per DOF
// (the reference's construct(), kernels/_cuda.py:407-420, through matmat_*_wcon
// kernels :837-996): row (i, k) holds the columns (j, d) of every node j that shares an
// element with node i, nodes ascending, components ascending.  The node neighbour
// lists come from the star: the smallest not yet listed node id among the star's
// element nodes, repeatedly (no per-thread sort buffer).

// next node id > last among the nodes of node i's star elements (INT32_MAX: none)
__device__ __forceinline__ int unbr_next(const et_umesh& m, int64_t i, int last) {
  int best = INT32_MAX;
  for (int p = m.star_ptr[i]; p < m.star_ptr[i + 1]; ++p) {
    const int e = m.el_of[m.star[p]];
    for (int f = m.elem_ptr[e]; f < m.elem_ptr[e + 1]; ++f) {
      const int v = m.elem_nodes[f];
      if (v > last && v < best) best = v;
    }
  }
  return best;
}

__global__ void unbr_count_kernel(const et_umesh m, int* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n_nodes) return;
  int c = 0, last = -1;
  for (;;) {
    const int v = unbr_next(m, i, last);
    if (v == INT32_MAX) break;
    ++c;
    last = v;
  }
  cnt[i] = c;
}

// node_off[i]: first CSR entry of node i's rows (dof rows of cnt[i] * dof entries each)
__global__ void unbr_fill_kernel(const et_umesh m, const int* __restrict__ cnt, const int64_t* __restrict__ node_off,
                                 int* __restrict__ row_ptr, int* __restrict__ cols) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m.n_nodes) return;
  const int dof = m.dof, c = cnt[i];
  const int64_t base = node_off[i];
  for (int k = 0; k < dof; ++k) row_ptr[i * dof + k] = (int)(base + (int64_t)k * c * dof);
  if (i == m.n_nodes - 1) row_ptr[m.n_nodes * dof] = (int)(base + (int64_t)dof * c * dof);
  int last = -1;
  for (int j = 0; j < c; ++j) {
    const int v = unbr_next(m, i, last);
    last = v;
    for (int k = 0; k < dof; ++k)
      for (int d = 0; d < dof; ++d) cols[base + (int64_t)k * c * dof + j * dof + d] = v * dof + d;
  }
}

// values of row (i, k): star order, element row K_e[(a, k), :] scaled by E_e, each column
// placed by a binary search of its node in the row's node list
template <int DOF>
__global__ __launch_bounds__(256) void uassemble_kernel(const et_umesh m, const double* __restrict__ E,
                                                        const uint8_t* __restrict__ cons,
                                                        const int* __restrict__ row_ptr,
                                                        const int* __restrict__ cols, double* __restrict__ vals) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= m.n_nodes * DOF) return;
  const int64_t i = row / DOF;
  const int k = (int)(row - i * DOF);
  const int a0 = row_ptr[row], a1 = row_ptr[row + 1];
  const int c = (a1 - a0) / DOF;
  for (int p = a0; p < a1; ++p) vals[p] = 0.0;
  if (cons != nullptr && cons[row]) {
    for (int p = a0; p < a1; ++p)
      if (cols[p] == row) vals[p] = 1.0;
    return;
  }
  for (int p = m.star_ptr[i]; p < m.star_ptr[i + 1]; ++p) {
    const int f = m.star[p];
    const int e = m.el_of[f];
    const int st = m.elem_ptr[e], size = m.elem_ptr[e + 1] - st;
    const int nc = size * DOF;
    const double* Kr = m.K + m.K_ptr[e] + (int64_t)((f - st) * DOF + k) * nc;
    const double w = E[e];
    for (int b = 0; b < size; ++b) {
      const int v = m.elem_nodes[st + b];
      int lo = 0, hi = c - 1;
      while (lo < hi) {  // the node's slot among the row's c nodes
        const int mid = (lo + hi) >> 1;
        if (cols[a0 + mid * DOF] / DOF < v)
          lo = mid + 1;
        else
          hi = mid;
      }
#pragma unroll
      for (int d = 0; d < DOF; ++d) {
        const int64_t col = (int64_t)v * DOF + d;
        if (cons == nullptr || !cons[col]) vals[a0 + lo * DOF + d] += Kr[b * DOF + d] * w;
      }
    }
  }
}

// ------------------------------------------------------------------ host ---

static int umesh_check(const et_umesh* m) {
  ET_REQUIRE(m != nullptr, "unstructured mesh: null descriptor");
  ET_REQUIRE(m->dof == 2 || m->dof == 3, "unstructured mesh: dof must be 2 or 3");
  ET_REQUIRE(m->n_nodes > 0 && m->n_el > 0, "unstructured mesh: empty mesh");
  ET_REQUIRE(m->n_nodes * m->dof < (int64_t)INT32_MAX, "unstructured mesh: too many DOFs for 32-bit indices");
  ET_REQUIRE(m->elem_ptr && m->elem_nodes && m->K && m->K_ptr && m->star_ptr && m->star && m->el_of,
             "unstructured mesh: null array in the descriptor");
  return ET_OK;
}

int umatvec(const et_umesh* m, int mode, const double* E, const uint8_t* cons, const double* u, const double* b,
            double* y, hipStream_t s) {
  ET_TRY(umesh_check(m));
  ET_REQUIRE(E && u && y && (mode == 0 || b), "unstructured matvec: null pointer");
  ET_REQUIRE(y != u, "unstructured matvec: y must not alias u");
  const unsigned grid = grid_for(m->n_nodes * m->dof);
  if (m->dof == 3) {
    if (mode == 0)
      umatvec_kernel<3, 0><<<grid, kBlock, 0, s>>>(*m, E, cons, u, b, y);
    else
      umatvec_kernel<3, 1><<<grid, kBlock, 0, s>>>(*m, E, cons, u, b, y);
  } else {
    if (mode == 0)
      umatvec_kernel<2, 0><<<grid, kBlock, 0, s>>>(*m, E, cons, u, b, y);
    else
      umatvec_kernel<2, 1><<<grid, kBlock, 0, s>>>(*m, E, cons, u, b, y);
  }
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int udiagonal(const et_umesh* m, const double* E, const uint8_t* cons, double* diag, hipStream_t s) {
  ET_TRY(umesh_check(m));
  ET_REQUIRE(E && diag, "unstructured diagonal: null pointer");
  const unsigned grid = grid_for(m->n_nodes * m->dof);
  if (m->dof == 3)
    udiag_kernel<3><<<grid, kBlock, 0, s>>>(*m, E, cons, diag);
  else
    udiag_kernel<2><<<grid, kBlock, 0, s>>>(*m, E, cons, diag);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int uenergy(const et_umesh* m, const double* U, const double* xt, double penal, double vmin, int mode, double* out,
            hipStream_t s) {
  ET_TRY(umesh_check(m));
  ET_REQUIRE(U && out && (mode == 0 || xt), "unstructured energy: null pointer");
  const unsigned grid = grid_for(m->n_el);
#define ET_UEN(D, M) uenergy_kernel<D, M><<<grid, kBlock, 0, s>>>(*m, U, xt, penal, vmin, out)
  if (m->dof == 3) {
    if (mode == 0) ET_UEN(3, 0); else ET_UEN(3, 1);
  } else {
    if (mode == 0) ET_UEN(2, 0); else ET_UEN(2, 1);
  }
#undef ET_UEN
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int ufea(const et_umesh* m, int nb, const double* B, const int64_t* B_ptr, const double* D, const int64_t* D_ptr,
         const double* E, const double* U, double* strain, double* stress, double* energy, double* vm,
         hipStream_t s) {
  ET_TRY(umesh_check(m));
  ET_REQUIRE(nb == (m->dof == 3 ? 6 : 3), "unstructured FEA: strain components must be 3 (2-D) or 6 (3-D)");
  ET_REQUIRE(B && B_ptr && D && D_ptr && E && U && strain && stress && energy, "unstructured FEA: null pointer");
  const unsigned grid = grid_for(m->n_el);
  if (m->dof == 3)
    ufea_kernel<3><<<grid, kBlock, 0, s>>>(*m, nb, B, B_ptr, D, D_ptr, E, U, strain, stress, energy, vm);
  else
    ufea_kernel<2><<<grid, kBlock, 0, s>>>(*m, nb, B, B_ptr, D, D_ptr, E, U, strain, stress, energy, vm);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int csr_spmv(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals, const double* x, double* y,
             hipStream_t s, int mode, const double* b) {
  ET_REQUIRE(n_rows > 0 && row_ptr && cols && vals && x && y && (mode == 0 || b), "CSR product: bad arguments");
  ET_REQUIRE(x != y, "CSR product: y must not alias x");
  ET_REQUIRE(n_rows < (int64_t)INT32_MAX, "CSR product: too many rows");
  if (mode == 0)
    csr_spmv_kernel<0><<<grid_for(n_rows * 8), kBlock, 0, s>>>(n_rows, row_ptr, cols, vals, x, b, y);
  else
    csr_spmv_kernel<1><<<grid_for(n_rows * 8), kBlock, 0, s>>>(n_rows, row_ptr, cols, vals, x, b, y);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int umesh_pattern(const et_umesh* m, int* row_ptr, int* cols, int64_t cap, int64_t* nnz, hipStream_t s) {
  ET_TRY(umesh_check(m));
  ET_REQUIRE(row_ptr && nnz, "assembly pattern: null pointer");
  const int64_t n = m->n_nodes;
  int* cnt = nullptr;
  int64_t* off = nullptr;
  ET_CHECK_HIP(hipMalloc(&cnt, (size_t)n * sizeof(int)));
  if (hipMalloc(&off, (size_t)(n + 1) * sizeof(int64_t)) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(cnt);
    set_error("assembly pattern: device allocation failed");
    return ET_EHIP;
  }
  auto run = [&]() -> int {
    unbr_count_kernel<<<grid_for(n), kBlock, 0, s>>>(*m, cnt);
    ET_CHECK_LAUNCH();
    std::vector<int> hc((size_t)n);
    ET_CHECK_HIP(hipMemcpyAsync(hc.data(), cnt, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> ho((size_t)n + 1, 0);
    const int64_t dd = (int64_t)m->dof * m->dof;
    for (int64_t i = 0; i < n; ++i) ho[i + 1] = ho[i] + hc[i] * dd;
    *nnz = ho[n];
    ET_REQUIRE(ho[n] < (int64_t)INT32_MAX, "assembly pattern: more than 2^31 entries");
    ET_CHECK_HIP(hipMemcpyAsync(off, ho.data(), (size_t)(n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    if (cols != nullptr && cap >= ho[n]) {
      unbr_fill_kernel<<<grid_for(n), kBlock, 0, s>>>(*m, cnt, off, row_ptr, cols);
      ET_CHECK_LAUNCH();
    }
    ET_CHECK_HIP(hipStreamSynchronize(s));
    return ET_OK;
  };
  const int st = run();
  (void)hipFree(cnt);
  (void)hipFree(off);
  return st;
}

int uassemble(const et_umesh* m, const double* E, const uint8_t* cons, const int* row_ptr, const int* cols,
              double* vals, hipStream_t s) {
  ET_TRY(umesh_check(m));
  ET_REQUIRE(E && row_ptr && cols && vals, "assembly: null pointer");
  const unsigned grid = grid_for(m->n_nodes * m->dof);
  if (m->dof == 3)
    uassemble_kernel<3><<<grid, kBlock, 0, s>>>(*m, E, cons, row_ptr, cols, vals);
  else
    uassemble_kernel<2><<<grid, kBlock, 0, s>>>(*m, E, cons, row_ptr, cols, vals);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

KrylovOp csr_op(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals) {
  KrylovOp op;
  op.n = n_rows;
  op.apply = [=](int mode, const double* x, const double* b, double* y, hipStream_t s) {
    return csr_spmv(n_rows, row_ptr, cols, vals, x, y, s, mode, b);
  };
  return op;
}

KrylovOp unstructured_op(const et_umesh* m, const double* E, const uint8_t* cons) {
  KrylovOp op;
  op.n = m->n_nodes * m->dof;
  const et_umesh mm = *m;
  op.apply = [mm, E, cons](int mode, const double* x, const double* b, double* y, hipStream_t s) {
    return umatvec(&mm, mode, E, cons, x, b, y, s);
  };
  return op;
}

}  // namespace et
