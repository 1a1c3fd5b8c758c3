// Unpreconditioned Krylov solvers on a matrix-free operator (the structured
// grid operator or the general-mesh operator, KrylovOp): conjugate gradients
// and restarted GMRES, the iterations of CuCG / CuGMRES (solvers/_gpu.py:11-107),
// which hand the stiffness operator to cupyx.scipy.sparse.linalg.cg / gmres.
//
// CG (cupyx cg semantics): atol = tol |b|; rho = r.r, beta = rho / rho_old,
// p = r + beta p, q = A p, alpha = rho / p.q, x += alpha p, r -= alpha q,
// stop when |r| <= atol.  Every iteration is four launches (p update, operator,
// x/r update with the p.q reduction finalised in-kernel, convergence check);
// the state lives on the device and every kernel returns at once after the
// stop, so the host enqueues chunks of iterations and reads the state once per
// chunk.
//
// GMRES(m) (cupyx gmres semantics): each cycle forms r = b - A x, stops if
// |r| <= atol, runs m Arnoldi steps with one classical Gram-Schmidt pass per
// step (h = V^T u, u -= V h, H[j+1, j] = |u|) and then solves the (m+1) x m
// least-squares problem on the host (Givens rotations) and updates x += V y.
// The convergence test is once per cycle, as in the reference.
//
// All reductions are two-stage with a fixed block count, so results are
// reproducible run to run.  The caller provides the workspace (krylov_work_size).
#include <algorithm>
#include <cmath>
#include <vector>

#include "blas.hpp"
#include "common.hpp"
#include "stiffness.hpp"

namespace et {

struct CgState {
  double rho, rho_old, atol2, rr, pq;
  int stop, iters, first;
};

constexpr int kCgChunk = 8;

__global__ void cg_init_kernel(const double* __restrict__ part_bb, const double* __restrict__ part_rr, int nb,
                               double tol, CgState* st) {
  double bb = 0.0, rr = 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) {
    bb += part_bb[t];
    rr += part_rr[t];
  }
  bb = block_sum256(bb);
  rr = block_sum256(rr);
  if (threadIdx.x == 0) {
    const double atol = tol * sqrt(bb);
    st->atol2 = atol * atol;
    st->rr = rr;
    st->rho = rr;
    st->rho_old = 0.0;
    st->iters = 0;
    st->first = 1;
    // |r| <= atol before any step stops at once (|b| = 0 is handled by the host)
    st->stop = (bb == 0.0 || rr <= st->atol2) ? 1 : 0;
  }
}

// p = r (first iteration) or p = r + (rho / rho_old) p
__global__ __launch_bounds__(256) void cg_p_kernel(int64_t n, const CgState* __restrict__ st,
                                                   const double* __restrict__ r, double* __restrict__ p) {
  if (st->stop) return;
  const bool first = st->first != 0;
  const double beta = first ? 0.0 : st->rho / st->rho_old;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = first ? r[t] : r[t] + beta * p[t];
}

__global__ __launch_bounds__(256) void cg_pq_kernel(int64_t n, const CgState* __restrict__ st,
                                                    const double* __restrict__ p, const double* __restrict__ q,
                                                    double* __restrict__ partial) {
  if (st->stop) return;
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    s += p[t] * q[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// alpha = rho / p.q (every block sums the same partials in the same order);
// x += alpha p, r -= alpha q, partial sums of r.r
__global__ __launch_bounds__(256) void cg_xr_kernel(int64_t n, const CgState* __restrict__ st,
                                                    const double* __restrict__ part_pq, int nb,
                                                    const double* __restrict__ p, const double* __restrict__ q,
                                                    double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ part_rr) {
  if (st->stop) return;
  __shared__ double s_alpha;
  double pq = 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) pq += part_pq[t];
  pq = block_sum256(pq);
  if (threadIdx.x == 0) s_alpha = st->rho / pq;
  __syncthreads();
  const double alpha = s_alpha;
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    x[t] += alpha * p[t];
    const double rv = r[t] - alpha * q[t];
    r[t] = rv;
    s += rv * rv;
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) part_rr[blockIdx.x] = s;
}

__global__ void cg_check_kernel(const double* __restrict__ part_rr, int nb, int maxiter, CgState* st) {
  if (st->stop) return;
  double rr = 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) rr += part_rr[t];
  rr = block_sum256(rr);
  if (threadIdx.x == 0) {
    st->iters += 1;
    st->first = 0;
    st->rr = rr;
    st->rho_old = st->rho;
    st->rho = rr;
    if (rr <= st->atol2 || st->iters >= maxiter || !(rr == rr)) st->stop = 1;
  }
}

namespace {
inline int64_t up256(int64_t bytes) { return (bytes + 255) / 256 * 256; }
}  // namespace

int64_t krylov_work_size(int64_t n, int restart) {
  const int64_t vec = up256(n * (int64_t)sizeof(double));
  const int64_t red = up256(kRedBlocks * (int64_t)sizeof(double));
  if (restart <= 0)  // CG: r, p, q, two reduction buffers, state + scalars
    return 3 * vec + 2 * red + up256(sizeof(CgState)) + up256(4 * sizeof(double));
  // GMRES: V (restart columns), u, r, reductions, H and scalars
  return up256((int64_t)restart * n * (int64_t)sizeof(double)) + 2 * vec +
         up256((int64_t)(restart + 2) * kRedBlocks * (int64_t)sizeof(double)) +
         up256((int64_t)(restart + 2) * (restart + 1) * (int64_t)sizeof(double));
}

namespace {

// carve the caller's workspace
struct Carve {
  char* p;
  template <typename T>
  T* take(int64_t count) {
    T* out = reinterpret_cast<T*>(p);
    p += up256(count * (int64_t)sizeof(T));
    return out;
  }
};

int residual_norm(const KrylovOp& op, const double* b, const double* x, double* r, double* partial, double* out,
                  hipStream_t s) {
  ET_TRY(op.apply(1, x, b, r, s));
  return dot(op.n, r, r, partial, out, s);
}

// the structured operator (fine-grid kernels when `fine`, a copy of which the
// operator keeps) as a Krylov operator
KrylovOp structured_op(const Grid& g, const double* K0, const double* E, const uint8_t* cons, const WhtTable* fine) {
  KrylovOp op;
  op.n = g.n_nodes() * g.dim;
  const bool has = fine != nullptr;
  WhtTable tab{};
  if (has) tab = *fine;
  op.apply = [=](int mode, const double* x, const double* b, double* y, hipStream_t s) {
    return level0_op<double>(mode, g, K0, E, cons, x, b, nullptr, nullptr, 1.0, nullptr, y, s, has ? &tab : nullptr);
  };
  return op;
}

}  // namespace

int cg_solve(const Grid& g, const double* K0, const double* E, const uint8_t* cons, const WhtTable* fine, const double* b,
             double* x, double tol, int maxiter, void* work, int64_t work_bytes, int* iters, double* residual,
             hipStream_t s) {
  return cg_solve_op(structured_op(g, K0, E, cons, fine), b, x, tol, maxiter, work, work_bytes, iters, residual,
                     s);
}

int cg_solve_op(const KrylovOp& op, const double* b, double* x, double tol, int maxiter, void* work,
                int64_t work_bytes, int* iters, double* residual, hipStream_t s) {
  const int64_t n = op.n;
  ET_REQUIRE(maxiter >= 1, "cg: maxiter must be >= 1");
  ET_REQUIRE(work_bytes >= krylov_work_size(n, 0), "cg: work buffer too small");
  Carve c{static_cast<char*>(work)};
  double* r = c.take<double>(n);
  double* p = c.take<double>(n);
  double* q = c.take<double>(n);
  double* part_a = c.take<double>(kRedBlocks);
  double* part_b = c.take<double>(kRedBlocks);
  CgState* st = c.take<CgState>(1);
  double* scal = c.take<double>(4);  // [0] |b|^2, [1] |b - A x|^2 after the loop
  {
    // |b| = 0: x = b = 0 at once, as cupyx returns
    ET_TRY(dot(n, b, b, part_a, scal, s));
    double bb0 = 0.0;
    ET_CHECK_HIP(hipMemcpyAsync(&bb0, scal, sizeof(double), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    if (bb0 == 0.0) {
      ET_CHECK_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
      if (iters) *iters = 0;
      if (residual) *residual = 0.0;
      return ET_OK;
    }
  }
  // r = b - A x, |b|^2, |r|^2
  ET_TRY(op.apply(1, x, b, r, s));
  ET_TRY(dot_partials(n, b, b, part_a, s));
  ET_TRY(dot_partials(n, r, r, part_b, s));
  cg_init_kernel<<<1, 256, 0, s>>>(part_a, part_b, kRedBlocks, tol, st);
  ET_CHECK_LAUNCH();
  CgState h{};
  for (int done = 0; done < maxiter;) {
    const int chunk = std::min(kCgChunk, maxiter - done);
    for (int k = 0; k < chunk; ++k) {
      cg_p_kernel<<<kRedBlocks, 256, 0, s>>>(n, st, r, p);
      ET_TRY(op.apply(0, p, nullptr, q, s));
      cg_pq_kernel<<<kRedBlocks, 256, 0, s>>>(n, st, p, q, part_a);
      cg_xr_kernel<<<kRedBlocks, 256, 0, s>>>(n, st, part_a, kRedBlocks, p, q, x, r, part_b);
      cg_check_kernel<<<1, 256, 0, s>>>(part_b, kRedBlocks, maxiter, st);
    }
    ET_CHECK_LAUNCH();
    done += chunk;
    ET_CHECK_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    if (h.stop) break;
  }
  ET_CHECK_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
  // the reported residual is the true one, |b - A x| / |b| (solvers/_gpu.py:43)
  ET_TRY(dot(n, b, b, part_a, scal, s));
  ET_TRY(residual_norm(op, b, x, r, part_b, scal + 1, s));
  double tail[2];
  ET_CHECK_HIP(hipMemcpyAsync(tail, scal, sizeof(tail), hipMemcpyDeviceToHost, s));
  ET_CHECK_HIP(hipStreamSynchronize(s));
  const double bb = tail[0], rr = tail[1];
  if (iters) *iters = h.iters;
  if (residual) *residual = bb > 0.0 ? std::sqrt(rr / bb) : 0.0;
  if (!(rr == rr)) {
    set_error("cg: non-finite residual");
    return ET_ENONFINITE;
  }
  if (!(h.rr <= h.atol2) && bb > 0.0) {
    set_error("cg: maxiter reached without meeting the tolerance");
    return ET_ENOCONV;
  }
  return ET_OK;
}

// ------------------------------------------------------------------ GMRES ---

// partial[k * nb + block] = sum over the block's slice of V[k] . u, k = blockIdx.y
__global__ __launch_bounds__(256) void gm_multidot_kernel(int64_t n, const double* __restrict__ V,
                                                          const double* __restrict__ u,
                                                          double* __restrict__ partial, int nb) {
  const double* v = V + (int64_t)blockIdx.y * n;
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    s += v[t] * u[t];
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * nb + blockIdx.x] = s;
}

// one block per k: out[k * stride] = sum of partial[k * nb ...]
__global__ void gm_finalize_kernel(const double* __restrict__ partial, int nb, double* __restrict__ out,
                                   int stride, int sqrt_it) {
  double s = 0.0;
  for (int t = threadIdx.x; t < nb; t += blockDim.x) s += partial[(int64_t)blockIdx.x * nb + t];
  s = block_sum256(s);
  if (threadIdx.x == 0) out[blockIdx.x * stride] = sqrt_it ? sqrt(s) : s;
}

// u -= sum_k V[k] h[k], k < kk
__global__ __launch_bounds__(256) void gm_orth_kernel(int64_t n, const double* __restrict__ V,
                                                      const double* __restrict__ h, int kk,
                                                      double* __restrict__ u) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < kk; ++k) s += V[(int64_t)k * n + t] * h[k];
    u[t] -= s;
  }
}

// v = u / *nrm (a zero norm, an exact breakdown, leaves v = 0)
__global__ __launch_bounds__(256) void gm_scale_kernel(int64_t n, const double* __restrict__ u,
                                                       const double* __restrict__ nrm, double* __restrict__ v) {
  const double d = *nrm;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    v[t] = d != 0.0 ? u[t] / d : 0.0;
}

// x += sum_k V[k] y[k]
__global__ __launch_bounds__(256) void gm_update_kernel(int64_t n, const double* __restrict__ V,
                                                        const double* __restrict__ y, int m,
                                                        double* __restrict__ x) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int k = 0; k < m; ++k) s += V[(int64_t)k * n + t] * y[k];
    x[t] += s;
  }
}

// least squares min |beta e1 - H y| for the (m+1) x m upper Hessenberg H
// (column-major, leading dimension m+1) by Givens rotations
static void hessenberg_lstsq(int m, std::vector<double> H, double beta, double* y) {
  const int ld = m + 1;
  std::vector<double> g(m + 1, 0.0);
  g[0] = beta;
  for (int j = 0; j < m; ++j) {
    const double a = H[j + j * ld], b = H[j + 1 + j * ld];
    const double r = std::hypot(a, b);
    const double c = r != 0.0 ? a / r : 1.0, sn = r != 0.0 ? b / r : 0.0;
    for (int k = j; k < m; ++k) {
      const double h1 = H[j + k * ld], h2 = H[j + 1 + k * ld];
      H[j + k * ld] = c * h1 + sn * h2;
      H[j + 1 + k * ld] = -sn * h1 + c * h2;
    }
    const double g1 = g[j], g2 = g[j + 1];
    g[j] = c * g1 + sn * g2;
    g[j + 1] = -sn * g1 + c * g2;
  }
  for (int j = m - 1; j >= 0; --j) {
    double s = g[j];
    for (int k = j + 1; k < m; ++k) s -= H[j + k * ld] * y[k];
    const double d = H[j + j * ld];
    y[j] = d != 0.0 ? s / d : 0.0;
  }
}

int gmres_solve(const Grid& g, const double* K0, const double* E, const uint8_t* cons, const WhtTable* fine,
                const double* b, double* x, double tol, int restart, int maxiter, void* work, int64_t work_bytes,
                int* iters, double* residual, hipStream_t s) {
  return gmres_solve_op(structured_op(g, K0, E, cons, fine), b, x, tol, restart, maxiter, work, work_bytes, iters,
                        residual, s);
}

int gmres_solve_op(const KrylovOp& op, const double* b, double* x, double tol, int restart, int maxiter,
                   void* work, int64_t work_bytes, int* iters, double* residual, hipStream_t s) {
  const int64_t n = op.n;
  ET_REQUIRE(restart >= 1 && restart <= 256, "gmres: restart must be in [1, 256]");
  ET_REQUIRE(maxiter >= 1, "gmres: maxiter must be >= 1");
  ET_REQUIRE(work_bytes >= krylov_work_size(n, restart), "gmres: work buffer too small");
  const int m = (int)std::min<int64_t>(restart, n);
  Carve c{static_cast<char*>(work)};
  double* V = c.take<double>((int64_t)restart * n);
  double* u = c.take<double>(n);
  double* r = c.take<double>(n);
  double* partial = c.take<double>((int64_t)(restart + 2) * kRedBlocks);
  double* Hd = c.take<double>((int64_t)(restart + 2) * (restart + 1));  // H (m+1 x m), then scalars
  double* sc = Hd + (int64_t)(m + 1) * m;                            // [0] |b|, [1] |r|, [2..] y
  const int nb = kRedBlocks;
  ET_TRY(dot(n, b, b, partial, sc, s));
  double bnorm2 = 0.0;
  ET_CHECK_HIP(hipMemcpyAsync(&bnorm2, sc, sizeof(double), hipMemcpyDeviceToHost, s));
  ET_CHECK_HIP(hipStreamSynchronize(s));
  if (bnorm2 == 0.0) {
    // |b| = 0: x = b = 0 at once, as cupyx returns
    ET_CHECK_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    if (iters) *iters = 0;
    if (residual) *residual = 0.0;
    return ET_OK;
  }
  const double bnorm = std::sqrt(bnorm2);
  const double atol = tol * bnorm;
  std::vector<double> H((size_t)(m + 1) * m), y(m);
  int it = 0;
  double rnorm = 0.0;
  const unsigned grid = (unsigned)std::min<int64_t>(kRedBlocks, (n + 255) / 256);
  while (true) {
    ET_TRY(op.apply(1, x, b, r, s));
    gm_multidot_kernel<<<dim3(nb, 1), 256, 0, s>>>(n, r, r, partial, nb);
    gm_finalize_kernel<<<1, 256, 0, s>>>(partial, nb, sc + 1, 1, 1);
    ET_CHECK_LAUNCH();
    ET_CHECK_HIP(hipMemcpyAsync(&rnorm, sc + 1, sizeof(double), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    if (!(rnorm == rnorm)) {
      set_error("gmres: non-finite residual");
      return ET_ENONFINITE;
    }
    if (bnorm == 0.0 || rnorm <= atol || it >= maxiter) break;
    gm_scale_kernel<<<grid, 256, 0, s>>>(n, r, sc + 1, V);
    for (int j = 0; j < m; ++j) {
      ET_TRY(op.apply(0, V + (int64_t)j * n, nullptr, u, s));
      // h = V[:, :j+1]^T u into column j of H, then u -= V h
      gm_multidot_kernel<<<dim3(nb, j + 1), 256, 0, s>>>(n, V, u, partial, nb);
      gm_finalize_kernel<<<j + 1, 256, 0, s>>>(partial, nb, Hd + (int64_t)j * (m + 1), 1, 0);
      gm_orth_kernel<<<grid, 256, 0, s>>>(n, V, Hd + (int64_t)j * (m + 1), j + 1, u);
      gm_multidot_kernel<<<dim3(nb, 1), 256, 0, s>>>(n, u, u, partial, nb);
      gm_finalize_kernel<<<1, 256, 0, s>>>(partial, nb, Hd + (int64_t)j * (m + 1) + j + 1, 1, 1);
      if (j + 1 < m) gm_scale_kernel<<<grid, 256, 0, s>>>(n, u, Hd + (int64_t)j * (m + 1) + j + 1, V + (int64_t)(j + 1) * n);
      ET_CHECK_LAUNCH();
    }
    ET_CHECK_HIP(hipMemcpyAsync(H.data(), Hd, H.size() * sizeof(double), hipMemcpyDeviceToHost, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
    hessenberg_lstsq(m, H, rnorm, y.data());
    ET_CHECK_HIP(hipMemcpyAsync(sc + 2, y.data(), m * sizeof(double), hipMemcpyHostToDevice, s));
    gm_update_kernel<<<grid, 256, 0, s>>>(n, V, sc + 2, m, x);
    ET_CHECK_LAUNCH();
    // the host copy of y must stay valid until the transfer has run
    ET_CHECK_HIP(hipStreamSynchronize(s));
    it += m;
  }
  if (iters) *iters = it;
  if (residual) *residual = bnorm > 0.0 ? rnorm / bnorm : 0.0;
  if (!(rnorm <= atol) && bnorm > 0.0) {
    set_error("gmres: maxiter reached without meeting the tolerance");
    return ET_ENOCONV;
  }
  return ET_OK;
}

}  // namespace et
