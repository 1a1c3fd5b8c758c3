// GMG-preconditioned conjugate gradients for A_hat(E) u = f on a structured
// grid: the behaviour of CuMultiGrid.solve (solvers/_multigrid_cuda.py:188-242)
// with the hierarchy built on the device.
//
//   level 0   matrix-free operator (fused Jacobi / residual kernels)
//   level l   Galerkin block stencil R S_{l-1} R^T, built every solve
//   level L   dense FP64 inverse (CoarseSolver)
//
// The PCG loop keeps alpha, beta, omega and the dot products in device
// memory; the host reads |r|^2 once per iteration for the convergence test.
// precision 1 runs the V-cycle in FP32 (operators cast from the FP64
// hierarchy) inside the FP64 outer loop.
#include <cmath>
#include <vector>

#include "blas.hpp"
#include "coarse.hpp"
#include "common.hpp"
#include "galerkin_elem.hpp"
#include "solver.hpp"
#include "stiffness.hpp"

namespace et {

// ------------------------------------------------------------ kernels ----

enum OpMode { OP_Y = 0, OP_RES = 1, OP_JAC = 2, OP_JAC_ADD = 3 };

template <typename T, int DIM, int MODE>
__device__ __forceinline__ void op_epilogue(int64_t n, const T (&ax)[DIM],
                                            const T* __restrict__ x, const T* __restrict__ b,
                                            const T* __restrict__ Dinv, T w,
                                            const T* __restrict__ xadd, T* __restrict__ y) {
#pragma unroll
  for (int r = 0; r < DIM; ++r) {
    const int64_t d = n * DIM + r;
    T v;
    if (MODE == OP_Y)
      v = ax[r];
    else if (MODE == OP_RES)
      v = b[d] - ax[r];
    else
      v = x[d] + w * Dinv[d] * (b[d] - ax[r]);
    if (MODE == OP_JAC_ADD) v += xadd[d];
    y[d] = v;
  }
}

// Level-0 operator, matrix free.  omega is read from device memory.
template <typename T, int DIM, int MODE>
__global__ __launch_bounds__(256) void op0_kernel(const Grid g, const T* __restrict__ K0,
                                                  const T* __restrict__ E,
                                                  const uint8_t* __restrict__ cons,
                                                  const T* __restrict__ x, const T* __restrict__ b,
                                                  const T* __restrict__ Dinv,
                                                  const double* __restrict__ omega, double wscale,
                                                  const T* __restrict__ xadd,
                                                  T* __restrict__ y) {
  constexpr int NE = DIM == 3 ? 8 : 4;
  constexpr int ND = NE * DIM;
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes()) return;
  const int nz = g.nz(), nx = g.nx();
  const int k = DIM == 3 ? (int)(n % nz) : 0;
  const int64_t t = DIM == 3 ? n / nz : n;
  const int i = (int)(t % nx), j = (int)(t / nx);
  T acc[DIM];
#pragma unroll
  for (int r = 0; r < DIM; ++r) acc[r] = T(0);
#pragma unroll
  for (int a = 0; a < NE; ++a) {
    const int ex = i - (DIM == 3 ? hx(a) : qx(a));
    const int ey = j - (DIM == 3 ? hy(a) : qy(a));
    const int ez = k - (DIM == 3 ? hz(a) : 0);
    bool valid = ex >= 0 && ex < g.ex && ey >= 0 && ey < g.ey;
    if (DIM == 3) valid = valid && ez >= 0 && ez < g.ez;
    if (!valid) continue;
    const int64_t e = DIM == 3 ? ((int64_t)ex * g.ey + ey) * g.ez + ez : (int64_t)ex * g.ey + ey;
    const T Ee = E[e];
    T s[DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r) s[r] = T(0);
#pragma unroll
    for (int b2 = 0; b2 < NE; ++b2) {
      const int bx = ex + (DIM == 3 ? hx(b2) : qx(b2));
      const int by = ey + (DIM == 3 ? hy(b2) : qy(b2));
      const int bz = ez + (DIM == 3 ? hz(b2) : 0);
      const int64_t m = DIM == 3 ? ((int64_t)by * nx + bx) * nz + bz : (int64_t)by * nx + bx;
      T um[DIM];
#pragma unroll
      for (int c = 0; c < DIM; ++c) um[c] = cons[m * DIM + c] ? T(0) : x[m * DIM + c];
#pragma unroll
      for (int r = 0; r < DIM; ++r)
#pragma unroll
        for (int c = 0; c < DIM; ++c) s[r] += K0[(DIM * a + r) * ND + DIM * b2 + c] * um[c];
    }
#pragma unroll
    for (int r = 0; r < DIM; ++r) acc[r] += Ee * s[r];
  }
#pragma unroll
  for (int r = 0; r < DIM; ++r)
    if (cons[n * DIM + r]) acc[r] = x[n * DIM + r];
  const T w = MODE >= OP_JAC ? (T)(*omega * wscale) : T(0);
  op_epilogue<T, DIM, MODE>(n, acc, x, b, Dinv, w, xadd, y);
}

// Coarse level operator: block stencil.
template <typename T, int DIM, int MODE>
__global__ __launch_bounds__(256) void opS_kernel(const Grid g, const T* __restrict__ S,
                                                  const T* __restrict__ x, const T* __restrict__ b,
                                                  const T* __restrict__ Dinv,
                                                  const double* __restrict__ omega, double wscale,
                                                  const T* __restrict__ xadd,
                                                  T* __restrict__ y) {
  constexpr int NO = DIM == 3 ? 27 : 9;
  const int64_t N = g.n_nodes();
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int nz = g.nz(), nx = g.nx();
  const int k = DIM == 3 ? (int)(n % nz) : 0;
  const int64_t t = DIM == 3 ? n / nz : n;
  const int i = (int)(t % nx), j = (int)(t / nx);
  T acc[DIM];
#pragma unroll
  for (int r = 0; r < DIM; ++r) acc[r] = T(0);
#pragma unroll
  for (int o = 0; o < NO; ++o) {
    const int oi = DIM == 3 ? o / 9 - 1 : o / 3 - 1;
    const int oj = DIM == 3 ? (o / 3) % 3 - 1 : o % 3 - 1;
    const int ok = DIM == 3 ? o % 3 - 1 : 0;
    const int mi = i + oi, mj = j + oj, mk = k + ok;
    if (mi < 0 || mi >= nx || mj < 0 || mj >= g.ny()) continue;
    if (DIM == 3 && (mk < 0 || mk >= nz)) continue;
    const int64_t m = DIM == 3 ? ((int64_t)mj * nx + mi) * nz + mk : (int64_t)mj * nx + mi;
#pragma unroll
    for (int q = 0; q < DIM; ++q) {
      const T xv = x[m * DIM + q];
#pragma unroll
      for (int r = 0; r < DIM; ++r) acc[r] += S[((int64_t)(o * DIM + r) * DIM + q) * N + n] * xv;
    }
  }
  const T w = MODE >= OP_JAC ? (T)(*omega * wscale) : T(0);
  op_epilogue<T, DIM, MODE>(n, acc, x, b, Dinv, w, xadd, y);
}

// y = omega*wscale * Dinv * b  (first Jacobi sweep from a zero guess)
template <typename T>
__global__ void scale_kernel(int64_t n, const T* __restrict__ Dinv, const T* __restrict__ b,
                             const double* __restrict__ omega, double wscale, T* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const T w = (T)(*omega * wscale);
  y[t] = w * Dinv[t] * b[t];
}

template <typename TI, typename TO>
__global__ void convert_kernel(int64_t n, const TI* __restrict__ a, TO* __restrict__ b) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (TO)a[t];
}

template <typename T>
__global__ void axpy_kernel(int64_t n, const T* __restrict__ x, T* __restrict__ y) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] += x[t];
}

template <typename T>
__global__ void recip_kernel(int64_t n, const double* __restrict__ d, T* __restrict__ out) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (T)(1.0 / d[t]);
}

// inverse diagonal of a block stencil (centre block diagonal)
template <typename T, int DIM>
__global__ void stencil_dinv_kernel(int64_t N, const double* __restrict__ S, T* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  constexpr int oc = DIM == 3 ? 13 : 4;
#pragma unroll
  for (int r = 0; r < DIM; ++r) out[n * DIM + r] = (T)(1.0 / S[((int64_t)(oc * DIM + r) * DIM + r) * N + n]);
}

// PCG: x += alpha p; r -= alpha q; partial sums of r.r   (alpha = rho / pq)
__global__ __launch_bounds__(256) void pcg_update_xr_kernel(int64_t n, const double* __restrict__ sc,
                                                            const double* __restrict__ p,
                                                            const double* __restrict__ q,
                                                            double* __restrict__ x,
                                                            double* __restrict__ r,
                                                            double* __restrict__ partial) {
  const double alpha = sc[SC_RHO] / sc[SC_PQ];
  double s = 0.0;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    x[t] += alpha * p[t];
    const double rv = r[t] - alpha * q[t];
    r[t] = rv;
    s += rv * rv;
  }
  s = block_sum256(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// p = z + beta p with beta = rho_new / rho (Fletcher-Reeves); then rho <- rho_new
__global__ __launch_bounds__(256) void pcg_update_p_kernel(int64_t n, const double* __restrict__ sc,
                                                           const double* __restrict__ z,
                                                           double* __restrict__ p) {
  const double beta = sc[SC_RHO_NEW] / sc[SC_RHO];
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    p[t] = z[t] + beta * p[t];
}

__global__ void pcg_rotate_kernel(double* sc) { sc[SC_RHO] = sc[SC_RHO_NEW]; }

// omega = min + d/2 exp((-clip(R, tol, 0.1) + tol) * 500), R = |r| / |b|
__global__ void pcg_omega_kernel(double* sc, double tol, double wmin, double dw) {
  const double R = sqrt(sc[SC_RR]) / sqrt(sc[SC_BB]);
  const double Rc = fmin(fmax(R, tol), 1e-1);
  sc[SC_OMEGA] = wmin + dw / 2.0 * exp((-Rc + tol) * 500.0);
}

__global__ void set_scalar_kernel(double* sc, int slot, double v) { sc[slot] = v; }

// ------------------------------------------------------------ host side ----

static int op_kind(int level, int mode) {
  if (level > 0) return ET_PROF_COARSE_OPS;
  return mode == OP_RES ? ET_PROF_L0_RES : (mode == OP_JAC_ADD ? ET_PROF_L0_JACADD : ET_PROF_L0_JAC);
}

template <typename T>
static int op_launch(const MGLevel<T>& L, const MGShared& sh, int mode, const T* x, const T* b,
                     const double* omega, double wscale, const T* xadd, T* y, hipStream_t s) {
  const Grid& g = L.g;
  sh.prof->begin(op_kind(L.level, mode), s);
  if (L.level == 0 && sizeof(T) == 4 && sh.mfma) {
    const int st = op0_mfma(mode, g, (int64_t)g.ey * g.ez, g.ez, sh.kscale, sh.K0f, sh.Ef, sh.cons,
                            (const float*)x, (const float*)b, (const float*)L.Dinv, omega, wscale,
                            (const float*)xadd, (float*)y, s);
    sh.prof->end(s);
    return st;
  }
  if (L.level == 0 && sh.tiled) {
    const int st = op0_tiled<T>(mode, g, (int64_t)g.ey * g.ez, g.ez, sh.K0<T>(), sh.E<T>(), sh.cons,
                                x, b, L.Dinv, omega, wscale, xadd, y, s);
    sh.prof->end(s);
    return st;
  }
  const unsigned grid = grid_for(g.n_nodes());
#define ET_OP0(DIM, MODE)                                                                   \
  op0_kernel<T, DIM, MODE><<<grid, kBlock, 0, s>>>(g, sh.K0<T>(), sh.E<T>(), sh.cons, x, b, \
                                                   L.Dinv, omega, wscale, xadd, y)
#define ET_OPS(DIM, MODE) \
  opS_kernel<T, DIM, MODE><<<grid, kBlock, 0, s>>>(g, L.S, x, b, L.Dinv, omega, wscale, xadd, y)
#define ET_DISPATCH(M)                \
  do {                                \
    if (L.level == 0) {               \
      if (g.dim == 3)                 \
        ET_OP0(3, M);                 \
      else                            \
        ET_OP0(2, M);                 \
    } else {                          \
      if (g.dim == 3)                 \
        ET_OPS(3, M);                 \
      else                            \
        ET_OPS(2, M);                 \
    }                                 \
  } while (0)
  switch (mode) {
    case OP_Y: ET_DISPATCH(OP_Y); break;
    case OP_RES: ET_DISPATCH(OP_RES); break;
    case OP_JAC: ET_DISPATCH(OP_JAC); break;
    default: ET_DISPATCH(OP_JAC_ADD); break;
  }
#undef ET_DISPATCH
#undef ET_OPS
#undef ET_OP0
  sh.prof->end(s);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

template <typename T>
int MGHierarchy<T>::vcycle(int l, MGShared& sh, hipStream_t s) {
  MGLevel<T>& L = lv[l];
  if (l == n_level) {
    sh.prof->begin(ET_PROF_COARSE_SOLVE, s);
    const int st = coarse->solve<T, T>(L.b, L.x, s);
    sh.prof->end(s);
    return st;
  }
  const double ws = std::pow(1.06, l);
  const double* om = sh.scal + SC_OMEGA;
  const int64_t n = L.n;
  if (n_smooth == 0) {
    ET_CHECK_HIP(hipMemsetAsync(L.x, 0, n * sizeof(T), s));
  } else {
    sh.prof->begin(l == 0 ? ET_PROF_L0_SCALE : ET_PROF_COARSE_OPS, s);
    scale_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, L.Dinv, L.b, om, ws, L.x);
    sh.prof->end(s);
    ET_CHECK_LAUNCH();
    for (int it = 1; it < n_smooth; ++it) {
      ET_TRY(op_launch<T>(L, sh, OP_JAC, L.x, L.b, om, ws, nullptr, L.t, s));
      std::swap(L.x, L.t);
    }
  }
  const int passes = (cycle_w && ((w_level_mask >> l) & 1)) ? 2 : 1;
  for (int pass = 0; pass < passes; ++pass) {
    ET_TRY(op_launch<T>(L, sh, OP_RES, L.x, L.b, om, ws, nullptr, L.r, s));
    sh.prof->begin(ET_PROF_TRANSFER, s);
    ET_TRY(restrict_op<T>(L.g, dof, L.r, lv[l + 1].b, s));
    sh.prof->end(s);
    ET_TRY(vcycle(l + 1, sh, s));
    sh.prof->begin(ET_PROF_TRANSFER, s);
    ET_TRY(prolong_op<T>(L.g, dof, lv[l + 1].x, L.e, s));
    sh.prof->end(s);
    if (n_smooth == 0) {
      axpy_kernel<T><<<grid_for(n), kBlock, 0, s>>>(n, L.e, L.x);
      ET_CHECK_LAUNCH();
    }
    for (int it = 0; it < n_smooth; ++it) {
      if (it + 1 < n_smooth) {
        ET_TRY(op_launch<T>(L, sh, OP_JAC, L.e, L.r, om, ws, nullptr, L.t, s));
        std::swap(L.e, L.t);
      } else {
        ET_TRY(op_launch<T>(L, sh, OP_JAC_ADD, L.e, L.r, om, ws, L.x, L.t, s));
        std::swap(L.x, L.t);
      }
    }
  }
  return ET_OK;
}

template <typename T>
int MGHierarchy<T>::alloc(const Grid& g0, int dof_, int n_level_, hipStream_t s) {
  dof = dof_;
  n_level = n_level_;
  lv.resize(n_level + 1);
  Grid g = g0;
  for (int l = 0; l <= n_level; ++l) {
    MGLevel<T>& L = lv[l];
    L.level = l;
    L.g = g;
    L.n = g.n_nodes() * dof;
    const size_t vb = (size_t)L.n * sizeof(T);
    for (T** p : {&L.b, &L.x, &L.r, &L.e, &L.t, &L.Dinv}) ET_CHECK_HIP(hipMalloc(p, vb));
    if (l > 0 && l < n_level) {
      const int NO = g.dim == 3 ? 27 : 9;
      ET_CHECK_HIP(hipMalloc(&L.S, (size_t)NO * dof * dof * g.n_nodes() * sizeof(T)));
    }
    g = g.coarse();
  }
  (void)s;
  return ET_OK;
}

template <typename T>
void MGHierarchy<T>::release() {
  for (auto& L : lv)
    for (T* p : {L.b, L.x, L.r, L.e, L.t, L.Dinv, L.S})
      if (p) (void)hipFree(p);
  lv.clear();
}

template struct MGHierarchy<double>;
template struct MGHierarchy<float>;

// ------------------------------------------------------------ et_mg ----

int MGSolver::profile_read(double* ms, int64_t* launches) {
  ET_CHECK_HIP(hipStreamSynchronize(stream));
  for (int k = 0; k < ET_PROF_KINDS; ++k) {
    ms[k] = 0.0;
    launches[k] = 0;
  }
  for (const auto& r : prof.recs) {
    float t = 0.f;
    ET_CHECK_HIP(hipEventElapsedTime(&t, r.a, r.b));
    if (r.kind >= 0 && r.kind < ET_PROF_KINDS) {
      ms[r.kind] += t;
      launches[r.kind] += 1;
    }
  }
  return ET_OK;
}

int MGSolver::create(const et_mg_params& p, const double* K0_host, hipStream_t s) {
  prm = p;
  sh.prof = &prof;
  stream = s;
  ET_TRY(make_grid(p.dim, p.nel, &g0));
  dof = p.dim;
  ET_REQUIRE(p.n_level >= 0 && p.n_level <= 12, "n_level out of range");
  ET_REQUIRE(p.n_smooth >= 0 && p.n_smooth <= 64, "n_smooth out of range");
  ET_REQUIRE(p.precision == 0 || p.precision == 1, "precision must be 0 (fp64) or 1 (mixed)");
  {
    Grid g = g0;
    for (int l = 0; l < p.n_level; ++l) {
      ET_REQUIRE(g.ex >= 2 && g.ey >= 2 && (g.dim == 2 || g.ez >= 2),
                 "n_level too deep for this mesh (a level would have fewer than one element)");
      g = g.coarse();
    }
  }
  n_dof = g0.n_nodes() * dof;
  n_elem = g0.n_elem();
  sh.tiled = g0.dim == 3 && tiled_supported(g0, sizeof(double));
  sh.mfma = p.precision == 1 && get_mixed_mfma() && mfma_supported(g0);
  sh.kscale = k0_scale(K0_host, (p.dim == 3 ? 24 * 24 : 8 * 8));
  const int nd = (p.dim == 3 ? 8 : 4) * dof;
  ET_CHECK_HIP(hipMalloc(&sh.K0d, nd * nd * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&sh.K0f, nd * nd * sizeof(float)));
  ET_CHECK_HIP(hipMemcpyAsync(sh.K0d, K0_host, nd * nd * sizeof(double), hipMemcpyHostToDevice, s));
  {
    std::vector<float> kf(nd * nd);
    for (int t = 0; t < nd * nd; ++t) kf[t] = (float)K0_host[t];
    ET_CHECK_HIP(hipMemcpyAsync(sh.K0f, kf.data(), nd * nd * sizeof(float), hipMemcpyHostToDevice, s));
    ET_CHECK_HIP(hipStreamSynchronize(s));
  }
  ET_CHECK_HIP(hipMalloc(&sh.scal, SC_COUNT * sizeof(double)));
  ET_CHECK_HIP(hipMalloc(&partial, kRedBlocks * sizeof(double)));
  ET_CHECK_HIP(hipHostMalloc(&host_scal, SC_COUNT * sizeof(double)));
  for (double** v : {&r, &z, &p_, &q, &diag64}) ET_CHECK_HIP(hipMalloc(v, n_dof * sizeof(double)));
  if (p.precision == 1) ET_CHECK_HIP(hipMalloc(&sh.Ef, n_elem * sizeof(float)));
  // FP64 stencils: level 0 (fine operator) and every coarse level
  const int NO = p.dim == 3 ? 27 : 9;
  S64.assign(p.n_level + 1, nullptr);
  elem_galerkin = p.n_level > 0 && galerkin_elem_supported(g0);
  if (elem_galerkin) {
    const std::vector<double> M = galerkin_child_tables(p.dim, K0_host);
    ET_CHECK_HIP(hipMalloc(&Mt, M.size() * sizeof(double)));
    ET_CHECK_HIP(hipMemcpy(Mt, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice));
    ET_CHECK_HIP(hipMalloc(&eflag, (size_t)n_elem));
  }
  {
    Grid g = g0;
    for (int l = 0; l <= p.n_level; ++l) {
      const bool need = p.n_level > 0 && !(l == 0 && elem_galerkin);
      if (need) ET_CHECK_HIP(hipMalloc(&S64[l], (size_t)NO * dof * dof * g.n_nodes() * sizeof(double)));
      if (l == p.n_level) coarse_grid = g;
      g = g.coarse();
    }
  }
  if (p.precision == 0)
    ET_TRY(h64.alloc(g0, dof, p.n_level, s));
  else
    ET_TRY(h32.alloc(g0, dof, p.n_level, s));
  ET_TRY(coarse.init((int)(coarse_grid.n_nodes() * dof), s));
  h64.coarse = &coarse;
  h32.coarse = &coarse;
  for (auto* h : {&h64.n_smooth, &h32.n_smooth}) *h = p.n_smooth;
  h64.cycle_w = h32.cycle_w = p.cycle_w;
  h64.w_level_mask = h32.w_level_mask = p.w_level_mask;
  max_omega = p.omega;
  min_omega = std::fmax(p.omega / 2.0, p.min_omega);
  d_omega = max_omega - min_omega;
  ET_CHECK_HIP(hipEventCreate(&ev0));
  ET_CHECK_HIP(hipEventCreate(&ev1));
  ET_CHECK_HIP(hipEventCreate(&ev2));
  return ET_OK;
}

void MGSolver::destroy() {
  prof.release();
  h64.release();
  h32.release();
  coarse.release();
  for (double* v : S64)
    if (v) (void)hipFree(v);
  for (void* v : {(void*)sh.K0d, (void*)sh.K0f, (void*)sh.scal, (void*)partial, (void*)r, (void*)z,
                  (void*)p_, (void*)q, (void*)diag64, (void*)sh.Ef, (void*)Mt, (void*)eflag})
    if (v) (void)hipFree(v);
  if (host_scal) (void)hipHostFree(host_scal);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (ev2) (void)hipEventDestroy(ev2);
}

template <typename T>
int MGSolver::setup_levels(MGHierarchy<T>& h) {
  hipStream_t s = stream;
  const int L = prm.n_level;
  // level 0: inverse diagonal
  ET_TRY(diagonal<double>(g0, sh.K0d, sh.Ed, sh.cons, diag64, s));
  recip_kernel<T><<<grid_for(n_dof), kBlock, 0, s>>>(n_dof, diag64, h.lv[0].Dinv);
  ET_CHECK_LAUNCH();
  if (L == 0) {
    return ET_EINVAL;  // a direct solve of the fine grid is not supported
  }
  if (!elem_galerkin) ET_TRY(stencil_from_operator(g0, sh.K0d, sh.Ed, sh.cons, S64[0], s));
  Grid g = g0;
  for (int l = 1; l <= L; ++l) {
    if (l == 1 && elem_galerkin)
      ET_TRY(galerkin_from_elements(g0, sh.K0d, Mt, sh.Ed, sh.cons, eflag, S64[1], s));
    else
      ET_TRY(galerkin(g, dof, S64[l - 1], S64[l], s));
    g = g.coarse();
    MGLevel<T>& lev = h.lv[l];
    const int64_t N = g.n_nodes();
    if (g.dim == 3)
      stencil_dinv_kernel<T, 3><<<grid_for(N), kBlock, 0, s>>>(N, S64[l], lev.Dinv);
    else
      stencil_dinv_kernel<T, 2><<<grid_for(N), kBlock, 0, s>>>(N, S64[l], lev.Dinv);
    ET_CHECK_LAUNCH();
    if (l < L) {
      const int NO = g.dim == 3 ? 27 : 9;
      const int64_t cnt = (int64_t)NO * dof * dof * N;
      convert_kernel<double, T><<<grid_for(cnt), kBlock, 0, s>>>(cnt, S64[l], lev.S);
      ET_CHECK_LAUNCH();
    }
  }
  ET_TRY(coarse.factor(coarse_grid, dof, S64[L], s));
  return ET_OK;
}

template <typename T>
int MGSolver::precondition(MGHierarchy<T>& h) {
  // z = V(r): level-0 rhs is the PCG residual, result goes to z
  hipStream_t s = stream;
  MGLevel<T>& L0 = h.lv[0];
  prof.begin(ET_PROF_PCG_VEC, s);
  convert_kernel<double, T><<<grid_for(n_dof), kBlock, 0, s>>>(n_dof, r, L0.b);
  prof.end(s);
  ET_CHECK_LAUNCH();
  ET_TRY(h.vcycle(0, sh, s));
  prof.begin(ET_PROF_PCG_VEC, s);
  convert_kernel<T, double><<<grid_for(n_dof), kBlock, 0, s>>>(n_dof, L0.x, z);
  prof.end(s);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int MGSolver::apply_A64(const double* x, double* y, int mode, const double* b) {
  hipStream_t s = stream;
  const unsigned grid = grid_for(g0.n_nodes());
  prof.begin(ET_PROF_MATVEC64, s);
  if (sh.tiled) {
    const int st = op0_tiled<double>(mode, g0, (int64_t)g0.ey * g0.ez, g0.ez, sh.K0d, sh.Ed, sh.cons,
                                     x, b, nullptr, sh.scal, 1.0, nullptr, y, s);
    prof.end(s);
    return st;
  }
  if (g0.dim == 3) {
    if (mode == OP_Y)
      op0_kernel<double, 3, OP_Y><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
    else
      op0_kernel<double, 3, OP_RES><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
  } else {
    if (mode == OP_Y)
      op0_kernel<double, 2, OP_Y><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
    else
      op0_kernel<double, 2, OP_RES><<<grid, kBlock, 0, s>>>(g0, sh.K0d, sh.Ed, sh.cons, x, b, nullptr, sh.scal, 1.0, nullptr, y);
  }
  prof.end(s);
  ET_CHECK_LAUNCH();
  return ET_OK;
}

int MGSolver::read_scalars() {
  ET_CHECK_HIP(hipMemcpyAsync(host_scal, sh.scal, SC_COUNT * sizeof(double), hipMemcpyDeviceToHost, stream));
  ET_CHECK_HIP(hipStreamSynchronize(stream));
  return ET_OK;
}

int MGSolver::solve(const double* E, const uint8_t* cons, const double* b, double* x, double tol,
                    int maxiter, int* iters, double* residual) {
  ET_REQUIRE(E && cons && b && x, "et_mg_solve: null pointer");
  ET_REQUIRE(maxiter >= 1, "et_mg_solve: maxiter must be >= 1");
  ET_REQUIRE(prm.n_level >= 1, "et_mg_solve: n_level must be >= 1");
  hipStream_t s = stream;
  sh.Ed = E;
  sh.cons = cons;
  ET_CHECK_HIP(hipEventRecord(ev0, s));
  prof.begin(ET_PROF_SETUP, s);
  if (prm.precision == 1) {
    convert_kernel<double, float><<<grid_for(n_elem), kBlock, 0, s>>>(n_elem, E, sh.Ef);
    ET_CHECK_LAUNCH();
    ET_TRY(setup_levels(h32));
  } else {
    ET_TRY(setup_levels(h64));
  }
  prof.end(s);
  ET_TRY(coarse.check_info(s));
  ET_CHECK_HIP(hipEventRecord(ev1, s));
  // omega <- min_omega at the start of every solve (GPU semantics)
  set_scalar_kernel<<<1, 1, 0, s>>>(sh.scal, SC_OMEGA, min_omega);
  ET_CHECK_LAUNCH();
  // r = b - A x ; |b|^2
  ET_TRY(apply_A64(x, r, OP_RES, b));
  ET_TRY(dot(n_dof, b, b, partial, sh.scal + SC_BB, s));
  auto prec = [&]() -> int {
    return prm.precision == 1 ? precondition(h32) : precondition(h64);
  };
  ET_TRY(prec());
  ET_CHECK_HIP(hipMemcpyAsync(p_, z, n_dof * sizeof(double), hipMemcpyDeviceToDevice, s));
  ET_TRY(dot(n_dof, r, z, partial, sh.scal + SC_RHO, s));
  int it = 0;
  double R = 0.0;
  int status = ET_OK;
  for (it = 1; it <= maxiter; ++it) {
    ET_TRY(apply_A64(p_, q, OP_Y, nullptr));
    prof.begin(ET_PROF_PCG_VEC, s);
    ET_TRY(dot(n_dof, p_, q, partial, sh.scal + SC_PQ, s));
    pcg_update_xr_kernel<<<kRedBlocks, 256, 0, s>>>(n_dof, sh.scal, p_, q, x, r, partial);
    ET_CHECK_LAUNCH();
    ET_TRY(reduce_finalize(partial, sh.scal + SC_RR, s));
    prof.end(s);
    ET_TRY(read_scalars());
    const double pq = host_scal[SC_PQ], rr = host_scal[SC_RR], bb = host_scal[SC_BB];
    R = std::sqrt(rr) / std::sqrt(bb);
    if (!std::isfinite(pq) || !std::isfinite(rr)) {
      status = ET_ENONFINITE;
      set_error("PCG: non-finite value (p.q or r.r)");
      break;
    }
    if (!(pq > 0.0)) {
      status = ET_EBREAKDOWN;
      set_error("PCG breakdown: p^T A p <= 0");
      break;
    }
    if (R < tol) break;
    pcg_omega_kernel<<<1, 1, 0, s>>>(sh.scal, tol, min_omega, d_omega);
    ET_CHECK_LAUNCH();
    ET_TRY(prec());
    prof.begin(ET_PROF_PCG_VEC, s);
    ET_TRY(dot(n_dof, r, z, partial, sh.scal + SC_RHO_NEW, s));
    pcg_update_p_kernel<<<kRedBlocks, 256, 0, s>>>(n_dof, sh.scal, z, p_);
    ET_CHECK_LAUNCH();
    pcg_rotate_kernel<<<1, 1, 0, s>>>(sh.scal);
    prof.end(s);
    ET_CHECK_LAUNCH();
  }
  if (it > maxiter) it = maxiter;
  ET_CHECK_HIP(hipEventRecord(ev2, s));
  ET_CHECK_HIP(hipEventSynchronize(ev2));
  float a = 0.f, c = 0.f;
  ET_CHECK_HIP(hipEventElapsedTime(&a, ev0, ev1));
  ET_CHECK_HIP(hipEventElapsedTime(&c, ev1, ev2));
  t_setup_ms = a;
  t_pcg_ms = c;
  if (iters) *iters = it;
  if (residual) *residual = R;
  return status;
}

}  // namespace et
