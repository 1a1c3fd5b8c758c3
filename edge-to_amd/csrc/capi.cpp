// extern "C" entry points declared in include/edgeto_amd.h.  Each one
// validates its arguments, converts them to the internal Grid description and
// forwards to the kernels; no exception crosses this boundary.
#include <cmath>
#include <new>
#include <string>

#include "blas.hpp"
#include "coarse.hpp"
#include "common.hpp"
#include "galerkin_elem.hpp"
#include "level1.hpp"
#include "solver.hpp"
#include "stiffness.hpp"

namespace et {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }
}  // namespace et

using et::Grid;

#define ET_GUARD(...)                                    \
  try {                                                  \
    __VA_ARGS__                                          \
  } catch (const std::bad_alloc&) {                      \
    et::set_error("host allocation failed");             \
    return ET_ENOMEM;                                    \
  } catch (...) {                                        \
    et::set_error("unexpected C++ exception");           \
    return ET_EINVAL;                                    \
  }

namespace {

// The 3-D fine-grid kernels need an element matrix with the brick's reflection
// structure (k0_structure); K0 arrives in device memory here, so it is read
// back first (this synchronises the stream), and its transformed-basis table is
// built from the same copy.  *fine = tab when the fine-grid kernels apply, else
// null (2-D grids always use the general kernel).
template <typename T>
static int fine_allowed(const Grid& g, const T* K0, hipStream_t stream, et::WhtTable* tab,
                        const et::WhtTable** fine) {
  *fine = nullptr;
  if (g.dim != 3) return ET_OK;
  T h[576];
  ET_CHECK_HIP(hipMemcpyAsync(h, K0, sizeof(h), hipMemcpyDeviceToHost, stream));
  ET_CHECK_HIP(hipStreamSynchronize(stream));
  double d[576];
  for (int t = 0; t < 576; ++t) d[t] = (double)h[t];
  if (et::k0_fast3d(et::k0_structure(3, d, sizeof(T) == 8 ? et::kK0Rtol : et::kK0RtolF32))) {
    et::wht_table(d, tab);
    *fine = tab;
  }
  return ET_OK;
}

template <typename T>
static int residual_impl(const char* name, int dim, const int* nel, const T* K0, const T* E,
                         const uint8_t* cons, const T* x, const T* b, T* r, hipStream_t stream) {
  et::Grid g;
  ET_TRY(et::make_grid(dim, nel, &g));
  ET_REQUIRE(K0 && E && x && b && r, name);
  ET_REQUIRE(r != x && r != b, "residual: r must not alias x or b");
  et::WhtTable tab;
  const et::WhtTable* fine = nullptr;
  ET_TRY(fine_allowed(g, K0, stream, &tab, &fine));
  return et::level0_op<T>(1, g, K0, E, cons, x, b, nullptr, nullptr, 1.0, nullptr, r, stream, fine);
}

template <typename T>
static int jacobi_impl(const char* name, int dim, const int* nel, const T* K0, const T* E,
                       const uint8_t* cons, const T* x, const T* b, const T* Dinv, double omega, T* y,
                       hipStream_t stream) {
  et::Grid g;
  ET_TRY(et::make_grid(dim, nel, &g));
  ET_REQUIRE(K0 && E && x && b && Dinv && y, name);
  ET_REQUIRE(y != x && y != b && y != Dinv, "jacobi: y must not alias x, b or Dinv");
  ET_REQUIRE(std::isfinite(omega), "jacobi: omega must be finite");
  et::WhtTable tab;
  const et::WhtTable* fine = nullptr;
  ET_TRY(fine_allowed(g, K0, stream, &tab, &fine));
  return et::level0_op<T>(2, g, K0, E, cons, x, b, Dinv, nullptr, omega, nullptr, y, stream, fine);
}

}  // namespace

extern "C" {

const char* et_last_error(void) { return et::last_error(); }
const char* et_version(void) { return "edgeto_amd 0.1 (gfx950)"; }

int et_tune_set(int param, int value) {
  ET_GUARD({
    if (param == ET_TUNE_ROW_STENCIL_NODES) {
      ET_REQUIRE(value >= 0, "et_tune_set: node count must be >= 0");
      et::set_row_stencil_nodes(value);
      return ET_OK;
    }
    ET_REQUIRE(value >= 0 && value <= 1024, "et_tune_set: value out of range");
    if (param == ET_TUNE_TILED_JC) {
      ET_REQUIRE(value >= 1, "et_tune_set: jc must be >= 1");
      et::set_tiled_jc(value);
    } else if (param == ET_TUNE_MFMA_JC) {
      ET_REQUIRE(value >= 1, "et_tune_set: jc must be >= 1");
      et::set_mfma_jc(value);
    } else if (param == ET_TUNE_MIXED_MFMA) {
      et::set_mixed_mfma(value);
    } else if (param == ET_TUNE_MX_F32 || param == ET_TUNE_MX_F64) {
      et::set_mx(param == ET_TUNE_MX_F64, value);
    } else if (param == ET_TUNE_MX_WGS_F32 || param == ET_TUNE_MX_WGS_F64) {
      ET_REQUIRE(value >= 1 && value <= 8, "et_tune_set: workgroups per CU must be in [1, 8]");
      et::set_mx_wgs(param == ET_TUNE_MX_WGS_F64, value);
    } else if (param == ET_TUNE_WHT_THREADS_F32 || param == ET_TUNE_WHT_THREADS_F64) {
      ET_REQUIRE(value == 256 || value == 512 || value == 1024,
                 "et_tune_set: workgroup size must be 256, 512 or 1024");
      et::set_wht_threads(param == ET_TUNE_WHT_THREADS_F64, value);
    } else if (param == ET_TUNE_WHT_SLICES_F32 || param == ET_TUNE_WHT_SLICES_F64) {
      ET_REQUIRE(value == 8 || value == 16, "et_tune_set: slices per layer must be 8 or 16");
      et::set_wht_slices(param == ET_TUNE_WHT_SLICES_F64, value);
    } else if (param == ET_TUNE_LEVEL1_MF) {
      et::set_level1_mf(value);
    } else if (param == ET_TUNE_FUSED_RZ) {
      et::set_fused_rz(value);
    } else if (param == ET_TUNE_COARSE_LIB) {
      ET_REQUIRE(value >= 0 && value <= 2, "et_tune_set: coarse factor path must be 0, 1 or 2");
      et::set_coarse_lib(value);
    } else if (param == ET_TUNE_WHT) {
      et::set_wht(value);
    } else if (param == ET_TUNE_WHT_WGS) {
      ET_REQUIRE(value >= 1 && value <= 4, "et_tune_set: workgroups per CU must be in [1, 4]");
      et::set_wht_wgs(value);
    } else if (param == ET_TUNE_FILTER_ROWS) {
      ET_REQUIRE(value == 4 || value == 8, "et_tune_set: filter tile rows must be 4 or 8");
      et::set_stencil_rows(value);
    } else {
      ET_REQUIRE(false, "et_tune_set: unknown parameter");
    }
    return ET_OK;
  })
}

int et_tune_get(int param) {
  if (param == ET_TUNE_TILED_JC) return et::get_tiled_jc();
  if (param == ET_TUNE_MFMA_JC) return et::get_mfma_jc();
  if (param == ET_TUNE_MIXED_MFMA) return et::get_mixed_mfma();
  if (param == ET_TUNE_MX_F32 || param == ET_TUNE_MX_F64) return et::get_mx(param == ET_TUNE_MX_F64);
  if (param == ET_TUNE_MX_WGS_F32 || param == ET_TUNE_MX_WGS_F64)
    return et::get_mx_wgs(param == ET_TUNE_MX_WGS_F64);
  if (param == ET_TUNE_WHT) return et::get_wht();
  if (param == ET_TUNE_COARSE_LIB) return et::get_coarse_lib();
  if (param == ET_TUNE_LEVEL1_MF) return et::get_level1_mf();
  if (param == ET_TUNE_FUSED_RZ) return et::get_fused_rz();
  if (param == ET_TUNE_WHT_SLICES_F32 || param == ET_TUNE_WHT_SLICES_F64)
    return et::get_wht_slices(param == ET_TUNE_WHT_SLICES_F64);
  if (param == ET_TUNE_ROW_STENCIL_NODES) return (int)et::get_row_stencil_nodes();
  if (param == ET_TUNE_WHT_THREADS_F32 || param == ET_TUNE_WHT_THREADS_F64)
    return et::get_wht_threads(param == ET_TUNE_WHT_THREADS_F64);
  if (param == ET_TUNE_WHT_WGS) return et::get_wht_wgs();
  if (param == ET_TUNE_FILTER_ROWS) return et::get_stencil_rows();
  return -1;
}

int et_k0_structure(int dim, const double* K0_host, int* flags) {
  ET_GUARD({
    ET_REQUIRE(dim == 2 || dim == 3, "et_k0_structure: dim must be 2 or 3");
    ET_REQUIRE(K0_host && flags, "et_k0_structure: null pointer");
    *flags = et::k0_structure(dim, K0_host, et::kK0Rtol);
    return ET_OK;
  })
}

int et_generate_elements(int dim, const int* nel, int32_t* elements, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(elements, "elements is null");
    return et::generate_elements(g, elements, stream);
  })
}

int et_matvec_f64(int dim, const int* nel, const double* K0, const double* E, const uint8_t* cons,
                  const double* u, double* y, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && E && u && y, "et_matvec_f64: null pointer");
    ET_REQUIRE(u != y, "et_matvec_f64: u and y must not alias");
    et::WhtTable tab;
    const et::WhtTable* fine = nullptr;
    ET_TRY(fine_allowed(g, K0, stream, &tab, &fine));
    return et::matvec<double>(g, K0, E, cons, u, y, stream, fine);
  })
}

int et_matvec_f32(int dim, const int* nel, const float* K0, const float* E, const uint8_t* cons,
                  const float* u, float* y, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && E && u && y, "et_matvec_f32: null pointer");
    ET_REQUIRE(u != y, "et_matvec_f32: u and y must not alias");
    et::WhtTable tab;
    const et::WhtTable* fine = nullptr;
    ET_TRY(fine_allowed(g, K0, stream, &tab, &fine));
    return et::matvec<float>(g, K0, E, cons, u, y, stream, fine);
  })
}

int et_residual_f64(int dim, const int* nel, const double* K0, const double* E, const uint8_t* cons,
                    const double* x, const double* b, double* r, hipStream_t stream) {
  ET_GUARD({ return residual_impl<double>("et_residual_f64: null pointer", dim, nel, K0, E, cons, x, b, r, stream); })
}
int et_residual_f32(int dim, const int* nel, const float* K0, const float* E, const uint8_t* cons,
                    const float* x, const float* b, float* r, hipStream_t stream) {
  ET_GUARD({ return residual_impl<float>("et_residual_f32: null pointer", dim, nel, K0, E, cons, x, b, r, stream); })
}
int et_jacobi_f64(int dim, const int* nel, const double* K0, const double* E, const uint8_t* cons,
                  const double* x, const double* b, const double* Dinv, double omega, double* y,
                  hipStream_t stream) {
  ET_GUARD({
    return jacobi_impl<double>("et_jacobi_f64: null pointer", dim, nel, K0, E, cons, x, b, Dinv, omega, y, stream);
  })
}
int et_jacobi_f32(int dim, const int* nel, const float* K0, const float* E, const uint8_t* cons,
                  const float* x, const float* b, const float* Dinv, double omega, float* y,
                  hipStream_t stream) {
  ET_GUARD({
    return jacobi_impl<float>("et_jacobi_f32: null pointer", dim, nel, K0, E, cons, x, b, Dinv, omega, y, stream);
  })
}

int et_matvec_f32_mfma(int dim, const int* nel, const double* K0_host, const float* K0,
                       const float* E, const uint8_t* cons, const float* u, float* y,
                       hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0_host && K0 && E && u && y, "et_matvec_f32_mfma: null pointer");
    ET_REQUIRE(u != y, "et_matvec_f32_mfma: u and y must not alias");
    ET_REQUIRE(et::mfma_supported(g), "et_matvec_f32_mfma: 3-D grids only (nz <= ~290)");
    ET_REQUIRE(et::k0_fast3d(et::k0_structure(3, K0_host, et::kK0Rtol)),
               "et_matvec_f32_mfma: K0 lacks the brick's reflection structure (use et_matvec_f32)");
    return et::op0_mfma(0, g, (int64_t)g.ey * g.ez, g.ez, et::k0_scale(K0_host, 576), K0, E, cons, u,
                        nullptr, nullptr, nullptr, 1.0, nullptr, y, stream);
  })
}

int et_diagonal_f64(int dim, const int* nel, const double* K0, const double* E,
                    const uint8_t* cons, double* diag, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && E && diag, "et_diagonal_f64: null pointer");
    return et::diagonal<double>(g, K0, E, cons, diag, stream);
  })
}

int et_elem_energy_f64(int dim, const int* nel, const double* K0, const double* U, double* dk,
                       hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && U && dk, "et_elem_energy_f64: null pointer");
    return et::elem_energy(g, K0, U, nullptr, 0.0, 0.0, 0, dk, stream);
  })
}

int et_sensitivity_f64(int dim, const int* nel, const double* K0, const double* U,
                       const double* xt, double penal, double void_, double* out,
                       hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && U && xt && out, "et_sensitivity_f64: null pointer");
    return et::elem_energy(g, K0, U, xt, penal, void_, 1, out, stream);
  })
}

int et_fea_integrals_f64(int dim, const int* nel, const double* K0, const double* D, const double* B,
                         const double* E, const double* U, double* strain, double* stress,
                         double* energy, double* von_mises, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && D && B && E && U && strain && stress && energy,
               "et_fea_integrals_f64: null pointer");
    return et::fea_integrals(g, K0, D, B, E, U, strain, stress, energy, von_mises, stream);
  })
}

int et_simp_forward_f64(int64_t n, const double* xt, double penal, double void_, int plain,
                        double* E, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(n > 0 && xt && E, "et_simp_forward_f64: bad arguments");
    return et::simp_forward(n, xt, penal, void_, plain, E, stream);
  })
}

int et_filter_normalizer_f64(int dim, const int* nel, int n_off, const int32_t* offsets,
                             const double* w, double* norm, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(offsets && w && norm, "et_filter_normalizer_f64: null pointer");
    return et::filter_normalizer(g, n_off, offsets, w, norm, stream);
  })
}

int et_filter_fwd_f64(int dim, const int* nel, int n_off, const int32_t* offsets, const double* w,
                      const double* x, double* y, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(offsets && w && x && y, "et_filter_fwd_f64: null pointer");
    ET_REQUIRE(x != y, "et_filter_fwd_f64: x and y must not alias");
    return et::filter_fwd(g, n_off, offsets, w, x, y, stream);
  })
}

int et_filter_adj_f64(int dim, const int* nel, int n_off, const int32_t* offsets, const double* w,
                      const double* norm, const double* v, double* y, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(offsets && w && norm && v && y, "et_filter_adj_f64: null pointer");
    ET_REQUIRE(v != y, "et_filter_adj_f64: v and y must not alias");
    return et::filter_adj(g, n_off, offsets, w, norm, v, y, stream);
  })
}

int et_simp_adjoint_f64(int dim, const int* nel, const double* K0, const double* U, const double* xt,
                        double penal, double void_, int n_off, const int32_t* offsets, const double* w,
                        const double* norm, int k0_flags, double* out, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && U && xt && offsets && w && norm && out, "et_simp_adjoint_f64: null pointer");
    ET_REQUIRE(out != xt && out != norm, "et_simp_adjoint_f64: out must not alias xt or norm");
    return et::simp_adjoint(g, K0, U, xt, penal, void_, n_off, offsets, w, norm, k0_flags, out, stream);
  })
}

int et_restrict_f64(int dim, const int* nel, int dof, const double* vf, double* vc,
                    hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(vf && vc && dof >= 1 && dof <= 3, "et_restrict_f64: bad arguments");
    return et::restrict_op<double>(g, dof, vf, vc, stream);
  })
}

int et_prolong_f64(int dim, const int* nel, int dof, const double* vc, double* vf,
                   hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(vf && vc && dof >= 1 && dof <= 3, "et_prolong_f64: bad arguments");
    return et::prolong_op<double>(g, dof, vc, vf, stream);
  })
}

int et_stencil_from_operator_f64(int dim, const int* nel, const double* K0, const double* E,
                                 const uint8_t* cons, double* S, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && E && S, "et_stencil_from_operator_f64: null pointer");
    return et::stencil_from_operator(g, K0, E, cons, S, stream);
  })
}

int et_galerkin_f64(int dim, const int* nel, int dof, const double* S_fine, double* S_coarse,
                    hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(S_fine && S_coarse, "et_galerkin_f64: null pointer");
    return et::galerkin(g, dof, S_fine, S_coarse, stream);
  })
}

int et_galerkin_elements_f64(int dim, const int* nel, const double* K0_host, const double* E,
                             const uint8_t* cons, double* S_coarse, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0_host && E && cons && S_coarse, "et_galerkin_elements_f64: null pointer");
    ET_REQUIRE(et::galerkin_elem_supported(g), "et_galerkin_elements_f64: element counts must be even");
    const int nd = (dim == 3 ? 8 : 4) * dim;
    const std::vector<double> M = et::galerkin_child_tables(dim, K0_host);
    double *Mt = nullptr, *K0 = nullptr;
    uint8_t* flag = nullptr;
    int st = ET_OK;
    if (hipMalloc(&Mt, M.size() * sizeof(double)) != hipSuccess ||
        hipMalloc(&K0, nd * nd * sizeof(double)) != hipSuccess || hipMalloc(&flag, g.n_elem()) != hipSuccess) {
      et::set_error("et_galerkin_elements_f64: device allocation failed");
      st = ET_ENOMEM;
    }
    if (st == ET_OK && (hipMemcpy(Mt, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemcpy(K0, K0_host, nd * nd * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)) {
      et::set_error("et_galerkin_elements_f64: copy failed");
      st = ET_EHIP;
    }
    if (st == ET_OK) st = et::galerkin_from_elements(g, K0, Mt, E, cons, flag, S_coarse, stream);
    if (hipStreamSynchronize(stream) != hipSuccess && st == ET_OK) st = ET_EHIP;
    (void)hipFree(Mt);
    (void)hipFree(K0);
    (void)hipFree(flag);
    return st;
  })
}

int et_stencil_apply_f64(int dim, const int* nel, int dof, const double* S, const double* x,
                         double* y, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(S && x && y && dof == dim, "et_stencil_apply_f64: bad arguments");
    return et::stencil_apply<double>(g, dof, S, x, y, stream);
  })
}

int et_dot_f64(int64_t n, const double* a, const double* b, double* out, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(n > 0 && a && b && out, "et_dot_f64: bad arguments");
    // partial sums live in a per-thread device scratch buffer kept for the process
    static thread_local double* partial = nullptr;
    if (partial == nullptr) ET_CHECK_HIP(hipMalloc(&partial, et::kRedBlocks * sizeof(double)));
    return et::dot(n, a, b, partial, out, stream);
  })
}

int et_sqdist_f64(int64_t n, const double* a, const double* b, double* out, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(n > 0 && a && b && out, "et_sqdist_f64: bad arguments");
    static thread_local double* partial = nullptr;
    if (partial == nullptr) ET_CHECK_HIP(hipMalloc(&partial, et::kRedBlocks * sizeof(double)));
    return et::sqdist(n, a, b, partial, out, stream);
  })
}

int64_t et_krylov_work_size(int64_t n_dof, int restart) { return et::krylov_work_size(n_dof, restart); }

int et_cg_f64(int dim, const int* nel, const double* K0, const double* E, const uint8_t* cons, const double* b,
              double* x, double tol, int maxiter, void* work, int64_t work_bytes, int* iters, double* residual,
              hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && E && b && x && work, "et_cg_f64: null pointer");
    ET_REQUIRE(b != x, "et_cg_f64: b and x must not alias");
    ET_REQUIRE(std::isfinite(tol) && tol >= 0, "et_cg_f64: tol must be finite and >= 0");
    et::WhtTable tab;
    const et::WhtTable* fine = nullptr;
    ET_TRY(fine_allowed(g, K0, stream, &tab, &fine));
    return et::cg_solve(g, K0, E, cons, fine, b, x, tol, maxiter, work, work_bytes, iters, residual, stream);
  })
}

int et_gmres_f64(int dim, const int* nel, const double* K0, const double* E, const uint8_t* cons, const double* b,
                 double* x, double tol, int restart, int maxiter, void* work, int64_t work_bytes, int* iters,
                 double* residual, hipStream_t stream) {
  ET_GUARD({
    Grid g;
    ET_TRY(et::make_grid(dim, nel, &g));
    ET_REQUIRE(K0 && E && b && x && work, "et_gmres_f64: null pointer");
    ET_REQUIRE(b != x, "et_gmres_f64: b and x must not alias");
    ET_REQUIRE(std::isfinite(tol) && tol >= 0, "et_gmres_f64: tol must be finite and >= 0");
    et::WhtTable tab;
    const et::WhtTable* fine = nullptr;
    ET_TRY(fine_allowed(g, K0, stream, &tab, &fine));
    return et::gmres_solve(g, K0, E, cons, fine, b, x, tol, restart, maxiter, work, work_bytes, iters, residual,
                           stream);
  })
}

int et_multimat_f64(int64_t n, int m, const double* rho, const double* E_mat, double penal, double void_, int rp,
                    int mode, double* out, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(n > 0 && rho && E_mat && out, "et_multimat_f64: bad arguments");
    ET_REQUIRE(mode >= 0 && mode <= 2, "et_multimat_f64: mode must be 0, 1 or 2");
    return et::multimat(n, m, rho, E_mat, penal, void_, rp, mode, out, stream);
  })
}

int64_t et_oc_multi_work_size(int64_t n, int m) { return et::oc_multi_work_size(n, m); }

int et_oc_update_multi_f64(int64_t n, int m, const double* x, const double* dc, double A, double vol, double move,
                           const double* mass, const double* rho_mat, double* x_new, void* work, int* steps,
                           hipStream_t stream) {
  ET_GUARD({ return et::oc_update_multi(n, m, x, dc, A, vol, move, mass, rho_mat, x_new, work, steps, stream); })
}

int64_t et_oc_work_size(int64_t n) { return et::oc_work_size(n); }

int et_oc_update_f64(int64_t n, const double* x, const double* dc, double A, double vol, double vf,
                     double move, double* x_new, void* work, int* steps, hipStream_t stream) {
  ET_GUARD({ return et::oc_update(n, x, dc, A, vol, vf, move, x_new, work, steps, stream); })
}

int et_oc_update_elem_f64(int64_t n, const double* x, const double* dc, const double* A_elem, double vol,
                          double vf, double move, double* x_new, void* work, int* steps, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(A_elem != nullptr, "et_oc_update_elem_f64: A_elem is null");
    return et::oc_update(n, x, dc, 0.0, vol, vf, move, x_new, work, steps, stream, A_elem);
  })
}

// ------------------------------------------------------------ general meshes

int et_umesh_star(int64_t n_nodes, int64_t n_el, const int* elem_ptr, const int* elem_nodes, int* star_ptr,
                  int* star, int* el_of, hipStream_t stream) {
  ET_GUARD({ return et::umesh_star(n_nodes, n_el, elem_ptr, elem_nodes, star_ptr, star, el_of, stream); })
}

int et_umatvec_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const double* u, double* y,
                   hipStream_t stream) {
  ET_GUARD({ return et::umatvec(mesh, 0, E, cons, u, nullptr, y, stream); })
}

int et_udiagonal_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, double* diag,
                     hipStream_t stream) {
  ET_GUARD({ return et::udiagonal(mesh, E, cons, diag, stream); })
}

int et_uelem_energy_f64(const et_umesh* mesh, const double* U, double* dk, hipStream_t stream) {
  ET_GUARD({ return et::uenergy(mesh, U, nullptr, 0.0, 0.0, 0, dk, stream); })
}

int et_usensitivity_f64(const et_umesh* mesh, const double* U, const double* xt, double penal, double void_,
                        double* out, hipStream_t stream) {
  ET_GUARD({ return et::uenergy(mesh, U, xt, penal, void_, 1, out, stream); })
}

int et_ufea_integrals_f64(const et_umesh* mesh, int nb, const double* B, const int64_t* B_ptr, const double* D,
                          const int64_t* D_ptr, const double* E, const double* U, double* strain, double* stress,
                          double* energy, double* vm, hipStream_t stream) {
  ET_GUARD({ return et::ufea(mesh, nb, B, B_ptr, D, D_ptr, E, U, strain, stress, energy, vm, stream); })
}

int et_csr_spmv_f64(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals, const double* x,
                    double* y, hipStream_t stream) {
  ET_GUARD({ return et::csr_spmv(n_rows, row_ptr, cols, vals, x, y, stream); })
}

int et_umesh_pattern(const et_umesh* mesh, int* row_ptr, int* cols, int64_t cap, int64_t* nnz, hipStream_t stream) {
  ET_GUARD({ return et::umesh_pattern(mesh, row_ptr, cols, cap, nnz, stream); })
}

int et_uassemble_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const int* row_ptr,
                     const int* cols, double* vals, hipStream_t stream) {
  ET_GUARD({ return et::uassemble(mesh, E, cons, row_ptr, cols, vals, stream); })
}

int et_csr_cg_f64(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals, const double* b,
                  double* x, double tol, int maxiter, void* work, int64_t work_bytes, int* iters, double* residual,
                  hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(n_rows > 0 && row_ptr && cols && vals && b && x && work, "et_csr_cg_f64: bad arguments");
    ET_REQUIRE(b != x, "et_csr_cg_f64: b and x must not alias");
    ET_REQUIRE(std::isfinite(tol) && tol >= 0, "et_csr_cg_f64: tol must be finite and >= 0");
    return et::cg_solve_op(et::csr_op(n_rows, row_ptr, cols, vals), b, x, tol, maxiter, work, work_bytes, iters,
                           residual, stream);
  })
}

int et_csr_gmres_f64(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals, const double* b,
                     double* x, double tol, int restart, int maxiter, void* work, int64_t work_bytes, int* iters,
                     double* residual, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(n_rows > 0 && row_ptr && cols && vals && b && x && work, "et_csr_gmres_f64: bad arguments");
    ET_REQUIRE(b != x, "et_csr_gmres_f64: b and x must not alias");
    ET_REQUIRE(std::isfinite(tol) && tol >= 0, "et_csr_gmres_f64: tol must be finite and >= 0");
    return et::gmres_solve_op(et::csr_op(n_rows, row_ptr, cols, vals), b, x, tol, restart, maxiter, work,
                              work_bytes, iters, residual, stream);
  })
}

int et_ucg_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const double* b, double* x, double tol,
               int maxiter, void* work, int64_t work_bytes, int* iters, double* residual, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(mesh && E && b && x && work, "et_ucg_f64: null pointer");
    ET_REQUIRE(b != x, "et_ucg_f64: b and x must not alias");
    ET_REQUIRE(std::isfinite(tol) && tol >= 0, "et_ucg_f64: tol must be finite and >= 0");
    return et::cg_solve_op(et::unstructured_op(mesh, E, cons), b, x, tol, maxiter, work, work_bytes, iters,
                           residual, stream);
  })
}

int et_ugmres_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const double* b, double* x,
                  double tol, int restart, int maxiter, void* work, int64_t work_bytes, int* iters,
                  double* residual, hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(mesh && E && b && x && work, "et_ugmres_f64: null pointer");
    ET_REQUIRE(b != x, "et_ugmres_f64: b and x must not alias");
    ET_REQUIRE(std::isfinite(tol) && tol >= 0, "et_ugmres_f64: tol must be finite and >= 0");
    return et::gmres_solve_op(et::unstructured_op(mesh, E, cons), b, x, tol, restart, maxiter, work, work_bytes,
                              iters, residual, stream);
  })
}

int et_oc_update_f64_noalloc(int64_t n, const double* x, const double* dc, double A, double vol,
                             double vf, double move, double* x_new, int* steps, hipStream_t stream) {
  ET_GUARD({ return et::oc_update_noalloc(n, x, dc, A, vol, vf, move, x_new, steps, stream); })
}

int64_t et_oc_noalloc_max_n(void) { return et::oc_noalloc_max_n(); }

int et_mg_create(et_mg** handle, const et_mg_params* params, const double* K0_host,
                 hipStream_t stream) {
  ET_GUARD({
    ET_REQUIRE(handle && params && K0_host, "et_mg_create: null pointer");
    *handle = nullptr;
    et_mg* h = new et_mg();
    const int st = h->impl.create(*params, K0_host, stream);
    if (st != ET_OK) {
      h->impl.destroy();
      delete h;
      return st;
    }
    *handle = h;
    return ET_OK;
  })
}

int et_mg_destroy(et_mg* handle) {
  ET_GUARD({
    if (handle) {
      handle->impl.destroy();
      delete handle;
    }
    return ET_OK;
  })
}

int et_mg_solve(et_mg* handle, const double* E, const uint8_t* cons, const double* b, double* x,
                double tol, int maxiter, int* iters, double* residual) {
  ET_GUARD({
    ET_REQUIRE(handle, "et_mg_solve: null handle");
    return handle->impl.solve(E, cons, b, x, tol, maxiter, iters, residual);
  })
}

int et_mg_last_timings(et_mg* handle, double* setup_ms, double* pcg_ms) {
  ET_GUARD({
    ET_REQUIRE(handle, "et_mg_last_timings: null handle");
    if (setup_ms) *setup_ms = handle->impl.t_setup_ms;
    if (pcg_ms) *pcg_ms = handle->impl.t_pcg_ms;
    return ET_OK;
  })
}

int et_mg_profile(et_mg* handle, int enable) {
  ET_GUARD({
    ET_REQUIRE(handle, "et_mg_profile: null handle");
    ET_CHECK_HIP(hipStreamSynchronize(handle->impl.stream));
    handle->impl.prof.reset();
    handle->impl.prof.mask = (unsigned)enable & ((1u << ET_PROF_KINDS) - 1u);
    return ET_OK;
  })
}

int et_mg_coarse_solve(et_mg* handle, const void* b, void* x) {
  ET_GUARD({
    ET_REQUIRE(handle && b && x, "et_mg_coarse_solve: null pointer");
    ET_REQUIRE(b != x, "et_mg_coarse_solve: b and x must not alias");
    return handle->impl.coarse_solve(b, x);
  })
}

int et_mg_apply_level(et_mg* handle, int level, const void* x, void* y) {
  ET_GUARD({
    ET_REQUIRE(handle && x && y, "et_mg_apply_level: null pointer");
    ET_REQUIRE(x != y, "et_mg_apply_level: x and y must not alias");
    return handle->impl.apply_level(level, x, y);
  })
}

int et_mg_profile_read(et_mg* handle, double* ms, int64_t* launches) {
  ET_GUARD({
    ET_REQUIRE(handle && ms && launches, "et_mg_profile_read: null pointer");
    return handle->impl.profile_read(ms, launches);
  })
}

}  // extern "C"
