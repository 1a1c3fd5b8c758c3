// Internal declarations shared between the kernel translation units, the
// solver and the C-ABI layer.
#pragma once
#include <functional>
#include "common.hpp"
#include "op_tiled.hpp"

namespace et {

int generate_elements(const Grid& g, int32_t* el, hipStream_t s);

// Structure of a row-major element matrix (host, FP64; k0_structure.cpp):
// K0S_SYM symmetric; K0S_REFLECT (3-D) blocks invariant under the brick's axis
// reflections; K0S_WHT (3-D) only the couplings the transformed-basis operator
// keeps; K0S_LEVEL1 (3-D) the matrix-free first coarse level's terms cover it.
// Entries count as zero / equal within rtol * max|K0|.
enum { K0S_SYM = 1, K0S_REFLECT = 2, K0S_WHT = 4, K0S_LEVEL1 = 8 };
int k0_structure(int dim, const double* K0, double rtol);
constexpr double kK0Rtol = 1e-12;  // FP64 element matrices
constexpr double kK0RtolF32 = 1e-6; // element matrices given in FP32
// the 3-D fine-grid operators (transformed-basis, LDS-tiled, FP16 matrix-core) apply
inline bool k0_fast3d(int flags) {
  return (flags & (K0S_SYM | K0S_REFLECT | K0S_WHT)) == (K0S_SYM | K0S_REFLECT | K0S_WHT);
}

// Coefficients of the transformed-basis fine-grid operator for one element matrix
// (op_wht.hip): the 45 non-zeros of G = H K0 H / 64, 15 per output component in
// the order its contraction uses them, each component padded to 16.  Built on the
// host (wht_table) by whoever holds K0 there; a null table stands for an element
// matrix without the brick's reflection structure (k0_fast3d false), which keeps 3-D
// grids on the general node-gather kernels.
struct WhtTable {
  double g[48];
};
void wht_table(const double* K0_host, WhtTable* t);

template <typename T>
int matvec(const Grid& g, const T* K0, const T* E, const uint8_t* cons, const T* u, T* y,
           hipStream_t s, const WhtTable* fine = nullptr);
template <typename T>
int diagonal(const Grid& g, const T* K0, const T* E, const uint8_t* cons, T* diag, hipStream_t s);

// LDS-tiled fine-grid H8 operator; mode: 0 y = Ax, 1 y = b - Ax,
// 2 y = x + w D^-1 (b - Ax), 3 the same plus xadd.  Element index of (ex, ey, ez)
// is ex*e_sx + ey*e_sy + ez.  omega/wscale only used by modes 2 and 3.
template <typename T>
int op0_tiled(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const T* K0, const T* E,
              const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
              double wscale, const T* xadd, T* y, hipStream_t s);
bool tiled_supported(const Grid& g, size_t elem_bytes);
// Matrix-core variant at working precision (FP32 or FP64 MFMA), same modes.
template <typename T>
int op0_mx(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const T* K0, const T* E,
           const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
           double wscale, const T* xadd, T* y, hipStream_t s);
bool mx_supported(const Grid& g, size_t elem_bytes);
void set_mx(int is_double, int on);
int get_mx(int is_double);
void set_mx_wgs(int is_double, int n);
int get_mx_wgs(int is_double);
// Transformed-basis (corner Walsh-Hadamard) element operator, FP32 or FP64, same modes.
template <typename T>
int op0_wht(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const WhtTable& tab, const T* E,
            const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
            double wscale, const T* xadd, T* y, hipStream_t s);
bool wht_supported(const Grid& g, size_t elem_bytes);
// y = A x (FP64) fused with the x.y partial sum of every workgroup (fixed order);
// writes wht_partials(g, sizeof(T), TM_YDOT) <= max_partial values to partial.
template <typename T>
int op0_wht_ydot(const Grid& g, int64_t e_sx, int64_t e_sy, const WhtTable& tab, const T* E, const uint8_t* cons,
                 const T* x, T* y, double* partial, int max_partial, hipStream_t s);
// number of per-workgroup partial sums written by the fused-dot variants
int wht_partials(const Grid& g, size_t elem_bytes, int mode);
// Final post-smoothing sweep y = xadd + Jacobi(x; b) with the per-workgroup
// partial sums of dotv.y (wht_partials(g, sizeof(T), TM_JAC_ADD_DOT) values).
template <typename T>
int op0_wht_jac_add_dot(const Grid& g, int64_t e_sx, int64_t e_sy, const WhtTable& tab, const T* E, const uint8_t* cons,
                        const T* x, const T* b, const T* Dinv, const double* omega, double wscale, const T* xadd,
                        T* y, const T* dotv, double* partial, int max_partial, hipStream_t s);
void set_wht(int on);
int get_wht();
void set_wht_wgs(int n);
int get_wht_wgs();
void set_wht_threads(int is_double, int nt);
int get_wht_threads(int is_double);
void set_wht_slices(int is_double, int ns);
// tiles of the calling host thread's launches ([float, double]; 0 = the process-wide setting)
void set_thread_wht_tiles(const int* nt, const int* ns);
int get_wht_slices(int is_double);
// Fine-grid dispatch: the transformed-basis operator where enabled and
// supported, else the matrix-core one where enabled, else the LDS-tiled one.
// tab: the element matrix's table (it has the reflection structure).
template <typename T>
int op0_fine(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, const T* K0, const WhtTable& tab, const T* E,
             const uint8_t* cons, const T* x, const T* b, const T* Dinv, const double* omega,
             double wscale, const T* xadd, T* y, hipStream_t s);
// whether op0_fine supports the fused zero-guess Jacobi mode (5) on this grid
bool fine_jac0_supported(const Grid& g, size_t elem_bytes);
bool fine_supported(const Grid& g, size_t elem_bytes);
// Any level-0 operator mode on any grid (fine-grid kernels in 3-D when supported).
// omega == nullptr means omega = 1.
template <typename T>
int level0_op(int mode, const Grid& g, const T* K0, const T* E, const uint8_t* cons, const T* x,
              const T* b, const T* Dinv, const double* omega, double wscale, const T* xadd, T* y,
              hipStream_t s, const WhtTable* fine = nullptr);
// Matrix-core FP32 variant (FP16 hi/lo MFMA); kscale: power of two >= max|K0|.
int op0_mfma(int mode, const Grid& g, int64_t e_sx, int64_t e_sy, float kscale, const float* K0,
             const float* E, const uint8_t* cons, const float* x, const float* b, const float* Dinv,
             const double* omega, double wscale, const float* xadd, float* y, hipStream_t s);
bool mfma_supported(const Grid& g);
void set_mfma_jc(int jc);
int get_mfma_jc();
void set_mixed_mfma(int on);
int get_mixed_mfma();
float k0_scale(const double* K0_host, int n);
void set_tiled_jc(int jc);
int get_tiled_jc();

// mode 0: -u^T K0 u; mode 1: SIMP sensitivity p*clip(xt)^(p-1) * (-u^T K0 u)
int elem_energy(const Grid& g, const double* K0, const double* U, const double* xt, double penal,
                double vmin, int mode, double* out, hipStream_t s);
// strain (n_elem x NB), stress (n_elem x NB), energy (n_elem), von Mises (n_elem, may be
// null) of every element; NB = 3 (2-D) or 6 (3-D)
int fea_integrals(const Grid& g, const double* K0, const double* D, const double* B, const double* E,
                  const double* U, double* strain, double* stress, double* energy, double* vm,
                  hipStream_t s);
int simp_forward(int64_t n, const double* xt, double penal, double vmin, int plain, double* E,
                 hipStream_t s);
int stencil_from_operator(const Grid& g, const double* K0, const double* E, const uint8_t* cons,
                          double* S, hipStream_t s);

// filter
int filter_normalizer(const Grid& g, int n_off, const int32_t* off, const double* w, double* norm,
                      hipStream_t s);
// rows of the 3-D filter stencils' tiles (4 or 8)
void set_stencil_rows(int ty);
int get_stencil_rows();
int filter_fwd(const Grid& g, int n_off, const int32_t* off, const double* w, const double* x,
               double* y, hipStream_t s);
int filter_adj(const Grid& g, int n_off, const int32_t* off, const double* w, const double* norm,
               const double* v, double* y, hipStream_t s);

// y = H^T [p clip(xt, void, 1)^(p-1) (-u_e^T K0 u_e)] in one LDS-halo stencil pass
// (k0_flags: k0_structure of K0; the transformed-basis energy needs K0S_WHT)
int simp_adjoint(const Grid& g, const double* K0, const double* U, const double* xt, double penal, double vmin,
                 int n_off, const int32_t* off, const double* w, const double* norm, int k0_flags, double* y,
                 hipStream_t s);

// transfers (g = fine grid)
template <typename T>
// with x0: also x0 = omega * wscale * Dinv * vc (first Jacobi sweep from zero on the coarse grid)
int restrict_op(const Grid& g, int dof, const T* vf, T* vc, hipStream_t s, const T* Dinv = nullptr,
                const double* omega = nullptr, double wscale = 1.0, T* x0 = nullptr);
template <typename T>
int prolong_op(const Grid& g, int dof, const T* vc, T* vf, hipStream_t s);

// stencils (double)
// work: galerkin_work_doubles(g) doubles for the three-pass 3-D product, or
// nullptr for the one-pass kernels (no scratch)
int galerkin(const Grid& g, int dof, const double* Sf, double* Sc, hipStream_t s, double* work = nullptr);
int64_t galerkin_work_doubles(const Grid& g);
template <typename T>
int stencil_apply(const Grid& g, int dof, const T* S, const T* x, T* y, hipStream_t s);

// multi-material SIMP models (multimat.hip): mode 0 E, 1 plain E, 2 dE/drho (n x m)
int multimat(int64_t n, int m, const double* rho, const double* Emat, double penal, double vmin, int rp, int mode,
             double* out, hipStream_t s);
int64_t oc_multi_work_size(int64_t n, int m);
int oc_update_multi(int64_t n, int m, const double* x, const double* dc, double A, double vol, double move,
                    const double* mass, const double* rho_mat, double* xn, void* work, int* steps, hipStream_t s);

// unpreconditioned Krylov solvers on the structured operator (krylov.hip)
int64_t krylov_work_size(int64_t n, int restart);
// A linear operator for the Krylov solvers: mode 0 y = A x, mode 1 y = b - A x (b may be
// null in mode 0), on the given stream; n unknowns.
struct KrylovOp {
  int64_t n = 0;
  std::function<int(int mode, const double* x, const double* b, double* y, hipStream_t s)> apply;
};
// general (unstructured) meshes (unstructured.hip)
int umesh_star(int64_t n_nodes, int64_t n_el, const int* elem_ptr, const int* elem_nodes, int* star_ptr, int* star,
               int* el_of, hipStream_t s);
int umatvec(const et_umesh* m, int mode, const double* E, const uint8_t* cons, const double* u, const double* b,
            double* y, hipStream_t s);
int udiagonal(const et_umesh* m, const double* E, const uint8_t* cons, double* diag, hipStream_t s);
int uenergy(const et_umesh* m, const double* U, const double* xt, double penal, double vmin, int mode, double* out,
            hipStream_t s);
int ufea(const et_umesh* m, int nb, const double* B, const int64_t* B_ptr, const double* D, const int64_t* D_ptr,
         const double* E, const double* U, double* strain, double* stress, double* energy, double* vm,
         hipStream_t s);
// mode 0: y = A x; 1: y = b - A x
int csr_spmv(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals, const double* x, double* y,
             hipStream_t s, int mode = 0, const double* b = nullptr);
int umesh_pattern(const et_umesh* m, int* row_ptr, int* cols, int64_t cap, int64_t* nnz, hipStream_t s);
int uassemble(const et_umesh* m, const double* E, const uint8_t* cons, const int* row_ptr, const int* cols,
              double* vals, hipStream_t s);
KrylovOp csr_op(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals);
KrylovOp unstructured_op(const et_umesh* m, const double* E, const uint8_t* cons);
int cg_solve_op(const KrylovOp& op, const double* b, double* x, double tol, int maxiter, void* work,
                int64_t work_bytes, int* iters, double* residual, hipStream_t s);
int gmres_solve_op(const KrylovOp& op, const double* b, double* x, double tol, int restart, int maxiter,
                   void* work, int64_t work_bytes, int* iters, double* residual, hipStream_t s);
int cg_solve(const Grid& g, const double* K0, const double* E, const uint8_t* cons, const WhtTable* fine, const double* b,
             double* x, double tol, int maxiter, void* work, int64_t work_bytes, int* iters, double* residual,
             hipStream_t s);
int gmres_solve(const Grid& g, const double* K0, const double* E, const uint8_t* cons, const WhtTable* fine,
                const double* b, double* x, double tol, int restart, int maxiter, void* work, int64_t work_bytes,
                int* iters, double* residual, hipStream_t s);

// optimality criteria
int64_t oc_work_size(int64_t n);
// Ae: per-element measures (general meshes) in place of the scalar A, or null
int oc_update(int64_t n, const double* x, const double* dc, double A, double vol, double vf,
              double move, double* xn, void* work, int* steps, hipStream_t s, const double* Ae = nullptr);
// oc_update on a work buffer in a static device array (no allocation); n <= oc_noalloc_max_n()
int oc_update_noalloc(int64_t n, const double* x, const double* dc, double A, double vol, double vf,
                      double move, double* xn, int* steps, hipStream_t s);
int64_t oc_noalloc_max_n();
int sqdist(int64_t n, const double* a, const double* b, double* partial, double* out,
           hipStream_t s);

}  // namespace et
