/*
 * edgeto_amd.h -- C ABI of the MI355X-native EDGE-TO SIMP hot path.
 *
 * Every pointer argument named as device memory is a HIP device pointer owned
 * by the caller; the library never frees caller memory.  It allocates
 * device memory only in et_mg_create (solver workspace) and, once per host
 * thread, a small reduction scratch on the first et_dot_f64 / et_sqdist_f64 /
 * et_oc_update_f64 call.  Every entry point returns an
 * int status (ET_OK = 0); on failure et_last_error() returns a message for the
 * calling thread.  Work is enqueued on the given stream and is asynchronous
 * unless stated otherwise.  No C++ exceptions cross this boundary.
 *
 * Indexing follows the reference exactly (bit-exact target):
 *   nodes     2-D j*nx + i,              3-D j*nx*nz + i*nz + k   (nx = nelx+1, ...)
 *   elements  2-D i*nely + j,            3-D i*nely*nelz + j*nelz + k
 *   DOFs      node*dof + d, interleaved
 * `nel` is the element count per axis (x, y[, z]); `dim` is 2 (Q4, dof 2) or
 * 3 (H8, dof 3).  K0 is the row-major element matrix (8x8 or 24x24).
 * Constraint masks are one byte per DOF (the reference's bool array).
 *
 * Reference citations are paths inside the EDGE-TO repository.
 */
#ifndef EDGETO_AMD_H
#define EDGETO_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ET_OK = 0,
  ET_EINVAL = 1,      /* bad argument (shape, dtype, null pointer, size limit) */
  ET_EHIP = 2,        /* HIP runtime error */
  ET_ENONFINITE = 3,  /* non-finite value detected */
  ET_EBREAKDOWN = 4,  /* PCG breakdown: p^T A p <= 0 */
  ET_ENOCONV = 5,     /* maxiter reached without meeting the tolerance */
  ET_ENOMEM = 6       /* device allocation failed */
};

/* Message of the last failing call on this thread (empty string if none). */
const char* et_last_error(void);

/* Library version string. */
const char* et_version(void);

/* Launch-shape and kernel-choice knobs (process-wide; read by subsequent launches).
 * ET_TUNE_TILED_JC: y planes marched per workgroup by the LDS-tiled fine-grid operator.
 * ET_TUNE_MFMA_JC: the same for the FP16-split matrix-core FP32 operator.
 * ET_TUNE_MIXED_MFMA: 1 = mixed-precision V-cycles use the FP16-split operator
 *   (default 0).  Read by et_mg_create.
 * ET_TUNE_MX_F32 / ET_TUNE_MX_F64: 1 = the fine-grid operator of that precision runs
 *   on the matrix cores at full precision (FP32 / FP64 MFMA), 0 = LDS-tiled vector FMAs.
 *   Read by et_mg_create and by each et_matvec call.
 * ET_TUNE_MX_WGS_F32 / ET_TUNE_MX_WGS_F64: workgroups per compute unit launched by the
 *   matrix-core operator (persistent grid).
 * ET_TUNE_WHT: 1 (default) = the fine-grid operator (both precisions) runs in the corner
 *   Walsh-Hadamard basis (45 FMAs per element); takes precedence over ET_TUNE_MX_*.
 * ET_TUNE_WHT_WGS: workgroups per compute unit launched by that operator for 16-slice
 *   tiles (8-slice tiles launch twice as many).
 * ET_TUNE_COARSE_LIB: factor path of the coarsest grid, read by et_mg_create:
 *   2 (default) = band Cholesky (in-house tile kernels, rocBLAS band updates) and
 *   L^-1 by rocSOLVER trtri, coarse solve by two triangular products; 1 = rocSOLVER
 *   potrf + potri; 0 = in-house dense blocked Cholesky and inverse kernels.
 * ET_TUNE_WHT_THREADS_F32 / _F64: workgroup size of the transformed-basis operator:
 *   1024 or 512 for 16-slice tiles, 512 or 256 for 8-slice tiles (other values map to
 *   the nearer of the two).
 * ET_TUNE_WHT_SLICES_F32 / _F64: 64-element z slices per layer of that operator's x tile
 *   in its residual and Jacobi modes, 16 (8 element columns for nelz <= 128) or 8 (half
 *   the tile: fewer registers per thread, more halo recomputation); defaults 16 (FP32)
 *   and 8 (FP64).  y = A x products always use 16-slice tiles.
 * ET_TUNE_ROW_STENCIL_NODES: coarse levels with at most this many nodes apply their
 *   stencil with one thread per (node, component) instead of one per node.
 * ET_TUNE_LEVEL1_MF: 1 (default) = the first coarse level of 3-D hierarchies with even
 *   element counts is applied matrix-free from the element densities; 0 = stored stencil.
 *   Read by et_mg_create.
 * ET_TUNE_FUSED_RZ: 1 (default) = in PCG with the transformed-basis operator, the
 *   last fine-grid sweep of each V-cycle also sums r.z per workgroup, r taken in the
 *   V-cycle's precision (FP32 V-cycles: their FP32 copy of r); no separate pass over r
 *   and z.  0 = separate reduction with r in FP64.  Read by each solve.
 * ET_TUNE_FILTER_ROWS: y rows of the 3-D filter stencils' tiles (64 elements along z each):
 *   8 (default, 512 threads) or 4 (256 threads). */
enum {
  ET_TUNE_TILED_JC = 0,
  ET_TUNE_MFMA_JC = 1,
  ET_TUNE_MIXED_MFMA = 2,
  ET_TUNE_MX_F32 = 3,
  ET_TUNE_MX_F64 = 4,
  ET_TUNE_MX_WGS_F32 = 5,
  ET_TUNE_MX_WGS_F64 = 6,
  ET_TUNE_WHT = 7,
  ET_TUNE_WHT_WGS = 8,
  ET_TUNE_COARSE_LIB = 9,
  ET_TUNE_WHT_THREADS_F32 = 10,
  ET_TUNE_WHT_THREADS_F64 = 11,
  ET_TUNE_ROW_STENCIL_NODES = 12,
  ET_TUNE_LEVEL1_MF = 13,
  ET_TUNE_FUSED_RZ = 14,
  ET_TUNE_WHT_SLICES_F32 = 15,
  ET_TUNE_WHT_SLICES_F64 = 16,
  ET_TUNE_FILTER_ROWS = 17
};
int et_tune_set(int param, int value);
int et_tune_get(int param);

/* Structure of an element matrix (host memory, FP64, row-major) that the fast
 * 3-D kernels rely on.  *flags receives the OR of
 *   ET_K0_SYMMETRIC  K0 = K0^T;
 *   ET_K0_REFLECT    3-D: with corners in bit order (x + 2y + 4z), every 3x3 block
 *                    K0[a][b] = R_a K0[0][a^b] R_a, R_a = diag((-1)^a_x, (-1)^a_y, (-1)^a_z)
 *                    (isotropic or axis-aligned orthotropic material on any brick);
 *   ET_K0_WHT        3-D: only the couplings the transformed-basis operator keeps are
 *                    non-zero in the corner Walsh-Hadamard basis;
 *   ET_K0_LEVEL1     3-D: the matrix-free first coarse level covers the matrix;
 * each within 1e-12 * max|K0|.  The reference takes any element matrix through its
 * stiffness= callable (geom/_mesh.py:113-135); the library routes a 3-D matrix without
 * SYMMETRIC | REFLECT | WHT to the general node-gather kernels (et_matvec_*,
 * et_residual_*, et_jacobi_*, et_mg_create) and one without LEVEL1 to a stored
 * first coarse level.  No counterpart in the reference. */
enum { ET_K0_SYMMETRIC = 1, ET_K0_REFLECT = 2, ET_K0_WHT = 4, ET_K0_LEVEL1 = 8 };
int et_k0_structure(int dim, const double* K0_host, int* flags);

/* ---------------------------------------------------------------- mesh --- */

/* Element connectivity (nel x 4|8, int32), bit-exact to
 * generate_elements_2d/3d (core/_cuda.py:1089-1135, core/_geom.py:5-59). */
int et_generate_elements(int dim, const int* nel, int32_t* elements, hipStream_t stream);

/* ------------------------------------------------ stiffness operator --- */

/* y = A_hat u with A_hat = M K(E) M + I_c: constrained rows copy u, constrained
 * columns are skipped.  Replaces mat_vec_node_basis_parallel_wcon_cuda_kernel
 * (core/_templated_cuda.cu:782-834) behind StructuredStiffnessKernel._matvec
 * (kernels/_cuda.py:173-204).  cons may be NULL (no constraints, :137-178).
 * 3-D grids (also et_residual_*, et_jacobi_*): K0 is read back to the host to
 * check its structure (et_k0_structure), which synchronises `stream`. */
int et_matvec_f64(int dim, const int* nel, const double* K0, const double* E,
                  const uint8_t* cons, const double* u, double* y, hipStream_t stream);
int et_matvec_f32(int dim, const int* nel, const float* K0, const float* E,
                  const uint8_t* cons, const float* u, float* y, hipStream_t stream);

/* The same operator for 3-D grids on the matrix cores: FP32 vectors, element
 * contraction in FP16 MFMA with a hi/lo split of both operands (~22 bits),
 * FP32 accumulation.  K0_host (host, FP64) gives the power-of-two K0 scale. */
int et_matvec_f32_mfma(int dim, const int* nel, const double* K0_host, const float* K0,
                       const float* E, const uint8_t* cons, const float* u, float* y,
                       hipStream_t stream);

/* r = b - A_hat x (the V-cycle residual at the fine level; the reference forms it
 * as b - A @ x with the matrix-free operator at level 0, solvers/_multigrid_cuda.py:143). */
int et_residual_f64(int dim, const int* nel, const double* K0, const double* E,
                    const uint8_t* cons, const double* x, const double* b, double* r,
                    hipStream_t stream);
int et_residual_f32(int dim, const int* nel, const float* K0, const float* E,
                    const uint8_t* cons, const float* x, const float* b, float* r,
                    hipStream_t stream);

/* One damped Jacobi sweep, y = x + omega * Dinv * (b - A_hat x), fused in one pass.
 * Replaces CuMultiGrid._jacobi_smoother (solvers/_multigrid_cuda.py:59-62) at the fine
 * level (x + omega * (b - A @ x) / D).  Dinv is the inverse of diag(A_hat). */
int et_jacobi_f64(int dim, const int* nel, const double* K0, const double* E,
                  const uint8_t* cons, const double* x, const double* b, const double* Dinv,
                  double omega, double* y, hipStream_t stream);
int et_jacobi_f32(int dim, const int* nel, const float* K0, const float* E,
                  const uint8_t* cons, const float* x, const float* b, const float* Dinv,
                  double omega, float* y, hipStream_t stream);

/* diag(A_hat); constrained entries are 1.  Replaces
 * get_diagonal_node_basis_cuda_kernel (core/_templated_cuda.cu:1068-1113). */
int et_diagonal_f64(int dim, const int* nel, const double* K0, const double* E,
                    const uint8_t* cons, double* diag, hipStream_t stream);

/* dk_e = -u_e^T K0 u_e on the unmasked U.  Replaces process_dk_kernel_cuda
 * (core/_templated_cuda.cu:180-194) behind process_grad (kernels/_cuda.py:152-156). */
int et_elem_energy_f64(int dim, const int* nel, const double* K0, const double* U,
                       double* dk, hipStream_t stream);

/* Fused SIMP sensitivity before the filter transpose:
 * out_e = p * clip(xt_e, void, 1)^(p-1) * (-u_e^T K0 u_e).  Replaces the pair
 * process_dk + SingleMaterial.__backward__ (MaterialModels.py:62-73) used in
 * CuTopOpt.solve_and_process (TopOpt/_cuda.py:293-295). */
int et_sensitivity_f64(int dim, const int* nel, const double* K0, const double* U,
                       const double* xt, double penal, double void_, double* out,
                       hipStream_t stream);

/* Element post-processing of a displacement U (n_dof, constrained DOFs included):
 *   strain_e = B u_e, stress_e = D strain_e, energy_e = 1/2 E_e u_e^T K0 u_e, and the
 *   von Mises stress of stress_e (plane stress in 2-D).
 * D: NB x NB and B: NB x (4|8)*dim row-major device arrays (NB = 3 in 2-D, 6 in 3-D;
 * strain order xx, yy[, zz], xy[, yz, zx] with engineering shears).  Outputs:
 * strain, stress (n_elem x NB), energy (n_elem), von_mises (n_elem; may be NULL).
 * Replaces FEA_Integrals_node_basis_parallel_cuda_kernel (core/_templated_cuda.cu:1681-1719)
 * and the von Mises step of TopOpt.FEA_integrals (TopOpt/_cuda.py:78-129). */
int et_fea_integrals_f64(int dim, const int* nel, const double* K0, const double* D, const double* B,
                         const double* E, const double* U, double* strain, double* stress,
                         double* energy, double* von_mises, hipStream_t stream);

/* E = clip(xt^p, void, 1) (plain: clip(xt, void, 1)), MaterialModels.py:47-60. */
int et_simp_forward_f64(int64_t n, const double* xt, double penal, double void_, int plain,
                        double* E, hipStream_t stream);

/* ------------------------------------------------------ density filter --- */
/* offsets: device int32 (n_off x dim); w: device weights (n_off).
 * Normaliser N_i = sum of in-bounds weights: get_filter_{2D,3D}_weights_kernel
 * (core/_templated_cuda.cu:1585-1630). */
int et_filter_normalizer_f64(int dim, const int* nel, int n_off, const int32_t* offsets,
                             const double* w, double* norm, hipStream_t stream);
/* y = H x: apply_filter_{2D,3D}_kernel (core/_templated_cuda.cu:1526-1583). */
int et_filter_fwd_f64(int dim, const int* nel, int n_off, const int32_t* offsets,
                      const double* w, const double* x, double* y, hipStream_t stream);
/* y = H^T v: apply_filter_{2D,3D}_transpose_kernel (core/_templated_cuda.cu:1632-1679). */
int et_filter_adj_f64(int dim, const int* nel, int n_off, const int32_t* offsets,
                      const double* w, const double* norm, const double* v, double* y,
                      hipStream_t stream);

/* The SIMP adjoint fused with the filter transpose, one kernel:
 *   out = H^T s,  s_e = p * clip(xt_e, void, 1)^(p-1) * (-u_e^T K0 u_e),
 * i.e. et_sensitivity_f64 followed by et_filter_adj_f64 (CuTopOpt.solve_and_process,
 * TopOpt/_cuda.py:293-300) with s_e / N_e formed once per element while it is staged
 * in LDS.  U: n_dof displacement (constrained DOFs included); xt: filtered design;
 * offsets, w, norm: the filter's table and normaliser.  k0_flags: et_k0_structure of
 * the same K0 (host check, once per matrix); with ET_K0_WHT set the element energy is
 * evaluated in the corner Walsh-Hadamard basis (~140 flops instead of 1152), otherwise as
 * the full product.  Asynchronous (no host round trip). */
int et_simp_adjoint_f64(int dim, const int* nel, const double* K0, const double* U, const double* xt,
                        double penal, double void_, int n_off, const int32_t* offsets, const double* w,
                        const double* norm, int k0_flags, double* out, hipStream_t stream);

/* ------------------------------------------------- multigrid transfers --- */
/* Restriction fine -> coarse (nel = fine element counts): restriction_{2d,3d}_kernel
 * (core/_templated_cuda.cu:232-364) behind apply_restriction_cuda (core/_mgm_cuda.py:223-268). */
int et_restrict_f64(int dim, const int* nel, int dof, const double* vf, double* vc,
                    hipStream_t stream);
/* Prolongation coarse -> fine: prolongation_{2d,3d}_kernel (core/_templated_cuda.cu:366-475)
 * behind apply_prolongation_cuda (core/_mgm_cuda.py:270-314). */
int et_prolong_f64(int dim, const int* nel, int dof, const double* vc, double* vf,
                   hipStream_t stream);

/* Galerkin coarse operators as dense block stencils.  Layout: for the 3^dim
 * neighbour offsets o (o = (di+1)*9 + (dj+1)*3 + (dk+1) in 3-D, (di+1)*3 + (dj+1)
 * in 2-D) and dof x dof block entries (r, c), plane (o*dof + r)*dof + c holds one
 * value per node: S[((o*dof + r)*dof + c) * n_nodes + node].
 * et_stencil_from_operator: stencil of A_hat on the fine grid (nel).
 * et_galerkin: S_coarse = R S_fine R^T (nel = fine element counts), replacing
 * get_restricted_{2d,3d}_{l0,l1p}_nnz_based (core/_mgm_cuda.cu:95-344, 430-649). */
int et_stencil_from_operator_f64(int dim, const int* nel, const double* K0, const double* E,
                                 const uint8_t* cons, double* S, hipStream_t stream);
int et_galerkin_f64(int dim, const int* nel, int dof, const double* S_fine, double* S_coarse,
                    hipStream_t stream);
/* First coarse operator R A_hat R^T built directly from the element densities
 * (no fine-grid stencil), for even element counts.  K0_host is the element
 * matrix in host memory.  Allocates and frees temporary device workspace and
 * synchronises `stream`. */
int et_galerkin_elements_f64(int dim, const int* nel, const double* K0_host, const double* E,
                             const uint8_t* cons, double* S_coarse, hipStream_t stream);
/* y = S x for a stencil operator on grid nel. */
int et_stencil_apply_f64(int dim, const int* nel, int dof, const double* S, const double* x,
                         double* y, hipStream_t stream);

/* ------------------------------------------------------- vector algebra --- */
/* *out = a . b (device scalar), deterministic two-stage reduction.
 * Replaces the cuBLAS dot / nrm2 calls of CuMultiGrid.solve (solvers/_multigrid_cuda.py:214-231). */
int et_dot_f64(int64_t n, const double* a, const double* b, double* out, hipStream_t stream);

/* *out = sum (a - b)^2 (device scalar); the RMS design change of the SIMP loop
 * (TopOpt/_cuda.py:350) is sqrt(*out / n). */
int et_sqdist_f64(int64_t n, const double* a, const double* b, double* out, hipStream_t stream);

/* ------------------------------------------------------ multi-material --- */
/* PenalizedMultiMaterial (rp = 0) / PenalizedMultiMaterialRP (rp = 1), MaterialModels.py:100-369,
 * for m <= 4 materials: rho n x m (row-major, device), E_mat m moduli (device).
 * mode 0: out[n] = E (__forward__); 1: out[n] = clip(rho . E_mat, void, 1) (plain);
 * 2: out[n x m] = dE/drho (__backward__). */
int et_multimat_f64(int64_t n, int m, const double* rho, const double* E_mat, double penal, double void_, int rp,
                    int mode, double* out, hipStream_t stream);
/* OC update with one multiplier per material, bisected in lockstep until every bracket closes
 * (TopOpt/_cuda.py:242-280 with n_sets = m; constraint PenalizedMultiMaterial.__con__).
 * x, dc, x_new: n x m device arrays; mass (host): m fractions, or 1 when rho_mat (host, m
 * material densities) is given (then one mass constraint for all materials).  work:
 * et_oc_multi_work_size(n, m) bytes, 8-byte aligned.  Synchronises `stream`. */
int64_t et_oc_multi_work_size(int64_t n, int m);
int et_oc_update_multi_f64(int64_t n, int m, const double* x, const double* dc, double A, double vol, double move,
                           const double* mass, const double* rho_mat, double* x_new, void* work, int* steps,
                           hipStream_t stream);

/* ------------------------------------------------------ Krylov solvers --- */
/* Unpreconditioned solvers of A_hat x = b on the matrix-free operator (the
 * iterations CuCG / CuGMRES run through cupyx cg / gmres, solvers/_gpu.py:11-107).
 * x holds the initial guess on entry.  atol = tol |b|.  work: device buffer of at
 * least et_krylov_work_size(n_dof, restart) bytes (restart 0 for CG), given in
 * work_bytes.  Synchronise `stream`.  *iters: CG iterations, or GMRES inner steps
 * (restart per cycle); *residual: |b - A x| / |b| of the returned x.  Status
 * ET_ENOCONV when maxiter ends the loop above the tolerance.
 * et_cg_f64: rho = r.r, p = r + (rho/rho_old) p, alpha = rho / p.A p, stop on |r| <= atol.
 * et_gmres_f64: cycles of `restart` Arnoldi steps (one classical Gram-Schmidt pass),
 * least squares by Givens rotations, convergence tested once per cycle. */
int64_t et_krylov_work_size(int64_t n_dof, int restart);
int et_cg_f64(int dim, const int* nel, const double* K0, const double* E, const uint8_t* cons,
              const double* b, double* x, double tol, int maxiter, void* work, int64_t work_bytes,
              int* iters, double* residual, hipStream_t stream);
int et_gmres_f64(int dim, const int* nel, const double* K0, const double* E, const uint8_t* cons,
                 const double* b, double* x, double tol, int restart, int maxiter, void* work,
                 int64_t work_bytes, int* iters, double* residual, hipStream_t stream);

/* ------------------------------------------------------ general meshes --- */
/* Unstructured meshes with element matrices stored per element and any mix of
 * element types (T3 / Q4 in 2-D, T4 / H8 in 3-D): CuGeneralMesh and the
 * Uniform / General stiffness kernels (geom/_mesh.py:146-253,
 * kernels/_cuda.py:305-815).  All arrays are device memory:
 *   elem_ptr   n_el + 1 offsets into elem_nodes (the reference's elements_ptr);
 *   elem_nodes the elements' node lists, concatenated (elements_flat);
 *   K, K_ptr   row-major (size * dof)^2 element matrices, concatenated, and their
 *              n_el + 1 offsets (Ks of a uniform mesh is K_ptr[e] = e * (size*dof)^2);
 *   star_ptr   n_nodes + 1 offsets into star (node_ids, with the end appended);
 *   star       positions into elem_nodes, ascending within each node: the stable
 *              argsort of elem_nodes (sorter);
 *   el_of      the element of each position of elem_nodes (el_ids).
 * DOFs are interleaved node * dof + d.  Operators follow the structured ones:
 * constrained rows are identity rows and constrained columns are skipped
 * (mat_vec_node_basis_parallel_{full,flat}_wcon, core/_templated_cuda.cu:669-780). */
typedef struct et_umesh {
  int64_t n_nodes;
  int64_t n_el;
  int dof;
  const int* elem_ptr;
  const int* elem_nodes;
  const double* K;
  const int64_t* K_ptr;
  const int* star_ptr;
  const int* star;
  const int* el_of;
} et_umesh;

/* Build star_ptr (n_nodes + 1), star and el_of (elem_ptr[n_el] entries each) on the
 * device from elem_ptr / elem_nodes; node ids outside [0, n_nodes) are ET_EINVAL.
 * Allocates temporaries and synchronises `stream` (setup; kernels/_cuda.py:320-331). */
int et_umesh_star(int64_t n_nodes, int64_t n_el, const int* elem_ptr, const int* elem_nodes,
                  int* star_ptr, int* star, int* el_of, hipStream_t stream);
/* y = A(E) u (mat_vec_node_basis_parallel_{full,flat}_wcon); cons may be NULL. */
int et_umatvec_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const double* u,
                   double* y, hipStream_t stream);
/* diag of A(E), 1 on constrained DOFs (get_diagonal_node_basis_{full,flat}, :1116-1213). */
int et_udiagonal_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, double* diag,
                     hipStream_t stream);
/* dk_e = -u_e^T K_e u_e (process_dk_{full,flat}, :196-231). */
int et_uelem_energy_f64(const et_umesh* mesh, const double* U, double* dk, hipStream_t stream);
/* p * clip(xt_e, void, 1)^(p-1) * (-u_e^T K_e u_e) (process_dk + SingleMaterial.__backward__). */
int et_usensitivity_f64(const et_umesh* mesh, const double* U, const double* xt, double penal,
                        double void_, double* out, hipStream_t stream);
/* strain_e = B_e u_e, stress_e = D_e strain_e, energy_e = 1/2 E_e u_e^T K_e u_e and von
 * Mises (FEA_Integrals_node_basis_parallel_{full,flat}, :1722-1800; TopOpt/_cuda.py:78-129).
 * B / D: row-major nb x (size*dof) and nb x nb per element with offsets B_ptr / D_ptr
 * (n_el + 1, int64); nb = 3 (2-D) or 6 (3-D); vm may be NULL. */
int et_ufea_integrals_f64(const et_umesh* mesh, int nb, const double* B, const int64_t* B_ptr,
                          const double* D, const int64_t* D_ptr, const double* E, const double* U,
                          double* strain, double* stress, double* energy, double* vm,
                          hipStream_t stream);
/* y = A x for a CSR matrix (row_ptr n_rows + 1, int32 columns): the general-mesh density
 * filter and its transpose (CuGeneralFilter, geom/_filters.py:278-298, as a stored CSR
 * transpose).  Each row sums in a fixed order. */
int et_csr_spmv_f64(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals,
                    const double* x, double* y, hipStream_t stream);
/* The assembled operator as CSR, one row per DOF, int32 offsets and columns (the
 * reference's construct(), kernels/_cuda.py:407-420 with matmat_*_wcon, :837-996):
 * row (i, k) lists the DOFs (j, d) of every node j sharing an element with node i, node
 * ids ascending.  et_umesh_pattern fills row_ptr (n_nodes * dof + 1) and *nnz, and cols
 * when cap >= *nnz (call it once with cols = NULL to size cols); synchronises `stream`.
 * et_uassemble_f64 writes vals (nnz) for densities E: A_hat = M K(E) M + I_c (constrained
 * rows and columns zero, 1 on the constrained diagonal), summed in star order. */
int et_umesh_pattern(const et_umesh* mesh, int* row_ptr, int* cols, int64_t cap, int64_t* nnz,
                     hipStream_t stream);
int et_uassemble_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const int* row_ptr,
                     const int* cols, double* vals, hipStream_t stream);
/* CG / GMRES(restart) on a CSR matrix (n_rows unknowns): the solvers of
 * et_cg_f64 / et_gmres_f64 on an assembled operator (CuCG / CuGMRES with
 * matrix_free=False, solvers/_gpu.py:40-52). */
int et_csr_cg_f64(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals,
                  const double* b, double* x, double tol, int maxiter, void* work, int64_t work_bytes,
                  int* iters, double* residual, hipStream_t stream);
int et_csr_gmres_f64(int64_t n_rows, const int* row_ptr, const int* cols, const double* vals,
                     const double* b, double* x, double tol, int restart, int maxiter, void* work,
                     int64_t work_bytes, int* iters, double* residual, hipStream_t stream);
/* CG / GMRES(restart) on the general-mesh operator: et_cg_f64 / et_gmres_f64 semantics
 * (CuCG / CuGMRES with a Uniform or General stiffness kernel, solvers/_gpu.py:11-107). */
int et_ucg_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const double* b, double* x,
               double tol, int maxiter, void* work, int64_t work_bytes, int* iters, double* residual,
               hipStream_t stream);
int et_ugmres_f64(const et_umesh* mesh, const double* E, const uint8_t* cons, const double* b,
                  double* x, double tol, int restart, int maxiter, void* work, int64_t work_bytes,
                  int* iters, double* residual, hipStream_t stream);

/* ------------------------------------------------- optimality criteria --- */
/* One OC design update (TopOpt/_cuda.py:242-280 with SingleMaterial.__ocP__ /
 * __con__, MaterialModels.py:75-90): bisection on lambda in [1e-9, 1e9] until
 * (l2-l1)/(l2+l1) <= 1e-4; the returned design is the one at the last midpoint.
 * A = element measure (structured: one value), vol = A * n, vf = volume fraction.
 * Runs entirely on the device (a fixed sequence of kernels, no host round trip):
 * with steps == NULL the call is asynchronous; with steps != NULL it synchronises
 * `stream` and *steps receives the number of bisection steps.  work: device buffer
 * of et_oc_work_size(n) bytes, 8-byte aligned; it begins with the bisection state
 * {double l1, l2, lmid, sabs; int32 done, steps} (40 bytes), so the step count
 * can also be read from work + 36 once the stream has passed the update. */
int64_t et_oc_work_size(int64_t n);
int et_oc_update_f64(int64_t n, const double* x, const double* dc, double A, double vol,
                     double vf, double move, double* x_new, void* work, int* steps,
                     hipStream_t stream);

/* et_oc_update_f64 without a work argument and without any device allocation:
 * the work buffer is a static device array of the library (one per GPU, kept for
 * the process; about 8.1 MiB), so n must not exceed et_oc_noalloc_max_n()
 * (1 048 576 elements; larger n returns ET_EINVAL before any launch).  steps as in
 * et_oc_update_f64 (NULL: asynchronous, no host synchronisation).  Calls that
 * overlap on one GPU share the array and race; each call re-initialises what it
 * reads.  No counterpart in the reference: for callers that cannot allocate
 * (code under graph capture, serving loops, allocation-free harnesses). */
/* et_oc_update_f64 with one element measure per element (A_elem, n values: a general
 * mesh's As, geom/_mesh.py:195-214) in place of the scalar A: ocP = x sqrt(-dc / A_e)
 * and the volume sum(x_new A_e) / vol (MaterialModels.py:75-90). */
int et_oc_update_elem_f64(int64_t n, const double* x, const double* dc, const double* A_elem,
                          double vol, double vf, double move, double* x_new, void* work, int* steps,
                          hipStream_t stream);

int et_oc_update_f64_noalloc(int64_t n, const double* x, const double* dc, double A, double vol,
                             double vf, double move, double* x_new, int* steps, hipStream_t stream);
int64_t et_oc_noalloc_max_n(void);

/* ------------------------------------------------------- GMG-PCG solver --- */
/* Opaque solver: the CuMultiGrid object (solvers/_multigrid_cuda.py:15-242) with
 * its hierarchy, workspace and coarse factor.  precision: 0 = FP64 throughout;
 * 1 = mixed (FP32 V-cycle, FP64 outer PCG, coarse solve and residual).
 * cycle_w != 0 selects the 'w' branch for levels whose bit is set in w_level_mask
 * (the reference only takes it for a lowercase 'w', :165). */
typedef struct et_mg et_mg;

typedef struct et_mg_params {
  int dim;
  int nel[3];
  int n_level;       /* number of coarsenings (grids = n_level + 1) */
  int n_smooth;      /* damped-Jacobi sweeps before and after the coarse correction */
  double omega;      /* reference `omega` (max omega) */
  double min_omega;  /* reference `min_omega`; min = max(omega/2, min_omega) */
  int cycle_w;
  int w_level_mask;
  int precision;
} et_mg_params;

int et_mg_create(et_mg** handle, const et_mg_params* params, const double* K0_host,
                 hipStream_t stream);
int et_mg_destroy(et_mg* handle);

/* Solve A_hat(E) x = b by GMG-preconditioned CG (solve(), :188-242).  E, cons,
 * b, x are device pointers (E: n_elem, cons: n_dof bytes, b, x: n_dof).  x holds
 * the initial guess on entry (the caller zeroes it when use_last is false) and
 * the solution on exit.  The hierarchy is rebuilt from E on every call, as the
 * reference does.  Synchronises `stream`.  Outputs: iterations, |r|/|b|. */
int et_mg_solve(et_mg* handle, const double* E, const uint8_t* cons, const double* b,
                double* x, double tol, int maxiter, int* iters, double* residual);

/* y = A_l x with the operator of multigrid level l (0 = fine grid, up to n_level - 1)
 * as the last et_mg_solve built it, in the V-cycle's precision (double for precision
 * 0, float for 1); x and y are device vectors of that level.  No counterpart in the
 * reference: exposes each level's operator for verification.  Synchronises. */
int et_mg_apply_level(et_mg* handle, int level, const void* x, void* y);

/* x = A_L^-1 b on the coarsest grid L = n_level of the hierarchy built by the
 * last et_mg_solve, in the V-cycle's precision (double for precision 0, float
 * for 1).  No counterpart in the reference (its coarsest-grid solve is
 * scipy's splu inside solvers/_multigrid_cuda.py:120-131): exposes the direct
 * solve for verification.  Synchronises. */
int et_mg_coarse_solve(et_mg* handle, const void* b, void* x);

/* Per-phase timings of the last solve in milliseconds (setup, pcg), via HIP events. */
int et_mg_last_timings(et_mg* handle, double* setup_ms, double* pcg_ms);

/* Kernel-class profiling with HIP events recorded around each launch on the
 * solver's stream.  et_mg_profile(h, mask) clears the records and times the
 * classes whose bit (1 << ET_PROF_*) is set in mask: -1 times every class, 0
 * disables.  Each timed launch costs two event records on the stream, so time
 * only the classes needed inside a measured region.
 * et_mg_profile_read synchronises and fills ms[k] / launches[k] for the
 * ET_PROF_KINDS classes below (accumulated since the last enable). */
enum {
  ET_PROF_L0_JAC = 0,      /* fine-grid damped-Jacobi sweep  y = x + w D^-1 (b - A x)   */
  ET_PROF_L0_RES = 1,      /* fine-grid residual             y = b - A x               */
  ET_PROF_L0_JACADD = 2,   /* last post-smoothing sweep fused with x += e             */
  ET_PROF_L0_SCALE = 3,    /* first sweep from a zero guess  y = w D^-1 b              */
  ET_PROF_MATVEC64 = 4,    /* outer FP64 operator (PCG q = A p, initial residual)     */
  ET_PROF_COARSE_OPS = 5,  /* stencil sweeps / residuals on levels >= 1                */
  ET_PROF_TRANSFER = 6,    /* restriction and prolongation                            */
  ET_PROF_COARSE_SOLVE = 7,/* dense coarsest-grid solve                               */
  ET_PROF_PCG_VEC = 8,     /* PCG dots, updates and precision conversions             */
  ET_PROF_SETUP = 9,       /* hierarchy construction and coarse factorisation         */
  ET_PROF_L0_JAC0 = 10,    /* first fine-grid sweep fused with its zero-guess scaling  */
  ET_PROF_KINDS = 11
};
int et_mg_profile(et_mg* handle, int enable);
int et_mg_profile_read(et_mg* handle, double* ms, int64_t* launches);

/* ------------------------------------------------------ batched engine --- */
/* Many independent structured SIMP problems optimised concurrently on one GPU
 * (the batched workload: each problem is CuTopOpt.optimize, TopOpt/_cuda.py:304-382,
 * with its own mesh, BCs, loads, filter radius, volume fraction and GMG depth).
 * The engine copies every array at creation and keeps each problem's state and its
 * GMG-PCG solver resident; et_batch_run runs, for every problem, the initial solve and
 * n_iter OC iterations, n_workers host threads each driving one problem at a time on
 * that problem's stream (largest problems first).  Between two solves a problem issues
 * one captured graph (fused SIMP adjoint, OC update, design change, filter, SIMP map),
 * captured the first time its pair of penalties occurs and replayed.  Synchronous.  No
 * counterpart in the reference (it optimises one problem per process). */
typedef struct et_batch et_batch;
typedef struct et_batch_problem {
  int dim;
  int nel[3];
  const double* K0;        /* host: element matrix ((4|8) dim)^2, row-major */
  const uint8_t* cons;     /* host: n_dof constraint bytes */
  const double* f;         /* host: n_dof loads */
  const double* x0;        /* host: n_elem initial design (clipped to [void, 1]) */
  int n_off;               /* filter taps */
  const int32_t* offsets;  /* host: n_off x dim */
  const double* weights;   /* host: n_off normalised hat weights */
  double A;                /* element measure */
  double volfrac, penal, void_, move, tol;
  int maxiter, n_level, n_smooth;
  /* transformed-basis tiles of this problem's sweeps (0 = library default) */
  int tile_threads_f32, tile_threads_f64, tile_slices_f32, tile_slices_f64;
  /* penalty continuation (NULL / 0: penal throughout): the densities of OC iteration k
   * use penal_schedule[min(k, n_schedule - 1)], those of the initial solve entry 0 --
   * the material model's penalty_schedule(p, k) evaluated by the caller
   * (MaterialModels.py:47-60) */
  const double* penal_schedule;
  int n_schedule;
  /* stop the problem after OC iteration k once the RMS design change is below ch_tol,
   * the relative compliance change below fun_tol and iteration k's penalty equals penal
   * (CuTopOpt.optimize, TopOpt/_cuda.py:366-373; zero tolerances: all n_iter iterations) */
  double ch_tol, fun_tol;
} et_batch_problem;
int et_batch_create(et_batch** handle, int n, const et_batch_problem* problems, int precision);
/* *wall_ms: host wall time of the whole run. */
int et_batch_run(et_batch* handle, int n_iter, int n_workers, double* wall_ms);
/* Results of problem i of the last run, for the k OC iterations it ran (k = n_iter
 * unless it stopped early): compliance (k + 1), PCG iterations (k + 1), wall time of
 * each OC iteration in ms (k), final design (n_elem, host).  Any pointer may be NULL. */
int et_batch_result(et_batch* handle, int i, double* compliance, int* pcg_iterations,
                    double* iteration_ms, double* x_host);
/* OC iterations problem i ran in the last run and whether it stopped on the convergence
 * test (the flag that CuTopOpt.optimize returns).  Any pointer may be NULL. */
int et_batch_status(et_batch* handle, int i, int* n_iterations, int* converged);
int et_batch_destroy(et_batch* handle);

#ifdef __cplusplus
}
#endif
#endif /* EDGETO_AMD_H */
