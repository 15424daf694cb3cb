/*
 * libcapytaine_amd -- C-ABI of the MI355X (gfx950) implementation of the
 * Capytaine boundary-element hot path: Delhommeau Green-function
 * evaluation, assembly of the influence matrices S and K, and their dense
 * LU solve.
 *
 * Conventions
 *   - Every function returns 0 (CPT_OK) on success or a negative CPT_ERR_*
 *     code; cpt_last_error() then describes the failure (thread-local).
 *   - Array arguments are DEVICE pointers (hipMalloc'd or torch CUDA/HIP
 *     tensors), owned by the caller, unless the parameter says "host".
 *   - Work is enqueued on the caller's stream and is asynchronous unless
 *     stated otherwise.  A handle owns only workspace and the device copy
 *     of the tabulation; one handle per host thread (or per stream).
 *   - Complex numbers are interleaved (re, im) doubles: numpy complex128.
 *   - Matrices are column-major ("Fortran order"), as in the reference.
 *   - Point/vector arrays are row-major (n, 3): numpy's default layout.
 *
 * Reference citations are relative to the capytaine 3.0.0 source tree.
 */
#ifndef CAPYTAINE_AMD_H
#define CAPYTAINE_AMD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPT_ABI_VERSION 1

enum {
    CPT_OK = 0,
    CPT_ERR_ARG = -1,         /* invalid argument */
    CPT_ERR_HIP = -2,         /* HIP runtime error */
    CPT_ERR_ALLOC = -3,       /* device allocation failed */
    CPT_ERR_STATE = -4,       /* call order violated (e.g. no tabulation) */
    CPT_ERR_UNSUPPORTED = -5, /* option of the reference not implemented here */
    CPT_ERR_NOT_CONVERGED = -6 /* iterative solve left above its tolerance (outputs still written) */
};

/* Values of the integer options, equal to the reference's Fortran
 * enumerations (libs/libDelhommeau/src/constants.f90:40-52). */
enum { CPT_LEGACY_GRID = 0, CPT_SCALED_NEMOH3_GRID = 1 };
/* finite_depth_method (constants.f90:45-47); FinGreen3D has an entry point of
 * its own, cpt_build_matrices_fingreen3d */
enum { CPT_FINGREEN3D_METHOD = -1, CPT_LEGACY_FINITE_DEPTH = 0, CPT_NEWER_FINITE_DEPTH = 1 };
enum { CPT_HIGH_FREQ = 0, CPT_LOW_FREQ = 1, CPT_LOW_FREQ_WITH_RANKINE_PART = 2 };

typedef struct cpt_context* cpt_handle;

/* Last error message of the calling thread ("" if none). */
const char* cpt_last_error(void);
/* CPT_ABI_VERSION of the library. */
int cpt_version(void);

/* Create / destroy a context bound to HIP device `device`.
 * Replaces the module state of the f2py extension
 * capytaine.green_functions.Delhommeau_float64 (meson.build:39-47,96-112). */
int cpt_create(int device, cpt_handle* out);
int cpt_destroy(cpt_handle h);

/* Tabulated Delhommeau integrals I1..I5 on the grid (r_range, z_range):
 * writes tab, shape (nr, nz, 5) column-major, the reference table layout.
 * Replaces delhommeau_integrals.construct_tabulation
 * (libs/libDelhommeau/src/Delhommeau_integrals.f90:190-204, with
 * numerical_integration :43-153), called from
 * src/capytaine/green_functions/delhommeau.py:259-261. */
int cpt_tabulate(cpt_handle h, const double* r_range, int nr, const double* z_range, int nz,
                 int nb_integration_points, double* tab, hipStream_t stream);

/* Copy a tabulation into the context (re-laid out for the assembly kernel).
 * Takes the arguments the reference passes to build_matrices as
 * `*self.all_tabulation_parameters` (delhommeau.py:263-267, 505):
 * nb_integration_points, grid_shape, r_range, z_range, tabulated_integrals.
 * nr <= 1 or nz <= 1 selects the untabulated evaluation (the pointers may
 * then be NULL): the wave term integrates over theta with
 * nb_integration_points Simpson points for each point, as the reference
 * does for such a table (Green_wave.f90:226-228), with the device's
 * transcendental functions.  Synchronises `stream`. */
int cpt_set_tabulation(cpt_handle h, const double* r_range, int nr, const double* z_range, int nz,
                       const double* tab, int grid_shape, int nb_integration_points, hipStream_t stream);

/* Assemble S and K for nb_wavenumbers wavenumbers in one launch.
 * Replaces matrices.build_matrices (libs/libDelhommeau/src/matrices.f90:28-325)
 * followed by matrices.add_diagonal_term (:329-364) when
 * diagonal_term_in_double_layer, and matrices.add_rankine_term_only
 * (:368-432) when free_surface is +inf; called from
 * src/capytaine/green_functions/delhommeau.py:498-515 and :377-390.
 *
 *   collocation_points  (n,3); dot_product_normals (n,3) if adjoint else (m,3)
 *   vertices (nv,3); faces (m,4) int32, 0-based (the reference passes faces+1)
 *   centers, normals (m,3); areas, radiuses (m)
 *   quad_points (m,nq,3); quad_weights (m,nq)
 *   wavenumbers, gf_singularities: HOST arrays (nb); gf_singularities after
 *       the overrides of delhommeau.py:466-490
 *   depth: +inf for infinite depth; finite_depth_method: CPT_*_FINITE_DEPTH
 *   prony: HOST array (nb, 2, nexp) of (lambda_k, a_k), the reference's
 *       prony_decomposition (delhommeau.py:492-495); unused in infinite depth
 *   early_dot_product: K is (n,m) if non-zero, else (3,n,m)
 *   S: complex128, S[i + ldS*j] (+ b*strideS for wavenumber b), OVERWRITTEN
 *   K: complex128, K[c + nk*(i + ldK*j)] (+ b*strideK), nk = 1 or 3, OVERWRITTEN
 * (the reference accumulates into caller-zeroed arrays: same values). */
int cpt_build_matrices(cpt_handle h, int nb_collocation_points, const double* collocation_points,
                       const double* dot_product_normals, int nb_vertices, const double* vertices, int nb_faces,
                       const int32_t* faces, const double* centers, const double* normals, const double* areas,
                       const double* radiuses, int nb_quad_points, const double* quad_points,
                       const double* quad_weights, int nb_wavenumbers, const double* wavenumbers,
                       const int* gf_singularities, double depth, int finite_depth_method, int nexp,
                       const double* prony, int adjoint_double_layer, int early_dot_product,
                       int diagonal_term_in_double_layer, double free_surface, void* S, int64_t ldS,
                       int64_t strideS, void* K, int64_t ldK, int64_t strideK, hipStream_t stream);

/* cpt_build_matrices with the per-wavenumber arguments in device memory:
 * wavenumbers (nb_wavenumbers doubles), gf_singularities (nb_wavenumbers
 * ints) and prony (nb_wavenumbers x 2 x nexp doubles, as above) are device
 * pointers read by the kernels when they run.  Nothing from host memory
 * enters the stream, so a call captured in a HIP graph can be replayed with
 * new values written to those arrays (the per-frequency graphs of the
 * sweeps).  The wavenumbers are not checked on the host: each must be 0,
 * positive or +inf, and gf_singularities one of the CPT_*_FREQ values; a
 * table must be set unless free_surface is infinite.  The same buffers of
 * the context as cpt_build_matrices are reused: a first call of the same
 * sizes outside the capture makes sure the capture allocates nothing.
 * No counterpart in the reference (its arguments are host arrays). */
int cpt_build_matrices_dev(cpt_handle h, int nb_collocation_points, const double* collocation_points,
                           const double* dot_product_normals, int nb_vertices, const double* vertices, int nb_faces,
                           const int32_t* faces, const double* centers, const double* normals, const double* areas,
                           const double* radiuses, int nb_quad_points, const double* quad_points,
                           const double* quad_weights, int nb_wavenumbers, const double* wavenumbers,
                           const int* gf_singularities, double depth, int finite_depth_method, int nexp,
                           const double* prony, int adjoint_double_layer, int early_dot_product,
                           int diagonal_term_in_double_layer, double free_surface, void* S, int64_t ldS,
                           int64_t strideS, void* K, int64_t ldK, int64_t strideK, hipStream_t stream);

/* The HAMS finite-depth Green function FinGreen3D: replaces build_matrices
 * with finite_depth_method = FINGREEN3D_METHOD (matrices.f90:95-209,
 * Green_wave.f90:545-616) as FinGreen3D.evaluate calls it (hams.py:158-209),
 * one wavenumber per call.  Mesh arrays, S and K as for cpt_build_matrices
 * (device memory, overwritten, strides ldS and ldK); roots: HOST array of
 * nb_roots >= 100 doubles, the roots of the dispersion relation (roots[0] =
 * wavenumber, then the evanescent ones, hams.py:152-156).  depth finite and
 * free_surface finite (the reference raises otherwise), wavenumber finite and
 * positive.  complex128 only.  S = -1/(4 pi) (Rankine + sign * free-surface
 * image + sum_q w_q G_wave), K likewise with the gradient, the diagonal term
 * as in cpt_build_matrices. */
int cpt_build_matrices_fingreen3d(cpt_handle h, int nb_collocation_points, const double* collocation_points,
                                  const double* dot_product_normals, int nb_vertices, const double* vertices,
                                  int nb_faces, const int32_t* faces, const double* centers, const double* normals,
                                  const double* areas, const double* radiuses, int nb_quad_points,
                                  const double* quad_points, const double* quad_weights, double wavenumber,
                                  double depth, int nb_roots, const double* roots, int gf_singularities,
                                  int adjoint_double_layer, int early_dot_product, int diagonal_term_in_double_layer,
                                  double free_surface, void* S, int64_t ldS, void* K, int64_t ldK, hipStream_t stream);

/* fingreen3d_routine (FinGreen3D.f90:71-117) at np points: rr, zf, zp device
 * arrays of np doubles (horizontal distance, the two depths), v = omega^2/g,
 * roots: DEVICE array of nb_roots >= 100 dispersion roots (roots[0] the
 * propagating one), depth h.  grn: complex128 (np, 3) row-major, [G, dG/dR,
 * dG/dzf] of the wave part (1/r and the free-surface image 1/r1 taken out). */
int cpt_fingreen3d_points(cpt_handle h, int np, const double* rr, const double* zf, const double* zp, double v,
                          int nb_roots, const double* roots, double depth, void* grn, hipStream_t stream);

/* One body of cpt_build_matrices_batched: the arrays cpt_build_matrices
 * takes, for a mesh against itself (collocation points and dot-product
 * normals are the panel centres and normals, as BEMSolver uses them). */
typedef struct {
    int nb_faces, nb_vertices, nb_quad_points, nb_wavenumbers;
    const double* vertices;       /* (nv,3) device */
    const int32_t* faces;         /* (m,4) device, 0-based */
    const double* centers;        /* (m,3) device */
    const double* normals;        /* (m,3) device */
    const double* areas;          /* (m) device */
    const double* radiuses;       /* (m) device */
    const double* quad_points;    /* (m,nq,3) device */
    const double* quad_weights;   /* (m,nq) device */
    const double* wavenumbers;    /* HOST (nb) */
    const int* gf_singularities;  /* HOST (nb), after the overrides of delhommeau.py:466-490 */
    void* S;                      /* complex128, S[i + ldS*j] + b*strideS, OVERWRITTEN */
    int64_t ldS, strideS;
    void* K;                      /* complex128, K[i + ldK*j] + b*strideK (early dot product), OVERWRITTEN */
    int64_t ldK, strideK;
} cpt_body;

/* S and K of many bodies, each at its own list of wavenumbers, in ONE
 * persistent launch walking a (body, row block, panel chunk) work-list.
 * Infinite depth and early dot product only; other options as in
 * cpt_build_matrices.  Equal, entry by entry, to one cpt_build_matrices call
 * per body.  Replaces the loop of BEMSolver.solve_all over bodies and
 * frequencies (src/capytaine/bem/solver.py:268-372) down to
 * matrices.build_matrices (libs/libDelhommeau/src/matrices.f90:28-325). */
int cpt_build_matrices_batched(cpt_handle h, int nb_bodies, const cpt_body* bodies, int adjoint_double_layer,
                               int diagonal_term_in_double_layer, double free_surface, hipStream_t stream);

/* Wave part of the infinite-depth Green function at np point pairs.
 * Replaces interface.vectorized_wave_part_infinite_depth
 * (libs/libDelhommeau/src/interface.f90:29-67) and
 * green_wave.wave_part_infinite_depth (Green_wave.f90:162-271), the hooks the
 * reference's tests call (pytest/test_green_functions_fortran_core.py:37-114).
 * x0i, x0j (np,3); g (np) complex128; nablag (np,3) complex128 row-major.
 * Uses the tabulation set by cpt_set_tabulation. */
int cpt_wave_part_infinite_depth(cpt_handle h, int np, const double* x0i, const double* x0j, double wavenumber,
                                 int gf_singularities, void* g, void* nablag, hipStream_t stream);

/* Batched LU factorisation with partial pivoting of `batch` complex128
 * n x n matrices A + b*strideA (column-major, leading dimension lda), in
 * place.  ipiv[b*strideP + i] (0-based) and info[b] as LAPACK zgetrf
 * (info = 1 + index of the first exactly-zero pivot, 0 if none); info = -1
 * when a multi-work-group panel leaf gave up waiting for its work-groups
 * (CPT_TUNE_PANEL_SPIN), the factor then being incomplete.
 * Replaces scipy.linalg.lu_factor (LAPACK zgetrf) as called by
 * src/capytaine/tools/block_circulant_matrices.py:361-378 from
 * src/capytaine/bem/engines.py:326-335. */
int cpt_lu_factor(cpt_handle h, void* A, int n, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP,
                  int* info, int batch, hipStream_t stream);

/* Solve A X = B for nrhs right-hand sides at once with the factors of
 * cpt_lu_factor; B (n x nrhs, column-major, ldb) is overwritten by X.  The
 * inverses of the factors' 64 x 64 diagonal blocks are formed in the
 * context's workspace (n x 64 complex128 per matrix; CPT_ERR_ALLOC if it
 * cannot grow); calls on one context share it, so order them on one stream.
 * Replaces scipy.linalg.lu_solve (LAPACK zgetrs),
 * src/capytaine/tools/block_circulant_matrices.py:380-385, which the
 * reference calls once per right-hand side (engines.py:338). */
int cpt_lu_solve(cpt_handle h, const void* LU, int n, int64_t lda, int64_t strideA, const int* ipiv,
                 int64_t strideP, void* B, int nrhs, int64_t ldb, int64_t strideB, int batch, hipStream_t stream);

/* cpt_lu_solve without any device allocation: the inverted diagonal blocks
 * go to a static device array of the library (32 MiB per device, held until
 * the process ends) instead of the context's workspace.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * n x batch must not exceed CPT_LU_SOLVE_NOALLOC_MAX_ROWS (CPT_ERR_ARG
 * otherwise).  Every call on one device shares that array: calls that
 * overlap on one device (other streams, other threads, other contexts) race
 * and must be ordered by the caller.  Same results as cpt_lu_solve. */
#define CPT_LU_SOLVE_NOALLOC_MAX_ROWS 32768
int cpt_lu_solve_noalloc(cpt_handle h, const void* LU, int n, int64_t lda, int64_t strideA, const int* ipiv,
                         int64_t strideP, void* B, int nrhs, int64_t ldb, int64_t strideB, int batch,
                         hipStream_t stream);

/* Mixed-precision solve of K X = B for `batch` complex128 systems with nrhs
 * right-hand sides each: K is rounded to complex64 and factorised with FP32
 * MFMA trailing updates; X is then refined with FP64 residuals r = B - K X
 * (against the unrounded complex128 K) and complex64 corrections until
 * max_j ||r_j||_2 / ||b_j||_2 <= tol or max_iter refinement sweeps.  K and B
 * are not modified.  Outputs the sweeps used and the final relative
 * residual.  Synchronises `stream` once per sweep.  With FP16-input tile
 * products (CPT_TUNE_GEMM_TILES 4), a solve left short of tol is factored
 * and refined again with FP32-input products; the sweeps of both are
 * reported and the context's CPT_TUNE_MIXED_RETRIES count grows by one.
 * Returns CPT_ERR_NOT_CONVERGED when the final residual exceeds tol (NaN
 * included) or when the complex64 factor has an exactly-zero pivot; X,
 * iterations and residual are written in that case too.  tol must be >= 0.
 * Replaces, with the north-star's mixed-precision scheme, the factor + solve
 * of DefaultMatrixEngine.linear_solver (src/capytaine/bem/engines.py:326-338,
 * LAPACK zgetrf/zgetrs through block_circulant_matrices.py:374-385). */
int cpt_lu_solve_mixed(cpt_handle h, const void* K, int n, int64_t lda, int64_t strideK, int batch, const void* B,
                       int nrhs, int64_t ldb, int64_t strideB, void* X, int64_t ldx, int64_t strideX, int max_iter,
                       double tol, int* iterations, double* residual, hipStream_t stream);

/* The two halves of cpt_lu_solve_mixed, for a factor reused across calls,
 * as the reference keeps the LU of K in place of K and solves each later
 * right-hand side with it (src/capytaine/bem/engines.py:326-338).
 *
 * cpt_mixed_factor_size: bytes of the factor buffer of `batch` matrices of
 * order n.  Layout (for inspection; offsets from the buffer): a 256-byte
 * header (int32 magic, n, batch, product form), then the complex64 LU
 * factors (batch x n x n, column-major, leading dimension n, L unit lower
 * and U upper as LAPACK cgetrf leaves them), then the pivots (batch x n
 * int32, 0-based) and info (batch int32), then the inverted 64 x 64 diagonal
 * blocks (n x 64 complex64 per matrix); each part starts on a 256-byte
 * boundary.
 * cpt_mixed_factor: factor K (complex128, read) into `factor` (caller-owned
 * device memory, 256-byte aligned, at least cpt_mixed_factor_size bytes),
 * asynchronously on `stream`.
 * cpt_mixed_refine: the refinement of cpt_lu_solve_mixed with that factor
 * (same K, unchanged since), same outputs, tolerance and errors; when
 * FP16-input products leave it short of tol, the factor is recomputed in
 * place with FP32-input products and later refinements use it.  Reads the
 * factor's header (synchronises `stream` first); CPT_ERR_ARG if the buffer
 * holds no factor of this n and batch. */
int cpt_mixed_factor_size(int n, int batch, size_t* bytes);
int cpt_mixed_factor(cpt_handle h, const void* K, int n, int64_t lda, int64_t strideK, int batch, void* factor,
                     size_t factor_bytes, hipStream_t stream);
int cpt_mixed_refine(cpt_handle h, const void* K, int n, int64_t lda, int64_t strideK, int batch, void* factor,
                     size_t factor_bytes, const void* B, int nrhs, int64_t ldb, int64_t strideB, void* X, int64_t ldx,
                     int64_t strideX, int max_iter, double tol, int* iterations, double* residual, hipStream_t stream);

/* One group of the grouped (variable-size) LU: `batch` matrices of order n
 * with the strides of cpt_lu_factor / cpt_lu_solve / cpt_lu_solve_mixed.
 * Fields a call does not use may be zero. */
typedef struct {
    int n, batch;
    void* A;            /* complex128: factored in place (factor), factors (solve), K (mixed, read) */
    int64_t lda, strideA;
    int* ipiv;          /* factor, solve */
    int64_t strideP;
    int* info;          /* factor: (batch) */
    void* B;            /* solve: right-hand sides, overwritten by X; mixed: read */
    int nrhs;
    int64_t ldb, strideB;
    void* X;            /* mixed: solution */
    int64_t ldx, strideX;
} cpt_lu_group;

/* Variable-size batched LU: each group is one cpt_lu_factor /
 * cpt_lu_solve / cpt_lu_solve_mixed batch; the groups run concurrently on
 * the context's lane streams (longest first), ordered after the work
 * already on `stream`, which waits for all of them.  The mixed form
 * synchronises once per refinement sweep for all groups together and
 * writes iterations[g] and residual[g] per group; it retries groups with
 * FP32-input products like cpt_lu_solve_mixed and returns
 * CPT_ERR_NOT_CONVERGED if any group ends above tol or with a singular
 * complex64 factor.  Replaces the per-problem
 * LAPACK calls of DefaultMatrixEngine.linear_solver
 * (src/capytaine/bem/engines.py:300-346) over the bodies and frequencies of
 * BEMSolver.solve_all (src/capytaine/bem/solver.py:268-372). */
int cpt_lu_factor_grouped(cpt_handle h, int ngroups, const cpt_lu_group* groups, hipStream_t stream);
int cpt_lu_solve_grouped(cpt_handle h, int ngroups, const cpt_lu_group* groups, hipStream_t stream);
int cpt_lu_solve_mixed_grouped(cpt_handle h, int ngroups, const cpt_lu_group* groups, int max_iter, double tol,
                               int* iterations, double* residual, hipStream_t stream);

/* C = alpha*A*B + (beta_zero ? 0 : C), complex128, column-major, batched.
 * Replaces the products `S @ sources` of src/capytaine/bem/solver.py:197-198
 * (numpy matmul), done here for all right-hand sides at once. */
int cpt_zgemm(cpt_handle h, int M, int N, int K, double alpha, const void* A, int64_t lda, int64_t strideA,
              const void* B, int64_t ldb, int64_t strideB, void* C, int64_t ldc, int64_t strideC, int beta_zero,
              int batch, hipStream_t stream);

/* complex64 counterparts for the single-precision Green function
 * (floating_point_precision="float32", whose S and K are complex64:
 * abstract_green_function.py:60-70).  cpt_clu_factor / cpt_clu_solve:
 * cpt_lu_factor / cpt_lu_solve on complex64 matrices (LAPACK cgetrf /
 * cgetrs, which scipy.linalg.lu_factor / lu_solve call for a complex64
 * matrix from block_circulant_matrices.py:361-385; the same pivot rule and
 * info); workspace N x 64 complex64 per matrix for the solve.  cpt_cgemm:
 * cpt_zgemm on complex64 (S @ sources, solver.py:197-198).  The tile
 * products take FP32 inputs whatever CPT_TUNE_GEMM_TILES says (its FP16
 * setting serves the mixed solve only). */
int cpt_clu_factor(cpt_handle h, void* A, int n, int64_t lda, int64_t strideA, int* ipiv, int64_t strideP,
                   int* info, int batch, hipStream_t stream);
int cpt_clu_solve(cpt_handle h, const void* LU, int n, int64_t lda, int64_t strideA, const int* ipiv,
                  int64_t strideP, void* B, int nrhs, int64_t ldb, int64_t strideB, int batch, hipStream_t stream);
int cpt_cgemm(cpt_handle h, int M, int N, int K, float alpha, const void* A, int64_t lda, int64_t strideA,
              const void* B, int64_t ldb, int64_t strideB, void* C, int64_t ldc, int64_t strideC, int beta_zero,
              int batch, hipStream_t stream);

/* Restarted GMRES without preconditioner, A x = b for each of the nrhs
 * columns of B (column-major, ldb) in turn, x0 = 0, into X (ldx): the
 * reference's linear_solver="gmres" (src/capytaine/bem/engines.py:74-89,
 * 340-341: scipy.sparse.linalg.gmres(A, b, atol=1e-6) and its defaults
 * rtol=1e-5, restart=20, maxiter=10 n restart cycles).  Stops when
 * ||b - A x||_2 <= max(atol, rtol ||b||_2), the true residual checked after
 * each restart cycle (the Arnoldi estimate ends a cycle early).  restart <= 0
 * and maxiter <= 0 take those defaults; restart is at most 128 and at most n.
 * iterations[j] = Arnoldi steps, residual[j] = final ||b - A x|| / ||b|| of
 * column j (host arrays of nrhs).  Returns CPT_ERR_NOT_CONVERGED when a
 * column misses the tolerance (the reference raises RuntimeError); X holds
 * the last iterates then.  Synchronises `stream` once per Arnoldi step.
 * cpt_zgmres: complex128; cpt_cgmres: complex64 (the single-precision Green
 * function's matrices), reductions in FP64 for both. */
int cpt_zgmres(cpt_handle h, const void* A, int n, int64_t lda, const void* B, int nrhs, int64_t ldb, void* X,
               int64_t ldx, int restart, int maxiter, double rtol, double atol, int* iterations, double* residual,
               hipStream_t stream);
int cpt_cgmres(cpt_handle h, const void* A, int n, int64_t lda, const void* B, int nrhs, int64_t ldb, void* X,
               int64_t ldx, int restart, int maxiter, double rtol, double atol, int* iterations, double* residual,
               hipStream_t stream);

/* Group Fourier transform over the n blocks of a symmetric mesh's
 * matrices or vectors: out_k = s * sum_j chi(k, j) in_j, with
 * chi(k, j) = exp(-2 pi i j k / n) for rotations (kind CPT_BLOCKS_CYCLIC;
 * exp(+...) and s = 1/n when `inverse`) or (-1)^popcount(j & k) for one or
 * two reflection planes (kind CPT_BLOCKS_XOR, n = 2 or 4; s = 1/n when
 * `inverse`).  Element (r, c) of block j of batch item b is
 * in[b*in_batch_stride + j*in_block_stride + c*ldi + r], r < rows, c < cols;
 * of output block k: out[b*out_batch_stride + k*out_block_stride + c*ldo + r].
 * complex128, 1 <= n <= 128; `in` and `out` must not overlap.
 * Replaces BlockCirculantMatrix / NestedBlockCirculantMatrix
 * .block_diagonalize and the FFTs of .solve
 * (src/capytaine/tools/block_circulant_matrices.py:131-160, :243-254),
 * called from engines.py:178-247 (_build_matrices_with_symmetries). */
enum { CPT_BLOCKS_CYCLIC = 0, CPT_BLOCKS_XOR = 1 };
int cpt_block_transform(cpt_handle h, int kind, int n, int inverse, int rows, int cols, int batch, const void* in,
                        int64_t in_batch_stride, int64_t in_block_stride, int64_t ldi, void* out,
                        int64_t out_batch_stride, int64_t out_block_stride, int64_t ldo, hipStream_t stream);

/* Live timing of the hot kernels, for measurement only (no counterpart in
 * the reference).  Between cpt_profile_begin and cpt_profile_end every
 * launch of the kernel classes below made through `h` is bracketed by HIP
 * events recorded on the stream it is launched on.  cpt_profile_end waits
 * for those events and writes, per class, the summed kernel time in ms, the
 * summed algorithmic work (flops for the product and panel classes, 8 M N K
 * per complex product; bytes of A, B and C moved for the skinny products;
 * collocation-point x panel x wavenumber pairs for the assembly) and the
 * number of launches, and, when `bytes` is not NULL, the algorithmic bytes
 * the launches read and write (operands once, results once): arrays of
 * CPT_TIMED_KERNELS.  A panel bracket spans the products launched inside
 * it, which are also counted in their class. */
enum {
    CPT_TIMED_ZGEMM = 0,     /* complex128 MFMA tile product (LU update) */
    CPT_TIMED_CGEMM = 1,     /* complex64 MFMA tile product (mixed precision) */
    CPT_TIMED_PANEL_Z = 2,   /* complex128 panel factorisation */
    CPT_TIMED_PANEL_C = 3,   /* complex64 panel factorisation */
    CPT_TIMED_ASSEMBLY = 4,  /* S/K assembly */
    CPT_TIMED_SKINNY_Z = 5,  /* complex128 product with <= 32 columns (solves, residuals, S X) */
    CPT_TIMED_SKINNY_C = 6,  /* complex64 product with <= 32 columns */
    CPT_TIMED_CGEMM_PACKED = 7, /* complex64 trailing update on FP16 operands packed once per block
                                   step (mixed precision, tile setting 4) */
    CPT_TIMED_KERNELS = 8
};
int cpt_profile_begin(cpt_handle h);
int cpt_profile_end(cpt_handle h, double* ms, double* work, int64_t* launches, double* bytes);

/* Precision of the following assemblies of `h` (cpt_build_matrices,
 * cpt_build_matrices_batched); `flags` is a sum of the following, 0 (the
 * default) being the plain FP64 assembly in complex128:
 *
 * CPT_ASSEMBLY_COMPLEX64: S and K are stored as complex64 (strides and
 * leading dimensions in entries), the FP64 value of each entry rounded once;
 * the single-precision Green function (Delhommeau with
 * floating_point_precision="float32", delhommeau.py:39, :130, whose matrices
 * are complex64: abstract_green_function.py:60-70).
 *
 * CPT_ASSEMBLY_FAR_FIELD_FP32, for the mixed-precision form: reads, for the
 * panel pairs outside the near-field
 * radius (r0 > 7 R and r1 > 7 R, the reference's switch to the one-point
 * Rankine formulas, Green_Rankine.f90:50 and :224), the infinite-depth wave
 * term's 3 x 3 stencil from a float copy of the table, with the reference's
 * stencil; the Lagrange weights, the interpolation and every other term are
 * FP64.  Far entries then differ from
 * the FP64 assembly by the table's float rounding (~1e-8 of max |S|, |K|);
 * near pairs are bit-equal.  Infinite depth only (ignored in finite depth).
 * No reference counterpart: the reference is FP64 (or FP32 throughout when
 * built in single precision). */
enum { CPT_ASSEMBLY_FAR_FIELD_FP32 = 1, CPT_ASSEMBLY_COMPLEX64 = 2 };
int cpt_set_assembly_precision(cpt_handle h, int flags);

/* Caller-owned device workspace for cpt_lu_solve_mixed and
 * cpt_lu_solve_mixed_grouped (the complex64 factors, pivots, residuals and
 * corrections), so that an allocator outside the library (torch's) holds
 * all large buffers.  Used by the following mixed solves of `h` while set
 * and at least as large as they need (cpt_mixed_workspace_size: the same
 * for a single batch, as one group); ptr = NULL unsets.  The buffer must
 * stay allocated until the mixed solve's work on its streams is done. */
int cpt_set_workspace(cpt_handle h, void* ptr, size_t bytes);
int cpt_mixed_workspace_size(cpt_handle h, int ngroups, const cpt_lu_group* groups, size_t* bytes);

/* Free the device workspace a context keeps between calls (it grows to the
 * largest call and is reused otherwise); the tabulation stays.  Waits for
 * the device.  No counterpart in the reference (host memory there). */
int cpt_release_workspace(cpt_handle h);

/* Tuning parameters of a context, set by the autotuner (no counterpart in
 * the reference, whose LAPACK and OpenMP loops have none).  Results depend
 * on them only through rounding: every value factors with LAPACK's partial
 * pivoting rule. */
enum {
    CPT_TUNE_LU_BLOCK = 0,       /* block width of the LU: 64, 128, 256, 512, 1024, or 0 = by size
                                    (256 from 2048 rows, 128 from 1024, else 64); 512 and 1024 with
                                    speculative panels only (else 256) */
    CPT_TUNE_PANEL_LEAF = 1,     /* panel columns factored by one work-group per matrix, 1..64 */
    CPT_TUNE_LOOKAHEAD = 2,      /* 0: in order; 1: next panel on a side stream.  2 (the next panel
                                    on CU-masked streams) returns CPT_ERR_UNSUPPORTED: the runtime
                                    refuses such streams on MI355X */
    CPT_TUNE_ASSEMBLY_CHUNK = 3, /* panels per assembly work-group, 0 = by size, else 1..1024 */
    /* 4 is unused (it was the CU count of look-ahead mode 2) */
    CPT_TUNE_GEMM_TILES = 5,     /* complex64 tiled products: 0: 128 x 64 tiles; 1: 128 x 128;
                                    2: 128 x 128, three real products per complex one (Gauss),
                                    one wavefront per SIMD; 3: the Gauss form on 128 x 64;
                                    4: FP16 inputs and FP32 accumulation, four real products
                                    per complex one on two FP16 planes per operand (the
                                    trailing updates on operands packed once per block step;
                                    the mixed solve's FP64 refinement takes more sweeps).
                                    complex128 products always use 128 x 64, four products;
                                    default 0 */
    CPT_TUNE_LOOKAHEAD_ACTIVE = 6, /* read only (cpt_get_tuning): the look-ahead mode factorisations
                                      run */
    CPT_TUNE_PANEL_THREADS = 7,    /* threads of a panel leaf's work-group: 512 (default) or 256
                                      (one wavefront per SIMD, up to 8192 rows: fits on a CU beside
                                      one work-group of the tile product) */
    CPT_TUNE_MIXED_RETRIES = 8,    /* read only: mixed solves (or groups) of this context factored
                                      again with FP32-input products after the FP16-input products
                                      left the refinement short of its tolerance */
    CPT_TUNE_PANEL_WGS = 9,        /* 1 (default): panel leaves of 32 (complex64) or 16 (complex128)
                                      columns factored by up to 32 work-groups per matrix that
                                      exchange each column's pivot candidates through device memory,
                                      where the leaf's rows fit their registers; 0: one work-group
                                      per matrix */
    CPT_TUNE_PANEL_SPIN = 10,      /* polls a multi-work-group leaf waits for each column's candidates
                                      before giving up (then info = -1): 0 = default (2^20); a
                                      negative value gives up at once (a test of that path) */
    CPT_TUNE_PIVOT_SPECULATION = 11, /* 1 (default): each panel is factored with its diagonal entries
                                        as pivots (no pivot search; the rows below a leaf's leading
                                        block solved at once) and every entry the elimination meets
                                        is checked against LAPACK's rule (|Re|+|Im| of an entry below
                                        the pivot larger than the pivot's, or a zero pivot, fails);
                                        a failing matrix's panel is restored and factored again with
                                        partial pivoting, so the factor is LAPACK's either way.
                                        0: pivot searches in every panel */
    CPT_TUNE_SPEC_REPAIRS = 12     /* read only: panels of this context factored again after a failed
                                      check (one per matrix and panel); reading it waits for the
                                      device */
};
int cpt_set_tuning(cpt_handle h, int param, int value);
int cpt_get_tuning(cpt_handle h, int param, int* value);

#ifdef __cplusplus
}
#endif

#endif /* CAPYTAINE_AMD_H */
