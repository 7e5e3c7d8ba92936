/* bem_abi.h -- C ABI of the MI355X dense PMCHWT library (libbem_amd.so).
 *
 * The reference (KirillSalnikov/BEM-CPP) drives its dense path through C++
 * functions with host buffers (pmchwt.h, assembly.h, solver.h, rhs.h,
 * farfield.h).  This ABI replaces those call sites with plain C entry points:
 * plain pointers and sizes, interleaved (re, im) doubles for complex numbers
 * (the memory layout of std::complex<double>), integer return codes.
 *
 * State lives on the device inside an opaque context (mesh, RWG basis, Z in
 * FP64 and FP32, LU factors and pivots, right-hand sides, solutions), so the
 * pipeline assemble -> factor -> solve -> far field never crosses PCIe except
 * for the small inputs and outputs.  One context is driven by one host thread;
 * work is ordered on the context's HIP stream and every call returns after its
 * work is complete, matching the reference's synchronous semantics.
 *
 * Return codes: 0 = success, > 0 = LAPACK-style info from a factorisation
 * (i+1 for the first zero pivot), < 0 = error (text in bem_last_error).  The
 * library never exits the process (the reference exits via CUDA_CHECK,
 * types.h:12-35).
 */
#ifndef BEM_ABI_H
#define BEM_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bem_ctx bem_ctx;

enum {
    BEM_OK = 0,
    BEM_ERR_ARG = -1,     /* invalid argument or size */
    BEM_ERR_STATE = -2,   /* call out of order (e.g. solve before factor) */
    BEM_ERR_HIP = -3,     /* HIP runtime error */
    BEM_ERR_NOMEM = -4,   /* device allocation failed */
    BEM_ERR_NOCONV = -5,  /* iterative refinement did not reach the tolerance */
    BEM_ERR_MESH = -6     /* mesh has no interior edge / non-finite coordinates */
};

/* Factorisation precision (bem_factor). */
enum {
    BEM_F64 = 0,     /* complex FP64 blocked LU (the reference's Zgetrf semantics) */
    BEM_F32_3M = 1   /* complex FP32 blocked LU, trailing updates as 3 real MFMA GEMMs */
};

/* Preconditioners of bem_solve_gmres and bem_precond_apply. */
enum {
    BEM_PRECOND_LU = 0,    /* the context's LU factors (bem_factor), possibly of an earlier matrix */
    BEM_PRECOND_BLOCK = 1, /* block-Jacobi / restricted additive Schwarz (bem_block_jacobi) */
    BEM_PRECOND_NONE = 2   /* none */
};

/* Assembly flags (bem_assemble). */
enum {
    BEM_ASM_Z64 = 1u, /* keep the FP64 copy of Z (needed by get_Z, the FP64 LU and refinement) */
    BEM_ASM_Z32 = 2u  /* build the FP32 copy of Z (needed by the FP32 LU) */
};

typedef struct {
    int iterations;       /* refinement sweeps performed (0 = plain solve) */
    int converged;        /* 1 if the relative residual reached ir_tol */
    double residual0;     /* max_j ||b_j - Z x_j|| / ||b_j|| after the plain solve (FP64 Z) */
    double residual;      /* the same after the last sweep */
    double t_solve_ms;    /* device time of triangular solves */
    double t_ir_ms;       /* device time of FP64 residuals + corrections */
} bem_solve_stats;

/* ---- context ----------------------------------------------------------- */
bem_ctx* bem_create(int device);
void bem_destroy(bem_ctx* ctx);
const char* bem_last_error(const bem_ctx* ctx);
/* Device-time breakdown of the last calls, in ms:
 * [0] assembly  [1] factorisation  [2] solve  [3] refinement  [4] far field
 * [5] rhs generation  [6] adjoint.  Returns the number written (<= n). */
int bem_timings(const bem_ctx* ctx, double* out, int n);
/* Per-kernel-class device time, measured with HIP events on the stream of
 * each launch (no counterpart in the reference; used for the roofline
 * figures).  mode 1: reset and enable for every class, 3: reset and enable
 * for the LU trailing-update GEMMs only (two events per launch cost the
 * device about 2 us each, so timing every launch slows a factorisation by
 * ~10 %), 0: reset and disable, 2: read.  For class i, out[4i..4i+3] =
 * launches, total ms, counted flops, counted bytes.  Classes: 0 GEMM
 * FP32-3M, 1 GEMM FP64, 2 assembly far pairs, 3 assembly near pairs,
 * 4 assembly gather, 5 pivot tournament, 6 row swaps, 7 panel, 8 triangular
 * solves, 9 far field, 10 RHS, 11 other, 12 LU trailing-update GEMM FP32-3M,
 * 13 LU trailing-update GEMM FP64.  Returns the number of classes. */
int bem_kernel_stats(bem_ctx* ctx, int mode, double* out, int n);

/* Times the GEMM tile configurations of the factorisation and the solves at
 * the current system size with nrhs right-hand sides and keeps the fastest
 * (no counterpart in the reference, which leaves tiling to cuSOLVER; the
 * factorisation also tunes itself on its first call for a size unless
 * BEM_AUTOTUNE=0).  out (nullable, nout doubles) receives the measured TFLOP/s: FP32
 * trailing-update shape per configuration, FP32 solve shape, FP64
 * trailing-update shape, FP64 solve shape (5 configurations each), then the
 * chosen (large, skinny) configuration per precision.  Returns the number of
 * values written, or < 0. */
int bem_autotune(bem_ctx* ctx, int nrhs, double* out, int nout);

/* Fixes the GEMM plan of one precision (prec 0 FP32, 1 FP64) instead of the
 * autotuner's: configuration `large` for products with enough tiles to fill
 * the device, `skinny` for the others; cap > 0 runs the trailing updates that
 * overlap the panel factorisation on a persistent grid of cap blocks (0: one
 * block per tile).  No counterpart in the reference; for experiments and
 * tests.  Requires a system size; returns 0 or < 0. */
int bem_gemm_plan(bem_ctx* ctx, int prec, int large, int skinny, long cap);

/* The GEMM plan of one precision (prec 0 FP32, 1 FP64) as doubles: large,
 * skinny and panel configuration, grid cap, tuned (0/1), chosen by trial
 * factorisations (0/1), then the trial factorisation time in ms of each
 * trailing-update configuration (5 values, 0 where none ran), then the
 * trailing update's configuration.  The FP32 trailing update runs on its own
 * kernel with 4 configurations (64 x 64 with 3 or 2 blocks per CU, 128 x 64,
 * 64 x 64 with a K slice of 32); FP64 on the large tile.  Systems of 8192
 * unknowns and more pick it on their first factorisation per precision by
 * factoring once with each (no counterpart in the reference).  Returns the
 * number of values written, or < 0. */
int bem_gemm_plan_get(bem_ctx* ctx, int prec, double* out, int nout);

/* Copies `bytes` bytes (rounded down to 16) device to device with one plain
 * vectorised kernel: a transfer of known size, the scale for the HBM
 * FETCH_SIZE / WRITE_SIZE counters of profiling runs (no counterpart in the
 * reference).  Returns 0 or < 0. */
int bem_device_copy(bem_ctx* ctx, double bytes);

/* ---- host-side generators (no device work) ------------------------------ */
/* Replace icosphere / hex_prism (mesh.h:21-27, mesh.cpp:11-154): shape 0 is
 * the icosphere of radius `param`, shape 1 the hexagonal prism of unit
 * equal-volume radius and aspect ratio `param`.  bem_mesh_counts gives the
 * vertex and triangle counts; bem_make_mesh fills verts (nv x 3) and tris
 * (nt x 3) in the reference's order. */
int bem_mesh_counts(int shape, int refinements, int* nv, int* nt);
int bem_make_mesh(int shape, int refinements, double param, double* verts, int* tris);
/* Replaces generate_orientations + sort_orientations_nearest (orient.h:41-47):
 * writes RT (n x 3 x 3 row-major, lab -> body) and weights w (n); returns n. */
int bem_orientations(int n_alpha, int n_beta, int n_gamma, int beta_sym, int gamma_sym, int gamma_mirror,
                     int sort, double* RT, double* w);

/* ---- output files (no device work) -------------------------------------- */
/* Replaces write_json (output.h:6-14, output.cpp:5-64): the result file of
 * the reference, byte for byte (M is 4 x 4 x ntheta, [(i*4+j)*ntheta + t];
 * theta in radians; timing = {assembly, solve, far field, total} seconds).
 * Returns 0, or BEM_ERR_ARG when the file cannot be written. */
int bem_write_json(const char* path, const double* M, const double* theta, int ntheta, double ka, double n_re,
                   double n_im, int refinements, int n_alpha, int n_beta, int n_gamma, const double* timing);
/* Replaces the per-orientation Mueller file of main.cpp:640-653
 * (`<out>.per_orient`): int32 header {n_orient, ntheta, n_alpha, n_beta},
 * then n_orient FP64 weights, then n_orient x 16 x ntheta FP64 Mueller
 * matrices without the 1/|k|^2 normalisation (the averaged matrix is
 * sum_o w_o M_o / |k|^2).  Returns 0, or BEM_ERR_ARG when the file cannot be
 * written. */
int bem_write_per_orient(const char* path, int n_orient, int ntheta, int n_alpha, int n_beta, const double* weights,
                         const double* M_orient);

/* ---- mesh and basis ---------------------------------------------------- */
/* Replaces the Mesh/RWG hand-off into assemble_pmchwt (pmchwt.h:12-19):
 * verts is nv x 3, tris is nt x 3 (0-based).  Builds the RWG basis in the
 * reference's std::map edge order (rwg.cpp:5-43) and uploads the quadrature
 * data (tri_quadrature, quadrature.h:14-70; orders 4, 7, 13, others -> 7). */
int bem_set_mesh(bem_ctx* ctx, const double* verts, int nv, const int* tris, int nt, int quad_order);
int bem_num_rwg(const bem_ctx* ctx);
/* RWG arrays as build_rwg returns them (rwg.h:7-16); any pointer may be NULL. */
int bem_get_rwg(const bem_ctx* ctx, int* tri_p, int* tri_m, double* free_p, double* free_m, double* length,
                double* area_p, double* area_m);

/* ---- assembly ---------------------------------------------------------- */
/* Replaces assemble_pmchwt (pmchwt.h:12-19, pmchwt.cu:6-56) and the two
 * assemble_L_K_cuda calls it makes (assembly.h:11-13): builds
 * Z = [[ee L_e + ei L_i, -(K_e+K_i)], [K_e+K_i, L_e/ee + L_i/ei]] on the device. */
int bem_assemble(bem_ctx* ctx, double kext_re, double kext_im, double kint_re, double kint_im, double eta_ext,
                 double eta_int, unsigned flags);
/* Copies Z (2N x 2N, FP64) to the host, row-major (row_major=1, the reference
 * layout of pmchwt.cu:43-49) or column-major. */
int bem_get_Z(bem_ctx* ctx, double* Z, int row_major);
/* count entries Z[rows[e]][cols[e]] of the FP64 matrix (0-based indices) into
 * out[2e], out[2e+1] (interleaved complex).  No reference counterpart: the
 * sampled form of bem_get_Z, for checks of large systems without copying Z. */
int bem_get_Z_entries(bem_ctx* ctx, const int* rows, const int* cols, int count, double* out);
/* Replaces BemFmmOperator::precompute_corrections (bem_fmm.h:165-166,
 * bem_fmm.cu:506-686): the near-field singular corrections that the fast (FMM
 * / pFFT) operators add to their point sums, for the context's mesh and
 * quadrature and the wavenumbers k_ext, k_int.  For every pair (m, n) of RWG
 * functions with halves on one triangle: C_L = the smooth same-triangle
 * quadrature (R = 0 taken at its exact limit ik/(4 pi)) plus the Graglia
 * static part minus the point sum of the full kernel, C_K = minus the point
 * sum's K term.  Output in the reference's CSR form (bem_fmm.h:56-64):
 * row_ptr (N + 1), col_idx (nnz, sorted per row), and the values of L_ext,
 * K_ext, L_int, K_int (nnz interleaved complex each; any may be NULL).  With
 * col_idx == NULL only row_ptr is written.  Returns nnz, or < 0. */
int bem_near_corrections(bem_ctx* ctx, double kext_re, double kext_im, double kint_re, double kint_im, int* row_ptr,
                         int* col_idx, double* L_ext, double* K_ext, double* L_int, double* K_int);
/* Replaces assemble_L_K_cuda (assembly.h:11-13): L and K (N x N row-major) for
 * one wavenumber.  Runs the same kernels with k_ext = k_int = k, eta = 1. */
int bem_assemble_LK(bem_ctx* ctx, double k_re, double k_im, double* L, double* K);

/* ---- dense LU ---------------------------------------------------------- */
/* Replaces lu_factorize_cuda (solver.h:11, solver.cu:7-48): factors the
 * assembled Z in the requested precision.  Returns info (> 0 singular). */
int bem_factor(bem_ctx* ctx, int precision);
/* Loads a caller matrix (n x n, interleaved complex, row-major or column-major)
 * as Z, replacing an assembled one; used by the lu_* drop-ins. */
int bem_set_Z(bem_ctx* ctx, const double* Z, int n, int row_major);
/* Copies the 1-based LAPACK-style pivot vector of the last factorisation
 * (ipiv of lu_factorize_cuda, solver.h:11) into ipiv[n], also after one that
 * returned info > 0 (solver.cu:39-41 copies them back whatever the info). */
int bem_get_pivots(const bem_ctx* ctx, int* ipiv);
/* The factors of the last FP64 factorisation, P Z = L U with unit L, as one
 * n x n interleaved complex matrix (L below, U on and above the diagonal) in
 * row-major order: what lu_factorize_cuda leaves in Z (solver.h:11), also
 * after a zero pivot (info > 0: U is singular, as Zgetrf leaves it). */
int bem_get_lu(bem_ctx* ctx, double* LU);
/* Loads factors and 1-based pivots in the bem_get_lu/bem_get_pivots layout as
 * the current FP64 factorisation: the (Z, ipiv) input of lu_solve_cuda
 * (solver.h:16-17). */
int bem_set_lu(bem_ctx* ctx, const double* LU, const int* ipiv, int n);
/* Replaces lu_solve_cuda / lu_solve_full (solver.h:16-22, solver.cu:51-142):
 * solves Z X = B.  B and X are n x nrhs column-major (one RHS per contiguous
 * column, solver.h:13-15; main.cpp:739-751); X may alias B.  B == NULL solves
 * the right-hand sides left on the device by bem_rhs_planewave; X == NULL
 * keeps the solution on the device only.  With the FP32 factorisation the
 * solution is refined against the FP64 Z until max_j ||b_j - Z x_j||/||b_j||
 * <= ir_tol or ir_max sweeps (ir_max = 0: no refinement).  stats may be NULL. */
int bem_solve(bem_ctx* ctx, const double* B, int nrhs, double* X, double ir_tol, int ir_max, bem_solve_stats* stats);
/* Replaces gmres_solve_paired (block_gmres.h:52-61) on the dense operator and
 * solve_gmres_aext (python/bem_core.py:1104-1123): restarted flexible GMRES
 * on the FP64 Z of the device for all right-hand sides at once (x_0 = 0),
 * preconditioned on the right by `precond` (BEM_PRECOND_LU: the context's
 * factors; BEM_PRECOND_BLOCK: bem_block_jacobi; BEM_PRECOND_NONE), until
 * max_j ||b_j - Z x_j||/||b_j|| <= tol or max_iter iterations (restart
 * length `restart`, 1..200).  The factors may belong to an earlier matrix:
 * bem_assemble without Z32, or bem_set_Z, keeps them for this call (and
 * bem_solve then refuses them).  With the factors of Z at m = 1 this is the
 * reference's A_ext preconditioner (2 A_ext when eta_ext = 1), reused over
 * refractive indices.  B, X, nrhs as for bem_solve; stats->iterations counts
 * the GMRES iterations (one matvec and one preconditioner solve each),
 * residual0 = 1, residual = the final true relative residual.  Returns 0,
 * BEM_ERR_NOCONV when the limit is reached first (X holds the last iterate),
 * or < 0. */
int bem_solve_gmres(bem_ctx* ctx, const double* B, int nrhs, double* X, double tol, int restart, int max_iter,
                    int precond, bem_solve_stats* stats);
/* Replaces the block-Jacobi / RAS build of NearFieldPrecond (precond.cu:407-895;
 * precond.h:38-50) on the dense path: the RWGs are split into spatial-cell
 * blocks (cell side = bounding box / (2.5 / radius_mult), blocks above
 * max_block_rwg bisected; max_block_rwg = 1 gives one block per RWG, the 2 x 2
 * self-interaction preconditioner of precond.cu:976-1090), each extended by
 * the RWGs within overlap_layers * 2 mean element extents (RAS), and each
 * extended block's 2B x 2B sub-matrix of the device's FP64 Z is factored on
 * the device.  Needs the mesh and the FP64 Z; build_ms (nullable) receives the
 * build time.  Returns the block count or < 0.  Cost: each block is inverted
 * by one workgroup, ~(2 Be)^3 operations for Be extended RWGs, so the build
 * time grows with the cube of the block size: the reference's automatic size
 * (AutoPrecondParams::compute, precond.cu:896-970: 180 at N = 9216) builds
 * the 80 blocks of a 2N = 18432 mesh in ~2 s, while max_block_rwg = 1500 on
 * that mesh gives 10 blocks extended to up to 3106 RWGs and takes minutes. */
int bem_block_jacobi(bem_ctx* ctx, double radius_mult, int max_block_rwg, int overlap_layers, double* build_ms);
/* The same partition on the host without a context (precond.cu:414-605):
 * returns the block count; *n_ext receives the total extended size; sizes
 * (nullable, 2 x count ints) receives own and extended sizes per block,
 * rwg_ext (nullable, *n_ext ints) the extended RWG lists (own first), in the
 * RWG order of bem_get_rwg. */
int bem_block_partition(const double* verts, int nv, const int* tris, int nt, int quad_order, double radius_mult,
                        int max_block_rwg, int overlap_layers, int* n_ext, int* sizes, int* rwg_ext);
/* Applies a preconditioner (BEM_PRECOND_LU or BEM_PRECOND_BLOCK) to R
 * (n x nrhs column-major complex): Zout = M^{-1} R (the apply of
 * precond.cu:1154-1265, 1529-1550).  Like bem_solve with B given, it replaces
 * the device right-hand sides (with R) and solutions (with Zout). */
int bem_precond_apply(bem_ctx* ctx, int precond, const double* R, int nrhs, double* Zout);

/* ---- excitation -------------------------------------------------------- */
/* Replaces compute_rhs_planewave (rhs.h:11-16, rhs.cpp:6-64), batched: RHS j
 * is the plane wave with direction k_hat[3j..3j+2] and polarisation
 * E0[3j..3j+2].  Keeps B (2N x nrhs) on the device; B_out (nullable) receives
 * it column-major. */
int bem_rhs_planewave(bem_ctx* ctx, double kext_re, double kext_im, double eta_ext, const double* k_hat,
                      const double* E0, int nrhs, double* B_out);

/* ---- far field and Mueller ---------------------------------------------- */
/* Replaces compute_farfield_batch_cuda (farfield.h:41-48, farfield.cu:267-346):
 * coeffs_J/M are n_calls x N (call c at offset c*N), r_hats n_orient x ndir x 3
 * with call c using orientation c/2, Fv_out n_calls x ndir x 3 complex. */
int bem_farfield(bem_ctx* ctx, const double* coeffs_J, const double* coeffs_M, const double* r_hats, double kext_re,
                 double kext_im, double eta_ext, int n_calls, int n_orient, int ndir, double* Fv_out);
/* No counterpart in the reference: bem_farfield for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results as
 * bem_farfield; it calls no hipMalloc or hipFree: the coefficients,
 * directions, the currents at the quadrature points, the partial sums and the
 * result are staged in one static device array of 32 MiB per GPU, held until
 * the process ends (the far field of one or two orientations fits up to
 * 2N = 18432 unknowns).  A call that needs more returns BEM_ERR_ARG without
 * launching anything.  Every call on a GPU shares its array: calls that
 * overlap on one GPU (from several contexts or threads) race. */
int bem_farfield_noalloc(bem_ctx* ctx, const double* coeffs_J, const double* coeffs_M, const double* r_hats,
                         double kext_re, double kext_im, double eta_ext, int n_calls, int n_orient, int ndir,
                         double* Fv_out);
/* The dense orientation loop of main.cpp:758-846 on the device solution of
 * bem_solve (column 2o = parallel, 2o+1 = perpendicular polarisation of
 * orientation o): far field at theta[t] in the lab xz-plane rotated by RT
 * (n_orient x 3 x 3 row-major), amplitudes S1..S4, Mueller matrix (Bohren &
 * Huffman 3.16, farfield.cu:464-508), optional gamma-mirror zeroing, divided
 * by |k|^2 and summed with weights w in orientation order.  M_avg is 4 x 4 x
 * ntheta ([(i*4+j)*ntheta + t], output.cpp layout).  M_orient (nullable)
 * receives the unweighted per-orientation matrices (n_orient x 16 x ntheta). */
int bem_orient_mueller(bem_ctx* ctx, double kext_re, double kext_im, double eta_ext, const double* RT,
                       const double* w, int n_orient, const double* theta, int ntheta, int gamma_mirror,
                       double* M_avg, double* M_orient);

/* ---- adjoint ----------------------------------------------------------- */
/* No counterpart in the reference (it has no derivatives).  Gradient of a
 * real loss of the averaged Mueller matrix of the last bem_orient_mueller
 * call, given dL/dM_avg (4 x 4 x ntheta), with respect to (Re m, Im m, ka)
 * for k_ext = ka, k_int = ka m, eta_ext = 1, eta_int = 1/|m|
 * (main.cpp:192-196).  Uses one adjoint solve with the existing
 * factorisation (Z^T = S Z S, S = diag(I, -I)), refined to the relative
 * residual the forward solution was refined to (bem_solve's ir_tol with
 * ir_max > 0, bem_solve_gmres' tol; 1e-12 otherwise or when tighter), and a
 * matrix-free contraction of dZ/dp.  grad receives 3 doubles.  Device memory for the call: a work
 * area of about 10 npad x nrhs (nrhs rounded up to 16) + 2 npad^2 doubles,
 * plus two extra triangle-pair record arrays (2 x 54 x T (T + 1) / 2
 * doubles; 16 GB at 2N = 18432); when the record arrays do not fit, the
 * contraction runs once per parameter instead (same result, about three
 * times the assembly passes).  Returns BEM_ERR_NOCONV, with grad written, when
 * the adjoint solve misses its tolerance within 30 refinement sweeps (the
 * environment variable BEM_ADJOINT_TOL, when set, replaces that tolerance).
 * The device right-hand sides and solutions are those of the forward solve
 * again when the call returns, whatever it returns. */
int bem_mueller_grad(bem_ctx* ctx, const double* dL_dM, double* grad);

#ifdef __cplusplus
}
#endif

#endif /* BEM_ABI_H */
