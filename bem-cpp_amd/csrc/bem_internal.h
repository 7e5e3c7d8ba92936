// bem_internal.h -- context layout and kernel-launcher interfaces shared by
// the C ABI (abi.cpp) and the HIP kernels (kernels/*.hip).  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "bem_abi.h"
#include "host/bem_host.h"

// Device-side geometry of one mesh.  Triangle t hosts three RWG "halves", one
// per local edge j (vertices j, j+1; free vertex j+2).
struct DevMesh {
    int T = 0, N = 0, Nq = 0;
    double4* qp = nullptr;   // [T*Nq] quadrature point (x, y, z) and jw = area * w_q
    int4* tv = nullptr;      // [T] vertex indices (w unused)
    double* hc = nullptr;    // [T*3] c = +-l / (2 A): f(r) = c (r - p)
    double* hp = nullptr;    // [T*3*3] free vertex p of each half
    int* hrwg = nullptr;     // [T*3] RWG index of each half (-1: boundary edge)
    int4* rwg_half = nullptr;  // [N] (T+, slot+, T-, slot-)
    int2* near = nullptr;    // [n_near] triangle pairs (t <= s) sharing a vertex
    int n_near = 0;
    double4* pv = nullptr;   // [T*Nq] Graglia P and V at the triangle's own points
};

// Problem constants of one assembly.
struct AsmParams {
    double kr[2], ki[2];  // k_ext, k_int
    double eta[2];        // eta_ext, eta_int
};

// Half-pair records as general combinations of the per-wavenumber parts
// U_w = c_a c_b lm_w (+ self static part), V_w = c_a c_b m0_w (+ static),
// K_w = c_a c_b km_w (forward: L_w = i k_w U_w - (4i/k_w) V_w):
//   P = sum_w aP_w U_w + bP_w V_w,  Q = sum_w aQ_w U_w + bQ_w V_w,
//   Ks = sum_w cK_w K_w,
// complex coefficients [w][re, im].  `deriv` selects the moments of dG/dk and
// dg/dk instead of G and g (no static part; r0[w] is dG/dk at R = 0 of a self
// pair, given the clamped R = 0 value i Re(k)/(4 pi) of the forward kernel).
struct RecCoef {
    double aP[2][2], bP[2][2], aQ[2][2], bQ[2][2], cK[2][2];
    double r0[2][2];
    int deriv;
};

// Up to three record sets written by one pass over the triangle pairs: the
// moments (which do not depend on the coefficients, apart from r0 at the
// R = 0 points of self pairs) are formed once and combined with q[j] into the
// record array H[j].  All q[j] share `deriv`.
constexpr int kMaxRecSets = 3;
struct RecCoefSet {
    RecCoef q[kMaxRecSets];
    double* H[kMaxRecSets];
    int n;
};

// Per-kernel-class device timing (bem_kernel_stats): HIP events around each
// launch on the context stream, resolved when the ABI call returns.
enum KernelClass {
    KS_GEMM32 = 0,   // complex GEMM, FP32 3M on MFMA
    KS_GEMM64,       // complex GEMM, FP64 4M on MFMA
    KS_TRIPAIR_FAR,  // assembly, non-touching triangle pairs
    KS_TRIPAIR_NEAR, // assembly, touching pairs + Graglia
    KS_GATHER,       // assembly, Z from half-pair records
    KS_TSLU,         // pivot tournament
    KS_LASWP,        // row interchanges
    KS_PANEL,        // panel factor (diagonal block, L21, U12 strip)
    KS_TRSM,         // triangular solves in LDS
    KS_FARFIELD,     // far field partial sums
    KS_RHS,          // plane-wave right-hand sides
    KS_OTHER,
    KS_UPDATE32,     // trailing-update GEMMs of the LU, FP32 3M (own kernel symbol)
    KS_UPDATE64,     // trailing-update GEMMs of the LU, FP64
    KS_COUNT
};

struct KernelTimer {
    bool on = false;
    unsigned mask = ~0u;  // classes that get events
    struct Rec {
        hipEvent_t a, b;
        hipStream_t s;
        int cls;
        double flops, bytes;
    };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double ms[KS_COUNT] = {}, flops[KS_COUNT] = {}, bytes[KS_COUNT] = {};
    long launches[KS_COUNT] = {};
};

// GEMM tile choice of one precision (configurations of kernels/gemm.h):
// `large` for products with at least `fill` blocks in it, else `skinny`.
struct GemmPlan {
    static constexpr int kMaxCfgs = 5;
    int large = 0, skinny = 2;
    int panel = 3;                           // products of the panel stream (smallest LDS footprint)
    int nt = 0;                              // FP32 trailing update: configuration of k_cupdate_nt
    long fill = 512;
    bool tuned = false;
    int npad = 0;                            // system size the plan was tuned for
    int cus = 256;                           // compute units of the device
    int occ[kMaxCfgs] = {1, 1, 1, 1, 1};     // resident blocks per CU of each configuration
    bool occ_known = false;                  // cus and occ queried from the device
    long cap = 0;                            // grid of capped launches; 0: from occ and cus
    int sk_cfg = -1;                         // split-K products: configuration timed fastest for
    int sk_shape[4] = {0, 0, 0, 0};          //   this (M, N, K, role)
    bool trialed = false;                    // trailing-update tile chosen by whole trial factorisations
    double trial_ms[kMaxCfgs] = {};          // their times per configuration (0: not run): FP32 over `nt`,
                                             //   FP64 over `large`
    double alone_tflops[kMaxCfgs] = {};      // trailing-update shape timed alone (autotuner)
    GemmPlan() = default;
    GemmPlan(int l, int s) : large(l), skinny(s) {}
};

// Block-Jacobi / restricted additive Schwarz preconditioner on the device
// (precond.hip): per block the extended RWG list and the inverse of its
// 2Be x 2Be sub-matrix of Z in a zero-padded ld x ld slot.
struct BlockJacobi {
    bool valid = false;
    int nb = 0, n2max = 0, ld = 0, N = 0;
    size_t bytes = 0;
    int *rwg = nullptr, *B = nullptr, *Be = nullptr, *roff = nullptr, *poff = nullptr, *piv = nullptr;
    double *Ar = nullptr, *Ai = nullptr;
    std::vector<std::vector<int>> own, ext;
};

struct bem_ctx {
    int device = 0;
    GemmPlan gemm_plan[2] = {GemmPlan(0, 2), GemmPlan(1, 2)};  // [0] FP32 3M, [1] FP64 (defaults before tuning)
    KernelTimer kt;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // panel stream of the look-ahead factorisation
    hipStream_t stream3 = nullptr;  // updates inside an outer panel off its critical path
    // factorisation stream layout (see factor() in kernels/lu.hip): 0 one
    // stream, 1 panel + update streams on all CUs
    int lu_streams = -1;            // -1: not chosen yet
    hipEvent_t ev[2] = {nullptr, nullptr};
    hipEvent_t lu_ev[6] = {};  // cross-stream ordering inside the factorisation
    std::string err;

    // mesh and basis (host copies)
    Mesh mesh;
    RWG rwg;
    TriQuad quad{};
    bool has_mesh = false;
    DevMesh dm;

    // system matrix: n = 2N, rows/cols padded to npad, planar complex, row-major, ld = npad
    int n = 0, npad = 0;
    bool has_Z64 = false, has_Z32 = false;
    double* Zr = nullptr;  // FP64 Z (kept for the refinement residual and get_Z)
    double* Zi = nullptr;
    float* Wr = nullptr;   // FP32 working copy / FP32 LU factors
    float* Wi = nullptr;
    double* Vr = nullptr;  // FP64 LU factors
    double* Vi = nullptr;
    void* inv32 = nullptr;  // inverses of the 256 x 256 diagonal blocks of L and U (FP32 factors;
    size_t inv32_bytes = 0; // the FP64 solves substitute with the diagonal blocks themselves)
    int* ipiv = nullptr;   // [npad] LAPACK-style swaps (0-based row swapped with row i)
    int* perm = nullptr;   // [npad] row permutation: row i of PZ is row perm[i] of Z
    BlockJacobi bj;        // block-Jacobi preconditioner (bem_block_jacobi)
    int factored = -1;     // -1 none, BEM_F64, BEM_F32_3M (factors usable for solves)
    int lu_avail = -1;     // precision of the last factorisation, also one that met a zero pivot (get_lu)
    bool factors_stale = false;  // the factors belong to an earlier matrix (preconditioner only)
    AsmParams asm_params{};
    bool asm_valid = false;

    // assembly scratch: per triangle pair (t <= s) the 9 half-pair values
    double* H = nullptr;
    size_t H_bytes = 0;

    // right-hand sides / solutions, FP64 planar row-major n x ldr
    int nrhs = 0, ldr = 0;
    double *Br = nullptr, *Bi = nullptr, *Xr = nullptr, *Xi = nullptr;
    size_t rhs_cap = 0;  // doubles per plane allocated
    bool has_B = false, has_X = false;
    double x_tol = 0.0;  // relative residual the device solution was refined to (0: not known)

    // plane waves of the last bem_rhs_planewave call (k_hat, E0, H0 per RHS)
    std::vector<double> rhs_prm;
    double rhs_k[2] = {0, 0}, rhs_eta = 1.0;
    bool rhs_planewave = false;

    // far field / Mueller of the last bem_orient_mueller call (for the adjoint)
    std::vector<double> last_RT, last_w, last_theta;
    int last_norient = 0, last_ntheta = 0, last_mirror = 0;
    double last_kext[2] = {0, 0}, last_eta_ext = 1.0;

    // scratch allocations owned by the kernels (grown on demand)
    void* scratch = nullptr;
    size_t scratch_bytes = 0;
    void* part = nullptr;  // partial products of split-K GEMMs
    size_t part_bytes = 0;
    void* noalloc_ptr = nullptr;  // device address of the static no-allocation array (bem_farfield_noalloc)

    double t_ms[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Error helpers ---------------------------------------------------------------
int bem_fail(bem_ctx* c, int code, const std::string& msg);
#define BEM_HIP(ctx, call)                                                                              \
    do {                                                                                                \
        hipError_t e_ = (call);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return bem_fail((ctx), -3, std::string(#call) + ": " + hipGetErrorString(e_) + " @" +       \
                                           __FILE__ + ":" + std::to_string(__LINE__));                  \
    } while (0)

// Device scratch of at least `bytes` (reallocated when too small).
int ctx_scratch(bem_ctx* c, size_t bytes, void** out);

// Launchers (kernels/*.hip).  All enqueue on c->stream and return 0 or < 0.
int asm_launch(bem_ctx* c, const AsmParams& p, bool want64, bool want32);
int lu_factor_launch(bem_ctx* c, int precision);
// D = (P^T L U)^{-1} S with the context's factors (FP32 or FP64), S and D
// planar FP64 npad x ldr (D must not alias S); work holds
// lu_precond_work_bytes(c) bytes.  Enqueued on the context stream.
size_t lu_precond_work_bytes(const bem_ctx* c);
int lu_precond_apply(bem_ctx* c, const double* Sr, const double* Si, double* Dr, double* Di, void* work);
// Block-Jacobi preconditioner (precond.hip): build from the FP64 Z and the
// mesh (returns the block count or < 0), free, apply to planar FP64 npad x ldr
// vectors (work: bj_work_bytes(c) bytes).
int bj_build_launch(bem_ctx* c, double radius_mult, int max_block_rwg, int overlap_layers, double* t_ms);
void bj_free(bem_ctx* c);
size_t bj_work_bytes(const bem_ctx* c);
int bj_apply(bem_ctx* c, const double* Rr, const double* Ri, double* Zr, double* Zi, void* work);
// Restarted flexible GMRES for Z X = B (Z the FP64 matrix on the device, B
// and X the context's right-hand sides and solutions) with the factors as
// right preconditioner (gmres.hip).
int gmres_solve_launch(bem_ctx* c, int nrhs, double tol, int restart, int max_iter, int precond, int* iters,
                       double* res0, double* res, double* t_ms);
int lu_solve_launch(bem_ctx* c, int nrhs, double ir_tol, int ir_max, int* iters, double* res0, double* res,
                    double* t_solve_ms, double* t_ir_ms);
int rhs_launch(bem_ctx* c, double kr, double ki, double eta, const double* khat, const double* E0, int nrhs);
int ff_launch(bem_ctx* c, const double* dJr, const double* dJi, const double* dMr, const double* dMi, int ldc,
              const double* drh, double kr, double ki, double eta, int n_calls, int n_orient, int ndir,
              double* dFv);
// bem_farfield without device allocation: all staging in the library's static
// per-GPU array of kNoAllocBytes (farfield.hip); ff_noalloc_init fetches its
// device address into c->noalloc_ptr (once per context, device current).
constexpr size_t kNoAllocBytes = 32u << 20;
int ff_noalloc_init(bem_ctx* c);
int ff_noalloc_launch(bem_ctx* c, const double* cJ, const double* cM, const double* rh, double kr, double ki,
                      double eta, int n_calls, int n_orient, int ndir, double* Fv);
int mueller_launch(bem_ctx* c, double kr, double ki, double eta, const double* RT, const double* w, int n_orient,
                   const double* theta, int ntheta, int mirror, double* M_avg, double* M_orient);
int adjoint_launch(bem_ctx* c, const double* dL_dM, double* grad);
// Times the GEMM tile configurations at the shapes of an npad system with
// nrhs right-hand sides and sets c->gemm_plan; out (nullable) receives the
// measured TFLOP/s per (precision, shape, configuration).
int gemm_autotune_launch(bem_ctx* c, int npad, int nrhs, double* out, int nout);
int gemm_ensure_tuned(bem_ctx* c, bool f32);
// Sets the plan of one precision (0 FP32, 1 FP64) to the given configurations
// and grid cap of the capped trailing updates (0: one block short of full
// occupancy per CU), replacing the autotuner's choice.
int gemm_set_plan(bem_ctx* c, int prec, int large, int skinny, long cap);
// Stores the context's plan of one precision in the per-process cache of
// tuned plans (after the factorisation trials changed it).
void gemm_cache_store(bem_ctx* c, bool f32);
// Fills cus and occ of the plan of one precision from the device (no timing).
void gemm_plan_occupancy(bem_ctx* c, bool f32);
// Chooses the factorisation stream layout (environment BEM_LU_STREAMS,
// BEM_LU_PANEL_CUS) and creates the CU-masked streams it needs.
int lu_setup_streams(bem_ctx* c);
// One copy kernel of n16 16-byte vectors between two fresh device buffers.
int device_copy_launch(bem_ctx* c, size_t n16);
// d B / d k_ext of the stored plane waves into (dBr, dBi) (npad x ldr).
int rhs_dk_launch(bem_ctx* c, double* dBr, double* dBi);
// Far-field part of the adjoint for the last bem_orient_mueller call: given
// dL/dM_avg (host, 16 x ntheta), writes g = dL/dX (holomorphic derivative,
// npad x ldr planar, device) and the explicit dL/dk_ext at fixed X.
int ff_adjoint_launch(bem_ctx* c, const double* dL_dM, double* gr, double* gi, double* explicit_dk);
// <Z_q, W> = sum_{m,n} Z_q[m][n] W[m][n] over the 2N x 2N matrix whose
// half-pair records are combined with q (PMCHWT block pattern of Z), W planar
// FP64 row-major with leading dimension ldw, for the nq coefficient sets q[j]
// (1 <= nq <= kMaxRecSets, one `deriv`): result j as (re, im) in
// out[2j], out[2j + 1].  Sets after the first need record arrays of their
// own: `extra` holds (nq - 1) * asm_record_bytes(c) bytes.
int asm_contract_launch(bem_ctx* c, const AsmParams& p, const RecCoef* q, int nq, double* extra, const double* Wr,
                        const double* Wi, long ldw, double* out);
// Gathers count entries Z64[rows[e]][cols[e]] (device index arrays, checked
// by the caller) into out (device, interleaved complex).
int z_entries_launch(bem_ctx* c, const int* rows, const int* cols, int count, double* out);
// Near-field singular corrections of the fast operators (assembly.hip):
// writes row_ptr (N + 1) and returns nnz; with col_idx also col_idx (nnz) and
// out[0..3] = L_ext, K_ext, L_int, K_int (nnz interleaved complex each, host;
// null entries skipped).
int near_corr_launch(bem_ctx* c, double ker, double kei, double kir, double kii, int* row_ptr, int* col_idx,
                     double* const out[4]);
// Bytes of one record array (T (T + 1) / 2 triangle-pair records).
size_t asm_record_bytes(const bem_ctx* c);

// Small helpers used by several launchers
int ctx_ensure_rhs(bem_ctx* c, int nrhs);
inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Kernel timing helpers (no-ops unless enabled).  kt_mark returns a record
// index to close with kt_close after the launch; both events go on stream s
// (default: the context stream).
inline int kt_mark(bem_ctx* c, int cls, hipStream_t s = nullptr) {
    KernelTimer& k = c->kt;
    if (!k.on || !((k.mask >> cls) & 1u)) return -1;
    KernelTimer::Rec r{};
    for (hipEvent_t* e : {&r.a, &r.b}) {
        if (!k.pool.empty()) {
            *e = k.pool.back();
            k.pool.pop_back();
        } else if (hipEventCreate(e) != hipSuccess) {
            return -1;
        }
    }
    r.cls = cls;
    r.s = s ? s : c->stream;
    (void)hipEventRecord(r.a, r.s);
    k.pending.push_back(r);
    return (int)k.pending.size() - 1;
}
inline void kt_close(bem_ctx* c, int idx, double flops, double bytes) {
    if (idx < 0) return;
    KernelTimer::Rec& r = c->kt.pending[idx];
    r.flops = flops;
    r.bytes = bytes;
    (void)hipEventRecord(r.b, r.s);
}
// Waits for the recorded launches and folds them into the totals.
void kt_resolve(bem_ctx* c);
