// gemm.h -- complex GEMM on CDNA4 matrix cores, planar operands.
//
//   C = C - A B   (mode 0)     C = A B   (mode 1)     C = C + A B   (mode 2)
//   C = -A B      (mode 3)
//
// mode | GEMM_MODE_URGENT raises the waves' issue priority (s_setprio), for
// the small products on the factorisation's critical path that share CUs
// with a running trailing update.
//
// A batch of independent products of one shape runs in one launch
// (blockIdx.z), operand z at the base pointers plus z times the batch strides.
// A is M x K, B is K x N, C is M x N; every matrix is stored as two planes
// (real, imaginary) in row-major order with its own leading dimension.
//
// FP32 uses the 3M form: three real products T1 = Ar Br, T2 = Ai Bi,
// T3 = (Ar + Ai)(Br + Bi) accumulate on v_mfma_f32_16x16x4f32, and the
// epilogue forms Re = T1 - T2, Im = T3 - T1 - T2 (3/4 of the real flops of the
// 4M form).  FP64 uses the 4M form on v_mfma_f64_16x16x4f64 with two
// accumulators (Re += Ar Br - Ai Bi, Im += Ar Bi + Ai Br).
//
// Tiling: 256 threads = 4 waves of 64 in a 2 x 2 arrangement; a K slice of BK
// is staged through LDS (k-major, double buffered, one barrier per slice);
// each wave owns a WM x WN sub-tile made of 16 x 16 MFMA tiles.
// MFMA 16x16x4 operand layout (wave64): A[i][k] in lane i + 16k, B[k][j] in
// lane j + 16k.  Result D[i][j]: f32 form in lane j + 16 (i / 4), register
// i % 4; f64 form in lane j + 16 (i % 4), register i / 4 (measured on gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

constexpr int GEMM_MODE_URGENT = 4;

typedef float v4f __attribute__((ext_vector_type(4)));
typedef double v4d __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma16x16x4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ v4d mfma16x16x4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Tile configurations, indexed per precision (0, 1, 2): block BM x BN, wave
// tile WM x WN (4 waves in 2 x 2), K slice BK.  MINB: workgroups per CU the
// register allocation must allow (two per CU let one block's prologue and C
// update overlap another's MFMA stream).  Which configuration a launch uses
// is decided by the context's plan (gemm_plan in bem_ctx), set by the
// autotuner from measured times at the context's shapes.
template <typename T, int CFG>
struct GemmCfg;
template <>
struct GemmCfg<float, 0> {
    static constexpr int BM = 128, BN = 128, BK = 16, WM = 64, WN = 64, PAD = 16, MINB = 1;
    typedef v4f acc_t;
    typedef float4 vec_t;  // 16-byte global loads
    static constexpr int VEC = 4;
};
template <>
struct GemmCfg<float, 1> {
    static constexpr int BM = 128, BN = 64, BK = 16, WM = 64, WN = 32, PAD = 16, MINB = 2;
    typedef v4f acc_t;
    typedef float4 vec_t;
    static constexpr int VEC = 4;
};
template <>
struct GemmCfg<float, 2> {
    static constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32, PAD = 16, MINB = 2;
    typedef v4f acc_t;
    typedef float4 vec_t;
    static constexpr int VEC = 4;
};
template <>
struct GemmCfg<double, 0> {
    static constexpr int BM = 128, BN = 64, BK = 16, WM = 64, WN = 32, PAD = 8, MINB = 1;
    typedef v4d acc_t;
    typedef double2 vec_t;
    static constexpr int VEC = 2;
};
template <>
struct GemmCfg<double, 1> {
    static constexpr int BM = 64, BN = 64, BK = 16, WM = 32, WN = 32, PAD = 8, MINB = 2;
    typedef v4d acc_t;
    typedef double2 vec_t;
    static constexpr int VEC = 2;
};
template <>
struct GemmCfg<double, 2> {
    static constexpr int BM = 64, BN = 32, BK = 16, WM = 32, WN = 16, PAD = 8, MINB = 2;
    typedef v4d acc_t;
    typedef double2 vec_t;
    static constexpr int VEC = 2;
};
// 64 x 64 with a K slice of 8: half the LDS of configuration 2 / 1, so more
// blocks fit on a CU beside the latency-bound panel kernels.
template <>
struct GemmCfg<float, 3> {
    static constexpr int BM = 64, BN = 64, BK = 8, WM = 32, WN = 32, PAD = 16, MINB = 2;
    typedef v4f acc_t;
    typedef float4 vec_t;
    static constexpr int VEC = 4;
};
template <>
struct GemmCfg<double, 3> {
    static constexpr int BM = 64, BN = 64, BK = 8, WM = 32, WN = 32, PAD = 8, MINB = 2;
    typedef v4d acc_t;
    typedef double2 vec_t;
    static constexpr int VEC = 2;
};
// 32 x 32 (one 16 x 16 MFMA tile per wave): products too small to give the
// device one block per CU with the other tiles (the block-inverse products
// and the late updates of the solves, the L11^{-1} doubling).
template <>
struct GemmCfg<float, 4> {
    static constexpr int BM = 32, BN = 32, BK = 16, WM = 16, WN = 16, PAD = 16, MINB = 4;
    typedef v4f acc_t;
    typedef float4 vec_t;
    static constexpr int VEC = 4;
};
template <>
struct GemmCfg<double, 4> {
    static constexpr int BM = 32, BN = 32, BK = 16, WM = 16, WN = 16, PAD = 8, MINB = 4;
    typedef v4d acc_t;
    typedef double2 vec_t;
    static constexpr int VEC = 2;
};
constexpr int kGemmCfgs = 5;
constexpr int kGemmSmall = 4;       // the 32 x 32 configuration
constexpr int kGemmTrialCfgs = 4;   // configurations tried for the trailing update

template <typename T>
__device__ __forceinline__ void vec_split(const typename GemmCfg<T, 0>::vec_t& v, T* out);
template <>
__device__ __forceinline__ void vec_split<float>(const float4& v, float* o) {
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
template <>
__device__ __forceinline__ void vec_split<double>(const double2& v, double* o) {
    o[0] = v.x; o[1] = v.y;
}

// ROLE 1 only names the instantiation (the LU trailing updates, so a
// profiler lists them apart from the other products of the same tile).
// ROLE 2 selects the 3M form in FP64: the refinement residual, whose
// rounding (a few eps of |Z||X|) only has to stay far below the FP32
// factors' error, takes three real products instead of four.
template <typename T, int CFG, int ROLE = 0>
__global__ __launch_bounds__(256, (GemmCfg<T, CFG>::MINB)) void k_cgemm(int M, int N, int K, const T* __restrict__ Ar,
                                               const T* __restrict__ Ai, long lda, const T* __restrict__ Br,
                                               const T* __restrict__ Bi, long ldb, T* __restrict__ Cr,
                                               T* __restrict__ Ci, long ldc, int mode, int nz, long sa, long sb,
                                               long sc) {
    typedef GemmCfg<T, CFG> C;
    constexpr int BM = C::BM, BN = C::BN, BK = C::BK, WM = C::WM, WN = C::WN, VEC = C::VEC;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int SA = BM + C::PAD, SB = BN + C::PAD;
    constexpr int AVT = BM * BK / VEC, BVT = BK * BN / VEC;  // 16-byte vectors per plane per slice
    constexpr int AV = (AVT + 255) / 256, BV = (BVT + 255) / 256;  // per thread
    typedef typename C::vec_t vec_t;
    typedef typename C::acc_t acc_t;
    constexpr bool M3 = sizeof(T) == 4 || ROLE == 2;  // 3M form (FP32 always; FP64 for ROLE 2)
    constexpr int NACC = M3 ? 3 : 2;

    __shared__ T sA[2][2][BK][SA];
    __shared__ T sB[2][2][BK][SB];

    if (mode & GEMM_MODE_URGENT) __builtin_amdgcn_s_setprio(3);
    mode &= 3;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm0 = (wave >> 1) * WM, wn0 = (wave & 1) * WN;
    const int ntx = (N + BN - 1) / BN, nty = (M + BM - 1) / BM;
    const int ntiles = ntx * nty * nz;
    const int nk = K / BK;

    // Tiles in row-major order over (z, row, column); a grid smaller than the
    // tile count makes every block loop (persistent launch).
    for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
        const int bz = tile / (ntx * nty), rem = tile - bz * (ntx * nty);
        const int row0 = (rem / ntx) * BM, col0 = (rem % ntx) * BN;
        const T* __restrict__ ar_ = Ar + bz * sa;
        const T* __restrict__ ai_ = Ai + bz * sa;
        const T* __restrict__ br_ = Br + bz * sb;
        const T* __restrict__ bi_ = Bi + bz * sb;

        vec_t ga[2][AV], gb[2][BV];
        auto load = [&](int k0) {
#pragma unroll
            for (int u = 0; u < AV; u++) {
                const int v = tid + 256 * u;
                const int r = v / (BK / VEC), kv = (v % (BK / VEC)) * VEC;
                const bool ok = (AVT % 256 == 0 || v < AVT) && row0 + r < M;
                const long o = (long)(row0 + r) * lda + k0 + kv;
                ga[0][u] = ok ? *reinterpret_cast<const vec_t*>(ar_ + o) : vec_t{};
                ga[1][u] = ok ? *reinterpret_cast<const vec_t*>(ai_ + o) : vec_t{};
            }
#pragma unroll
            for (int u = 0; u < BV; u++) {
                const int v = tid + 256 * u;
                const int kr = v / (BN / VEC), nv = (v % (BN / VEC)) * VEC;
                const bool ok = (BVT % 256 == 0 || v < BVT) && col0 + nv < N;
                const long o = (long)(k0 + kr) * ldb + col0 + nv;
                gb[0][u] = ok ? *reinterpret_cast<const vec_t*>(br_ + o) : vec_t{};
                gb[1][u] = ok ? *reinterpret_cast<const vec_t*>(bi_ + o) : vec_t{};
            }
        };
        auto store = [&](int buf) {
#pragma unroll
            for (int u = 0; u < AV; u++) {
                const int v = tid + 256 * u;
                if (AVT % 256 != 0 && v >= AVT) continue;
                const int r = v / (BK / VEC), kv = (v % (BK / VEC)) * VEC;
#pragma unroll
                for (int p = 0; p < 2; p++) {
                    T e[VEC];
                    vec_split<T>(ga[p][u], e);
#pragma unroll
                    for (int q = 0; q < VEC; q++) sA[buf][p][kv + q][r] = e[q];
                }
            }
#pragma unroll
            for (int u = 0; u < BV; u++) {
                const int v = tid + 256 * u;
                if (BVT % 256 != 0 && v >= BVT) continue;
                const int kr = v / (BN / VEC), nv = (v % (BN / VEC)) * VEC;
#pragma unroll
                for (int p = 0; p < 2; p++) *reinterpret_cast<vec_t*>(&sB[buf][p][kr][nv]) = gb[p][u];
            }
        };

        acc_t acc[NACC][TM][TN];
#pragma unroll
        for (int a = 0; a < NACC; a++)
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) acc[a][i][j] = acc_t{};

        load(0);
        store(0);
        __syncthreads();
        for (int kt = 0; kt < nk; kt++) {
            const int buf = kt & 1;
            if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                const int ka = kk + (lane >> 4);
                T ar[TM], ai[TM], br[TN], bi[TN];
#pragma unroll
                for (int i = 0; i < TM; i++) {
                    ar[i] = sA[buf][0][ka][wm0 + i * 16 + (lane & 15)];
                    ai[i] = sA[buf][1][ka][wm0 + i * 16 + (lane & 15)];
                }
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    br[j] = sB[buf][0][ka][wn0 + j * 16 + (lane & 15)];
                    bi[j] = sB[buf][1][ka][wn0 + j * 16 + (lane & 15)];
                }
                if constexpr (M3) {
                    T as[TM], bs[TN];
#pragma unroll
                    for (int i = 0; i < TM; i++) as[i] = ar[i] + ai[i];
#pragma unroll
                    for (int j = 0; j < TN; j++) bs[j] = br[j] + bi[j];
#pragma unroll
                    for (int i = 0; i < TM; i++)
#pragma unroll
                        for (int j = 0; j < TN; j++) {
                            acc[0][i][j] = mfma16x16x4(ar[i], br[j], acc[0][i][j]);
                            acc[1][i][j] = mfma16x16x4(ai[i], bi[j], acc[1][i][j]);
                            acc[2][i][j] = mfma16x16x4(as[i], bs[j], acc[2][i][j]);
                        }
                } else {
#pragma unroll
                    for (int i = 0; i < TM; i++)
#pragma unroll
                        for (int j = 0; j < TN; j++) {
                            acc[0][i][j] = mfma16x16x4(ar[i], br[j], acc[0][i][j]);
                            acc[0][i][j] = mfma16x16x4(ai[i], -bi[j], acc[0][i][j]);
                            acc[1][i][j] = mfma16x16x4(ar[i], bi[j], acc[1][i][j]);
                            acc[1][i][j] = mfma16x16x4(ai[i], br[j], acc[1][i][j]);
                        }
                }
            }
            if (kt + 1 < nk) store(buf ^ 1);
            __syncthreads();  // also ends the reads of this tile before the next tile's first store
        }

        T* __restrict__ cr_ = Cr + bz * sc;
        T* __restrict__ ci_ = Ci + bz * sc;
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const int col = col0 + wn0 + j * 16 + (lane & 15);
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int row = row0 + wm0 + i * 16 + (sizeof(T) == 4 ? 4 * (lane >> 4) + r : (lane >> 4) + 4 * r);
                    if (row >= M || col >= N) continue;
                    T vr, vi;
                    if constexpr (M3) {
                        vr = acc[0][i][j][r] - acc[1][i][j][r];
                        vi = acc[2][i][j][r] - acc[0][i][j][r] - acc[1][i][j][r];
                    } else {
                        vr = acc[0][i][j][r];
                        vi = acc[1][i][j][r];
                    }
                    const long o = (long)row * ldc + col;
                    if (mode == 0) {
                        cr_[o] -= vr;
                        ci_[o] -= vi;
                    } else if (mode == 1) {
                        cr_[o] = vr;
                        ci_[o] = vi;
                    } else if (mode == 3) {
                        cr_[o] = -vr;
                        ci_[o] = -vi;
                    } else {
                        cr_[o] += vr;
                        ci_[o] += vi;
                    }
                }
            }
    }
}

// Host launcher of configuration cfg.  K must be a multiple of BK; N a
// multiple of the vector width (4 floats / 2 doubles) and every leading
// dimension a multiple of it.  nb products with batch strides sa, sb, sc.
// cap > 0 limits the grid to cap blocks, which then loop over the tiles.
template <typename T, int CFG>
inline void cgemm_cfg(hipStream_t s, int M, int N, int K, const T* Ar, const T* Ai, long lda, const T* Br,
                      const T* Bi, long ldb, T* Cr, T* Ci, long ldc, int mode, int nb, long sa, long sb, long sc,
                      long cap, int role) {
    typedef GemmCfg<T, CFG> C;
    const long tiles = (long)((N + C::BN - 1) / C::BN) * ((M + C::BM - 1) / C::BM) * nb;
    const long grid = cap > 0 && cap < tiles ? cap : tiles;
    if (role == 1)
        hipLaunchKernelGGL((k_cgemm<T, CFG, 1>), dim3((unsigned)grid), dim3(256), 0, s, M, N, K, Ar, Ai, lda, Br, Bi,
                           ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc);
    else if constexpr (sizeof(T) == 8 && CFG != 0) {  // the 128 x 64 tile would spill in the 3M form
        if (role == 2)
            hipLaunchKernelGGL((k_cgemm<T, CFG, 2>), dim3((unsigned)grid), dim3(256), 0, s, M, N, K, Ar, Ai, lda, Br,
                               Bi, ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc);
        else
            hipLaunchKernelGGL((k_cgemm<T, CFG, 0>), dim3((unsigned)grid), dim3(256), 0, s, M, N, K, Ar, Ai, lda, Br,
                               Bi, ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc);
    }
    else
        hipLaunchKernelGGL((k_cgemm<T, CFG, 0>), dim3((unsigned)grid), dim3(256), 0, s, M, N, K, Ar, Ai, lda, Br, Bi,
                           ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc);
}

template <typename T>
inline void cgemm_any(int cfg, hipStream_t s, int M, int N, int K, const T* Ar, const T* Ai, long lda, const T* Br,
                      const T* Bi, long ldb, T* Cr, T* Ci, long ldc, int mode, int nb = 1, long sa = 0, long sb = 0,
                      long sc = 0, long cap = 0, int role = 0) {
    if (M <= 0 || N <= 0 || K <= 0 || nb <= 0) return;
    switch (cfg) {
        case 1: cgemm_cfg<T, 1>(s, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc, cap, role); break;
        case 2: cgemm_cfg<T, 2>(s, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc, cap, role); break;
        case 3: cgemm_cfg<T, 3>(s, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc, cap, role); break;
        case 4: cgemm_cfg<T, 4>(s, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc, cap, role); break;
        default: cgemm_cfg<T, 0>(s, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, nb, sa, sb, sc, cap, role); break;
    }
}

// Blocks of configuration cfg that fit on one CU (registers, LDS, waves).
template <typename T, int CFG>
inline int cgemm_occ_cfg() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, reinterpret_cast<const void*>(k_cgemm<T, CFG>), 256, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}
template <typename T>
inline int cgemm_occupancy(int cfg) {
    switch (cfg) {
        case 1: return cgemm_occ_cfg<T, 1>();
        case 2: return cgemm_occ_cfg<T, 2>();
        case 3: return cgemm_occ_cfg<T, 3>();
        case 4: return cgemm_occ_cfg<T, 4>();
        default: return cgemm_occ_cfg<T, 0>();
    }
}
static_assert(kGemmCfgs <= GemmPlan::kMaxCfgs, "GemmPlan holds one occupancy per configuration");

// Grid cap that leaves one block of the configuration free on every CU: a
// trailing update launched this way never takes all of a CU's LDS, so the
// panel kernels running beside it on the other stream find room at once.
inline long gemm_headroom_cap(const GemmPlan& p, int cfg) {
    if (p.cap > 0) return p.cap;
    const int per_cu = p.occ[cfg] > 1 ? p.occ[cfg] - 1 : 1;
    return (long)p.cus * per_cu;
}

// Blocks of configuration cfg for an M x N product.
template <typename T>
inline long gemm_blocks(int cfg, int M, int N) {
    int bm, bn;
    switch (cfg) {
        case 1: bm = GemmCfg<T, 1>::BM; bn = GemmCfg<T, 1>::BN; break;
        case 2: bm = GemmCfg<T, 2>::BM; bn = GemmCfg<T, 2>::BN; break;
        case 3: bm = GemmCfg<T, 3>::BM; bn = GemmCfg<T, 3>::BN; break;
        case 4: bm = GemmCfg<T, 4>::BM; bn = GemmCfg<T, 4>::BN; break;
        default: bm = GemmCfg<T, 0>::BM; bn = GemmCfg<T, 0>::BN; break;
    }
    return (long)((M + bm - 1) / bm) * ((N + bn - 1) / bn);
}

// The plan's configuration for an M x N product: the large-product tile when
// it still gives at least `fill` blocks, else the skinny-product tile while
// that gives one block per CU, else the 32 x 32 tile.
template <typename T>
inline int gemm_choose(const bem_ctx* c, int M, int N) {
    const GemmPlan& p = c->gemm_plan[sizeof(T) == 4 ? 0 : 1];
    if (gemm_blocks<T>(p.large, M, N) >= p.fill) return p.large;
    return gemm_blocks<T>(p.skinny, M, N) >= p.cus ? p.skinny : kGemmSmall;
}

// ---------------------------------------------------------------------------
// FP32 products with a short K (K = 16 KQ, at most 128): C (M x N) op= A B in
// the 3M form, A k-contiguous (row-major, 16-byte aligned rows), B row-major.
// No LDS and no barrier: every wave owns one 16 x 16 tile of C over the whole
// K and issues all of its operand loads before the first MFMA (A as 16-byte
// pieces of its row, B as single values of its column).  For the products on
// the factorisation's panel chain (the rank-32 updates of a panel step, the
// L11^{-1} doubling), which run beside the trailing update: there a kernel
// that stages operands through LDS with a barrier per slice slows down about
// three times.  The MFMA takes the B fragment first, so each lane ends with
// four consecutive columns of one row of C.  Batched over blockIdx.y with
// strides sa, sb, sc.  mode & 3: 0 C -= AB, 1 C = AB, 2 C += AB, 3 C = -AB;
// mode & GEMM_MODE_URGENT raises the waves' issue priority.
// ---------------------------------------------------------------------------
template <int KQ>
__global__ __launch_bounds__(256) void k_cgemm_rd(int M, int N, const float* __restrict__ Ar,
                                                  const float* __restrict__ Ai, long lda,
                                                  const float* __restrict__ Br, const float* __restrict__ Bi,
                                                  long ldb, float* __restrict__ Cr, float* __restrict__ Ci, long ldc,
                                                  int mode, long sa, long sb, long sc) {
    if (mode & GEMM_MODE_URGENT) __builtin_amdgcn_s_setprio(3);
    const int ntc = N / 16, ntiles = (M / 16) * ntc;
    const int tile = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (tile >= ntiles) return;
    const long z = blockIdx.y;
    Ar += z * sa;
    Ai += z * sa;
    Br += z * sb;
    Bi += z * sb;
    Cr += z * sc;
    Ci += z * sc;
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    const int m0 = (tile / ntc) * 16, n0 = (tile % ntc) * 16;
    const int op = mode & 3;
    v4f a0[KQ], a1[KQ];
    float b0[KQ][4], b1[KQ][4];
    const long oa = (long)(m0 + r) * lda + 4 * g;
#pragma unroll
    for (int q = 0; q < KQ; q++) {
        a0[q] = *reinterpret_cast<const v4f*>(Ar + oa + 16 * q);
        a1[q] = *reinterpret_cast<const v4f*>(Ai + oa + 16 * q);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const long ob = (long)(16 * q + 4 * g + e) * ldb + n0 + r;
            b0[q][e] = Br[ob];
            b1[q][e] = Bi[ob];
        }
    }
    const long oc = (long)(m0 + r) * ldc + n0 + 4 * g;
    v4f cr = v4f{0.f, 0.f, 0.f, 0.f}, ci = cr;
    if (op == 0 || op == 2) {
        cr = *reinterpret_cast<const v4f*>(Cr + oc);
        ci = *reinterpret_cast<const v4f*>(Ci + oc);
    }
    __builtin_amdgcn_sched_barrier(0);  // every load in flight before the first MFMA
    v4f t1 = v4f{0.f, 0.f, 0.f, 0.f}, t2 = t1, t3 = t1;
#pragma unroll
    for (int q = 0; q < KQ; q++) {
        const v4f a2 = a0[q] + a1[q];
#pragma unroll
        for (int e = 0; e < 4; e++) {
            t1 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[q][e], a0[q][e], t1, 0, 0, 0);
            t2 = __builtin_amdgcn_mfma_f32_16x16x4f32(b1[q][e], a1[q][e], t2, 0, 0, 0);
            t3 = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[q][e] + b1[q][e], a2[e], t3, 0, 0, 0);
        }
    }
    const v4f pr = t1 - t2, pi = t3 - t1 - t2;
    v4f outr, outi;
    if (op == 0) {
        outr = cr - pr;
        outi = ci - pi;
    } else if (op == 1) {
        outr = pr;
        outi = pi;
    } else if (op == 2) {
        outr = cr + pr;
        outi = ci + pi;
    } else {
        outr = -pr;
        outi = -pi;
    }
    *reinterpret_cast<v4f*>(Cr + oc) = outr;
    *reinterpret_cast<v4f*>(Ci + oc) = outi;
}

// Whether k_cgemm_rd takes the product (FP32, K in {32, 64, 128}).
inline bool cgemm_rd_fits(int M, int N, int K, const float* Ar, const float* Ai, long lda, const float* Cr,
                          const float* Ci, long ldc, long sa = 0, long sc = 0) {
    auto al = [](const float* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    return M > 0 && N > 0 && M % 16 == 0 && N % 16 == 0 && (K == 32 || K == 64 || K == 128) && lda % 4 == 0 &&
           ldc % 4 == 0 && sa % 4 == 0 && sc % 4 == 0 && al(Ar) && al(Ai) && al(Cr) && al(Ci);
}

inline void cgemm_rd(hipStream_t s, int M, int N, int K, const float* Ar, const float* Ai, long lda, const float* Br,
                     const float* Bi, long ldb, float* Cr, float* Ci, long ldc, int mode, int nb = 1, long sa = 0,
                     long sb = 0, long sc = 0) {
    const long tiles = (long)(M / 16) * (N / 16);
    const dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)nb);
    if (K == 32)
        hipLaunchKernelGGL(k_cgemm_rd<2>, grid, dim3(256), 0, s, M, N, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, sa,
                           sb, sc);
    else if (K == 64)
        hipLaunchKernelGGL(k_cgemm_rd<4>, grid, dim3(256), 0, s, M, N, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, sa,
                           sb, sc);
    else
        hipLaunchKernelGGL(k_cgemm_rd<8>, grid, dim3(256), 0, s, M, N, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, sa,
                           sb, sc);
}

// Launch flags of cgemm_ctx: GEMM_HEADROOM caps the grid as
// gemm_headroom_cap does (an update running beside the panel stream);
// GEMM_PANEL marks a product of the panel stream's critical path: the plan's
// panel configuration at raised issue priority; GEMM_UPDATE: an LU trailing
// update (its own kernel symbol and timing class).
constexpr int GEMM_HEADROOM = 1, GEMM_PANEL = 2, GEMM_UPDATE = 4;

// GEMM with optional per-launch timing, on stream s (default: the context
// stream), tile configuration from the context's plan.  Counted work: 3M
// form 6 M N K real flops (FP32), 4M form 8 M N K (FP64); bytes = operands
// read once plus C read and written.
template <typename T>
inline void cgemm_ctx(bem_ctx* c, int M, int N, int K, const T* Ar, const T* Ai, long lda, const T* Br, const T* Bi,
                      long ldb, T* Cr, T* Ci, long ldc, int mode, hipStream_t s = nullptr, int flags = 0) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    if (!s) s = c->stream;
    const bool upd = flags & GEMM_UPDATE;
    const int k = kt_mark(c, sizeof(T) == 4 ? (upd ? KS_UPDATE32 : KS_GEMM32) : (upd ? KS_UPDATE64 : KS_GEMM64), s);
    const GemmPlan& plan = c->gemm_plan[sizeof(T) == 4 ? 0 : 1];
    const int cfg = (flags & GEMM_PANEL) ? plan.panel : gemm_choose<T>(c, M, N);
    const long cap = (flags & GEMM_HEADROOM) ? gemm_headroom_cap(plan, cfg) : 0;
    if (flags & GEMM_PANEL) mode |= GEMM_MODE_URGENT;
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        if ((flags & GEMM_PANEL) && cgemm_rd_fits(M, N, K, Ar, Ai, lda, Cr, Ci, ldc)) {
            cgemm_rd(s, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode);
            done = true;
        }
    }
    if (!done) cgemm_any<T>(cfg, s, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, 1, 0, 0, 0, cap, upd ? 1 : 0);
    const double mnk = (double)M * N * K;
    const double bytes = 2.0 * sizeof(T) * ((double)M * K + (double)K * N + (mode == 1 ? 1.0 : 2.0) * M * N);
    kt_close(c, k, (sizeof(T) == 4 ? 6.0 : 8.0) * mnk, bytes);
}

// Sum of S split-K partial products (planes of M x N at P + 2 s M N, imaginary
// plane M N further) into C: C -= sum (mode 0) or C = sum (mode 1), summed in
// split order so results do not depend on scheduling.
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_sum(long count, int N, int S, const T* __restrict__ P,
                                                    T* __restrict__ Cr, T* __restrict__ Ci, long ldc, int mode) {
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= count) return;
    T vr = T(0), vi = T(0);
    for (int s = 0; s < S; s++) {
        vr += P[2 * s * count + idx];
        vi += P[(2 * s + 1) * count + idx];
    }
    const long o = (idx / N) * ldc + idx % N;
    if (mode == 0) {
        Cr[o] -= vr;
        Ci[o] -= vi;
    } else {
        Cr[o] = vr;
        Ci[o] = vi;
    }
}

// Splits for a product whose output tiles leave the device's block slots
// unevenly used (few tiles, long K): the smallest S in {1, 2, 4, 8, 16} whose
// last round of blocks is at least 90 % full (else the fullest), with every
// K slice a multiple of 16 and at least 512 long.
template <typename T>
inline int gemm_splits(bem_ctx* c, int cfg, int M, int N, int K) {
    const GemmPlan& p = c->gemm_plan[sizeof(T) == 4 ? 0 : 1];
    const long tiles = gemm_blocks<T>(cfg, M, N);
    const long slots = (long)p.cus * (p.occ[cfg] > 0 ? p.occ[cfg] : 1);
    int best = 1;
    double best_eff = 0.0;
    for (int S = 1; S <= 16; S *= 2) {
        if (S > 1 && (K % (16 * S) != 0 || K / S < 512)) break;
        const long blocks = tiles * S;
        const double eff = (double)blocks / (double)(slots * ((blocks + slots - 1) / slots));
        if (eff > best_eff + 1e-9) {
            best = S;
            best_eff = eff;
        }
        if (eff >= 0.9) break;
    }
    return best;
}

// One split-K product on configuration cfg with S splits: S batched partial
// products into P (2 S M N values), summed into C by k_splitk_sum.
template <typename T>
inline void splitk_launch(bem_ctx* c, int cfg, int S, int M, int N, int K, const T* Ar, const T* Ai, long lda,
                          const T* Br, const T* Bi, long ldb, T* P, T* Cr, T* Ci, long ldc, int mode, int role) {
    if (S == 1) {
        cgemm_any<T>(cfg, c->stream, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode, 1, 0, 0, 0, 0, role);
        return;
    }
    const long count = (long)M * N;
    const int Ks = K / S;
    cgemm_any<T>(cfg, c->stream, M, N, Ks, Ar, Ai, lda, Br, Bi, ldb, P, P + count, N, 1, S, (long)Ks, (long)Ks * ldb,
                 2 * count, 0, role);
    hipLaunchKernelGGL(k_splitk_sum<T>, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, c->stream, count, N, S,
                       (const T*)P, Cr, Ci, ldc, mode);
}

// C = C - A B (mode 0) or C = A B (mode 1) for products with few output tiles
// and a long K (the refinement residual), with K split over S batched
// products into the context's partial buffer, then summed in split order
// (k_splitk_sum).  The configuration is chosen once per shape by timing
// each one (with its own S) into a scratch output; the choice changes speed
// only.
template <typename T>
inline int cgemm_splitk_ctx(bem_ctx* c, int M, int N, int K, const T* Ar, const T* Ai, long lda, const T* Br,
                            const T* Bi, long ldb, T* Cr, T* Ci, long ldc, int mode, int role = 0) {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    gemm_plan_occupancy(c, sizeof(T) == 4);
    GemmPlan& plan = c->gemm_plan[sizeof(T) == 4 ? 0 : 1];
    const long count = (long)M * N;
    int Smax = 1;
    for (int cfg = 0; cfg < kGemmCfgs; cfg++) Smax = std::max(Smax, gemm_splits<T>(c, cfg, M, N, K));
    const size_t bytes = (2 * (size_t)Smax + 2) * count * sizeof(T);  // partials + a scratch output
    if (c->part_bytes < bytes) {
        if (c->part) (void)hipFree(c->part);
        c->part = nullptr;
        c->part_bytes = 0;
        if (hipMalloc(&c->part, bytes) != hipSuccess) return bem_fail(c, BEM_ERR_NOMEM, "split-K partial products");
        c->part_bytes = bytes;
    }
    T* P = static_cast<T*>(c->part);
    if (plan.sk_cfg < 0 || plan.sk_shape[0] != M || plan.sk_shape[1] != N || plan.sk_shape[2] != K ||
        plan.sk_shape[3] != role) {
        T* Xr = P + 2 * (size_t)Smax * count;  // scratch output, ld N
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            if (e0) (void)hipEventDestroy(e0);
            return bem_fail(c, BEM_ERR_HIP, "split-K tuning: events");
        }
        float best_ms = 0.0f;
        int best = gemm_choose<T>(c, M, N);
        for (int cfg = 0; cfg < kGemmCfgs; cfg++) {
            const int S = gemm_splits<T>(c, cfg, M, N, K);
            (void)hipEventRecord(e0, c->stream);
            splitk_launch<T>(c, cfg, S, M, N, K, Ar, Ai, lda, Br, Bi, ldb, P, Xr, Xr + count, N, 1, role);
            (void)hipEventRecord(e1, c->stream);
            if (hipEventSynchronize(e1) != hipSuccess) break;
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms > 0.0f && (best_ms == 0.0f || ms < best_ms)) {
                best_ms = ms;
                best = cfg;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        plan.sk_cfg = best;
        plan.sk_shape[0] = M;
        plan.sk_shape[1] = N;
        plan.sk_shape[2] = K;
        plan.sk_shape[3] = role;
    }
    const int cfg = plan.sk_cfg;
    const int S = gemm_splits<T>(c, cfg, M, N, K);
    const double fl = (sizeof(T) == 4 || (role == 2 && cfg != 0) ? 6.0 : 8.0) * (double)M * N * K;
    const double b = 2.0 * sizeof(T) * ((double)M * K + (double)K * N + (2.0 * S + (mode == 1 ? 1.0 : 2.0)) * count);
    const int k = kt_mark(c, sizeof(T) == 4 ? KS_GEMM32 : KS_GEMM64, c->stream);
    splitk_launch<T>(c, cfg, S, M, N, K, Ar, Ai, lda, Br, Bi, ldb, P, Cr, Ci, ldc, mode, role);
    kt_close(c, k, fl, b);
    return 0;
}

// ---------------------------------------------------------------------------
// Short-K products, K = 4 * kShortKW: C (M x N) op= A B with one 16 x 16
// output tile per block and the block's four waves each taking a quarter of
// K.  The triangular solves' diagonal-block products (256 x ncol x 256) and
// their last updates have too few output tiles and too short a K for the
// LDS-staged kernel to hide its load latency; here every wave issues all its
// loads at once, straight into MFMA operand registers.  Lane (r, g) =
// (lane % 16, lane / 16) holds A[m0 + r][k + 4g + i] (one 16-byte row chunk)
// and B[k + 4g + i][n0 + r] for MFMA i = 0..3 of each 16-wide K step: A and
// B see the same order of k.  The four wave tiles are summed through LDS in
// wave order.  Blocks b with one value of b % 8 (one XCD) take a contiguous
// run of the row-major tile order, so the column tiles of a row share that
// XCD's L2 copy of A's rows.  FP32: 3M form; FP64: 4M form (as k_cgemm).
// ---------------------------------------------------------------------------
constexpr int kShortKW = 64;  // K per wave

template <typename T>
__device__ __forceinline__ void load_row4(const T* __restrict__ p, T o[4]);
template <>
__device__ __forceinline__ void load_row4<float>(const float* __restrict__ p, float o[4]) {
    const float4 v = *reinterpret_cast<const float4*>(p);
    o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
}
template <>
__device__ __forceinline__ void load_row4<double>(const double* __restrict__ p, double o[4]) {
    const double2 v0 = *reinterpret_cast<const double2*>(p), v1 = *reinterpret_cast<const double2*>(p + 2);
    o[0] = v0.x; o[1] = v0.y; o[2] = v1.x; o[3] = v1.y;
}

template <typename T>
__global__ __launch_bounds__(256) void k_cgemm_short(int M, int N, const T* __restrict__ Ar, const T* __restrict__ Ai,
                                                     long lda, const T* __restrict__ Br, const T* __restrict__ Bi,
                                                     long ldb, T* __restrict__ Cr, T* __restrict__ Ci, long ldc,
                                                     int mode, T* __restrict__ Ctr, T* __restrict__ Cti, long ldt) {
    constexpr bool M3 = sizeof(T) == 4;
    constexpr int NACC = M3 ? 3 : 2;
    constexpr int NS = kShortKW / 16;  // K steps per wave
    typedef typename std::conditional<M3, v4f, v4d>::type acc_t;
    __shared__ T red[4][2][16][17];
    const int nct = N / 16, G = (M / 16) * nct;
    const int per = (G + 7) / 8;
    const int f = (blockIdx.x % 8) * per + blockIdx.x / 8;
    if (f >= G) return;
    const int m0 = (f / nct) * 16, n0 = (f % nct) * 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int r = lane & 15, g = lane >> 4;
    T ar[NS][4], ai[NS][4], br[NS][4], bi[NS][4];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int k = w * kShortKW + 16 * s + 4 * g;
        const long oa = (long)(m0 + r) * lda + k;
        load_row4<T>(Ar + oa, ar[s]);
        load_row4<T>(Ai + oa, ai[s]);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const long ob = (long)(k + i) * ldb + n0 + r;
            br[s][i] = Br[ob];
            bi[s][i] = Bi[ob];
        }
    }
    // every load in flight before the first MFMA waits on its operands
    __builtin_amdgcn_sched_barrier(0);
    acc_t acc[NACC];
#pragma unroll
    for (int a = 0; a < NACC; a++) acc[a] = acc_t{};
#pragma unroll
    for (int s = 0; s < NS; s++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if constexpr (M3) {
                acc[0] = mfma16x16x4(ar[s][i], br[s][i], acc[0]);
                acc[1] = mfma16x16x4(ai[s][i], bi[s][i], acc[1]);
                acc[2] = mfma16x16x4(ar[s][i] + ai[s][i], br[s][i] + bi[s][i], acc[2]);
            } else {
                acc[0] = mfma16x16x4(ar[s][i], br[s][i], acc[0]);
                acc[0] = mfma16x16x4(ai[s][i], -bi[s][i], acc[0]);
                acc[1] = mfma16x16x4(ar[s][i], bi[s][i], acc[1]);
                acc[1] = mfma16x16x4(ai[s][i], br[s][i], acc[1]);
            }
        }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int row = M3 ? 4 * g + q : g + 4 * q;
        if constexpr (M3) {
            red[w][0][row][r] = acc[0][q] - acc[1][q];
            red[w][1][row][r] = acc[2][q] - acc[0][q] - acc[1][q];
        } else {
            red[w][0][row][r] = acc[0][q];
            red[w][1][row][r] = acc[1][q];
        }
    }
    __syncthreads();
    const int row = threadIdx.x / 16, col = threadIdx.x % 16;
    const T vr = ((red[0][0][row][col] + red[1][0][row][col]) + red[2][0][row][col]) + red[3][0][row][col];
    const T vi = ((red[0][1][row][col] + red[1][1][row][col]) + red[2][1][row][col]) + red[3][1][row][col];
    const long o = (long)(m0 + row) * ldc + n0 + col;
    if (mode == 0) {
        Cr[o] -= vr;
        Ci[o] -= vi;
    } else if (mode == 1) {
        Cr[o] = vr;
        Ci[o] = vi;
    } else if (mode == 3) {
        Cr[o] = -vr;
        Ci[o] = -vi;
    } else {
        Cr[o] += vr;
        Ci[o] += vi;
    }
    if (Ctr) {  // mode 1 only: C^T as well, consecutive threads on consecutive rows of C
        const int tr = threadIdx.x % 16, tc = threadIdx.x / 16;
        const T ur = ((red[0][0][tr][tc] + red[1][0][tr][tc]) + red[2][0][tr][tc]) + red[3][0][tr][tc];
        const T ui = ((red[0][1][tr][tc] + red[1][1][tr][tc]) + red[2][1][tr][tc]) + red[3][1][tr][tc];
        const long ot = (long)(n0 + tc) * ldt + m0 + tr;
        Ctr[ot] = ur;
        Cti[ot] = ui;
    }
}

// Whether k_cgemm_short takes the product: K = 4 kShortKW, M and N multiples
// of 16, A's rows 16-byte aligned.
template <typename T>
inline bool cgemm_short_fits(int M, int N, int K, const T* Ar, const T* Ai, long lda) {
    return K == 4 * kShortKW && M > 0 && N > 0 && M % 16 == 0 && N % 16 == 0 && lda % (16 / sizeof(T)) == 0 &&
           reinterpret_cast<uintptr_t>(Ar) % 16 == 0 && reinterpret_cast<uintptr_t>(Ai) % 16 == 0;
}

// Short-K product on the context stream with per-launch timing; products
// the kernel does not take go to cgemm_ctx.  With Ctr (mode 1, and only for
// products the kernel takes) C^T is written to (Ctr, Cti) too, N x M with
// leading dimension ldt.
template <typename T>
inline void cgemm_short_ctx(bem_ctx* c, int M, int N, int K, const T* Ar, const T* Ai, long lda, const T* Br,
                            const T* Bi, long ldb, T* Cr, T* Ci, long ldc, int mode, T* Ctr = nullptr,
                            T* Cti = nullptr, long ldt = 0) {
    if (!cgemm_short_fits<T>(M, N, K, Ar, Ai, lda)) {
        cgemm_ctx<T>(c, M, N, K, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci, ldc, mode);
        return;
    }
    const int k = kt_mark(c, sizeof(T) == 4 ? KS_GEMM32 : KS_GEMM64, c->stream);
    const long tiles = (long)(M / 16) * (N / 16);
    const unsigned grid = (unsigned)(((tiles + 7) / 8) * 8);
    hipLaunchKernelGGL(k_cgemm_short<T>, dim3(grid), dim3(256), 0, c->stream, M, N, Ar, Ai, lda, Br, Bi, ldb, Cr, Ci,
                       ldc, mode, mode == 1 ? Ctr : nullptr, Cti, ldt);
    const double mnk = (double)M * N * K;
    const double bytes = 2.0 * sizeof(T) * ((double)M * K + (double)K * N + (mode == 1 ? 1.0 : 2.0) * M * N);
    kt_close(c, k, (sizeof(T) == 4 ? 6.0 : 8.0) * mnk, bytes);
}
