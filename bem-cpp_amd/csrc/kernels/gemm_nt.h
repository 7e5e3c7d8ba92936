// gemm_nt.h -- the FP32 trailing update of the LU on the matrix cores,
//
//   C (M x N) -= A (M x K) Bt (N x K)^T,
//
// planar complex FP32 in the 3M form (T1 = Ar Br, T2 = Ai Bi,
// T3 = (Ar + Ai)(Br + Bi); Re = T1 - T2, Im = T3 - T1 - T2), with both
// operands k-contiguous: A is the panel L21 as it lies in the matrix, Bt the
// transpose of U12 (written by k_u12_out beside the copy of U12 into the
// matrix).
//
// Why both operands k-contiguous: v_mfma_f32_16x16x4f32 takes one k per lane
// group g = lane / 16.  Four consecutive MFMAs s = 0..3 use k = 4g + s, so a
// lane's operands for the four are four consecutive floats of one row of A or
// of Bt: one 16-byte LDS read (ds_read_b128) instead of four, and the global
// tiles go to LDS as they are (no transposing stores).
//
// Output orientation: the MFMA's first operand is the Bt fragment and its
// second the A fragment, so the result D = (A Bt^T)^T of a 16 x 16 tile sits
// in lane m + 16 (n / 4), register n % 4: each lane holds four consecutive
// columns of one row of C, read and written as one 16-byte vector.  The C
// tile is loaded at the start of the tile's K loop, so its latency hides
// behind the matrix-core work, and stored after it.
//
// Tiling: 256 threads = 4 waves in a (BM / WM) x (BN / WN) arrangement, each
// wave a WM x WN sub-tile of 16 x 16 MFMA tiles; K slices of BK staged in LDS
// ([row][k], padded to a stride of BK + 4 floats: the 16-byte reads of the 16
// rows of a lane group fall in distinct banks), double buffered, one barrier
// per slice.  Tiles are spread XCD-aware: block b runs on XCD b % 8, and the
// blocks of one XCD take one contiguous run of the tile order, which walks
// GROUP tile rows at a time column by column, so the XCD's L2 holds the A and
// Bt rows its resident blocks share.
//
// Requirements (checked by the launcher): M % BM == 0, N % BN == 0,
// K % BK == 0, lda and ldb multiples of 4, ldc a multiple of 4, 16-byte
// aligned planes.  The sum over k runs in the same order for every tile
// configuration of this kernel, so the choice changes speed only.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

typedef float v4f_nt __attribute__((ext_vector_type(4)));

template <int BM_, int BN_, int WM_, int WN_, int BK_, int MINB_>
struct NtCfg {
    static constexpr int BM = BM_, BN = BN_, WM = WM_, WN = WN_, BK = BK_, MINB = MINB_;
    static constexpr int SK = BK + 4;  // LDS row stride (floats)
};

constexpr int kNtGroup = 8;  // tile rows walked together (L2 reuse of Bt)

template <class C>
__global__ __launch_bounds__(256, C::MINB) void k_cupdate_nt(int M, int N, int K, const float* __restrict__ Ar,
                                                             const float* __restrict__ Ai, long lda,
                                                             const float* __restrict__ Btr,
                                                             const float* __restrict__ Bti, long ldb,
                                                             float* __restrict__ Cr, float* __restrict__ Ci, long ldc) {
    constexpr int BM = C::BM, BN = C::BN, WM = C::WM, WN = C::WN, BK = C::BK, SK = C::SK;
    constexpr int TM = WM / 16, TN = WN / 16;
    constexpr int WGN = BN / WN;  // waves along n
    static_assert((BM / WM) * (BN / WN) == 4, "four waves per block");
    constexpr int AV = BM * BK / 4 / 256, BV = BN * BK / 4 / 256;  // 16-byte vectors per thread per plane
    static_assert(AV * 256 * 4 == BM * BK && BV * 256 * 4 == BN * BK, "slices split evenly over the threads");
    constexpr int KV = BK / 4;  // 16-byte vectors per row of a slice
    typedef v4f_nt f4;

    __shared__ f4 sA[2][2][BM][SK / 4];
    __shared__ f4 sB[2][2][BN][SK / 4];

    // tiles of this block (XCD-aware, grouped order; a grid smaller than the
    // tile count loops, every XCD over its own run of tiles)
    const int ntx = N / BN, nty = M / BM, ntiles = ntx * nty;
    const int per = (ntiles + 7) / 8, xcd = blockIdx.x % 8, stride = gridDim.x / 8;
    for (int jt = blockIdx.x / 8; jt < per; jt += stride) {
        const int f = xcd * per + jt;
        if (f >= ntiles) break;
        const int gsz = kNtGroup * ntx;
        const int g0 = (f / gsz) * kNtGroup;
        const int grows = min(kNtGroup, nty - g0);
        const int fin = f % gsz;
        const int row0 = (g0 + fin % grows) * BM, col0 = (fin / grows) * BN;

        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        const int wm0 = (wave / WGN) * WM, wn0 = (wave % WGN) * WN;
        const int r16 = lane & 15, g4 = lane >> 4;

        // C tile of this lane: rows row0 + wm0 + 16 i + r16, columns
        // col0 + wn0 + 16 j + 4 g4 .. + 3; loaded now, used after the K loop
        f4 cr[TM][TN], ci[TM][TN];
#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const long o = (long)(row0 + wm0 + 16 * i + r16) * ldc + col0 + wn0 + 16 * j + 4 * g4;
                cr[i][j] = *reinterpret_cast<const f4*>(Cr + o);
                ci[i][j] = *reinterpret_cast<const f4*>(Ci + o);
            }

        // this thread's 16-byte pieces of a slice: row v / KV, k (v % KV) * 4
        const float* pa[2][AV];
        const float* pb[2][BV];
        int sa_row[AV], sa_kq[AV], sb_row[BV], sb_kq[BV];
#pragma unroll
        for (int u = 0; u < AV; u++) {
            const int v = tid + 256 * u;
            sa_row[u] = v / KV;
            sa_kq[u] = v % KV;
            const long o = (long)(row0 + sa_row[u]) * lda + 4 * sa_kq[u];
            pa[0][u] = Ar + o;
            pa[1][u] = Ai + o;
        }
#pragma unroll
        for (int u = 0; u < BV; u++) {
            const int v = tid + 256 * u;
            sb_row[u] = v / KV;
            sb_kq[u] = v % KV;
            const long o = (long)(col0 + sb_row[u]) * ldb + 4 * sb_kq[u];
            pb[0][u] = Btr + o;
            pb[1][u] = Bti + o;
        }

        f4 ga[2][AV], gb[2][BV];
#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int u = 0; u < AV; u++) ga[p][u] = *reinterpret_cast<const f4*>(pa[p][u]);
#pragma unroll
            for (int u = 0; u < BV; u++) gb[p][u] = *reinterpret_cast<const f4*>(pb[p][u]);
        }
#pragma unroll
        for (int p = 0; p < 2; p++) {
#pragma unroll
            for (int u = 0; u < AV; u++) sA[0][p][sa_row[u]][sa_kq[u]] = ga[p][u];
#pragma unroll
            for (int u = 0; u < BV; u++) sB[0][p][sb_row[u]][sb_kq[u]] = gb[p][u];
        }

        f4 acc[3][TM][TN];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int i = 0; i < TM; i++)
#pragma unroll
                for (int j = 0; j < TN; j++) acc[a][i][j] = f4{0.f, 0.f, 0.f, 0.f};

        __syncthreads();
        const int nk = K / BK;
        for (int kt = 0; kt < nk; kt++) {
            const int buf = kt & 1;
            const bool more = kt + 1 < nk;
            if (more) {
                const int k1 = (kt + 1) * BK;
#pragma unroll
                for (int p = 0; p < 2; p++) {
#pragma unroll
                    for (int u = 0; u < AV; u++) ga[p][u] = *reinterpret_cast<const f4*>(pa[p][u] + k1);
#pragma unroll
                    for (int u = 0; u < BV; u++) gb[p][u] = *reinterpret_cast<const f4*>(pb[p][u] + k1);
                }
            }
#pragma unroll
            for (int kk = 0; kk < BK / 16; kk++) {
                f4 a0[TM], a1[TM], a2[TM], b0[TN], b1[TN], b2[TN];
#pragma unroll
                for (int i = 0; i < TM; i++) {
                    a0[i] = sA[buf][0][wm0 + 16 * i + r16][4 * kk + g4];
                    a1[i] = sA[buf][1][wm0 + 16 * i + r16][4 * kk + g4];
                    a2[i] = a0[i] + a1[i];
                }
#pragma unroll
                for (int j = 0; j < TN; j++) {
                    b0[j] = sB[buf][0][wn0 + 16 * j + r16][4 * kk + g4];
                    b1[j] = sB[buf][1][wn0 + 16 * j + r16][4 * kk + g4];
                    b2[j] = b0[j] + b1[j];
                }
#pragma unroll
                for (int s = 0; s < 4; s++)
#pragma unroll
                    for (int i = 0; i < TM; i++)
#pragma unroll
                        for (int j = 0; j < TN; j++) {
                            acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b0[j][s], a0[i][s], acc[0][i][j], 0, 0, 0);
                            acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b1[j][s], a1[i][s], acc[1][i][j], 0, 0, 0);
                            acc[2][i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b2[j][s], a2[i][s], acc[2][i][j], 0, 0, 0);
                        }
            }
            if (more) {
#pragma unroll
                for (int p = 0; p < 2; p++) {
#pragma unroll
                    for (int u = 0; u < AV; u++) sA[buf ^ 1][p][sa_row[u]][sa_kq[u]] = ga[p][u];
#pragma unroll
                    for (int u = 0; u < BV; u++) sB[buf ^ 1][p][sb_row[u]][sb_kq[u]] = gb[p][u];
                }
            }
            __syncthreads();
        }

#pragma unroll
        for (int i = 0; i < TM; i++)
#pragma unroll
            for (int j = 0; j < TN; j++) {
                const f4 t1 = acc[0][i][j], t2 = acc[1][i][j], t3 = acc[2][i][j];
                const f4 r = cr[i][j] - (t1 - t2);
                const f4 m = ci[i][j] - (t3 - t1 - t2);
                const long o = (long)(row0 + wm0 + 16 * i + r16) * ldc + col0 + wn0 + 16 * j + 4 * g4;
                *reinterpret_cast<f4*>(Cr + o) = r;
                *reinterpret_cast<f4*>(Ci + o) = m;
            }
    }  // tiles
}

// Tile configurations of the update (GemmPlan::nt): all sum k in the same
// order, so they give the same bits.  (A persistent grid of 2.5 blocks per CU,
// leaving slots free for the panel kernels of the other stream, measured
// 171 ms against 150 ms for the C5 factorisation: not kept.)
constexpr int kNtCfgs = 4;
typedef NtCfg<64, 64, 32, 32, 16, 3> NtCfg0;
typedef NtCfg<64, 64, 32, 32, 16, 2> NtCfg1;
typedef NtCfg<128, 64, 64, 32, 16, 1> NtCfg2;
typedef NtCfg<64, 64, 32, 32, 32, 2> NtCfg3;
// Not an update configuration: the narrow products of the triangular solves
// (N = right-hand sides), where 32 x 32 tiles give four times the blocks.
typedef NtCfg<32, 32, 16, 16, 32, 4> NtCfg4;

// Whether configuration cfg takes the product (shape and alignment).
inline bool cupdate_nt_fits(int cfg, int M, int N, int K, const float* Ar, const float* Ai, long lda,
                            const float* Btr, const float* Bti, long ldb, const float* Cr, const float* Ci, long ldc) {
    const int bm = cfg == 2 ? 128 : cfg == 4 ? 32 : 64, bn = cfg == 4 ? 32 : 64, bk = cfg >= 3 ? 32 : 16;
    auto al = [](const float* p) { return reinterpret_cast<uintptr_t>(p) % 16 == 0; };
    return M > 0 && N > 0 && M % bm == 0 && N % bn == 0 && K % bk == 0 && lda % 4 == 0 && ldb % 4 == 0 &&
           ldc % 4 == 0 && al(Ar) && al(Ai) && al(Btr) && al(Bti) && al(Cr) && al(Ci);
}

template <class C>
inline void cupdate_nt_cfg(hipStream_t s, int M, int N, int K, const float* Ar, const float* Ai, long lda,
                           const float* Btr, const float* Bti, long ldb, float* Cr, float* Ci, long ldc, long cap) {
    const long tiles = (long)(M / C::BM) * (N / C::BN);
    long grid = (tiles + 7) / 8 * 8;
    if (cap > 0 && cap < grid) grid = cap < 8 ? 8 : cap / 8 * 8;  // whole XCD rounds
    hipLaunchKernelGGL((k_cupdate_nt<C>), dim3((unsigned)grid), dim3(256), 0, s, M, N, K, Ar, Ai, lda, Btr, Bti, ldb,
                       Cr, Ci, ldc);
}

// C -= A Bt^T with configuration cfg (cupdate_nt_fits must hold); cap > 0
// limits the grid (rounded to a multiple of 8 blocks) and the blocks loop.
inline void cupdate_nt(int cfg, hipStream_t s, int M, int N, int K, const float* Ar, const float* Ai, long lda,
                       const float* Btr, const float* Bti, long ldb, float* Cr, float* Ci, long ldc, long cap = 0) {
    switch (cfg) {
        case 1: cupdate_nt_cfg<NtCfg1>(s, M, N, K, Ar, Ai, lda, Btr, Bti, ldb, Cr, Ci, ldc, cap); break;
        case 2: cupdate_nt_cfg<NtCfg2>(s, M, N, K, Ar, Ai, lda, Btr, Bti, ldb, Cr, Ci, ldc, cap); break;
        case 3: cupdate_nt_cfg<NtCfg3>(s, M, N, K, Ar, Ai, lda, Btr, Bti, ldb, Cr, Ci, ldc, cap); break;
        case 4: cupdate_nt_cfg<NtCfg4>(s, M, N, K, Ar, Ai, lda, Btr, Bti, ldb, Cr, Ci, ldc, cap); break;
        default: cupdate_nt_cfg<NtCfg0>(s, M, N, K, Ar, Ai, lda, Btr, Bti, ldb, Cr, Ci, ldc, cap); break;
    }
}
