// Dense complex linear algebra for gfx950: MFMA complex GEMM, blocked LU with
// partial pivoting (cooperative LDS-resident panel), triangular solves,
// FP64 residual, GEMV and precision conversion.
//
// All matrices are row-major complex (interleaved re, im).  The LU is the
// right-looking blocked algorithm:
//   for each panel of nb columns:
//     factor rows [k, n) x cols [k, k+nb) with row pivoting   (k_panel)
//     apply the panel's row swaps to the other columns         (k_laswp)
//     U12 = L11^{-1} A12                                       (k_trsm_l)
//     A22 -= L21 U12                                           (MFMA GEMM)
#include <math.h>
#include <stdint.h>

#include <atomic>
#include <cstring>
#include <mutex>
#include <vector>

#include "gemm_h3.hpp"
#include "gemm_h3_dev.hpp"
#include "mba_common.hpp"
#include "mba_guard.hpp"
#include "upload.hpp"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

// ============================================================== GEMM
// Complex product as a real MFMA: for one complex k, the real operands are
//   A'(i, 0..1) = (Re a_ik, Im a_ik)
//   B'(0, 2j..2j+1) = (Re b_kj, Im b_kj),  B'(1, 2j..2j+1) = (-Im b_kj, Re b_kj)
// so that C'(i, 2j) = Re c_ij and C'(i, 2j+1) = Im c_ij: the real output tile is
// exactly the row-major complex memory layout of C.
//
// FP32: v_mfma_f32_32x32x2f32, one instruction = 32 rows x 16 complex columns
// x 1 complex k.  Block tile 128 x (32 WN) complex, 4 waves in 2 x 2, each wave
// 64 x (16 WN) (2 x WN instructions), k-chunks of 16 staged in LDS (double
// buffer).  A is staged k-major as separate re / im planes so a wave's operand
// read is 32 consecutive floats; B is staged row-major real.
// XCD = true orders tiles for the 8 XCDs (workgroup L runs on XCD L % 8): each
// XCD takes a contiguous range of tiles, walked in groups of GROUP_M row tiles
// so an XCD's L2 keeps the group's A row blocks while its B columns stream.
constexpr int G32_BM = 128, G32_BK = 16, G32_GROUP_M = 8, NUM_XCD = 8;

__device__ __forceinline__ void tile_of(int64_t L, int64_t nbm, int64_t nbn, int64_t& tm, int64_t& tn)
{
    const int64_t T = nbm * nbn;
    const int64_t per = (T + NUM_XCD - 1) / NUM_XCD;
    const int64_t xcd = L % NUM_XCD, loc = L / NUM_XCD;
    int64_t v = xcd * per + loc;  // contiguous range per XCD
    if (v >= T) {
        tm = -1;
        tn = -1;
        return;
    }
    const int64_t group = v / (G32_GROUP_M * nbn);
    const int64_t first_m = group * G32_GROUP_M;
    const int64_t gm = (nbm - first_m) < G32_GROUP_M ? (nbm - first_m) : G32_GROUP_M;
    const int64_t in = v - group * G32_GROUP_M * nbn;
    tm = first_m + in % gm;
    tn = in / gm;
}

template <int WN, bool XCD>
__global__ __launch_bounds__(256) void k_cgemm_f32(int64_t M, int64_t N, int64_t K, float alpha,
                                                   const cpx<float>* __restrict__ A, int64_t lda,
                                                   const cpx<float>* __restrict__ B, int64_t ldb,
                                                   float beta, cpx<float>* __restrict__ C,
                                                   int64_t ldc)
{
    constexpr int BN = 32 * WN;       // complex columns per block
    constexpr int BPT = BN / 16;      // complex B elements per thread per chunk
    __shared__ alignas(16) float Are[2][G32_BK][G32_BM];
    __shared__ alignas(16) float Aim[2][G32_BK][G32_BM];
    __shared__ alignas(16) float Bs[2][G32_BK][2 * BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int64_t nbm = (M + G32_BM - 1) / G32_BM, nbn = (N + BN - 1) / BN;
    int64_t tmi, tni;
    if (XCD) {
        tile_of(blockIdx.x, nbm, nbn, tmi, tni);
        if (tmi < 0) return;
    } else {
        tmi = blockIdx.x / nbn;
        tni = blockIdx.x % nbn;
    }
    const int64_t m0 = tmi * G32_BM, n0 = tni * BN;

    // A: 128 rows x 16 complex: thread -> row tid/2, complex k (tid&1)*8 .. +8
    const int ar = tid >> 1, ak = (tid & 1) * 8;
    // B: 16 rows x BN complex: thread -> row tid/16, complex cols (tid&15)*BPT .. +BPT
    const int bkr = tid >> 4, bc = (tid & 15) * BPT;
    float2 ra[8], rb[BPT];

    auto load_chunk = [&](int64_t k0) {
        const int64_t gi = m0 + ar;
        const bool rowok = gi < M;
        const float2* pa = (const float2*)(A + (rowok ? gi : 0) * lda);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int64_t gk = k0 + ak + q;
            ra[q] = (rowok && gk < K) ? pa[gk] : make_float2(0.f, 0.f);
        }
        const int64_t gk = k0 + bkr;
        const bool kok = gk < K;
        const float2* pb = (const float2*)(B + (kok ? gk : 0) * ldb);
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int64_t gj = n0 + bc + q;
            rb[q] = (kok && gj < N) ? pb[gj] : make_float2(0.f, 0.f);
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            Are[buf][ak + q][ar] = ra[q].x;
            Aim[buf][ak + q][ar] = ra[q].y;
        }
#pragma unroll
        for (int q = 0; q < BPT; q += 2)
            *(float4*)&Bs[buf][bkr][2 * (bc + q)] = make_float4(rb[q].x, rb[q].y, rb[q + 1].x, rb[q + 1].y);
    };

    floatx16 acc[2][WN];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

    const int hi = lane >> 5, jj = lane & 31;
    const int bcol = hi ? (jj ^ 1) : jj;
    const float bsign = (hi && !(jj & 1)) ? -1.f : 1.f;
    const int nchunks = (int)((K + G32_BK - 1) / G32_BK);
    // C tile prefetched into registers so its HBM latency overlaps the MFMA loop
    float cpre[2][WN][16];
    if (beta != 0.f) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < WN; ++j) {
                const int64_t gjr = 2 * (n0 + wn * 16 * WN + j * 16) + jj;
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int64_t gi = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
                    cpre[i][j][r] = (gi < M && gjr < 2 * N) ? ((const float*)(C + gi * ldc))[gjr] : 0.f;
                }
            }
    }
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) load_chunk((int64_t)(c + 1) * G32_BK);
#pragma unroll
        for (int kk = 0; kk < G32_BK; ++kk) {
            float a[2], b[WN];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int row = wm * 64 + i * 32 + jj;
                a[i] = hi ? Aim[buf][kk][row] : Are[buf][kk][row];
            }
#pragma unroll
            for (int j = 0; j < WN; ++j) b[j] = bsign * Bs[buf][kk][wn * 32 * WN + j * 32 + bcol];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < WN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (c + 1 < nchunks) {
            __syncthreads();
            store_chunk(buf ^ 1);
            __syncthreads();
        }
    }
    // epilogue: register r of lane -> row 8*(r/4) + 4*hi + r%4, real column jj
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < WN; ++j) {
            const int64_t gjr = 2 * (n0 + wn * 16 * WN + j * 16) + jj;  // real column
            if (gjr >= 2 * N) continue;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
                if (gi < M) {
                    float* p = (float*)(C + gi * ldc) + gjr;
                    float v = alpha * acc[i][j][r];
                    if (beta != 0.f) v += beta * cpre[i][j][r];
                    *p = v;
                }
            }
        }
}

// 3M complex product on real MFMAs: with A = Ar + j Ai, B = Br + j Bi,
//   T1 = Ar Br, T2 = Ai Bi, T3 = (Ar + Ai)(Br + Bi),
//   Re C = T1 - T2, Im C = T3 - T1 - T2:
// three real products instead of the four of the real embedding (25 % fewer
// MFMAs; the imaginary part's rounding error grows by a small constant, which
// the refinement absorbs).  v_mfma_f32_32x32x2f32 takes 2 k per instruction:
// lane l supplies A[row l % 32][k l / 32] and B[k l / 32][col l % 32]; each
// lane reads (Ar, Ai) and (Br, Bi) once per k-pair and forms the sums in
// registers.  Block tile 128 x 64 complex, 4 waves in 2 x 2 (64 x 32 each:
// 2 x 1 instructions per product), k-chunks of 16 in LDS (double buffer), re
// and im planes stored k-major (A) and row-major (B).
constexpr int G3M_BN = 64;

// LDS of one tile: Ar, Ai [2][G32_BK][rows], Br, Bi [2][G32_BK][G3M_BN] floats
constexpr int G3M_LDS_FLOATS = 4 * G32_BK * G32_BM + 4 * G32_BK * G3M_BN;

template <int BMT = G32_BM>
__device__ __forceinline__ void cgemm3m_tile(int64_t M, int64_t N, int64_t K, float alpha,
                                             const cpx<float>* __restrict__ A, int64_t lda,
                                             const cpx<float>* __restrict__ B, int64_t ldb, float beta,
                                             cpx<float>* __restrict__ C, int64_t ldc, int64_t m0, int64_t n0,
                                             float* lds)
{
    static_assert(BMT == 128 || BMT == 64, "128 or 64 rows");
    constexpr int BN = G3M_BN;
    constexpr int MI = BMT / 64;               // 32-row blocks per wave
    constexpr int TPR = 256 / BMT, KPT = G32_BK / TPR;  // A staging: threads per row, k per thread
    float(*Ar)[G32_BK][BMT] = (float(*)[G32_BK][BMT])lds;
    float(*Ai)[G32_BK][BMT] = (float(*)[G32_BK][BMT])(lds + 2 * G32_BK * BMT);
    float(*Br)[G32_BK][BN] = (float(*)[G32_BK][BN])(lds + 4 * G32_BK * BMT);
    float(*Bi)[G32_BK][BN] = (float(*)[G32_BK][BN])(lds + 4 * G32_BK * BMT + 2 * G32_BK * BN);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // A: BMT rows x 16 complex k: thread -> row tid / TPR, k (tid % TPR) * KPT .. + KPT
    const int ar = tid / TPR, ak = (tid % TPR) * KPT;
    // B: 16 k x 64 complex: thread -> k tid/16, cols (tid&15)*4 .. +4
    const int bkr = tid >> 4, bc = (tid & 15) * 4;
    float2 ra[KPT], rb[4];
    auto load_chunk = [&](int64_t k0) {
        const int64_t gi = m0 + ar;
        const bool rowok = gi < M;
        const float2* pa = (const float2*)(A + (rowok ? gi : 0) * lda);
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            const int64_t gk = k0 + ak + q;
            ra[q] = (rowok && gk < K) ? pa[gk] : make_float2(0.f, 0.f);
        }
        const int64_t gk = k0 + bkr;
        const bool kok = gk < K;
        const float2* pb = (const float2*)(B + (kok ? gk : 0) * ldb);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t gj = n0 + bc + q;
            rb[q] = (kok && gj < N) ? pb[gj] : make_float2(0.f, 0.f);
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            Ar[buf][ak + q][ar] = ra[q].x;
            Ai[buf][ak + q][ar] = ra[q].y;
        }
        *(float4*)&Br[buf][bkr][bc] = make_float4(rb[0].x, rb[1].x, rb[2].x, rb[3].x);
        *(float4*)&Bi[buf][bkr][bc] = make_float4(rb[0].y, rb[1].y, rb[2].y, rb[3].y);
    };
    floatx16 t1[MI], t2[MI], t3[MI];
#pragma unroll
    for (int i = 0; i < MI; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) t1[i][r] = t2[i][r] = t3[i][r] = 0.f;
    const int hi = lane >> 5, jj = lane & 31;
    const int nchunks = (int)((K + G32_BK - 1) / G32_BK);
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) load_chunk((int64_t)(c + 1) * G32_BK);
#pragma unroll
        for (int kk = 0; kk < G32_BK; kk += 2) {
            const int k = kk + hi;
            const float br = Br[buf][k][wn * 32 + jj], bi = Bi[buf][k][wn * 32 + jj];
            const float bs = br + bi;
#pragma unroll
            for (int i = 0; i < MI; ++i) {
                const int row = wm * (BMT / 2) + i * 32 + jj;
                const float a_r = Ar[buf][k][row], a_i = Ai[buf][k][row];
                t1[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r, br, t1[i], 0, 0, 0);
                t2[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a_i, bi, t2[i], 0, 0, 0);
                t3[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r + a_i, bs, t3[i], 0, 0, 0);
            }
        }
        if (c + 1 < nchunks) {
            __syncthreads();
            store_chunk(buf ^ 1);
            __syncthreads();
        }
    }
    // epilogue: register r of lane -> row 8*(r/4) + 4*hi + r%4, complex column jj.
    // All C loads of a 32-row block are issued before its stores (one pointer
    // for both, so the compiler would otherwise serialise load after store).
    const int64_t gj = n0 + wn * 32 + jj;
    if (gj < N) {
#pragma unroll
        for (int i = 0; i < MI; ++i) {
            float2 cv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = m0 + wm * (BMT / 2) + i * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
                cv[r] = (beta != 0.f && gi < M) ? *(const float2*)(C + gi * ldc + gj) : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = m0 + wm * (BMT / 2) + i * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
                if (gi < M) {
                    float re = alpha * (t1[i][r] - t2[i][r]);
                    float im = alpha * (t3[i][r] - t1[i][r] - t2[i][r]);
                    if (beta != 0.f) {
                        re += beta * cv[r].x;
                        im += beta * cv[r].y;
                    }
                    *(float2*)(C + gi * ldc + gj) = make_float2(re, im);
                }
            }
        }
    }
}

template <bool XCD, int BMT = G32_BM>
__global__ __launch_bounds__(256, 3) void k_cgemm3m_f32(int64_t M, int64_t N, int64_t K, float alpha,
                                                     const cpx<float>* __restrict__ A, int64_t lda,
                                                     const cpx<float>* __restrict__ B, int64_t ldb,
                                                     float beta, cpx<float>* __restrict__ C, int64_t ldc)
{
    const int64_t nbm = (M + BMT - 1) / BMT, nbn = (N + G3M_BN - 1) / G3M_BN;
    int64_t tmi, tni;
    if (XCD) {
        tile_of(blockIdx.x, nbm, nbn, tmi, tni);
        if (tmi < 0) return;
    } else {
        tmi = blockIdx.x / nbn;
        tni = blockIdx.x % nbn;
    }
    __shared__ alignas(16) float lds[4 * G32_BK * BMT + 4 * G32_BK * G3M_BN];
    cgemm3m_tile<BMT>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, tmi * BMT, tni * G3M_BN, lds);
}

// ---- grouped products: job j owns tiles [tile_off[j], tile_off[j + 1]),
// row-major over its (row tile, column tile) grid.  With XCD the launch is
// padded to a multiple of the XCD count and block b takes tile
// (b % NUM_XCD) * (blocks / NUM_XCD) + b / NUM_XCD, so each XCD (blocks are
// dealt to XCDs round-robin) works on a contiguous run of tiles and keeps its
// jobs' operands in its own L2.
struct GemmJobDev {
    const void* a;
    const void* b;
    void* c;
    int64_t m, n, k, lda, ldb, ldc;
    double alpha, beta;
};
static_assert(sizeof(GemmJobDev) == sizeof(mba_gemm_job), "descriptor layouts must match");

__device__ __forceinline__ int grouped_tile(bool xcd, int total, int& t)
{
    const int b = (int)blockIdx.x;
    t = xcd ? (b % NUM_XCD) * ((int)gridDim.x / NUM_XCD) + b / NUM_XCD : b;
    return t < total;
}

template <bool XCD>
__global__ __launch_bounds__(256, 3) void k_cgemm3m_grouped(const GemmJobDev* __restrict__ jobs,
                                                         const int* __restrict__ tile_off, int njob)
{
    int t;
    if (!grouped_tile(XCD, tile_off[njob], t)) return;
    int lo = 0, hi = njob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tile_off[mid] <= t) lo = mid; else hi = mid;
    }
    const GemmJobDev& J = jobs[lo];
    const int64_t nbn = (J.n + G3M_BN - 1) / G3M_BN;
    const int64_t u = t - tile_off[lo];
    __shared__ alignas(16) float lds[G3M_LDS_FLOATS];
    cgemm3m_tile(J.m, J.n, J.k, (float)J.alpha, (const cpx<float>*)J.a, J.lda, (const cpx<float>*)J.b, J.ldb,
                 (float)J.beta, (cpx<float>*)J.c, J.ldc, (u / nbn) * G32_BM, (u % nbn) * G3M_BN, lds);
}

int device_cus();

// 128 x 64 tiles, or 64 x 64 when 128-row tiles would give fewer than two
// workgroups per CU (narrow LU products late in a factorisation): twice the
// workgroups, each element's MFMA sequence unchanged (bit-identical results)
template <bool XCD>
int launch_cgemm3m_f32(int64_t M, int64_t N, int64_t K, float alpha, const void* A, int64_t lda, const void* B,
                       int64_t ldb, float beta, void* C, int64_t ldc, hipStream_t st)
{
    const int64_t nbn = (N + 63) / 64;
    const int64_t tiles128 = nbn * ((M + G32_BM - 1) / G32_BM);
    const int cus = device_cus();
    const bool small = cus > 0 && tiles128 < 2 * (int64_t)cus;
    const int64_t tiles = small ? nbn * ((M + 63) / 64) : tiles128;
    const int64_t blocks = XCD ? ((tiles + NUM_XCD - 1) / NUM_XCD) * NUM_XCD : tiles;
    if (small)
        hipLaunchKernelGGL((k_cgemm3m_f32<XCD, 64>), dim3((unsigned)blocks), dim3(256), 0, st, M, N, K, alpha,
                           (const cpx<float>*)A, lda, (const cpx<float>*)B, ldb, beta, (cpx<float>*)C, ldc);
    else
        hipLaunchKernelGGL((k_cgemm3m_f32<XCD, 128>), dim3((unsigned)blocks), dim3(256), 0, st, M, N, K, alpha,
                           (const cpx<float>*)A, lda, (const cpx<float>*)B, ldb, beta, (cpx<float>*)C, ldc);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// FP32 variants: 0 = 128x64 tile, 1 = 128x64 XCD-ordered, 2 = 128x128, 3 = 128x128 XCD-ordered,
// 4 = 128x64 3M, 5 = 128x64 3M XCD-ordered
template <int WN, bool XCD>
int launch_cgemm_f32(int64_t M, int64_t N, int64_t K, float alpha, const void* A, int64_t lda,
                     const void* B, int64_t ldb, float beta, void* C, int64_t ldc, hipStream_t st)
{
    const int64_t BN = 32 * WN;
    const int64_t tiles = ((N + BN - 1) / BN) * ((M + G32_BM - 1) / G32_BM);
    const int64_t blocks = XCD ? ((tiles + NUM_XCD - 1) / NUM_XCD) * NUM_XCD : tiles;
    hipLaunchKernelGGL((k_cgemm_f32<WN, XCD>), dim3((unsigned)blocks), dim3(256), 0, st, M, N, K, alpha,
                       (const cpx<float>*)A, lda, (const cpx<float>*)B, ldb, beta, (cpx<float>*)C, ldc);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// Process-wide defaults behind the mba_*_set_* calls.  Atomic, so setting one
// while other threads run calls is not a data race; the LU reads its own
// (update, panel, solve) choices once at the start of each call.
std::atomic<int> g_gemm_f32_variant{4};  // tile choice of the FP32 GEMM (set by mba_gemm_set_variant)
// trailing update of the complex64 LU: 0 = FP32 MFMA (the variant above),
// 1 = FP32-accurate split-FP16 MFMA product (gemm_h3.hip); set by mba_lu_set_update
std::atomic<int> g_lu_update_f32{0};
// the split product pays off once the trailing block has this many rows and columns
constexpr int64_t SPLIT_MIN_DIM = 2048;
// panels of the complex64 LU: 0 = partial pivoting (LAPACK's row order), 1 =
// tournament pivoting (64-column panels of >= 2048 rows); set by mba_lu_set_panel
std::atomic<int> g_lu_panel_f32{0};

int cgemm_f32(int variant, int64_t M, int64_t N, int64_t K, float alpha, const void* A, int64_t lda,
              const void* B, int64_t ldb, float beta, void* C, int64_t ldc, hipStream_t st)
{
    switch (variant) {
    case 4: return launch_cgemm3m_f32<false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
    case 5: return launch_cgemm3m_f32<true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
    case 1: return launch_cgemm_f32<2, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
    case 2: return launch_cgemm_f32<4, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
    case 3: return launch_cgemm_f32<4, true>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
    default: return launch_cgemm_f32<2, false>(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, st);
    }
}

// FP64: v_mfma_f64_16x16x4f64, one instruction = 16 rows x 8 complex columns
// x 2 complex k.  Block tile 64 x 32 complex, 4 waves in 2 x 2, each wave
// 32 x 16 (2 x 2 instructions), k-chunks of 8 complex.
constexpr int G64_BM = 64, G64_BN = 32, G64_BK = 8;

__device__ __forceinline__ void cgemm_f64_tile(int64_t M, int64_t N, int64_t K, double alpha,
                                               const cpx<double>* __restrict__ A, int64_t lda,
                                               const cpx<double>* __restrict__ B, int64_t ldb, double beta,
                                               cpx<double>* __restrict__ C, int64_t ldc, int64_t m0, int64_t n0)
{
    __shared__ alignas(16) double Are[2][G64_BK][G64_BM];
    __shared__ alignas(16) double Aim[2][G64_BK][G64_BM];
    __shared__ alignas(16) double Bs[2][G64_BK][2 * G64_BN];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // A: 64 rows x 8 complex: thread -> row tid/4, complex k (tid&3)*2 .. +2
    const int ar = tid >> 2, ak = (tid & 3) * 2;
    // B: 8 rows x 32 complex: thread -> row tid/32, complex col tid&31
    const int bkr = tid >> 5, bc = tid & 31;
    double2 ra[2], rb;

    auto load_chunk = [&](int64_t k0) {
        const int64_t gi = m0 + ar;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t gk = k0 + ak + q;
            double2 v = make_double2(0.0, 0.0);
            if (gi < M && gk < K) v = *(const double2*)(A + gi * lda + gk);
            ra[q] = v;
        }
        const int64_t gk = k0 + bkr, gj = n0 + bc;
        rb = make_double2(0.0, 0.0);
        if (gk < K && gj < N) rb = *(const double2*)(B + gk * ldb + gj);
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            Are[buf][ak + q][ar] = ra[q].x;
            Aim[buf][ak + q][ar] = ra[q].y;
        }
        *(double2*)&Bs[buf][bkr][2 * bc] = rb;
    };

    doublex4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[i][j][r] = 0.0;

    // lane -> real k index kr = lane/16 (complex k = kr/2, part q = kr%2), column jj = lane%16
    const int kr = lane >> 4, jj = lane & 15;
    const int q = kr & 1, kc = kr >> 1;
    const int bcol = q ? (jj ^ 1) : jj;
    const double bsign = (q && !(jj & 1)) ? -1.0 : 1.0;
    const int nchunks = (int)((K + G64_BK - 1) / G64_BK);
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) load_chunk((int64_t)(c + 1) * G64_BK);
#pragma unroll
        for (int ks = 0; ks < G64_BK; ks += 2) {
            double a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const int row = wm * 32 + i * 16 + jj;
                a[i] = q ? Aim[buf][ks + kc][row] : Are[buf][ks + kc][row];
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = bsign * Bs[buf][ks + kc][wn * 32 + j * 16 + bcol];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (c + 1 < nchunks) {
            __syncthreads();
            store_chunk(buf ^ 1);
            __syncthreads();
        }
    }
    // epilogue: register r of lane -> row lane/16 + 4*r, real column lane%16
    // (v_mfma_f64_16x16x4f64 interleaves its output rows across registers).
    // C loads are all issued before the stores (one pointer for both).
    double cv[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t gjr = 2 * (n0 + wn * 16 + j * 8) + jj;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gi = m0 + wm * 32 + i * 16 + kr + 4 * r;
                cv[i][j][r] = (beta != 0.0 && gjr < 2 * N && gi < M) ? *((const double*)(C + gi * ldc) + gjr) : 0.0;
            }
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t gjr = 2 * (n0 + wn * 16 + j * 8) + jj;
            if (gjr >= 2 * N) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t gi = m0 + wm * 32 + i * 16 + kr + 4 * r;
                if (gi < M) {
                    double* p = (double*)(C + gi * ldc) + gjr;
                    double v = alpha * acc[i][j][r];
                    if (beta != 0.0) v += beta * cv[i][j][r];
                    *p = v;
                }
            }
        }
}

__global__ __launch_bounds__(256) void k_cgemm_f64(int64_t M, int64_t N, int64_t K, double alpha,
                                                   const cpx<double>* __restrict__ A, int64_t lda,
                                                   const cpx<double>* __restrict__ B, int64_t ldb,
                                                   double beta, cpx<double>* __restrict__ C,
                                                   int64_t ldc)
{
    cgemm_f64_tile(M, N, K, alpha, A, lda, B, ldb, beta, C, ldc, (int64_t)blockIdx.y * G64_BM,
                   (int64_t)blockIdx.x * G64_BN);
}

template <bool XCD>
__global__ __launch_bounds__(256) void k_cgemm_f64_grouped(const GemmJobDev* __restrict__ jobs,
                                                           const int* __restrict__ tile_off, int njob)
{
    int t;
    if (!grouped_tile(XCD, tile_off[njob], t)) return;
    int lo = 0, hi = njob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (tile_off[mid] <= t) lo = mid; else hi = mid;
    }
    const GemmJobDev& J = jobs[lo];
    const int64_t nbn = (J.n + G64_BN - 1) / G64_BN;
    const int64_t u = t - tile_off[lo];
    cgemm_f64_tile(J.m, J.n, J.k, J.alpha, (const cpx<double>*)J.a, J.lda, (const cpx<double>*)J.b, J.ldb, J.beta,
                   (cpx<double>*)J.c, J.ldc, (u / nbn) * G64_BM, (u % nbn) * G64_BN);
}

// Tiles of a grouped job and the launch of a device-resident job table.
inline int64_t grouped_tiles(int prec, int64_t m, int64_t n)
{
    if (m <= 0 || n <= 0) return 0;
    return prec == MBA_F32 ? ((m + G32_BM - 1) / G32_BM) * ((n + G3M_BN - 1) / G3M_BN)
                           : ((m + G64_BM - 1) / G64_BM) * ((n + G64_BN - 1) / G64_BN);
}

int launch_grouped(int prec, bool xcd, const GemmJobDev* d_jobs, const int* d_tile_off, int njob, int64_t tiles,
                   hipStream_t st)
{
    if (njob <= 0 || tiles <= 0) return MBA_OK;
    const int64_t blocks = xcd ? ((tiles + NUM_XCD - 1) / NUM_XCD) * NUM_XCD : tiles;
    if (blocks > 0x7fffffff) return MBA_ERR_SIZE;
    if (prec == MBA_F32) {
        if (xcd)
            hipLaunchKernelGGL(k_cgemm3m_grouped<true>, dim3((unsigned)blocks), dim3(256), 0, st, d_jobs, d_tile_off,
                               njob);
        else
            hipLaunchKernelGGL(k_cgemm3m_grouped<false>, dim3((unsigned)blocks), dim3(256), 0, st, d_jobs,
                               d_tile_off, njob);
    } else {
        if (xcd)
            hipLaunchKernelGGL(k_cgemm_f64_grouped<true>, dim3((unsigned)blocks), dim3(256), 0, st, d_jobs,
                               d_tile_off, njob);
        else
            hipLaunchKernelGGL(k_cgemm_f64_grouped<false>, dim3((unsigned)blocks), dim3(256), 0, st, d_jobs,
                               d_tile_off, njob);
    }
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

template <class T>
int gemm(int64_t M, int64_t N, int64_t K, double alpha, const void* A, int64_t lda, const void* B,
         int64_t ldb, double beta, void* C, int64_t ldc, hipStream_t st);

template <>
int gemm<float>(int64_t M, int64_t N, int64_t K, double alpha, const void* A, int64_t lda,
                const void* B, int64_t ldb, double beta, void* C, int64_t ldc, hipStream_t st)
{
    if (M <= 0 || N <= 0) return MBA_OK;
    return cgemm_f32(g_gemm_f32_variant.load(std::memory_order_relaxed), M, N, K, (float)alpha, A, lda, B, ldb, (float)beta, C, ldc, st);
}

template <>
int gemm<double>(int64_t M, int64_t N, int64_t K, double alpha, const void* A, int64_t lda,
                 const void* B, int64_t ldb, double beta, void* C, int64_t ldc, hipStream_t st)
{
    if (M <= 0 || N <= 0) return MBA_OK;
    dim3 grid((unsigned)((N + G64_BN - 1) / G64_BN), (unsigned)((M + G64_BM - 1) / G64_BM));
    hipLaunchKernelGGL(k_cgemm_f64, grid, dim3(256), 0, st, M, N, K, alpha, (const cpx<double>*)A,
                       lda, (const cpx<double>*)B, ldb, beta, (cpx<double>*)C, ldc);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// ============================================================== LU panel
// Cooperative kernel: G workgroups (one per CU, all resident) each hold R
// consecutive rows of the m x nb panel in LDS.  Per column j:
//   local arg-max of |Re| + |Im| (ties: lowest row) and the candidate row are
//   published; the owner of row j publishes row j; every workgroup gathers the
//   G candidates, picks the same winner, swaps locally, scales, and applies
//   the rank-1 update.
// Messages are self-validating 64-bit words (epoch << 32 | payload) moved with
// relaxed agent-scope atomics: a reader waits until every word it needs carries
// the current step's epoch, so no cache write-back / invalidate fences are
// needed and no shared counter is contended.  Epochs k + j + 1 are unique per
// step.  Buffers are double-buffered on j parity: a writer reuses a buffer two
// steps later, after every workgroup has published the step in between, which
// each does only after finishing its reads of the first.  Waits are bounded;
// on expiry the kernel sets *err, stops waiting and runs to completion (its
// result is garbage and the host reports a timeout).
constexpr int PANEL_THREADS = 512;
constexpr long SPIN_LIMIT = 1L << 22;

template <class T>
struct PanelArgs {
    cpx<T>* A;
    int64_t lda;
    int64_t k;       // first row / column of the panel
    int64_t m;       // rows in the panel (n - k)
    int nb;          // columns in the panel
    int R;           // rows per workgroup
    int32_t* ipiv;   // global [n]
    int32_t* info;
    int* err;
    uint64_t* meta;  // [2][G] (epoch, candidate row)
    uint64_t* cval;  // [2][G][VW] (epoch, candidate |.|_1 bits)
    uint64_t* cand;  // [2][G][nb][2 VW] candidate rows
    uint64_t* diag;  // [2][nb][2 VW] row j
};

template <class T>
__device__ __forceinline__ T cabs1(cpx<T> v) { return fabs(v.re) + fabs(v.im); }

__device__ __forceinline__ void tag_put(uint64_t* p, uint32_t epoch, uint32_t payload)
{
    __hip_atomic_store(p, ((uint64_t)epoch << 32) | payload, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ uint32_t tag_get(uint64_t* p, uint32_t epoch, int* err)
{
    long spins = 0;
    while (true) {
        uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if ((uint32_t)(w >> 32) == epoch) return (uint32_t)w;
        // the error flag costs a memory round trip: look at it every 64 polls
        if ((spins & 63) == 63 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return (uint32_t)w;
        if (++spins > SPIN_LIMIT) {
            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return (uint32_t)w;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// T <-> 32-bit words (VW words per value)
__device__ __forceinline__ void put_val(uint64_t* p, uint32_t e, float v) { tag_put(p, e, __float_as_uint(v)); }
__device__ __forceinline__ void put_val(uint64_t* p, uint32_t e, double v)
{
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    tag_put(p, e, (uint32_t)(b >> 32));
    tag_put(p + 1, e, (uint32_t)b);
}
__device__ __forceinline__ void get_val(uint64_t* p, uint32_t e, int* err, float& v)
{
    v = __uint_as_float(tag_get(p, e, err));
}
__device__ __forceinline__ void get_val(uint64_t* p, uint32_t e, int* err, double& v)
{
    unsigned long long hi = tag_get(p, e, err), lo = tag_get(p + 1, e, err);
    v = __longlong_as_double((long long)((hi << 32) | lo));
}

// ---- arg-max of (v desc, r asc) carrying q
// Within a wave: DPP lane moves (quad permutes lane ^ 1 and lane ^ 2, then the
// half-row and row mirrors) leave every lane with the best of its 16-lane
// row; the four rows' results are combined from lane reads.  Every lane of
// the wave gets the result.  Call with all 64 lanes active.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float x) { return __int_as_float(dpp_i<CTRL>(__float_as_int(x))); }
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double x)
{
    const long long b = __double_as_longlong(x);
    const int lo = dpp_i<CTRL>((int)(b & 0xffffffffLL)), hi = dpp_i<CTRL>((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <int CTRL>
__device__ __forceinline__ int32_t dpp_mov(int32_t x) { return dpp_i<CTRL>(x); }

__device__ __forceinline__ int32_t lane_read(int32_t x, int l) { return __builtin_amdgcn_readlane(x, l); }
__device__ __forceinline__ float lane_read(float x, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}
__device__ __forceinline__ double lane_read(double x, int l)
{
    const long long b = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <class T>
__device__ __forceinline__ void better(T& v, int32_t& r, int32_t& q, T v2, int32_t r2, int32_t q2)
{
    if (v2 > v || (v2 == v && r2 < r)) {
        v = v2;
        r = r2;
        q = q2;
    }
}

template <int CTRL, class T>
__device__ __forceinline__ void dpp_step(T& v, int32_t& r, int32_t& q)
{
    better(v, r, q, dpp_mov<CTRL>(v), dpp_mov<CTRL>(r), dpp_mov<CTRL>(q));
}

template <class T>
__device__ __forceinline__ void wave_argmax(T& v, int32_t& r, int32_t& q)
{
    dpp_step<0xB1>(v, r, q);   // quad_perm [1, 0, 3, 2]
    dpp_step<0x4E>(v, r, q);   // quad_perm [2, 3, 0, 1]
    dpp_step<0x141>(v, r, q);  // row_half_mirror
    dpp_step<0x140>(v, r, q);  // row_mirror
    T bv = lane_read(v, 0);
    int32_t br = lane_read(r, 0), bq = lane_read(q, 0);
#pragma unroll
    for (int l = 16; l < 64; l += 16) better(bv, br, bq, lane_read(v, l), lane_read(r, l), lane_read(q, l));
    v = bv;
    r = br;
    q = bq;
}

// Block-wide: the waves' results meet in LDS after one barrier and every
// thread combines them itself.  red_* hold 2 x NW entries; two calls not
// separated by another barrier must use different slots.
template <int NW, class T>
__device__ __forceinline__ void block_argmax(T& v, int32_t& r, int32_t& q, T* red_v, int32_t* red_r,
                                             int32_t* red_q, int slot)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    wave_argmax(v, r, q);
    T* rv = red_v + slot * NW;
    int32_t* rr = red_r + slot * NW;
    int32_t* rq = red_q + slot * NW;
    if (lane == 0) {
        rv[wave] = v;
        rr[wave] = r;
        rq[wave] = q;
    }
    __syncthreads();
    T pv[NW];
    int32_t pr[NW], pq[NW];
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        pv[w] = rv[w];
        pr[w] = rr[w];
        pq[w] = rq[w];
    }
    v = pv[0];
    r = pr[0];
    q = pq[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) better(v, r, q, pv[w], pr[w], pq[w]);
}

// The same arg-max for FP32 values in one 64-bit key per lane: the value's
// bits (a non-negative float ranks as its bit pattern) over the complement
// of the row (the smaller row wins a tie); key 0 is "no candidate" (the empty
// value -1, and NaN, which never wins).  One DPP chain on two words instead
// of three values, one LDS word per wave.
__device__ __forceinline__ uint64_t am_key(float v, int32_t r)
{
    if (!(v >= 0.f)) return 0ull;
    return ((uint64_t)__float_as_uint(v) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)r);
}
__device__ __forceinline__ void am_unkey(uint64_t k, float& v, int32_t& r, int32_t none)
{
    if (k == 0ull) {
        v = -1.f;
        r = none;
    } else {
        v = __uint_as_float((uint32_t)(k >> 32));
        r = (int32_t)(0xFFFFFFFFu - (uint32_t)k);
    }
}
template <int CTRL>
__device__ __forceinline__ void key_step(uint64_t& k)
{
    const uint32_t lo = (uint32_t)dpp_i<CTRL>((int)(uint32_t)k), hi = (uint32_t)dpp_i<CTRL>((int)(uint32_t)(k >> 32));
    const uint64_t o = ((uint64_t)hi << 32) | lo;
    k = o > k ? o : k;
}
__device__ __forceinline__ uint64_t wave_max_key(uint64_t k)
{
    key_step<0xB1>(k);   // quad_perm [1, 0, 3, 2]
    key_step<0x4E>(k);   // quad_perm [2, 3, 0, 1]
    key_step<0x141>(k);  // row_half_mirror
    key_step<0x140>(k);  // row_mirror
    const uint32_t lo = (uint32_t)k, hi = (uint32_t)(k >> 32);
    uint64_t best = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, 0) << 32) |
                    (uint32_t)__builtin_amdgcn_readlane((int)lo, 0);
#pragma unroll
    for (int l = 16; l < 64; l += 16) {
        const uint64_t o = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)hi, l) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)lo, l);
        best = o > best ? o : best;
    }
    return best;
}
// block-wide key maximum: one barrier; red holds 2 x NW words, two calls not
// separated by another barrier must use different slots
template <int NW>
__device__ __forceinline__ uint64_t block_max_key(uint64_t k, uint64_t* red, int slot)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    k = wave_max_key(k);
    if (lane == 0) red[slot * NW + wave] = k;
    __syncthreads();
    uint64_t b = red[slot * NW];
#pragma unroll
    for (int w = 1; w < NW; ++w) {
        const uint64_t o = red[slot * NW + w];
        b = o > b ? o : b;
    }
    return b;
}

// Waits until all N tagged words carry `epoch`; the N loads are issued
// together, so a wait costs one memory round trip, not N.  Only words still
// stale are re-read.  Bounded like tag_get.
template <int N>
__device__ __forceinline__ void tag_get_n(uint64_t* const (&p)[N], uint32_t epoch, int* err, uint32_t (&out)[N])
{
    uint64_t w[N];
#pragma unroll
    for (int i = 0; i < N; ++i) w[i] = __hip_atomic_load(p[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    long spins = 0;
    while (true) {
        bool ok = true;
#pragma unroll
        for (int i = 0; i < N; ++i) ok = ok && (uint32_t)(w[i] >> 32) == epoch;
        if (ok) break;
        if ((spins & 63) == 63 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
        if (++spins > SPIN_LIMIT) {
            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int i = 0; i < N; ++i)
            if ((uint32_t)(w[i] >> 32) != epoch)
                w[i] = __hip_atomic_load(p[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = (uint32_t)w[i];
}

// Waits until the N tagged words carry `epoch` (null pointers read as 0);
// returns false when it gave up (timeout or *err set elsewhere).
template <int N>
__device__ __forceinline__ bool tag_get_opt(uint64_t* const (&p)[N], uint32_t epoch, int* err, uint32_t (&out)[N])
{
    uint64_t w[N];
#pragma unroll
    for (int i = 0; i < N; ++i)
        w[i] = p[i] ? __hip_atomic_load(p[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ((uint64_t)epoch << 32);
    long spins = 0;
    bool good = true;
    while (true) {
        bool ok = true;
#pragma unroll
        for (int i = 0; i < N; ++i) ok = ok && (uint32_t)(w[i] >> 32) == epoch;
        if (ok) break;
        if ((spins & 63) == 63 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            good = false;
            break;
        }
        if (++spins > SPIN_LIMIT) {
            __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            good = false;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
#pragma unroll
        for (int i = 0; i < N; ++i)
            if ((uint32_t)(w[i] >> 32) != epoch)
                w[i] = __hip_atomic_load(p[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
#pragma unroll
    for (int i = 0; i < N; ++i) out[i] = (uint32_t)w[i];
    return good;
}

template <class T>
__device__ __forceinline__ T val_of(const uint32_t* wds);
template <>
__device__ __forceinline__ float val_of<float>(const uint32_t* wds) { return __uint_as_float(wds[0]); }
template <>
__device__ __forceinline__ double val_of<double>(const uint32_t* wds)
{
    return __longlong_as_double((long long)(((unsigned long long)wds[0] << 32) | wds[1]));
}

// Publishes local row r (panel-local row index pr) of the LDS panel, all nb
// columns, into the message row `dst` (thread c writes column c).
template <class T, int NT = PANEL_THREADS>
__device__ __forceinline__ void put_row(uint64_t* dst, uint32_t epoch, const cpx<T>* row, int nb)
{
    constexpr int VW = sizeof(T) / 4;
    for (int c = threadIdx.x; c < nb; c += NT) {
        const cpx<T> x = row ? row[c] : cpx<T>{(T)0, (T)0};
        put_val(dst + (2 * c) * VW, epoch, x.re);
        put_val(dst + (2 * c + 1) * VW, epoch, x.im);
    }
}

// Panel body for workgroup g of G (the workgroups owning this panel's rows).
//
// Per column j the critical path is two memory round trips: (1) the G
// candidates (row, |.|_1) of column j, (2) the winner's row and row j, both
// published in full.  The candidates of column j + 1 are formed and
// published right after column j is scaled and column j + 1 updated, and the
// two rows that may be read next (the local candidate and row j + 1) are
// completed and published before the bulk rank-1 update of the other rows,
// so that update overlaps the exchange.  Buffers alternate on j parity: data
// of column j + 1 is written (during column j) only after this workgroup has
// read every workgroup's column-j candidate, and each workgroup publishes its
// column-j candidate only after finishing its column j - 1 reads, so the
// column j - 1 buffers are no longer read when they are overwritten.
template <class T, int NT = PANEL_THREADS>
__device__ __forceinline__ void panel_body(const PanelArgs<T>& P, const int G, const int g)
{
    constexpr int VW = sizeof(T) / 4;  // 32-bit words per real
    constexpr int32_t NONE = 0x7fffffff;
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* rows = (cpx<T>*)smem_raw;  // [R][nb]
    __shared__ T red_v[2 * NT / 64];
    __shared__ int32_t red_r[2 * NT / 64];
    __shared__ int32_t red_q[2 * NT / 64];
    __shared__ uint64_t red_k[2 * NT / 64];  // FP32: packed arg-max keys
    constexpr bool KEYED = sizeof(T) == 4;
    __shared__ cpx<T> s_prow[128];
    __shared__ cpx<T> s_drow[128];

    const int tid = threadIdx.x;
    const int nb = P.nb, R = P.R;
    const int64_t lr0 = (int64_t)g * R;  // first local row of this workgroup
    const int nr = (int)((P.m - lr0) < R ? (P.m - lr0 > 0 ? P.m - lr0 : 0) : R);
    const int64_t row_words = (int64_t)nb * 2 * VW;
    auto owns = [&](int64_t pr) { return pr >= lr0 && pr < lr0 + nr; };

    for (int idx = tid; idx < nr * nb; idx += NT) {
        int r = idx / nb, c = idx % nb;
        rows[idx] = P.A[(P.k + lr0 + r) * P.lda + P.k + c];
    }
    __syncthreads();

    // column 0: local arg-max, then publish candidate, candidate row, row 0
    {
        T bv = (T)-1;
        int32_t br = NONE;
        for (int r = tid; r < nr; r += NT) {
            const T v = cabs1(rows[r * nb]);
            if (v > bv) {
                bv = v;
                br = (int32_t)(lr0 + r);
            }
        }
        if constexpr (KEYED) {
            am_unkey(block_max_key<NT / 64>(am_key((float)bv, br), red_k, 0), *(float*)&bv, br, NONE);
        } else {
            int32_t bq = 0;
            block_argmax<NT / 64>(bv, br, bq, red_v, red_r, red_q, 0);
        }
        const int32_t lb = br;
        const uint32_t e0 = (uint32_t)(P.k + 1);
        if (tid == 0) {
            tag_put(P.meta + g, e0, (uint32_t)lb);
            put_val(P.cval + (int64_t)g * VW, e0, bv);
        }
        put_row<T, NT>(P.cand + (int64_t)g * row_words, e0, lb != NONE ? rows + (lb - lr0) * nb : nullptr, nb);
        if (owns(0)) put_row<T, NT>(P.diag, e0, rows + (0 - lr0) * nb, nb);
    }

    for (int j = 0; j < nb; ++j) {
        const int par = j & 1;
        const uint32_t epoch = (uint32_t)(P.k + j + 1);
        // --- (1) gather the G candidates of column j (one slot per thread)
        T v = (T)-1;
        int32_t r = NONE, q = 0;
        for (int c = tid; c < G; c += NT) {
            uint64_t* pp[1 + VW];
            pp[0] = P.meta + par * G + c;
#pragma unroll
            for (int w = 0; w < VW; ++w) pp[1 + w] = P.cval + ((int64_t)par * G + c) * VW + w;
            uint32_t wd[1 + VW];
            tag_get_n<1 + VW>(pp, epoch, P.err, wd);
            better(v, r, q, val_of<T>(wd + 1), (int32_t)wd[0], (int32_t)c);
        }
        if constexpr (KEYED) {
            // the winner's workgroup is the owner of its row: rows [g R, g R + R)
            am_unkey(block_max_key<NT / 64>(am_key((float)v, r), red_k, 1), *(float*)&v, r, NONE);
            q = r == NONE ? 0 : r / R;
        } else {
            block_argmax<NT / 64>(v, r, q, red_v, red_r, red_q, 1);
        }
        const int wg = q;
        // no candidate at all (every value NaN): no interchange, the column
        // is reported singular below, and no row index leaves the panel
        const int32_t p = r == NONE ? (int32_t)j : r;
        // --- (2) the winner's row and row j
        {
            uint64_t* pw = P.cand + ((int64_t)par * G + wg) * row_words;
            uint64_t* dw = P.diag + (int64_t)par * row_words;
            for (int c = tid; c < nb; c += NT) {
                uint64_t* pp[4 * VW];
#pragma unroll
                for (int w = 0; w < 2 * VW; ++w) {
                    pp[w] = pw + 2 * c * VW + w;
                    pp[2 * VW + w] = dw + 2 * c * VW + w;
                }
                uint32_t wd[4 * VW];
                tag_get_n<4 * VW>(pp, epoch, P.err, wd);
                s_prow[c] = cpx<T>{val_of<T>(wd), val_of<T>(wd + VW)};
                s_drow[c] = cpx<T>{val_of<T>(wd + 2 * VW), val_of<T>(wd + 3 * VW)};
            }
        }
        __syncthreads();
        // every thread forms the pivot's inverse itself (same value everywhere)
        cpx<T> inv;
        {
            const cpx<T> pv = s_prow[j];
            const T d = pv.re * pv.re + pv.im * pv.im;
            if (d == (T)0) {
                inv.re = 0;
                inv.im = 0;
            } else {
                inv.re = pv.re / d;
                inv.im = -pv.im / d;
            }
            if (tid == 0 && g == 0) {
                if (d == (T)0 || r == NONE) atomicCAS(P.info, 0, (int32_t)(P.k + j + 1));
                P.ipiv[P.k + j] = (int32_t)(P.k + p);
            }
        }
        const int j1 = j + 1;
        // --- (3) scale column j below the diagonal; update column j + 1 and
        // take its local arg-max.  The interchange of rows j and p is folded
        // in: place p (if here, p > j) takes old row j, whose columns j and
        // j + 1 come from s_drow; the interchange writes the other columns
        // (disjoint from what (3) writes), so no barrier separates them.
        T bv = (T)-1;
        int32_t br = NONE;
        const cpx<T> u1 = j1 < nb ? s_prow[j1] : cpx<T>{(T)0, (T)0};
        for (int rr = tid; rr < nr; rr += NT) {
            const int64_t pr = lr0 + rr;
            if (pr > j) {
                const bool moved = pr == p;
                const cpx<T> a = moved ? s_drow[j] : rows[rr * nb + j];
                cpx<T> l;
                l.re = a.re * inv.re - a.im * inv.im;
                l.im = a.re * inv.im + a.im * inv.re;
                rows[rr * nb + j] = l;
                if (j1 < nb) {
                    cpx<T> b = moved ? s_drow[j1] : rows[rr * nb + j1];
                    b.re -= l.re * u1.re - l.im * u1.im;
                    b.im -= l.re * u1.im + l.im * u1.re;
                    rows[rr * nb + j1] = b;
                    const T av = cabs1(b);
                    if (av > bv) {
                        bv = av;
                        br = (int32_t)pr;
                    }
                }
            }
        }
        if (p != j) {
            if (owns(p))
                for (int c = tid; c < nb; c += NT)
                    if (c != j && c != j1) rows[(p - lr0) * nb + c] = s_drow[c];
            if (owns(j))
                for (int c = tid; c < nb; c += NT) rows[(j - lr0) * nb + c] = s_prow[c];
        }
        if (j1 >= nb) break;
        if constexpr (KEYED) {  // its barrier orders (3) before (4)
            am_unkey(block_max_key<NT / 64>(am_key((float)bv, br), red_k, 0), *(float*)&bv, br, NONE);
        } else {
            int32_t bq = 0;
            block_argmax<NT / 64>(bv, br, bq, red_v, red_r, red_q, 0);
        }
        const int32_t lb = br;
        const uint32_t e1 = epoch + 1;
        const int par1 = j1 & 1;
        if (tid == 0) {
            tag_put(P.meta + par1 * G + g, e1, (uint32_t)lb);
            put_val(P.cval + ((int64_t)par1 * G + g) * VW, e1, bv);
        }
        // --- (4) complete and publish the two rows column j + 1 may read:
        // the local candidate lb and row j + 1 (thread c owns column c)
        const bool own1 = owns(j1);
        for (int c = tid; c < nb; c += NT) {
            const cpx<T> u = s_prow[c];
            auto fin = [&](int64_t pr) {
                cpx<T>* row = rows + (pr - lr0) * nb;
                if (c > j1) {
                    const cpx<T> l = row[j];
                    cpx<T> a = row[c];
                    a.re -= l.re * u.re - l.im * u.im;
                    a.im -= l.re * u.im + l.im * u.re;
                    row[c] = a;
                }
                return row[c];
            };
            cpx<T> xc = cpx<T>{(T)0, (T)0};
            if (lb != NONE) xc = fin(lb);
            uint64_t* cw = P.cand + ((int64_t)par1 * G + g) * row_words;
            put_val(cw + (2 * c) * VW, e1, xc.re);
            put_val(cw + (2 * c + 1) * VW, e1, xc.im);
            if (own1) {
                const cpx<T> xd = (lb == j1) ? xc : fin(j1);
                uint64_t* dw = P.diag + (int64_t)par1 * row_words;
                put_val(dw + (2 * c) * VW, e1, xd.re);
                put_val(dw + (2 * c + 1) * VW, e1, xd.im);
            }
        }
        // --- (5) rank-1 update of columns > j + 1 for the other rows: the
        // (row, column) pairs are dealt to the threads in row-major order, so
        // every thread has work whatever the number of columns left; UB pairs
        // per step, all 3 UB LDS reads of a step issued before the first is
        // waited for (clamped indices, no branches)
        {
            constexpr int UB = 8;
            const int c0 = j1 + 1, ncol = nb - c0;
            const int total = nr * ncol;
            const float rinv = ncol > 0 ? 1.f / (float)ncol : 0.f;
            for (int e0 = tid; e0 < total; e0 += NT * UB) {
                int rq[UB], cq[UB];
                cpx<T> l[UB], a[UB], u[UB];
#pragma unroll
                for (int q = 0; q < UB; ++q) {
                    const int e = e0 + q * NT;
                    const int ec = e < total ? e : 0;
                    int rr = (int)((float)ec * rinv);  // row = ec / ncol, corrected below
                    int cc = ec - rr * ncol;
                    if (cc < 0) {
                        --rr;
                        cc += ncol;
                    } else if (cc >= ncol) {
                        ++rr;
                        cc -= ncol;
                    }
                    rq[q] = rr;
                    cq[q] = c0 + cc;
                    l[q] = rows[rr * nb + j];
                    a[q] = rows[rr * nb + c0 + cc];
                    u[q] = s_prow[c0 + cc];
                }
#pragma unroll
                for (int q = 0; q < UB; ++q) {
                    const int64_t pr = lr0 + rq[q];
                    if (e0 + q * NT < total && pr > j && pr != lb && pr != j1) {
                        a[q].re -= l[q].re * u[q].re - l[q].im * u[q].im;
                        a[q].im -= l[q].re * u[q].im + l[q].im * u[q].re;
                        rows[rq[q] * nb + cq[q]] = a[q];
                    }
                }
            }
        }
        // no barrier here: the next column's gather only reads global words,
        // and the barrier of its arg-max comes before anything that touches
        // rows or s_prow / s_drow, so this update overlaps that round trip
    }
    __syncthreads();
    for (int idx = tid; idx < nr * nb; idx += NT) {
        int r = idx / nb, c = idx % nb;
        P.A[(P.k + lr0 + r) * P.lda + P.k + c] = rows[idx];
    }
}
template <class T>
__global__ __launch_bounds__(PANEL_THREADS) void k_panel(PanelArgs<T> P)
{
    panel_body<T>(P, gridDim.x, blockIdx.x);
}

// The panels of several matrices in one launch: matrix i owns the
// workgroups [wg_off[i], wg_off[i+1]) and its own message slots, so the
// per-column exchange latency is paid once for all of them.
template <class T>
__global__ __launch_bounds__(PANEL_THREADS) void k_panel_batched(const PanelArgs<T>* __restrict__ Ps,
                                                                 const int* __restrict__ wg_off, int nmat)
{
    int lo = 0, hi = nmat;  // wg_off[lo] <= blockIdx.x < wg_off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (wg_off[mid] <= (int)blockIdx.x) lo = mid; else hi = mid;
    }
    const PanelArgs<T> P = Ps[lo];
    panel_body<T>(P, wg_off[lo + 1] - wg_off[lo], (int)blockIdx.x - wg_off[lo]);
}

// ============================================================== tournament-pivoting panel
// complex64 panels of 64 columns and m >= 2048 rows (workgroup g owns local
// rows [g R, g R + R), 64 <= R <= 128):
//   1. each workgroup picks <= 64 candidate rows of its own by partial
//      pivoting on its rows (all rows when it has <= 64 and G > 1);
//   2. binary tree: at level l workgroup g (g % 2^(l+1) == 0) stacks its
//      candidates with those of g + 2^l (read as row ids from a tagged
//      message, the rows themselves from A) and keeps the 64 rows that
//      partial pivoting on the stack picks; the other workgroup publishes its
//      candidates and leaves the tree;
//   3. the root's last elimination is the LU of the 64 chosen rows in pivot
//      order (L11 \ U11); it writes ipiv as the equivalent sequence of
//      interchanges and publishes the pivot rows and L11 \ U11;
//   4. every workgroup solves L21 U11 = A21 for its non-pivot rows and writes
//      them to their final places; workgroup 0 also writes L11 \ U11 to rows
//      k .. k + 63 and moves the displaced rows of that range.
// Rows are read from A (never written before step 4, which starts only after
// the whole tree has read), so messages carry row ids only.  A row leaving
// rows k .. k + 63 goes to the former place of a pivot row, whose owner never
// uses that place's old contents.  Tagged words and bounded waits as in the
// partial-pivoting panel; epochs: (k / 64 + 1) * 16 + level, + 15 for the
// root message.
constexpr int TP_NB = 64;
constexpr int TP_RMAX = 128;
constexpr int TP_PITCH = TP_NB + 2;  // LDS row pitch (complex): lanes (r, h) hit distinct banks
constexpr int TP_THREADS = 256;
constexpr int64_t TP_MIN_M = 2048;
constexpr int TP_CAND_W = TP_NB + 1;                    // words per candidate set: count, ids
constexpr int TP_ROOT_W = TP_NB + TP_NB * TP_NB * 2 + 1;  // pivot ids, L11 \ U11 (re, im words), ready

struct TpArgs {
    cpx<float>* A;
    int64_t lda;
    int64_t k;      // first row / column of the panel
    int64_t m;      // rows (n - k)
    int R;          // rows per workgroup
    int32_t* ipiv;  // global [n]
    int32_t* info;
    int* err;
    uint64_t* cand;  // [G][TP_CAND_W]
    uint64_t* root;  // [TP_ROOT_W]
    uint32_t epoch0;
};

// One thread of a workgroup waits for one tagged word, polling every ~0.4 us:
// long waits (the tree, the root message) then cost one load per workgroup
// per poll instead of one per waiting word and thread, which would flood the
// memory system while the tree works.  The words of the message are read
// afterwards with the usual bounded waits (they are then nearly always there).
__device__ __forceinline__ void tp_wait_word(uint64_t* p, uint32_t epoch, int* err)
{
    if (threadIdx.x == 0) {
        long spins = 0;
        while ((uint32_t)(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) != epoch) {
            if ((spins & 15) == 15 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
            if (++spins > SPIN_LIMIT) {
                __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(16);
        }
    }
    __syncthreads();
}

// Partial pivoting on F[nrows][TP_NB] (row ids ids[], panel-local): for
// j < min(nrows, 64) the unselected row of largest |.|_1 in column j (ties:
// smaller id) becomes pivot j (sel[j] = its F row); the other unselected rows
// get their multiplier in column j and their columns > j updated.  Returns
// the first column with an exactly zero pivot, or -1.
__device__ int tp_gepp(cpx<float>* F, const int32_t* ids, int nrows, int32_t* sel, int32_t* flag, float* red_v,
                       int32_t* red_r, int32_t* red_q)
{
    const int tid = threadIdx.x;
    const int npiv = nrows < TP_NB ? nrows : TP_NB;
    int zero_col = -1;
    for (int r = tid; r < nrows; r += TP_THREADS) flag[r] = 0;
    __syncthreads();
    for (int j = 0; j < npiv; ++j) {
        float v = -1.f;
        int32_t id = 0x7fffffff, q = -1;
        for (int r = tid; r < nrows; r += TP_THREADS)
            if (!flag[r]) {
                const float a = cabs1(F[r * TP_PITCH + j]);
                if (a > v || (a == v && ids[r] < id)) {
                    v = a;
                    id = ids[r];
                    q = r;
                }
            }
        block_argmax<TP_THREADS / 64>(v, id, q, red_v, red_r, red_q, 0);
        const int s = q;
        const cpx<float> piv = F[s * TP_PITCH + j];
        const float d = piv.re * piv.re + piv.im * piv.im;
        cpx<float> inv = {0.f, 0.f};
        if (d == 0.f) {
            if (zero_col < 0) zero_col = j;
        } else {
            inv.re = piv.re / d;
            inv.im = -piv.im / d;
        }
        if (tid == 0) {
            sel[j] = s;
            flag[s] = 1;
        }
        // thread pair per row, columns of one parity each; the fixed-trip
        // loop (columns <= j masked) lets every LDS read issue up front
        const int r = tid >> 1, h = tid & 1;
        if (r < nrows && r != s && !flag[r]) {
            cpx<float>* Fr = F + r * TP_PITCH;
            const cpx<float>* Fs = F + s * TP_PITCH;
            const cpx<float> a = Fr[j];
            cpx<float> l;
            l.re = a.re * inv.re - a.im * inv.im;
            l.im = a.re * inv.im + a.im * inv.re;
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                cpx<float> u[TP_NB / 4], x[TP_NB / 4];
#pragma unroll
                for (int i = 0; i < TP_NB / 4; ++i) {
                    const int c = 2 * (i + half * TP_NB / 4) + h;
                    if (c > j) {
                        u[i] = Fs[c];
                        x[i] = Fr[c];
                    }
                }
#pragma unroll
                for (int i = 0; i < TP_NB / 4; ++i) {
                    const int c = 2 * (i + half * TP_NB / 4) + h;
                    if (c > j) {
                        x[i].re -= l.re * u[i].re - l.im * u[i].im;
                        x[i].im -= l.re * u[i].im + l.im * u[i].re;
                        Fr[c] = x[i];
                    }
                }
            }
            if (h == 0) Fr[j] = l;  // both lanes of the row (one wave) have read Fr[j]
        }
        __syncthreads();
    }
    return zero_col;
}

// panel rows (original values) -> F rows [0, n): row r is ids[r], or row0 + r
// when ids is null.  All loads of a thread are issued before its LDS stores,
// so the copy costs one memory latency.
__device__ __forceinline__ void tp_load_rows(const TpArgs& P, const int32_t* ids, int64_t row0, int n, cpx<float>* F)
{
    constexpr int PER = TP_RMAX * TP_NB / TP_THREADS;
    cpx<float> v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int idx = threadIdx.x + TP_THREADS * i;
        const int r = idx / TP_NB, c = idx % TP_NB;
        if (r < n) {
            const int64_t row = ids ? (int64_t)ids[r] : row0 + r;
            v[i] = P.A[(P.k + row) * P.lda + P.k + c];
        }
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int idx = threadIdx.x + TP_THREADS * i;
        const int r = idx / TP_NB, c = idx % TP_NB;
        if (r < n) F[r * TP_PITCH + c] = v[i];
    }
}

__global__ __launch_bounds__(TP_THREADS) void k_panel_tp(TpArgs P)
{
    extern __shared__ unsigned char smem_raw[];
    cpx<float>* F = (cpx<float>*)smem_raw;       // [TP_RMAX][TP_PITCH]
    cpx<float>* U = F + TP_RMAX * TP_PITCH;      // [TP_NB][TP_PITCH]: L11 \ U11 rows in pivot order
    __shared__ int32_t ids[TP_RMAX], sel[TP_NB], flag[TP_RMAX], cids[TP_NB], pid[TP_NB], piv_of[TP_RMAX];
    __shared__ int32_t top[TP_NB], wt[TP_NB];
    __shared__ float red_v[2 * TP_THREADS / 64];
    __shared__ int32_t red_r[2 * TP_THREADS / 64], red_q[2 * TP_THREADS / 64];
    __shared__ int s_cnt;
    __shared__ cpx<float> s_invd[TP_NB];

    const int tid = threadIdx.x;
    const int G = gridDim.x, g = blockIdx.x;
    const int64_t lr0 = (int64_t)g * P.R;
    const int nr = (int)((P.m - lr0) < P.R ? (P.m - lr0) : P.R);  // >= 1: G = ceil-free split, see host

    // ---- 1. local candidates
    if (G == 1 || nr > TP_NB) {
        for (int r = tid; r < nr; r += TP_THREADS) ids[r] = (int32_t)(lr0 + r);
        tp_load_rows(P, nullptr, lr0, nr, F);
        __syncthreads();
        tp_gepp(F, ids, nr, sel, flag, red_v, red_r, red_q);
        const int cnt = nr < TP_NB ? nr : TP_NB;
        if (tid < cnt) cids[tid] = ids[sel[tid]];
        if (tid == 0) s_cnt = cnt;
    } else {
        if (tid < nr) cids[tid] = (int32_t)(lr0 + tid);
        if (tid == 0) s_cnt = nr;
    }
    __syncthreads();

    // ---- 2. tree
    for (int level = 0; (1 << level) < G; ++level) {
        const uint32_t e = P.epoch0 + (uint32_t)level;
        if ((g >> level) & 1) {  // publish and leave
            uint64_t* w = P.cand + (int64_t)g * TP_CAND_W;
            const int cnt = s_cnt;
            if (tid < TP_NB) tag_put(w + 1 + tid, e, tid < cnt ? (uint32_t)cids[tid] : 0u);
            __syncthreads();
            if (tid == 0) tag_put(w, e, (uint32_t)cnt);
            break;
        }
        const int partner = g + (1 << level);
        if (partner >= G) continue;
        // partner's count and ids: all 65 words in one round trip (bounded
        // waits; after a timeout the words are clamped into range)
        uint64_t* w = P.cand + (int64_t)partner * TP_CAND_W;
        __shared__ int32_t s_pw[TP_CAND_W];
        tp_wait_word(w, e, P.err);  // the count, written after the ids
        if (tid < TP_CAND_W) s_pw[tid] = (int32_t)tag_get(w + tid, e, P.err);
        __syncthreads();
        const int pc = s_pw[0] < 0 ? 0 : (s_pw[0] < TP_NB ? s_pw[0] : TP_NB);
        const int own = s_cnt;
        if (tid < own) ids[tid] = cids[tid];
        if (tid < pc) {
            const int32_t v = s_pw[1 + tid];
            ids[own + tid] = (v >= 0 && v < P.m) ? v : 0;
        }
        __syncthreads();
        tp_load_rows(P, ids, 0, own + pc, F);
        __syncthreads();
        tp_gepp(F, ids, own + pc, sel, flag, red_v, red_r, red_q);
        const int cnt = (own + pc) < TP_NB ? (own + pc) : TP_NB;
        if (tid < cnt) cids[tid] = ids[sel[tid]];
        if (tid == 0) s_cnt = cnt;
        __syncthreads();
    }

    // ---- 3. root: interchanges, zero pivots, message
    const uint32_t eR = P.epoch0 + 15u;
    if (g == 0) {
        // F rows sel[j] hold L11 \ U11 of the chosen rows (the last elimination
        // was on the final stack, or the local one when G == 1)
        for (int idx = tid; idx < TP_NB * TP_NB; idx += TP_THREADS) {
            const int j = idx / TP_NB, c = idx % TP_NB;
            U[j * TP_PITCH + c] = F[sel[j] * TP_PITCH + c];
        }
        if (tid < TP_NB) pid[tid] = cids[tid];
        __syncthreads();
        if (tid == 0) {
            for (int i = 0; i < TP_NB; ++i) {
                top[i] = i;  // row now at top place i
                wt[i] = i;   // place of original top row i
            }
            int zc = -1;
            for (int j = 0; j < TP_NB; ++j) {
                const int32_t x = pid[j];
                const int32_t src = x < TP_NB ? wt[x] : x;
                P.ipiv[P.k + j] = (int32_t)(P.k + src);
                const int32_t rj = top[j];
                if (src < TP_NB) top[src] = rj;
                if (rj < TP_NB) wt[rj] = src;
                top[j] = x;
                if (x < TP_NB) wt[x] = j;
                const cpx<float> d = U[j * TP_PITCH + j];
                if (zc < 0 && d.re == 0.f && d.im == 0.f) zc = j;
            }
            if (zc >= 0) atomicCAS(P.info, 0, (int32_t)(P.k + zc + 1));
        }
        // publish pivot ids and L11 \ U11
        if (tid < TP_NB) tag_put(P.root + tid, eR, (uint32_t)pid[tid]);
        for (int idx = tid; idx < TP_NB * TP_NB; idx += TP_THREADS) {
            const cpx<float> x = U[(idx / TP_NB) * TP_PITCH + idx % TP_NB];
            tag_put(P.root + TP_NB + 2 * idx, eR, __float_as_uint(x.re));
            tag_put(P.root + TP_NB + 2 * idx + 1, eR, __float_as_uint(x.im));
        }
        __syncthreads();
        if (tid == 0) tag_put(P.root + TP_ROOT_W - 1, eR, 1u);  // ready
    } else {
        tp_wait_word(P.root + TP_ROOT_W - 1, eR, P.err);
        // every word of the root message in one round trip: thread t takes
        // values t + 256 i (i < 16) and pivot id t (t < 64)
        constexpr int VPT = TP_NB * TP_NB / TP_THREADS;
        uint64_t* pp[2 * VPT];
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const int idx = tid + TP_THREADS * i;
            pp[2 * i] = P.root + TP_NB + 2 * idx;
            pp[2 * i + 1] = P.root + TP_NB + 2 * idx + 1;
        }
        uint32_t wd[2 * VPT];
        tag_get_n<2 * VPT>(pp, eR, P.err, wd);
#pragma unroll
        for (int i = 0; i < VPT; ++i) {
            const int idx = tid + TP_THREADS * i;
            U[(idx / TP_NB) * TP_PITCH + idx % TP_NB] =
                cpx<float>{__uint_as_float(wd[2 * i]), __uint_as_float(wd[2 * i + 1])};
        }
        if (tid < TP_NB) {
            const int32_t v = (int32_t)tag_get(P.root + tid, eR, P.err);
            pid[tid] = (v >= 0 && v < P.m) ? v : 0;
        }
    }
    // ---- 4. L21 U11 = A21 for the non-pivot rows of this workgroup
    for (int r = tid; r < nr; r += TP_THREADS) piv_of[r] = -1;
    __syncthreads();
    if (tid < TP_NB && pid[tid] >= lr0 && pid[tid] < lr0 + nr) piv_of[pid[tid] - lr0] = tid;
    if (tid < TP_NB) {
        const cpx<float> d = U[tid * TP_PITCH + tid];
        const float dd = d.re * d.re + d.im * d.im;
        s_invd[tid] = dd == 0.f ? cpx<float>{0.f, 0.f} : cpx<float>{d.re / dd, -d.im / dd};
    }
    tp_load_rows(P, nullptr, lr0, nr, F);
    __syncthreads();
    // forward substitution along the row, thread pair (row, column parity)
    {
        for (int r0 = 0; r0 < nr; r0 += TP_THREADS / 2) {
            const int r = r0 + (tid >> 1), h = tid & 1;
            if (r < nr && piv_of[r] < 0) {
                cpx<float>* x = F + r * TP_PITCH;
                for (int c = 0; c < TP_NB; ++c) {
                    const cpx<float> a = x[c], di = s_invd[c];
                    cpx<float> l;
                    l.re = a.re * di.re - a.im * di.im;
                    l.im = a.re * di.im + a.im * di.re;
                    const cpx<float>* u = U + c * TP_PITCH;
#pragma unroll
                    for (int half = 0; half < 2; ++half) {
                        cpx<float> uu[TP_NB / 4], yy[TP_NB / 4];
#pragma unroll
                        for (int i = 0; i < TP_NB / 4; ++i) {
                            const int c2 = 2 * (i + half * TP_NB / 4) + h;
                            if (c2 > c) {
                                uu[i] = u[c2];
                                yy[i] = x[c2];
                            }
                        }
#pragma unroll
                        for (int i = 0; i < TP_NB / 4; ++i) {
                            const int c2 = 2 * (i + half * TP_NB / 4) + h;
                            if (c2 > c) {
                                yy[i].re -= l.re * uu[i].re - l.im * uu[i].im;
                                yy[i].im -= l.re * uu[i].im + l.im * uu[i].re;
                                x[c2] = yy[i];
                            }
                        }
                    }
                    if (h == 0) x[c] = l;  // both lanes (one wave) have read x[c]
                }
            }
        }
    }
    __syncthreads();
    // write: non-pivot rows to their final places, L11 \ U11 to rows k .. k + 63
    for (int idx = tid; idx < nr * TP_NB; idx += TP_THREADS) {
        const int r = idx / TP_NB, c = idx % TP_NB;
        if (piv_of[r] >= 0) continue;
        const int64_t x = lr0 + r;
        const int64_t dst = x < TP_NB ? (int64_t)wt[x] : x;  // only workgroup 0 has x < 64 (R >= 64)
        P.A[(P.k + dst) * P.lda + P.k + c] = F[r * TP_PITCH + c];
    }
    if (g == 0)
        for (int idx = tid; idx < TP_NB * TP_NB; idx += TP_THREADS) {
            const int j = idx / TP_NB, c = idx % TP_NB;
            P.A[(P.k + j) * P.lda + P.k + c] = U[j * TP_PITCH + c];
        }
}

// ============================================================== row swaps
// Plan (one thread): the panel's sequential swaps (k+j <-> ipiv[k+j]) act on
// the row set S = {k..k+nb-1} U {ipiv}; slots 0..nb-1 are rows k..k+nb-1, the
// other pivot rows get slots nb.. in order of first appearance.  plan[j] =
// (slot_a, slot_b) per swap; slot_row[s] = row of slot s; returns |S| in
// plan_n.
__device__ void swap_plan_body(const int32_t* __restrict__ ipiv, int64_t k, int nb, int32_t* __restrict__ slot_row,
                               int32_t* __restrict__ plan, int32_t* __restrict__ plan_n)
{
    int ns = nb;
    for (int s = 0; s < nb; ++s) slot_row[s] = (int32_t)(k + s);
    for (int j = 0; j < nb; ++j) {
        int32_t p = ipiv[k + j];
        int sp;
        if (p < k + nb) {
            sp = (int)(p - k);
        } else {
            sp = -1;
            for (int s = nb; s < ns; ++s)
                if (slot_row[s] == p) {
                    sp = s;
                    break;
                }
            if (sp < 0) {
                sp = ns;
                slot_row[ns++] = p;
            }
        }
        plan[2 * j] = j;
        plan[2 * j + 1] = sp;
    }
    *plan_n = ns;
}

// The same plan by 128 threads (nb <= 128), thread j for swap j: a far pivot
// row (>= k + nb) takes a new slot at its first appearance among the swaps
// (found by scanning the earlier pivots in LDS), numbered nb + (first
// appearances before it, from wave ballots); later appearances reuse it.
__global__ __launch_bounds__(128) void k_swap_plan(const int32_t* __restrict__ ipiv, int64_t k, int nb,
                                                   int32_t* __restrict__ slot_row, int32_t* __restrict__ plan,
                                                   int32_t* __restrict__ plan_n)
{
    __shared__ int32_t pv[132], first_of[128], slot_of[128];  // pv[128..131] = -1 (never a row)
    __shared__ int s_cnt0;
    const int j = threadIdx.x, lane = j & 63, wave = j >> 6;
    const bool act = j < nb;
    const int32_t p = act ? ipiv[k + j] : -1;
    pv[j] = p;
    if (j < 4) pv[128 + j] = -1;
    __syncthreads();
    const bool far = act && p >= k + nb;
    int f = j;  // first swap with the same pivot row; four LDS reads per step
    if (far)
        for (int i = 0; i < j; i += 4) {
            const int32_t q0 = pv[i], q1 = pv[i + 1], q2 = pv[i + 2], q3 = pv[i + 3];
            if (q0 == p) { f = i; break; }
            if (i + 1 < j && q1 == p) { f = i + 1; break; }
            if (i + 2 < j && q2 == p) { f = i + 2; break; }
            if (i + 3 < j && q3 == p) { f = i + 3; break; }
        }
    const bool is_first = far && f == j;
    first_of[j] = f;
    const unsigned long long bal = __ballot(is_first);
    const int below = __popcll(bal & ((1ull << lane) - 1ull));
    if (wave == 0 && lane == 0) s_cnt0 = __popcll(bal);
    __syncthreads();
    const int rank = below + (wave == 1 ? s_cnt0 : 0);
    if (is_first) {
        slot_of[j] = nb + rank;
        slot_row[nb + rank] = p;
    }
    if (act) slot_row[j] = (int32_t)(k + j);
    __syncthreads();
    if (act) {
        plan[2 * j] = j;
        plan[2 * j + 1] = far ? slot_of[first_of[j]] : (int32_t)(p - k);
    }
    if (wave == 1 && lane == 63) {
        // total = first appearances in both waves
        *plan_n = nb + s_cnt0 + __popcll(bal);
    }
}

// Each block: LASWP_COLS complex columns of the column range; the |S| rows are
// staged in LDS and written to their final rows (the interchanges composed
// into one permutation of the slots).
constexpr int LASWP_COLS = 16;

template <class T>
__device__ void laswp_body(cpx<T>* __restrict__ A, int64_t lda, int64_t c_lo, int64_t c_hi, int64_t skip_lo,
                           int64_t skip_hi, int nb, const int32_t* __restrict__ slot_row,
                           const int32_t* __restrict__ plan, const int32_t* __restrict__ plan_n, int64_t blk,
                           cpx<T>* buf)
{
    // 256 threads: thread = (slot group s0 = tid / 16, column c = tid % 16);
    // |S| <= 2 nb <= 256 slots, so 16 slots per thread
    constexpr int QS = 256 / (256 / LASWP_COLS);
    __shared__ int32_t srow[256], pl[256], cur[256], dst[256];
    const int ns = *plan_n;
    // map block -> columns, skipping [skip_lo, skip_hi)
    int64_t c0 = c_lo + blk * LASWP_COLS;
    if (c0 >= skip_lo) c0 += (skip_hi - skip_lo);
    if (c0 >= c_hi) return;
    int64_t cend = c0 + LASWP_COLS;
    if (c0 < skip_lo && cend > skip_lo) cend = skip_lo;
    if (cend > c_hi) cend = c_hi;
    const int w = (int)(cend - c0);
    const int tid = threadIdx.x;
    if (tid < ns) {
        srow[tid] = slot_row[tid];
        cur[tid] = tid;
    }
    if (tid < 2 * nb) pl[tid] = plan[tid];
    __syncthreads();
    // the rows' values, all loads of a thread in flight together
    const int c = tid % LASWP_COLS, s0 = tid / LASWP_COLS;
    for (int q0 = 0; q0 < QS; q0 += 8) {
        cpx<T> v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int sl = s0 + (256 / LASWP_COLS) * (q0 + q);
            v[q] = (sl < ns && c < w) ? A[(int64_t)srow[sl] * lda + c0 + c] : cpx<T>{(T)0, (T)0};
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int sl = s0 + (256 / LASWP_COLS) * (q0 + q);
            if (sl < ns && c < w) buf[sl * LASWP_COLS + c] = v[q];
        }
    }
    // the panel's interchanges, in order, on slot labels (one thread): the
    // same for every column, so done once instead of once per column
    if (tid == 0)
        for (int j = 0; j < nb; ++j) {
            const int sa = pl[2 * j], sb = pl[2 * j + 1];
            const int t = cur[sa];
            cur[sa] = cur[sb];
            cur[sb] = t;
        }
    __syncthreads();
    if (tid < ns) dst[cur[tid]] = tid;  // place tid now holds original slot cur[tid]
    __syncthreads();
#pragma unroll
    for (int q = 0; q < QS; ++q) {
        const int sl = s0 + (256 / LASWP_COLS) * q;
        if (sl < ns && c < w) A[(int64_t)srow[dst[sl]] * lda + c0 + c] = buf[sl * LASWP_COLS + c];
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_laswp(cpx<T>* __restrict__ A, int64_t lda, int64_t c_lo,
                                               int64_t c_hi, int64_t skip_lo, int64_t skip_hi,
                                               int nb, const int32_t* __restrict__ slot_row,
                                               const int32_t* __restrict__ plan,
                                               const int32_t* __restrict__ plan_n)
{
    extern __shared__ unsigned char smem_raw[];
    laswp_body<T>(A, lda, c_lo, c_hi, skip_lo, skip_hi, nb, slot_row, plan, plan_n, (int64_t)blockIdx.x,
                  (cpx<T>*)smem_raw);  // buffer [2 nb][LASWP_COLS]
}

// ============================================================== TRSM (L unit lower)
// X = L11^{-1} X for X = rows [k, k+nb) x columns [c0, c0 + 64) of A, L11 the
// unit lower panel block.  X is staged in LDS and solved in blocks of 16 rows:
// each block first subtracts the contribution of the solved blocks above it
// (thread = (row, 4 columns)), then a 16 x 16 forward substitution runs with
// one thread per column -- two barriers per block instead of two per row.
// Columns per workgroup of the panel triangular solve (FP32 / FP64 complex).
// 32 columns keep X + Ls at 48 KiB (complex64), so three workgroups share a
// CU by LDS; the complex64 kernels are also held to three waves per SIMD (168
// VGPRs; unbounded, the compiler takes 198 and fits only two workgroups).
template <class T>
struct TrsmCols {
    static constexpr int value = 32;
};
constexpr int TRSM_RB = 16;

// X = L11^{-1} X for the nb x (columns of this workgroup) block X of rows
// [k, k + nb): X in LDS, L11 read in row blocks of TRSM_RB rows staged in LDS
// (Ls [TRSM_RB][nb]).  Per block: X(block) -= L(block, done) X(done) by all
// threads, then the forward substitution inside the block, one thread per
// column.  Global loads are issued in batches (all of a thread's loads
// before its LDS stores), so each staging step costs one memory latency.
template <class T>
__device__ void trsm_body(cpx<T>* __restrict__ A, int64_t lda, int64_t k, int nb, int64_t col_begin,
                          int64_t col_end, int64_t blk, cpx<T>* X /* LDS [nb][TC] */,
                          cpx<T>* Ls /* LDS [TRSM_RB][nb] */)
{
    constexpr int TC = TrsmCols<T>::value;
    constexpr int CPT = TC / 16;  // columns per thread in the update role
    constexpr int XB = 16;        // X loads in flight per thread
    const int tid = threadIdx.x;
    const int64_t c0 = col_begin + blk * TC;
    const int w = (int)((col_end - c0) < TC ? (col_end - c0) : TC);
    for (int base = 0; base < nb * TC; base += 256 * XB) {
        cpx<T> v[XB];
#pragma unroll
        for (int q = 0; q < XB; ++q) {
            const int idx = base + tid + 256 * q;
            const int r = idx / TC, c = idx % TC;
            v[q] = (idx < nb * TC && c < w) ? A[(k + r) * lda + c0 + c] : cpx<T>{(T)0, (T)0};
        }
#pragma unroll
        for (int q = 0; q < XB; ++q) {
            const int idx = base + tid + 256 * q;
            if (idx < nb * TC) X[idx] = v[q];
        }
    }
    const int rr = tid >> 4, cg = (tid & 15) * CPT;  // update role: row rr of the block, columns cg..
    constexpr int LPT = TRSM_RB * 128 / 256;          // L values per thread (nb <= 128)
    for (int b0 = 0; b0 < nb; b0 += TRSM_RB) {
        const int rb = (nb - b0) < TRSM_RB ? (nb - b0) : TRSM_RB;
        const int width = b0 + rb;  // L columns this block needs
        __syncthreads();            // previous block done with Ls (and X loaded)
        {
            cpx<T> lv[LPT];
#pragma unroll
            for (int q = 0; q < LPT; ++q) {
                const int idx = tid + 256 * q;
                const int i = idx / width, t = idx % width;
                lv[q] = idx < rb * width ? A[(k + b0 + i) * lda + k + t] : cpx<T>{(T)0, (T)0};
            }
#pragma unroll
            for (int q = 0; q < LPT; ++q) {
                const int idx = tid + 256 * q;
                if (idx < rb * width) Ls[(idx / width) * nb + idx % width] = lv[q];
            }
        }
        __syncthreads();
        if (b0 > 0 && rr < rb) {
            const cpx<T>* Lr = Ls + rr * nb;
            T sr[CPT], si[CPT];
#pragma unroll
            for (int q = 0; q < CPT; ++q) sr[q] = si[q] = 0;
            for (int t0 = 0; t0 < b0; t0 += 4) {  // b0 is a multiple of TRSM_RB (16)
#pragma unroll
                for (int tt = 0; tt < 4; ++tt) {
                    const int t = t0 + tt;
                    const cpx<T> l = Lr[t];
#pragma unroll
                    for (int q = 0; q < CPT; ++q) {
                        const cpx<T> x = X[t * TC + cg + q];
                        sr[q] += l.re * x.re - l.im * x.im;
                        si[q] += l.re * x.im + l.im * x.re;
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < CPT; ++q) {
                cpx<T> x = X[(b0 + rr) * TC + cg + q];
                x.re -= sr[q];
                x.im -= si[q];
                X[(b0 + rr) * TC + cg + q] = x;
            }
        }
        __syncthreads();
        if (tid < TC && rb == TRSM_RB) {
            // full block: the column segment in registers, L rows read from
            // LDS (one address per wave)
            const int c = tid;
            cpx<T> y[TRSM_RB];
#pragma unroll
            for (int i = 0; i < TRSM_RB; ++i) y[i] = X[(b0 + i) * TC + c];
#pragma unroll
            for (int i = 1; i < TRSM_RB; ++i) {
                const cpx<T>* Lr = Ls + i * nb + b0;
#pragma unroll
                for (int t = 0; t < i; ++t) {
                    const cpx<T> l = Lr[t];
                    y[i].re -= l.re * y[t].re - l.im * y[t].im;
                    y[i].im -= l.re * y[t].im + l.im * y[t].re;
                }
            }
#pragma unroll
            for (int i = 1; i < TRSM_RB; ++i) X[(b0 + i) * TC + c] = y[i];
        } else if (tid < TC) {
            const int c = tid;
            for (int i = 1; i < rb; ++i) {
                const cpx<T>* Lr = Ls + i * nb + b0;
                cpx<T> x = X[(b0 + i) * TC + c];
                for (int t = 0; t < i; ++t) {
                    const cpx<T> l = Lr[t], y = X[(b0 + t) * TC + c];
                    x.re -= l.re * y.re - l.im * y.im;
                    x.im -= l.re * y.im + l.im * y.re;
                }
                X[(b0 + i) * TC + c] = x;
            }
        }
    }
    __syncthreads();
    for (int idx = tid; idx < nb * TC; idx += 256) {
        int r = idx / TC, cc = idx % TC;
        if (cc < w) A[(k + r) * lda + c0 + cc] = X[idx];
    }
}

// dynamic LDS of the panel triangular solve: X [nb][TC] + Ls [TRSM_RB][nb]
template <class T>
constexpr size_t trsm_lds_bytes(int nb)
{
    return ((size_t)nb * TrsmCols<T>::value + (size_t)TRSM_RB * nb) * sizeof(cpx<T>);
}

template <class T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 3 : 1) void k_trsm_l(cpx<T>* __restrict__ A, int64_t lda, int64_t k,
                                                int nb, int64_t col_begin, int64_t col_end)
{
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* X = (cpx<T>*)smem_raw;
    trsm_body<T>(A, lda, k, nb, col_begin, col_end, (int64_t)blockIdx.x, X, X + (size_t)nb * TrsmCols<T>::value);
}

// ---- complex64 panel triangular solve on MFMA (panels of 64 or 128 rows).
// Block triangular solve with inverted diagonal blocks (the scheme analysed by
// Du Croz and Higham, IMA J. Numer. Anal. 12 (1992), and used by blocked
// GPU trsm routines): for L11 = [D0 0; L21 D1] (64-row blocks),
//   Y0 = D0^{-1} X0,   Y1 = D1^{-1} (X1 - L21 Y0).
// k_trsm_prep inverts D0 and D1 (one workgroup each) and copies L21 (a third
// workgroup), all as MFMA A operands: planes [kk][i] (real, then imaginary)
// of 64 x 64 floats, kk the reduction index.  k_trsm_mfma then gives each
// workgroup 64 columns of X in LDS and runs the three products on
// v_mfma_f32_32x32x2f32 (three real products per complex one, 3M).
constexpr int TM_PLANE = 64 * 64;          // floats per plane
constexpr int TM_OPER = 2 * TM_PLANE;      // floats per operand (re, im)
constexpr int TM_SET = 3 * TM_OPER;        // D0^{-1}, D1^{-1}, L21 of one panel
constexpr int TM_SLOTS = 16;               // one set per panel of a super panel
constexpr int TM_WS_FLOATS = TM_SLOTS * TM_SET;
constexpr int TM_LD = 65;                  // padded LDS row of the inversion
constexpr size_t TM_PREP_LDS = (4 * 64 * TM_LD + 2 * 3 * 16 * 17) * sizeof(float);

// Inverse of the 64 x 64 unit lower triangle at (k0, k0): 16 x 16 diagonal
// blocks by forward substitution (lane c of wave w solves column c of block
// w), then the blocks below the diagonal one block diagonal at a time,
//   X_ij = -X_ii sum_{t=j}^{i-1} L_it X_tj      (i > j, 16-row blocks).
__device__ void tm_tri_inv64(const cpx<float>* __restrict__ A, int64_t lda, int64_t k0, float* __restrict__ out,
                             float* smem)
{
    float* Lr = smem;
    float* Li = Lr + 64 * TM_LD;
    float* Xr = Li + 64 * TM_LD;
    float* Xi = Xr + 64 * TM_LD;
    float* Tr = Xi + 64 * TM_LD;  // [3][16][17]
    float* Ti = Tr + 3 * 16 * 17;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    {
        cpx<float> v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int idx = tid + 256 * q, i = idx >> 6, t = idx & 63;
            v[q] = A[(k0 + i) * lda + k0 + t];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int idx = tid + 256 * q, i = idx >> 6, t = idx & 63;
            Lr[i * TM_LD + t] = v[q].re;
            Li[i * TM_LD + t] = v[q].im;
            Xr[i * TM_LD + t] = 0.f;
            Xi[i * TM_LD + t] = 0.f;
        }
    }
    __syncthreads();
    if (lane < 16) {
        const int c = lane, b = 16 * w;
        float xr[16], xi[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            xr[r] = r == c ? 1.f : 0.f;
            xi[r] = 0.f;
        }
#pragma unroll
        for (int r = 1; r < 16; ++r) {
            float sr = 0.f, si = 0.f;
#pragma unroll
            for (int t = 0; t < r; ++t) {
                const float lr = Lr[(b + r) * TM_LD + b + t], li = Li[(b + r) * TM_LD + b + t];
                sr += lr * xr[t] - li * xi[t];
                si += lr * xi[t] + li * xr[t];
            }
            if (r > c) {
                xr[r] = -sr;
                xi[r] = -si;
            }
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            Xr[(b + r) * TM_LD + b + c] = xr[r];
            Xi[(b + r) * TM_LD + b + c] = xi[r];
        }
    }
    __syncthreads();
    const int r = tid >> 4, c = tid & 15;
    for (int s = 1; s < 4; ++s) {
        for (int q = 0; q + s < 4; ++q) {  // block (q + s, q): S = sum_t L_(q+s),t X_t,q
            const int ri = 16 * (q + s) + r, cj = 16 * q + c;
            float sr = 0.f, si = 0.f;
            for (int tb = q; tb < q + s; ++tb) {
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int t = 16 * tb + u;
                    const float lr = Lr[ri * TM_LD + t], li = Li[ri * TM_LD + t];
                    const float xr = Xr[t * TM_LD + cj], xi = Xi[t * TM_LD + cj];
                    sr += lr * xr - li * xi;
                    si += lr * xi + li * xr;
                }
            }
            Tr[(q * 16 + r) * 17 + c] = sr;
            Ti[(q * 16 + r) * 17 + c] = si;
        }
        __syncthreads();
        for (int q = 0; q + s < 4; ++q) {  // X_(q+s),q = -X_(q+s),(q+s) S
            const int ri = 16 * (q + s) + r, d = 16 * (q + s);
            float yr = 0.f, yi = 0.f;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const float dr = Xr[ri * TM_LD + d + u], di = Xi[ri * TM_LD + d + u];
                const float tr = Tr[(q * 16 + u) * 17 + c], ti = Ti[(q * 16 + u) * 17 + c];
                yr += dr * tr - di * ti;
                yi += dr * ti + di * tr;
            }
            Xr[ri * TM_LD + 16 * q + c] = -yr;
            Xi[ri * TM_LD + 16 * q + c] = -yi;
        }
        __syncthreads();
    }
    for (int idx = tid; idx < 64 * 64; idx += 256) {
        const int kk = idx >> 6, i = idx & 63;
        out[idx] = Xr[i * TM_LD + kk];
        out[TM_PLANE + idx] = Xi[i * TM_LD + kk];
    }
}

// grid: 1 workgroup (nb = 64) or 3 (nb = 128: D0, D1, L21)
__global__ __launch_bounds__(256) void k_trsm_prep(const cpx<float>* __restrict__ A, int64_t lda, int64_t k,
                                                   float* __restrict__ ws)
{
    extern __shared__ unsigned char smem_raw[];
    float* smem = (float*)smem_raw;
    const int b = blockIdx.x;
    if (b < 2) {
        tm_tri_inv64(A, lda, k + 64 * b, ws + (size_t)b * TM_OPER, smem);
        return;
    }
    // L21 = rows [k + 64, k + 128) x columns [k, k + 64), transposed through LDS
    float* Lr = smem;
    float* Li = Lr + 64 * TM_LD;
    const int tid = threadIdx.x;
    {
        cpx<float> v[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int idx = tid + 256 * q, i = idx >> 6, t = idx & 63;
            v[q] = A[(k + 64 + i) * lda + k + t];
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int idx = tid + 256 * q, i = idx >> 6, t = idx & 63;
            Lr[i * TM_LD + t] = v[q].re;
            Li[i * TM_LD + t] = v[q].im;
        }
    }
    __syncthreads();
    float* out = ws + 2 * TM_OPER;
    for (int idx = tid; idx < 64 * 64; idx += 256) {
        const int kk = idx >> 6, i = idx & 63;
        out[idx] = Lr[i * TM_LD + kk];
        out[TM_PLANE + idx] = Li[i * TM_LD + kk];
    }
}

// A operands of wave row block bi for kk in [0, 2 * steps): lane (jj, hi) of
// step s holds P[kk = 2 s + hi][bi * 32 + jj] (P an operand of k_trsm_prep).
__device__ __forceinline__ void tm_load(const float* __restrict__ P, int bi, int steps, float (&ar)[32],
                                        float (&ai)[32])
{
    const int lane = threadIdx.x & 63, hi = lane >> 5, jj = lane & 31;
    const float* pr = P + hi * 64 + bi * 32 + jj;
    const float* pi = pr + TM_PLANE;
#pragma unroll
    for (int s = 0; s < 32; ++s) {
        ar[s] = s < steps ? pr[s * 128] : 0.f;
        ai[s] = s < steps ? pi[s * 128] : 0.f;
    }
}

// acc (32 x 32 block of wave (bi, bj)) += sign * P[bi rows] X[k rows] over
// kk in [0, 2 * steps): X planes in LDS with rows of 64 floats from row x0.
template <bool NEG>
__device__ __forceinline__ void tm_product(floatx16& cr, floatx16& ci, const float (&ar)[32], const float (&ai)[32],
                                           const float* Xr, const float* Xi, int x0, int bj, int steps)
{
    const int lane = threadIdx.x & 63, hi = lane >> 5, jj = lane & 31;
    // 3M: t1 = Cr + sum ar br, t2 = sum ai bi, t3 = Cr + Ci + sum (ar + ai)(br + bi);
    // Re = t1 - t2, Im = t3 - t1 - t2 (three chains that alternate, so no MFMA
    // waits on the one before it)
    floatx16 t1 = cr, t2, t3;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        t2[r] = 0.f;
        t3[r] = cr[r] + ci[r];
    }
#pragma unroll
    for (int s = 0; s < 32; ++s) {
        if (s < steps) {
            const int kr = x0 + 2 * s + hi;
            const float br = Xr[kr * 64 + bj * 32 + jj], bim = Xi[kr * 64 + bj * 32 + jj];
            const float a_r = NEG ? -ar[s] : ar[s], a_i = NEG ? -ai[s] : ai[s];
            t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r, br, t1, 0, 0, 0);
            t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_i, bim, t2, 0, 0, 0);
            t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r + a_i, br + bim, t3, 0, 0, 0);
        }
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        cr[r] = t1[r] - t2[r];
        ci[r] = t3[r] - t1[r] - t2[r];
    }
}

// X = L11^{-1} X for rows [k, k + NB) and columns [c_lo, c_hi), 64 columns per
// workgroup; ws holds the operands k_trsm_prep wrote for this panel.
template <int NB>
__global__ __launch_bounds__(256, 2) void k_trsm_mfma(cpx<float>* __restrict__ A, int64_t lda, int64_t k,
                                                      int64_t c_lo, int64_t c_hi, const float* __restrict__ ws)
{
    extern __shared__ unsigned char smem_raw[];
    float* Xr = (float*)smem_raw;  // [NB][64]
    float* Xi = Xr + NB * 64;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, hi = lane >> 5, jj = lane & 31;
    const int bi = w >> 1, bj = w & 1;
    const int64_t c0 = c_lo + (int64_t)blockIdx.x * 64;
    const int wc = (int)((c_hi - c0) < 64 ? (c_hi - c0) : 64);
    cpx<float>* X = A + k * lda + c0;
    // operands of the first two products in flight with the X loads, the
    // third's during the first product
    const int tri = 16 * (bi + 1);  // steps of the triangular D^{-1} blocks
    float a0r[32], a0i[32], a2r[32], a2i[32], a1r[32], a1i[32];
    tm_load(ws, bi, tri, a0r, a0i);
    if (NB == 128) tm_load(ws + 2 * TM_OPER, bi, 32, a2r, a2i);
    {
        constexpr int PER = NB * 64 / 256;
        float2 v[PER];
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int idx = tid + 256 * q, r = idx >> 6, c = idx & 63;
            v[q] = c < wc ? *(const float2*)(X + r * lda + c) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int idx = tid + 256 * q;
            Xr[idx] = v[q].x;
            Xi[idx] = v[q].y;
        }
    }
    __syncthreads();
    const int col = bj * 32 + jj;
    auto row_of = [&](int r) { return bi * 32 + 8 * (r >> 2) + 4 * hi + (r & 3); };
    floatx16 cr, ci;
#pragma unroll
    for (int r = 0; r < 16; ++r) cr[r] = ci[r] = 0.f;
    if (NB == 128) tm_load(ws + TM_OPER, bi, tri, a1r, a1i);
    // Y0 = D0^{-1} X0 (D0^{-1} is lower triangular: row block bi needs kk < 32 (bi + 1))
    tm_product<false>(cr, ci, a0r, a0i, Xr, Xi, 0, bj, tri);
    if (col < wc) {
#pragma unroll
        for (int r = 0; r < 16; ++r) *(float2*)(X + row_of(r) * lda + col) = make_float2(cr[r], ci[r]);
    }
    if (NB == 128) {
        __syncthreads();  // every wave is done reading X0
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            Xr[row_of(r) * 64 + col] = cr[r];
            Xi[row_of(r) * 64 + col] = ci[r];
            cr[r] = Xr[(64 + row_of(r)) * 64 + col];
            ci[r] = Xi[(64 + row_of(r)) * 64 + col];
        }
        __syncthreads();  // Y0 in LDS
        // X1 -= L21 Y0
        tm_product<true>(cr, ci, a2r, a2i, Xr, Xi, 0, bj, 32);
#pragma unroll
        for (int r = 0; r < 16; ++r) {  // each wave writes back only its own block of X1
            Xr[(64 + row_of(r)) * 64 + col] = cr[r];
            Xi[(64 + row_of(r)) * 64 + col] = ci[r];
            cr[r] = ci[r] = 0.f;
        }
        __syncthreads();
        // Y1 = D1^{-1} X1
        tm_product<false>(cr, ci, a1r, a1i, Xr, Xi, 64, bj, tri);
        if (col < wc) {
#pragma unroll
            for (int r = 0; r < 16; ++r)
                *(float2*)(X + (64 + row_of(r)) * lda + col) = make_float2(cr[r], ci[r]);
        }
    }
}

// ---- the same three steps for many matrices in one launch each: job j owns
// workgroups [wg_off[j], wg_off[j + 1]).
template <class T>
struct SwapJob {
    cpx<T>* A;
    int64_t lda;
    int64_t c_hi;     // columns [0, c_hi) are swapped, except the panel [k, k + nb)
    int64_t k;
    int nb;
    const int32_t* ipiv;
    int32_t* slot_row;
    int32_t* plan;
    int32_t* plan_n;
};

template <class T>
struct TrsmJob {
    cpx<T>* A;
    int64_t lda;
    int64_t k;
    int nb;
    int64_t c_lo, c_hi;
};

__device__ __forceinline__ int job_of(const int* __restrict__ wg_off, int njob, int b)
{
    int lo = 0, hi = njob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (wg_off[mid] <= b) lo = mid; else hi = mid;
    }
    return lo;
}

template <class T>
__global__ void k_swap_plan_batched(const SwapJob<T>* __restrict__ jobs)
{
    if (threadIdx.x != 0) return;
    const SwapJob<T> J = jobs[blockIdx.x];
    swap_plan_body(J.ipiv, J.k, J.nb, J.slot_row, J.plan, J.plan_n);
}

template <class T>
__global__ __launch_bounds__(256) void k_laswp_batched(const SwapJob<T>* __restrict__ jobs,
                                                       const int* __restrict__ wg_off, int njob)
{
    extern __shared__ unsigned char smem_raw[];
    const int j = job_of(wg_off, njob, (int)blockIdx.x);
    const SwapJob<T> J = jobs[j];
    laswp_body<T>(J.A, J.lda, 0, J.c_hi, J.k, J.k + J.nb, J.nb, J.slot_row, J.plan, J.plan_n,
                  (int64_t)blockIdx.x - wg_off[j], (cpx<T>*)smem_raw);
}

template <class T>
__global__ __launch_bounds__(256, sizeof(T) == 4 ? 3 : 1) void k_trsm_batched(const TrsmJob<T>* __restrict__ jobs,
                                                      const int* __restrict__ wg_off, int njob)
{
    extern __shared__ unsigned char smem_raw[];
    const int j = job_of(wg_off, njob, (int)blockIdx.x);
    const TrsmJob<T> J = jobs[j];
    cpx<T>* X = (cpx<T>*)smem_raw;
    trsm_body<T>(J.A, J.lda, J.k, J.nb, J.c_lo, J.c_hi, (int64_t)blockIdx.x - wg_off[j], X,
                 X + (size_t)J.nb * TrsmCols<T>::value);
}

__global__ void k_zero_u64(unsigned long long* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ULL;
}

__global__ void k_copy_i32(const int32_t* __restrict__ src, int32_t* __restrict__ dst, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_zero_i32(int32_t* p, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ============================================================== workspace
struct LuWs {
    int* err;
    uint64_t* meta;
    uint64_t* cval;
    uint64_t* cand;
    uint64_t* diag;
    uint64_t* tp_cand;  // tournament panel: candidate sets, root message
    uint64_t* tp_root;
    int32_t* slot_row;
    int32_t* plan;
    int32_t* plan_n;
    void* split;        // operands of the split-FP16 trailing update (complex64 only)
    size_t split_bytes;
    float* tm;          // operands of the MFMA panel triangular solve (complex64 only)
    size_t total;
    size_t msg_words;  // words from meta to the end of diag (zeroed per factorisation)
};

constexpr int MAX_G = 1024;

__host__ __device__ inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

LuWs lu_ws_layout(int prec, int64_t n, char* base)
{
    const size_t VW = prec == MBA_F32 ? 1 : 2;
    const int nbmax = 128;
    LuWs w;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += al256(bytes);
        return p;
    };
    w.err = (int*)take(256);
    const size_t msg0 = off;
    w.meta = (uint64_t*)take(2 * MAX_G * 8);
    w.cval = (uint64_t*)take(2 * MAX_G * VW * 8);
    w.cand = (uint64_t*)take(2 * MAX_G * (size_t)nbmax * 2 * VW * 8);
    w.diag = (uint64_t*)take(2 * (size_t)nbmax * 2 * VW * 8);
    w.tp_cand = (uint64_t*)take((size_t)MAX_G * TP_CAND_W * 8);
    w.tp_root = (uint64_t*)take((size_t)TP_ROOT_W * 8);
    w.msg_words = (off - msg0) / 8;
    w.slot_row = (int32_t*)take(2 * nbmax * 4);
    w.plan = (int32_t*)take(2 * nbmax * 4);
    w.plan_n = (int32_t*)take(256);
    w.split_bytes = prec == MBA_F32 ? mba_h3::workspace_bytes(n, n, mba_h3::KMAX) : 0;
    w.split = w.split_bytes ? (void*)take(w.split_bytes) : nullptr;
    w.tm = prec == MBA_F32 ? (float*)take((size_t)TM_WS_FLOATS * sizeof(float)) : nullptr;
    w.total = off;
    (void)n;
    return w;
}

int device_cus()
{
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return cus;
}

// Launch of a kernel whose workgroups wait on each other (panel exchanges,
// grid barriers): every workgroup must be resident at once.  The grid is
// checked against what the device holds at once (occupancy per CU x CUs, the
// same check a cooperative launch makes) and launched as an ordinary kernel.
// None of these kernels uses the runtime's grid synchronisation, and an
// ordinary launch keeps the runtime's cooperative queue (and what a profiler
// hooks onto it at exit) out of the process.  Workgroups of a grid that fits
// are all dispatched as the CUs free up from earlier work, so the waits end;
// each wait is bounded in any case.
int launch_resident(const void* fn, unsigned grid, unsigned block, void** args, size_t shm, hipStream_t st)
{
    const int cus = device_cus();
    int occ = 0;
    if (cus <= 0 || hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, (int)block, shm) != hipSuccess)
        return MBA_ERR_LAUNCH;
    if ((int64_t)grid > (int64_t)occ * cus) return MBA_ERR_SIZE;
    return hipLaunchKernel(fn, dim3(grid), dim3(block), args, shm, st) == hipSuccess ? MBA_OK : MBA_ERR_LAUNCH;
}

// Optional per-phase timing: HIP events recorded on the factorisation stream
// around each phase of every panel, summed after the final synchronisation.
struct PhaseTimer {
    bool on = false;
    hipStream_t st = nullptr;
    std::vector<hipEvent_t> ev[5];
    void mark(int phase)
    {
        if (!on) return;
        hipEvent_t e;
        if (hipEventCreate(&e) == hipSuccess) {
            (void)hipEventRecord(e, st);
            ev[phase].push_back(e);
        }
    }
    // phases: 0 panel, 1 row swaps, 2 triangular solve, 3 FP32 / FP64 MFMA
    // products and the operand splits of the split-FP16 products, 4 split-FP16
    // trailing-update products (k_cgemm_h3 alone)
    void collect(double* ms, int64_t* launches)
    {
        for (int p = 0; p < 5; ++p) {
            double tot = 0.0;
            for (size_t i = 0; i + 1 < ev[p].size(); i += 2) {
                float t = 0.f;
                if (hipEventElapsedTime(&t, ev[p][i], ev[p][i + 1]) == hipSuccess) tot += t;
            }
            if (ms) ms[p] = tot;
            if (launches) launches[p] = (int64_t)(ev[p].size() / 2);
            for (auto e : ev[p]) (void)hipEventDestroy(e);
            ev[p].clear();
        }
    }
};

template <class T>
int lu_factor(void* a, int64_t n, int64_t lda, int32_t* ipiv, int32_t* info, void* ws_ptr,
              size_t ws_bytes, hipStream_t st, double* phase_ms = nullptr,
              int64_t* phase_launches = nullptr, double* gemm_flops = nullptr, int max_wg = 0,
              int32_t* err_out = nullptr)
{
    PhaseTimer tm;
    tm.on = phase_ms != nullptr;
    tm.st = st;
    double gflops = 0.0, gflops_split = 0.0;
    const int prec = sizeof(T) == 4 ? MBA_F32 : MBA_F64;
    LuWs ws = lu_ws_layout(prec, n, (char*)ws_ptr);
    if (ws_bytes < ws.total) return MBA_ERR_ALLOC;
    cpx<T>* A = (cpx<T>*)a;
    const int cus = device_cus();
    if (cus <= 0) return MBA_ERR_LAUNCH;
    const size_t esz = sizeof(cpx<T>);
    // dynamic LDS of the panel; the static arrays add up to 11 KiB (FP64), the
    // total stays under the 160 KiB of a CU
    const size_t lds_budget = 140 * 1024;
    // panel width: largest of 128/64/32 whose per-workgroup row share fits the
    // LDS budget, with the workgroup count the panel launches will use (one
    // per CU, at most max_wg when the caller caps it, at most MAX_G)
    int64_t gcap = cus;
    if (max_wg > 0 && max_wg < gcap) gcap = max_wg;
    if (gcap > MAX_G) gcap = MAX_G;
    int NB = 128;
    while (NB > 16 && (size_t)((n + gcap - 1) / gcap) * NB * esz > lds_budget) NB /= 2;
    if ((size_t)((n + gcap - 1) / gcap) * NB * esz > lds_budget) return MBA_ERR_SIZE;
    const int upd_kind = g_lu_update_f32.load(std::memory_order_relaxed);
    const bool tournament = sizeof(T) == 4 && g_lu_panel_f32.load(std::memory_order_relaxed) == 1 && NB >= TP_NB;
    if (tournament) NB = TP_NB;
    // kernel attributes: set once per process (per element type); call_once
    // makes concurrent first calls from several host threads safe
    static std::once_flag attr_once;
    std::call_once(attr_once, [&] {
        (void)hipFuncSetAttribute((const void*)k_panel<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_budget);
        (void)hipFuncSetAttribute((const void*)k_trsm_l<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)trsm_lds_bytes<T>(128));
        if (sizeof(T) == 4) {
            (void)hipFuncSetAttribute((const void*)k_panel_tp, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)((TP_RMAX + TP_NB) * TP_PITCH * sizeof(cpx<float>)));
            (void)hipFuncSetAttribute((const void*)k_trsm_prep, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)TM_PREP_LDS);
            (void)hipFuncSetAttribute((const void*)k_trsm_mfma<64>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(64 * 64 * 2 * sizeof(float)));
            (void)hipFuncSetAttribute((const void*)k_trsm_mfma<128>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)(128 * 64 * 2 * sizeof(float)));
        }
    });
    // err := 0, every message word := epoch 0 (epochs of this factorisation are k + j + 1 >= 1)
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, (unsigned long long*)ws.err, 32);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((ws.msg_words + 255) / 256)), dim3(256), 0, st,
                       (unsigned long long*)ws.meta, (int)ws.msg_words);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, info, 1);
    MBA_LAUNCH_CHECK();
    // A[r0:r0+M, c0:c0+N] -= A[r0:r0+M, k0:k0+kk] A[k0:k0+kk, c0:c0+N]
    auto do_gemm_rect = [&](int64_t r0, int64_t M, int64_t c0, int64_t N, int64_t k0, int64_t kk) -> int {
        if (M <= 0 || N <= 0 || kk <= 0) return MBA_OK;
        tm.mark(3);
        int rc = gemm<T>(M, N, kk, -1.0, A + r0 * lda + k0, lda, A + k0 * lda + c0, lda, 1.0,
                         A + r0 * lda + c0, lda, st);
        if (rc != MBA_OK) return rc;
        tm.mark(3);
        gflops += 8.0 * (double)M * (double)N * (double)kk;
        return MBA_OK;
    };
    // One panel of nb columns at k: factor, then apply its row interchanges to
    // the column ranges [0, k) and [k + nb, swap_hi).
    auto do_panel = [&](int64_t k, int nb, int64_t swap_hi) -> int {
        const int64_t m = n - k;
        const int tp_cap = max_wg > 0 && max_wg < cus ? max_wg : cus;
        if (tournament && nb == TP_NB && m >= TP_MIN_M && m <= (int64_t)TP_RMAX * tp_cap) {
            // G workgroups of R rows, 64 <= R <= 128, none empty
            int64_t Gt = m / TP_NB < tp_cap ? m / TP_NB : tp_cap;
            const int64_t Rt = (m + Gt - 1) / Gt;
            Gt = (m + Rt - 1) / Rt;
            TpArgs TP;
            TP.A = (cpx<float>*)A;
            TP.lda = lda;
            TP.k = k;
            TP.m = m;
            TP.R = (int)Rt;
            TP.ipiv = ipiv;
            TP.info = info;
            TP.err = ws.err;
            TP.cand = ws.tp_cand;
            TP.root = ws.tp_root;
            TP.epoch0 = (uint32_t)((k / TP_NB + 1) * 16);
            void* targs[] = {&TP};
            tm.mark(0);
            const int lr = launch_resident((const void*)k_panel_tp, (unsigned)Gt, TP_THREADS, targs,
                                           (TP_RMAX + TP_NB) * TP_PITCH * sizeof(cpx<float>), st);
            if (lr != MBA_OK) return lr;
            tm.mark(0);
        } else {
            // >= 32 rows per workgroup: fewer candidates to gather and less
            // spread between workgroups for short panels (measured best)
            int G = (int)((m + 31) / 32);
            if (G > cus) G = cus;
            if (max_wg > 0 && G > max_wg) G = max_wg;
            if (G > MAX_G) G = MAX_G;
            int R = (int)((m + G - 1) / G);
            G = (int)((m + R - 1) / R);
            PanelArgs<T> P;
            P.A = A;
            P.lda = lda;
            P.k = k;
            P.m = m;
            P.nb = nb;
            P.R = R;
            P.ipiv = ipiv;
            P.info = info;
            P.err = ws.err;
            P.meta = ws.meta;
            P.cval = ws.cval;
            P.cand = ws.cand;
            P.diag = ws.diag;
            void* args[] = {&P};
            size_t shm = (size_t)R * nb * esz;
            tm.mark(0);
            const int lr = launch_resident((const void*)k_panel<T>, (unsigned)G, PANEL_THREADS, args, shm, st);
            if (lr != MBA_OK) return lr;
            tm.mark(0);
        }
        const int64_t right = swap_hi - (k + nb);
        if (k > 0 || right > 0) {
            tm.mark(1);
            hipLaunchKernelGGL(k_swap_plan, dim3(1), dim3(128), 0, st, ipiv, k, nb, ws.slot_row,
                               ws.plan, ws.plan_n);
            MBA_LAUNCH_CHECK();
            const int64_t blocks = (k + LASWP_COLS - 1) / LASWP_COLS + (right + LASWP_COLS - 1) / LASWP_COLS;
            // blocks enumerate [0, k) then [k+nb, swap_hi); the kernel skips the panel range
            hipLaunchKernelGGL(k_laswp<T>, dim3((unsigned)blocks), dim3(256),
                               (size_t)2 * nb * LASWP_COLS * esz, st, A, lda, (int64_t)0, swap_hi,
                               k, k + nb, nb, ws.slot_row, ws.plan, ws.plan_n);
            MBA_LAUNCH_CHECK();
            tm.mark(1);
        }
        return MBA_OK;
    };
    // rows [r0, r0 + nb) x columns [c_lo, c_hi): X = L(r0 block)^{-1} X.  slot:
    // the panel's place in its super panel; its inverted diagonal blocks are
    // formed by the first call that has columns to solve and kept for the
    // second-half panels' later call past the super panel.
    bool tm_ready[TM_SLOTS] = {};
    auto do_trsm = [&](int64_t r0, int nb, int64_t c_lo, int64_t c_hi, int slot) -> int {
        if (c_lo >= c_hi) return MBA_OK;
        tm.mark(2);
        if (sizeof(T) == 4 && (nb == 64 || nb == 128) && slot >= 0 && slot < TM_SLOTS) {
            // complex64: inverted diagonal blocks, then three MFMA products per
            // 64 columns
            float* set = ws.tm + (size_t)slot * TM_SET;
            if (!tm_ready[slot]) {
                hipLaunchKernelGGL(k_trsm_prep, dim3(nb == 128 ? 3u : 1u), dim3(256), TM_PREP_LDS, st,
                                   (const cpx<float>*)A, lda, r0, set);
                MBA_LAUNCH_CHECK();
                tm_ready[slot] = true;
            }
            const dim3 g((unsigned)((c_hi - c_lo + 63) / 64));
            if (nb == 128)
                hipLaunchKernelGGL(k_trsm_mfma<128>, g, dim3(256), (size_t)128 * 64 * 2 * sizeof(float), st,
                                   (cpx<float>*)A, lda, r0, c_lo, c_hi, (const float*)set);
            else
                hipLaunchKernelGGL(k_trsm_mfma<64>, g, dim3(256), (size_t)64 * 64 * 2 * sizeof(float), st,
                                   (cpx<float>*)A, lda, r0, c_lo, c_hi, (const float*)set);
            MBA_LAUNCH_CHECK();
        } else {
            constexpr int TC = TrsmCols<T>::value;
            hipLaunchKernelGGL(k_trsm_l<T>, dim3((unsigned)((c_hi - c_lo + TC - 1) / TC)), dim3(256),
                               trsm_lds_bytes<T>(nb), st, A, lda, r0, nb, c_lo, c_hi);
            MBA_LAUNCH_CHECK();
        }
        tm.mark(2);
        return MBA_OK;
    };
    // Super panels of P panels of NB columns (left-looking inside, right-looking
    // across super panels).  Panel p: its columns are first updated by the
    // super panel's earlier panels (rows [kp, n)), then factored (its row
    // interchanges applied to every column outside it, including the earlier
    // panels' L columns), then its U rows to its right get the delayed update
    // by the earlier panels and the triangular solve.  One trailing update with
    // k = P NB follows.  Row interchanges commute with the delayed (linear)
    // updates because they are applied to every column before them.
    // P NB = 1024 with the split-FP16 trailing product (a wider k spreads the
    // C read / write over more MFMA work: 316 -> 359 TFLOP/s at 20 480^2, k 512
    // -> 1024), 256 otherwise (NB = 64 with tournament panels, else 128 while
    // the LDS allows).  With the split product the super panel is two halves,
    // each left-looking on its own, so the narrow FP32 products only span one
    // half's earlier panels: when the second half starts, the first half's
    // contribution to its columns (rows [kh, n)) is one split product of
    // k = P NB / 2; the second half's U rows past the super panel are formed
    // after all its interchanges (the panels never read those columns), the
    // first half's part again as one product.
    const bool split_upd = sizeof(T) == 4 && upd_kind == 1;
    const int P = (split_upd ? 8 : 2) * (tournament ? 2 : 1);
    const int HALF = split_upd ? P / 2 : P;
    // C[r0 : r0 + M, c0 : c0 + N] -= A[r0 :, k0 : k0 + kk] A[k0 : k0 + kk, c0 :], on
    // the split-FP16 product when both sides are large enough, else FP32
    auto do_gemm_wide = [&](int64_t r0, int64_t M, int64_t c0, int64_t N, int64_t k0, int64_t kk) -> int {
        if (M <= 0 || N <= 0 || kk <= 0) return MBA_OK;
        if (M + N < SPLIT_MIN_DIM) return do_gemm_rect(r0, M, c0, N, k0, (int)kk);
        tm.mark(3);
        int rc2 = mba_h3::split(M, N, kk, A + r0 * lda + k0, lda, A + k0 * lda + c0, lda, ws.split,
                                ws.split_bytes, st);
        if (rc2 != MBA_OK) return rc2;
        tm.mark(3);
        tm.mark(4);
        rc2 = mba_h3::multiply(M, N, kk, -1.f, 1.f, A + r0 * lda + c0, lda, ws.split, ws.split_bytes, st);
        if (rc2 != MBA_OK) return rc2;
        tm.mark(4);
        gflops_split += 8.0 * (double)M * (double)N * (double)kk;
        return MBA_OK;
    };
    for (int64_t k = 0; k < n; k += (int64_t)P * NB) {
        int64_t kp = k;
        int rc;
        for (bool& r : tm_ready) r = false;
        const int64_t kh = k + (int64_t)HALF * NB;  // first column of the second half
        const int64_t ce = (k + (int64_t)P * NB) < n ? (k + (int64_t)P * NB) : n;  // end of the super panel
        bool two = false;  // the super panel is split into halves
        int np = 0;
        int64_t pr0[64];
        int pnb[64];
        for (int p = 0; p < P && kp < n; ++p) {
            const int nbp = (int)((n - kp) < NB ? (n - kp) : NB);
            if (p == HALF && split_upd) {
                // first half -> every column of the second half, rows [kp, n)
                if ((rc = do_gemm_wide(kp, n - kp, kp, ce - kp, k, kp - k)) != MBA_OK) return rc;
                two = true;
            }
            const int64_t kb = two ? kh : k;  // first column of this panel's half
            if (kp > kb && (rc = do_gemm_rect(kp, n - kp, kp, nbp, kb, (int)(kp - kb))) != MBA_OK) return rc;
            if ((rc = do_panel(kp, nbp, n)) != MBA_OK) return rc;
            const int64_t r0 = kp, right = kp + nbp;
            pr0[np] = r0;
            pnb[np++] = nbp;
            kp = right;
            if (right >= n) break;
            // this panel's U rows (its half's earlier panels): the first half's
            // over every column to the right, the second half's inside the
            // super panel only
            const int64_t ue = two ? ce : n;
            if ((rc = do_gemm_rect(r0, nbp, right, ue - right, kb, (int)(r0 - kb))) != MBA_OK) return rc;
            if ((rc = do_trsm(r0, nbp, right, ue, np - 1)) != MBA_OK) return rc;
        }
        if (two && ce < n) {
            // the second half's U rows past the super panel, once all its
            // interchanges are known: the first half's part in one product,
            // then one panel at a time inside the half
            if ((rc = do_gemm_wide(kh, ce - kh, ce, n - ce, k, kh - k)) != MBA_OK) return rc;
            for (int q = 0; q < np; ++q) {
                const int64_t r0 = pr0[q];
                if (r0 < kh) continue;
                if ((rc = do_gemm_rect(r0, pnb[q], ce, n - ce, kh, (int)(r0 - kh))) != MBA_OK) return rc;
                if ((rc = do_trsm(r0, pnb[q], ce, n, q)) != MBA_OK) return rc;
            }
        }
        const int64_t c3 = kp;  // first column after the super panel
        if (c3 >= n) break;
        const int64_t k3 = c3 - k;
        if (sizeof(T) == 4 && upd_kind == 1 && n - c3 >= SPLIT_MIN_DIM) {
            const int64_t m3 = n - c3;
            tm.mark(3);
            rc = mba_h3::split(m3, m3, k3, A + c3 * lda + k, lda, A + k * lda + c3, lda, ws.split, ws.split_bytes, st);
            if (rc != MBA_OK) return rc;
            tm.mark(3);
            tm.mark(4);
            rc = mba_h3::multiply(m3, m3, k3, -1.f, 1.f, A + c3 * lda + c3, lda, ws.split, ws.split_bytes, st);
            if (rc != MBA_OK) return rc;
            tm.mark(4);
            gflops_split += 8.0 * (double)m3 * (double)m3 * (double)k3;
        } else if ((rc = do_gemm_rect(c3, n - c3, c3, n - c3, k, (int)k3)) != MBA_OK) {
            return rc;
        }
    }
    if (err_out) {
        // asynchronous form: hand the wait flag to the caller, no host synchronisation
        hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, st, (const int32_t*)ws.err, err_out, 1);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    }
    int herr = 0;
    if (hipMemcpyAsync(&herr, ws.err, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return MBA_ERR_LAUNCH;
    if (hipStreamSynchronize(st) != hipSuccess) return MBA_ERR_LAUNCH;
    tm.collect(phase_ms, phase_launches);
    if (gemm_flops) {
        gemm_flops[0] = gflops;
        gemm_flops[1] = gflops_split;
    }
    return herr ? MBA_ERR_TIMEOUT : MBA_OK;
}

// ============================================================== triangular solves
// ipiv -> permutation gather list: after the sequential swaps, position i holds
// original row perm[i].  One wave; the permutation lives in LDS (n <= PERM_LDS)
// and pivots are read 64 at a time and broadcast lane by lane.
constexpr int PERM_LDS = 36 * 1024;

__device__ void ipiv_perm_body(const int32_t* __restrict__ ipiv, int64_t n, int32_t* __restrict__ perm)
{
    __shared__ int32_t pl[PERM_LDS];
    const int lane = threadIdx.x;
    const bool in_lds = n <= PERM_LDS;
    int32_t* P = in_lds ? pl : perm;
    for (int64_t i = lane; i < n; i += 64) P[i] = (int32_t)i;
    __syncthreads();
    for (int64_t c = 0; c < n; c += 64) {
        const int32_t v = (c + lane < n) ? ipiv[c + lane] : 0;
        // 64 interchanges at once when no two of the real ones (p != j) touch
        // a common place: then their order does not matter
        const int64_t jl = c + lane;
        const int32_t pl_ = (jl < n) ? v : (int32_t)jl;
        const bool real = pl_ != jl;
        bool conf = false;
        for (int q = 0; q < 64; ++q) {
            const int32_t pq = __builtin_amdgcn_readlane(pl_, q);
            const int64_t jq = c + q;
            if (q != lane && real && pq != jq && (pq == jl || pq == pl_ || jq == pl_)) conf = true;
        }
        if (!__any(conf)) {
            if (real) {
                const int32_t a = P[jl], b = P[pl_];
                P[jl] = b;
                P[pl_] = a;
            }
            continue;
        }
        // in order, skipping the interchanges with p == j (often about half of
        // them): walk the set bits of the wave's mask of real interchanges
        uint64_t mv = __ballot(real);
        while (mv) {
            const int q = __builtin_ctzll(mv);
            mv &= mv - 1;
            const int32_t p = __builtin_amdgcn_readlane(pl_, q);  // uniform q: a scalar read, no LDS trip
            if (lane == 0) {
                const int64_t j = c + q;
                int32_t t = P[j];
                P[j] = P[p];
                P[p] = t;
            }
        }
    }
    __syncthreads();
    if (in_lds)
        for (int64_t i = lane; i < n; i += 64) perm[i] = pl[i];
}

__global__ __launch_bounds__(64) void k_ipiv_perm(const int32_t* __restrict__ ipiv, int64_t n,
                                                  int32_t* __restrict__ perm)
{
    ipiv_perm_body(ipiv, n, perm);
}

template <class T>
__global__ void k_permute_rows(const cpx<T>* __restrict__ src, int64_t ld_src, const int32_t* perm,
                               int64_t n, int64_t nrhs, cpx<T>* __restrict__ dst, int64_t ld_dst)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= n * nrhs) return;
    int64_t i = id / nrhs, c = id % nrhs;
    const int64_t si = perm ? (int64_t)perm[i] : i;
    dst[i * ld_dst + c] = src[si * ld_src + c];
}

// Diagonal block solve, one workgroup per chunk of 4 right-hand sides: the
// bs x bs block is staged transposed in LDS (column t contiguous); column
// sweep: x_t is final, then every (row, rhs) thread updates its entry.
// lower: unit diagonal, forward; upper: divides by the diagonal, backward.
constexpr int DS_RHS = 4;

template <class T, bool LOWER>
__global__ __launch_bounds__(256) void k_diag_solve(const cpx<T>* __restrict__ A, int64_t lda,
                                                    int64_t r0, int bs, cpx<T>* __restrict__ B,
                                                    int64_t ldb, int64_t nrhs)
{
    extern __shared__ unsigned char smem_raw[];
    cpx<T>* Lt = (cpx<T>*)smem_raw;  // [bs][bs]: Lt[t * bs + i] = A[r0 + i][r0 + t]
    __shared__ cpx<T> X[128][DS_RHS];
    const int tid = threadIdx.x;
    const int64_t c0 = (int64_t)blockIdx.x * DS_RHS;
    const int nc = (int)((nrhs - c0) < DS_RHS ? (nrhs - c0) : DS_RHS);
    for (int idx = tid; idx < bs * bs; idx += 256) {
        int i = idx / bs, t = idx % bs;
        Lt[t * bs + i] = A[(r0 + i) * lda + r0 + t];
    }
    for (int idx = tid; idx < bs * DS_RHS; idx += 256) {
        int i = idx / DS_RHS, c = idx % DS_RHS;
        cpx<T> v;
        v.re = 0;
        v.im = 0;
        if (c < nc) v = B[(r0 + i) * ldb + c0 + c];
        X[i][c] = v;
    }
    __syncthreads();
    const int i = tid / DS_RHS, c = tid % DS_RHS;  // 64 rows per pass
    for (int s = 0; s < bs; ++s) {
        const int t = LOWER ? s : bs - 1 - s;
        if (!LOWER) {
            if (tid < DS_RHS) {
                cpx<T> d = Lt[t * bs + t];
                cpx<T> x = X[t][tid];
                T dd = d.re * d.re + d.im * d.im;
                cpx<T> q;
                q.re = (x.re * d.re + x.im * d.im) / dd;
                q.im = (x.im * d.re - x.re * d.im) / dd;
                X[t][tid] = q;
            }
            __syncthreads();
        }
        const cpx<T> xt = X[t][c];
        for (int ii = i; ii < bs; ii += 256 / DS_RHS) {
            if (LOWER ? (ii > t) : (ii < t)) {
                cpx<T> l = Lt[t * bs + ii];
                cpx<T> x = X[ii][c];
                x.re -= l.re * xt.re - l.im * xt.im;
                x.im -= l.re * xt.im + l.im * xt.re;
                X[ii][c] = x;
            }
        }
        __syncthreads();
    }
    for (int idx = tid; idx < bs * DS_RHS; idx += 256) {
        int ii = idx / DS_RHS, cc = idx % DS_RHS;
        if (cc < nc) B[(r0 + ii) * ldb + c0 + cc] = X[ii][cc];
    }
}

// Block update B[rows] -= A[rows, cols r0..r0+bs) * B[r0..r0+bs), one wave per
// (row, rhs); lanes stride the bs columns.
template <class T>
__global__ __launch_bounds__(256) void k_block_update(const cpx<T>* __restrict__ A, int64_t lda,
                                                      int64_t row_lo, int64_t row_hi, int64_t r0,
                                                      int bs, cpx<T>* __restrict__ B, int64_t ldb,
                                                      int64_t nrhs)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int64_t nrow = row_hi - row_lo;
    if (wid >= nrow * nrhs) return;
    const int64_t i = row_lo + wid / nrhs, c = wid % nrhs;
    const cpx<T>* Ar = A + i * lda + r0;
    T sr = 0, si = 0;
    for (int t = lane; t < bs; t += 64) {
        cpx<T> a = Ar[t], y = B[(r0 + t) * ldb + c];
        sr += a.re * y.re - a.im * y.im;
        si += a.re * y.im + a.im * y.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off);
        si += __shfl_xor(si, off);
    }
    if (lane == 0) {
        cpx<T> b = B[i * ldb + c];
        b.re -= sr;
        b.im -= si;
        B[i * ldb + c] = b;
    }
}

// ---- fused blocked triangular solves (nrhs <= TS_RHS) ----------------------
// Diagonal blocks of TS_BS rows are inverted once per factorisation
// (k_tri_inv); a solve is then one launch per block: workgroup 0 updates the
// next block's rows with the current block's solution and applies that
// block's inverse, the other workgroups update the remaining rows (one wave
// per row).  Lower: unit diagonal, forward; upper: backward.
constexpr int TS_BS = 64;
constexpr int TS_RHS = 4;

template <class T, bool LOWER>
__device__ void tri_inv_body(const cpx<T>* __restrict__ A, int64_t lda, int64_t n, cpx<T>* __restrict__ inv,
                             int64_t blk, cpx<T>* smem)
{
    cpx<T>* Lt = smem;                       // [TS_BS][TS_BS + 1], Lt[t][i] = A[r0+i][r0+t]
    cpx<T>* X = Lt + TS_BS * (TS_BS + 1);    // [TS_BS][TS_BS + 1], X[i][c]
    const int tid = threadIdx.x;
    const int64_t r0 = blk * TS_BS;
    const int bs = (int)((n - r0) < TS_BS ? (n - r0) : TS_BS);
    constexpr int LD = TS_BS + 1;
    for (int idx = tid; idx < TS_BS * TS_BS; idx += 256) {
        const int i = idx / TS_BS, t = idx % TS_BS;
        cpx<T> v;
        v.re = 0;
        v.im = 0;
        if (i < bs && t < bs) v = A[(r0 + i) * lda + r0 + t];
        Lt[t * LD + i] = v;
        cpx<T> e;
        e.re = (i == t && i < bs) ? (T)1 : (T)0;
        e.im = 0;
        X[i * LD + t] = e;
    }
    __syncthreads();
    for (int s = 0; s < bs; ++s) {
        const int t = LOWER ? s : bs - 1 - s;
        if (!LOWER) {
            if (tid < bs) {
                cpx<T> d = Lt[t * LD + t];
                cpx<T> x = X[t * LD + tid];
                T dd = d.re * d.re + d.im * d.im;
                cpx<T> q;
                q.re = (x.re * d.re + x.im * d.im) / dd;
                q.im = (x.im * d.re - x.re * d.im) / dd;
                X[t * LD + tid] = q;
            }
            __syncthreads();
        }
        for (int idx = tid; idx < bs * bs; idx += 256) {
            const int i = idx / bs, c = idx % bs;
            if (LOWER ? (i > t) : (i < t)) {
                const cpx<T> l = Lt[t * LD + i], xt = X[t * LD + c];
                cpx<T> x = X[i * LD + c];
                x.re -= l.re * xt.re - l.im * xt.im;
                x.im -= l.re * xt.im + l.im * xt.re;
                X[i * LD + c] = x;
            }
        }
        __syncthreads();
    }
    cpx<T>* out = inv + blk * TS_BS * TS_BS;
    for (int idx = tid; idx < TS_BS * TS_BS; idx += 256) out[idx] = X[(idx / TS_BS) * LD + idx % TS_BS];
}

template <class T, bool LOWER>
__global__ __launch_bounds__(256) void k_tri_inv(const cpx<T>* __restrict__ A, int64_t lda, int64_t n,
                                                 cpx<T>* __restrict__ inv)
{
    extern __shared__ unsigned char smem_raw[];
    tri_inv_body<T, LOWER>(A, lda, n, inv, (int64_t)blockIdx.x, (cpx<T>*)smem_raw);
}

// ---- solve tables of many factorisations: one launch for the permutations
// (one wave per job), one for the inverted diagonal blocks (job j owns
// workgroups [wg_off[j], wg_off[j + 1]): its L blocks, then its U blocks).
struct PrepJobDev {
    const void* lu;
    int64_t n, lda;
    const int32_t* ipiv;
    void* aux;
    size_t aux_bytes;
};
static_assert(sizeof(PrepJobDev) == sizeof(mba_lu_prepare_job), "descriptor layouts must match");

__global__ __launch_bounds__(64) void k_ipiv_perm_batched(const PrepJobDev* __restrict__ jobs)
{
    const PrepJobDev& J = jobs[blockIdx.x];
    ipiv_perm_body(J.ipiv, J.n, (int32_t*)J.aux);
}

template <class T>
__global__ __launch_bounds__(256) void k_tri_inv_batched(const PrepJobDev* __restrict__ jobs,
                                                         const int* __restrict__ wg_off, int njob)
{
    extern __shared__ unsigned char smem_raw[];
    int lo = 0, hi = njob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (wg_off[mid] <= (int)blockIdx.x) lo = mid; else hi = mid;
    }
    const PrepJobDev& J = jobs[lo];
    const int64_t nblk = (J.n + TS_BS - 1) / TS_BS;
    const int64_t u = (int64_t)blockIdx.x - wg_off[lo];
    cpx<T>* linv = (cpx<T>*)((char*)J.aux + al256((size_t)J.n * 4));
    cpx<T>* uinv = (cpx<T>*)((char*)linv + al256((size_t)nblk * TS_BS * TS_BS * sizeof(cpx<T>)));
    if (u < nblk)
        tri_inv_body<T, true>((const cpx<T>*)J.lu, J.lda, J.n, linv, u, (cpx<T>*)smem_raw);
    else
        tri_inv_body<T, false>((const cpx<T>*)J.lu, J.lda, J.n, uinv, u - nblk, (cpx<T>*)smem_raw);
}

// Tile of up to 64 rows handled by one workgroup: thread = (row tid/4, part
// tid%4); each part covers 16 consecutive columns (a 128/256-byte segment of
// the row) and the four partial sums are combined with two lane shuffles.
// out[i][c] = base[i][c] - sum_q M[i][q] v[q][c]  (sub = true)
// out[i][c] =               sum_q M[i][q] v[q][c]  (sub = false)
template <class T, bool SUB>
__device__ __forceinline__ void tile_matvec(const cpx<T>* __restrict__ M, int64_t ldm, int64_t row0,
                                            int nrows, int ncols, const cpx<T>* __restrict__ v,
                                            cpx<T>* __restrict__ out, int nrhs)
{
    const int r = threadIdx.x >> 2, part = threadIdx.x & 3;
    const bool active = r < nrows;
    const cpx<T>* Mr = M + (active ? (int64_t)r : 0) * ldm;
    cpx<T> m[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int col = part * 16 + q;
        if (active && col < ncols) {
            m[q] = Mr[col];
        } else {
            m[q].re = 0;
            m[q].im = 0;
        }
    }
    for (int c = 0; c < nrhs; ++c) {
        T sr = 0, si = 0;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
            const int col = part * 16 + q;
            if (col < ncols) {
                const cpx<T> y = v[col * nrhs + c];
                sr += m[q].re * y.re - m[q].im * y.im;
                si += m[q].re * y.im + m[q].im * y.re;
            }
        }
        sr += __shfl_xor(sr, 1);
        si += __shfl_xor(si, 1);
        sr += __shfl_xor(sr, 2);
        si += __shfl_xor(si, 2);
        if (active && part == 0) {
            cpx<T>* o = out + ((int64_t)row0 + r) * nrhs + c;
            if (SUB) {
                cpx<T> b = *o;
                b.re -= sr;
                b.im -= si;
                *o = b;
            } else {
                store_cpx(o, sr, si);
            }
        }
    }
}

// LOWER: current block [r0, r0 + bs) has its solution in x; next block
// [r0 + bs, r0 + bs + 64).  UPPER: current block [r0, r0 + bs); next block
// [r0 - 64, r0).  Workgroup 0 updates the next block and applies its inverse;
// workgroup b >= 1 updates the b-th 64-row tile of the remaining rows.
// first != 0: only x of block r0 = inv * t.
template <class T, bool LOWER>
__global__ __launch_bounds__(256) void k_tri_step(const cpx<T>* __restrict__ A, int64_t lda, int64_t n,
                                                  int64_t r0, int bs, const cpx<T>* __restrict__ inv,
                                                  cpx<T>* __restrict__ t, cpx<T>* __restrict__ x,
                                                  int nrhs, int first)
{
    if (first) {
        if (blockIdx.x == 0)
            tile_matvec<T, false>(inv + (r0 / TS_BS) * TS_BS * TS_BS, TS_BS, r0, bs, bs, t + r0 * nrhs, x,
                                  nrhs);
        return;
    }
    int64_t nx0, nx1, rest0, rest1;  // next block rows, remaining rows
    if (LOWER) {
        nx0 = r0 + bs;
        nx1 = nx0 + TS_BS < n ? nx0 + TS_BS : n;
        rest0 = nx1;
        rest1 = n;
    } else {
        nx1 = r0;
        nx0 = r0 - TS_BS > 0 ? r0 - TS_BS : 0;
        rest0 = 0;
        rest1 = nx0;
    }
    const cpx<T>* xc = x + r0 * nrhs;
    if (blockIdx.x == 0) {
        const int nn = (int)(nx1 - nx0);
        tile_matvec<T, true>(A + nx0 * lda + r0, lda, nx0, nn, bs, xc, t, nrhs);
        __syncthreads();
        tile_matvec<T, false>(inv + (nx0 / TS_BS) * TS_BS * TS_BS, TS_BS, nx0, nn, nn, t + nx0 * nrhs, x,
                              nrhs);
        return;
    }
    const int64_t i0 = rest0 + (int64_t)(blockIdx.x - 1) * 64;
    if (i0 >= rest1) return;
    const int nn = (int)((rest1 - i0) < 64 ? (rest1 - i0) : 64);
    tile_matvec<T, true>(A + i0 * lda + r0, lda, i0, nn, bs, xc, t, nrhs);
}

// ---- persistent triangular solves (complex64, nrhs <= TS_RHS) ---------------
// One launch per triangle.  Workgroup q (in start order: a ticket drawn from
// a counter) owns block row i = q (forward, unit lower) or nblk - 1 - q
// (backward, upper).  It subtracts A(i, j) x_j over the solved blocks j in
// solve order, each as soon as x_j appears in a tagged buffer (each 32-bit
// real travels with the call's epoch in one 64-bit word, so a reader needs no
// fence and polls exactly the words it reads), then applies the inverted
// diagonal block and publishes x_i.  A workgroup waits only on lower tickets,
// which started before it, so every wait ends whatever the dispatch order.
// Waits are bounded all the same: on expiry *err is set, waiters stop at their
// next look at it and the affected blocks come out as NaN.
//
// Bulk blocks (all but the one solved last before i): wave w takes rows
// 16w..16w+15, lane l column l of the 64-column block, the next block's loads
// in flight during the wait; the lane partial sums are reduced once by a
// halving butterfly (16 rows -> 1 row per lane quad in 6 exchanges).  The last
// block and the inverted diagonal block sit in registers before the wait
// (lane = row, wave = 16-column quarter), so the critical path per block is
// one tagged read, two small products and three barriers.
constexpr int PS_MAX_NR = TS_RHS;

// t = P b, tagged words := 0, tickets and error word := 0
__global__ void k_ps_prep(const cpx<float>* __restrict__ b, int64_t ldb, const int32_t* __restrict__ perm,
                          int64_t n, int nrhs, cpx<float>* __restrict__ t, uint64_t* __restrict__ tag,
                          int64_t tag_words, int* __restrict__ ctl)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n * nrhs) {
        const int64_t i = id / nrhs, c = id % nrhs;
        t[i * nrhs + c] = b[(int64_t)perm[i] * ldb + c];
    }
    if (id < tag_words) tag[id] = 0;
    if (id < 4) ctl[id] = 0;
}

// rhs [n][ldr], out [n][ldo]; tag [n][NR][2] words; ctl[0] ticket, ctl[2] error
template <int NR, bool LOWER>
__global__ __launch_bounds__(256) void k_trisolve_persist(const cpx<float>* __restrict__ A, int64_t lda, int64_t n,
                                                          const cpx<float>* __restrict__ inv,
                                                          const cpx<float>* __restrict__ rhs, int64_t ldr,
                                                          cpx<float>* __restrict__ out, int64_t ldo,
                                                          uint64_t* __restrict__ tag, uint32_t epoch,
                                                          int* __restrict__ ctl)
{
    __shared__ int s_ticket, s_bad;
    __shared__ cpx<float> s_bulk[TS_BS][NR];
    __shared__ cpx<float> s_part[4][TS_BS][NR];
    __shared__ cpx<float> s_vec[TS_BS][NR];
    const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
    int* err = ctl + 2;
    const int nblk = (int)((n + TS_BS - 1) / TS_BS);
    if (tid == 0) {
        s_ticket = atomicAdd(ctl + (LOWER ? 0 : 1), 1);
        s_bad = 0;
    }
    __syncthreads();
    const int q = s_ticket;
    if (q >= nblk) return;  // grid and block count agree; kept as a guard
    const int i = LOWER ? q : nblk - 1 - q;
    const int64_t r0 = (int64_t)i * TS_BS;
    bool bad = false;

    // critical-path operands: last block jl, inverted diagonal block (row l, quarter w)
    const int jl = LOWER ? i - 1 : i + 1;
    const bool has_last = jl >= 0 && jl < nblk;
    cpx<float> lb[16], iv[16];
    {
        const bool rv = r0 + l < n;
        const cpx<float>* ar = A + (rv ? r0 + l : 0) * lda + (int64_t)(has_last ? jl : 0) * TS_BS + 16 * w;
        const cpx<float>* ir = inv + (int64_t)i * TS_BS * TS_BS + l * TS_BS + 16 * w;
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const bool cv = has_last && rv && (int64_t)jl * TS_BS + 16 * w + cc < n;
            lb[cc] = cv ? ar[cc] : cpx<float>{0.f, 0.f};
            iv[cc] = ir[cc];
        }
    }
    cpx<float> rv_[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) rv_[k] = (w == 0 && r0 + l < n) ? rhs[(r0 + l) * ldr + k] : cpx<float>{0.f, 0.f};

    // bulk: blocks s = 0 .. nbulk - 1 in solve order
    const int nbulk = LOWER ? (i > 1 ? i - 1 : 0) : (nblk - 2 - i > 0 ? nblk - 2 - i : 0);
    cpx<float> acc[16][NR];
#pragma unroll
    for (int rr = 0; rr < 16; ++rr)
#pragma unroll
        for (int k = 0; k < NR; ++k) acc[rr][k] = cpx<float>{0.f, 0.f};
    const int64_t rowb = r0 + 16 * w;
    auto load_rows = [&](cpx<float>(&a)[16], int s) {
        const int j = LOWER ? s : nblk - 1 - s;
        const int64_t c = (int64_t)j * TS_BS + l;
#pragma unroll
        for (int rr = 0; rr < 16; ++rr)
            a[rr] = (rowb + rr < n && c < n) ? A[(rowb + rr) * lda + c] : cpx<float>{0.f, 0.f};
    };
    cpx<float> a[16], an[16];
    if (nbulk > 0) load_rows(a, 0);
    for (int s = 0; s < nbulk; ++s) {
        if (s + 1 < nbulk) load_rows(an, s + 1);
        const int j = LOWER ? s : nblk - 1 - s;
        const int64_t c = (int64_t)j * TS_BS + l;
        uint64_t* pp[2 * NR];
#pragma unroll
        for (int m = 0; m < 2 * NR; ++m) pp[m] = c < n ? tag + c * NR * 2 + m : nullptr;
        uint32_t wd[2 * NR] = {};
        if (!bad && !tag_get_opt<2 * NR>(pp, epoch, err, wd)) bad = true;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const float xr = __uint_as_float(wd[2 * k]), xi = __uint_as_float(wd[2 * k + 1]);
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) {
                acc[rr][k].re += a[rr].re * xr - a[rr].im * xi;
                acc[rr][k].im += a[rr].re * xi + a[rr].im * xr;
            }
        }
        if (s + 1 < nbulk) {
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) a[rr] = an[rr];
        }
    }
    // halving butterfly: after the step of width h, lanes keep rows [0, h) of
    // their half; lane quad (l >> 2) ends with row l >> 2 of the wave
#pragma unroll
    for (int h = 8, sh = 32; h >= 1; h >>= 1, sh >>= 1) {
        const bool up = (l & sh) != 0;
#pragma unroll
        for (int rr = 0; rr < h; ++rr)
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const cpx<float> lo = acc[rr][k], hi = acc[rr + h][k];
                const cpx<float> send = up ? lo : hi, keep = up ? hi : lo;
                acc[rr][k].re = keep.re + __shfl_xor(send.re, sh);
                acc[rr][k].im = keep.im + __shfl_xor(send.im, sh);
            }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        acc[0][k].re += __shfl_xor(acc[0][k].re, 2);
        acc[0][k].im += __shfl_xor(acc[0][k].im, 2);
        acc[0][k].re += __shfl_xor(acc[0][k].re, 1);
        acc[0][k].im += __shfl_xor(acc[0][k].im, 1);
    }
    if ((l & 3) == 0)
#pragma unroll
        for (int k = 0; k < NR; ++k) s_bulk[16 * w + (l >> 2)][k] = acc[0][k];

    // last block: wave w reads the 16 NR complex of its quarter, one or two
    // tagged words per lane, and broadcasts them by lane reads
    constexpr int XW = 32 * NR;
    constexpr int XPL = (XW + 63) / 64;
    cpx<float> d[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) d[k] = cpx<float>{0.f, 0.f};
    if (has_last) {
        const int64_t v0 = (int64_t)jl * TS_BS + 16 * w;
        uint64_t* pp[XPL];
#pragma unroll
        for (int m = 0; m < XPL; ++m) {
            const int idx = l + 64 * m;
            pp[m] = (idx < XW && v0 + idx / (2 * NR) < n) ? tag + v0 * NR * 2 + idx : nullptr;
        }
        uint32_t xv[XPL] = {};
        if (!bad && !tag_get_opt<XPL>(pp, epoch, err, xv)) bad = true;
#pragma unroll
        for (int cc = 0; cc < 16; ++cc)
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const int ir = (cc * NR + k) * 2, ii = ir + 1;
                const float xr = __uint_as_float(
                    (uint32_t)__builtin_amdgcn_readlane((int)xv[ir / 64], ir % 64));
                const float xi = __uint_as_float(
                    (uint32_t)__builtin_amdgcn_readlane((int)xv[ii / 64], ii % 64));
                d[k].re += lb[cc].re * xr - lb[cc].im * xi;
                d[k].im += lb[cc].re * xi + lb[cc].im * xr;
            }
    }
#pragma unroll
    for (int k = 0; k < NR; ++k) s_part[w][l][k] = d[k];
    if (bad) s_bad = 1;
    __syncthreads();
    if (w == 0) {
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            cpx<float> sv = rv_[k];
            sv.re -= s_bulk[l][k].re + s_part[0][l][k].re + s_part[1][l][k].re + s_part[2][l][k].re +
                     s_part[3][l][k].re;
            sv.im -= s_bulk[l][k].im + s_part[0][l][k].im + s_part[1][l][k].im + s_part[2][l][k].im +
                     s_part[3][l][k].im;
            s_vec[l][k] = r0 + l < n ? sv : cpx<float>{0.f, 0.f};
        }
    }
    __syncthreads();
    // x_i = inv(A_ii) s: quarter w of the columns, then the four partials
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        cpx<float> p = {0.f, 0.f};
#pragma unroll
        for (int cc = 0; cc < 16; ++cc) {
            const cpx<float> sv = s_vec[16 * w + cc][k];
            p.re += iv[cc].re * sv.re - iv[cc].im * sv.im;
            p.im += iv[cc].re * sv.im + iv[cc].im * sv.re;
        }
        s_part[w][l][k] = p;
    }
    __syncthreads();
    if (w == 0 && r0 + l < n) {
        const bool nan = s_bad != 0;
        const int64_t row = r0 + l;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            cpx<float> x;
            x.re = s_part[0][l][k].re + s_part[1][l][k].re + s_part[2][l][k].re + s_part[3][l][k].re;
            x.im = s_part[0][l][k].im + s_part[1][l][k].im + s_part[2][l][k].im + s_part[3][l][k].im;
            if (nan) x = cpx<float>{__int_as_float(0x7fc00000), __int_as_float(0x7fc00000)};
            put_val(tag + (row * NR + k) * 2, epoch, x.re);
            put_val(tag + (row * NR + k) * 2 + 1, epoch, x.im);
            out[row * ldo + k] = x;
        }
    }
}

size_t solve_aux_bytes(int prec, int64_t n)
{
    const size_t esz = prec == MBA_F32 ? 8 : 16;
    const int64_t nblk = (n + TS_BS - 1) / TS_BS;
    return al256((size_t)n * 4) + 2 * al256((size_t)nblk * TS_BS * TS_BS * esz);
}

template <class T>
int lu_prepare_batched(const mba_lu_prepare_job* jobs, int njob, void* scratch, size_t scratch_bytes, hipStream_t st)
{
    std::vector<int> wg(njob + 1, 0);
    for (int j = 0; j < njob; ++j) wg[j + 1] = wg[j] + 2 * (int)((jobs[j].n + TS_BS - 1) / TS_BS);
    const size_t o_w = al256((size_t)njob * sizeof(PrepJobDev));
    if (scratch_bytes < o_w + (size_t)(njob + 1) * sizeof(int)) return MBA_ERR_ALLOC;
    PrepJobDev* dj = (PrepJobDev*)scratch;
    int* dw = (int*)((char*)scratch + o_w);
    if (mba_host::upload_tables(dj, jobs, (size_t)njob * sizeof(PrepJobDev), st) != 0 ||
        mba_host::upload_tables(dw, wg.data(), (size_t)(njob + 1) * sizeof(int), st) != 0)
        return MBA_ERR_LAUNCH;
    const size_t shm = 2 * (size_t)TS_BS * (TS_BS + 1) * sizeof(cpx<T>);
    static std::once_flag attr_once;  // per element type (template instance)
    std::call_once(attr_once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tri_inv_batched<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)shm);
    });
    hipLaunchKernelGGL(k_ipiv_perm_batched, dim3((unsigned)njob), dim3(64), 0, st, (const PrepJobDev*)dj);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tri_inv_batched<T>, dim3((unsigned)wg[njob]), dim3(256), shm, st, (const PrepJobDev*)dj,
                       (const int*)dw, njob);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

template <class T>
int lu_prepare(const void* lu, int64_t n, int64_t lda, const int32_t* ipiv, void* aux, hipStream_t st)
{
    const size_t esz = sizeof(cpx<T>);
    const int64_t nblk = (n + TS_BS - 1) / TS_BS;
    int32_t* perm = (int32_t*)aux;
    cpx<T>* linv = (cpx<T>*)((char*)aux + al256((size_t)n * 4));
    cpx<T>* uinv = (cpx<T>*)((char*)linv + al256((size_t)nblk * TS_BS * TS_BS * esz));
    hipLaunchKernelGGL(k_ipiv_perm, dim3(1), dim3(64), 0, st, ipiv, n, perm);
    MBA_LAUNCH_CHECK();
    const size_t shm = 2 * (size_t)TS_BS * (TS_BS + 1) * esz;
    static std::once_flag attr_once;  // per element type (template instance)
    std::call_once(attr_once, [&] {
        (void)hipFuncSetAttribute((const void*)k_tri_inv<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
        (void)hipFuncSetAttribute((const void*)k_tri_inv<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm);
    });
    hipLaunchKernelGGL((k_tri_inv<T, true>), dim3((unsigned)nblk), dim3(256), shm, st, (const cpx<T>*)lu, lda, n, linv);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_tri_inv<T, false>), dim3((unsigned)nblk), dim3(256), shm, st, (const cpx<T>*)lu, lda, n, uinv);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// ---- many right-hand sides, many matrices: one launch per triangle --------
// Workgroup = (job, group of TSC right-hand sides).  It walks the 64-row
// blocks in order (forward for L, backward for U): t = rhs - A(block, done) x,
// then x(block) = inv(A_bb) t with the inverse prepared by lu_prepare.  The
// solution is kept in the job's work array [n, nrhs]; the forward pass reads
// the right-hand side through the row permutation, the backward pass writes
// the result back into b.
//
// Thread = (row i = tid / 4 of the block, k-lane kq = tid % 4).  The update
// A(block, done) x runs over chunks of CK columns: the chunk of x (CK x TSC) is
// staged in LDS, each thread holds its row's A entries k = 4 u + kq in
// registers and accumulates all TSC columns; the next chunk's loads are issued
// before the current chunk is multiplied.  The four k-lanes of a row are then
// summed with lane exchanges and lane kq owns columns 4 kq .. 4 kq + 3.
constexpr int TSC = 16;

struct TsJobDev {
    const void* lu;
    int64_t n, lda;
    const void* aux;
    void* b;
    int64_t ldb, nrhs;
    void* work;
};
static_assert(sizeof(TsJobDev) == sizeof(mba_lu_solve_job), "descriptor layouts must match");

template <class T>
struct TsGeom {
    static constexpr int CK = sizeof(T) == 4 ? 64 : 32;            // columns per staged chunk
    static constexpr int KPT = CK / 4;                               // chunk columns per thread
    static constexpr int LDC = TSC + 16 / (int)sizeof(cpx<T>);       // LDS row, 16 bytes of padding
    static constexpr int XPT = CK * TSC / 256;                       // staged elements per thread
};

template <class T>
__device__ __forceinline__ void quad_sum(T* v)
{
#pragma unroll
    for (int c = 0; c < TSC; ++c) {
        v[c] += __shfl_xor(v[c], 1);
        v[c] += __shfl_xor(v[c], 2);
    }
}

// Loads of one chunk: the thread's A entries (row Ar, columns k0 + 4 u + kq)
// and its staged x elements (rows k0 + (tid + 256 e) / TSC, column col); zero
// beyond qb.
template <class T>
__device__ __forceinline__ void ts_fetch(cpx<T> (&an)[TsGeom<T>::KPT], cpx<T> (&xn)[TsGeom<T>::XPT],
                                         const cpx<T>* __restrict__ Ar, const cpx<T>* __restrict__ X, int64_t k0,
                                         int64_t qb, int kq, bool rok, int tid, bool col_ok, int64_t nrhs,
                                         int64_t col)
{
#pragma unroll
    for (int u = 0; u < TsGeom<T>::KPT; ++u) {
        const int64_t k = k0 + 4 * u + kq;
        an[u].re = 0;
        an[u].im = 0;
        if (rok && k < qb) an[u] = Ar[k];
    }
#pragma unroll
    for (int e = 0; e < TsGeom<T>::XPT; ++e) {
        const int64_t kk = k0 + (tid + 256 * e) / TSC;
        xn[e].re = 0;
        xn[e].im = 0;
        if (kk < qb && col_ok) xn[e] = X[kk * nrhs + col];
    }
}

template <class T, bool LOWER>
__global__ __launch_bounds__(256) void k_trisolve_cols(const TsJobDev* __restrict__ jobs,
                                                       const int* __restrict__ wg_off, int njob)
{
    using G = TsGeom<T>;
    __shared__ cpx<T> xs[G::CK * G::LDC];
    __shared__ cpx<T> tb[TS_BS * G::LDC];
    int lo = 0, hi = njob;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (wg_off[mid] <= (int)blockIdx.x) lo = mid; else hi = mid;
    }
    const TsJobDev J = jobs[lo];
    const int64_t n = J.n, lda = J.lda, nrhs = J.nrhs;
    const int64_t c0 = (int64_t)(blockIdx.x - wg_off[lo]) * TSC;
    const cpx<T>* A = (const cpx<T>*)J.lu;
    const int32_t* perm = (const int32_t*)J.aux;
    const int64_t nblk = (n + TS_BS - 1) / TS_BS;
    const size_t esz = sizeof(cpx<T>);
    const cpx<T>* linv = (const cpx<T>*)((const char*)J.aux + al256((size_t)n * 4));
    const cpx<T>* uinv = (const cpx<T>*)((const char*)linv + al256((size_t)nblk * TS_BS * TS_BS * esz));
    const cpx<T>* inv = LOWER ? linv : uinv;
    cpx<T>* X = (cpx<T>*)J.work;
    cpx<T>* B = (cpx<T>*)J.b;
    const int tid = threadIdx.x, i = tid >> 2, kq = tid & 3;
    const int sc = tid % TSC;  // staged column: (tid + 256 e) % TSC == tid % TSC
    const bool sc_ok = c0 + sc < nrhs;
    cpx<T> zero;
    zero.re = 0;
    zero.im = 0;
    for (int64_t s = 0; s < nblk; ++s) {
        const int64_t bi = LOWER ? s : nblk - 1 - s;
        const int64_t r0 = bi * TS_BS;
        const int bs = (int)((n - r0) < TS_BS ? (n - r0) : TS_BS);
        const bool rok = i < bs;
        const int64_t row = r0 + (rok ? i : 0);
        const int64_t qa = LOWER ? 0 : r0 + bs, qb = LOWER ? r0 : n;
        T pr[TSC], pi[TSC];
#pragma unroll
        for (int c = 0; c < TSC; ++c) pr[c] = pi[c] = 0;
        const cpx<T>* Ar = A + row * lda;
        if (qa < qb) {
            cpx<T> an[G::KPT], xn[G::XPT];
            ts_fetch<T>(an, xn, Ar, X, qa, qb, kq, rok, tid, sc_ok, nrhs, c0 + sc);
            for (int64_t k0 = qa; k0 < qb; k0 += G::CK) {
                __syncthreads();  // the previous chunk's readers are done
#pragma unroll
                for (int e = 0; e < G::XPT; ++e) xs[((tid + 256 * e) / TSC) * G::LDC + sc] = xn[e];
                cpx<T> a[G::KPT];
#pragma unroll
                for (int u = 0; u < G::KPT; ++u) a[u] = an[u];
                __syncthreads();
                if (k0 + G::CK < qb) ts_fetch<T>(an, xn, Ar, X, k0 + G::CK, qb, kq, rok, tid, sc_ok, nrhs, c0 + sc);
#pragma unroll
                for (int u = 0; u < G::KPT; ++u) {
                    const cpx<T>* xr = xs + (4 * u + kq) * G::LDC;
#pragma unroll
                    for (int c = 0; c < TSC; ++c) {
                        const cpx<T> x = xr[c];
                        pr[c] += a[u].re * x.re - a[u].im * x.im;
                        pi[c] += a[u].re * x.im + a[u].im * x.re;
                    }
                }
            }
            quad_sum(pr);
            quad_sum(pi);
        }
        // t = rhs - update for the lane's own columns
        cpx<T> rhs[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t col = c0 + 4 * kq + q;
            rhs[q] = zero;
            if (rok && col < nrhs) rhs[q] = LOWER ? B[(int64_t)perm[row] * J.ldb + col] : X[row * nrhs + col];
        }
#pragma unroll
        for (int c = 0; c < TSC; ++c)
            if ((c >> 2) == kq) {
                cpx<T> v;
                v.re = rhs[c & 3].re - pr[c];
                v.im = rhs[c & 3].im - pi[c];
                tb[i * G::LDC + c] = v;
            }
        __syncthreads();
        // x(block) = inv(A_bb) t; inverse rows and columns beyond bs are zero
        T xr[TSC], xi[TSC];
#pragma unroll
        for (int c = 0; c < TSC; ++c) xr[c] = xi[c] = 0;
        const cpx<T>* Mi = inv + (bi * TS_BS + i) * TS_BS;
#pragma unroll 4
        for (int u = 0; u < TS_BS / 4; ++u) {
            const int t = 4 * u + kq;
            const cpx<T> m = Mi[t];
            const cpx<T>* tr = tb + t * G::LDC;
#pragma unroll
            for (int c = 0; c < TSC; ++c) {
                const cpx<T> v = tr[c];
                xr[c] += m.re * v.re - m.im * v.im;
                xi[c] += m.re * v.im + m.im * v.re;
            }
        }
        quad_sum(xr);
        quad_sum(xi);
#pragma unroll
        for (int c = 0; c < TSC; ++c)
            if ((c >> 2) == kq && rok && c0 + c < nrhs) store_cpx(X + row * nrhs + c0 + c, xr[c], xi[c]);
        __syncthreads();  // block rows of x visible to the next block's update; tb free
    }
    if (!LOWER) {
        for (int64_t r = i; r < n; r += TS_BS)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t col = c0 + 4 * kq + q;
                if (col < nrhs) B[r * J.ldb + col] = X[r * nrhs + col];
            }
    }
}

// FP32 form of the same walk on MFMA.  The update P = A(block, chunk) x(chunk)
// (64 rows x 16 complex columns) is a real product in the embedding
//   A' = [Re A, -Im A] (k' = 2 s + {0, 1}),  B' = [[Re x, Im x], [Im x, -Re x]],
// so P'[row][2 c] = Re P, P'[row][2 c + 1] = Im P: one v_mfma_f32_32x32x2f32 per
// complex k.  Wave w: rows 32 (w & 1) .. +32, k-half w >> 1 of each 32-column
// chunk; the two k-halves are summed through LDS.  The lane's A entries come
// from global memory into registers (the next chunk's loads are issued before
// the current chunk's products), x chunks are staged in LDS.  The diagonal
// block step x = inv(A_bb) t uses the same product with A = the inverse.
// Workgroups are dealt to XCDs so that the workgroups of one job share an L2.
constexpr int TSM_CK = 32;

__device__ __forceinline__ void tsm_fetch(float2 (&an)[TSM_CK / 2], float2 (&xn)[TSM_CK * TSC / 256],
                                          const cpx<float>* __restrict__ Ar, bool rok, int64_t kb,
                                          const cpx<float>* __restrict__ X, int64_t k0, int64_t qb, int tid,
                                          bool col_ok, int64_t nrhs, int64_t col)
{
#pragma unroll
    for (int u = 0; u < TSM_CK / 2; ++u) {
        const int64_t k = kb + u;
        an[u] = make_float2(0.f, 0.f);
        if (rok && k < qb) an[u] = *(const float2*)(Ar + k);
    }
#pragma unroll
    for (int e = 0; e < TSM_CK * TSC / 256; ++e) {
        const int64_t kk = k0 + (tid + 256 * e) / TSC;
        xn[e] = make_float2(0.f, 0.f);
        if (kk < qb && col_ok) xn[e] = *(const float2*)(X + kk * nrhs + col);
    }
}

template <bool LOWER>
__global__ __launch_bounds__(256, 2) void k_trisolve_mfma(const TsJobDev* __restrict__ jobs,
                                                          const int* __restrict__ wg_off, int njob)
{
    constexpr int LDX = TSC + 2;  // LDS row of 16 complex + 16 bytes
    constexpr int KH = TSM_CK / 2;
    constexpr int XPT = TSM_CK * TSC / 256;
    __shared__ cpx<float> xs[TSM_CK * LDX];
    __shared__ cpx<float> tb[TS_BS * LDX];
    __shared__ float red[2][16][64];
    const int nbk = (int)gridDim.x;
    const int b = (int)(blockIdx.x % NUM_XCD) * (nbk / NUM_XCD) + (int)(blockIdx.x / NUM_XCD);
    if (b >= wg_off[njob]) return;
    int lo = 0, hi_ = njob;
    while (hi_ - lo > 1) {
        const int mid = (lo + hi_) >> 1;
        if (wg_off[mid] <= b) lo = mid; else hi_ = mid;
    }
    const TsJobDev J = jobs[lo];
    const int64_t n = J.n, lda = J.lda, nrhs = J.nrhs;
    const int64_t c0 = (int64_t)(b - wg_off[lo]) * TSC;
    const cpx<float>* A = (const cpx<float>*)J.lu;
    const int32_t* perm = (const int32_t*)J.aux;
    const int64_t nblk = (n + TS_BS - 1) / TS_BS;
    const cpx<float>* linv = (const cpx<float>*)((const char*)J.aux + al256((size_t)n * 4));
    const cpx<float>* uinv =
        (const cpx<float>*)((const char*)linv + al256((size_t)nblk * TS_BS * TS_BS * sizeof(cpx<float>)));
    const cpx<float>* inv = LOWER ? linv : uinv;
    cpx<float>* X = (cpx<float>*)J.work;
    cpx<float>* B = (cpx<float>*)J.b;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rt = wave & 1, kh = wave >> 1;
    const int hi = lane >> 5, l32 = lane & 31;
    const int arow = 32 * rt + l32;       // block row of the lane's A operand
    const int bc = l32 >> 1, bq = l32 & 1;  // complex column and part of the lane's B operand / output
    const bool bim = (hi ^ bq) != 0;
    const float bsg = (hi & bq) ? -1.f : 1.f;
    const int sc = tid % TSC;
    const bool sc_ok = c0 + sc < nrhs;
    const bool oc_ok = c0 + bc < nrhs;
    for (int64_t st = 0; st < nblk; ++st) {
        const int64_t bi = LOWER ? st : nblk - 1 - st;
        const int64_t r0 = bi * TS_BS;
        const int bs = (int)((n - r0) < TS_BS ? (n - r0) : TS_BS);
        const bool rok = arow < bs;
        const cpx<float>* Ar = A + (r0 + (rok ? arow : 0)) * lda;
        const int64_t qa = LOWER ? 0 : r0 + bs, qb = LOWER ? r0 : n;
        floatx16 acc;
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
        if (qa < qb) {
            float2 an[KH], xn[XPT];
            tsm_fetch(an, xn, Ar, rok, qa + kh * KH, X, qa, qb, tid, sc_ok, nrhs, c0 + sc);
            for (int64_t k0 = qa; k0 < qb; k0 += TSM_CK) {
                __syncthreads();  // the previous chunk's readers are done
#pragma unroll
                for (int e = 0; e < XPT; ++e) {
                    cpx<float> v;
                    v.re = xn[e].x;
                    v.im = xn[e].y;
                    xs[((tid + 256 * e) / TSC) * LDX + sc] = v;
                }
                float2 a[KH];
#pragma unroll
                for (int u = 0; u < KH; ++u) a[u] = an[u];
                __syncthreads();
                if (k0 + TSM_CK < qb)
                    tsm_fetch(an, xn, Ar, rok, k0 + TSM_CK + kh * KH, X, k0 + TSM_CK, qb, tid, sc_ok, nrhs, c0 + sc);
#pragma unroll
                for (int u = 0; u < KH; ++u) {
                    const cpx<float> v = xs[(kh * KH + u) * LDX + bc];
                    const float bv = bsg * (bim ? v.im : v.re);
                    const float av = hi ? -a[u].y : a[u].x;
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
                }
            }
        }
        // t = rhs - P into tb (waves of k-half 0 after adding k-half 1)
        if (kh == 1)
#pragma unroll
            for (int r = 0; r < 16; ++r) red[rt][r][lane] = acc[r];
        __syncthreads();
        if (kh == 0) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * rt + 8 * (r >> 2) + 4 * hi + (r & 3);
                float rhs = 0.f, p = acc[r] + red[rt][r][lane];
                if (row < bs && oc_ok) {
                    const float* src = LOWER ? (const float*)(B + (int64_t)perm[r0 + row] * J.ldb + c0 + bc)
                                             : (const float*)(X + (r0 + row) * nrhs + c0 + bc);
                    rhs = src[bq];
                } else {
                    p = 0.f;
                }
                ((float*)(tb + row * LDX + bc))[bq] = rhs - p;
            }
        }
        __syncthreads();
        // x(block) = inv(A_bb) t; the inverse is zero beyond bs
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[r] = 0.f;
        {
            constexpr int KI = TS_BS / 2;  // block columns per k-half
            const cpx<float>* Mi = inv + (bi * TS_BS + arow) * TS_BS + kh * KI;
#pragma unroll 8
            for (int u = 0; u < KI; ++u) {
                const float2 m = *(const float2*)(Mi + u);
                const cpx<float> v = tb[(kh * KI + u) * LDX + bc];
                const float bv = bsg * (bim ? v.im : v.re);
                const float av = hi ? -m.y : m.x;
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv, acc, 0, 0, 0);
            }
        }
        if (kh == 1)
#pragma unroll
            for (int r = 0; r < 16; ++r) red[rt][r][lane] = acc[r];
        __syncthreads();
        if (kh == 0 && oc_ok) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int row = 32 * rt + 8 * (r >> 2) + 4 * hi + (r & 3);
                if (row < bs) ((float*)(X + (r0 + row) * nrhs + c0 + bc))[bq] = acc[r] + red[rt][r][lane];
            }
        }
        __syncthreads();  // block rows of x visible to the next block's update; tb, red free
    }
    if (!LOWER) {
        for (int64_t idx = tid; idx < n * TSC; idx += 256) {
            const int64_t r = idx / TSC, c = idx % TSC;
            if (c0 + c < nrhs) B[r * J.ldb + c0 + c] = X[r * nrhs + c0 + c];
        }
    }
}

template <class T>
int lu_solve_batched(const mba_lu_solve_job* jobs, int njob, void* scratch, size_t scratch_bytes, hipStream_t st)
{
    std::vector<int> wg(njob + 1, 0);
    for (int m = 0; m < njob; ++m) wg[m + 1] = wg[m] + (int)((jobs[m].nrhs + TSC - 1) / TSC);
    const size_t need = al256((size_t)njob * sizeof(TsJobDev)) + (size_t)(njob + 1) * sizeof(int);
    if (scratch_bytes < need) return MBA_ERR_ALLOC;
    char* base = (char*)scratch;
    TsJobDev* dj = (TsJobDev*)base;
    int* dw = (int*)(base + al256((size_t)njob * sizeof(TsJobDev)));
    if (mba_host::upload_tables(dj, jobs, (size_t)njob * sizeof(TsJobDev), st) != 0 ||
        mba_host::upload_tables(dw, wg.data(), (size_t)(njob + 1) * sizeof(int), st) != 0)
        return MBA_ERR_LAUNCH;
    if (sizeof(T) == 4) {
        const unsigned blocks = (unsigned)(((wg[njob] + NUM_XCD - 1) / NUM_XCD) * NUM_XCD);
        hipLaunchKernelGGL(k_trisolve_mfma<true>, dim3(blocks), dim3(256), 0, st, dj, dw, njob);
        MBA_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_trisolve_mfma<false>, dim3(blocks), dim3(256), 0, st, dj, dw, njob);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    }
    hipLaunchKernelGGL((k_trisolve_cols<T, true>), dim3((unsigned)wg[njob]), dim3(256), 0, st, dj, dw, njob);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL((k_trisolve_cols<T, false>), dim3((unsigned)wg[njob]), dim3(256), 0, st, dj, dw, njob);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// Solve with the prepared aux (perm + inverted diagonal blocks), nrhs <= TS_RHS.
template <class T>
int lu_solve_fused(const void* lu, int64_t n, int64_t lda, const void* aux, void* b, int64_t nrhs,
                   int64_t ldb, void* ws_ptr, size_t ws_bytes, hipStream_t st)
{
    const size_t esz = sizeof(cpx<T>);
    if (ws_bytes < 2 * al256((size_t)n * nrhs * esz)) return MBA_ERR_ALLOC;
    const int64_t nblk = (n + TS_BS - 1) / TS_BS;
    const int32_t* perm = (const int32_t*)aux;
    const cpx<T>* linv = (const cpx<T>*)((const char*)aux + al256((size_t)n * 4));
    const cpx<T>* uinv = (const cpx<T>*)((const char*)linv + al256((size_t)nblk * TS_BS * TS_BS * esz));
    cpx<T>* t = (cpx<T>*)ws_ptr;
    cpx<T>* x = (cpx<T>*)((char*)ws_ptr + al256((size_t)n * nrhs * esz));
    const cpx<T>* A = (const cpx<T>*)lu;
    const int64_t tot = n * nrhs;
    hipLaunchKernelGGL(k_permute_rows<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       (const cpx<T>*)b, ldb, perm, n, nrhs, t, nrhs);
    MBA_LAUNCH_CHECK();
    const int nr = (int)nrhs;
    // forward: L y = P b  (y in x)
    for (int64_t bi = 0; bi < nblk; ++bi) {
        const int64_t r0 = bi * TS_BS;
        const int bs = (int)((n - r0) < TS_BS ? (n - r0) : TS_BS);
        if (bi == 0)
            hipLaunchKernelGGL((k_tri_step<T, true>), dim3(1), dim3(256), 0, st, A, lda, n, r0, bs,
                               linv, t, x, nr, 1);
        const int64_t rest = n - (r0 + bs) - TS_BS;
        const int64_t blocks = 1 + (rest > 0 ? (rest + 63) / 64 : 0);
        if (r0 + bs < n)
            hipLaunchKernelGGL((k_tri_step<T, true>), dim3((unsigned)blocks), dim3(256), 0, st, A, lda, n,
                               r0, bs, linv, t, x, nr, 0);
        MBA_LAUNCH_CHECK();
    }
    // backward: U z = y  (t <- y, result in x)
    hipLaunchKernelGGL(k_permute_rows<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       (const cpx<T>*)x, nrhs, nullptr, n, nrhs, t, nrhs);
    MBA_LAUNCH_CHECK();
    for (int64_t bi = nblk - 1; bi >= 0; --bi) {
        const int64_t r0 = bi * TS_BS;
        const int bs = (int)((n - r0) < TS_BS ? (n - r0) : TS_BS);
        if (bi == nblk - 1)
            hipLaunchKernelGGL((k_tri_step<T, false>), dim3(1), dim3(256), 0, st, A, lda, n, r0, bs,
                               uinv, t, x, nr, 1);
        const int64_t rest = r0 - TS_BS;
        const int64_t blocks = 1 + (rest > 0 ? (rest + 63) / 64 : 0);
        if (r0 > 0)
            hipLaunchKernelGGL((k_tri_step<T, false>), dim3((unsigned)blocks), dim3(256), 0, st, A, lda, n,
                               r0, bs, uinv, t, x, nr, 0);
        MBA_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_permute_rows<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       (const cpx<T>*)x, nrhs, nullptr, n, nrhs, (cpx<T>*)b, ldb);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// complex64 triangular solves of mba_lu_solve for nrhs <= TS_RHS: 1 = one
// persistent launch per triangle (default), 0 = one launch per 64-row block;
// set by mba_lu_set_solve
std::atomic<int> g_lu_solve_f32{1};

size_t persist_ws_bytes(int64_t n, int64_t nrhs)
{
    return 2 * al256((size_t)n * nrhs * 8) + al256((size_t)n * nrhs * 16) + 256;
}

template <int NR>
void launch_persist(const cpx<float>* A, int64_t lda, int64_t n, const cpx<float>* linv, const cpx<float>* uinv,
                    const cpx<float>* t, cpx<float>* x, cpx<float>* b, int64_t ldb, uint64_t* tag, int* ctl,
                    hipStream_t st)
{
    const unsigned nblk = (unsigned)((n + TS_BS - 1) / TS_BS);
    hipLaunchKernelGGL((k_trisolve_persist<NR, true>), dim3(nblk), dim3(256), 0, st, A, lda, n, linv, t,
                       (int64_t)NR, x, (int64_t)NR, tag, 1u, ctl);
    hipLaunchKernelGGL((k_trisolve_persist<NR, false>), dim3(nblk), dim3(256), 0, st, A, lda, n, uinv,
                       (const cpx<float>*)x, (int64_t)NR, b, ldb, tag, 2u, ctl);
}

// P A = L U: t = P b, L y = t (y in x), U z = y (z into b).  Asynchronous; a
// wait that gave up leaves NaN in the affected rows.
int lu_solve_persist(const void* lu, int64_t n, int64_t lda, const void* aux, void* b, int64_t nrhs, int64_t ldb,
                     void* ws_ptr, size_t ws_bytes, hipStream_t st)
{
    if (ws_bytes < persist_ws_bytes(n, nrhs)) return MBA_ERR_ALLOC;
    const size_t esz = sizeof(cpx<float>);
    const int64_t nblk = (n + TS_BS - 1) / TS_BS;
    const int32_t* perm = (const int32_t*)aux;
    const cpx<float>* linv = (const cpx<float>*)((const char*)aux + al256((size_t)n * 4));
    const cpx<float>* uinv = (const cpx<float>*)((const char*)linv + al256((size_t)nblk * TS_BS * TS_BS * esz));
    const size_t tb = al256((size_t)n * nrhs * esz);
    cpx<float>* t = (cpx<float>*)ws_ptr;
    cpx<float>* x = (cpx<float>*)((char*)ws_ptr + tb);
    uint64_t* tag = (uint64_t*)((char*)ws_ptr + 2 * tb);
    int* ctl = (int*)((char*)tag + al256((size_t)n * nrhs * 16));
    const int64_t tag_words = n * nrhs * 2;
    hipLaunchKernelGGL(k_ps_prep, dim3((unsigned)((tag_words + 255) / 256)), dim3(256), 0, st,
                       (const cpx<float>*)b, ldb, perm, n, (int)nrhs, t, tag, tag_words, ctl);
    MBA_LAUNCH_CHECK();
    const cpx<float>* A = (const cpx<float>*)lu;
    cpx<float>* B = (cpx<float>*)b;
    switch (nrhs) {
    case 1: launch_persist<1>(A, lda, n, linv, uinv, t, x, B, ldb, tag, ctl, st); break;
    case 2: launch_persist<2>(A, lda, n, linv, uinv, t, x, B, ldb, tag, ctl, st); break;
    case 3: launch_persist<3>(A, lda, n, linv, uinv, t, x, B, ldb, tag, ctl, st); break;
    default: launch_persist<4>(A, lda, n, linv, uinv, t, x, B, ldb, tag, ctl, st); break;
    }
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

template <class T>
int lu_solve(const void* lu, int64_t n, int64_t lda, const int32_t* perm, void* b, int64_t nrhs,
             int64_t ldb, void* ws_ptr, size_t ws_bytes, hipStream_t st)
{
    const size_t esz = sizeof(cpx<T>);
    if (ws_bytes < (size_t)n * nrhs * esz) return MBA_ERR_ALLOC;
    cpx<T>* tmp = (cpx<T>*)ws_ptr;
    const cpx<T>* A = (const cpx<T>*)lu;
    cpx<T>* B = (cpx<T>*)b;
    const int64_t tot = n * nrhs;
    hipLaunchKernelGGL(k_permute_rows<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, B, ldb,
                       perm, n, nrhs, tmp, nrhs);
    MBA_LAUNCH_CHECK();
    const int BS = 64;
    const size_t dshm = (size_t)BS * BS * esz;
    static std::once_flag attr_once;  // per element type (template instance)
    std::call_once(attr_once, [&] {
        (void)hipFuncSetAttribute((const void*)k_diag_solve<T, true>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)dshm);
        (void)hipFuncSetAttribute((const void*)k_diag_solve<T, false>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)dshm);
    });
    const unsigned dgrid = (unsigned)((nrhs + DS_RHS - 1) / DS_RHS);
    // forward: L y = P b (tmp holds P b, ld = nrhs)
    for (int64_t r0 = 0; r0 < n; r0 += BS) {
        const int bs = (int)((n - r0) < BS ? (n - r0) : BS);
        hipLaunchKernelGGL((k_diag_solve<T, true>), dim3(dgrid), dim3(256), dshm, st, A, lda, r0, bs,
                           tmp, nrhs, nrhs);
        MBA_LAUNCH_CHECK();
        const int64_t lo = r0 + bs;
        if (lo < n) {
            if (nrhs >= 32) {
                int rc = gemm<T>(n - lo, nrhs, bs, -1.0, A + lo * lda + r0, lda, tmp + r0 * nrhs, nrhs,
                                 1.0, tmp + lo * nrhs, nrhs, st);
                if (rc != MBA_OK) return rc;
            } else {
                const int64_t waves = (n - lo) * nrhs;
                hipLaunchKernelGGL(k_block_update<T>, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256),
                                   0, st, A, lda, lo, n, r0, bs, tmp, nrhs, nrhs);
                MBA_LAUNCH_CHECK();
            }
        }
    }
    // backward: U x = y
    const int64_t nblk = (n + BS - 1) / BS;
    for (int64_t bi = nblk - 1; bi >= 0; --bi) {
        const int64_t r0 = bi * BS;
        const int bs = (int)((n - r0) < BS ? (n - r0) : BS);
        hipLaunchKernelGGL((k_diag_solve<T, false>), dim3(dgrid), dim3(256), dshm, st, A, lda, r0, bs,
                           tmp, nrhs, nrhs);
        MBA_LAUNCH_CHECK();
        if (r0 > 0) {
            if (nrhs >= 32) {
                int rc = gemm<T>(r0, nrhs, bs, -1.0, A + r0, lda, tmp + r0 * nrhs, nrhs, 1.0, tmp, nrhs, st);
                if (rc != MBA_OK) return rc;
            } else {
                const int64_t waves = r0 * nrhs;
                hipLaunchKernelGGL(k_block_update<T>, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256),
                                   0, st, A, lda, (int64_t)0, r0, r0, bs, tmp, nrhs, nrhs);
                MBA_LAUNCH_CHECK();
            }
        }
    }
    // copy back into b (ld = ldb)
    hipLaunchKernelGGL(k_permute_rows<T>, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, tmp, nrhs,
                       nullptr, n, nrhs, B, ldb);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// ============================================================== residual / gemv
// r = v - Z x (complex128), one wave per (row, rhs).
__global__ __launch_bounds__(256) void k_residual_c128(const cpx<double>* __restrict__ Z, int64_t n,
                                                       int64_t ldz, const cpx<double>* __restrict__ x,
                                                       const cpx<double>* __restrict__ v,
                                                       cpx<double>* __restrict__ r, int64_t nrhs)
{
    const int lane = threadIdx.x & 63;
    const int64_t wid = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (wid >= n * nrhs) return;
    const int64_t i = wid / nrhs, c = wid % nrhs;
    const cpx<double>* Zr = Z + i * ldz;
    double sr = 0, si = 0;
    for (int64_t t = lane; t < n; t += 64) {
        cpx<double> a = Zr[t], y = x[t * nrhs + c];
        sr += a.re * y.re - a.im * y.im;
        si += a.re * y.im + a.im * y.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off);
        si += __shfl_xor(si, off);
    }
    if (lane == 0) {
        cpx<double> b = v[i * nrhs + c];
        store_cpx(r + i * nrhs + c, b.re - sr, b.im - si);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_gemv(const cpx<T>* __restrict__ A, int64_t n, int64_t lda,
                                              const cpx<T>* __restrict__ x, cpx<T>* __restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    if (i >= n) return;
    const cpx<T>* Ar = A + i * lda;
    T sr = 0, si = 0;
    for (int64_t t = lane; t < n; t += 64) {
        cpx<T> a = Ar[t], b = x[t];
        sr += a.re * b.re - a.im * b.im;
        si += a.re * b.im + a.im * b.re;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off);
        si += __shfl_xor(si, off);
    }
    if (lane == 0) store_cpx(y + i, sr, si);
}

__global__ void k_c128_to_c64(const cpx<double>* __restrict__ s, int64_t m, int64_t n, int64_t lds,
                              cpx<float>* __restrict__ d, int64_t ldd)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= m * n) return;
    int64_t i = id / n, j = id % n;
    cpx<double> v = s[i * lds + j];
    store_cpx(d + i * ldd + j, (float)v.re, (float)v.im);
}

// ---- batched refinement helpers ---------------------------------------------
struct ConvJobDev {
    const cpx<double>* src;
    cpx<float>* dst;
    int64_t count;
};

// grid (x: blocks per job, y: job); each block strides over its job's elements
__global__ void k_c128_to_c64_batched(const ConvJobDev* __restrict__ jobs)
{
    const ConvJobDev J = jobs[blockIdx.y];
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < J.count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const cpx<double> v = J.src[i];
        store_cpx(J.dst + i, (float)v.re, (float)v.im);
    }
}

struct RefJobDev {
    cpx<double>* j;
    int64_t ldj;
    const cpx<float>* d;
    int64_t ldd;
    int64_t n, nrhs;
};
constexpr int RF_SLABS = 8;    // row slabs per job (partial sums)
constexpr int RF_COLS = 64;    // columns per workgroup
constexpr int RF_THREADS = 256;

// grid (x: 64-column group, y: row slab, z: job).  J += D over the slab's rows
// and the partial sums |D_c|^2, |J_c|^2 (FP64) of its columns: thread (p, c)
// takes rows p, p + 4, ... of column c; the four row phases are added in a
// fixed order.  part: [job][slab][ncol_max] pairs.
__global__ __launch_bounds__(RF_THREADS) void k_refine_partial(const RefJobDev* __restrict__ jobs, int ncol_max,
                                                                double2* __restrict__ part)
{
    __shared__ double2 s_p[RF_THREADS / RF_COLS][RF_COLS];
    const RefJobDev J = jobs[blockIdx.z];
    const int tc = threadIdx.x % RF_COLS, tp = threadIdx.x / RF_COLS;
    const int64_t c = (int64_t)blockIdx.x * RF_COLS + tc;
    const int64_t rows = (J.n + RF_SLABS - 1) / RF_SLABS;
    const int64_t r0 = (int64_t)blockIdx.y * rows, r1 = (r0 + rows) < J.n ? (r0 + rows) : J.n;
    double dd = 0.0, jj = 0.0;
    if (c < J.nrhs) {
        for (int64_t r = r0 + tp; r < r1; r += RF_THREADS / RF_COLS) {
            const cpx<float> d = J.d[r * J.ldd + c];
            cpx<double> v = J.j[r * J.ldj + c];
            v.re += (double)d.re;
            v.im += (double)d.im;
            J.j[r * J.ldj + c] = v;
            dd += (double)d.re * d.re + (double)d.im * d.im;
            jj += v.re * v.re + v.im * v.im;
        }
    }
    s_p[tp][tc] = make_double2(dd, jj);
    __syncthreads();
    if (tp == 0 && c < J.nrhs) {
        double2 a = s_p[0][tc];
#pragma unroll
        for (int q = 1; q < RF_THREADS / RF_COLS; ++q) {
            a.x += s_p[q][tc].x;
            a.y += s_p[q][tc].y;
        }
        part[((int64_t)blockIdx.z * RF_SLABS + blockIdx.y) * ncol_max + c] = a;
    }
}

// one workgroup per job: slabs added in order per column, the ratio, the max
__global__ __launch_bounds__(RF_THREADS) void k_refine_ratio(const RefJobDev* __restrict__ jobs, int ncol_max,
                                                              const double2* __restrict__ part,
                                                              double* __restrict__ ratio_out)
{
    __shared__ double s_m[RF_THREADS];
    const RefJobDev J = jobs[blockIdx.x];
    double m = 0.0;
    for (int64_t c = threadIdx.x; c < J.nrhs; c += RF_THREADS) {
        double dd = 0.0, jj = 0.0;
        for (int s = 0; s < RF_SLABS; ++s) {
            const double2 a = part[((int64_t)blockIdx.x * RF_SLABS + s) * ncol_max + c];
            dd += a.x;
            jj += a.y;
        }
        const double r = jj > 0.0 ? sqrt(dd / jj) : (dd > 0.0 ? __builtin_inf() : 0.0);
        m = r > m ? r : m;
    }
    s_m[threadIdx.x] = m;
    __syncthreads();
    for (int w = RF_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s_m[threadIdx.x] = s_m[threadIdx.x + w] > s_m[threadIdx.x] ? s_m[threadIdx.x + w]
                                                                                        : s_m[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) ratio_out[blockIdx.x] = s_m[0];
}

// ============================================================== batched LU
// Several matrices factored in lock step: the same super-panel schedule as
// lu_factor, with each panel phase's panels of all matrices in shared
// launches of k_panel_batched (as many matrices per launch as fit one
// workgroup per CU); row interchanges, triangular solves and GEMMs are
// per-matrix launches.  The whole panel schedule (arguments and workgroup
// offsets) is built on the host and copied to the device once.
template <class T>
int lu_factor_batched(const mba_lu_matrix* mats, int nmat, void* scratch, size_t scratch_bytes, int32_t* err_out,
                      hipStream_t st)
{
    const int prec = sizeof(T) == 4 ? MBA_F32 : MBA_F64;
    const size_t esz = sizeof(cpx<T>);
    const int cus = device_cus();
    if (cus <= 0) return MBA_ERR_LAUNCH;
    const int NB = 128;
    const size_t lds_budget = 140 * 1024;
    const int rmax = (int)(lds_budget / (NB * esz));
    std::vector<LuWs> ws(nmat);
    int64_t nmax = 0;
    for (int i = 0; i < nmat; ++i) {
        ws[i] = lu_ws_layout(prec, mats[i].n, (char*)mats[i].workspace);
        if (mats[i].workspace_bytes < ws[i].total) return MBA_ERR_ALLOC;
        nmax = mats[i].n > nmax ? mats[i].n : nmax;
    }
    static std::once_flag attr_once;  // per element type (template instance)
    std::call_once(attr_once, [&] {
        (void)hipFuncSetAttribute((const void*)k_panel_batched<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds_budget);
        (void)hipFuncSetAttribute((const void*)k_trsm_batched<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)trsm_lds_bytes<T>(NB));
        (void)hipFuncSetAttribute((const void*)k_laswp_batched<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)(2 * NB * LASWP_COLS * esz));
    });
    // ---- host schedule: per panel phase, launches of (first entry, count, shm)
    struct Launch {
        size_t first, count, woff;
        int wg;
        size_t shm;
    };
    std::vector<PanelArgs<T>> args;
    std::vector<int> woffs;
    std::vector<std::vector<Launch>> phases;  // phase p = 2 s + half
    const int64_t steps = (nmax + 2 * NB - 1) / (2 * NB);
    for (int64_t s = 0; s < steps; ++s)
        for (int half = 0; half < 2; ++half) {
            std::vector<Launch> ls;
            Launch cur{args.size(), 0, woffs.size(), 0, 0};
            woffs.push_back(0);
            for (int i = 0; i < nmat; ++i) {
                const int64_t n = mats[i].n;
                const int64_t k = 2 * NB * s + (int64_t)half * NB;
                if (k >= n) continue;
                const int nb = (int)((n - k) < NB ? (n - k) : NB);
                const int64_t m = n - k;
                int G = (int)((m + rmax - 1) / rmax);
                if (G > MAX_G) return MBA_ERR_SIZE;
                const int R = (int)((m + G - 1) / G);
                G = (int)((m + R - 1) / R);
                if (G > cus) return MBA_ERR_SIZE;
                if (cur.wg + G > cus) {
                    ls.push_back(cur);
                    cur = Launch{args.size(), 0, woffs.size(), 0, 0};
                    woffs.push_back(0);
                }
                PanelArgs<T> P;
                P.A = (cpx<T>*)mats[i].a;
                P.lda = mats[i].lda;
                P.k = k;
                P.m = m;
                P.nb = nb;
                P.R = R;
                P.ipiv = mats[i].ipiv;
                P.info = mats[i].info;
                P.err = ws[i].err;
                P.meta = ws[i].meta;
                P.cval = ws[i].cval;
                P.cand = ws[i].cand;
                P.diag = ws[i].diag;
                args.push_back(P);
                cur.count += 1;
                cur.wg += G;
                woffs.push_back(cur.wg);
                const size_t shm = (size_t)R * nb * esz;
                cur.shm = shm > cur.shm ? shm : cur.shm;
            }
            if (cur.count > 0) ls.push_back(cur);
            else woffs.pop_back();
            phases.push_back(ls);
        }
    // ---- host schedule of the row swaps and triangular solves: per phase
    // (2 s + half) one batched launch over the matrices that have work
    struct JobRange {
        size_t first = 0, count = 0, woff = 0;
        int wg = 0;
        int nb_max = 0;
    };
    std::vector<SwapJob<T>> sjobs;
    std::vector<TrsmJob<T>> tjobs;
    std::vector<int> swoff, twoff;
    std::vector<JobRange> srange(2 * steps), trange(2 * steps);
    auto open_range = [](JobRange& r, size_t first, std::vector<int>& wo) {
        r.first = first;
        r.woff = wo.size();
        wo.push_back(0);
    };
    auto add_swap = [&](JobRange& r, int i, int64_t k, int nb) {
        const int64_t n = mats[i].n;
        const int64_t right = n - (k + nb);
        if (k == 0 && right <= 0) return;
        SwapJob<T> J;
        J.A = (cpx<T>*)mats[i].a;
        J.lda = mats[i].lda;
        J.c_hi = n;
        J.k = k;
        J.nb = nb;
        J.ipiv = mats[i].ipiv;
        J.slot_row = ws[i].slot_row;
        J.plan = ws[i].plan;
        J.plan_n = ws[i].plan_n;
        sjobs.push_back(J);
        r.count += 1;
        r.wg += (int)((k + LASWP_COLS - 1) / LASWP_COLS + (right + LASWP_COLS - 1) / LASWP_COLS);
        swoff.push_back(r.wg);
        r.nb_max = nb > r.nb_max ? nb : r.nb_max;
    };
    auto add_trsm = [&](JobRange& r, int i, int64_t k, int nb, int64_t c_lo) {
        const int64_t n = mats[i].n;
        if (c_lo >= n) return;
        TrsmJob<T> J;
        J.A = (cpx<T>*)mats[i].a;
        J.lda = mats[i].lda;
        J.k = k;
        J.nb = nb;
        J.c_lo = c_lo;
        J.c_hi = n;
        tjobs.push_back(J);
        r.count += 1;
        r.wg += (int)((n - c_lo + TrsmCols<T>::value - 1) / TrsmCols<T>::value);
        twoff.push_back(r.wg);
        r.nb_max = nb > r.nb_max ? nb : r.nb_max;
    };
    for (int64_t s = 0; s < steps; ++s) {
        const int64_t k = 2 * NB * s;
        for (int half = 0; half < 2; ++half) {
            JobRange& sr = srange[2 * s + half];
            JobRange& tr = trange[2 * s + half];
            open_range(sr, sjobs.size(), swoff);
            open_range(tr, tjobs.size(), twoff);
            for (int i = 0; i < nmat; ++i) {
                const int64_t n = mats[i].n;
                if (k >= n) continue;
                const int nb1 = (int)((n - k) < NB ? (n - k) : NB);
                const int64_t k2 = k + nb1;
                if (half == 0) {
                    add_swap(sr, i, k, nb1);
                    if (k2 < n) add_trsm(tr, i, k, nb1, k2);
                } else if (k2 < n) {
                    const int nb2 = (int)((n - k2) < NB ? (n - k2) : NB);
                    add_swap(sr, i, k2, nb2);
                    add_trsm(tr, i, k2, nb2, k2 + nb2);
                }
            }
        }
    }
    // trailing products per step: 0 = next panel's columns (after the first
    // panel), 1 = second panel's rows right of it, 2 = the trailing matrix
    struct GemmRange {
        size_t first = 0, count = 0, toff = 0;
        int64_t tiles = 0;
    };
    std::vector<GemmJobDev> gjobs;
    std::vector<int> gtoff;
    std::vector<GemmRange> grange(3 * steps);
    const int gv = g_gemm_f32_variant.load(std::memory_order_relaxed);
    const bool grouped = prec == MBA_F64 || gv == 4 || gv == 5;
    const bool gxcd = prec == MBA_F32 && gv == 5;
    if (grouped) {
        for (int64_t s = 0; s < steps; ++s) {
            const int64_t k = 2 * NB * s;
            for (int g = 0; g < 3; ++g) {
                GemmRange& r = grange[3 * s + g];
                r.first = gjobs.size();
                r.toff = gtoff.size();
                gtoff.push_back(0);
                for (int i = 0; i < nmat; ++i) {
                    const int64_t n = mats[i].n;
                    if (k >= n) continue;
                    const int nb1 = (int)((n - k) < NB ? (n - k) : NB);
                    const int64_t k2 = k + nb1;
                    if (k2 >= n) continue;
                    const int nb2 = (int)((n - k2) < NB ? (n - k2) : NB);
                    const int64_t c3 = k2 + nb2;
                    int64_t r0, M, c0, N, k0, kk;
                    if (g == 0) {
                        r0 = k2; M = n - k2; c0 = k2; N = nb2; k0 = k; kk = nb1;
                    } else if (g == 1) {
                        r0 = k2; M = nb2; c0 = c3; N = n - c3; k0 = k; kk = nb1;
                    } else {
                        r0 = c3; M = n - c3; c0 = c3; N = n - c3; k0 = k; kk = nb1 + nb2;
                    }
                    if (M <= 0 || N <= 0 || kk <= 0) continue;
                    cpx<T>* A = (cpx<T>*)mats[i].a;
                    const int64_t lda = mats[i].lda;
                    GemmJobDev J;
                    J.a = A + r0 * lda + k0;
                    J.b = A + k0 * lda + c0;
                    J.c = A + r0 * lda + c0;
                    J.m = M;
                    J.n = N;
                    J.k = kk;
                    J.lda = J.ldb = J.ldc = lda;
                    J.alpha = -1.0;
                    J.beta = 1.0;
                    gjobs.push_back(J);
                    r.count += 1;
                    r.tiles += grouped_tiles(prec, M, N);
                    if (r.tiles > 0x7fffffff) return MBA_ERR_SIZE;
                    gtoff.push_back((int)r.tiles);
                }
            }
        }
    }
    const size_t o_woff = al256(args.size() * sizeof(PanelArgs<T>));
    const size_t o_sj = o_woff + al256(woffs.size() * sizeof(int));
    const size_t o_swoff = o_sj + al256(sjobs.size() * sizeof(SwapJob<T>));
    const size_t o_tj = o_swoff + al256(swoff.size() * sizeof(int));
    const size_t o_twoff = o_tj + al256(tjobs.size() * sizeof(TrsmJob<T>));
    const size_t o_gj = o_twoff + al256(twoff.size() * sizeof(int));
    const size_t o_gtoff = o_gj + al256(gjobs.size() * sizeof(GemmJobDev));
    const size_t need = o_gtoff + gtoff.size() * sizeof(int);
    if (need > scratch_bytes) return MBA_ERR_ALLOC;
    char* base = (char*)scratch;
    PanelArgs<T>* d_args = (PanelArgs<T>*)base;
    int* d_woff = (int*)(base + o_woff);
    SwapJob<T>* d_sj = (SwapJob<T>*)(base + o_sj);
    int* d_swoff = (int*)(base + o_swoff);
    TrsmJob<T>* d_tj = (TrsmJob<T>*)(base + o_tj);
    int* d_twoff = (int*)(base + o_twoff);
    GemmJobDev* d_gj = (GemmJobDev*)(base + o_gj);
    int* d_gtoff = (int*)(base + o_gtoff);
    auto up = [&](void* dst, const void* src, size_t bytes) {
        return mba_host::upload_tables(dst, src, bytes, st) == 0;
    };
    if (!up(d_args, args.data(), args.size() * sizeof(PanelArgs<T>)) ||
        !up(d_woff, woffs.data(), woffs.size() * sizeof(int)) ||
        !up(d_sj, sjobs.data(), sjobs.size() * sizeof(SwapJob<T>)) ||
        !up(d_swoff, swoff.data(), swoff.size() * sizeof(int)) ||
        !up(d_tj, tjobs.data(), tjobs.size() * sizeof(TrsmJob<T>)) ||
        !up(d_twoff, twoff.data(), twoff.size() * sizeof(int)) ||
        !up(d_gj, gjobs.data(), gjobs.size() * sizeof(GemmJobDev)) ||
        !up(d_gtoff, gtoff.data(), gtoff.size() * sizeof(int)))
        return MBA_ERR_LAUNCH;
    for (int i = 0; i < nmat; ++i) {
        hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, (unsigned long long*)ws[i].err, 32);
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((ws[i].msg_words + 255) / 256)), dim3(256), 0, st,
                           (unsigned long long*)ws[i].meta, (int)ws[i].msg_words);
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, st, mats[i].info, 1);
    }
    MBA_LAUNCH_CHECK();
    auto gemm_rect = [&](int i, int64_t r0, int64_t M, int64_t c0, int64_t N, int64_t k0, int64_t kk) -> int {
        if (M <= 0 || N <= 0 || kk <= 0) return MBA_OK;
        cpx<T>* A = (cpx<T>*)mats[i].a;
        const int64_t lda = mats[i].lda;
        return gemm<T>(M, N, kk, -1.0, A + r0 * lda + k0, lda, A + k0 * lda + c0, lda, 1.0, A + r0 * lda + c0, lda,
                       st);
    };
    auto swaps = [&](int p) -> int {
        const JobRange& r = srange[p];
        if (r.count == 0) return MBA_OK;
        const SwapJob<T>* jj = d_sj + r.first;
        hipLaunchKernelGGL(k_swap_plan_batched<T>, dim3((unsigned)r.count), dim3(64), 0, st, jj);
        hipLaunchKernelGGL(k_laswp_batched<T>, dim3((unsigned)r.wg), dim3(256), (size_t)2 * r.nb_max * LASWP_COLS * esz,
                           st, jj, (const int*)(d_swoff + r.woff), (int)r.count);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    };
    auto trsm = [&](int p) -> int {
        const JobRange& r = trange[p];
        if (r.count == 0) return MBA_OK;
        hipLaunchKernelGGL(k_trsm_batched<T>, dim3((unsigned)r.wg), dim3(256), trsm_lds_bytes<T>(r.nb_max), st,
                           (const TrsmJob<T>*)(d_tj + r.first), (const int*)(d_twoff + r.woff), (int)r.count);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    };
    auto run_phase = [&](int p) -> int {
        for (const Launch& L : phases[p]) {
            const PanelArgs<T>* pa = d_args + L.first;
            const int* wo = d_woff + L.woff;
            int nm = (int)L.count;
            void* kargs[] = {(void*)&pa, (void*)&wo, (void*)&nm};
            const int lr = launch_resident((const void*)k_panel_batched<T>, (unsigned)L.wg, PANEL_THREADS, kargs,
                                           L.shm, st);
            if (lr != MBA_OK) return lr;
        }
        return MBA_OK;
    };
    // trailing products of step s, group g: one grouped launch, or one launch
    // per matrix for the tile variants without a grouped form
    auto products = [&](int64_t s, int g) -> int {
        if (grouped) {
            const GemmRange& r = grange[3 * s + g];
            return launch_grouped(prec, gxcd, d_gj + r.first, d_gtoff + r.toff, (int)r.count, r.tiles, st);
        }
        const int64_t k = 2 * NB * s;
        for (int i = 0; i < nmat; ++i) {
            const int64_t n = mats[i].n;
            if (k >= n) continue;
            const int nb1 = (int)((n - k) < NB ? (n - k) : NB);
            const int64_t k2 = k + nb1;
            if (k2 >= n) continue;
            const int nb2 = (int)((n - k2) < NB ? (n - k2) : NB);
            const int64_t c3 = k2 + nb2;
            int rc2 = MBA_OK;
            if (g == 0) rc2 = gemm_rect(i, k2, n - k2, k2, nb2, k, nb1);
            else if (g == 1) rc2 = gemm_rect(i, k2, nb2, c3, n - c3, k, nb1);
            else rc2 = gemm_rect(i, c3, n - c3, c3, n - c3, k, nb1 + nb2);
            if (rc2 != MBA_OK) return rc2;
        }
        return MBA_OK;
    };
    int rc;
    for (int64_t s = 0; s < steps; ++s) {
        if ((rc = run_phase((int)(2 * s))) != MBA_OK) return rc;
        if ((rc = swaps((int)(2 * s))) != MBA_OK) return rc;
        if ((rc = trsm((int)(2 * s))) != MBA_OK) return rc;
        if ((rc = products(s, 0)) != MBA_OK) return rc;
        if ((rc = run_phase((int)(2 * s + 1))) != MBA_OK) return rc;
        if ((rc = swaps((int)(2 * s + 1))) != MBA_OK) return rc;
        if ((rc = products(s, 1)) != MBA_OK) return rc;
        if ((rc = trsm((int)(2 * s + 1))) != MBA_OK) return rc;
        if ((rc = products(s, 2)) != MBA_OK) return rc;
    }
    for (int i = 0; i < nmat; ++i)
        hipLaunchKernelGGL(k_copy_i32, dim3(1), dim3(64), 0, st, (const int32_t*)ws[i].err, err_out + i, 1);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

// ============================================================== persistent batched LU
// Every factorisation of a batch (complex64) in ONE launch of resident
// workgroups (launch_resident): the
// workgroups stay resident and step through a host-built table of phases,
// each phase a set of independent workgroup tasks over all matrices, phases
// separated by a grid barrier (release / acquire at agent scope, so the
// stores of one phase are visible to every XCD in the next).
//   * Panels are PB_NB = 32 columns: the panel rows of every matrix then fit
//     the LDS of the resident workgroups together, so one panel phase serves
//     the whole batch with short rows per workgroup (the exchange protocol of
//     k_panel); the panel's first workgroup also writes the row-swap plan.
//   * Super panels of PB_P = 8 panels: a panel's columns are first updated by
//     the super panel's earlier panels (left-looking), after the panel the
//     row interchanges go to every other column, and its U rows over every
//     column to its right get the same update and the triangular solve; one
//     trailing product with k = 256 per matrix closes the super panel.
// Waits (panel exchange, grid barrier) are bounded: on expiry a flag is set,
// the kernel runs to completion and every matrix reports the timeout.
constexpr int PB_NB = 32, PB_P = 8, PB_THREADS = 256;
constexpr int PB_DST = 128;  // offset of the slot destinations in the plan words (2 nb <= 64 pairs before it)
constexpr int PB_WG_LDS = 80 * 1024;  // per workgroup: two workgroups share a CU

enum PbKind : int {
    PB_INIT = 0,
    PB_GEMM = 1,
    PB_PANEL = 2,
    PB_LASWP = 3,
    PB_TRSM = 4,
    PB_FINAL = 5,
    PB_SPLIT = 10, // trailing product, split FP16: operand splits
    PB_H3 = 11     // trailing product, split FP16: product tiles
};

// Split-FP16 trailing product of one matrix: C[c3:, c3:] -= A B with A =
// L[c3:, k:c3], B = U[k:c3, c3:], operand records in the matrix's split
// workspace.  Split tasks: Mpad / 4 row tasks, then Npad / SC column tasks.
struct PbSplitJob {
    const cpx<float>* A;
    const cpx<float>* B;
    cpx<float>* C;
    int64_t ld, M, N;
    int K, Kpad, nrec;
    int64_t Mpad, Npad;
    _Float16* Ap;
    _Float16* Bp;
    int32_t* ea;
    int32_t* eb;
};
constexpr int PB_H3_KM = 256;  // largest k of the split trailing product (32 x 8 panels)
constexpr int PB_H3_BM = 128;  // product tile rows (4 waves)

// One triangular-solve job of the persistent solves (complex64): the factor
// and its prepared solve tables (row permutation, inverted 64 x 64 diagonal
// blocks of L and U), right-hand sides b [n, ldb] (overwritten with the
// solution) and a work array [n, nrhs].
struct PbSolveJob {
    const cpx<float>* lu;
    int64_t n, lda;
    const int32_t* perm;
    const cpx<float>* linv;
    const cpx<float>* uinv;
    cpx<float>* b;
    int64_t ldb, nrhs;
    cpx<float>* work;
    uint32_t* flags;  // [2 passes][ngrp][nblk] block done (zero at launch)
    int nblk, ngrp;   // 64-row blocks, groups of PB_SCW right-hand sides
};
constexpr int PB_SCW = 64;  // right-hand sides per solve task

// One task of the dataflow solves: block i of job `job`, group g of
// right-hand sides, forward (pass 0) or backward (pass 1).
struct PbDfTask {
    int32_t job, i;
    int16_t pass, g;
};

struct PbPhase {
    int kind;
    int njob;      // jobs in this phase
    int nvb;       // workgroup tasks in this phase
    int sub;       // products: 0 panel columns, 1 U rows, 2 trailing (timing only)
    int64_t job0;  // first job in the kind's job table
    int64_t off0;  // first entry in the kind's offset table (njob + 1 entries)
};

struct PbPanelJob {
    PanelArgs<float> P;
    int32_t* slot_row;
    int32_t* plan;
    int32_t* plan_n;
    int64_t swaps;  // the row interchanges go to other columns (the plan is needed)
};

struct PbMat {
    uint64_t* zero[4];  // message words to clear: meta, cval, cand (used part), diag
    int64_t nzero[4];
    int* err;
    int32_t* info;
    int32_t* err_out;
};

struct PbTables {
    const PbPhase* phases;
    int nphase;
    int nmat;
    const PbMat* mats;
    const PbPanelJob* panels;
    const int* panel_off;
    const SwapJob<float>* swaps;
    const int* swap_off;
    const TrsmJob<float>* trsms;
    const int* trsm_off;
    const GemmJobDev* gemms;
    const int* gemm_off;
    const PbSplitJob* hjobs;  // split-FP16 trailing products
    const int* h_off;
    unsigned* barrier;  // barrier words (PB_BAR_WORDS, zero at launch)
    int* gerr;          // barrier timeout flag, zero at launch
    uint64_t* tstamp;   // optional [3 nphase + 1]: wall clock (100 MHz) of workgroup 0 at each phase
                        // start and at the end, then per phase the last workgroup's end of work
                        // (before its barrier) and the last workgroup's release by that barrier
};

// Row interchanges of one panel (nb <= PB_NB) on 128 columns outside the
// panel: thread (h, c) holds column c's values of slots [32 h, 32 h + 32) in
// registers; all loads precede all stores (one barrier), and slot s goes to
// row slot_row[dst[s]] (dst composed once by the panel's first workgroup).
constexpr int PB_SWAP_COLS = 128;

__device__ __forceinline__ void pb_laswp(cpx<float>* __restrict__ A, int64_t lda, int64_t n, int64_t kp, int nb,
                                         const int32_t* __restrict__ slot_row, const int32_t* __restrict__ dst,
                                         const int32_t* __restrict__ plan_n, int64_t blk)
{
    const int tid = threadIdx.x, h = tid / PB_SWAP_COLS, c = tid % PB_SWAP_COLS;
    const int ns = *plan_n;
    int64_t col = blk * PB_SWAP_COLS + c;  // columns [0, n - nb) with the panel's nb columns skipped
    if (col >= kp) col += nb;
    const bool ok = col < n;
    cpx<float> v[32];
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int sl = 32 * h + q;
        v[q] = (ok && sl < ns) ? A[(int64_t)slot_row[sl] * lda + col] : cpx<float>{0.f, 0.f};
    }
    __builtin_amdgcn_s_waitcnt(0);  // every load of the block has returned before any store
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 32; ++q) {
        const int sl = 32 * h + q;
        if (ok && sl < ns) A[(int64_t)slot_row[dst[sl]] * lda + col] = v[q];
    }
}

// X = L11^{-1} X for rows [kp, kp + PB_NB) of a full panel and 256 columns
// from c_lo + 256 blk: one column per thread in registers, L11 in LDS.  No
// per-row conditions: the loads read a clamped column and the stores are one
// branch per thread (per-row selects turn into branches around every load and
// the register allocator then spills the column).  Partial panels take
// trsm_body.
__device__ __forceinline__ void pb_trsm_full(cpx<float>* __restrict__ A, int64_t lda, int64_t kp, int64_t c_lo,
                                             int64_t c_hi, int64_t blk, cpx<float>* Ls /* LDS [PB_NB][PB_NB] */)
{
    const int tid = threadIdx.x;
    for (int idx = tid; idx < PB_NB * PB_NB; idx += PB_THREADS) Ls[idx] = A[(kp + idx / PB_NB) * lda + kp + idx % PB_NB];
    const int64_t col = c_lo + blk * PB_THREADS + tid;
    const bool ok = col < c_hi;
    cpx<float>* q = A + kp * lda + (ok ? col : c_lo);
    cpx<float> x[PB_NB];
#pragma unroll
    for (int r = 0; r < PB_NB; ++r) x[r] = q[(int64_t)r * lda];
    __syncthreads();
#pragma unroll
    for (int t = 0; t < PB_NB - 1; ++t) {
        const cpx<float> xt = x[t];
#pragma unroll
        for (int i = t + 1; i < PB_NB; ++i) {
            const cpx<float> l = Ls[i * PB_NB + t];
            x[i].re -= l.re * xt.re - l.im * xt.im;
            x[i].im -= l.re * xt.im + l.im * xt.re;
        }
    }
    if (ok) {
#pragma unroll
        for (int r = 1; r < PB_NB; ++r) q[(int64_t)r * lda] = x[r];
    }
}

// After swap_plan_body: dst[s] = the slot whose row slot s's values go to
// (the panel's interchanges applied in order to slot labels).  One thread.
__device__ __forceinline__ void pb_swap_dst(int nb, const int32_t* __restrict__ plan, int ns, int32_t* __restrict__ dst)
{
    int32_t cur[2 * PB_NB];
    for (int q = 0; q < ns; ++q) cur[q] = q;
    for (int j = 0; j < nb; ++j) {
        const int a = plan[2 * j], b = plan[2 * j + 1];
        const int32_t t = cur[a];
        cur[a] = cur[b];
        cur[b] = t;
    }
    for (int q = 0; q < ns; ++q) dst[cur[q]] = q;  // place q holds original slot cur[q]
}

// Grid barrier number `epoch` (1, 2, ...): every workgroup adds one to the
// arrival count and waits until it reaches epoch x (workgroups).  The count
// only grows, so no reset races with early arrivals at the next barrier.
// Grid barriers.  The L2 of each XCD is not coherent with the others for
// ordinary device memory, so a phase's stores reach another XCD only through
// an L2 write-back on the writer's XCD and an L2 invalidate on the reader's.
// Both are whole-cache operations: one per XCD per barrier, not one per
// workgroup (64 of them per XCD and barrier measured ~100 us per phase).
//   arrive: every wave drains its stores (they are in its XCD's L2); the last
//     arriver of the XCD writes that L2 back and adds the XCD to the global count;
//   leave: that workgroup waits for every XCD, invalidates its L2 (and its
//     CU's L1) and raises the XCD's flag; the others wait for the flag and
//     invalidate their CU's L1.
// Counters only grow (barrier e waits for e x members), so no reset races
// with early arrivals at the next barrier.  Each word sits on its own 128-byte
// line; agent-scope atomics are performed beyond the L2s.  Waits are bounded:
// after a timeout (gerr set) barriers stop waiting and the remaining phases
// are skipped, so a failed launch ends after one bounded wait.
constexpr int PB_LINE = 32;                           // words per 128-byte line
constexpr int PB_BAR_GLOBAL = 0;                      // XCDs arrived (all barriers)
constexpr int PB_BAR_START = 1 * PB_LINE;             // workgroups registered (start barrier)
constexpr int PB_BAR_MEMBERS = 2 * PB_LINE;           // [8] workgroups per XCD
constexpr int PB_BAR_ARRIVE = 10 * PB_LINE;           // [8] arrivals per XCD
constexpr int PB_BAR_DONE = 18 * PB_LINE;             // [8] barrier number whose L2 invalidate is done
constexpr int PB_BAR_TICKET = 26 * PB_LINE;           // [8] next task ticket of each XCD queue (solves)
constexpr int PB_BAR_FIN = 34 * PB_LINE;              // workgroups done with the solve tasks
constexpr int PB_BAR_WORDS = 35 * PB_LINE;
constexpr int PB_MAX_XCD = 8;

__device__ __forceinline__ unsigned pb_ld(unsigned* p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool pb_wait_ge(unsigned* p, unsigned target, int* gerr)
{
    long spins = 0;
    while (pb_ld(p) < target) {
        if ((spins & 63) == 63 && __hip_atomic_load(gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return false;
        if (++spins > SPIN_LIMIT) {
            __hip_atomic_fetch_or(gerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    return true;
}

// XCD of the calling workgroup (HW_REG_XCC_ID, bits [3:0])
__device__ __forceinline__ int pb_xcc() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15; }

struct PbBarrier {
    int xcc;           // this workgroup's XCD
    unsigned members;  // workgroups on it
    unsigned nxcd;     // XCDs with workgroups
    unsigned epoch;    // barriers passed
};

// Start: every workgroup registers on its XCD; one barrier with a full fence
// per workgroup (once per launch) makes the counts visible.
__device__ __forceinline__ void pb_barrier_init(PbBarrier& B, unsigned* bar, unsigned nwg, int* gerr)
{
    __shared__ unsigned s_members, s_nxcd;
    __shared__ int s_xcc;
    if (threadIdx.x == 0) {
        int x = pb_xcc();
        if (x >= PB_MAX_XCD) {
            __hip_atomic_fetch_or(gerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            x = 0;
        }
        __hip_atomic_fetch_add(bar + PB_BAR_MEMBERS + x * PB_LINE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __threadfence();
        __hip_atomic_fetch_add(bar + PB_BAR_START, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pb_wait_ge(bar + PB_BAR_START, nwg, gerr);
        __threadfence();
        unsigned nx = 0;
        for (int q = 0; q < PB_MAX_XCD; ++q) nx += pb_ld(bar + PB_BAR_MEMBERS + q * PB_LINE) > 0 ? 1u : 0u;
        s_xcc = x;
        s_members = pb_ld(bar + PB_BAR_MEMBERS + x * PB_LINE);
        s_nxcd = nx;
    }
    __syncthreads();
    B.xcc = s_xcc;
    B.members = s_members;
    B.nxcd = s_nxcd;
    B.epoch = 0;
}

__device__ __forceinline__ void pb_grid_sync(PbBarrier& B, unsigned* bar, int* gerr)
{
    const unsigned e = ++B.epoch;
    __builtin_amdgcn_s_waitcnt(0);  // this wave's stores of the phase are in the L2
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned* arrive = bar + PB_BAR_ARRIVE + B.xcc * PB_LINE;
        unsigned* done = bar + PB_BAR_DONE + B.xcc * PB_LINE;
        const unsigned a =
            __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (a == e * B.members) {
            // last of this XCD: write its L2 back, then count the XCD
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(bar + PB_BAR_GLOBAL, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pb_wait_ge(bar + PB_BAR_GLOBAL, e * B.nxcd, gerr);
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // this XCD's L2 (and this CU's L1)
            __hip_atomic_store(done, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            pb_wait_ge(done, e, gerr);
            // this CU's L1 (the agent-scope invalidate: the workgroup-scope
            // one leaves the L1 alone when a workgroup's waves share a CU)
            asm volatile("buffer_inv sc1" ::: "memory");
        }
    }
    __syncthreads();
}

// Phase bodies, each its own function so that every one gets its own
// register allocation (inlined into one switch the largest body's pressure
// spread to all of them).
__device__ __noinline__ void pb_phase_panel(const PbTables& T, const PbPhase& ph, int b)
{
    if (b >= ph.nvb) return;
    const int* off = T.panel_off + ph.off0;
    const int j = job_of(off, ph.njob, b);
    const PbPanelJob& J = T.panels[ph.job0 + j];
    const int g = b - off[j];
    panel_body<float, PB_THREADS>(J.P, off[j + 1] - off[j], g);
    if (g == 0 && J.swaps) {
        // ipiv of this panel was written by this workgroup's thread 0: the
        // slot plan, then the slots' destinations (at plan + PB_DST, after
        // the 2 nb pairs)
        __syncthreads();
        if (threadIdx.x == 0) {
            swap_plan_body(J.P.ipiv, J.P.k, J.P.nb, J.slot_row, J.plan, J.plan_n);
            pb_swap_dst(J.P.nb, J.plan, *J.plan_n, J.plan + PB_DST);
        }
    }
}

__device__ __noinline__ void pb_phase_laswp(const PbTables& T, const PbPhase& ph, int b, int nwg)
{
    const int* off = T.swap_off + ph.off0;
    for (int vb = b; vb < ph.nvb; vb += nwg) {
        const int j = job_of(off, ph.njob, vb);
        const SwapJob<float>& J = T.swaps[ph.job0 + j];
        pb_laswp(J.A, J.lda, J.c_hi, J.k, J.nb, J.slot_row, J.plan + PB_DST, J.plan_n, (int64_t)(vb - off[j]));
        __syncthreads();
    }
}

__device__ __noinline__ void pb_phase_trsm(const PbTables& T, const PbPhase& ph, int b, int nwg, cpx<float>* lds)
{
    const int* off = T.trsm_off + ph.off0;
    for (int vb = b; vb < ph.nvb; vb += nwg) {
        const int j = job_of(off, ph.njob, vb);
        const TrsmJob<float>& J = T.trsms[ph.job0 + j];
        if (J.nb == PB_NB)
            pb_trsm_full(J.A, J.lda, J.k, J.c_lo, J.c_hi, (int64_t)(vb - off[j]), lds);
        else
            trsm_body<float>(J.A, J.lda, J.k, J.nb, J.c_lo, J.c_hi, (int64_t)(vb - off[j]), lds,
                             lds + (size_t)J.nb * TrsmCols<float>::value);
        __syncthreads();
    }
}

// tasks dealt so that each XCD (workgroup b runs on XCD b % 8; the grid is a
// multiple of 8) walks a contiguous run of tiles
__device__ __noinline__ void pb_phase_gemm(const PbTables& T, const PbPhase& ph, int b, int nwg, float* lds)
{
    const int* off = T.gemm_off + ph.off0;
    const int per = (ph.nvb + NUM_XCD - 1) / NUM_XCD;
    const int span = per * NUM_XCD;
    for (int vb = b; vb < span; vb += nwg) {
        const int t = (vb % NUM_XCD) * per + vb / NUM_XCD;
        if (t < ph.nvb) {
            const int j = job_of(off, ph.njob, t);
            const GemmJobDev& J = T.gemms[ph.job0 + j];
            const int64_t nbn = (J.n + G3M_BN - 1) / G3M_BN;
            const int64_t u = t - off[j];
            cgemm3m_tile(J.m, J.n, J.k, (float)J.alpha, (const cpx<float>*)J.a, J.lda, (const cpx<float>*)J.b, J.ldb,
                         (float)J.beta, (cpx<float>*)J.c, J.ldc, (u / nbn) * G32_BM, (u % nbn) * G3M_BN, lds);
        }
        __syncthreads();
    }
}

__device__ __noinline__ void pb_phase_init(const PbTables& T, int b, int nwg)
{
    const int64_t gt = (int64_t)b * PB_THREADS + threadIdx.x, gs = (int64_t)nwg * PB_THREADS;
    for (int i = 0; i < T.nmat; ++i) {
        const PbMat& M = T.mats[i];
        for (int z = 0; z < 4; ++z)
            for (int64_t w = gt; w < M.nzero[z]; w += gs) M.zero[z][w] = 0ull;
        if (gt == 0) {
            *M.err = 0;
            *M.info = 0;
        }
    }
}

__device__ __noinline__ void pb_phase_split(const PbTables& T, const PbPhase& ph, int b, int nwg, void* lds)
{
    const int* off = T.h_off + ph.off0;
    for (int vb = b; vb < ph.nvb; vb += nwg) {
        const int j = job_of(off, ph.njob, vb);
        const PbSplitJob& J = T.hjobs[ph.job0 + j];
        const int64_t u = vb - off[j], nrt = J.Mpad / 4;
        if (u < nrt)
            mba_h3::dev::split_rows_task<PB_H3_KM>(J.A, J.ld, J.M, J.K, J.Kpad, J.Mpad, J.Ap, J.ea, u);
        else
            mba_h3::dev::split_cols_task<PB_H3_KM>(J.B, J.ld, J.K, J.Kpad, J.N, J.Npad, J.Bp, J.eb, u - nrt,
                                                   *(mba_h3::dev::SplitColsLds<PB_H3_KM>*)lds);
        __syncthreads();
    }
}

__device__ __noinline__ void pb_phase_h3(const PbTables& T, const PbPhase& ph, int b, int nwg, void* lds)
{
    // tiles dealt so that each XCD walks a contiguous run (as pb_phase_gemm)
    const int* off = T.h_off + ph.off0;
    const int per = (ph.nvb + NUM_XCD - 1) / NUM_XCD;
    const int span = per * NUM_XCD;
    for (int vb = b; vb < span; vb += nwg) {
        const int t = (vb % NUM_XCD) * per + vb / NUM_XCD;
        if (t < ph.nvb) {
            const int j = job_of(off, ph.njob, t);
            const PbSplitJob& J = T.hjobs[ph.job0 + j];
            const int64_t nbn = (J.N + mba_h3::dev::BN - 1) / mba_h3::dev::BN;
            const int64_t u = t - off[j];
            static_assert(sizeof(mba_h3::dev::TileLdsDma<PB_H3_BM>) <= sizeof(mba_h3::dev::TileLds<PB_H3_BM>),
                          "the product's LDS fits the region sized for it");
            mba_h3::dev::tile_h3_dma<PB_H3_BM, PB_THREADS>(J.M, J.N, J.nrec, J.Ap, J.ea, J.Bp, J.eb, -1.f, 1.f, J.C,
                                                           J.ld, u / nbn, u % nbn,
                                                           *(mba_h3::dev::TileLdsDma<PB_H3_BM>*)lds);
        }
        __syncthreads();
    }
}

// ---- persistent solves, dataflow inside each XCD.  Every job (a prepared
// factorisation with its right-hand sides) is assigned to one XCD, whose
// workgroups take its tasks from that XCD's ticket counter.  A task (job, i,
// g, pass) streams its block row: forward, t_i = (P b)_i - sum_{j<i} L_ij x_j
// over the blocks x_j (in `work`) as their flags come up, then x_i =
// inv(L_ii) t_i in place and its flag; backward (after the job's last forward
// block), z_i = x_i - sum_{j>i} U_ij z_j over the finished blocks (in b), then
// z_i = inv(U_ii) (.).  Tickets are dealt in an order where every task waits
// only on tasks with lower tickets (forward by ascending block, then backward
// by descending block), taken by running workgroups, so every wait ends.  All
// data a job's tasks exchange stays in one XCD's L2: a producer drains its
// stores before raising the flag, a consumer reads the blocks of other tasks
// with agent-scope loads, which miss its CU's L1 (a line of a block can sit
// in that L1 from before the block was final: the rows around a block
// boundary share lines).  Invalidating the L1 instead needs the agent-scope
// `buffer_inv sc1` (the workgroup-scope `sc0` leaves the L1 alone unless a
// workgroup's waves are split over CUs), which also drops L2 lines.

// Wait until flag j is up, then return the end of the run of raised flags
// j, j + step, ... (at most 64, stopping before `end`); j when the wait
// failed.  Wave 0 polls 64 flags per load; every thread's earlier stores are
// drained first, so the caller's own stores are visible to all its threads.
__device__ __forceinline__ int pb_flag_run(const uint32_t* f, int j, int end, int step, int* gerr)
{
    __shared__ int s_k;
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (threadIdx.x < 64) {
        const int q = j + step * (int)threadIdx.x;
        const bool inr = step > 0 ? q < end : q > end;
        long spins = 0;
        int k = j;
        while (true) {
            const uint32_t v = inr ? __hip_atomic_load(f + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0u;
            const uint64_t m = __ballot(v != 0u);
            if (m & 1ull) {
                k = j + step * (~m == 0ull ? 64 : __builtin_ctzll(~m));
                break;
            }
            if ((spins & 63) == 63 && __hip_atomic_load(gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
                break;
            if (++spins > SPIN_LIMIT) {
                if (threadIdx.x == 0) __hip_atomic_fetch_or(gerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (threadIdx.x == 0) s_k = k;
    }
    __syncthreads();
    return s_k;
}

// this workgroup's global writes are complete and visible to all its threads
// (one CU: its L1 sees its own stores)
__device__ __forceinline__ void pb_local_fence()
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
}

__device__ __forceinline__ void pb_flag_raise(uint32_t* f)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the block's stores are in the XCD's L2
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(f, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// C = alpha A B + beta C for one 64 x 64 tile (M, N <= 64, any K; B written
// by this workgroup or before the launch), 3M product on FP32 MFMA: four
// waves, each a 32 x 32 quarter.  Not inlined: its registers are its own.
constexpr int DF_LDS_FLOATS = 8 * G32_BK * 64;
__device__ __noinline__ void pb_df_product(int64_t M, int64_t N, int64_t K, float alpha, const cpx<float>* A,
                                           int64_t lda, const cpx<float>* B, int64_t ldb, float beta, cpx<float>* C,
                                           int64_t ldc, float* lds)
{
    float(*Ar)[G32_BK][64] = (float(*)[G32_BK][64])lds;
    float(*Ai)[G32_BK][64] = (float(*)[G32_BK][64])(lds + 2 * G32_BK * 64);
    float(*Br)[G32_BK][64] = (float(*)[G32_BK][64])(lds + 4 * G32_BK * 64);
    float(*Bi)[G32_BK][64] = (float(*)[G32_BK][64])(lds + 6 * G32_BK * 64);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    // A: 64 rows x 16 k: thread -> row tid/4, k (tid&3)*4 .. +4
    const int ar = tid >> 2, ak = (tid & 3) * 4;
    // B: 16 k x 64 columns: thread -> k tid/16, columns (tid&15)*4 .. +4
    const int bkr = tid >> 4, bc = (tid & 15) * 4;
    float2 ra[4], rb[4];
    auto load_chunk = [&](int64_t k0) {
        const bool rowok = ar < M;
        const float2* pa = (const float2*)(A + (rowok ? ar : 0) * lda);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int64_t gk = k0 + ak + q;
            ra[q] = (rowok && gk < K) ? pa[gk] : make_float2(0.f, 0.f);
        }
        const int64_t gk = k0 + bkr;
        const bool kok = gk < K;
        const cpx<float>* pb = B + (kok ? gk : 0) * ldb;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int gj = bc + q;
            rb[q] = (kok && gj < N) ? *(const float2*)(pb + gj) : make_float2(0.f, 0.f);
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Ar[buf][ak + q][ar] = ra[q].x;
            Ai[buf][ak + q][ar] = ra[q].y;
        }
        *(float4*)&Br[buf][bkr][bc] = make_float4(rb[0].x, rb[1].x, rb[2].x, rb[3].x);
        *(float4*)&Bi[buf][bkr][bc] = make_float4(rb[0].y, rb[1].y, rb[2].y, rb[3].y);
    };
    floatx16 t1, t2, t3;
#pragma unroll
    for (int r = 0; r < 16; ++r) t1[r] = t2[r] = t3[r] = 0.f;
    const int hi = lane >> 5, jj = lane & 31;
    const int nchunks = (int)((K + G32_BK - 1) / G32_BK);
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) load_chunk((int64_t)(c + 1) * G32_BK);
#pragma unroll
        for (int kk = 0; kk < G32_BK; kk += 2) {
            const int k = kk + hi;
            const float br = Br[buf][k][wn * 32 + jj], bi = Bi[buf][k][wn * 32 + jj];
            const float a_r = Ar[buf][k][wm * 32 + jj], a_i = Ai[buf][k][wm * 32 + jj];
            t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r, br, t1, 0, 0, 0);
            t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_i, bi, t2, 0, 0, 0);
            t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r + a_i, br + bi, t3, 0, 0, 0);
        }
        if (c + 1 < nchunks) {
            __syncthreads();
            store_chunk(buf ^ 1);
            __syncthreads();
        }
    }
    // epilogue: register r of lane -> row 8*(r/4) + 4*hi + r%4, column jj
    const int gj = wn * 32 + jj;
    if (gj < N) {
        float2 cv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int gi = wm * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            cv[r] = (beta != 0.f && gi < M) ? *(const float2*)(C + gi * ldc + gj) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int gi = wm * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            if (gi < M) {
                float re = alpha * (t1[r] - t2[r]);
                float im = alpha * (t3[r] - t1[r] - t2[r]);
                if (beta != 0.f) {
                    re += beta * cv[r].x;
                    im += beta * cv[r].y;
                }
                *(float2*)(C + gi * ldc + gj) = make_float2(re, im);
            }
        }
    }
}

// X -= A_i,j X_j over a task's finished blocks, streamed: one product whose
// k range runs over the blocks j = j0, j0 + step, ... (nb of them; A_i,j at
// A + r_j, block j of the source at src + r_j ld_src), each block's loads
// issued once its flag is up.  The k order is the same whatever the timing,
// so the result is reproducible bit for bit.  The source blocks were written
// by other CUs during the launch: agent-scope loads, past this CU's L1.
// Tile 64 x 64 (M, N <= 64), 3M on FP32 MFMA, four waves of 32 x 32.
// Returns false when a wait failed (nothing written).
__device__ __noinline__ bool pb_df_stream(const cpx<float>* A, int64_t lda, const cpx<float>* src, int64_t lds_,
                                          cpx<float>* X, int64_t ldx, int64_t M, int64_t N, int nb, int j0,
                                          int step, int64_t n, const uint32_t* fl, int* gerr, float* lds)
{
    float(*Ar)[G32_BK][64] = (float(*)[G32_BK][64])lds;
    float(*Ai)[G32_BK][64] = (float(*)[G32_BK][64])(lds + 2 * G32_BK * 64);
    float(*Br)[G32_BK][64] = (float(*)[G32_BK][64])(lds + 4 * G32_BK * 64);
    float(*Bi)[G32_BK][64] = (float(*)[G32_BK][64])(lds + 6 * G32_BK * 64);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int ar = tid >> 2, ak = (tid & 3) * 4;
    const int bkr = tid >> 4, bc = (tid & 15) * 4;
    constexpr int CPB = TS_BS / G32_BK;  // chunks per block
    float2 ra[4], rb[4];
    int ready = 0;  // blocks of the sequence whose flags were seen
    // chunk q: block s = q / CPB of the sequence, rows [k0, k0 + 16) of it
    auto load_chunk = [&](int q) {
        const int j = j0 + step * (q / CPB);
        const int64_t r_j = (int64_t)j * TS_BS, kend = (r_j + TS_BS) < n ? (r_j + TS_BS) : n;
        const int64_t k0 = r_j + (int64_t)(q % CPB) * G32_BK;
        const bool rowok = ar < M;
        const float2* pa = (const float2*)(A + (rowok ? ar : 0) * lda);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t gk = k0 + ak + u;
            ra[u] = (rowok && gk < kend) ? pa[gk] : make_float2(0.f, 0.f);
        }
        const int64_t gk = k0 + bkr;
        const bool kok = gk < kend;
        const cpx<float>* pb = src + (kok ? gk : 0) * lds_;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int gj = bc + u;
            if (kok && gj < N) {
                const uint64_t w =
                    __hip_atomic_load((const uint64_t*)(pb + gj), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                rb[u] = make_float2(__uint_as_float((uint32_t)w), __uint_as_float((uint32_t)(w >> 32)));
            } else {
                rb[u] = make_float2(0.f, 0.f);
            }
        }
    };
    // before the first chunk of block s: wait for its flag (and learn how
    // many after it are up already); uniform over the workgroup
    auto ensure = [&](int q) -> bool {
        if (q % CPB != 0 || q / CPB < ready) return true;
        const int s = q / CPB, j = j0 + step * s;
        const int end = j0 + step * nb;
        const int k = pb_flag_run(fl, j, end, step, gerr);
        if (k == j) return false;
        ready = s + (k - j) * step;
        return true;
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            Ar[buf][ak + u][ar] = ra[u].x;
            Ai[buf][ak + u][ar] = ra[u].y;
        }
        *(float4*)&Br[buf][bkr][bc] = make_float4(rb[0].x, rb[1].x, rb[2].x, rb[3].x);
        *(float4*)&Bi[buf][bkr][bc] = make_float4(rb[0].y, rb[1].y, rb[2].y, rb[3].y);
    };
    floatx16 t1, t2, t3;
#pragma unroll
    for (int r = 0; r < 16; ++r) t1[r] = t2[r] = t3[r] = 0.f;
    const int hi = lane >> 5, jj = lane & 31;
    const int nchunks = nb * CPB;
    if (!ensure(0)) return false;
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    for (int c = 0; c < nchunks; ++c) {
        const int buf = c & 1;
        if (c + 1 < nchunks) {
            if (!ensure(c + 1)) return false;
            load_chunk(c + 1);
        }
#pragma unroll
        for (int kk = 0; kk < G32_BK; kk += 2) {
            const int k = kk + hi;
            const float br = Br[buf][k][wn * 32 + jj], bi = Bi[buf][k][wn * 32 + jj];
            const float a_r = Ar[buf][k][wm * 32 + jj], a_i = Ai[buf][k][wm * 32 + jj];
            t1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r, br, t1, 0, 0, 0);
            t2 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_i, bi, t2, 0, 0, 0);
            t3 = __builtin_amdgcn_mfma_f32_32x32x2f32(a_r + a_i, br + bi, t3, 0, 0, 0);
        }
        if (c + 1 < nchunks) {
            __syncthreads();
            store_chunk(buf ^ 1);
            __syncthreads();
        }
    }
    // X -= product; register r of lane -> row 8*(r/4) + 4*hi + r%4, column jj
    const int gj = wn * 32 + jj;
    if (gj < N) {
        float2 cv[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int gi = wm * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            cv[r] = gi < M ? *(const float2*)(X + gi * ldx + gj) : make_float2(0.f, 0.f);
        }
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int gi = wm * 32 + 8 * (r >> 2) + 4 * hi + (r & 3);
            if (gi < M)
                *(float2*)(X + gi * ldx + gj) =
                    make_float2(cv[r].x - (t1[r] - t2[r]), cv[r].y - (t3[r] - t1[r] - t2[r]));
        }
    }
    return true;
}

__device__ __forceinline__ void pb_df_task(const PbSolveJob& J, const PbDfTask& t, float* lds, int* gerr)
{
    const int i = t.i, g = t.g;
    const bool fwd = t.pass == 0;
    const int64_t r_i = (int64_t)i * TS_BS, bs_i = (J.n - r_i) < TS_BS ? (J.n - r_i) : TS_BS;
    const int64_t c0 = (int64_t)g * PB_SCW, cw = (J.nrhs - c0) < PB_SCW ? (J.nrhs - c0) : PB_SCW;
    const uint32_t* ffwd = J.flags + (int64_t)g * J.nblk;
    const uint32_t* fbwd = J.flags + ((int64_t)J.ngrp + g) * J.nblk;
    cpx<float>* Wi = J.work + r_i * J.nrhs + c0;  // forward output x_i (ld nrhs)
    cpx<float>* Bi = J.b + r_i * J.ldb + c0;      // backward output z_i (ld ldb)
    // this pass: output block, finished blocks (base, ld), their flags, diagonal inverse
    cpx<float>* X = fwd ? Wi : Bi;
    const int64_t ldx = fwd ? J.nrhs : J.ldb;
    const cpx<float>* src = fwd ? J.work : J.b;
    const uint32_t* fl = fwd ? ffwd : fbwd;
    const cpx<float>* inv = (fwd ? J.linv : J.uinv) + (int64_t)i * TS_BS * TS_BS;
    if (fwd) {
        for (int64_t e = threadIdx.x; e < bs_i * cw; e += PB_THREADS) {
            const int64_t r = e / cw, c = e % cw;
            Wi[r * J.nrhs + c] = J.b[(int64_t)J.perm[r_i + r] * J.ldb + c0 + c];
        }
    } else {
        // the job's forward pass is done once its last block is
        if (pb_flag_run(ffwd, J.nblk - 1, J.nblk, 1, gerr) == J.nblk - 1) return;
        // x_i was written on another CU: agent-scope loads (past this CU's L1)
        for (int64_t e = threadIdx.x; e < bs_i * cw; e += PB_THREADS) {
            const int64_t r = e / cw, c = e % cw;
            const uint64_t w = __hip_atomic_load((const uint64_t*)(Wi + r * J.nrhs + c), __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
            Bi[r * J.ldb + c] = cpx<float>{__uint_as_float((uint32_t)w), __uint_as_float((uint32_t)(w >> 32))};
        }
    }
    const int nb = fwd ? i : J.nblk - 1 - i;
    if (nb > 0 && !pb_df_stream(J.lu + r_i * J.lda, J.lda, src + c0, ldx, X, ldx, bs_i, cw, nb, fwd ? 0 : J.nblk - 1,
                                fwd ? 1 : -1, J.n, fl, gerr, lds))
        return;
    pb_local_fence();
    // in place: the single tile reads all of X before its epilogue writes it
    pb_df_product(bs_i, cw, bs_i, 1.f, inv, TS_BS, X, ldx, 0.f, X, ldx, lds);
    pb_flag_raise((uint32_t*)fl + i);
}

// Tables of the dataflow solves (one launch).
struct PbDfTables {
    const PbSolveJob* sjobs;
    const PbDfTask* dtasks;  // solve tasks, queue by queue in ticket order
    const int* d_off;        // [PB_MAX_XCD + 1] each queue's range of dtasks
    int32_t* err_out;        // written by the last workgroup: nonzero after a timed-out wait
    unsigned* barrier;       // barrier words (PB_BAR_WORDS, zero at launch): registration, tickets
    int* gerr;               // zero at launch
};

__device__ __forceinline__ void pb_df_serve(const PbDfTables& T, const PbBarrier& B, unsigned* bar, float* lds)
{
    __shared__ int s_t;
    // queue x is served by XCD x; a queue of an XCD without workgroups by one
    // that has some (all of a queue on one XCD: its flags and blocks stay in
    // that XCD's L2)
    int present[PB_MAX_XCD], np = 0;
    for (int x = 0; x < PB_MAX_XCD; ++x)
        if (pb_ld(bar + PB_BAR_MEMBERS + x * PB_LINE) > 0) present[np++] = x;
    __shared__ int s_fail;  // one reading for the whole workgroup (uniform loop bounds)
    if (threadIdx.x == 0) s_fail = __hip_atomic_load(T.gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const bool failed = s_fail != 0;
    for (int x = 0; x < PB_MAX_XCD && !failed && np > 0; ++x) {
        const int owner = pb_ld(bar + PB_BAR_MEMBERS + x * PB_LINE) > 0 ? x : present[x % np];
        if (owner != B.xcc) continue;
        const int lo = T.d_off[x], cnt = T.d_off[x + 1] - lo;
        unsigned* ticket = bar + PB_BAR_TICKET + x * PB_LINE;
        while (true) {
            if (threadIdx.x == 0)
                s_t = (int)__hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __syncthreads();
            const int t = s_t;
            __syncthreads();
            if (t >= cnt) break;
            const PbDfTask task = T.dtasks[lo + t];
            pb_df_task(T.sjobs[task.job], task, lds, T.gerr);
            __syncthreads();
        }
    }
    // the last workgroup to finish reports the error word
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned f =
            __hip_atomic_fetch_add(bar + PB_BAR_FIN, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) + 1u;
        if (f == gridDim.x)
            T.err_out[0] = __hip_atomic_load(T.gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// All workgroups resident (they register their XCDs at the start barrier).
__global__ __launch_bounds__(PB_THREADS, 2) void k_lu_solve_dataflow(PbDfTables T)
{
    __shared__ alignas(16) float lds[DF_LDS_FLOATS];
    PbBarrier B;
    pb_barrier_init(B, T.barrier, gridDim.x, T.gerr);
    pb_df_serve(T, B, T.barrier, lds);
}

__global__ __launch_bounds__(PB_THREADS, 2) void k_lu_batched_persistent(PbTables T)
{
    extern __shared__ unsigned char smem_raw[];
    const int b = blockIdx.x, nwg = gridDim.x, tid = threadIdx.x;
    PbBarrier B;
    pb_barrier_init(B, T.barrier, (unsigned)nwg, T.gerr);
    __shared__ int s_failed;
    for (int p = 0; p < T.nphase; ++p) {
        const PbPhase ph = T.phases[p];
        if (tid == 0) {
            s_failed = __hip_atomic_load(T.gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (T.tstamp && b == 0) T.tstamp[p] = __builtin_amdgcn_s_memrealtime();
        }
        __syncthreads();
        const int kind = (s_failed && ph.kind != PB_FINAL) ? -1 : ph.kind;
        if (kind == PB_INIT) {
            pb_phase_init(T, b, nwg);
        } else if (kind == PB_PANEL) {
            pb_phase_panel(T, ph, b);
        } else if (kind == PB_LASWP) {
            pb_phase_laswp(T, ph, b, nwg);
        } else if (kind == PB_TRSM) {
            pb_phase_trsm(T, ph, b, nwg, (cpx<float>*)smem_raw);
        } else if (kind == PB_GEMM) {
            pb_phase_gemm(T, ph, b, nwg, (float*)smem_raw);
        } else if (kind == PB_SPLIT) {
            pb_phase_split(T, ph, b, nwg, smem_raw);
        } else if (kind == PB_H3) {
            pb_phase_h3(T, ph, b, nwg, smem_raw);
        } else if (kind == PB_FINAL && b == 0) {
            for (int i = tid; i < T.nmat; i += PB_THREADS) {
                const int e = __hip_atomic_load(T.mats[i].err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
                              __hip_atomic_load(T.gerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                T.mats[i].err_out[0] = e;
            }
        }
        if (T.tstamp) {
            __syncthreads();
            if (tid == 0)
                __hip_atomic_fetch_max(T.tstamp + T.nphase + 1 + p, (uint64_t)__builtin_amdgcn_s_memrealtime(),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (p + 1 < T.nphase) {
            pb_grid_sync(B, T.barrier, T.gerr);
            if (T.tstamp && tid == 0)
                __hip_atomic_fetch_max(T.tstamp + 2 * T.nphase + 1 + p, (uint64_t)__builtin_amdgcn_s_memrealtime(),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (T.tstamp && b == 0 && tid == 0) T.tstamp[T.nphase] = __builtin_amdgcn_s_memrealtime();
}

// Host schedule of the persistent launch (job tables and phases), built for a
// given number of resident workgroups; `dry` only counts (scratch sizing).
struct PbSchedule {
    std::vector<PbPhase> phases;
    std::vector<PbMat> mats;
    std::vector<PbPanelJob> panels;
    std::vector<int> panel_off;
    std::vector<SwapJob<float>> swaps;
    std::vector<int> swap_off;
    std::vector<TrsmJob<float>> trsms;
    std::vector<int> trsm_off;
    std::vector<GemmJobDev> gemms;
    std::vector<int> gemm_off;
    std::vector<PbSplitJob> hjobs;
    std::vector<int> h_off;
    int rmax = 0;           // largest panel rows per workgroup
    std::vector<int> gmax;  // per matrix: largest panel workgroup count
};

struct PbMatIn {
    cpx<float>* a;
    int64_t n, lda;
    int32_t* ipiv;
    int32_t* info;
    LuWs ws;
};

inline int pb_build(const std::vector<PbMatIn>& M, int nwg, int rmax_lds, int P, int trailing, PbSchedule& S)
{
    const int nmat = (int)M.size();
    int64_t nmax = 0;
    for (const auto& m : M) nmax = m.n > nmax ? m.n : nmax;
    S.gmax.assign(nmat, 0);
    S.phases.push_back(PbPhase{PB_INIT, 0, 0, 0, 0, 0});
    auto open = [&](int kind, size_t job0, std::vector<int>& off) {
        PbPhase ph{kind, 0, 0, 0, (int64_t)job0, (int64_t)off.size()};
        off.push_back(0);
        return ph;
    };
    auto close = [&](PbPhase& ph, std::vector<int>& off) {
        if (ph.njob > 0) S.phases.push_back(ph);
        else off.pop_back();
    };
    auto gemm_job = [&](PbPhase& ph, int i, int64_t r0, int64_t Mr, int64_t c0, int64_t Nc, int64_t k0,
                        int64_t kk) -> int {
        if (Mr <= 0 || Nc <= 0 || kk <= 0) return MBA_OK;
        GemmJobDev J;
        const int64_t lda = M[i].lda;
        J.a = M[i].a + r0 * lda + k0;
        J.b = M[i].a + k0 * lda + c0;
        J.c = M[i].a + r0 * lda + c0;
        J.m = Mr;
        J.n = Nc;
        J.k = kk;
        J.lda = J.ldb = J.ldc = lda;
        J.alpha = -1.0;
        J.beta = 1.0;
        S.gemms.push_back(J);
        const int64_t tiles = (int64_t)ph.nvb + grouped_tiles(MBA_F32, Mr, Nc);
        if (tiles > 0x7fffffff) return MBA_ERR_SIZE;
        ph.nvb = (int)tiles;
        ph.njob += 1;
        S.gemm_off.push_back(ph.nvb);
        return MBA_OK;
    };
    const int64_t SP = (int64_t)PB_NB * P;
    for (int64_t k = 0; k < nmax; k += SP) {
        for (int p = 0; p < P; ++p) {
            const int64_t kp = k + (int64_t)PB_NB * p;
            if (kp >= nmax) break;
            // [A] panel columns <- the super panel's earlier panels
            if (p > 0) {
                PbPhase ph = open(PB_GEMM, S.gemms.size(), S.gemm_off);
                ph.sub = 0;
                for (int i = 0; i < nmat; ++i) {
                    const int64_t n = M[i].n;
                    if (kp >= n) continue;
                    const int64_t nbp = (n - kp) < PB_NB ? (n - kp) : PB_NB;
                    const int rc = gemm_job(ph, i, kp, n - kp, kp, nbp, k, kp - k);
                    if (rc != MBA_OK) return rc;
                }
                close(ph, S.gemm_off);
            }
            // [B] panels: rows per workgroup R, as few as fit the resident
            // workgroups, at least 32, at most rmax_lds; matrices that do not
            // fit go to another panel phase
            {
                std::vector<int> todo;
                for (int i = 0; i < nmat; ++i)
                    if (kp < M[i].n) todo.push_back(i);
                size_t t0 = 0;
                while (t0 < todo.size()) {
                    // largest prefix of todo that fits with R = rmax_lds
                    size_t t1 = t0;
                    int64_t gsum = 0;
                    while (t1 < todo.size()) {
                        const int64_t m = M[todo[t1]].n - kp;
                        const int64_t g = (m + rmax_lds - 1) / rmax_lds;
                        if (gsum + g > nwg) break;
                        gsum += g;
                        ++t1;
                    }
                    if (t1 == t0) return MBA_ERR_SIZE;  // one panel exceeds the launch
                    // smallest R >= 32 whose workgroup total fits
                    int64_t msum = 0;
                    for (size_t t = t0; t < t1; ++t) msum += M[todo[t]].n - kp;
                    int64_t R = (msum + nwg - 1) / nwg;
                    if (R < 32) R = 32;
                    for (;; ++R) {
                        if (R >= rmax_lds) {
                            R = rmax_lds;
                            break;
                        }
                        int64_t gs = 0;
                        for (size_t t = t0; t < t1; ++t) gs += (M[todo[t]].n - kp + R - 1) / R;
                        if (gs <= nwg) break;
                    }
                    PbPhase ph = open(PB_PANEL, S.panels.size(), S.panel_off);
                    for (size_t t = t0; t < t1; ++t) {
                        const int i = todo[t];
                        const int64_t n = M[i].n, m = n - kp;
                        const int nbp = (int)((n - kp) < PB_NB ? (n - kp) : PB_NB);
                        int64_t G = (m + R - 1) / R;
                        const int64_t Ri = (m + G - 1) / G;
                        G = (m + Ri - 1) / Ri;
                        if (G > MAX_G) return MBA_ERR_SIZE;
                        S.rmax = (int)Ri > S.rmax ? (int)Ri : S.rmax;
                        S.gmax[i] = (int)G > S.gmax[i] ? (int)G : S.gmax[i];
                        PbPanelJob J;
                        J.P.A = M[i].a;
                        J.P.lda = M[i].lda;
                        J.P.k = kp;
                        J.P.m = m;
                        J.P.nb = nbp;
                        J.P.R = (int)Ri;
                        J.P.ipiv = M[i].ipiv;
                        J.P.info = M[i].info;
                        J.P.err = M[i].ws.err;
                        J.P.meta = M[i].ws.meta;
                        J.P.cval = M[i].ws.cval;
                        J.P.cand = M[i].ws.cand;
                        J.P.diag = M[i].ws.diag;
                        J.slot_row = M[i].ws.slot_row;
                        J.plan = M[i].ws.plan;
                        J.plan_n = M[i].ws.plan_n;
                        J.swaps = (kp > 0 || kp + nbp < n) ? 1 : 0;
                        S.panels.push_back(J);
                        ph.njob += 1;
                        ph.nvb += (int)G;
                        S.panel_off.push_back(ph.nvb);
                    }
                    close(ph, S.panel_off);
                    t0 = t1;
                }
            }
            // [C] row interchanges on every other column
            {
                PbPhase ph = open(PB_LASWP, S.swaps.size(), S.swap_off);
                for (int i = 0; i < nmat; ++i) {
                    const int64_t n = M[i].n;
                    if (kp >= n) continue;
                    const int nbp = (int)((n - kp) < PB_NB ? (n - kp) : PB_NB);
                    const int64_t right = n - (kp + nbp);
                    if (kp == 0 && right <= 0) continue;
                    SwapJob<float> J;
                    J.A = M[i].a;
                    J.lda = M[i].lda;
                    J.c_hi = n;
                    J.k = kp;
                    J.nb = nbp;
                    J.ipiv = M[i].ipiv;
                    J.slot_row = M[i].ws.slot_row;
                    J.plan = M[i].ws.plan;
                    J.plan_n = M[i].ws.plan_n;
                    S.swaps.push_back(J);
                    ph.njob += 1;
                    ph.nvb += (int)((n - nbp + PB_SWAP_COLS - 1) / PB_SWAP_COLS);
                    S.swap_off.push_back(ph.nvb);
                }
                close(ph, S.swap_off);
            }
            // [D] U rows right of the panel: the earlier panels' part, then L11^{-1}
            if (p > 0) {
                PbPhase ph = open(PB_GEMM, S.gemms.size(), S.gemm_off);
                ph.sub = 1;
                for (int i = 0; i < nmat; ++i) {
                    const int64_t n = M[i].n;
                    if (kp >= n) continue;
                    const int64_t nbp = (n - kp) < PB_NB ? (n - kp) : PB_NB;
                    const int rc = gemm_job(ph, i, kp, nbp, kp + nbp, n - kp - nbp, k, kp - k);
                    if (rc != MBA_OK) return rc;
                }
                close(ph, S.gemm_off);
            }
            {
                PbPhase ph = open(PB_TRSM, S.trsms.size(), S.trsm_off);
                for (int i = 0; i < nmat; ++i) {
                    const int64_t n = M[i].n;
                    if (kp >= n) continue;
                    const int nbp = (int)((n - kp) < PB_NB ? (n - kp) : PB_NB);
                    if (kp + nbp >= n) continue;
                    TrsmJob<float> J;
                    J.A = M[i].a;
                    J.lda = M[i].lda;
                    J.k = kp;
                    J.nb = nbp;
                    J.c_lo = kp + nbp;
                    J.c_hi = n;
                    S.trsms.push_back(J);
                    ph.njob += 1;
                    const int tc = nbp == PB_NB ? PB_THREADS : TrsmCols<float>::value;
                    ph.nvb += (int)((n - kp - nbp + tc - 1) / tc);
                    S.trsm_off.push_back(ph.nvb);
                }
                close(ph, S.trsm_off);
            }
        }
        // [E] trailing product of the super panel, k = its width: on the
        // split-FP16 product (operand splits, then product tiles) or the 3M one
        if (trailing == 1) {
            PbPhase ps = open(PB_SPLIT, S.hjobs.size(), S.h_off);
            ps.sub = 2;
            for (int i = 0; i < nmat; ++i) {
                const int64_t n = M[i].n;
                const int64_t c3 = (k + SP) < n ? (k + SP) : n;
                if (k >= n || c3 >= n) continue;
                PbSplitJob J;
                const int64_t lda = M[i].lda, Mr = n - c3, K = c3 - k;
                J.A = M[i].a + c3 * lda + k;
                J.B = M[i].a + k * lda + c3;
                J.C = M[i].a + c3 * lda + c3;
                J.ld = lda;
                J.M = Mr;
                J.N = Mr;
                J.K = (int)K;
                if (!M[i].ws.split || M[i].ws.split_bytes < mba_h3::dev::parts_bytes(Mr, Mr, K)) return MBA_ERR_ALLOC;
                const mba_h3::dev::Parts Pt = mba_h3::dev::parts(Mr, Mr, K, M[i].ws.split);
                J.Kpad = Pt.Kpad;
                J.nrec = Pt.nrec;
                J.Mpad = Pt.Mpad;
                J.Npad = Pt.Npad;
                J.Ap = Pt.Ap;
                J.Bp = Pt.Bp;
                J.ea = Pt.ea;
                J.eb = Pt.eb;
                S.hjobs.push_back(J);
                const int64_t tasks = (int64_t)ps.nvb + Pt.Mpad / 4 + Pt.Npad / mba_h3::dev::SC;
                if (tasks > 0x7fffffff) return MBA_ERR_SIZE;
                ps.nvb = (int)tasks;
                ps.njob += 1;
                S.h_off.push_back(ps.nvb);
            }
            if (ps.njob > 0) {
                const size_t first = (size_t)ps.job0;
                close(ps, S.h_off);
                // the product tiles of the same jobs
                PbPhase pm{PB_H3, 0, 0, 2, (int64_t)first, (int64_t)S.h_off.size()};
                S.h_off.push_back(0);
                for (size_t q = first; q < S.hjobs.size(); ++q) {
                    const PbSplitJob& J = S.hjobs[q];
                    const int64_t tiles = (int64_t)pm.nvb + ((J.M + PB_H3_BM - 1) / PB_H3_BM) *
                                                                ((J.N + mba_h3::dev::BN - 1) / mba_h3::dev::BN);
                    if (tiles > 0x7fffffff) return MBA_ERR_SIZE;
                    pm.nvb = (int)tiles;
                    pm.njob += 1;
                    S.h_off.push_back(pm.nvb);
                }
                S.phases.push_back(pm);
            } else {
                close(ps, S.h_off);
            }
        } else {
            PbPhase ph = open(PB_GEMM, S.gemms.size(), S.gemm_off);
            ph.sub = 2;
            for (int i = 0; i < nmat; ++i) {
                const int64_t n = M[i].n;
                const int64_t c3 = (k + SP) < n ? (k + SP) : n;
                if (k >= n || c3 >= n) continue;
                const int rc = gemm_job(ph, i, c3, n - c3, c3, n - c3, k, c3 - k);
                if (rc != MBA_OK) return rc;
            }
            close(ph, S.gemm_off);
        }
    }
    S.phases.push_back(PbPhase{PB_FINAL, 0, 0, 0, 0, 0});
    return MBA_OK;
}

int pb_resident(int cus, size_t dyn_lds, int* per_cu)
{
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_lu_batched_persistent, PB_THREADS,
                                                     dyn_lds) != hipSuccess)
        return MBA_ERR_LAUNCH;
    if (occ < 1) return MBA_ERR_LAUNCH;
    *per_cu = occ < 2 ? occ : 2;
    (void)cus;
    return MBA_OK;
}

size_t pb_dyn_lds()
{
    hipFuncAttributes fa;
    size_t stat = 8 * 1024;
    if (hipFuncGetAttributes(&fa, (const void*)k_lu_batched_persistent) == hipSuccess) stat = fa.sharedSizeBytes;
    const size_t dyn = PB_WG_LDS > stat + 256 ? PB_WG_LDS - stat - 256 : 0;
    return dyn & ~(size_t)255;
}

template <class V>
size_t pb_bytes(const std::vector<V>& v) { return al256(v.size() * sizeof(V)); }

// Dynamic LDS and resident workgroup count of the persistent launch (the
// same for every table it runs).
int pb_launch_shape(size_t* dyn_out, int* nwg_out)
{
    const int cus = device_cus();
    if (cus <= 0) return MBA_ERR_LAUNCH;
    static std::once_flag attr_once;
    std::call_once(attr_once, [&] {
        (void)hipFuncSetAttribute((const void*)k_lu_batched_persistent, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)pb_dyn_lds());
    });
    const size_t dyn = pb_dyn_lds();
    if (dyn < (size_t)G3M_LDS_FLOATS * sizeof(float) || dyn < (size_t)PB_NB * PB_NB * sizeof(cpx<float>) ||
        dyn < trsm_lds_bytes<float>(PB_NB))
        return MBA_ERR_LAUNCH;
    int per_cu = 0;
    const int rc = pb_resident(cus, dyn, &per_cu);
    if (rc != MBA_OK) return rc;
    // a multiple of the XCD count, so workgroup b stays on XCD b % 8 in every task loop
    const int nwg = (cus * per_cu) / NUM_XCD * NUM_XCD;
    if (nwg < NUM_XCD) return MBA_ERR_LAUNCH;
    *dyn_out = dyn;
    *nwg_out = nwg;
    return MBA_OK;
}

int lu_factor_batched_persistent(const mba_lu_matrix* mats, int nmat, const mba_batched_lu_options* opts,
                                 void* scratch, size_t scratch_bytes, int32_t* err_out, hipStream_t st,
                                 double* kind_ms = nullptr, int64_t* kind_phases = nullptr)
{
    const int P = (opts && opts->super_panels > 0) ? opts->super_panels : PB_P;
    const int trailing = opts ? opts->trailing : 0;
    if (P < 1 || P > PB_P || (trailing != 0 && trailing != 1)) return MBA_ERR_ARG;
    size_t dyn = 0;
    int nwg = 0;
    int rc = pb_launch_shape(&dyn, &nwg);
    if (rc != MBA_OK) return rc;
    const int rmax_lds = (int)(dyn / (PB_NB * sizeof(cpx<float>)));
    std::vector<PbMatIn> in(nmat);
    for (int i = 0; i < nmat; ++i) {
        in[i].a = (cpx<float>*)mats[i].a;
        in[i].n = mats[i].n;
        in[i].lda = mats[i].lda;
        in[i].ipiv = mats[i].ipiv;
        in[i].info = mats[i].info;
        in[i].ws = lu_ws_layout(MBA_F32, mats[i].n, (char*)mats[i].workspace);
        if (mats[i].workspace_bytes < in[i].ws.total) return MBA_ERR_ALLOC;
    }
    PbSchedule S;
    if ((rc = pb_build(in, nwg, rmax_lds, P, trailing, S)) != MBA_OK) return rc;
    // message words each matrix's panels use: meta and cval (two parities of
    // MAX_G), candidate rows of its largest panel workgroup count, row j
    S.mats.resize(nmat);
    const std::vector<int>& gmax = S.gmax;
    for (int i = 0; i < nmat; ++i) {
        PbMat& m = S.mats[i];
        const LuWs& w = in[i].ws;
        m.zero[0] = w.meta;
        m.nzero[0] = 2 * MAX_G;
        m.zero[1] = w.cval;
        m.nzero[1] = 2 * MAX_G;
        m.zero[2] = w.cand;
        m.nzero[2] = (int64_t)2 * gmax[i] * PB_NB * 2;
        m.zero[3] = w.diag;
        m.nzero[3] = (int64_t)2 * PB_NB * 2;
        m.err = w.err;
        m.info = mats[i].info;
        m.err_out = err_out + i;
    }
    // one upload: barrier words and the timeout flag, then the tables
    const size_t o_gerr = al256(PB_BAR_WORDS * sizeof(unsigned));
    const size_t o_ph = o_gerr + 256;
    const size_t o_mat = o_ph + pb_bytes(S.phases);
    const size_t o_pj = o_mat + pb_bytes(S.mats);
    const size_t o_po = o_pj + pb_bytes(S.panels);
    const size_t o_sj = o_po + pb_bytes(S.panel_off);
    const size_t o_so = o_sj + pb_bytes(S.swaps);
    const size_t o_tj = o_so + pb_bytes(S.swap_off);
    const size_t o_to = o_tj + pb_bytes(S.trsms);
    const size_t o_gj = o_to + pb_bytes(S.trsm_off);
    const size_t o_go = o_gj + pb_bytes(S.gemms);
    const size_t o_hj = o_go + pb_bytes(S.gemm_off);
    const size_t o_ho = o_hj + pb_bytes(S.hjobs);
    const size_t o_ts = o_ho + pb_bytes(S.h_off);
    const size_t need = o_ts + al256((3 * S.phases.size() + 1) * sizeof(uint64_t));
    if (need > scratch_bytes) return MBA_ERR_ALLOC;
    std::vector<unsigned char> blob(need, 0);
    auto put = [&](size_t off, const auto& v) {
        if (!v.empty()) std::memcpy(blob.data() + off, v.data(), v.size() * sizeof(v[0]));
    };
    put(o_ph, S.phases);
    put(o_mat, S.mats);
    put(o_pj, S.panels);
    put(o_po, S.panel_off);
    put(o_sj, S.swaps);
    put(o_so, S.swap_off);
    put(o_tj, S.trsms);
    put(o_to, S.trsm_off);
    put(o_gj, S.gemms);
    put(o_go, S.gemm_off);
    put(o_hj, S.hjobs);
    put(o_ho, S.h_off);
    char* base = (char*)scratch;
    if (mba_host::upload_tables(base, blob.data(), need, st) != 0) return MBA_ERR_LAUNCH;
    PbTables T;
    T.phases = (const PbPhase*)(base + o_ph);
    T.nphase = (int)S.phases.size();
    T.nmat = nmat;
    T.mats = (const PbMat*)(base + o_mat);
    T.panels = (const PbPanelJob*)(base + o_pj);
    T.panel_off = (const int*)(base + o_po);
    T.swaps = (const SwapJob<float>*)(base + o_sj);
    T.swap_off = (const int*)(base + o_so);
    T.trsms = (const TrsmJob<float>*)(base + o_tj);
    T.trsm_off = (const int*)(base + o_to);
    T.gemms = (const GemmJobDev*)(base + o_gj);
    T.gemm_off = (const int*)(base + o_go);
    T.hjobs = (const PbSplitJob*)(base + o_hj);
    T.h_off = (const int*)(base + o_ho);
    T.barrier = (unsigned*)base;
    T.gerr = (int*)(base + o_gerr);
    T.tstamp = kind_ms ? (uint64_t*)(base + o_ts) : nullptr;
    void* kargs[] = {&T};
    if ((rc = launch_resident((const void*)k_lu_batched_persistent, (unsigned)nwg, PB_THREADS, kargs, dyn, st)) !=
        MBA_OK)
        return rc;
    if (kind_ms) {
        // per phase kind: summed work time (phase start to the last
        // workgroup's end of work), summed barrier time (from there to the
        // next phase's start) and phase counts
        const size_t np = S.phases.size();
        std::vector<uint64_t> ts(3 * np + 1);
        if (hipMemcpyAsync(ts.data(), base + o_ts, ts.size() * sizeof(uint64_t), hipMemcpyDeviceToHost, st) !=
                hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return MBA_ERR_LAUNCH;
        for (int q = 0; q < 16; ++q) kind_ms[q] = 0.0;
        for (int q = 0; q < 6; ++q)
            if (kind_phases) kind_phases[q] = 0;
        for (size_t q = 0; q < np; ++q) {
            // a phase runs from the last workgroup's release by the previous
            // barrier (the first phase: workgroup 0's start) to the last
            // workgroup's end of work; its barrier from there to the last release
            const int kd = (S.phases[q].kind == PB_SPLIT || S.phases[q].kind == PB_H3) ? PB_GEMM : S.phases[q].kind;
            const uint64_t t0 = q > 0 && ts[2 * np + q] > ts[q] ? ts[2 * np + q] : ts[q];
            const uint64_t t1 = ts[np + 1 + q] > t0 ? ts[np + 1 + q] : t0;
            const uint64_t t2 = q + 1 < np ? (ts[2 * np + 1 + q] > t1 ? ts[2 * np + 1 + q] : t1) : (ts[np] > t1 ? ts[np] : t1);
            kind_ms[kd] += (double)(t1 - t0) * 1e-5;  // 100 MHz ticks -> ms
            kind_ms[6 + kd] += (double)(t2 - t1) * 1e-5;
            if (kd == PB_GEMM && S.phases[q].sub >= 0 && S.phases[q].sub < 3)  // the split product counts as trailing
                kind_ms[12 + S.phases[q].sub] += (double)(t1 - t0) * 1e-5;
            if (kind_phases) kind_phases[kd] += 1;
        }
    }
    return MBA_OK;
}

// Solves of prepared complex64 factorisations, all jobs in one launch of
// resident workgroups (k_lu_solve_dataflow).  Host side: the jobs are dealt to the eight
// XCD queues (largest first, each to the least loaded queue; the load of a job
// is its nblk^2 x groups block products), and each queue's tickets are
// ordered forward by ascending block, then backward by descending block, the
// jobs of a queue interleaved step by step.
struct PbSolvePlan {
    std::vector<PbSolveJob> sj;
    std::vector<PbDfTask> tasks;
    std::vector<int> d_off;
    std::vector<size_t> flag0;  // per job: first flag word
    size_t nflag = 0;
    size_t o_gerr = 0, o_sj = 0, o_dt = 0, o_doff = 0, o_flag = 0, total = 0;
};

int pb_solve_plan(const mba_lu_solve_job* jobs, int njob, bool dry, PbSolvePlan& S)
{
    for (int q = 0; q < njob; ++q) {
        const mba_lu_solve_job& J = jobs[q];
        if (J.n <= 0 || J.nrhs <= 0) continue;
        PbSolveJob d{};
        d.lu = (const cpx<float>*)J.lu;
        d.n = J.n;
        d.lda = J.lda;
        d.nblk = (int)((J.n + TS_BS - 1) / TS_BS);
        d.perm = (const int32_t*)J.aux;
        if (!dry) {
            d.linv = (const cpx<float>*)((const char*)J.aux + al256((size_t)J.n * 4));
            // the inverse tables are laid out as lu_prepare writes them
            d.uinv = (const cpx<float>*)((const char*)d.linv + al256((size_t)d.nblk * TS_BS * TS_BS * 8));
        }
        d.b = (cpx<float>*)J.b;
        d.ldb = J.ldb;
        d.nrhs = J.nrhs;
        d.work = (cpx<float>*)J.work;
        const int64_t ngrp = (J.nrhs + PB_SCW - 1) / PB_SCW;
        if (ngrp > 0x7fff || (int64_t)d.nblk * ngrp > (1 << 24)) return MBA_ERR_SIZE;
        d.ngrp = (int)ngrp;
        S.nflag += (size_t)2 * d.ngrp * d.nblk;
        S.sj.push_back(d);
    }
    const int nj = (int)S.sj.size();
    // queues: jobs by descending load, each to the least loaded queue
    std::vector<int> order(nj), queue(nj);
    for (int q = 0; q < nj; ++q) order[q] = q;
    auto load = [&](int q) { return (double)S.sj[q].nblk * S.sj[q].nblk * S.sj[q].ngrp; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return load(a) > load(b); });
    double qload[PB_MAX_XCD] = {};
    for (int q : order) {
        int best = 0;
        for (int x = 1; x < PB_MAX_XCD; ++x) best = qload[x] < qload[best] ? x : best;
        queue[q] = best;
        qload[best] += load(q);
    }
    size_t ntask = S.nflag;
    if (ntask > 0x7fffffff) return MBA_ERR_SIZE;
    size_t flag_at = 0;
    S.flag0.assign(nj, 0);
    for (int q = 0; q < nj; ++q) {
        S.flag0[q] = flag_at;
        flag_at += (size_t)2 * S.sj[q].ngrp * S.sj[q].nblk;
    }
    S.d_off.assign(PB_MAX_XCD + 1, 0);
    if (!dry) {
        S.tasks.reserve(ntask);
        for (int x = 0; x < PB_MAX_XCD; ++x) {
            S.d_off[x] = (int)S.tasks.size();
            int nb_max = 0;
            for (int q = 0; q < nj; ++q)
                if (queue[q] == x) nb_max = S.sj[q].nblk > nb_max ? S.sj[q].nblk : nb_max;
            for (int pass = 0; pass < 2; ++pass)
                for (int st = 0; st < nb_max; ++st)
                    for (int q = 0; q < nj; ++q) {
                        if (queue[q] != x || st >= S.sj[q].nblk) continue;
                        const int i = pass == 0 ? st : S.sj[q].nblk - 1 - st;
                        for (int g = 0; g < S.sj[q].ngrp; ++g) S.tasks.push_back(PbDfTask{q, i, (int16_t)pass, (int16_t)g});
                    }
        }
        S.d_off[PB_MAX_XCD] = (int)S.tasks.size();
    }
    S.o_gerr = al256(PB_BAR_WORDS * sizeof(unsigned));
    S.o_sj = S.o_gerr + 256;
    S.o_dt = S.o_sj + al256((size_t)nj * sizeof(PbSolveJob));
    S.o_doff = S.o_dt + al256(ntask * sizeof(PbDfTask));
    S.o_flag = S.o_doff + al256((PB_MAX_XCD + 1) * sizeof(int));
    S.total = S.o_flag + al256(S.nflag * sizeof(uint32_t));
    return MBA_OK;
}

size_t pb_solve_scratch(const mba_lu_solve_job* jobs, int32_t n_jobs)
{
    PbSolvePlan S;
    if (pb_solve_plan(jobs, n_jobs, true, S) != MBA_OK) return 0;
    return S.total;
}

int lu_solve_batched_persistent(const mba_lu_solve_job* jobs, int njob, void* scratch, size_t scratch_bytes,
                                int32_t* err_out, hipStream_t st)
{
    int dev = 0, cus = 0, occ = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_lu_solve_dataflow, PB_THREADS, 0) !=
            hipSuccess)
        return MBA_ERR_LAUNCH;
    if (occ < 1) return MBA_ERR_LAUNCH;
    const int nwg = cus * (occ < 4 ? occ : 4);
    PbSolvePlan S;
    int rc = pb_solve_plan(jobs, njob, false, S);
    if (rc != MBA_OK) return rc;
    if (S.total > scratch_bytes) return MBA_ERR_ALLOC;
    char* base = (char*)scratch;
    for (size_t q = 0; q < S.sj.size(); ++q) S.sj[q].flags = (uint32_t*)(base + S.o_flag) + S.flag0[q];
    std::vector<unsigned char> blob(S.total, 0);  // barrier words, error word and flags start at zero
    std::memcpy(blob.data() + S.o_sj, S.sj.data(), S.sj.size() * sizeof(PbSolveJob));
    std::memcpy(blob.data() + S.o_dt, S.tasks.data(), S.tasks.size() * sizeof(PbDfTask));
    std::memcpy(blob.data() + S.o_doff, S.d_off.data(), S.d_off.size() * sizeof(int));
    if (mba_host::upload_tables(base, blob.data(), blob.size(), st) != 0) return MBA_ERR_LAUNCH;
    PbDfTables T{};
    T.sjobs = (const PbSolveJob*)(base + S.o_sj);
    T.dtasks = (const PbDfTask*)(base + S.o_dt);
    T.d_off = (const int*)(base + S.o_doff);
    T.err_out = err_out;
    T.barrier = (unsigned*)base;
    T.gerr = (int*)(base + S.o_gerr);
    void* kargs[] = {&T};
    return launch_resident((const void*)k_lu_solve_dataflow, (unsigned)nwg, PB_THREADS, kargs, 0, st);
}

// Upper bound of the persistent schedule's scratch for n_mat matrices of order <= n_max.
size_t pb_scratch_bound(int32_t n_mat, int64_t n_max)
{
    const size_t np = (size_t)((n_max + PB_NB - 1) / PB_NB);  // panels per matrix
    const size_t sp = np;  // super panels: at most one per panel (super panels of one panel)
    const size_t nm = (size_t)n_mat;
    const size_t phases = 2 + 2 * sp + np * (4 + nm);  // panel phases may split, at most once per matrix
    const size_t pj = nm * np, sj = nm * np, tj = nm * np, gj = nm * (2 * np + sp), hj = nm * sp;
    auto a = [](size_t x) { return al256(x); };
    return a(PB_BAR_WORDS * sizeof(unsigned)) + 256 + a((3 * phases + 1) * sizeof(uint64_t)) +
           a(phases * sizeof(PbPhase)) + a(nm * sizeof(PbMat)) +
           a(pj * sizeof(PbPanelJob)) +
           a((pj + phases) * sizeof(int)) + a(sj * sizeof(SwapJob<float>)) + a((sj + phases) * sizeof(int)) +
           a(tj * sizeof(TrsmJob<float>)) + a((tj + phases) * sizeof(int)) + a(gj * sizeof(GemmJobDev)) +
           a((gj + phases) * sizeof(int)) + a(hj * sizeof(PbSplitJob)) + a((hj + 2 * phases) * sizeof(int));
}

}  // namespace

// ============================================================== C ABI
extern "C" {

size_t mba_lu_workspace_bytes(int prec, int64_t n)
{
    if ((prec != MBA_F32 && prec != MBA_F64) || n <= 0) return 0;
    return lu_ws_layout(prec, n, nullptr).total;
}

int mba_lu_factor(int prec, void* a, int64_t n, int64_t lda, int32_t* ipiv, int32_t* info,
                  void* workspace, size_t workspace_bytes, mba_stream_t stream)
{
    if (!a || !ipiv || !info || !workspace) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n <= 0 || n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
    if (lda < n) return MBA_ERR_ARG;
    return mba_guard([&] {
        if (prec == MBA_F32) return lu_factor<float>(a, n, lda, ipiv, info, workspace, workspace_bytes, S(stream));
        return lu_factor<double>(a, n, lda, ipiv, info, workspace, workspace_bytes, S(stream));
    });
}

int mba_lu_factor_async(int prec, void* a, int64_t n, int64_t lda, int32_t* ipiv, int32_t* info,
                        void* workspace, size_t workspace_bytes, int32_t max_wg, int32_t* err_out,
                        mba_stream_t stream)
{
    if (!a || !ipiv || !info || !workspace || !err_out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n <= 0 || n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
    if (lda < n || max_wg < 0) return MBA_ERR_ARG;
    return mba_guard([&] {
        if (prec == MBA_F32)
            return lu_factor<float>(a, n, lda, ipiv, info, workspace, workspace_bytes, S(stream), nullptr, nullptr,
                                    nullptr, max_wg, err_out);
        return lu_factor<double>(a, n, lda, ipiv, info, workspace, workspace_bytes, S(stream), nullptr, nullptr,
                                 nullptr, max_wg, err_out);
    });
}

size_t mba_lu_solve_batched_scratch_bytes(int32_t n_jobs)
{
    if (n_jobs <= 0) return 0;
    return al256((size_t)n_jobs * sizeof(TsJobDev)) + (size_t)(n_jobs + 1) * sizeof(int);
}

int mba_lu_solve_batched(int prec, const mba_lu_solve_job* jobs, int32_t n_jobs, void* scratch,
                         size_t scratch_bytes, mba_stream_t stream)
{
    if (!jobs || !scratch || n_jobs <= 0) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    int64_t wgs = 0;
    for (int m = 0; m < n_jobs; ++m) {
        const mba_lu_solve_job& j = jobs[m];
        if (!j.lu || !j.aux || !j.b || !j.work) return MBA_ERR_ARG;
        if (j.n <= 0 || j.nrhs <= 0) return MBA_ERR_SIZE;
        if (j.lda < j.n || j.ldb < j.nrhs) return MBA_ERR_ARG;
        wgs += (j.nrhs + TSC - 1) / TSC;
    }
    if (wgs > 0x7fffffff) return MBA_ERR_SIZE;
    return mba_guard([&] {
        if (prec == MBA_F32) return lu_solve_batched<float>(jobs, n_jobs, scratch, scratch_bytes, S(stream));
        return lu_solve_batched<double>(jobs, n_jobs, scratch, scratch_bytes, S(stream));
    });
}

size_t mba_gemm_grouped_scratch_bytes(int32_t n_jobs)
{
    if (n_jobs <= 0) return 0;
    return al256((size_t)n_jobs * sizeof(GemmJobDev)) + (size_t)(n_jobs + 1) * sizeof(int);
}

int mba_gemm_grouped(int prec, const mba_gemm_job* jobs, int32_t n_jobs, void* scratch, size_t scratch_bytes,
                     mba_stream_t stream)
{
    if (!jobs || !scratch || n_jobs <= 0) return MBA_ERR_ARG;
    return mba_guard([&]() -> int {
        if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
        if (scratch_bytes < mba_gemm_grouped_scratch_bytes(n_jobs)) return MBA_ERR_ALLOC;
        std::vector<int> toff(n_jobs + 1, 0);
        int64_t tiles = 0;
        for (int j = 0; j < n_jobs; ++j) {
            const mba_gemm_job& J = jobs[j];
            if (J.m < 0 || J.n < 0 || J.k < 0) return MBA_ERR_SIZE;
            if (J.m > 0 && J.n > 0) {
                if (!J.a || !J.b || !J.c) return MBA_ERR_ARG;
                if (J.lda < J.k || J.ldb < J.n || J.ldc < J.n) return MBA_ERR_ARG;
            }
            tiles += grouped_tiles(prec, J.m, J.n);
            if (tiles > 0x7fffffff - NUM_XCD) return MBA_ERR_SIZE;
            toff[j + 1] = (int)tiles;
        }
        hipStream_t st = S(stream);
        char* base = (char*)scratch;
        GemmJobDev* dj = (GemmJobDev*)base;
        int* dt = (int*)(base + al256((size_t)n_jobs * sizeof(GemmJobDev)));
        if (mba_host::upload_tables(dj, jobs, (size_t)n_jobs * sizeof(GemmJobDev), st) != 0 ||
            mba_host::upload_tables(dt, toff.data(), (size_t)(n_jobs + 1) * sizeof(int), st) != 0)
            return MBA_ERR_LAUNCH;
        return launch_grouped(prec, prec == MBA_F32 && g_gemm_f32_variant.load(std::memory_order_relaxed) == 5, dj,
                              dt, n_jobs, tiles, st);
    });
}

size_t mba_lu_batched_scratch_bytes(int prec, int32_t n_mat, int64_t n_max)
{
    if (n_mat <= 0 || n_max <= 0) return 0;
    const size_t phases = 2 * (size_t)((n_max + 255) / 256);
    const size_t ent = (size_t)n_mat * phases;
    const size_t pa = prec == MBA_F32 ? sizeof(PanelArgs<float>) : sizeof(PanelArgs<double>);
    const size_t sj = prec == MBA_F32 ? sizeof(SwapJob<float>) : sizeof(SwapJob<double>);
    const size_t tj = prec == MBA_F32 ? sizeof(TrsmJob<float>) : sizeof(TrsmJob<double>);
    // panel entries and their launch offsets (at most one launch per entry
    // plus one per phase), then swap and solve jobs with one offset per job
    // and one per phase
    // and the grouped products: three groups per step (3/2 per phase)
    const size_t gent = (size_t)n_mat * 3 * (phases / 2);
    return al256(ent * pa) + al256((2 * ent + phases + 8) * sizeof(int)) + al256(ent * sj) +
           al256((ent + phases + 8) * sizeof(int)) + al256(ent * tj) + al256((ent + phases + 8) * sizeof(int)) +
           al256(gent * sizeof(GemmJobDev)) + (gent + 3 * (phases / 2) + 8) * sizeof(int);
}

int mba_lu_factor_batched(int prec, const mba_lu_matrix* mats, int32_t n_mat, void* scratch, size_t scratch_bytes,
                          int32_t* err_out, mba_stream_t stream)
{
    if (!mats || !scratch || !err_out || n_mat <= 0) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    for (int i = 0; i < n_mat; ++i) {
        const mba_lu_matrix& m = mats[i];
        if (!m.a || !m.ipiv || !m.info || !m.workspace) return MBA_ERR_ARG;
        if (m.n <= 0 || m.n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
        if (m.lda < m.n) return MBA_ERR_ARG;
    }
    return mba_guard([&] {
        if (prec == MBA_F32)
            return lu_factor_batched<float>(mats, n_mat, scratch, scratch_bytes, err_out, S(stream));
        return lu_factor_batched<double>(mats, n_mat, scratch, scratch_bytes, err_out, S(stream));
    });
}

size_t mba_lu_batched_persistent_scratch_bytes(int prec, int32_t n_mat, int64_t n_max)
{
    if (n_mat <= 0 || n_max <= 0) return 0;
    const size_t a = pb_scratch_bound(n_mat, n_max), b = mba_lu_batched_scratch_bytes(prec, n_mat, n_max);
    return a > b ? a : b;
}

int mba_lu_factor_batched_persistent(int prec, const mba_lu_matrix* mats, int32_t n_mat, void* scratch,
                                     size_t scratch_bytes, int32_t* err_out, mba_stream_t stream)
{
    if (!mats || !scratch || !err_out || n_mat <= 0) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    for (int i = 0; i < n_mat; ++i) {
        const mba_lu_matrix& m = mats[i];
        if (!m.a || !m.ipiv || !m.info || !m.workspace) return MBA_ERR_ARG;
        if (m.n <= 0 || m.n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
        if (m.lda < m.n) return MBA_ERR_ARG;
    }
    return mba_guard([&] {
        if (prec == MBA_F32)
            return lu_factor_batched_persistent(mats, n_mat, nullptr, scratch, scratch_bytes, err_out, S(stream));
        return lu_factor_batched<double>(mats, n_mat, scratch, scratch_bytes, err_out, S(stream));
    });
}

size_t mba_lu_solve_batched_persistent_scratch_bytes(const mba_lu_solve_job* jobs, int32_t n_jobs)
{
    if (!jobs || n_jobs <= 0) return 0;
    const size_t a = pb_solve_scratch(jobs, n_jobs), b = mba_lu_solve_batched_scratch_bytes(n_jobs);
    return a == 0 ? 0 : (a > b ? a : b);
}

int mba_lu_solve_batched_persistent(int prec, const mba_lu_solve_job* jobs, int32_t n_jobs, void* scratch,
                                    size_t scratch_bytes, int32_t* err_out, mba_stream_t stream)
{
    if (!jobs || !scratch || !err_out || n_jobs <= 0) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    for (int q = 0; q < n_jobs; ++q) {
        const mba_lu_solve_job& J = jobs[q];
        if (!J.lu || !J.aux || !J.b || !J.work || J.n < 0 || J.nrhs < 0) return MBA_ERR_ARG;
        if (J.n > 0 && (J.lda < J.n || J.ldb < J.nrhs)) return MBA_ERR_ARG;
    }
    return mba_guard([&] {
        if (prec == MBA_F32)
            return lu_solve_batched_persistent(jobs, n_jobs, scratch, scratch_bytes, err_out, S(stream));
        hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, S(stream), err_out, 1);
        if (hipGetLastError() != hipSuccess) return (int)MBA_ERR_LAUNCH;
        return lu_solve_batched<double>(jobs, n_jobs, scratch, scratch_bytes, S(stream));
    });
}

int mba_lu_factor_batched_persistent_opts(int prec, const mba_lu_matrix* mats, int32_t n_mat,
                                          const mba_batched_lu_options* opts, void* scratch, size_t scratch_bytes,
                                          int32_t* err_out, mba_stream_t stream)
{
    if (!mats || !scratch || !err_out || n_mat <= 0) return MBA_ERR_ARG;
    if (prec != MBA_F32) return MBA_ERR_ARG;
    for (int i = 0; i < n_mat; ++i) {
        const mba_lu_matrix& m = mats[i];
        if (!m.a || !m.ipiv || !m.info || !m.workspace) return MBA_ERR_ARG;
        if (m.n <= 0 || m.n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
        if (m.lda < m.n) return MBA_ERR_ARG;
    }
    return mba_guard([&] {
        return lu_factor_batched_persistent(mats, n_mat, opts, scratch, scratch_bytes, err_out, S(stream));
    });
}

int mba_lu_factor_batched_persistent_profiled(const mba_lu_matrix* mats, int32_t n_mat,
                                              const mba_batched_lu_options* opts, void* scratch,
                                              size_t scratch_bytes, int32_t* err_out, double* kind_ms,
                                              int64_t* kind_phases, mba_stream_t stream)
{
    if (!mats || !scratch || !err_out || !kind_ms || n_mat <= 0) return MBA_ERR_ARG;
    for (int i = 0; i < n_mat; ++i) {
        const mba_lu_matrix& m = mats[i];
        if (!m.a || !m.ipiv || !m.info || !m.workspace) return MBA_ERR_ARG;
        if (m.n <= 0 || m.n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
        if (m.lda < m.n) return MBA_ERR_ARG;
    }
    return mba_guard([&] {
        return lu_factor_batched_persistent(mats, n_mat, opts, scratch, scratch_bytes, err_out, S(stream), kind_ms,
                                            kind_phases);
    });
}

int mba_lu_factor_profiled(int prec, void* a, int64_t n, int64_t lda, int32_t* ipiv, int32_t* info,
                           void* workspace, size_t workspace_bytes, double* phase_ms,
                           int64_t* phase_launches, double* gemm_flops, mba_stream_t stream)
{
    if (!a || !ipiv || !info || !workspace || !phase_ms || !phase_launches || !gemm_flops)
        return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n <= 0 || n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
    if (lda < n) return MBA_ERR_ARG;
    return mba_guard([&] {
        if (prec == MBA_F32)
            return lu_factor<float>(a, n, lda, ipiv, info, workspace, workspace_bytes, S(stream), phase_ms,
                                    phase_launches, gemm_flops);
        return lu_factor<double>(a, n, lda, ipiv, info, workspace, workspace_bytes, S(stream), phase_ms,
                                 phase_launches, gemm_flops);
    });
}

int mba_lu_pivots_to_perm(const int32_t* ipiv, int64_t n, int32_t* perm, mba_stream_t stream)
{
    if (!ipiv || !perm) return MBA_ERR_ARG;
    if (n <= 0 || n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
    hipLaunchKernelGGL(k_ipiv_perm, dim3(1), dim3(64), 0, S(stream), ipiv, n, perm);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

size_t mba_lu_solve_workspace_bytes(int prec, int64_t n, int64_t nrhs)
{
    if ((prec != MBA_F32 && prec != MBA_F64) || n <= 0 || nrhs <= 0) return 0;
    if (prec == MBA_F32) return persist_ws_bytes(n, nrhs);
    return 2 * al256((size_t)n * nrhs * 16);
}

size_t mba_lu_solve_aux_bytes(int prec, int64_t n)
{
    if ((prec != MBA_F32 && prec != MBA_F64) || n <= 0) return 0;
    return solve_aux_bytes(prec, n);
}

int mba_lu_solve_prepare(int prec, const void* lu, int64_t n, int64_t lda, const int32_t* ipiv, void* aux,
                         size_t aux_bytes, mba_stream_t stream)
{
    if (!lu || !ipiv || !aux) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n <= 0 || n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
    if (lda < n) return MBA_ERR_ARG;
    if (aux_bytes < solve_aux_bytes(prec, n)) return MBA_ERR_ALLOC;
    if (prec == MBA_F32) return lu_prepare<float>(lu, n, lda, ipiv, aux, S(stream));
    return lu_prepare<double>(lu, n, lda, ipiv, aux, S(stream));
}

size_t mba_lu_solve_prepare_batched_scratch_bytes(int32_t n_jobs)
{
    if (n_jobs <= 0) return 0;
    return al256((size_t)n_jobs * sizeof(PrepJobDev)) + (size_t)(n_jobs + 1) * sizeof(int);
}

int mba_lu_solve_prepare_batched(int prec, const mba_lu_prepare_job* jobs, int32_t n_jobs, void* scratch,
                                 size_t scratch_bytes, mba_stream_t stream)
{
    if (!jobs || !scratch || n_jobs <= 0) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    int64_t wgs = 0;
    for (int j = 0; j < n_jobs; ++j) {
        const mba_lu_prepare_job& J = jobs[j];
        if (!J.lu || !J.ipiv || !J.aux) return MBA_ERR_ARG;
        if (J.n <= 0 || J.n > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
        if (J.lda < J.n) return MBA_ERR_ARG;
        if (J.aux_bytes < solve_aux_bytes(prec, J.n)) return MBA_ERR_ALLOC;
        wgs += 2 * ((J.n + TS_BS - 1) / TS_BS);
        if (wgs > 0x7fffffff) return MBA_ERR_SIZE;
    }
    return mba_guard([&] {
        if (prec == MBA_F32) return lu_prepare_batched<float>(jobs, n_jobs, scratch, scratch_bytes, S(stream));
        return lu_prepare_batched<double>(jobs, n_jobs, scratch, scratch_bytes, S(stream));
    });
}

int mba_lu_solve(int prec, const void* lu, int64_t n, int64_t lda, const void* aux, void* b,
                 int64_t nrhs, int64_t ldb, void* workspace, size_t workspace_bytes,
                 mba_stream_t stream)
{
    if (!lu || !aux || !b || !workspace) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n <= 0 || nrhs <= 0) return MBA_ERR_SIZE;
    if (lda < n || ldb < nrhs) return MBA_ERR_ARG;
    const int32_t* perm = (const int32_t*)aux;
    if (nrhs <= TS_RHS) {
        if (prec == MBA_F32 && g_lu_solve_f32.load(std::memory_order_relaxed) == 1)
            return lu_solve_persist(lu, n, lda, aux, b, nrhs, ldb, workspace, workspace_bytes, S(stream));
        if (prec == MBA_F32)
            return lu_solve_fused<float>(lu, n, lda, aux, b, nrhs, ldb, workspace, workspace_bytes, S(stream));
        return lu_solve_fused<double>(lu, n, lda, aux, b, nrhs, ldb, workspace, workspace_bytes, S(stream));
    }
    if (prec == MBA_F32)
        return lu_solve<float>(lu, n, lda, perm, b, nrhs, ldb, workspace, workspace_bytes, S(stream));
    return lu_solve<double>(lu, n, lda, perm, b, nrhs, ldb, workspace, workspace_bytes, S(stream));
}

int mba_gemm(int prec, int64_t m, int64_t n, int64_t k, double alpha, const void* a, int64_t lda,
             const void* b, int64_t ldb, double beta, void* c, int64_t ldc, mba_stream_t stream)
{
    if (!a || !b || !c) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (m < 0 || n < 0 || k < 0) return MBA_ERR_SIZE;
    if (lda < k || ldb < n || ldc < n) return MBA_ERR_ARG;
    if (prec == MBA_F32) return gemm<float>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, S(stream));
    return gemm<double>(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, S(stream));
}

int mba_gemm_set_variant(int prec, int variant)
{
    if (prec != MBA_F32) return MBA_ERR_ARG;
    if (variant < 0 || variant > 5) return MBA_ERR_ARG;
    g_gemm_f32_variant = variant;
    return MBA_OK;
}

int mba_gemm_split_f16_set_tile(int tile) { return mba_h3::set_tile(tile); }

int mba_lu_set_panel(int kind)
{
    if (kind != 0 && kind != 1) return MBA_ERR_ARG;
    g_lu_panel_f32 = kind;
    return MBA_OK;
}

int mba_lu_set_solve(int kind)
{
    if (kind != 0 && kind != 1) return MBA_ERR_ARG;
    g_lu_solve_f32 = kind;
    return MBA_OK;
}

int mba_lu_set_update(int kind)
{
    if (kind != 0 && kind != 1) return MBA_ERR_ARG;
    g_lu_update_f32 = kind;
    return MBA_OK;
}

size_t mba_gemm_split_f16_workspace_bytes(int64_t m, int64_t n, int64_t k)
{
    if (m <= 0 || n <= 0 || k <= 0 || k > mba_h3::KMAX) return 0;
    return mba_h3::workspace_bytes(m, n, k);
}

int mba_gemm_split_f16(int64_t m, int64_t n, int64_t k, double alpha, const void* a, int64_t lda, const void* b,
                       int64_t ldb, double beta, void* c, int64_t ldc, void* workspace, size_t workspace_bytes,
                       mba_stream_t stream)
{
    if (!a || !b || !c || !workspace) return MBA_ERR_ARG;
    if (m < 0 || n < 0 || k < 1 || k > mba_h3::KMAX) return MBA_ERR_SIZE;
    if (lda < k || ldb < n || ldc < n) return MBA_ERR_ARG;
    return mba_h3::gemm(m, n, k, (float)alpha, a, lda, b, ldb, (float)beta, c, ldc, workspace, workspace_bytes,
                        S(stream));
}

int mba_gemm_variant(int prec, int variant, int64_t m, int64_t n, int64_t k, double alpha,
                     const void* a, int64_t lda, const void* b, int64_t ldb, double beta, void* c,
                     int64_t ldc, mba_stream_t stream)
{
    if (!a || !b || !c) return MBA_ERR_ARG;
    if (prec != MBA_F32 || variant < 0 || variant > 5) return MBA_ERR_ARG;
    if (m < 0 || n < 0 || k < 0) return MBA_ERR_SIZE;
    if (lda < k || ldb < n || ldc < n) return MBA_ERR_ARG;
    if (m == 0 || n == 0) return MBA_OK;
    return cgemm_f32(variant, m, n, k, (float)alpha, a, lda, b, ldb, (float)beta, c, ldc, S(stream));
}

int mba_residual_c128(const void* z, int64_t n, int64_t ldz, const void* x, const void* v, void* r,
                      int64_t nrhs, mba_stream_t stream)
{
    if (!z || !x || !v || !r) return MBA_ERR_ARG;
    if (n <= 0 || nrhs <= 0) return MBA_ERR_SIZE;
    if (ldz < n) return MBA_ERR_ARG;
    const int64_t waves = n * nrhs;
    hipLaunchKernelGGL(k_residual_c128, dim3((unsigned)((waves * 64 + 255) / 256)), dim3(256), 0,
                       S(stream), (const cpx<double>*)z, n, ldz, (const cpx<double>*)x,
                       (const cpx<double>*)v, (cpx<double>*)r, nrhs);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_gemv(int prec, const void* a, int64_t n, int64_t lda, const void* x, void* y,
             mba_stream_t stream)
{
    if (!a || !x || !y) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n <= 0) return MBA_ERR_SIZE;
    if (lda < n) return MBA_ERR_ARG;
    dim3 grid((unsigned)((n * 64 + 255) / 256));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_gemv<float>, grid, dim3(256), 0, S(stream), (const cpx<float>*)a, n, lda,
                           (const cpx<float>*)x, (cpx<float>*)y);
    else
        hipLaunchKernelGGL(k_gemv<double>, grid, dim3(256), 0, S(stream), (const cpx<double>*)a, n, lda,
                           (const cpx<double>*)x, (cpx<double>*)y);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

size_t mba_convert_c128_c64_batched_scratch_bytes(int32_t n_jobs)
{
    if (n_jobs <= 0) return 0;
    return (size_t)n_jobs * sizeof(ConvJobDev);
}

int mba_convert_c128_c64_batched(const mba_convert_job* jobs, int32_t n_jobs, void* scratch, size_t scratch_bytes,
                                 mba_stream_t stream)
{
    if (!jobs || !scratch || n_jobs <= 0 || n_jobs > 65535) return MBA_ERR_ARG;
    for (int q = 0; q < n_jobs; ++q)
        if (jobs[q].count < 0 || (jobs[q].count > 0 && (!jobs[q].src || !jobs[q].dst))) return MBA_ERR_ARG;
    if (scratch_bytes < mba_convert_c128_c64_batched_scratch_bytes(n_jobs)) return MBA_ERR_ALLOC;
    return mba_guard([&]() -> int {
        std::vector<ConvJobDev> dj(n_jobs);
        for (int q = 0; q < n_jobs; ++q)
            dj[q] = ConvJobDev{(const cpx<double>*)jobs[q].src, (cpx<float>*)jobs[q].dst, jobs[q].count};
        hipStream_t st = S(stream);
        if (mba_host::upload_tables(scratch, dj.data(), dj.size() * sizeof(ConvJobDev), st) != 0)
            return MBA_ERR_LAUNCH;
        hipLaunchKernelGGL(k_c128_to_c64_batched, dim3(64, (unsigned)n_jobs), dim3(256), 0, st,
                           (const ConvJobDev*)scratch);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    });
}

size_t mba_refine_update_batched_scratch_bytes(const mba_refine_job* jobs, int32_t n_jobs)
{
    if (!jobs || n_jobs <= 0) return 0;
    int64_t ncol = 0;
    for (int q = 0; q < n_jobs; ++q) ncol = jobs[q].nrhs > ncol ? jobs[q].nrhs : ncol;
    ncol = (ncol + RF_COLS - 1) / RF_COLS * RF_COLS;
    return al256((size_t)n_jobs * sizeof(RefJobDev)) + (size_t)n_jobs * RF_SLABS * ncol * sizeof(double2);
}

int mba_refine_update_batched(const mba_refine_job* jobs, int32_t n_jobs, double* ratio_out, void* scratch,
                              size_t scratch_bytes, mba_stream_t stream)
{
    if (!jobs || !ratio_out || !scratch || n_jobs <= 0 || n_jobs > 65535) return MBA_ERR_ARG;
    int64_t ncol = 0;
    for (int q = 0; q < n_jobs; ++q) {
        const mba_refine_job& J = jobs[q];
        if (J.n < 0 || J.nrhs < 0 || (J.n > 0 && J.nrhs > 0 && (!J.j || !J.d))) return MBA_ERR_ARG;
        if (J.n > 0 && (J.ldj < J.nrhs || J.ldd < J.nrhs)) return MBA_ERR_ARG;
        ncol = J.nrhs > ncol ? J.nrhs : ncol;
    }
    ncol = (ncol + RF_COLS - 1) / RF_COLS * RF_COLS;
    if (ncol == 0 || ncol / RF_COLS > 65535) return MBA_ERR_SIZE;
    if (scratch_bytes < mba_refine_update_batched_scratch_bytes(jobs, n_jobs)) return MBA_ERR_ALLOC;
    return mba_guard([&]() -> int {
        std::vector<RefJobDev> dj(n_jobs);
        for (int q = 0; q < n_jobs; ++q)
            dj[q] = RefJobDev{(cpx<double>*)jobs[q].j, jobs[q].ldj, (const cpx<float>*)jobs[q].d, jobs[q].ldd,
                              jobs[q].n, jobs[q].nrhs};
        hipStream_t st = S(stream);
        RefJobDev* d_jobs = (RefJobDev*)scratch;
        double2* part = (double2*)((char*)scratch + al256((size_t)n_jobs * sizeof(RefJobDev)));
        if (mba_host::upload_tables(d_jobs, dj.data(), dj.size() * sizeof(RefJobDev), st) != 0)
            return MBA_ERR_LAUNCH;
        hipLaunchKernelGGL(k_refine_partial, dim3((unsigned)(ncol / RF_COLS), RF_SLABS, (unsigned)n_jobs),
                           dim3(RF_THREADS), 0, st, (const RefJobDev*)d_jobs, (int)ncol, part);
        MBA_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_refine_ratio, dim3((unsigned)n_jobs), dim3(RF_THREADS), 0, st, (const RefJobDev*)d_jobs,
                           (int)ncol, (const double2*)part, ratio_out);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    });
}

int mba_convert_c128_c64(const void* src, int64_t m, int64_t n, int64_t lds, void* dst, int64_t ldd,
                         mba_stream_t stream)
{
    if (!src || !dst) return MBA_ERR_ARG;
    if (m <= 0 || n <= 0) return MBA_ERR_SIZE;
    if (lds < n || ldd < n) return MBA_ERR_ARG;
    const int64_t tot = m * n;
    hipLaunchKernelGGL(k_c128_to_c64, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, S(stream),
                       (const cpx<double>*)src, m, n, lds, (cpx<float>*)dst, ldd);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
