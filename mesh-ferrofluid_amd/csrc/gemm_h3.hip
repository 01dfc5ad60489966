// Complex64 GEMM with FP32 accuracy on the FP16 matrix cores of gfx950.
//
// On CDNA4 the dense FP16 MFMA rate is 16x the FP32 MFMA rate, so an FP32
// product can be carried by FP16 pieces:
//   every value x (already scaled by a power of two) is split into
//   hi = fp16(x) and lo = fp16(x - hi)   (x - hi is exact in FP32),
//   x*y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y      (lo_x lo_y ~ 2^-22 x y dropped)
// with FP32 accumulation in the MFMA.  hi + lo holds x to 2^-22 relative.
// Rows of A and columns of B are scaled by powers of two so their largest
// component lies in [2^14, 2^15): FP16 keeps every value down to 2^-29 of the
// row (column) maximum in its normal range, products stay below 2^30, and the
// scale 2^(ea_i + eb_j) is undone exactly in the epilogue.
//
// The complex product uses the real embedding
//   Re C = Ar Br - Ai Bi,  Im C = Ar Bi + Ai Br
// (4 real products x 3 piece products = 12 MFMAs per 32 x 32 block per 16 k).
//
// Operand format (written by the split kernels): per row of A (per column of
// B) and per chunk of 16 k, one 128-byte record of four planes of 16 halves:
// hi_re, hi_im, lo_re, lo_im.  A: [Mpad][Kpad / 16] records; B (transposed):
// [Npad][Kpad / 16] records; pad rows / columns / k are zero.
#include <math.h>
#include <atomic>
#include <stdint.h>

#include "gemm_h3.hpp"
#include "gemm_h3_dev.hpp"
#include "mba_common.hpp"

namespace mba_h3 {
namespace {

using namespace dev;

std::atomic<int> g_tile{0};  // process-wide default (mba_gemm_split_f16_set_tile)

__global__ __launch_bounds__(256) void k_split_rows(const cpx<float>* __restrict__ A, int64_t lda, int64_t M,
                                                    int K, int Kpad, int64_t Mpad, _Float16* __restrict__ Ap,
                                                    int32_t* __restrict__ ea)
{
    split_rows_task<KMAX>(A, lda, M, K, Kpad, Mpad, Ap, ea, (int64_t)blockIdx.x);
}

__global__ __launch_bounds__(256) void k_split_cols(const cpx<float>* __restrict__ B, int64_t ldb, int K, int Kpad,
                                                    int64_t N, int64_t Npad, _Float16* __restrict__ Bp,
                                                    int32_t* __restrict__ eb)
{
    __shared__ SplitColsLds<KMAX> L;
    split_cols_task<KMAX>(B, ldb, K, Kpad, N, Npad, Bp, eb, (int64_t)blockIdx.x, L);
}

template <int BMV, int NT>
__global__ __launch_bounds__(NT, 2) void k_cgemm_h3(int64_t M, int64_t N, int nrec,
                                                      const _Float16* __restrict__ Ap,
                                                      const int32_t* __restrict__ ea,
                                                      const _Float16* __restrict__ Bp,
                                                      const int32_t* __restrict__ eb, float alpha, float beta,
                                                      cpx<float>* __restrict__ C, int64_t ldc)
{
    const int64_t nbm = (M + BMV - 1) / BMV, nbn = (N + BN - 1) / BN;
    int64_t tm, tn;
    if (!tile_of(blockIdx.x, nbm, nbn, tm, tn)) return;
    if constexpr (NT == 512) {
        // operand records staged by direct global -> LDS loads
        __shared__ TileLdsDma<BMV> L;
        tile_h3_dma<BMV, NT>(M, N, nrec, Ap, ea, Bp, eb, alpha, beta, C, ldc, tm, tn, L);
    } else {
        __shared__ TileLds<BMV> L;
        tile_h3<BMV, NT>(M, N, nrec, Ap, ea, Bp, eb, alpha, beta, C, ldc, tm, tn, L);
    }
}

}  // namespace

size_t workspace_bytes(int64_t M, int64_t N, int64_t K) { return parts_bytes(M, N, K); }

int split(int64_t M, int64_t N, int64_t K, const void* A, int64_t lda, const void* B, int64_t ldb, void* ws,
          size_t ws_bytes, hipStream_t st)
{
    if (M <= 0 || N <= 0) return MBA_OK;
    if (K <= 0 || K > KMAX) return MBA_ERR_SIZE;
    if (ws_bytes < workspace_bytes(M, N, K)) return MBA_ERR_ALLOC;
    const Parts P = parts(M, N, K, ws);
    hipLaunchKernelGGL(k_split_rows, dim3((unsigned)((P.Mpad + 3) / 4)), dim3(256), 0, st, (const cpx<float>*)A, lda,
                       M, (int)K, P.Kpad, P.Mpad, P.Ap, P.ea);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_split_cols, dim3((unsigned)(P.Npad / SC)), dim3(256), 0, st, (const cpx<float>*)B, ldb,
                       (int)K, P.Kpad, N, P.Npad, P.Bp, P.eb);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int multiply(int64_t M, int64_t N, int64_t K, float alpha, float beta, void* C, int64_t ldc, const void* ws,
             size_t ws_bytes, hipStream_t st)
{
    if (M <= 0 || N <= 0) return MBA_OK;
    if (K <= 0 || K > KMAX) return MBA_ERR_SIZE;
    if (ws_bytes < workspace_bytes(M, N, K)) return MBA_ERR_ALLOC;
    const Parts P = parts(M, N, K, (void*)ws);
    if (g_tile.load(std::memory_order_relaxed) == 0) {
        const int64_t tiles = ((M + 255) / 256) * ((N + BN - 1) / BN);
        const int64_t blocks = (tiles + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
        hipLaunchKernelGGL((k_cgemm_h3<256, 512>), dim3((unsigned)blocks), dim3(512), 0, st, M, N, P.nrec, P.Ap, P.ea,
                           P.Bp, P.eb, alpha, beta, (cpx<float>*)C, ldc);
    } else {
        const int64_t tiles = ((M + 127) / 128) * ((N + BN - 1) / BN);
        const int64_t blocks = (tiles + NUM_XCD - 1) / NUM_XCD * NUM_XCD;
        hipLaunchKernelGGL((k_cgemm_h3<128, 256>), dim3((unsigned)blocks), dim3(256), 0, st, M, N, P.nrec, P.Ap, P.ea,
                           P.Bp, P.eb, alpha, beta, (cpx<float>*)C, ldc);
    }
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int set_tile(int tile)
{
    if (tile != 0 && tile != 1) return MBA_ERR_ARG;
    g_tile = tile;
    return MBA_OK;
}

int gemm(int64_t M, int64_t N, int64_t K, float alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
         float beta, void* C, int64_t ldc, void* ws, size_t ws_bytes, hipStream_t st)
{
    int rc = split(M, N, K, A, lda, B, ldb, ws, ws_bytes, st);
    if (rc != MBA_OK) return rc;
    return multiply(M, N, K, alpha, beta, C, ldc, ws, ws_bytes, st);
}

}  // namespace mba_h3
