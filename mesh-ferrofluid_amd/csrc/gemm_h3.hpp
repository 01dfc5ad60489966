// Complex64 GEMM with FP32 accuracy on FP16 MFMA (split operands); see gemm_h3.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mba_h3 {

constexpr int KMAX = 1024;  // largest inner dimension

// Device workspace for gemm() with these sizes (split operands and scales).
size_t workspace_bytes(int64_t M, int64_t N, int64_t K);

// C = alpha A B + beta C for row-major complex64 A [M x K], B [K x N],
// C [M x N], 1 <= K <= KMAX.  A and B are split into the workspace first;
// C must not overlap A or B.
int gemm(int64_t M, int64_t N, int64_t K, float alpha, const void* A, int64_t lda, const void* B, int64_t ldb,
         float beta, void* C, int64_t ldc, void* ws, size_t ws_bytes, hipStream_t st);

// Workgroup tile of the product: 0 = 256 x 128 (8 waves, one workgroup per
// CU; default), 1 = 128 x 128 (4 waves, two per CU).  Process-wide.
int set_tile(int tile);

// The two stages of gemm(): split A [M x K] and B [K x N] into the workspace,
// then C = alpha A B + beta C from the split operands.
int split(int64_t M, int64_t N, int64_t K, const void* A, int64_t lda, const void* B, int64_t ldb, void* ws,
          size_t ws_bytes, hipStream_t st);
int multiply(int64_t M, int64_t N, int64_t K, float alpha, float beta, void* C, int64_t ldc, const void* ws,
             size_t ws_bytes, hipStream_t st);

}  // namespace mba_h3
