// Device code of the split-FP16 complex product (FP32-accurate complex64
// GEMM on the FP16 matrix cores): the operand splits and the product tile,
// shared by the standalone launches (gemm_h3.hip) and the persistent batched
// LU (dense.hip).  Record format and scaling: see gemm_h3.hip.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace mba_h3 {
namespace dev {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int KC = 16;                  // k per chunk (one v_mfma_f32_32x32x16_f16)
constexpr int REC_H = 4 * KC;           // halves per record
constexpr int BM = 256, BN = 128;       // largest workgroup tile (complex); operand padding unit
// Tile variants: 256 x 128 with 8 waves (4 x 2, 64 x 64 each, one workgroup
// per CU) or 128 x 128 with 4 waves (2 x 2, two workgroups per CU, so one
// workgroup's C read / write overlaps the other's MFMA work).
constexpr int PITCH_H = REC_H + 8;      // LDS row pitch in halves (144 B: conflict-free 16-byte reads)
constexpr int NUM_XCD = 8, GROUP_M = 8;
constexpr int SCALE_TOP = 15;           // scaled maxima in [2^14, 2^15)

// exponent e with max = m 2^e, m in [0.5, 1); 0 for max == 0
__device__ __forceinline__ int scale_exp(float mx)
{
    if (!(mx > 0.f)) return 0;
    int e;
    (void)frexpf(mx, &e);
    return e - SCALE_TOP;  // value = scaled * 2^ea, scaled = value * 2^-ea
}

__device__ __forceinline__ void split2(float x, _Float16& h, _Float16& l)
{
    h = (_Float16)x;
    l = (_Float16)(x - (float)h);
}

// ---- A: one wave per row (K <= KM: lane holds 4 consecutive k of
// each 256-wide block)
template <int KM>
__device__ __forceinline__ void split_rows_task(const cpx<float>* __restrict__ A, int64_t lda, int64_t M, int K,
                                                int Kpad, int64_t Mpad, _Float16* __restrict__ Ap,
                                                int32_t* __restrict__ ea, int64_t rb)
{
    constexpr int NBLK = KM / 256;
    typedef _Float16 half4 __attribute__((ext_vector_type(4)));
    const int lane = threadIdx.x & 63;
    const int64_t row = rb * 4 + (threadIdx.x >> 6);
    if (row >= Mpad) return;
    const int nrec = Kpad / KC;
    float mx = 0.f;
    cpx<float> v[NBLK][4];
#pragma unroll
    for (int b = 0; b < NBLK; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = 256 * b + 4 * lane + q;
            v[b][q] = (row < M && k < K) ? A[row * lda + k] : cpx<float>{0.f, 0.f};
            mx = fmaxf(mx, fmaxf(fabsf(v[b][q].re), fabsf(v[b][q].im)));
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off));
    const int e = scale_exp(mx);
    if (lane == 0) ea[row] = e;
#pragma unroll
    for (int b = 0; b < NBLK; ++b) {
        const int kc = 16 * b + lane / 4;
        if (kc >= nrec) break;
        _Float16 h[2][4], l[2][4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            split2(ldexpf(v[b][q].re, -e), h[0][q], l[0][q]);
            split2(ldexpf(v[b][q].im, -e), h[1][q], l[1][q]);
        }
        _Float16* rec = Ap + (row * nrec + kc) * REC_H + (lane & 3) * 4;
        *(half4*)(rec + 0 * KC) = half4{h[0][0], h[0][1], h[0][2], h[0][3]};
        *(half4*)(rec + 1 * KC) = half4{h[1][0], h[1][1], h[1][2], h[1][3]};
        *(half4*)(rec + 2 * KC) = half4{l[0][0], l[0][1], l[0][2], l[0][3]};
        *(half4*)(rec + 3 * KC) = half4{l[1][0], l[1][1], l[1][2], l[1][3]};
    }
}

// ---- B (K x N, row-major) -> transposed records.  Workgroup: SC columns;
// the K x SC block is read once into LDS (all 32 loads of a thread in flight
// together), then the column maxima and the split records come from LDS.
constexpr int SC = 16;

template <int KM>
struct SplitColsLds {
    cpx<float> blk[KM * SC];  // [k][c]; rows k >= K are zero
    float red[256 / SC][SC];
    int s_e[SC];
};

template <int KM>
__device__ __forceinline__ void split_cols_task(const cpx<float>* __restrict__ B, int64_t ldb, int K, int Kpad,
                                                int64_t N, int64_t Npad, _Float16* __restrict__ Bp,
                                                int32_t* __restrict__ eb, int64_t cb, SplitColsLds<KM>& L)
{
    constexpr int LPT = KM * SC / 256;  // block elements per thread
    cpx<float>* blk = L.blk;
    float(*red)[SC] = L.red;
    int* s_e = L.s_e;
    const int tid = threadIdx.x;
    const int64_t col0 = cb * SC;
    const int nrec = Kpad / KC;
    {
        cpx<float> v[LPT];
#pragma unroll
        for (int i = 0; i < LPT; ++i) {
            const int e = tid + 256 * i, k = e / SC;
            const int64_t col = col0 + e % SC;
            v[i] = (k < K && col < N) ? B[(int64_t)k * ldb + col] : cpx<float>{0.f, 0.f};
        }
#pragma unroll
        for (int i = 0; i < LPT; ++i) blk[tid + 256 * i] = v[i];
    }
    __syncthreads();
    {
        const int c = tid % SC, g = tid / SC;
        float mx = 0.f;
        for (int k = g; k < K; k += 256 / SC) {
            const cpx<float> x = blk[k * SC + c];
            mx = fmaxf(mx, fmaxf(fabsf(x.re), fabsf(x.im)));
        }
        red[g][c] = mx;
    }
    __syncthreads();
    if (tid < SC) {
        float mx = 0.f;
#pragma unroll
        for (int g = 0; g < 256 / SC; ++g) mx = fmaxf(mx, red[g][tid]);
        const int e = scale_exp(mx);
        s_e[tid] = e;
        if (col0 + tid < Npad) eb[col0 + tid] = e;
    }
    __syncthreads();
    for (int pi = tid; pi < SC * nrec; pi += 256) {
        const int c = pi % SC, kc = pi / SC;
        const int64_t col = col0 + c;
        if (col >= Npad) continue;
        const int e = s_e[c];
        _Float16 h[2][KC], l[2][KC];
#pragma unroll
        for (int q = 0; q < KC; ++q) {
            const cpx<float> x = blk[(kc * KC + q) * SC + c];
            split2(ldexpf(x.re, -e), h[0][q], l[0][q]);
            split2(ldexpf(x.im, -e), h[1][q], l[1][q]);
        }
        half8* rec = (half8*)(Bp + (col * nrec + kc) * REC_H);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            rec[2 * p + 0] = half8{h[p][0], h[p][1], h[p][2], h[p][3], h[p][4], h[p][5], h[p][6], h[p][7]};
            rec[2 * p + 1] = half8{h[p][8], h[p][9], h[p][10], h[p][11], h[p][12], h[p][13], h[p][14], h[p][15]};
            rec[4 + 2 * p + 0] = half8{l[p][0], l[p][1], l[p][2], l[p][3], l[p][4], l[p][5], l[p][6], l[p][7]};
            rec[4 + 2 * p + 1] = half8{l[p][8], l[p][9], l[p][10], l[p][11], l[p][12], l[p][13], l[p][14], l[p][15]};
        }
    }
}

// tile L -> (row tile, column tile): each XCD (workgroup L runs on XCD L % 8)
// takes a contiguous range of tiles, walked in groups of GROUP_M row tiles so
// the group's A records stay in the XCD's L2 while B records stream.
__device__ __forceinline__ bool tile_of(int64_t L, int64_t nbm, int64_t nbn, int64_t& tm, int64_t& tn)
{
    const int64_t T = nbm * nbn;
    const int64_t per = (T + NUM_XCD - 1) / NUM_XCD;
    const int64_t v = (L % NUM_XCD) * per + L / NUM_XCD;
    if (v >= T) return false;
    const int64_t group = v / (GROUP_M * nbn);
    const int64_t first_m = group * GROUP_M;
    const int64_t gm = (nbm - first_m) < GROUP_M ? (nbm - first_m) : GROUP_M;
    const int64_t in = v - group * GROUP_M * nbn;
    tm = first_m + in % gm;
    tn = in / gm;
    return true;
}

// C[M x N] = alpha * 2^(ea_i + eb_j) * (A' B')_ij + beta * C   (complex64, row-major)
// LDS of one tile: As [2][BMV * PITCH_H], Bs [2][BN * PITCH_H] halves
template <int BMV>
struct TileLds {
    alignas(16) _Float16 As[2][BMV * PITCH_H];
    alignas(16) _Float16 Bs[2][BN * PITCH_H];
};

// One BMV x BN tile (row tile tm, column tile tn) with NT threads
template <int BMV, int NT>
__device__ __forceinline__ void tile_h3(int64_t M, int64_t N, int nrec, const _Float16* __restrict__ Ap,
                                        const int32_t* __restrict__ ea, const _Float16* __restrict__ Bp,
                                        const int32_t* __restrict__ eb, float alpha, float beta,
                                        cpx<float>* __restrict__ C, int64_t ldc, int64_t tm, int64_t tn,
                                        TileLds<BMV>& L)
{
    constexpr int RSTEP = NT / 8, QA = BMV / RSTEP, QB = BN / RSTEP;
    auto& As = L.As;
    auto& Bs = L.Bs;
    const int64_t m0 = tm * BMV, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int l32 = lane & 31, oct = lane >> 5;

    // global -> register staging: A BMV records x 8 parts of 16 B, B BN
    // records x 8 parts; part = tid % 8, record tid / 8 + RSTEP q
    const int part = tid & 7, rsub = tid >> 3;  // rsub in [0, RSTEP)
    half8 ga[QA], gb[QB];
    auto gload = [&](int rc) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int64_t row = m0 + rsub + RSTEP * q;
            ga[q] = *(const half8*)(Ap + (row * nrec + rc) * REC_H + part * 8);
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int64_t col = n0 + rsub + RSTEP * q;
            gb[q] = *(const half8*)(Bp + (col * nrec + rc) * REC_H + part * 8);
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < QA; ++q) *(half8*)&As[buf][(rsub + RSTEP * q) * PITCH_H + part * 8] = ga[q];
#pragma unroll
        for (int q = 0; q < QB; ++q) *(half8*)&Bs[buf][(rsub + RSTEP * q) * PITCH_H + part * 8] = gb[q];
    };

    floatx16 acc[2][2][2];  // [i][j][re/im]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][0][r] = acc[i][j][1][r] = 0.f;

    gload(0);
    sstore(0);
    __syncthreads();
    for (int rc = 0; rc < nrec; ++rc) {
        const int buf = rc & 1;
        if (rc + 1 < nrec) gload(rc + 1);
        // A operands of this wave: rows wm*64 + i*32 + l32, planes p, k-octet oct
        half8 a[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int p = 0; p < 4; ++p)
                a[i][p] = *(const half8*)&As[buf][(wm * 64 + i * 32 + l32) * PITCH_H + p * KC + oct * 8];
        // B operands of both 32-column blocks, -Bi (hi, lo) by a sign flip of
        // the packed halves
        half8 b[2][4], nbh[2], nbl[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int p = 0; p < 4; ++p)
                b[j][p] = *(const half8*)&Bs[buf][(wn * 64 + j * 32 + l32) * PITCH_H + p * KC + oct * 8];
            nbh[j] = -b[j][1];
            nbl[j] = -b[j][3];
        }
        // six steps (hi x hi, hi x lo, lo x hi, each as a real and an
        // imaginary pair); every step issues one MFMA into each of the eight
        // accumulators, so consecutive MFMAs never depend on each other.
        // Each accumulator receives its 12 products in the same order as a
        // block-by-block loop would give (bit-identical results).
#define MBA_H3_STEP(AR, BR, AI, BI)                                                                           \
        _Pragma("unroll") for (int j = 0; j < 2; ++j) _Pragma("unroll") for (int i = 0; i < 2; ++i) {        \
            acc[i][j][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[i][AR], BR, acc[i][j][0], 0, 0, 0);       \
            acc[i][j][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[i][AI], BI, acc[i][j][1], 0, 0, 0);       \
        }
        MBA_H3_STEP(0, b[j][0], 0, b[j][1])  // hi x hi: Ar Br, Ar Bi
        MBA_H3_STEP(1, nbh[j], 1, b[j][0])   //          -Ai Bi, Ai Br
        MBA_H3_STEP(0, b[j][2], 0, b[j][3])  // hi x lo
        MBA_H3_STEP(1, nbl[j], 1, b[j][2])
        MBA_H3_STEP(2, b[j][0], 2, b[j][1])  // lo x hi
        MBA_H3_STEP(3, nbh[j], 3, b[j][0])
#undef MBA_H3_STEP
        if (rc + 1 < nrec) sstore(buf ^ 1);
        __syncthreads();
    }
    // epilogue: register r of lane -> row 8 (r / 4) + 4 oct + r % 4, column l32.
    // Per 32 x 32 block all C loads are issued before any store (C is read
    // and written through one pointer, so the compiler would otherwise wait
    // for each load in turn).
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int64_t gj = n0 + wn * 64 + j * 32 + l32;
        const bool jok = gj < N;
        const int ej = jok ? eb[gj] : 0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int64_t gi0 = m0 + wm * 64 + i * 32 + 4 * oct;
            float2 cv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = gi0 + 8 * (r >> 2) + (r & 3);
                cv[r] = (beta != 0.f && jok && gi < M) ? *(const float2*)(C + gi * ldc + gj) : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = gi0 + 8 * (r >> 2) + (r & 3);
                if (jok && gi < M) {
                    // scale the product, not alpha: 2^(ea + eb) alone may leave FP32's range
                    const int e = ea[gi] + ej;
                    float re = ldexpf(alpha * acc[i][j][0][r], e), im = ldexpf(alpha * acc[i][j][1][r], e);
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


// ---- the same tile with the operand records staged by direct global -> LDS
// loads (global_load_lds_dwordx4: each lane's 16 bytes land at LDS base +
// 16 lane), no VGPR round trip and no LDS stores.  Records sit unpadded (128
// B per row) with the 16-byte parts of row R permuted by part ^ (R & 7), so
// the operand reads of eight consecutive rows at one part hit eight distinct
// 16-byte bank groups; a lane loads the global part that belongs at its
// fixed LDS slot.  The loads of chunk rc + 1 go to the buffer that chunk rc - 1
// read (every wave is past the barrier that ended it); a chunk's records are
// complete for all waves after each wave's vector memory counter reaches 0
// and the barrier.  Same MFMA order and epilogue as tile_h3: bit-identical
// results.
typedef __attribute__((address_space(3))) void lds_void_t;
constexpr int DREC_H = REC_H;  // halves per LDS row (no padding)

template <int BMV>
struct TileLdsDma {
    alignas(16) _Float16 As[2][BMV * DREC_H];
    alignas(16) _Float16 Bs[2][BN * DREC_H];
};

template <int BMV, int NT>
__device__ __forceinline__ void tile_h3_dma(int64_t M, int64_t N, int nrec, const _Float16* __restrict__ Ap,
                                            const int32_t* __restrict__ ea, const _Float16* __restrict__ Bp,
                                            const int32_t* __restrict__ eb, float alpha, float beta,
                                            cpx<float>* __restrict__ C, int64_t ldc, int64_t tm, int64_t tn,
                                            TileLdsDma<BMV>& L)
{
    static_assert((NT == 512 || NT == 256) && BMV == NT / 2, "2 x 2 or 4 x 2 waves of 64 x 64");
    constexpr int RSTEP = NT / 8;  // record rows staged per load instruction of all waves
    constexpr int QA = BMV / RSTEP, QB = BN / RSTEP;
    const int64_t m0 = tm * BMV, n0 = tn * BN;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave >> 1, wn = wave & 1;
    const int l32 = lane & 31, oct = lane >> 5;
    // instruction q of this wave stages rows RSTEP q + 8 wave + lane / 8; lane
    // writes LDS slot lane % 8 of its row, i.e. global part (lane % 8) ^ (lane / 8)
    const int rsub = 8 * wave + (lane >> 3), gpart = (lane & 7) ^ (lane >> 3);
    // addresses: a wave-uniform base (the tile's first record row; +RSTEP q
    // rows for instruction q, +rc records for chunk rc) plus one 32-bit lane
    // offset in bytes (this lane's row and part within the first RSTEP rows)
    const uint32_t la = (uint32_t)((rsub * (int64_t)nrec * REC_H + gpart * 8) * 2);
    const int64_t qstride = (int64_t)RSTEP * nrec * REC_H;
    const _Float16* Ab = Ap + m0 * nrec * REC_H;
    const _Float16* Bb = Bp + n0 * nrec * REC_H;
    auto dma = [&](int rc, _Float16* As, _Float16* Bs) {
#pragma unroll
        for (int q = 0; q < QA; ++q)
            __builtin_amdgcn_global_load_lds((const void*)((const char*)(Ab + q * qstride + rc * REC_H) + la),
                                             (lds_void_t*)(As + (RSTEP * q + 8 * wave) * DREC_H), 16, 0, 0);
#pragma unroll
        for (int q = 0; q < QB; ++q)
            __builtin_amdgcn_global_load_lds((const void*)((const char*)(Bb + q * qstride + rc * REC_H) + la),
                                             (lds_void_t*)(Bs + (RSTEP * q + 8 * wave) * DREC_H), 16, 0, 0);
    };
    // operand reads: rows R = w* 64 + i 32 + l32 (R & 7 = l32 & 7), part
    // 2 p + oct at slot (2 p + oct) ^ (l32 & 7): one lane offset per p, the
    // row block i and the buffer are immediate offsets
    int offA[4], offB[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        const int slot = ((2 * p + oct) ^ (l32 & 7)) * 8;
        offA[p] = (wm * 64 + l32) * DREC_H + slot;
        offB[p] = (wn * 64 + l32) * DREC_H + slot;
    }

    floatx16 acc[2][2][2];  // [i][j][re/im]
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][0][r] = acc[i][j][1][r] = 0.f;

    auto compute = [&](const _Float16* As, const _Float16* Bs) {
        half8 a[2][4];
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int p = 0; p < 4; ++p) a[i][p] = *(const half8*)&As[offA[p] + i * 32 * DREC_H];
        half8 b[2][4], nbh[2], nbl[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
#pragma unroll
            for (int p = 0; p < 4; ++p) b[j][p] = *(const half8*)&Bs[offB[p] + j * 32 * DREC_H];
            nbh[j] = -b[j][1];
            nbl[j] = -b[j][3];
        }
#define MBA_H3_STEP(AR, BR, AI, BI)                                                                           \
        _Pragma("unroll") for (int j = 0; j < 2; ++j) _Pragma("unroll") for (int i = 0; i < 2; ++i) {        \
            acc[i][j][0] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[i][AR], BR, acc[i][j][0], 0, 0, 0);       \
            acc[i][j][1] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[i][AI], BI, acc[i][j][1], 0, 0, 0);       \
        }
        MBA_H3_STEP(0, b[j][0], 0, b[j][1])  // hi x hi: Ar Br, Ar Bi
        MBA_H3_STEP(1, nbh[j], 1, b[j][0])   //          -Ai Bi, Ai Br
        MBA_H3_STEP(0, b[j][2], 0, b[j][3])  // hi x lo
        MBA_H3_STEP(1, nbl[j], 1, b[j][2])
        MBA_H3_STEP(2, b[j][0], 2, b[j][1])  // lo x hi
        MBA_H3_STEP(3, nbh[j], 3, b[j][0])
#undef MBA_H3_STEP
    };
    // a chunk's records are visible to every wave once each wave's loads of
    // it have landed (vector memory counter 0) and the barrier has passed
    auto landed = [&]() {
        __builtin_amdgcn_s_waitcnt(0x3f70);  // vmcnt(0)
        __syncthreads();
    };

    dma(0, L.As[0], L.Bs[0]);
    landed();
    for (int rc = 0; rc < nrec; ++rc) {
        const int buf = rc & 1;
        if (rc + 1 < nrec) dma(rc + 1, L.As[buf ^ 1], L.Bs[buf ^ 1]);
        compute(L.As[buf], L.Bs[buf]);
        landed();
    }
    // epilogue: as tile_h3
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const int64_t gj = n0 + wn * 64 + j * 32 + l32;
        const bool jok = gj < N;
        const int ej = jok ? eb[gj] : 0;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int64_t gi0 = m0 + wm * 64 + i * 32 + 4 * oct;
            float2 cv[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = gi0 + 8 * (r >> 2) + (r & 3);
                cv[r] = (beta != 0.f && jok && gi < M) ? *(const float2*)(C + gi * ldc + gj) : make_float2(0.f, 0.f);
            }
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int64_t gi = gi0 + 8 * (r >> 2) + (r & 3);
                if (jok && gi < M) {
                    const int e = ea[gi] + ej;
                    float re = ldexpf(alpha * acc[i][j][0][r], e), im = ldexpf(alpha * acc[i][j][1][r], e);
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


// Workspace layout of the split operands for M x N x K: A records [Mpad][Kpad
// / 16], B records [Npad][Kpad / 16] (128 bytes each), exponents ea [Mpad],
// eb [Npad]; pads of BM rows and BN columns.
inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

struct Parts {
    _Float16 *Ap, *Bp;
    int32_t *ea, *eb;
    int Kpad, nrec;
    int64_t Mpad, Npad;
};

inline size_t parts_bytes(int64_t M, int64_t N, int64_t K)
{
    const int64_t Kpad = (K + KC - 1) / KC * KC;
    const int64_t Mpad = (M + BM - 1) / BM * BM, Npad = (N + BN - 1) / BN * BN;
    return al256((size_t)Mpad * Kpad * 4 * 2) + al256((size_t)Npad * Kpad * 4 * 2) + al256((size_t)Mpad * 4) +
           al256((size_t)Npad * 4);
}

inline Parts parts(int64_t M, int64_t N, int64_t K, void* ws)
{
    Parts P;
    P.Kpad = (int)((K + KC - 1) / KC * KC);
    P.nrec = P.Kpad / KC;
    P.Mpad = (M + BM - 1) / BM * BM;
    P.Npad = (N + BN - 1) / BN * BN;
    char* p = (char*)ws;
    P.Ap = (_Float16*)p;
    p += al256((size_t)P.Mpad * P.Kpad * 4 * 2);
    P.Bp = (_Float16*)p;
    p += al256((size_t)P.Npad * P.Kpad * 4 * 2);
    P.ea = (int32_t*)p;
    p += al256((size_t)P.Mpad * 4);
    P.eb = (int32_t*)p;
    return P;
}

}  // namespace dev
}  // namespace mba_h3
