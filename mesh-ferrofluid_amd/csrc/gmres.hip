// GMRES building blocks for gfx950: dense matrix-vector products and
// classical Gram-Schmidt steps over a row-major basis.
//
// The matrix and the Krylov basis are FP32, complex (interleaved re, im) or
// real.  The vectors the products read and write (x, b, y, w) are FP32 or
// FP64 (vprec): GMRES keeps its iterate and its restart residual in FP64 and
// everything else in FP32.  The matvec is HBM bound (the matrix is read once
// per product) and accumulates in FP64: the FP32 x FP32 products are exact in
// FP64, so b - A x carries no accumulation error above the FP64 level (an
// FP32 sum over a row of an EFIE matrix leaves a floor near 1e-6 relative),
// and the FP64 FMAs stay far below the FP64 VALU rate at HBM speed.  Inner
// products accumulate per thread in FP32 and across threads and blocks in
// FP64 with a fixed reduction order (deterministic).
#include <math.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace {

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

typedef float f4v __attribute__((ext_vector_type(4)));

typedef double d2v __attribute__((ext_vector_type(2)));

// x[4t .. 4t+3] as doubles; xvec: x is 16-byte aligned (vector loads).
__device__ __forceinline__ void ld4(const float* p, int xvec, double v[4])
{
    if (xvec) {
        const f4v q = *reinterpret_cast<const f4v*>(p);
        v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
    } else {
        v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
    }
}
__device__ __forceinline__ void ld4(const double* p, int xvec, double v[4])
{
    if (xvec) {
        const d2v q0 = *reinterpret_cast<const d2v*>(p), q1 = *reinterpret_cast<const d2v*>(p + 2);
        v[0] = q0.x; v[1] = q0.y; v[2] = q1.x; v[3] = q1.y;
    } else {
        v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
    }
}

// y[i] = (b[i] -) sum_t A[i, t] x[t].  One wave per MV_ROWS consecutive rows:
// each x chunk is read once for all of them (x stays in L2; the matrix streams
// from HBM with 16-byte non-temporal loads).
// L: row length in floats; ldf: row stride in floats (multiple of 4 when vec).
constexpr int MV_ROWS = 4;

template <bool CPLX, class VT>
__global__ __launch_bounds__(256) void k_mv32(const float* __restrict__ A, int64_t rows, int64_t L,
                                              int64_t ldf, const VT* __restrict__ x,
                                              const VT* __restrict__ b, VT* __restrict__ y, int vec,
                                              int xvec)
{
    const int lane = threadIdx.x & 63;
    const int64_t i0 = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * MV_ROWS;
    if (i0 >= rows) return;
    const int nr = (int)((rows - i0) < MV_ROWS ? (rows - i0) : MV_ROWS);
    const float* Ar[MV_ROWS];
#pragma unroll
    for (int r = 0; r < MV_ROWS; ++r) Ar[r] = A + (i0 + (r < nr ? r : 0)) * ldf;
    double sr[MV_ROWS], si[MV_ROWS];
#pragma unroll
    for (int r = 0; r < MV_ROWS; ++r) sr[r] = si[r] = 0.0;
    int64_t done = 0;
    if (vec) {
        const int64_t nv = L >> 2;  // float4 per row
        for (int64_t t = lane; t < nv; t += 64) {
            f4v a[MV_ROWS];
#pragma unroll
            for (int r = 0; r < MV_ROWS; ++r)
                a[r] = __builtin_nontemporal_load(reinterpret_cast<const f4v*>(Ar[r]) + t);
            double v[4];
            ld4(x + 4 * t, xvec, v);
#pragma unroll
            for (int r = 0; r < MV_ROWS; ++r) {
                if (CPLX) {
                    sr[r] += (double)a[r].x * v[0] - (double)a[r].y * v[1] + (double)a[r].z * v[2] -
                             (double)a[r].w * v[3];
                    si[r] += (double)a[r].x * v[1] + (double)a[r].y * v[0] + (double)a[r].z * v[3] +
                             (double)a[r].w * v[2];
                } else {
                    sr[r] += (double)a[r].x * v[0] + (double)a[r].y * v[1] + (double)a[r].z * v[2] +
                             (double)a[r].w * v[3];
                }
            }
        }
        done = nv << 2;
    }
    // scalar part (unaligned rows, or the tail of fewer than 4 floats)
    if (CPLX) {
        for (int64_t e = (done >> 1) + lane; 2 * e < L; e += 64) {
            const double xr = x[2 * e], xi = x[2 * e + 1];
#pragma unroll
            for (int r = 0; r < MV_ROWS; ++r) {
                const double ar = Ar[r][2 * e], ai = Ar[r][2 * e + 1];
                sr[r] += ar * xr - ai * xi;
                si[r] += ar * xi + ai * xr;
            }
        }
    } else {
        for (int64_t e = done + lane; e < L; e += 64) {
            const double xv = x[e];
#pragma unroll
            for (int r = 0; r < MV_ROWS; ++r) sr[r] += (double)Ar[r][e] * xv;
        }
    }
#pragma unroll
    for (int r = 0; r < MV_ROWS; ++r) {
        for (int off = 32; off > 0; off >>= 1) {
            sr[r] += __shfl_xor(sr[r], off);
            if (CPLX) si[r] += __shfl_xor(si[r], off);
        }
    }
    if (lane < nr) {
        // lane r writes row i0 + r (select with a uniform unrolled loop)
        double vr = 0.0, vi = 0.0;
#pragma unroll
        for (int r = 0; r < MV_ROWS; ++r)
            if (lane == r) {
                vr = sr[r];
                vi = si[r];
            }
        const int64_t i = i0 + lane;
        if (CPLX) {
            y[2 * i] = (VT)(b ? (double)b[2 * i] - vr : vr);
            y[2 * i + 1] = (VT)(b ? (double)b[2 * i + 1] - vi : vi);
        } else {
            y[i] = (VT)(b ? (double)b[i] - vr : vr);
        }
    }
}

constexpr int GM_THREADS = 256;
constexpr int GM_EPT = 4;  // elements per thread in the dot kernel
constexpr int GM_CHUNK = GM_THREADS * GM_EPT;

__device__ __forceinline__ double block_sum(double v, double* red)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int q = 0; q < GM_THREADS / 64; ++q) s += red[q];
    return s;  // valid in thread 0
}

// part[blk][j] = sum over the block's elements of conj(V[j, e]) w[e]   (j < nvec)
// part[blk][nvec] = sum |w[e]|^2                                         (re only)
template <bool CPLX, class VT>
__global__ __launch_bounds__(GM_THREADS) void k_gm_dots(const float* __restrict__ V, int64_t ldv,
                                                        int nvec, const VT* __restrict__ w, int64_t n,
                                                        double2* __restrict__ part)
{
    __shared__ double red[2][GM_THREADS / 64];
    const int64_t e0 = (int64_t)blockIdx.x * GM_CHUNK;
    VT wr[GM_EPT], wi[GM_EPT];
    int64_t idx[GM_EPT];
#pragma unroll
    for (int q = 0; q < GM_EPT; ++q) {
        idx[q] = e0 + threadIdx.x + (int64_t)q * GM_THREADS;
        const bool ok = idx[q] < n;
        wr[q] = ok ? (CPLX ? w[2 * idx[q]] : w[idx[q]]) : (VT)0;
        wi[q] = (ok && CPLX) ? w[2 * idx[q] + 1] : (VT)0;
    }
    double2* out = part + (int64_t)blockIdx.x * (nvec + 1);
    for (int j = 0; j <= nvec; ++j) {
        VT ar = 0, ai = 0;
        if (j < nvec) {
            const float* Vj = V + (int64_t)j * ldv;
#pragma unroll
            for (int q = 0; q < GM_EPT; ++q) {
                if (idx[q] < n) {
                    if (CPLX) {
                        const VT vr = Vj[2 * idx[q]], vi = Vj[2 * idx[q] + 1];
                        ar += vr * wr[q] + vi * wi[q];
                        ai += vr * wi[q] - vi * wr[q];
                    } else {
                        ar += (VT)Vj[idx[q]] * wr[q];
                    }
                }
            }
        } else {
#pragma unroll
            for (int q = 0; q < GM_EPT; ++q) ar += wr[q] * wr[q] + wi[q] * wi[q];
        }
        const double sr = block_sum((double)ar, red[0]);
        const double si = CPLX ? block_sum((double)ai, red[1]) : 0.0;
        if (threadIdx.x == 0) out[j] = make_double2(sr, si);
    }
}

// out[j] = sum_blk part[blk][j], one wave per j, fixed order.
__global__ __launch_bounds__(64) void k_gm_reduce(const double2* __restrict__ part, int nblk, int nout,
                                                  double2* __restrict__ out)
{
    const int j = blockIdx.x, lane = threadIdx.x;
    double sr = 0.0, si = 0.0;
    for (int b = lane; b < nblk; b += 64) {
        double2 p = part[(int64_t)b * nout + j];
        sr += p.x;
        si += p.y;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off);
        si += __shfl_xor(si, off);
    }
    if (lane == 0) out[j] = make_double2(sr, si);
}

constexpr int GM_MAXVEC = 512;

// w[e] += sign * sum_j V[j, e] c[j]   (coefficients rounded to VT)
template <bool CPLX, class VT>
__global__ __launch_bounds__(256) void k_gm_update(const float* __restrict__ V, int64_t ldv, int nvec,
                                                   const double2* __restrict__ c, double sign,
                                                   VT* __restrict__ w, int64_t n)
{
    __shared__ VT cr[GM_MAXVEC], ci[GM_MAXVEC];
    for (int j = threadIdx.x; j < nvec; j += blockDim.x) {
        double2 v = c[j];
        cr[j] = (VT)(sign * v.x);
        ci[j] = (VT)(sign * v.y);
    }
    __syncthreads();
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    if (CPLX) {
        VT ar = w[2 * e], ai = w[2 * e + 1];
        for (int j = 0; j < nvec; ++j) {
            const float* Vj = V + (int64_t)j * ldv;
            const VT vr = Vj[2 * e], vi = Vj[2 * e + 1];
            ar += vr * cr[j] - vi * ci[j];
            ai += vr * ci[j] + vi * cr[j];
        }
        w[2 * e] = ar;
        w[2 * e + 1] = ai;
    } else {
        VT acc = w[e];
        for (int j = 0; j < nvec; ++j) acc += (VT)V[(int64_t)j * ldv + e] * cr[j];
        w[e] = acc;
    }
}

// dst (FP32) = src / sqrt(nrm2[0].x)   (zero when the norm is zero)
template <class VT>
__global__ __launch_bounds__(256) void k_gm_scale(const VT* __restrict__ src, int64_t L,
                                                  const double2* __restrict__ nrm2, float* __restrict__ dst)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= L) return;
    const double q = nrm2[0].x;
    const double s = q > 0.0 ? 1.0 / sqrt(q) : 0.0;
    dst[e] = (float)((double)src[e] * s);
}

inline int64_t nblk_dots(int64_t n) { return (n + GM_CHUNK - 1) / GM_CHUNK; }

inline bool field_ok(int field) { return field == MBA_COMPLEX || field == MBA_REAL; }
inline bool prec_ok(int p) { return p == MBA_F32 || p == MBA_F64; }

}  // namespace

extern "C" {

int mba_mv32(int field, const void* a, int64_t rows, int64_t cols, int64_t lda, int vprec, const void* x,
             const void* b, void* y, mba_stream_t stream)
{
    if (!a || !x || !y) return MBA_ERR_ARG;
    if (!field_ok(field) || !prec_ok(vprec)) return MBA_ERR_ARG;
    if (rows <= 0 || cols <= 0) return MBA_ERR_SIZE;
    if (lda < cols) return MBA_ERR_ARG;
    const int64_t f = field == MBA_COMPLEX ? 2 : 1;
    const int64_t L = f * cols, ldf = f * lda;
    const int vec = (((uintptr_t)a & 15) == 0) && (ldf % 4 == 0);
    const int xvec = ((uintptr_t)x & 15) == 0;
    const int64_t waves = (rows + MV_ROWS - 1) / MV_ROWS;
    dim3 grid((unsigned)((waves * 64 + 255) / 256));
#define MBA_MV(C, VT)                                                                                   \
    hipLaunchKernelGGL((k_mv32<C, VT>), grid, dim3(256), 0, S(stream), (const float*)a, rows, L, ldf,   \
                       (const VT*)x, (const VT*)b, (VT*)y, vec, xvec)
    if (field == MBA_COMPLEX) {
        if (vprec == MBA_F32) MBA_MV(true, float); else MBA_MV(true, double);
    } else {
        if (vprec == MBA_F32) MBA_MV(false, float); else MBA_MV(false, double);
    }
#undef MBA_MV
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

size_t mba_gmres_workspace_bytes(int64_t n, int32_t max_vec)
{
    if (n <= 0 || max_vec < 0) return 0;
    return (size_t)nblk_dots(n) * (size_t)(max_vec + 1) * sizeof(double2);
}

int mba_gmres_dots(int field, const void* basis, int64_t ldb, int32_t nvec, int vprec, const void* w,
                   int64_t n, void* out, void* ws, size_t ws_bytes, mba_stream_t stream)
{
    if (!w || !out || !ws || (nvec > 0 && !basis)) return MBA_ERR_ARG;
    if (!field_ok(field) || !prec_ok(vprec)) return MBA_ERR_ARG;
    if (n <= 0 || nvec < 0 || nvec > GM_MAXVEC) return MBA_ERR_SIZE;
    if (nvec > 0 && ldb < n) return MBA_ERR_ARG;
    if (ws_bytes < mba_gmres_workspace_bytes(n, nvec)) return MBA_ERR_ALLOC;
    const int64_t f = field == MBA_COMPLEX ? 2 : 1;
    const int64_t nb = nblk_dots(n);
#define MBA_DOTS(C, VT)                                                                                  \
    hipLaunchKernelGGL((k_gm_dots<C, VT>), dim3((unsigned)nb), dim3(GM_THREADS), 0, S(stream),           \
                       (const float*)basis, f * ldb, nvec, (const VT*)w, n, (double2*)ws)
    if (field == MBA_COMPLEX) {
        if (vprec == MBA_F32) MBA_DOTS(true, float); else MBA_DOTS(true, double);
    } else {
        if (vprec == MBA_F32) MBA_DOTS(false, float); else MBA_DOTS(false, double);
    }
#undef MBA_DOTS
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gm_reduce, dim3((unsigned)(nvec + 1)), dim3(64), 0, S(stream),
                       (const double2*)ws, (int)nb, nvec + 1, (double2*)out);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_gmres_update(int field, const void* basis, int64_t ldb, int32_t nvec, const void* coef,
                     double sign, int vprec, void* w, int64_t n, mba_stream_t stream)
{
    if (!w || (nvec > 0 && (!basis || !coef))) return MBA_ERR_ARG;
    if (!field_ok(field) || !prec_ok(vprec)) return MBA_ERR_ARG;
    if (n <= 0 || nvec < 0 || nvec > GM_MAXVEC) return MBA_ERR_SIZE;
    if (nvec == 0) return MBA_OK;
    if (ldb < n) return MBA_ERR_ARG;
    const int64_t f = field == MBA_COMPLEX ? 2 : 1;
    dim3 grid((unsigned)((n + 255) / 256));
#define MBA_UPD(C, VT)                                                                                   \
    hipLaunchKernelGGL((k_gm_update<C, VT>), grid, dim3(256), 0, S(stream), (const float*)basis, f * ldb, \
                       nvec, (const double2*)coef, sign, (VT*)w, n)
    if (field == MBA_COMPLEX) {
        if (vprec == MBA_F32) MBA_UPD(true, float); else MBA_UPD(true, double);
    } else {
        if (vprec == MBA_F32) MBA_UPD(false, float); else MBA_UPD(false, double);
    }
#undef MBA_UPD
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_gmres_scale(int field, int vprec, const void* src, int64_t n, const void* nrm2, void* dst,
                    mba_stream_t stream)
{
    if (!src || !nrm2 || !dst) return MBA_ERR_ARG;
    if (!field_ok(field) || !prec_ok(vprec)) return MBA_ERR_ARG;
    if (n <= 0) return MBA_ERR_SIZE;
    const int64_t L = (field == MBA_COMPLEX ? 2 : 1) * n;
    dim3 grid((unsigned)((L + 255) / 256));
    if (vprec == MBA_F32)
        hipLaunchKernelGGL(k_gm_scale<float>, grid, dim3(256), 0, S(stream), (const float*)src, L,
                           (const double2*)nrm2, (float*)dst);
    else
        hipLaunchKernelGGL(k_gm_scale<double>, grid, dim3(256), 0, S(stream), (const double*)src, L,
                           (const double2*)nrm2, (float*)dst);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
