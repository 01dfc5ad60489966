// Mesh-graph network kernels for gfx950: BF16 MFMA GEMMs with gathered /
// concatenated operands and fused epilogues, deterministic segment sums, the
// relative-L2 loss and Adam.  Used for both inference and training.
//
// Numerics: every GEMM operand is rounded to BF16 (round to nearest even) when
// it is staged in LDS; products accumulate in FP32 on v_mfma_f32_32x32x8bf16;
// activations, gradients and the master weights stay FP32 in HBM.
#include <math.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace {

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t bf16_rne(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// Virtual row-major matrix: up to 3 column segments (optionally row-gathered)
// followed by an optional column of ones (the bias input).
struct VSeg {
    const float* p;
    const int32_t* idx;
    int64_t ld;
    int w;
};
struct VMat {
    VSeg s[3];
    int nseg;
    int ones;
};

__device__ __forceinline__ float vget(const VMat& V, int64_t r, int c)
{
    int off = 0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        if (q < V.nseg) {
            const VSeg& s = V.s[q];
            if (c < off + s.w) {
                const int64_t rr = s.idx ? (int64_t)s.idx[r] : r;
                return s.p[rr * s.ld + (c - off)];
            }
            off += s.w;
        }
    }
    return (V.ones && c == off) ? 1.f : 0.f;
}

struct GemmArgs {
    VMat A;
    int a_trans;  // Aop[m, k] = a_trans ? A(k, m) : A(m, k)
    const float* B;
    int64_t ldb;
    int b_trans;  // Bop[k, n] = b_trans ? B[n * ldb + k] : B[k * ldb + n]
    int64_t M, N, K;
    int64_t kchunk;     // split-K: block z covers [z kchunk, (z+1) kchunk)
    float* Y;
    int64_t ldy;
    int64_t ystride_z;  // partial outputs of split z at Y + z * ystride_z
    float* Z;           // optional: store the pre-activation
    const float* R;     // optional residual added after the activation
    const float* Zp;    // optional: multiply by silu'(Zp) before everything else
    int act;            // 0 none, 1 SiLU
    int accumulate;     // Y += v instead of Y = v
    int split;          // split-K: plain store of the partial sums
};

__device__ __forceinline__ float sigm(float z) { return 1.f / (1.f + expf(-z)); }

constexpr int GBM = 64, GBN = 128, GBK = 16, GPAD = 4;

__global__ __launch_bounds__(256) void k_mgn_gemm(GemmArgs g)
{
    __shared__ uint16_t As[GBM][GBK + GPAD];
    __shared__ uint16_t Bs[GBN][GBK + GPAD];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int64_t m0 = (int64_t)blockIdx.x * GBM, n0 = (int64_t)blockIdx.y * GBN;
    const int64_t kb = (int64_t)blockIdx.z * g.kchunk;
    const int64_t ke = (kb + g.kchunk < g.K) ? kb + g.kchunk : g.K;
    f16v acc[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[j][r] = 0.f;

    for (int64_t k0 = kb; k0 < ke; k0 += GBK) {
        // A tile: thread -> (m = tid / 4, k = 4 (tid % 4) .. +3)
        {
            const int mm = tid >> 2, kk = (tid & 3) * 4;
            const int64_t m = m0 + mm;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t k = k0 + kk + q;
                float v = 0.f;
                if (m < g.M && k < ke) v = g.a_trans ? vget(g.A, k, (int)m) : vget(g.A, m, (int)k);
                As[mm][kk + q] = bf16_rne(v);
            }
        }
        // B tile: thread -> (n = tid / 2, k = 8 (tid % 2) .. +7)
        {
            const int nn = tid >> 1, kk = (tid & 1) * 8;
            const int64_t n = n0 + nn;
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int64_t k = k0 + kk + q;
                float v = 0.f;
                if (n < g.N && k < ke) v = g.b_trans ? g.B[n * g.ldb + k] : g.B[k * g.ldb + n];
                Bs[nn][kk + q] = bf16_rne(v);
            }
        }
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < GBK; ks += 8) {
            const int kq = ks + 4 * (lane >> 5);
            s4v a;
            const uint16_t* ap = &As[wm * 32 + (lane & 31)][kq];
            a[0] = (short)ap[0];
            a[1] = (short)ap[1];
            a[2] = (short)ap[2];
            a[3] = (short)ap[3];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                s4v b;
                const uint16_t* bp = &Bs[wn * 64 + j * 32 + (lane & 31)][kq];
                b[0] = (short)bp[0];
                b[1] = (short)bp[1];
                b[2] = (short)bp[2];
                b[3] = (short)bp[3];
                acc[j] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc[j], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // epilogue: acc[j][r] -> row 8 (r / 4) + 4 (lane / 32) + r % 4, col lane % 32
    float* Y = g.Y + (int64_t)blockIdx.z * g.ystride_z;
    const bool split = g.split != 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int64_t m = m0 + wm * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int64_t n = n0 + wn * 64 + j * 32 + (lane & 31);
            if (m >= g.M || n >= g.N) continue;
            const int64_t o = m * g.ldy + n;
            float v = acc[j][r];
            if (split) {
                Y[o] = v;
                continue;
            }
            if (g.Zp) {
                const float z = g.Zp[o], s = sigm(z);
                v *= s * (1.f + z * (1.f - s));
            }
            if (g.Z) g.Z[o] = v;
            if (g.act == 1) v = v * sigm(v);
            if (g.R) v += g.R[o];
            Y[o] = g.accumulate ? Y[o] + v : v;
        }
    }
}

// Y[m, n] (=/+=) sum_z P[z][m, n], fixed order.
__global__ void k_mgn_split_reduce(const float* __restrict__ P, int nz, int64_t M, int64_t N, int64_t ldp,
                                   int64_t stride_z, float* __restrict__ Y, int64_t ldy, int accumulate)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= M * N) return;
    const int64_t m = id / N, n = id % N;
    float s = 0.f;
    for (int z = 0; z < nz; ++z) s += P[z * stride_z + m * ldp + n];
    Y[m * ldy + n] = accumulate ? Y[m * ldy + n] + s : s;
}

// out[i, c] (=/+=) sum_{q in [ptr[i], ptr[i+1])} in[idx[q] * ldin + coff + c]
__global__ void k_mgn_segsum(const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx, int64_t rows,
                             const float* __restrict__ in, int64_t ldin, int coff, int width,
                             float* __restrict__ out, int64_t ldout, int accumulate)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= rows * width) return;
    const int64_t i = id / width;
    const int c = (int)(id % width);
    float s = 0.f;
    for (int64_t q = ptr[i]; q < ptr[i + 1]; ++q) s += in[(int64_t)idx[q] * ldin + coff + c];
    float* o = out + i * ldout + c;
    *o = accumulate ? *o + s : s;
}

// Per graph g (rows [goff[g], goff[g+1]) of pred / target [*, 2]):
// loss[g] = |p - t|^2 / |t|^2, grad = 2 (p - t) / (|t|^2 G).
__global__ __launch_bounds__(256) void k_mgn_loss(const float* __restrict__ pred, const float* __restrict__ tgt,
                                                  const int64_t* __restrict__ goff, int G,
                                                  float* __restrict__ loss, float* __restrict__ grad)
{
    __shared__ double red[2][4];
    const int gi = blockIdx.x;
    const int64_t a = goff[gi], b = goff[gi + 1];
    double num = 0.0, den = 0.0;
    for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x)
        for (int c = 0; c < 2; ++c) {
            const double d = (double)pred[2 * i + c] - (double)tgt[2 * i + c];
            num += d * d;
            den += (double)tgt[2 * i + c] * (double)tgt[2 * i + c];
        }
    for (int off = 32; off > 0; off >>= 1) {
        num += __shfl_xor(num, off);
        den += __shfl_xor(den, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = num;
        red[1][w] = den;
    }
    __syncthreads();
    num = red[0][0] + red[0][1] + red[0][2] + red[0][3];
    den = red[1][0] + red[1][1] + red[1][2] + red[1][3];
    if (threadIdx.x == 0) loss[gi] = (float)(num / den);
    if (grad) {
        const double sc = 2.0 / (den * (double)G);
        for (int64_t i = a + threadIdx.x; i < b; i += blockDim.x)
            for (int c = 0; c < 2; ++c)
                grad[2 * i + c] = (float)(sc * ((double)pred[2 * i + c] - (double)tgt[2 * i + c]));
    }
}

__global__ void k_mgn_adam(float* __restrict__ p, const float* __restrict__ gr, float* __restrict__ m,
                           float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps, float c1,
                           float c2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.f - b1) * g;
    const float vi = b2 * v[i] + (1.f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

}  // namespace

extern "C" {

int mba_mgn_gemm(const mba_mgn_vmat* A, int a_trans, const float* B, int64_t ldb, int b_trans, int64_t M,
                 int64_t N, int64_t K, float* Y, int64_t ldy, float* Z, const float* R, const float* Zp, int act,
                 int accumulate, int32_t split_k, float* partial, mba_stream_t stream)
{
    if (!A || !B || !Y) return MBA_ERR_ARG;
    if (M <= 0 || N <= 0 || K <= 0) return MBA_ERR_SIZE;
    if (A->nseg < 0 || A->nseg > 3 || act < 0 || act > 1) return MBA_ERR_ARG;
    if (split_k < 1 || split_k > 4096) return MBA_ERR_SIZE;
    if (split_k > 1 && (!partial || Z || R || Zp || act)) return MBA_ERR_ARG;
    GemmArgs g;
    for (int q = 0; q < 3; ++q) {
        g.A.s[q].p = q < A->nseg ? A->p[q] : nullptr;
        g.A.s[q].idx = q < A->nseg ? A->idx[q] : nullptr;
        g.A.s[q].ld = q < A->nseg ? A->ld[q] : 0;
        g.A.s[q].w = q < A->nseg ? A->w[q] : 0;
        if (q < A->nseg && (!A->p[q] || A->w[q] <= 0 || A->ld[q] < A->w[q])) return MBA_ERR_ARG;
    }
    g.A.nseg = A->nseg;
    g.A.ones = A->ones;
    g.a_trans = a_trans;
    g.B = B;
    g.ldb = ldb;
    g.b_trans = b_trans;
    g.M = M;
    g.N = N;
    g.K = K;
    int64_t kchunk = (K + split_k - 1) / split_k;
    kchunk = ((kchunk + GBK - 1) / GBK) * GBK;
    const int nz = (int)((K + kchunk - 1) / kchunk);
    g.kchunk = kchunk;
    g.Y = split_k > 1 ? partial : Y;
    g.ldy = split_k > 1 ? N : ldy;
    g.ystride_z = M * N;
    g.Z = Z;
    g.R = R;
    g.Zp = Zp;
    g.act = act;
    g.accumulate = accumulate;
    g.split = split_k > 1;
    dim3 grid((unsigned)((M + GBM - 1) / GBM), (unsigned)((N + GBN - 1) / GBN), (unsigned)nz);
    hipLaunchKernelGGL(k_mgn_gemm, grid, dim3(256), 0, S(stream), g);
    MBA_LAUNCH_CHECK();
    if (split_k > 1) {
        const int64_t tot = M * N;
        hipLaunchKernelGGL(k_mgn_split_reduce, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, S(stream),
                           (const float*)partial, nz, M, N, N, M * N, Y, ldy, accumulate);
        MBA_LAUNCH_CHECK();
    }
    return MBA_OK;
}

int mba_mgn_segsum(const int64_t* ptr, const int32_t* idx, int64_t rows, const float* in, int64_t ldin,
                   int32_t coff, int32_t width, float* out, int64_t ldout, int accumulate, mba_stream_t stream)
{
    if (!ptr || !idx || !in || !out) return MBA_ERR_ARG;
    if (rows <= 0 || width <= 0) return MBA_ERR_SIZE;
    const int64_t tot = rows * width;
    hipLaunchKernelGGL(k_mgn_segsum, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, S(stream), ptr, idx, rows,
                       in, ldin, coff, width, out, ldout, accumulate);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_mgn_loss(const float* pred, const float* target, const int64_t* graph_off, int32_t n_graphs, float* loss,
                 float* grad, mba_stream_t stream)
{
    if (!pred || !target || !graph_off || !loss) return MBA_ERR_ARG;
    if (n_graphs <= 0) return MBA_ERR_SIZE;
    hipLaunchKernelGGL(k_mgn_loss, dim3((unsigned)n_graphs), dim3(256), 0, S(stream), pred, target, graph_off,
                       n_graphs, loss, grad);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_mgn_adam(float* param, const float* grad, float* m, float* v, int64_t n, float lr, float beta1,
                 float beta2, float eps, int64_t step, mba_stream_t stream)
{
    if (!param || !grad || !m || !v) return MBA_ERR_ARG;
    if (n <= 0 || step <= 0) return MBA_ERR_SIZE;
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mgn_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, S(stream), param, grad, m, v,
                       n, lr, beta1, beta2, eps, c1, c2);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
