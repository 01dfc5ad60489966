// Mesh tables on the GPU: near-pair CSR lists and a diagonal update.
//
// Near pairs (the criterion of the near-singular treatment): test face a is
// near source face b when |o_a - o_b| < eta (rho_a + rho_b), with o the
// centroid and rho the largest vertex distance from it, in FP64.  The CSR is
// over source faces b with ascending a.  Brute force over all face pairs in
// two passes (count, fill) around a single-block scan: F^2 distance tests,
// a few milliseconds at F = 40k, and no host round trip for the geometry.
#include <math.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace {

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

template <class T>
__global__ void k_face_spheres(const T* __restrict__ verts, const int32_t* __restrict__ faces, int64_t F,
                               double4* __restrict__ sph)
{
#pragma clang fp contract(off)
    // separate multiplies and adds (no FMA), so the centroid, radius and
    // distances round exactly as the same FP64 formulas evaluated on the host
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double v[3][3];
    for (int t = 0; t < 3; ++t)
        for (int a = 0; a < 3; ++a) v[t][a] = (double)verts[3 * (int64_t)faces[3 * f + t] + a];
    double o[3];
    for (int a = 0; a < 3; ++a) o[a] = (v[0][a] + v[1][a] + v[2][a]) / 3.0;
    double rho = 0.0;
    for (int t = 0; t < 3; ++t) {
        const double d0 = v[t][0] - o[0], d1 = v[t][1] - o[1], d2 = v[t][2] - o[2];
        rho = fmax(rho, sqrt(d0 * d0 + d1 * d1 + d2 * d2));
    }
    sph[f] = make_double4(o[0], o[1], o[2], rho);
}

constexpr int NEAR_TILE = 256;

// FILL = false: cnt[b] = number of near a;  FILL = true: idx[ptr[b] + c] = a.
template <bool FILL>
__global__ __launch_bounds__(NEAR_TILE) void k_near(const double4* __restrict__ sph, int64_t F, double eta,
                                                    int32_t* __restrict__ cnt, const int64_t* __restrict__ ptr,
                                                    int32_t* __restrict__ idx)
{
#pragma clang fp contract(off)
    __shared__ double4 tile[NEAR_TILE];
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double4 sb = b < F ? sph[b] : make_double4(0, 0, 0, 0);
    int32_t c = 0;
    int64_t out = (FILL && b < F) ? ptr[b] : 0;
    for (int64_t a0 = 0; a0 < F; a0 += NEAR_TILE) {
        __syncthreads();
        if (a0 + threadIdx.x < F) tile[threadIdx.x] = sph[a0 + threadIdx.x];
        __syncthreads();
        const int na = (int)((F - a0) < NEAR_TILE ? (F - a0) : NEAR_TILE);
        if (b < F) {
            for (int t = 0; t < na; ++t) {
                const double4 sa = tile[t];
                const double d0 = sa.x - sb.x, d1 = sa.y - sb.y, d2 = sa.z - sb.z;
                const double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                if (d < eta * (sa.w + sb.w)) {
                    if (FILL) idx[out + c] = (int32_t)(a0 + t);
                    ++c;
                }
            }
        }
    }
    if (!FILL && b < F) cnt[b] = c;
}

// ptr[0] = 0, ptr[b + 1] = ptr[b] + cnt[b]; one block, fixed order.
__global__ __launch_bounds__(1024) void k_scan(const int32_t* __restrict__ cnt, int64_t F,
                                               int64_t* __restrict__ ptr)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t chunk = (F + 1023) / 1024;
    const int64_t lo = t * chunk, hi = (lo + chunk < F) ? lo + chunk : F;
    int64_t s = 0;
    for (int64_t i = lo; i < hi; ++i) s += cnt[i];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        int64_t run = 0;
        for (int q = 0; q < 1024; ++q) {
            const int64_t v = part[q];
            part[q] = run;
            run += v;
        }
    }
    __syncthreads();
    int64_t run = part[t];
    if (t == 0) ptr[0] = 0;
    for (int64_t i = lo; i < hi; ++i) {
        run += cnt[i];
        ptr[i + 1] = run;
    }
}

template <class T>
__global__ void k_add_diag(T* __restrict__ a, int64_t n, int64_t lda, double alpha, const T* __restrict__ d)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    a[i * lda + i] += (T)(alpha * (double)d[i]);
}

}  // namespace

extern "C" {

size_t mba_near_workspace_bytes(int64_t n_faces)
{
    if (n_faces <= 0) return 0;
    return (size_t)n_faces * (sizeof(double4) + sizeof(int32_t)) + 64;
}

int mba_near_pairs_count(int prec, const void* verts, const int32_t* faces, int64_t n_verts, int64_t n_faces,
                         double eta, int64_t* near_ptr, void* workspace, size_t workspace_bytes,
                         mba_stream_t stream)
{
    if (!verts || !faces || !near_ptr || !workspace) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_verts <= 0 || n_faces > ((int64_t)1 << 30)) return MBA_ERR_SIZE;
    if (!(eta > 0.0)) return MBA_ERR_ARG;
    if (workspace_bytes < mba_near_workspace_bytes(n_faces)) return MBA_ERR_ALLOC;
    double4* sph = (double4*)workspace;
    int32_t* cnt = (int32_t*)(sph + n_faces);
    const unsigned g = (unsigned)((n_faces + 255) / 256);
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_face_spheres<float>, dim3(g), dim3(256), 0, S(stream), (const float*)verts, faces,
                           n_faces, sph);
    else
        hipLaunchKernelGGL(k_face_spheres<double>, dim3(g), dim3(256), 0, S(stream), (const double*)verts,
                           faces, n_faces, sph);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_near<false>, dim3((unsigned)((n_faces + NEAR_TILE - 1) / NEAR_TILE)), dim3(NEAR_TILE),
                       0, S(stream), (const double4*)sph, n_faces, eta, cnt, (const int64_t*)nullptr,
                       (int32_t*)nullptr);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, S(stream), (const int32_t*)cnt, n_faces, near_ptr);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_near_pairs_fill(int64_t n_faces, double eta, const int64_t* near_ptr, int32_t* near_idx,
                        const void* workspace, size_t workspace_bytes, mba_stream_t stream)
{
    if (!near_ptr || !near_idx || !workspace) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_faces > ((int64_t)1 << 30)) return MBA_ERR_SIZE;
    if (!(eta > 0.0)) return MBA_ERR_ARG;
    if (workspace_bytes < mba_near_workspace_bytes(n_faces)) return MBA_ERR_ALLOC;
    const double4* sph = (const double4*)workspace;
    hipLaunchKernelGGL(k_near<true>, dim3((unsigned)((n_faces + NEAR_TILE - 1) / NEAR_TILE)), dim3(NEAR_TILE),
                       0, S(stream), sph, n_faces, eta, (int32_t*)nullptr, near_ptr, near_idx);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_add_diag(int prec, void* a, int64_t n, int64_t lda, double alpha, const void* d, mba_stream_t stream)
{
    if (!a || !d) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n <= 0) return MBA_ERR_SIZE;
    if (lda < n) return MBA_ERR_ARG;
    const unsigned g = (unsigned)((n + 255) / 256);
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_add_diag<float>, dim3(g), dim3(256), 0, S(stream), (float*)a, n, lda, alpha,
                           (const float*)d);
    else
        hipLaunchKernelGGL(k_add_diag<double>, dim3(g), dim3(256), 0, S(stream), (double*)a, n, lda, alpha,
                           (const double*)d);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
