// Surface Helmholtz-decomposition sums of the mesh ferrofluid step
// (src/mesh_ferrofluid/mesh_Hemholtz_decomposition.py:52-143).  At a target
// point x and for each velocity field b:
//
//   u(x) = sum_f sum_q W_fq [ (n_f . u_fq) g + (n_f x u_fq) x g ],
//   g = grad_x G(x, y_fq) = (y_fq - x) / (4 pi |y_fq - x|^3)      (grad_G_y_3d,
//       src/utils/BEM_utils.py:141-172, called with x = y_fq, y = target),
//   W_fq = 2 A_f w_q r1_q                                          (PLANAR
//       single_integration_on_panels_3d, BEM_utils.py:254-310),
//   y_fq = (1 - r1) v0 + (r1 - r2) v1 + r2 v2, u_fq the same interpolation of
//       the vertex velocities (linear_interplate_from_unit_panel_to_general,
//       BEM_utils.py:210-251),
//
// i.e. the reference's u1 + u2 (its two sums are only used added).  The
// reference has no singularity mask here; Gauss points lie inside the faces.
//
// Layout: workgroup = 256 targets (one per thread) x one slice of the source
// points (face, quadrature point).  Source tiles of 256 are built in LDS from
// the face vertices and velocities (y, W / 4 pi, n . u, n x u) and read by all
// threads at one address.  Per pair: offset (3), R^2 (5), 1/R (rsqrt), 1/R^3
// scaled (3), n x u x d (9), (n . u) d + that (6), accumulate (6).  Source
// slices give enough workgroups; partial sums are added in slice order.

#include "mba_common.hpp"

namespace {

constexpr int HD_THREADS = 256;
constexpr int HD_TILE = 256;
constexpr int64_t HD_TARGET_WGS = 2048;

template <class T>
struct HdSrc {
    T y[3];
    T s;     // W / (4 pi)
    T a;     // n . u
    T c[3];  // n x u
};

template <class T>
__global__ __launch_bounds__(HD_THREADS) void k_hd_sum(const T* __restrict__ tgt, int64_t n_tgt,
                                                       const T* __restrict__ verts, const int32_t* __restrict__ faces,
                                                       const T* __restrict__ vel, int64_t n_verts,
                                                       const T* __restrict__ rule, int nq, int64_t n_src,
                                                       int64_t slice, T* __restrict__ out, int64_t out_slice_stride)
{
    __shared__ HdSrc<T> tile[HD_TILE];
    const int tid = threadIdx.x;
    const int64_t t = (int64_t)blockIdx.x * HD_THREADS + tid;
    const int64_t b = blockIdx.z;
    const T* vb = vel + b * n_verts * 3;
    const int64_t s_begin = (int64_t)blockIdx.y * slice;
    const int64_t s_end = (s_begin + slice) < n_src ? (s_begin + slice) : n_src;
    const bool live = t < n_tgt;
    const T x0 = live ? tgt[3 * t] : (T)0, x1 = live ? tgt[3 * t + 1] : (T)0, x2 = live ? tgt[3 * t + 2] : (T)0;
    T u0 = 0, u1 = 0, u2 = 0;
    for (int64_t s0 = s_begin; s0 < s_end; s0 += HD_TILE) {
        const int ns = (int)((s_end - s0) < HD_TILE ? (s_end - s0) : HD_TILE);
        __syncthreads();  // the previous tile's readers are done
        if (tid < ns) {
            const int64_t src = s0 + tid;
            const int64_t f = src / nq;
            const int q = (int)(src - f * nq);
            const T r1 = rule[3 * q], r2 = rule[3 * q + 1], w = rule[3 * q + 2];
            const int64_t i0 = faces[3 * f], i1 = faces[3 * f + 1], i2 = faces[3 * f + 2];
            const T c0 = (T)1 - r1, c1 = r1 - r2, c2 = r2;
            T p[3][3], e[3][3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                p[0][d] = verts[3 * i0 + d];
                p[1][d] = verts[3 * i1 + d];
                p[2][d] = verts[3 * i2 + d];
                e[0][d] = vb[3 * i0 + d];
                e[1][d] = vb[3 * i1 + d];
                e[2][d] = vb[3 * i2 + d];
            }
            // normal = normalise((v1 - v0) x (v2 - v1)), area = |.| / 2
            const T ax = p[1][0] - p[0][0], ay = p[1][1] - p[0][1], az = p[1][2] - p[0][2];
            const T bx = p[2][0] - p[1][0], by = p[2][1] - p[1][1], bz = p[2][2] - p[1][2];
            const T nx = ay * bz - az * by, ny = az * bx - ax * bz, nz = ax * by - ay * bx;
            const T nn = sqrt(nx * nx + ny * ny + nz * nz);
            const T area = (T)0.5 * nn;
            const T inv = (T)1 / nn;
            const T n0 = nx * inv, n1 = ny * inv, n2 = nz * inv;
            HdSrc<T> S;
            T uv[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                S.y[d] = c0 * p[0][d] + c1 * p[1][d] + c2 * p[2][d];
                uv[d] = c0 * e[0][d] + c1 * e[1][d] + c2 * e[2][d];
            }
            S.s = (T)2 * area * w * r1 * (T)MBA_INV4PI;
            S.a = n0 * uv[0] + n1 * uv[1] + n2 * uv[2];
            S.c[0] = n1 * uv[2] - n2 * uv[1];
            S.c[1] = n2 * uv[0] - n0 * uv[2];
            S.c[2] = n0 * uv[1] - n1 * uv[0];
            tile[tid] = S;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < ns; ++k) {
            const HdSrc<T>& S = tile[k];
            const T dx = S.y[0] - x0, dy = S.y[1] - x1, dz = S.y[2] - x2;
            const T r2 = dx * dx + dy * dy + dz * dz;
            const T rinv = rsqrt_t(r2);
            const T g = S.s * rinv * rinv * rinv;
            // (n . u) d + (n x u) x d, scaled by W / (4 pi R^3)
            const T v0 = S.a * dx + (S.c[1] * dz - S.c[2] * dy);
            const T v1 = S.a * dy + (S.c[2] * dx - S.c[0] * dz);
            const T v2 = S.a * dz + (S.c[0] * dy - S.c[1] * dx);
            u0 += g * v0;
            u1 += g * v1;
            u2 += g * v2;
        }
    }
    if (live) {
        T* o = out + (int64_t)blockIdx.y * out_slice_stride + (b * n_tgt + t) * 3;
        o[0] = u0;
        o[1] = u1;
        o[2] = u2;
    }
}

// out[i] = sum over slices s (in order) of part[s][i].
template <class T>
__global__ void k_hd_slices(const T* __restrict__ part, int64_t n, int n_slices, T* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T v = 0;
    for (int s = 0; s < n_slices; ++s) v += part[(int64_t)s * n + i];
    out[i] = v;
}

void hd_split(int64_t n_tgt, int64_t n_src, int64_t batch, int64_t& n_slices, int64_t& slice)
{
    const int64_t tblocks = ((n_tgt + HD_THREADS - 1) / HD_THREADS) * batch;
    const int64_t tiles = (n_src + HD_TILE - 1) / HD_TILE;
    int64_t s = (HD_TARGET_WGS + tblocks - 1) / tblocks;
    if (s > tiles) s = tiles;
    if (s < 1) s = 1;
    const int64_t per = (tiles + s - 1) / s;
    slice = per * HD_TILE;
    n_slices = (n_src + slice - 1) / slice;
}

template <class T>
int hd(const void* tgt, int64_t n_tgt, const void* verts, const int32_t* faces, int64_t n_verts, int64_t n_faces,
       const void* vel, int64_t batch, const void* rule, int nq, void* out, void* ws, size_t ws_bytes,
       hipStream_t st)
{
    const int64_t n_src = n_faces * nq;
    int64_t ns, slice;
    hd_split(n_tgt, n_src, batch, ns, slice);
    const int64_t n = batch * n_tgt * 3;
    const dim3 grid((unsigned)((n_tgt + HD_THREADS - 1) / HD_THREADS), (unsigned)ns, (unsigned)batch);
    if (ns == 1) {
        hipLaunchKernelGGL(k_hd_sum<T>, grid, dim3(HD_THREADS), 0, st, (const T*)tgt, n_tgt, (const T*)verts, faces,
                           (const T*)vel, n_verts, (const T*)rule, nq, n_src, slice, (T*)out, (int64_t)0);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    }
    if (!ws || ws_bytes < (size_t)ns * n * sizeof(T)) return MBA_ERR_ALLOC;
    hipLaunchKernelGGL(k_hd_sum<T>, grid, dim3(HD_THREADS), 0, st, (const T*)tgt, n_tgt, (const T*)verts, faces,
                       (const T*)vel, n_verts, (const T*)rule, nq, n_src, slice, (T*)ws, n);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_hd_slices<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const T*)ws, n, (int)ns,
                       (T*)out);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // namespace

extern "C" {

size_t mba_helmholtz_decomp_workspace_bytes(int prec, int64_t n_targets, int64_t n_faces, int32_t nq, int64_t batch)
{
    if ((prec != MBA_F32 && prec != MBA_F64) || n_targets <= 0 || n_faces <= 0 || nq <= 0 || batch <= 0) return 0;
    int64_t ns, slice;
    hd_split(n_targets, n_faces * nq, batch, ns, slice);
    if (ns == 1) return 0;
    return (size_t)ns * (size_t)(batch * n_targets * 3) * (prec == MBA_F32 ? 4 : 8);
}

int mba_helmholtz_decomp(int prec, const void* targets, int64_t n_targets, const void* verts, const int32_t* faces,
                         int64_t n_verts, int64_t n_faces, const void* vert_vel, int64_t batch, const void* rule,
                         int32_t nq, void* out, void* workspace, size_t workspace_bytes, mba_stream_t stream)
{
    if (!targets || !verts || !faces || !vert_vel || !rule || !out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_targets <= 0 || n_verts <= 0 || n_faces <= 0 || nq <= 0 || batch <= 0) return MBA_ERR_SIZE;
    if (batch > 65535 || (n_targets + HD_THREADS - 1) / HD_THREADS > 0x7fffffff) return MBA_ERR_SIZE;
    hipStream_t st = (hipStream_t)stream;
    if (prec == MBA_F32)
        return hd<float>(targets, n_targets, verts, faces, n_verts, n_faces, vert_vel, batch, rule, nq, out,
                         workspace, workspace_bytes, st);
    return hd<double>(targets, n_targets, verts, faces, n_verts, n_faces, vert_vel, batch, rule, nq, out, workspace,
                      workspace_bytes, st);
}

}  // extern "C"
