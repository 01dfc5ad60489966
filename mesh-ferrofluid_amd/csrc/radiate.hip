// Scattered magnetic field radiated by the vertex currents of a MoM solution,
// sampled at arbitrary points (the post-processing of
// demo/demo_Helmholtz_MOM_3d.py:246-257):
//
//   H(p) = sum_v A_v  J_v x grad_y G(p, v),
//   grad_y G(p, v) = e^{-jkR} (1 + jkR) (p - v) / (4 pi R^3),   R = |p - v|,
//
// with pairs closer than 1e-6 left out (the singularity mask of gradG_y,
// src/utils/functions.py:844-868, whose wavenumber sign gives e^{-jkR}).
//
// Layout: workgroup = 256 points (one per thread) x one slice of the
// vertices.  Vertex tiles of 256 are staged in LDS as (position, A_v J_v / 4 pi)
// and read by every thread at the same address (broadcast).  Per pair: the
// offset, R^2, 1/R (rsqrt), the phase kR in revolutions (hardware sin/cos in
// FP32, the exact-reduction series in FP64), the complex factor
// e^{-jkR} (1 + jkR) / R^3, the cross product with the real offset and one
// complex multiply-add per component.  When the points alone give too few
// workgroups, the vertices are split into S slices whose partial sums go to a
// workspace and are added in slice order by a second kernel (the result does
// not depend on timing).

#include "mba_common.hpp"

namespace {

constexpr int RAD_THREADS = 256;
constexpr int RAD_TILE = 256;
constexpr int64_t RAD_TARGET_WGS = 2048;  // about 8 workgroups per CU

template <class T>
struct RadVert {
    T x, y, z;
    T jr[3], ji[3];  // A_v J_v / (4 pi)
};

template <class T>
__global__ __launch_bounds__(RAD_THREADS) void k_radiate_h(const T* __restrict__ pts, int64_t n_pts,
                                                           const T* __restrict__ verts,
                                                           const cpx<T>* __restrict__ js,
                                                           const T* __restrict__ areas, int64_t n_verts,
                                                           int64_t slice, double wavenumber,
                                                           cpx<T>* __restrict__ out, int64_t out_slice_stride)
{
    __shared__ RadVert<T> tile[RAD_TILE];
    const int tid = threadIdx.x;
    const int64_t p = (int64_t)blockIdx.x * RAD_THREADS + tid;
    const int64_t v_begin = (int64_t)blockIdx.y * slice;
    const int64_t v_end = (v_begin + slice) < n_verts ? (v_begin + slice) : n_verts;
    const bool live = p < n_pts;
    const T px = live ? pts[3 * p] : (T)0, py = live ? pts[3 * p + 1] : (T)0, pz = live ? pts[3 * p + 2] : (T)0;
    const T k_rev = (T)(wavenumber / (2.0 * MBA_PI));  // phase in revolutions per unit length
    const T kk = (T)wavenumber;
    T hr[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int64_t v0 = v_begin; v0 < v_end; v0 += RAD_TILE) {
        const int nv = (int)((v_end - v0) < RAD_TILE ? (v_end - v0) : RAD_TILE);
        __syncthreads();  // the previous tile's readers are done
        if (tid < nv) {
            const int64_t v = v0 + tid;
            RadVert<T> r;
            r.x = verts[3 * v];
            r.y = verts[3 * v + 1];
            r.z = verts[3 * v + 2];
            const T a = areas[v] * (T)MBA_INV4PI;
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const cpx<T> j = js[3 * v + d];
                r.jr[d] = a * j.re;
                r.ji[d] = a * j.im;
            }
            tile[tid] = r;
        }
        __syncthreads();
#pragma unroll 4
        for (int q = 0; q < nv; ++q) {
            const RadVert<T>& w = tile[q];
            const T dx = px - w.x, dy = py - w.y, dz = pz - w.z;
            const T r2 = dx * dx + dy * dy + dz * dz;
            // |p - v| < 1e-6: masked pair (1/R = 0 zeroes its terms)
            const T rinv = r2 < (T)1e-12 ? (T)0 : rsqrt_t(r2);
            const T R = r2 * rinv;
            T c, s;
            cis_rev(k_rev * R, c, s);  // e^{-jkR} = c - j s
            const T kr = kk * R;
            const T f = rinv * rinv * rinv;
            // e^{-jkR} (1 + jkR) / R^3
            const T gr = (c + kr * s) * f, gi = (kr * c - s) * f;
            // A J x (p - v), complex x real
            const T xr0 = w.jr[1] * dz - w.jr[2] * dy, xi0 = w.ji[1] * dz - w.ji[2] * dy;
            const T xr1 = w.jr[2] * dx - w.jr[0] * dz, xi1 = w.ji[2] * dx - w.ji[0] * dz;
            const T xr2 = w.jr[0] * dy - w.jr[1] * dx, xi2 = w.ji[0] * dy - w.ji[1] * dx;
            hr[0] += gr * xr0 - gi * xi0;
            hi[0] += gr * xi0 + gi * xr0;
            hr[1] += gr * xr1 - gi * xi1;
            hi[1] += gr * xi1 + gi * xr1;
            hr[2] += gr * xr2 - gi * xi2;
            hi[2] += gr * xi2 + gi * xr2;
        }
    }
    if (live) {
        cpx<T>* o = out + (int64_t)blockIdx.y * out_slice_stride + 3 * p;
#pragma unroll
        for (int d = 0; d < 3; ++d) store_cpx(o + d, hr[d], hi[d]);
    }
}

// out[i] = sum over slices s (in order) of part[s][i], i over 3 n_pts complex entries.
template <class T>
__global__ void k_radiate_sum(const cpx<T>* __restrict__ part, int64_t n, int n_slices, cpx<T>* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    T re = 0, im = 0;
    for (int s = 0; s < n_slices; ++s) {
        const cpx<T> v = part[(int64_t)s * n + i];
        re += v.re;
        im += v.im;
    }
    store_cpx(out + i, re, im);
}

// Vertex slices: enough workgroups to fill the chip, whole tiles per slice.
void rad_split(int64_t n_pts, int64_t n_verts, int64_t& n_slices, int64_t& slice)
{
    const int64_t pblocks = (n_pts + RAD_THREADS - 1) / RAD_THREADS;
    const int64_t tiles = (n_verts + RAD_TILE - 1) / RAD_TILE;
    int64_t s = (RAD_TARGET_WGS + pblocks - 1) / pblocks;
    if (s > tiles) s = tiles;
    if (s < 1) s = 1;
    const int64_t tiles_per = (tiles + s - 1) / s;
    slice = tiles_per * RAD_TILE;
    n_slices = (n_verts + slice - 1) / slice;
}

template <class T>
int radiate(const void* pts, int64_t n_pts, const void* verts, const void* js, const void* areas, int64_t n_verts,
            double k, void* out, void* ws, size_t ws_bytes, hipStream_t st)
{
    int64_t ns, slice;
    rad_split(n_pts, n_verts, ns, slice);
    const int64_t n = 3 * n_pts;
    const dim3 grid((unsigned)((n_pts + RAD_THREADS - 1) / RAD_THREADS), (unsigned)ns);
    if (ns == 1) {
        hipLaunchKernelGGL(k_radiate_h<T>, grid, dim3(RAD_THREADS), 0, st, (const T*)pts, n_pts, (const T*)verts,
                           (const cpx<T>*)js, (const T*)areas, n_verts, slice, k, (cpx<T>*)out, (int64_t)0);
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    }
    if (!ws || ws_bytes < (size_t)ns * n * sizeof(cpx<T>)) return MBA_ERR_ALLOC;
    hipLaunchKernelGGL(k_radiate_h<T>, grid, dim3(RAD_THREADS), 0, st, (const T*)pts, n_pts, (const T*)verts,
                       (const cpx<T>*)js, (const T*)areas, n_verts, slice, k, (cpx<T>*)ws, n);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_radiate_sum<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (const cpx<T>*)ws, n,
                       (int)ns, (cpx<T>*)out);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // namespace

extern "C" {

size_t mba_radiate_h_workspace_bytes(int prec, int64_t n_points, int64_t n_verts)
{
    if ((prec != MBA_F32 && prec != MBA_F64) || n_points <= 0 || n_verts <= 0) return 0;
    int64_t ns, slice;
    rad_split(n_points, n_verts, ns, slice);
    if (ns == 1) return 0;
    return (size_t)ns * 3 * (size_t)n_points * (prec == MBA_F32 ? 8 : 16);
}

int mba_radiate_h(int prec, const void* points, int64_t n_points, const void* verts, const void* vert_js,
                  const void* vert_areas, int64_t n_verts, double wavenumber, void* out, void* workspace,
                  size_t workspace_bytes, mba_stream_t stream)
{
    if (!points || !verts || !vert_js || !vert_areas || !out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_points <= 0 || n_verts <= 0) return MBA_ERR_SIZE;
    if ((n_points + RAD_THREADS - 1) / RAD_THREADS > 0x7fffffff) return MBA_ERR_SIZE;
    hipStream_t st = (hipStream_t)stream;
    if (prec == MBA_F32)
        return radiate<float>(points, n_points, verts, vert_js, vert_areas, n_verts, wavenumber, out, workspace,
                              workspace_bytes, st);
    return radiate<double>(points, n_points, verts, vert_js, vert_areas, n_verts, wavenumber, out, workspace,
                           workspace_bytes, st);
}

}  // extern "C"
