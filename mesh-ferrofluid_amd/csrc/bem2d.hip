// 2-D boundary elements on closed curves of line segments ("faces" [F, 2],
// vertices [V, 2]): the Galerkin single and double layers of the reference's
// 2-D solver (src/mesh_BEM/utils/integrations_2d.py:18-232), their packing to
// vertex space (src/utils/BEM_utils.py:314-357 with dim = 2) and the
// Dirichlet right-hand side (src/mesh_BEM/rhs_layer/rhs_layer_2d.py:66-130).
//
// Layer entry, row = test segment fy, column = source segment fx:
//   out[fy, fx, iy, ix] = A_y A_x sum_{a,b} w_a w_b J phi_iy(r1y) phi_ix(r1x) f(x, y)
//   x = (1 - r1x) X0 + r1x X1 on fx, y = (1 - r1y) Y0 + r1y Y1 on fy,
//   r1x = g_a, r1y = g_b, J = 1; for two segments sharing both vertices
//   (SAME_EDGE) r1y = g_a g_b and J = g_a (the reference's map of the
//   half square y < x).  f = G (single) or grad_y G . n_fy (double), with
//   LAPLACE:   G = (1 / 2 pi) log(1 / d),  grad_y G = (x - y) / (2 pi d^2);
//   HELMHOLTZ: the reference's large-argument Hankel form of the second kind,
//     H_m(z) = sqrt(2 / (pi z)) e^{-j (z - pi/4 - m pi/2)},
//     G = (Im H_0 - j Re H_0) / 4,  grad_y G = (Im H_1 - j Re H_1) k (x - y) / (4 d)
//   (BEM_utils.py:9-138).
//
// Kernel layout: one thread per source segment fx (x of the grid), the
// workgroup's BEM2D_ROWS test segments staged in LDS (y of the grid); every
// thread walks the Q x Q rule for each staged row and writes the nphi x nphi
// block of (fy, fx): consecutive threads write consecutive blocks.
#include <math.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace {

constexpr int BEM2D_ROWS = 16;
constexpr int BEM2D_QMAX = 12;

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

template <class T>
__device__ __forceinline__ T log_t(T x);
template <>
__device__ __forceinline__ float log_t<float>(float x) { return __logf(x); }
template <>
__device__ __forceinline__ double log_t<double>(double x) { return log(x); }

template <class T>
__device__ __forceinline__ T sqrt_t(T x);
template <>
__device__ __forceinline__ float sqrt_t<float>(float x) { return __fsqrt_rn(x); }
template <>
__device__ __forceinline__ double sqrt_t<double>(double x) { return sqrt(x); }

// (Im H_m - j Re H_m) / 4 for H_m(z) = s (cos(z - th) - j sin(z - th)),
// s = sqrt(2 / (pi z)), th = pi/4 + m pi/2: re = -s sin(z - th) / 4,
// im = -s cos(z - th) / 4.  The phase is taken in revolutions (cis_rev).
template <class T>
__device__ __forceinline__ void hankel_green(T z, int m, T& re, T& im)
{
    const T s = sqrt_t<T>((T)(2.0 / MBA_PI) / z);
    const T t = z * (T)(0.5 / MBA_PI) - (T)0.125 - (T)0.25 * (T)m;  // (z - th) / 2 pi
    T c, sn;
    cis_rev(t, c, sn);
    re = (T)-0.25 * s * sn;
    im = (T)-0.25 * s * c;
}

// KIND: 0 Laplace single, 1 Laplace double, 2 Helmholtz single, 3 Helmholtz double
template <class T, int KIND>
__global__ __launch_bounds__(256) void k_bem2d_layer(const T* __restrict__ verts, const int32_t* __restrict__ faces,
                                                     int64_t F, const T* __restrict__ gp, const T* __restrict__ gw,
                                                     int Q, T k, int nphi, T* __restrict__ out)
{
    constexpr bool CPLX = KIND >= 2;
    constexpr bool DOUBLE = (KIND & 1) != 0;
    __shared__ T ys[BEM2D_ROWS][7];  // Y0, Y1, normal, length
    __shared__ int32_t yi[BEM2D_ROWS][2];
    __shared__ T sg[2][BEM2D_QMAX];
    const int tid = threadIdx.x;
    const int64_t fy0 = (int64_t)blockIdx.y * BEM2D_ROWS;
    if (tid < Q) {
        sg[0][tid] = gp[tid];
        sg[1][tid] = gw[tid];
    }
    for (int t = tid; t < BEM2D_ROWS; t += blockDim.x) {
        const int64_t fy = fy0 + t;
        if (fy < F) {
            const int32_t a = faces[2 * fy], b = faces[2 * fy + 1];
            yi[t][0] = a;
            yi[t][1] = b;
            const T y0x = verts[2 * (int64_t)a], y0y = verts[2 * (int64_t)a + 1];
            const T y1x = verts[2 * (int64_t)b], y1y = verts[2 * (int64_t)b + 1];
            const T ex = y1x - y0x, ey = y1y - y0y;
            const T len = sqrt_t<T>(ex * ex + ey * ey);
            ys[t][0] = y0x;
            ys[t][1] = y0y;
            ys[t][2] = y1x;
            ys[t][3] = y1y;
            ys[t][4] = ey / len;
            ys[t][5] = -ex / len;
            ys[t][6] = len;
        }
    }
    __syncthreads();
    const int64_t fx = (int64_t)blockIdx.x * blockDim.x + tid;
    if (fx >= F) return;
    const int32_t xa = faces[2 * fx], xb = faces[2 * fx + 1];
    const T x0x = verts[2 * (int64_t)xa], x0y = verts[2 * (int64_t)xa + 1];
    const T x1x = verts[2 * (int64_t)xb], x1y = verts[2 * (int64_t)xb + 1];
    const T lx = sqrt_t<T>((x1x - x0x) * (x1x - x0x) + (x1y - x0y) * (x1y - x0y));
    const int rows = (int)((F - fy0) < BEM2D_ROWS ? (F - fy0) : BEM2D_ROWS);
    for (int t = 0; t < rows; ++t) {
        const int64_t fy = fy0 + t;
        const int shared = (yi[t][0] == xa) + (yi[t][0] == xb) + (yi[t][1] == xa) + (yi[t][1] == xb);
        const bool same_edge = shared == 2;
        // x - y = b0 + r1x ex - r1y ey (b0 = X0 - Y0): relative to the segment
        // origins, so coincident and touching segments lose no digits;
        // (x - y) . n_fy = b0 . n + r1x (ex x ey) / |ey|, the second term
        // exactly 0 for a segment with itself (ey . n_fy = 0 by construction)
        const T eyx = ys[t][2] - ys[t][0], eyy = ys[t][3] - ys[t][1];
        const T b0x = x0x - ys[t][0], b0y = x0y - ys[t][1];
        const T exx = x1x - x0x, exy = x1y - x0y;
        const T nx = ys[t][4], ny = ys[t][5];
        const T s0 = b0x * nx + b0y * ny;
        const T sx = (exx * eyy - exy * eyx) / ys[t][6];
        T acc[2][2][2] = {};  // [iy][ix][re/im]
        for (int b = 0; b < Q; ++b) {
            const T gb = sg[0][b], wb = sg[1][b];
            for (int a = 0; a < Q; ++a) {
                const T ga = sg[0][a];
                const T r1x = ga;
                const T r1y = same_edge ? ga * gb : gb;
                const T w = wb * sg[1][a] * (same_edge ? ga : (T)1);
                const T dx = b0x + r1x * exx - r1y * eyx, dy = b0y + r1x * exy - r1y * eyy;
                const T d2 = dx * dx + dy * dy;
                T fre, fim = 0;
                if (KIND == 0) {
                    fre = (T)(-0.25 / MBA_PI) * log_t<T>(d2);  // (1 / 2 pi) log(1 / d)
                } else if (KIND == 1) {
                    fre = (T)(0.5 / MBA_PI) * (s0 + r1x * sx) / d2;
                } else {
                    const T d = sqrt_t<T>(d2);
                    T hr, hi;
                    hankel_green<T>(k * d, DOUBLE ? 1 : 0, hr, hi);
                    if (DOUBLE) {
                        const T proj = k * (s0 + r1x * sx) / d;
                        fre = hr * proj;
                        fim = hi * proj;
                    } else {
                        fre = hr;
                        fim = hi;
                    }
                }
                fre *= w;
                fim *= w;
                if (nphi == 1) {
                    acc[0][0][0] += fre;
                    if (CPLX) acc[0][0][1] += fim;
                } else {
                    const T py[2] = {(T)1 - r1y, r1y}, px[2] = {(T)1 - r1x, r1x};
#pragma unroll
                    for (int iy = 0; iy < 2; ++iy)
#pragma unroll
                        for (int ix = 0; ix < 2; ++ix) {
                            const T p = py[iy] * px[ix];
                            acc[iy][ix][0] += fre * p;
                            if (CPLX) acc[iy][ix][1] += fim * p;
                        }
                }
            }
        }
        const T s = ys[t][6] * lx;
        const int64_t base = (fy * F + fx) * nphi * nphi;
        for (int iy = 0; iy < nphi; ++iy)
            for (int ix = 0; ix < nphi; ++ix) {
                const int64_t o = base + iy * nphi + ix;
                if (CPLX) {
                    out[2 * o] = s * acc[iy][ix][0];
                    out[2 * o + 1] = s * acc[iy][ix][1];
                } else {
                    out[o] = s * acc[iy][ix][0];
                }
            }
    }
}

// [F, F, nphi, nphi] -> [V, V]: A[vy, vx] = sum over the segment-vertex slots
// (2 f + t) of vy and vx, in slot order (vinc_ptr / vinc_idx: CSR vertex ->
// slots, ascending).  PLANAR (nphi == 1) spreads each entry as U / 4.  W =
// reals per entry (1 real, 2 complex).
template <class T, int W>
__global__ void k_bem2d_pack(const T* __restrict__ U, int64_t F, int nphi, const int64_t* __restrict__ vptr,
                             const int32_t* __restrict__ vidx, int64_t V, T* __restrict__ A)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= V * V) return;
    const int64_t vy = id / V, vx = id % V;
    T acc[W] = {};
    for (int64_t p = vptr[vy]; p < vptr[vy + 1]; ++p) {
        const int64_t sy = vidx[p], fy = sy / 2, iy = sy % 2;
        for (int64_t q = vptr[vx]; q < vptr[vx + 1]; ++q) {
            const int64_t sx = vidx[q], fx = sx / 2, ix = sx % 2;
            const int64_t o = nphi == 1 ? (fy * F + fx) : ((fy * F + fx) * 2 + iy) * 2 + ix;
#pragma unroll
            for (int c = 0; c < W; ++c) acc[c] += nphi == 1 ? U[W * o + c] / (T)4 : U[W * o + c];
        }
    }
#pragma unroll
    for (int c = 0; c < W; ++c) A[W * id + c] = acc[c];
}

// rhs[v] = c_m g[v] va[v] + c_k sum_vy Kt[vy, v] g[vy]  (W reals per value; g and Kt of the same kind)
template <class T, int W>
__global__ void k_bem2d_rhs(const T* __restrict__ Kt, int64_t V, const T* __restrict__ g, const T* __restrict__ va,
                            T cm, T ck, T* __restrict__ rhs)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    T ar = 0, ai = 0;
    if (ck != (T)0)
        for (int64_t y = 0; y < V; ++y) {
            if (W == 1) {
                ar += Kt[y * V + v] * g[y];
            } else {
                const T kr = Kt[2 * (y * V + v)], ki = Kt[2 * (y * V + v) + 1];
                const T gr = g[2 * y], gi = g[2 * y + 1];
                ar += kr * gr - ki * gi;
                ai += kr * gi + ki * gr;
            }
        }
    if (W == 1) {
        rhs[v] = cm * g[v] * va[v] + ck * ar;
    } else {
        rhs[2 * v] = cm * g[2 * v] * va[v] + ck * ar;
        rhs[2 * v + 1] = cm * g[2 * v + 1] * va[v] + ck * ai;
    }
}

template <class T>
int layer2d(int kind, const void* verts, const int32_t* faces, int64_t F, const void* gp, const void* gw, int Q,
            double k, int nphi, void* out, hipStream_t st)
{
    dim3 grid((unsigned)((F + 255) / 256), (unsigned)((F + BEM2D_ROWS - 1) / BEM2D_ROWS));
    const T* v = (const T*)verts;
    const T* p = (const T*)gp;
    const T* w = (const T*)gw;
    switch (kind) {
    case 0:
        hipLaunchKernelGGL((k_bem2d_layer<T, 0>), grid, dim3(256), 0, st, v, faces, F, p, w, Q, (T)k, nphi, (T*)out);
        break;
    case 1:
        hipLaunchKernelGGL((k_bem2d_layer<T, 1>), grid, dim3(256), 0, st, v, faces, F, p, w, Q, (T)k, nphi, (T*)out);
        break;
    case 2:
        hipLaunchKernelGGL((k_bem2d_layer<T, 2>), grid, dim3(256), 0, st, v, faces, F, p, w, Q, (T)k, nphi, (T*)out);
        break;
    default:
        hipLaunchKernelGGL((k_bem2d_layer<T, 3>), grid, dim3(256), 0, st, v, faces, F, p, w, Q, (T)k, nphi, (T*)out);
        break;
    }
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // namespace

extern "C" {

int mba_bem2d_layer(int prec, int bem_type, int layer, int order_type, const void* verts, const int32_t* faces,
                    int64_t n_verts, int64_t n_faces, const void* gp, const void* gw, int32_t nq, double wavenumber,
                    void* out, mba_stream_t stream)
{
    if (!verts || !faces || !gp || !gw || !out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (bem_type < 0 || bem_type > 2 || (layer != 1 && layer != 2) || (order_type != 0 && order_type != 1))
        return MBA_ERR_ARG;
    if (n_verts <= 0 || n_faces <= 0 || nq < 1 || nq > BEM2D_QMAX) return MBA_ERR_SIZE;
    if (bem_type != 0 && !(wavenumber > 0.0)) return MBA_ERR_ARG;
    const int kind = (bem_type == 0 ? 0 : 2) + (layer == 2 ? 1 : 0);
    const int nphi = order_type == 1 ? 2 : 1;
    if (prec == MBA_F32)
        return layer2d<float>(kind, verts, faces, n_faces, gp, gw, nq, wavenumber, nphi, out, S(stream));
    return layer2d<double>(kind, verts, faces, n_faces, gp, gw, nq, wavenumber, nphi, out, S(stream));
}

int mba_bem2d_pack(int prec, int is_complex, const void* unpacked, int64_t n_faces, int32_t nphi,
                   const int64_t* vinc_ptr, const int32_t* vinc_idx, int64_t n_verts, void* out, mba_stream_t stream)
{
    if (!unpacked || !vinc_ptr || !vinc_idx || !out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (nphi != 1 && nphi != 2) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_verts <= 0) return MBA_ERR_SIZE;
    dim3 grid((unsigned)((n_verts * n_verts + 255) / 256));
    hipStream_t st = S(stream);
    if (prec == MBA_F32) {
        if (is_complex)
            hipLaunchKernelGGL((k_bem2d_pack<float, 2>), grid, dim3(256), 0, st, (const float*)unpacked, n_faces, nphi,
                               vinc_ptr, vinc_idx, n_verts, (float*)out);
        else
            hipLaunchKernelGGL((k_bem2d_pack<float, 1>), grid, dim3(256), 0, st, (const float*)unpacked, n_faces, nphi,
                               vinc_ptr, vinc_idx, n_verts, (float*)out);
    } else {
        if (is_complex)
            hipLaunchKernelGGL((k_bem2d_pack<double, 2>), grid, dim3(256), 0, st, (const double*)unpacked, n_faces,
                               nphi, vinc_ptr, vinc_idx, n_verts, (double*)out);
        else
            hipLaunchKernelGGL((k_bem2d_pack<double, 1>), grid, dim3(256), 0, st, (const double*)unpacked, n_faces,
                               nphi, vinc_ptr, vinc_idx, n_verts, (double*)out);
    }
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_bem2d_dirichlet_rhs(int prec, int is_complex, const void* k_tight, int64_t n_verts, const void* g,
                            const void* vert_areas, double c_mass, double c_k, void* rhs, mba_stream_t stream)
{
    if (!g || !vert_areas || !rhs || (c_k != 0.0 && !k_tight)) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_verts <= 0) return MBA_ERR_SIZE;
    dim3 grid((unsigned)((n_verts + 255) / 256));
    hipStream_t st = S(stream);
    if (prec == MBA_F32) {
        if (is_complex)
            hipLaunchKernelGGL((k_bem2d_rhs<float, 2>), grid, dim3(256), 0, st, (const float*)k_tight, n_verts,
                               (const float*)g, (const float*)vert_areas, (float)c_mass, (float)c_k, (float*)rhs);
        else
            hipLaunchKernelGGL((k_bem2d_rhs<float, 1>), grid, dim3(256), 0, st, (const float*)k_tight, n_verts,
                               (const float*)g, (const float*)vert_areas, (float)c_mass, (float)c_k, (float*)rhs);
    } else {
        if (is_complex)
            hipLaunchKernelGGL((k_bem2d_rhs<double, 2>), grid, dim3(256), 0, st, (const double*)k_tight, n_verts,
                               (const double*)g, (const double*)vert_areas, c_mass, c_k, (double*)rhs);
        else
            hipLaunchKernelGGL((k_bem2d_rhs<double, 1>), grid, dim3(256), 0, st, (const double*)k_tight, n_verts,
                               (const double*)g, (const double*)vert_areas, c_mass, c_k, (double*)rhs);
    }
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
