// Nearest-vertex velocity transfer after remeshing: the GPU step of
// remesh_and_transfer_velocity (src/utils/mesh_utils.py:800-815), which takes
// torch.cdist between the new and the old vertices, argmin over the old ones
// and gathers the velocities [B, N_old, 3] -> [B, N_new, 3].
//
// Here: nearest[i] = the old vertex j of smallest |p_i - q_j|^2 (ties: the
// smallest j, as argmin's first occurrence), the squared distance formed from
// coordinate differences in the input precision.  Workgroup = 256 new
// vertices (one per thread) x one slice of the old vertices, staged 256 at a
// time in LDS and read at one address by all threads.  With fewer than
// ~2048 workgroups from the points alone the old vertices are split into S
// slices; each slice leaves (distance, index) per point in a workspace and a
// second kernel keeps the best in slice order (strict <, so the smallest
// index wins ties and the result does not depend on timing), then gathers
// the velocities.
#include "mba_common.hpp"

namespace {

constexpr int NN_THREADS = 256;
constexpr int NN_TILE = 256;
constexpr int64_t NN_TARGET_WGS = 2048;

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }
inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

template <class T>
__global__ __launch_bounds__(NN_THREADS) void k_nearest(const T* __restrict__ p, int64_t np, const T* __restrict__ q,
                                                        int64_t nq, int64_t slice, T* __restrict__ best_d,
                                                        int32_t* __restrict__ best_j)
{
    __shared__ T tile[3 * NN_TILE];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * NN_THREADS + tid;
    const int64_t j_begin = (int64_t)blockIdx.y * slice;
    const int64_t j_end = (j_begin + slice) < nq ? (j_begin + slice) : nq;
    const bool live = i < np;
    const T px = live ? p[3 * i] : (T)0, py = live ? p[3 * i + 1] : (T)0, pz = live ? p[3 * i + 2] : (T)0;
    T bd = (T)INFINITY;
    int32_t bj = (int32_t)j_begin;
    for (int64_t j0 = j_begin; j0 < j_end; j0 += NN_TILE) {
        const int n = (int)((j_end - j0) < NN_TILE ? (j_end - j0) : NN_TILE);
        __syncthreads();  // the previous tile's readers are done
        if (tid < n) {
            tile[3 * tid] = q[3 * (j0 + tid)];
            tile[3 * tid + 1] = q[3 * (j0 + tid) + 1];
            tile[3 * tid + 2] = q[3 * (j0 + tid) + 2];
        }
        __syncthreads();
#pragma unroll 4
        for (int t = 0; t < n; ++t) {
            const T dx = px - tile[3 * t], dy = py - tile[3 * t + 1], dz = pz - tile[3 * t + 2];
            const T d = dx * dx + dy * dy + dz * dz;
            if (d < bd) {
                bd = d;
                bj = (int32_t)(j0 + t);
            }
        }
    }
    if (live) {
        best_d[(int64_t)blockIdx.y * np + i] = bd;
        best_j[(int64_t)blockIdx.y * np + i] = bj;
    }
}

// best over the slices in slice order, then new_vel[b][i] = vel[b][nearest[i]]
template <class T>
__global__ __launch_bounds__(256) void k_nearest_finish(const T* __restrict__ best_d,
                                                        const int32_t* __restrict__ best_j, int nslice, int64_t np,
                                                        const T* __restrict__ vel, int64_t nq, int batch,
                                                        int32_t* __restrict__ nearest, T* __restrict__ new_vel)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    T bd = best_d[i];
    int32_t bj = best_j[i];
    for (int s = 1; s < nslice; ++s) {
        const T d = best_d[(int64_t)s * np + i];
        if (d < bd) {
            bd = d;
            bj = best_j[(int64_t)s * np + i];
        }
    }
    nearest[i] = bj;
    for (int b = 0; b < batch; ++b)
        for (int d = 0; d < 3; ++d) new_vel[((int64_t)b * np + i) * 3 + d] = vel[((int64_t)b * nq + bj) * 3 + d];
}

void plan(int64_t np, int64_t nq, int& nslice, int64_t& slice)
{
    const int64_t wgs = (np + NN_THREADS - 1) / NN_THREADS;
    int64_t s = (NN_TARGET_WGS + wgs - 1) / wgs;
    const int64_t max_s = (nq + NN_TILE - 1) / NN_TILE;  // at least one tile per slice
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    slice = (nq + s - 1) / s;
    slice = (slice + NN_TILE - 1) / NN_TILE * NN_TILE;
    nslice = (int)((nq + slice - 1) / slice);
}

template <class T>
int nearest_transfer(const T* p, int64_t np, const T* q, int64_t nq, const T* vel, int batch, int32_t* nearest,
                     T* new_vel, void* ws, size_t ws_bytes, hipStream_t st)
{
    int nslice;
    int64_t slice;
    plan(np, nq, nslice, slice);
    const size_t need = al256((size_t)nslice * np * sizeof(T)) + al256((size_t)nslice * np * 4);
    if (ws_bytes < need) return MBA_ERR_ALLOC;
    T* bd = (T*)ws;
    int32_t* bj = (int32_t*)((char*)ws + al256((size_t)nslice * np * sizeof(T)));
    const unsigned gx = (unsigned)((np + NN_THREADS - 1) / NN_THREADS);
    hipLaunchKernelGGL(k_nearest<T>, dim3(gx, (unsigned)nslice), dim3(NN_THREADS), 0, st, p, np, q, nq, slice, bd, bj);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_nearest_finish<T>, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, st, (const T*)bd,
                       (const int32_t*)bj, nslice, np, vel, nq, batch, nearest, new_vel);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // namespace

extern "C" {

size_t mba_nearest_transfer_workspace_bytes(int prec, int64_t n_points, int64_t n_verts)
{
    if ((prec != MBA_F32 && prec != MBA_F64) || n_points <= 0 || n_verts <= 0) return 0;
    int nslice;
    int64_t slice;
    plan(n_points, n_verts, nslice, slice);
    return al256((size_t)nslice * n_points * (prec == MBA_F32 ? 4 : 8)) + al256((size_t)nslice * n_points * 4);
}

int mba_nearest_transfer(int prec, const void* points, int64_t n_points, const void* verts, int64_t n_verts,
                         const void* vel, int32_t batch, int32_t* nearest, void* new_vel, void* workspace,
                         size_t workspace_bytes, mba_stream_t stream)
{
    if (!points || !verts || !nearest || !workspace) return MBA_ERR_ARG;
    if (batch > 0 && (!vel || !new_vel)) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_points <= 0 || n_verts <= 0 || n_verts > INT32_MAX || batch < 0) return MBA_ERR_SIZE;
    if (prec == MBA_F32)
        return nearest_transfer<float>((const float*)points, n_points, (const float*)verts, n_verts,
                                       (const float*)vel, batch, nearest, (float*)new_vel, workspace,
                                       workspace_bytes, S(stream));
    return nearest_transfer<double>((const double*)points, n_points, (const double*)verts, n_verts,
                                    (const double*)vel, batch, nearest, (double*)new_vel, workspace, workspace_bytes,
                                    S(stream));
}

}  // extern "C"
