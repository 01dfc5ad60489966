// Unique edges of a triangle mesh on the GPU (replaces get_edge_unique,
// src/utils/mesh_utils.py:921-977).
//
// Local edge e of face f joins faces[f][(e+1)%3] and faces[f][(e+2)%3]
// (mesh_utils.py:940-943); its key is max * V + min (mesh_utils.py:952-954).
// The reference takes torch.unique(sorted=False), whose edge order depends on
// the backend; here the order is fixed: edges ascend by key.
//   1. keys[3f+e] and slots 3f+e
//   2. stable LSD radix sort of (key, slot) pairs (hipCUB), only the bits a
//      key can use
//   3. head flags (first slot of each key) and their inclusive scan: the edge
//      index of every sorted slot
//   4. face_edge_indices[slot] = edge index; at heads: edges[id] = (max, min),
//      adjacent faces = the faces of the first and second slot of the key
//      (ascending, as the sort is stable), -1 for a boundary edge.
#include <stdint.h>

#include <hipcub/hipcub.hpp>

#include "mba_common.hpp"

namespace {

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }
inline size_t al256(size_t x) { return (x + 255) & ~(size_t)255; }

__global__ void k_edge_keys(const int32_t* __restrict__ faces, int64_t F, int64_t V, int64_t* __restrict__ keys,
                            int32_t* __restrict__ slots)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * F) return;
    const int64_t f = i / 3;
    const int e = (int)(i % 3);
    const int64_t a = faces[3 * f + (e + 1) % 3], b = faces[3 * f + (e + 2) % 3];
    keys[i] = (a > b ? a : b) * V + (a > b ? b : a);
    slots[i] = (int32_t)i;
}

__global__ void k_edge_heads(const int64_t* __restrict__ keys, int64_t n, int32_t* __restrict__ head)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_edge_fill(const int64_t* __restrict__ keys, const int32_t* __restrict__ slots,
                            const int32_t* __restrict__ scan, int64_t n, int64_t V, int32_t* __restrict__ edges,
                            int32_t* __restrict__ fe_idx, int32_t* __restrict__ adj, int64_t* __restrict__ n_edges)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t id = scan[i] - 1;
    fe_idx[slots[i]] = id;
    if (i == 0 || keys[i] != keys[i - 1]) {
        edges[2 * (int64_t)id] = (int32_t)(keys[i] / V);
        edges[2 * (int64_t)id + 1] = (int32_t)(keys[i] % V);
        adj[2 * (int64_t)id] = slots[i] / 3;
        adj[2 * (int64_t)id + 1] = (i + 1 < n && keys[i + 1] == keys[i]) ? slots[i + 1] / 3 : -1;
    }
    if (i == n - 1) *n_edges = scan[i];
}

int end_bit_of(int64_t V)
{
    const unsigned long long top = (unsigned long long)V * (unsigned long long)V;
    int b = 1;
    while (b < 63 && (1ull << b) < top) ++b;
    return b;
}

struct EdgeWs {
    int64_t *k0, *k1;
    int32_t *s0, *s1, *head, *scan;
    void* tmp;
    size_t tmp_bytes, total;
};

EdgeWs edge_ws(int64_t F, int64_t V, char* base, size_t sort_bytes, size_t scan_bytes)
{
    const size_t n = (size_t)(3 * F);
    EdgeWs w;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char* p = base ? base + off : nullptr;
        off += al256(bytes);
        return p;
    };
    w.k0 = (int64_t*)take(n * 8);
    w.k1 = (int64_t*)take(n * 8);
    w.s0 = (int32_t*)take(n * 4);
    w.s1 = (int32_t*)take(n * 4);
    w.head = (int32_t*)take(n * 4);
    w.scan = (int32_t*)take(n * 4);
    w.tmp_bytes = sort_bytes > scan_bytes ? sort_bytes : scan_bytes;
    w.tmp = take(w.tmp_bytes);
    w.total = off;
    (void)V;
    return w;
}

bool temp_sizes(int64_t F, int64_t V, size_t& sort_bytes, size_t& scan_bytes)
{
    const int n = (int)(3 * F);
    sort_bytes = scan_bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, (int64_t*)nullptr, (int64_t*)nullptr,
                                           (int32_t*)nullptr, (int32_t*)nullptr, n, 0, end_bit_of(V)) != hipSuccess)
        return false;
    if (hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (int32_t*)nullptr, (int32_t*)nullptr, n) != hipSuccess)
        return false;
    return true;
}

}  // namespace

extern "C" {

size_t mba_edge_unique_workspace_bytes(int64_t n_faces, int64_t n_verts)
{
    if (n_faces <= 0 || n_verts <= 0 || 3 * n_faces > INT32_MAX) return 0;
    size_t sb, cb;
    if (!temp_sizes(n_faces, n_verts, sb, cb)) return 0;
    return edge_ws(n_faces, n_verts, nullptr, sb, cb).total;
}

int mba_edge_unique(const int32_t* faces, int64_t n_faces, int64_t n_verts, int32_t* edges,
                    int32_t* face_edge_indices, int32_t* edge_adjacent_faces, int64_t* n_edges, void* workspace,
                    size_t workspace_bytes, mba_stream_t stream)
{
    if (!faces || !edges || !face_edge_indices || !edge_adjacent_faces || !n_edges || !workspace) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_verts <= 0 || 3 * n_faces > INT32_MAX || n_verts > INT32_MAX) return MBA_ERR_SIZE;
    size_t sb, cb;
    if (!temp_sizes(n_faces, n_verts, sb, cb)) return MBA_ERR_LAUNCH;
    EdgeWs w = edge_ws(n_faces, n_verts, (char*)workspace, sb, cb);
    if (workspace_bytes < w.total) return MBA_ERR_ALLOC;
    hipStream_t st = S(stream);
    const int64_t n = 3 * n_faces;
    const unsigned grid = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_edge_keys, dim3(grid), dim3(256), 0, st, faces, n_faces, n_verts, w.k0, w.s0);
    MBA_LAUNCH_CHECK();
    size_t tb = w.tmp_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(w.tmp, tb, w.k0, w.k1, w.s0, w.s1, (int)n, 0, end_bit_of(n_verts), st) !=
        hipSuccess)
        return MBA_ERR_LAUNCH;
    hipLaunchKernelGGL(k_edge_heads, dim3(grid), dim3(256), 0, st, w.k1, n, w.head);
    MBA_LAUNCH_CHECK();
    tb = w.tmp_bytes;
    if (hipcub::DeviceScan::InclusiveSum(w.tmp, tb, w.head, w.scan, (int)n, st) != hipSuccess) return MBA_ERR_LAUNCH;
    hipLaunchKernelGGL(k_edge_fill, dim3(grid), dim3(256), 0, st, w.k1, w.s1, w.scan, n, n_verts, edges,
                       face_edge_indices, edge_adjacent_faces, n_edges);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
