// Loop subdivision of a triangle mesh on the host: the refinement step of
// remesh_and_transfer_velocity (src/utils/mesh_utils.py:728-815), which calls
// open3d's TriangleMesh.subdivide_loop.  open3d is not available here; this
// follows Loop's published scheme (C. Loop, "Smooth Subdivision Surfaces
// Based on Triangles", M.S. thesis, University of Utah, 1987):
//   edge vertex   interior edge (v0, v1) with opposite vertices v2, v3:
//                 3/8 (v0 + v1) + 1/8 (v2 + v3); any other edge: midpoint
//   old vertex    interior (every incident edge has two faces), valence n:
//                 (1 - n b) v + b sum(neighbours),
//                 b = (5/8 - (3/8 + cos(2 pi / n) / 4)^2) / n;
//                 on the boundary with exactly two boundary neighbours (edges
//                 of one face): 3/4 v + 1/8 (b0 + b1); otherwise (more
//                 boundary edges, or an edge of three or more faces) unchanged
//   faces         (a, b, c) -> (a, ab, ca), (ab, b, bc), (ca, bc, c), (ab, bc, ca)
// Output: the old vertices (moved) in their order, then one vertex per unique
// edge in ascending (min, max) order; face 4 f + i comes from face f.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <vector>

#include "../../include/mesh_bem_amd.h"
#include "mba_guard.hpp"

namespace {

struct FaceEdge {
    int64_t key;  // min * V + max
    int32_t face;
    int8_t e;  // edge (faces[f][e], faces[f][(e + 1) % 3])
};

}  // namespace

namespace {

int loop_subdivide(const double* verts, int64_t n_verts, const int32_t* faces, int64_t n_faces,
                   double* new_verts, int32_t* new_faces, int64_t* n_new_verts)
{
    if (!verts || !faces || !n_new_verts) return MBA_ERR_ARG;
    if (n_verts <= 0 || n_faces <= 0 || n_verts > INT32_MAX / 2 || 4 * n_faces > INT32_MAX) return MBA_ERR_SIZE;
    const int64_t V = n_verts, F = n_faces;
    std::vector<FaceEdge> fe((size_t)(3 * F));
    for (int64_t f = 0; f < F; ++f) {
        for (int e = 0; e < 3; ++e) {
            const int64_t a = faces[3 * f + e], b = faces[3 * f + (e + 1) % 3];
            if (a < 0 || a >= V || b < 0 || b >= V || a == b) return MBA_ERR_ARG;
            fe[(size_t)(3 * f + e)] = FaceEdge{(a < b ? a : b) * V + (a < b ? b : a), (int32_t)f, (int8_t)e};
        }
    }
    std::stable_sort(fe.begin(), fe.end(), [](const FaceEdge& x, const FaceEdge& y) { return x.key < y.key; });
    // unique edges: [first, first + count) of fe per edge
    std::vector<int64_t> first;
    first.reserve((size_t)(3 * F / 2 + 1));
    for (size_t i = 0; i < fe.size(); ++i)
        if (i == 0 || fe[i].key != fe[i - 1].key) first.push_back((int64_t)i);
    const int64_t E = (int64_t)first.size();
    first.push_back((int64_t)fe.size());
    *n_new_verts = V + E;
    if (!new_verts) return MBA_OK;  // size query
    if (!new_faces) return MBA_ERR_ARG;

    std::vector<int32_t> edge_of((size_t)(3 * F));  // face-edge -> unique edge
    // per vertex: incident edges, boundary edges (one face), other edges (three or more faces)
    std::vector<int32_t> degree((size_t)V, 0), bcount((size_t)V, 0), ncount((size_t)V, 0);
    std::vector<double> nsum((size_t)(3 * V), 0.0), bsum((size_t)(3 * V), 0.0);
    auto P = [&](int64_t v, int d) { return verts[3 * v + d]; };
    for (int64_t u = 0; u < E; ++u) {
        const int64_t i0 = first[(size_t)u], i1 = first[(size_t)u + 1];
        const int64_t key = fe[(size_t)i0].key;
        const int64_t a = key / V, b = key % V;
        for (int64_t i = i0; i < i1; ++i) edge_of[(size_t)(3 * fe[(size_t)i].face + fe[(size_t)i].e)] = (int32_t)u;
        double* out = new_verts + 3 * (V + u);
        if (i1 - i0 == 2) {
            const FaceEdge& f0 = fe[(size_t)i0];
            const FaceEdge& f1 = fe[(size_t)i0 + 1];
            const int64_t c0 = faces[3 * (int64_t)f0.face + (f0.e + 2) % 3];
            const int64_t c1 = faces[3 * (int64_t)f1.face + (f1.e + 2) % 3];
            for (int d = 0; d < 3; ++d) out[d] = 0.375 * (P(a, d) + P(b, d)) + 0.125 * (P(c0, d) + P(c1, d));
        } else {
            for (int d = 0; d < 3; ++d) out[d] = 0.5 * (P(a, d) + P(b, d));
        }
        ++degree[(size_t)a];
        ++degree[(size_t)b];
        for (int d = 0; d < 3; ++d) {
            nsum[(size_t)(3 * a + d)] += P(b, d);
            nsum[(size_t)(3 * b + d)] += P(a, d);
        }
        if (i1 - i0 == 1) {
            ++bcount[(size_t)a];
            ++bcount[(size_t)b];
            for (int d = 0; d < 3; ++d) {
                bsum[(size_t)(3 * a + d)] += P(b, d);
                bsum[(size_t)(3 * b + d)] += P(a, d);
            }
        } else if (i1 - i0 > 2) {
            ++ncount[(size_t)a];
            ++ncount[(size_t)b];
        }
    }
    for (int64_t v = 0; v < V; ++v) {
        double* out = new_verts + 3 * v;
        const int n = degree[(size_t)v];
        if (ncount[(size_t)v] > 0) {
            for (int d = 0; d < 3; ++d) out[d] = P(v, d);
        } else if (n > 0 && bcount[(size_t)v] == 0) {
            const double t = 0.375 + 0.25 * cos(2.0 * M_PI / n);
            const double beta = (0.625 - t * t) / n;
            for (int d = 0; d < 3; ++d) out[d] = (1.0 - n * beta) * P(v, d) + beta * nsum[(size_t)(3 * v + d)];
        } else if (bcount[(size_t)v] == 2) {
            for (int d = 0; d < 3; ++d) out[d] = 0.75 * P(v, d) + 0.125 * bsum[(size_t)(3 * v + d)];
        } else {
            for (int d = 0; d < 3; ++d) out[d] = P(v, d);
        }
    }
    for (int64_t f = 0; f < F; ++f) {
        const int32_t a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
        const int32_t ab = (int32_t)(V + edge_of[(size_t)(3 * f)]);
        const int32_t bc = (int32_t)(V + edge_of[(size_t)(3 * f + 1)]);
        const int32_t ca = (int32_t)(V + edge_of[(size_t)(3 * f + 2)]);
        const int32_t quad[4][3] = {{a, ab, ca}, {ab, b, bc}, {ca, bc, c}, {ab, bc, ca}};
        for (int q = 0; q < 4; ++q)
            for (int d = 0; d < 3; ++d) new_faces[12 * f + 3 * q + d] = quad[q][d];
    }
    return MBA_OK;
}

}  // namespace

extern "C" {

int mba_loop_subdivide(const double* verts, int64_t n_verts, const int32_t* faces, int64_t n_faces,
                       double* new_verts, int32_t* new_faces, int64_t* n_new_verts)
{
    return mba_guard([&] { return loop_subdivide(verts, n_verts, faces, n_faces, new_verts, new_faces, n_new_verts); });
}

}  // extern "C"
