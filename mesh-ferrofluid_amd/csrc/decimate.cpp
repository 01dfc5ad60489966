// Quadric edge-collapse decimation on the host: the simplify_quadric_decimation
// step of remesh_and_transfer_velocity (src/utils/mesh_utils.py:793-794), which
// calls open3d.  open3d is not available here; this follows Garland and
// Heckbert, "Surface Simplification Using Quadric Error Metrics" (SIGGRAPH
// 1997):
//   - each face contributes the quadric K = p p^T of its plane
//     p = (n, -n.v0) (unit normal) to its three vertices;
//   - an edge (a, b) collapses to the point v minimising v^T (Q_a + Q_b) v
//     (the 3x3 system; when it is singular, the best of a, b and the
//     midpoint), at that cost;
//   - the cheapest edge collapses first (ties: smaller vertex pair), b merges
//     into a, Q_a += Q_b, and the edges around a are re-queued;
//   - a collapse is skipped when it would break the 2-manifold (the link
//     condition: a and b share exactly the two vertices opposite the edge),
//     flip a face (normal turns by more than 90 degrees) or leave a degenerate
//     face; vertices on a boundary (an edge of one face) and edges of three or
//     more faces are left alone.
// Stops when at most target_faces faces remain or no edge can collapse.
// Output: surviving vertices and faces in their original order, renumbered.
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include <algorithm>
#include <queue>
#include <unordered_map>
#include <vector>

#include "../../include/mesh_bem_amd.h"
#include "mba_guard.hpp"

namespace {

struct Quad {
    double q[10] = {0};  // a2 ab ac ad b2 bc bd c2 cd d2
    void add_plane(double a, double b, double c, double d)
    {
        q[0] += a * a; q[1] += a * b; q[2] += a * c; q[3] += a * d;
        q[4] += b * b; q[5] += b * c; q[6] += b * d;
        q[7] += c * c; q[8] += c * d;
        q[9] += d * d;
    }
    void add(const Quad& o)
    {
        for (int i = 0; i < 10; ++i) q[i] += o.q[i];
    }
    double eval(const double* v) const
    {
        const double x = v[0], y = v[1], z = v[2];
        return q[0] * x * x + 2 * q[1] * x * y + 2 * q[2] * x * z + 2 * q[3] * x + q[4] * y * y + 2 * q[5] * y * z +
               2 * q[6] * y + q[7] * z * z + 2 * q[8] * z + q[9];
    }
};

struct Cand {
    double cost;
    int32_t a, b;  // a < b
    uint32_t sa, sb;  // vertex versions when queued
    double p[3];
};
struct CandCmp {
    bool operator()(const Cand& x, const Cand& y) const
    {
        if (x.cost != y.cost) return x.cost > y.cost;
        if (x.a != y.a) return x.a > y.a;
        return x.b > y.b;
    }
};

void cross3(const double* u, const double* v, double* w)
{
    w[0] = u[1] * v[2] - u[2] * v[1];
    w[1] = u[2] * v[0] - u[0] * v[2];
    w[2] = u[0] * v[1] - u[1] * v[0];
}

}  // namespace

namespace {

int decimate_quadric(const double* verts, int64_t n_verts, const int32_t* faces, int64_t n_faces,
                       int64_t target_faces, double* out_verts, int32_t* out_faces, int64_t* n_out_verts,
                       int64_t* n_out_faces)
{
    if (!verts || !faces || !out_verts || !out_faces || !n_out_verts || !n_out_faces) return MBA_ERR_ARG;
    if (n_verts <= 0 || n_faces <= 0 || n_verts > INT32_MAX || n_faces > INT32_MAX) return MBA_ERR_SIZE;
    const int64_t V = n_verts, F = n_faces;
    std::vector<double> P(verts, verts + 3 * V);
    std::vector<int32_t> Fv(faces, faces + 3 * F);
    for (int64_t i = 0; i < 3 * F; ++i)
        if (Fv[i] < 0 || Fv[i] >= V) return MBA_ERR_ARG;
    for (int64_t f = 0; f < F; ++f)
        if (Fv[3 * f] == Fv[3 * f + 1] || Fv[3 * f + 1] == Fv[3 * f + 2] || Fv[3 * f] == Fv[3 * f + 2])
            return MBA_ERR_ARG;
    std::vector<char> falive((size_t)F, 1), valive((size_t)V, 1), locked((size_t)V, 0);
    std::vector<uint32_t> ver((size_t)V, 0);
    std::vector<std::vector<int32_t>> vf((size_t)V);
    std::vector<Quad> Q((size_t)V);
    for (int64_t f = 0; f < F; ++f) {
        const double* a = &P[3 * Fv[3 * f]];
        const double* b = &P[3 * Fv[3 * f + 1]];
        const double* c = &P[3 * Fv[3 * f + 2]];
        double e1[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]}, e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
        double n[3];
        cross3(e1, e2, n);
        const double len = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (len > 0) {
            n[0] /= len; n[1] /= len; n[2] /= len;
            const double d = -(n[0] * a[0] + n[1] * a[1] + n[2] * a[2]);
            for (int k = 0; k < 3; ++k) Q[(size_t)Fv[3 * f + k]].add_plane(n[0], n[1], n[2], d);
        }
        for (int k = 0; k < 3; ++k) vf[(size_t)Fv[3 * f + k]].push_back((int32_t)f);
    }
    // edge face counts: lock vertices on boundary / non-manifold edges
    {
        std::unordered_map<int64_t, int> cnt;
        cnt.reserve((size_t)(3 * F));
        for (int64_t f = 0; f < F; ++f)
            for (int k = 0; k < 3; ++k) {
                const int64_t a = Fv[3 * f + k], b = Fv[3 * f + (k + 1) % 3];
                ++cnt[(a < b ? a : b) * V + (a < b ? b : a)];
            }
        for (const auto& kv : cnt)
            if (kv.second != 2) {
                locked[(size_t)(kv.first / V)] = 1;
                locked[(size_t)(kv.first % V)] = 1;
            }
    }
    auto neighbours = [&](int32_t v, std::vector<int32_t>& out) {
        out.clear();
        for (int32_t f : vf[(size_t)v]) {
            if (!falive[(size_t)f]) continue;
            for (int k = 0; k < 3; ++k) {
                const int32_t w = Fv[3 * (int64_t)f + k];
                if (w != v && std::find(out.begin(), out.end(), w) == out.end()) out.push_back(w);
            }
        }
    };
    auto make = [&](int32_t a, int32_t b, Cand& c) {
        if (a > b) std::swap(a, b);
        c.a = a;
        c.b = b;
        c.sa = ver[(size_t)a];
        c.sb = ver[(size_t)b];
        Quad q = Q[(size_t)a];
        q.add(Q[(size_t)b]);
        // minimise v^T q v: A v = -b with A = [[a2 ab ac][ab b2 bc][ac bc c2]]
        const double A00 = q.q[0], A01 = q.q[1], A02 = q.q[2], A11 = q.q[4], A12 = q.q[5], A22 = q.q[7];
        const double r0 = -q.q[3], r1 = -q.q[6], r2 = -q.q[8];
        const double det = A00 * (A11 * A22 - A12 * A12) - A01 * (A01 * A22 - A12 * A02) + A02 * (A01 * A12 - A11 * A02);
        const double scale = fabs(A00) + fabs(A11) + fabs(A22);
        bool ok = false;
        if (scale > 0 && fabs(det) > 1e-10 * scale * scale * scale) {
            c.p[0] = (r0 * (A11 * A22 - A12 * A12) - A01 * (r1 * A22 - A12 * r2) + A02 * (r1 * A12 - A11 * r2)) / det;
            c.p[1] = (A00 * (r1 * A22 - A12 * r2) - r0 * (A01 * A22 - A12 * A02) + A02 * (A01 * r2 - r1 * A02)) / det;
            c.p[2] = (A00 * (A11 * r2 - r1 * A12) - A01 * (A01 * r2 - r1 * A02) + r0 * (A01 * A12 - A11 * A02)) / det;
            ok = std::isfinite(c.p[0]) && std::isfinite(c.p[1]) && std::isfinite(c.p[2]);
        }
        if (ok) {
            c.cost = q.eval(c.p);
        } else {
            const double* pa = &P[3 * (int64_t)a];
            const double* pb = &P[3 * (int64_t)b];
            const double mid[3] = {0.5 * (pa[0] + pb[0]), 0.5 * (pa[1] + pb[1]), 0.5 * (pa[2] + pb[2])};
            const double* opts[3] = {pa, pb, mid};
            double best = INFINITY;
            for (int i = 0; i < 3; ++i) {
                const double e = q.eval(opts[i]);
                if (e < best) {
                    best = e;
                    c.p[0] = opts[i][0];
                    c.p[1] = opts[i][1];
                    c.p[2] = opts[i][2];
                }
            }
            c.cost = best;
        }
        if (c.cost < 0) c.cost = 0;  // rounding
    };
    std::priority_queue<Cand, std::vector<Cand>, CandCmp> pq;
    {
        std::vector<int32_t> nb;
        for (int32_t v = 0; v < (int32_t)V; ++v) {
            if (locked[(size_t)v]) continue;
            neighbours(v, nb);
            for (int32_t w : nb)
                if (w > v && !locked[(size_t)w]) {
                    Cand c;
                    make(v, w, c);
                    pq.push(c);
                }
        }
    }
    int64_t live_faces = F;
    std::vector<int32_t> na, nbv, shared;
    while (live_faces > target_faces && !pq.empty()) {
        const Cand c = pq.top();
        pq.pop();
        const int32_t a = c.a, b = c.b;
        if (!valive[(size_t)a] || !valive[(size_t)b] || ver[(size_t)a] != c.sa || ver[(size_t)b] != c.sb) continue;
        // link condition
        neighbours(a, na);
        neighbours(b, nbv);
        if (std::find(na.begin(), na.end(), b) == na.end()) continue;
        shared.clear();
        for (int32_t w : na)
            if (std::find(nbv.begin(), nbv.end(), w) != nbv.end()) shared.push_back(w);
        int n_edge_faces = 0;
        for (int32_t f : vf[(size_t)a]) {
            if (!falive[(size_t)f]) continue;
            const int32_t* t = &Fv[3 * (int64_t)f];
            if (t[0] == b || t[1] == b || t[2] == b) ++n_edge_faces;
        }
        if (n_edge_faces != 2 || shared.size() != 2) continue;
        // faces that survive: no flip, no degeneracy
        bool bad = false;
        for (int side = 0; side < 2 && !bad; ++side) {
            const int32_t v = side ? b : a;
            for (int32_t f : vf[(size_t)v]) {
                if (!falive[(size_t)f]) continue;
                const int32_t* t = &Fv[3 * (int64_t)f];
                if ((t[0] == a || t[1] == a || t[2] == a) && (t[0] == b || t[1] == b || t[2] == b)) continue;
                double pn[3][3], po[3][3];
                for (int k = 0; k < 3; ++k)
                    for (int d = 0; d < 3; ++d) {
                        po[k][d] = P[3 * (int64_t)t[k] + d];
                        pn[k][d] = t[k] == v ? c.p[d] : po[k][d];
                    }
                double e1[3], e2[3], n0[3], n1[3];
                for (int d = 0; d < 3; ++d) {
                    e1[d] = po[1][d] - po[0][d];
                    e2[d] = po[2][d] - po[0][d];
                }
                cross3(e1, e2, n0);
                for (int d = 0; d < 3; ++d) {
                    e1[d] = pn[1][d] - pn[0][d];
                    e2[d] = pn[2][d] - pn[0][d];
                }
                cross3(e1, e2, n1);
                const double l0 = sqrt(n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2]);
                const double l1 = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
                if (!(l1 > 1e-12 * (l0 > 0 ? l0 : 1.0)) || n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2] <= 0) bad = true;
            }
        }
        if (bad) continue;
        // collapse b into a
        for (int32_t f : vf[(size_t)b]) {
            if (!falive[(size_t)f]) continue;
            int32_t* t = &Fv[3 * (int64_t)f];
            if (t[0] == a || t[1] == a || t[2] == a) {
                falive[(size_t)f] = 0;
                --live_faces;
            } else {
                for (int k = 0; k < 3; ++k)
                    if (t[k] == b) t[k] = a;
                vf[(size_t)a].push_back(f);
            }
        }
        valive[(size_t)b] = 0;
        vf[(size_t)b].clear();
        for (int d = 0; d < 3; ++d) P[3 * (int64_t)a + d] = c.p[d];
        Q[(size_t)a].add(Q[(size_t)b]);
        ++ver[(size_t)a];
        ++ver[(size_t)b];
        {
            std::vector<int32_t>& L = vf[(size_t)a];
            L.erase(std::remove_if(L.begin(), L.end(), [&](int32_t f) { return !falive[(size_t)f]; }), L.end());
        }
        neighbours(a, na);
        for (int32_t w : na) {
            if (locked[(size_t)w]) continue;
            Cand nc;
            make(a, w, nc);
            pq.push(nc);
        }
    }
    // compact in original order
    std::vector<int32_t> map((size_t)V, -1);
    int64_t nv = 0;
    for (int64_t v = 0; v < V; ++v) {
        if (!valive[(size_t)v] || vf[(size_t)v].empty()) continue;
        bool used = false;
        for (int32_t f : vf[(size_t)v])
            if (falive[(size_t)f]) {
                used = true;
                break;
            }
        if (!used) continue;
        map[(size_t)v] = (int32_t)nv;
        for (int d = 0; d < 3; ++d) out_verts[3 * nv + d] = P[3 * v + d];
        ++nv;
    }
    int64_t nf = 0;
    for (int64_t f = 0; f < F; ++f) {
        if (!falive[(size_t)f]) continue;
        for (int k = 0; k < 3; ++k) out_faces[3 * nf + k] = map[(size_t)Fv[3 * f + k]];
        ++nf;
    }
    *n_out_verts = nv;
    *n_out_faces = nf;
    return MBA_OK;
}

}  // namespace

extern "C" {

int mba_decimate_quadric(const double* verts, int64_t n_verts, const int32_t* faces, int64_t n_faces,
                         int64_t target_faces, double* out_verts, int32_t* out_faces, int64_t* n_out_verts,
                         int64_t* n_out_faces)
{
    return mba_guard([&] { return decimate_quadric(verts, n_verts, faces, n_faces, target_faces, out_verts, out_faces, n_out_verts, n_out_faces); });
}

}  // extern "C"
