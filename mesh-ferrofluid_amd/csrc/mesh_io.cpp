// Host-side mesh readers and the BEAST .in writer (no GPU code).
//
// Formats of LoadSingleMesh (src/utils/io.py:60-139), chosen by extension:
//   .obj   "v x y z" and "f a b c ..." lines; index forms i, i/j, i//k,
//          i/j/k; negative indices count back from the last vertex read;
//          polygons split into the fan (p0, p_i, p_i+1).  (The reference uses
//          pywavefront with collect_faces=True, which triangulates the same
//          way.)
//   .ply   ascii, binary_little_endian or binary_big_endian; vertex x, y, z of
//          any scalar type; face list "vertex_indices" / "vertex_index" of any
//          count / index types, polygons as fans; other elements and
//          properties skipped.  (The reference uses open3d.)
//   .json  {"verts": [...], "faces": [...]}: nested or flat number arrays,
//          reshaped to [-1, 3] (io.py:88-96).
//   .dae   COLLADA: every <geometry>'s <mesh> gives one part; positions from
//          the <vertices> POSITION source, faces from <triangles>,
//          <polylist> (with <vcount>, fans) and <polygons> (<p> per polygon),
//          each read through its VERTEX input offset and input stride; scene
//          transforms are not applied (the reference reads the raw geometry
//          of trimesh's scene, io.py:110-137).
//   .in    BEAST: "1", "NV NF", NV lines "x y z", NF lines of 1-based "a b c"
//          (demo/generate_in_file.py:12-40); indices are returned 0-based.
// Writer: the same .in layout, vertices printed with 17 significant digits
// (exact round trip of float64 and of float32 values),
// faces made 1-based when the smallest index is 0 (as save_beast_in does).
#include <ctype.h>
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/mesh_bem_amd.h"
#include "mba_guard.hpp"

namespace {

struct Part {
    std::vector<double> v;   // [n_verts * 3]
    std::vector<int32_t> f;  // [n_faces * 3]
};

bool read_file(const char* path, std::string& out)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) return false;
    if (fseek(fp, 0, SEEK_END) != 0) {
        fclose(fp);
        return false;
    }
    long n = ftell(fp);
    if (n < 0) {
        fclose(fp);
        return false;
    }
    rewind(fp);
    out.resize((size_t)n);
    size_t got = n ? fread(&out[0], 1, (size_t)n, fp) : 0;
    fclose(fp);
    return got == (size_t)n;
}

std::string lower_ext(const char* path)
{
    const char* dot = strrchr(path, '.');
    std::string e = dot ? dot : "";
    for (auto& c : e) c = (char)tolower((unsigned char)c);
    return e;
}

// x as an integer index: a finite double in [-2^53, 2^53] truncates to int64
// without undefined behaviour; anything else maps to an out-of-range value
// that the index checks reject.
int64_t as_index(double x)
{
    if (!(x >= -9007199254740992.0 && x <= 9007199254740992.0)) return INT64_MIN;
    return (int64_t)x;
}

// Triangle fan of a polygon; false when an index does not fit int32 (such
// an index cannot name a vertex of a mesh this library accepts).
bool fan(const std::vector<int64_t>& poly, std::vector<int32_t>& f)
{
    for (int64_t i : poly)
        if (i < INT32_MIN || i > INT32_MAX) return false;
    for (size_t i = 1; i + 1 < poly.size(); ++i) {
        f.push_back((int32_t)poly[0]);
        f.push_back((int32_t)poly[i]);
        f.push_back((int32_t)poly[i + 1]);
    }
    return true;
}

bool check_indices(const Part& p)
{
    const int64_t nv = (int64_t)(p.v.size() / 3);
    for (int32_t i : p.f)
        if (i < 0 || i >= nv) return false;
    return true;
}

// ---------------------------------------------------------------- OBJ
int parse_obj(const std::string& s, Part& p)
{
    size_t pos = 0;
    std::vector<int64_t> poly;
    while (pos < s.size()) {
        size_t eol = s.find('\n', pos);
        if (eol == std::string::npos) eol = s.size();
        const char* a = s.c_str() + pos;
        const char* e = s.c_str() + eol;
        while (a < e && (*a == ' ' || *a == '\t')) ++a;
        if (e - a >= 2 && a[0] == 'v' && (a[1] == ' ' || a[1] == '\t')) {
            char* q = (char*)a + 1;
            for (int c = 0; c < 3; ++c) {
                char* r;
                double x = strtod(q, &r);
                if (r == q || r > e) return MBA_ERR_MESH;
                p.v.push_back(x);
                q = r;
            }
        } else if (e - a >= 2 && a[0] == 'f' && (a[1] == ' ' || a[1] == '\t')) {
            poly.clear();
            const char* q = a + 1;
            while (q < e) {
                while (q < e && (*q == ' ' || *q == '\t' || *q == '\r')) ++q;
                if (q >= e) break;
                char* r;
                long long idx = strtoll(q, &r, 10);
                if (r == q) return MBA_ERR_MESH;
                const int64_t nv = (int64_t)(p.v.size() / 3);
                poly.push_back(idx > 0 ? idx - 1 : nv + idx);
                q = r;
                while (q < e && *q != ' ' && *q != '\t' && *q != '\r') ++q;  // skip /vt/vn
            }
            if (poly.size() < 3 || !fan(poly, p.f)) return MBA_ERR_MESH;
        }
        pos = eol + 1;
    }
    return MBA_OK;
}

// ---------------------------------------------------------------- PLY
enum PlyType { P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64, P_BAD };

PlyType ply_type(const std::string& t)
{
    if (t == "char" || t == "int8") return P_I8;
    if (t == "uchar" || t == "uint8") return P_U8;
    if (t == "short" || t == "int16") return P_I16;
    if (t == "ushort" || t == "uint16") return P_U16;
    if (t == "int" || t == "int32") return P_I32;
    if (t == "uint" || t == "uint32") return P_U32;
    if (t == "float" || t == "float32") return P_F32;
    if (t == "double" || t == "float64") return P_F64;
    return P_BAD;
}

int ply_size(PlyType t)
{
    switch (t) {
    case P_I8:
    case P_U8: return 1;
    case P_I16:
    case P_U16: return 2;
    case P_I32:
    case P_U32:
    case P_F32: return 4;
    case P_F64: return 8;
    default: return 0;
    }
}

struct PlyProp {
    std::string name;
    bool list = false;
    PlyType count = P_BAD, item = P_BAD;
};
struct PlyElem {
    std::string name;
    int64_t n = 0;
    std::vector<PlyProp> props;
};

struct PlyReader {
    const std::string& s;
    size_t pos;
    int fmt;  // 0 ascii, 1 little endian, 2 big endian
    bool ok = true;
    PlyReader(const std::string& str, size_t p, int f) : s(str), pos(p), fmt(f) {}
    double bin(PlyType t)
    {
        const int n = ply_size(t);
        if (pos + (size_t)n > s.size()) {
            ok = false;
            return 0.0;
        }
        unsigned char b[8];
        memcpy(b, s.data() + pos, (size_t)n);
        pos += (size_t)n;
        if (fmt == 2)
            for (int i = 0; i < n / 2; ++i) {
                unsigned char t2 = b[i];
                b[i] = b[n - 1 - i];
                b[n - 1 - i] = t2;
            }
        switch (t) {
        case P_I8: return (double)(int8_t)b[0];
        case P_U8: return (double)b[0];
        case P_I16: { int16_t x; memcpy(&x, b, 2); return (double)x; }
        case P_U16: { uint16_t x; memcpy(&x, b, 2); return (double)x; }
        case P_I32: { int32_t x; memcpy(&x, b, 4); return (double)x; }
        case P_U32: { uint32_t x; memcpy(&x, b, 4); return (double)x; }
        case P_F32: { float x; memcpy(&x, b, 4); return (double)x; }
        case P_F64: { double x; memcpy(&x, b, 8); return x; }
        default: ok = false; return 0.0;
        }
    }
    double ascii()
    {
        while (pos < s.size() && isspace((unsigned char)s[pos])) ++pos;
        if (pos >= s.size()) {
            ok = false;
            return 0.0;
        }
        char* r;
        const char* a = s.c_str() + pos;
        double x = strtod(a, &r);
        if (r == a) {
            ok = false;
            return 0.0;
        }
        pos += (size_t)(r - a);
        return x;
    }
    double get(PlyType t) { return fmt == 0 ? ascii() : bin(t); }
};

int parse_ply(const std::string& s, Part& p)
{
    if (s.compare(0, 3, "ply") != 0) return MBA_ERR_MESH;
    size_t hend = s.find("end_header");
    if (hend == std::string::npos) return MBA_ERR_MESH;
    size_t body = s.find('\n', hend);
    if (body == std::string::npos) return MBA_ERR_MESH;
    ++body;
    int fmt = -1;
    std::vector<PlyElem> elems;
    size_t pos = 0;
    while (pos < hend) {
        size_t eol = s.find('\n', pos);
        std::string line = s.substr(pos, eol - pos);
        pos = eol + 1;
        char w0[64] = {0}, w1[64] = {0}, w2[64] = {0}, w3[64] = {0}, w4[128] = {0};
        int nw = sscanf(line.c_str(), "%63s %63s %63s %63s %127s", w0, w1, w2, w3, w4);
        if (nw <= 0) continue;
        std::string k = w0;
        if (k == "format") {
            std::string f = w1;
            fmt = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
        } else if (k == "element" && nw >= 3) {
            PlyElem e;
            e.name = w1;
            e.n = strtoll(w2, nullptr, 10);
            elems.push_back(e);
        } else if (k == "property" && !elems.empty()) {
            PlyProp pr;
            if (std::string(w1) == "list" && nw >= 5) {
                pr.list = true;
                pr.count = ply_type(w2);
                pr.item = ply_type(w3);
                pr.name = w4;
            } else if (nw >= 3) {
                pr.item = ply_type(w1);
                pr.name = w2;
            }
            if (pr.item == P_BAD || (pr.list && pr.count == P_BAD)) return MBA_ERR_MESH;
            elems.back().props.push_back(pr);
        }
    }
    if (fmt < 0) return MBA_ERR_MESH;
    PlyReader r(s, body, fmt);
    std::vector<int64_t> poly;
    for (const PlyElem& e : elems) {
        const bool is_v = e.name == "vertex", is_f = e.name == "face";
        for (int64_t i = 0; i < e.n && r.ok; ++i) {
            double xyz[3] = {0, 0, 0};
            for (const PlyProp& pr : e.props) {
                if (pr.list) {
                    const int64_t c = as_index(r.get(pr.count));
                    if (c < 0) return MBA_ERR_MESH;
                    poly.clear();
                    for (int64_t j = 0; j < c && r.ok; ++j) poly.push_back(as_index(r.get(pr.item)));
                    if (is_f && (pr.name == "vertex_indices" || pr.name == "vertex_index")) {
                        if (poly.size() < 3 || !fan(poly, p.f)) return MBA_ERR_MESH;
                    }
                } else {
                    const double x = r.get(pr.item);
                    if (is_v) {
                        if (pr.name == "x") xyz[0] = x;
                        else if (pr.name == "y") xyz[1] = x;
                        else if (pr.name == "z") xyz[2] = x;
                    }
                }
            }
            if (is_v) p.v.insert(p.v.end(), xyz, xyz + 3);
        }
    }
    return r.ok ? MBA_OK : MBA_ERR_MESH;
}

// ---------------------------------------------------------------- JSON
// Numbers of the value of "key" (a possibly nested array), flattened.
bool json_numbers(const std::string& s, const char* key, std::vector<double>& out)
{
    const std::string k = std::string("\"") + key + "\"";
    size_t at = s.find(k);
    if (at == std::string::npos) return false;
    size_t p = s.find(':', at + k.size());
    if (p == std::string::npos) return false;
    ++p;
    while (p < s.size() && isspace((unsigned char)s[p])) ++p;
    if (p >= s.size() || s[p] != '[') return false;
    int depth = 0;
    while (p < s.size()) {
        const char c = s[p];
        if (c == '[') {
            ++depth;
            ++p;
        } else if (c == ']') {
            --depth;
            ++p;
            if (depth == 0) return true;
        } else if (c == ',' || isspace((unsigned char)c)) {
            ++p;
        } else {
            char* r;
            double x = strtod(s.c_str() + p, &r);
            if (r == s.c_str() + p) return false;
            out.push_back(x);
            p = (size_t)(r - s.c_str());
        }
    }
    return false;
}

int parse_json(const std::string& s, Part& p)
{
    std::vector<double> f;
    if (!json_numbers(s, "verts", p.v) || !json_numbers(s, "faces", f)) return MBA_ERR_MESH;
    if (p.v.size() % 3 || f.size() % 3) return MBA_ERR_MESH;
    for (double x : f) {
        if (!(x >= 0.0 && x <= (double)INT32_MAX)) return MBA_ERR_MESH;
        p.f.push_back((int32_t)x);
    }
    return MBA_OK;
}

// ---------------------------------------------------------------- BEAST .in
int parse_in(const std::string& s, Part& p)
{
    const char* a = s.c_str();
    char* r;
    long long one = strtoll(a, &r, 10);
    if (r == a || one != 1) return MBA_ERR_MESH;
    a = r;
    long long nv = strtoll(a, &r, 10);
    if (r == a) return MBA_ERR_MESH;
    a = r;
    long long nf = strtoll(a, &r, 10);
    if (r == a || nv < 0 || nf < 0) return MBA_ERR_MESH;
    // every value the counts announce takes at least two bytes of the file
    // (a digit and a separator): larger counts are malformed, and bounding
    // them here keeps 3 * count from overflowing and the resizes below from
    // asking for more memory than the file could describe
    const long long cap = (long long)(s.size() / 6);
    if (nv > cap || nf > cap) return MBA_ERR_MESH;
    a = r;
    p.v.resize((size_t)nv * 3);
    for (auto& x : p.v) {
        x = strtod(a, &r);
        if (r == a) return MBA_ERR_MESH;
        a = r;
    }
    p.f.resize((size_t)nf * 3);
    for (auto& i : p.f) {
        long long x = strtoll(a, &r, 10);
        if (r == a || x < 1 || x > nv) return MBA_ERR_MESH;
        i = (int32_t)(x - 1);
        a = r;
    }
    return MBA_OK;
}

// ---------------------------------------------------------------- COLLADA
// Minimal tag scanner: enough for the <mesh> subtree of COLLADA geometry.
struct Tag {
    std::string name;
    std::unordered_map<std::string, std::string> attr;
    size_t body = 0, end = 0;  // content range [body, end)
    bool self_closing = false;
};

// Next tag named `name` (start tag) at or after pos, within [pos, lim).
bool find_tag(const std::string& s, const std::string& name, size_t pos, size_t lim, Tag& t)
{
    while (true) {
        size_t a = s.find("<" + name, pos);
        if (a == std::string::npos || a >= lim) return false;
        const char nc = s[a + 1 + name.size()];
        if (!(nc == ' ' || nc == '>' || nc == '/' || nc == '\t' || nc == '\n' || nc == '\r')) {
            pos = a + 1;
            continue;
        }
        size_t gt = s.find('>', a);
        if (gt == std::string::npos) return false;
        t = Tag();
        t.name = name;
        t.self_closing = s[gt - 1] == '/';
        // attributes
        size_t q = a + 1 + name.size();
        while (q < gt) {
            while (q < gt && isspace((unsigned char)s[q])) ++q;
            size_t eq = s.find('=', q);
            if (eq == std::string::npos || eq >= gt) break;
            std::string key = s.substr(q, eq - q);
            while (!key.empty() && isspace((unsigned char)key.back())) key.pop_back();
            size_t q1 = s.find_first_of("\"'", eq);
            if (q1 == std::string::npos || q1 >= gt) break;
            size_t q2 = s.find(s[q1], q1 + 1);
            if (q2 == std::string::npos) break;
            t.attr[key] = s.substr(q1 + 1, q2 - q1 - 1);
            q = q2 + 1;
        }
        t.body = gt + 1;
        if (t.self_closing) {
            t.end = t.body;
            return true;
        }
        size_t close = s.find("</" + name, t.body);
        if (close == std::string::npos) return false;
        t.end = close;
        return true;
    }
}

void numbers(const std::string& s, size_t a, size_t b, std::vector<double>& out)
{
    const char* p = s.c_str() + a;
    const char* e = s.c_str() + b;
    while (p < e) {
        while (p < e && isspace((unsigned char)*p)) ++p;
        if (p >= e) break;
        char* r;
        double x = strtod(p, &r);
        if (r == p) break;
        out.push_back(x);
        p = r;
    }
}

std::string strip_hash(const std::string& x) { return (!x.empty() && x[0] == '#') ? x.substr(1) : x; }

int parse_dae(const std::string& s, std::vector<Part>& parts)
{
    size_t pos = 0;
    Tag geom;
    while (find_tag(s, "geometry", pos, s.size(), geom)) {
        pos = geom.end;
        Tag mesh;
        if (!find_tag(s, "mesh", geom.body, geom.end, mesh)) continue;
        // sources: id -> float_array values
        std::unordered_map<std::string, std::vector<double>> src;
        size_t sp = mesh.body;
        Tag source;
        while (find_tag(s, "source", sp, mesh.end, source)) {
            sp = source.end;
            Tag fa;
            if (find_tag(s, "float_array", source.body, source.end, fa)) {
                std::vector<double> vals;
                numbers(s, fa.body, fa.end, vals);
                src[source.attr["id"]] = vals;
            }
        }
        // vertices id -> POSITION source
        Tag vert;
        if (!find_tag(s, "vertices", mesh.body, mesh.end, vert)) return MBA_ERR_MESH;
        std::string pos_src;
        size_t ip = vert.body;
        Tag in;
        while (find_tag(s, "input", ip, vert.end, in)) {
            ip = in.body;
            if (in.attr["semantic"] == "POSITION") pos_src = strip_hash(in.attr["source"]);
        }
        if (!src.count(pos_src)) return MBA_ERR_MESH;
        Part p;
        p.v = src[pos_src];
        if (p.v.size() % 3) return MBA_ERR_MESH;
        const char* prims[3] = {"triangles", "polylist", "polygons"};
        for (int k = 0; k < 3; ++k) {
            size_t pp = mesh.body;
            Tag prim;
            while (find_tag(s, prims[k], pp, mesh.end, prim)) {
                pp = prim.end;
                int stride = 0, voff = -1;
                size_t qp = prim.body;
                while (find_tag(s, "input", qp, prim.end, in)) {
                    qp = in.body;
                    const int off = atoi(in.attr["offset"].c_str());
                    if (off < 0 || off > 4096) return MBA_ERR_MESH;
                    if (off + 1 > stride) stride = off + 1;
                    if (in.attr["semantic"] == "VERTEX") voff = off;
                }
                if (voff < 0 || stride <= 0) return MBA_ERR_MESH;
                std::vector<double> counts;
                if (k == 1) {
                    Tag vc;
                    if (!find_tag(s, "vcount", prim.body, prim.end, vc)) return MBA_ERR_MESH;
                    numbers(s, vc.body, vc.end, counts);
                }
                std::vector<int64_t> poly;
                size_t ppos = prim.body;
                Tag ptag;
                size_t ci = 0;
                while (find_tag(s, "p", ppos, prim.end, ptag)) {
                    ppos = ptag.end;
                    std::vector<double> idx;
                    numbers(s, ptag.body, ptag.end, idx);
                    const size_t nvtx = idx.size() / (size_t)stride;
                    if (k == 0) {
                        for (size_t t = 0; t + 3 <= nvtx; t += 3)
                            for (int c = 0; c < 3; ++c) {
                                const int64_t x = as_index(idx[(t + c) * stride + voff]);
                                if (x < 0 || x > INT32_MAX) return MBA_ERR_MESH;
                                p.f.push_back((int32_t)x);
                            }
                    } else if (k == 1) {
                        size_t at = 0;
                        for (; ci < counts.size(); ++ci) {
                            const int64_t cc = as_index(counts[ci]);
                            if (cc < 0) return MBA_ERR_MESH;
                            const size_t c = (size_t)cc;
                            if (at + c > nvtx) return MBA_ERR_MESH;
                            poly.clear();
                            for (size_t j = 0; j < c; ++j) poly.push_back(as_index(idx[(at + j) * stride + voff]));
                            if (poly.size() >= 3 && !fan(poly, p.f)) return MBA_ERR_MESH;
                            at += c;
                        }
                    } else {
                        poly.clear();
                        for (size_t j = 0; j < nvtx; ++j) poly.push_back(as_index(idx[j * stride + voff]));
                        if (poly.size() >= 3 && !fan(poly, p.f)) return MBA_ERR_MESH;
                    }
                }
            }
        }
        parts.push_back(std::move(p));
    }
    return parts.empty() ? MBA_ERR_MESH : MBA_OK;
}

// Reads every part of the mesh file at path into malloc'ed arrays.
int mesh_read(const char* path, mba_mesh_part** parts, int32_t* n_parts)
{
    const std::string ext = lower_ext(path);
    if (ext != ".obj" && ext != ".ply" && ext != ".json" && ext != ".dae" && ext != ".in") return MBA_ERR_ARG;
    std::string s;
    if (!read_file(path, s)) return MBA_ERR_IO;
    std::vector<Part> ps;
    int rc;
    if (ext == ".dae") {
        rc = parse_dae(s, ps);
    } else {
        Part p;
        rc = ext == ".obj" ? parse_obj(s, p) : ext == ".ply" ? parse_ply(s, p) : ext == ".json" ? parse_json(s, p)
                                                                                    : parse_in(s, p);
        if (rc == MBA_OK) ps.push_back(std::move(p));
    }
    if (rc != MBA_OK) return rc;
    if (ps.size() > (size_t)INT32_MAX) return MBA_ERR_SIZE;
    for (const Part& p : ps)
        if (!check_indices(p)) return MBA_ERR_MESH;
    mba_mesh_part* out = (mba_mesh_part*)calloc(ps.size(), sizeof(mba_mesh_part));
    if (!out) return MBA_ERR_ALLOC;
    for (size_t i = 0; i < ps.size(); ++i) {
        out[i].n_verts = (int64_t)(ps[i].v.size() / 3);
        out[i].n_faces = (int64_t)(ps[i].f.size() / 3);
        out[i].verts = (double*)malloc(ps[i].v.size() * sizeof(double) + 1);
        out[i].faces = (int32_t*)malloc(ps[i].f.size() * sizeof(int32_t) + 1);
        if (!out[i].verts || !out[i].faces) {
            mba_mesh_free(out, (int32_t)ps.size());
            return MBA_ERR_ALLOC;
        }
        // an empty part has no source arrays (data() may be null)
        if (!ps[i].v.empty()) memcpy(out[i].verts, ps[i].v.data(), ps[i].v.size() * sizeof(double));
        if (!ps[i].f.empty()) memcpy(out[i].faces, ps[i].f.data(), ps[i].f.size() * sizeof(int32_t));
    }
    *parts = out;
    *n_parts = (int32_t)ps.size();
    return MBA_OK;
}

}  // namespace

extern "C" {

int mba_mesh_read(const char* path, mba_mesh_part** parts, int32_t* n_parts)
{
    if (!path || !parts || !n_parts) return MBA_ERR_ARG;
    *parts = nullptr;
    *n_parts = 0;
    return mba_guard([&] { return mesh_read(path, parts, n_parts); }, MBA_ERR_MESH);
}

void mba_mesh_free(mba_mesh_part* parts, int32_t n_parts)
{
    if (!parts) return;
    for (int32_t i = 0; i < n_parts; ++i) {
        free(parts[i].verts);
        free(parts[i].faces);
    }
    free(parts);
}

int mba_mesh_write_beast_in(const char* path, const double* verts, int64_t n_verts, const int32_t* faces,
                            int64_t n_faces)
{
    if (!path || (n_verts > 0 && !verts) || (n_faces > 0 && !faces)) return MBA_ERR_ARG;
    if (n_verts < 0 || n_faces < 0) return MBA_ERR_SIZE;
    int32_t mn = 0;
    for (int64_t i = 0; i < 3 * n_faces; ++i) mn = (i == 0 || faces[i] < mn) ? faces[i] : mn;
    const int32_t shift = (n_faces > 0 && mn == 0) ? 1 : 0;
    FILE* fp = fopen(path, "w");
    if (!fp) return MBA_ERR_IO;
    fprintf(fp, "1\n%lld %lld\n", (long long)n_verts, (long long)n_faces);
    for (int64_t i = 0; i < n_verts; ++i)
        fprintf(fp, "%.17g %.17g %.17g\n", verts[3 * i], verts[3 * i + 1], verts[3 * i + 2]);
    for (int64_t i = 0; i < n_faces; ++i)
        fprintf(fp, "%d %d %d\n", faces[3 * i] + shift, faces[3 * i + 1] + shift, faces[3 * i + 2] + shift);
    const int rc = ferror(fp) ? MBA_ERR_IO : MBA_OK;
    if (fclose(fp) != 0) return MBA_ERR_IO;
    return rc;
}

}  // extern "C"
