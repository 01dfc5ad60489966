// Ferrofluid pressure step around the Laplace BEM solve
// (MeshPressure.solve, src/mesh_ferrofluid/mesh_pressure.py:54-163).
//
// k_pressure_bc (one thread per vertex j, gathering its incident faces in the
// vertex -> face-slot CSR order):
//   vertex normal n_j = normalise(sum_f A_f n_f)            (mesh_utils.py:242-274)
//   cot Laplacian of pytorch3d's cot_laplacian (the reference's surface-tension
//   term, mesh_pressure.py:111-124): for face f with j = f[t], a = f[(t+1)%3],
//   b = f[(t+2)%3]: s += cot(angle at b) (v_a - v_j) + cot(angle at a) (v_b - v_j),
//   S = sum_f A_f (Heron, A^2 clamped at 1e-12);
//   lv = 0.5 s / S, kappa = 0.5 |lv| sign(lv . n_j)
//   Dirichlet p_j = -rho g . v_j - P_mag_j + sigma kappa (0 at NEUMANN vertices)
//   Neumann dp/dn_j = (rho / dt) (vel_j . n_j)          (0 at DIRICHLET vertices)
// k_pressure_update (one thread per vertex):
//   grad p_j = dpdn_j n_j + (1/3) sum_f g_f,
//   g_f = sum_t p_{f[t]} (n_f x (v_{f[t+2]} - v_{f[t+1]})) / (2 A_f)
//   (compute_mesh_face_tangent_gradient / compute_mesh_vert_tangent_gradient,
//   mesh_utils.py:501-615); new_vel_j = vel_j - (dt / rho) grad p_j.
// Batches b < B of vel / flags / P_mag / p / dpdn share one mesh.
#include <math.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace {

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

template <class T>
__device__ __forceinline__ vec3<T> ld3(const T* p, int64_t i)
{
    return vec3<T>{p[3 * i], p[3 * i + 1], p[3 * i + 2]};
}
template <class T>
__device__ __forceinline__ vec3<T> sub(vec3<T> a, vec3<T> b)
{
    return vec3<T>{a.x - b.x, a.y - b.y, a.z - b.z};
}
template <class T>
__device__ __forceinline__ vec3<T> crs(vec3<T> a, vec3<T> b)
{
    return vec3<T>{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T>
__device__ __forceinline__ T nrm(vec3<T> a)
{
    return sqrt(a.x * a.x + a.y * a.y + a.z * a.z);
}

// face area (cross product, as compute_face_areas) and unit normal
// normalise((v1 - v0) x (v2 - v1))
template <class T>
__device__ __forceinline__ void face_geom(const T* verts, const int32_t* faces, int64_t f, vec3<T> v[3], T& area,
                                          vec3<T>& n)
{
    for (int t = 0; t < 3; ++t) v[t] = ld3(verts, faces[3 * f + t]);
    const vec3<T> c = crs(sub(v[1], v[0]), sub(v[2], v[1]));
    const T l = nrm(c);
    area = (T)0.5 * l;
    n = vec3<T>{c.x / l, c.y / l, c.z / l};
}

template <class T>
__global__ void k_pressure_bc(const T* __restrict__ verts, const int32_t* __restrict__ faces, int64_t V,
                              const int64_t* __restrict__ vptr, const int32_t* __restrict__ vidx,
                              const T* __restrict__ vel, const uint8_t* __restrict__ flags,
                              const T* __restrict__ pmag, int64_t B, T rho, T gy, T sigma, T dt,
                              T* __restrict__ vnormal, T* __restrict__ kappa, T* __restrict__ p_dir,
                              T* __restrict__ dpdn)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= V) return;
    const vec3<T> vj = ld3(verts, j);
    vec3<T> an{0, 0, 0}, s{0, 0, 0};
    T Ssum = 0;
    for (int64_t q = vptr[j]; q < vptr[j + 1]; ++q) {
        const int64_t slot = vidx[q], f = slot / 3;
        const int t = (int)(slot % 3);
        vec3<T> v[3], n;
        T area;
        face_geom(verts, faces, f, v, area, n);
        an.x += area * n.x;
        an.y += area * n.y;
        an.z += area * n.z;
        // Heron's area and the cotangents as pytorch3d forms them:
        // side A opposite v0, B opposite v1, C opposite v2; cot_k = (sum of the
        // other two squared sides - own squared side) / (4 area)
        const T A = nrm(sub(v[1], v[2])), Bs = nrm(sub(v[0], v[2])), C = nrm(sub(v[0], v[1]));
        const T hs = (T)0.5 * (A + Bs + C);
        T a2 = hs * (hs - A) * (hs - Bs) * (hs - C);
        a2 = a2 > (T)1e-12 ? a2 : (T)1e-12;
        const T ar = sqrt(a2);
        const T A2 = A * A, B2 = Bs * Bs, C2 = C * C;
        const T cot[3] = {(B2 + C2 - A2) / ar / (T)4, (A2 + C2 - B2) / ar / (T)4, (A2 + B2 - C2) / ar / (T)4};
        Ssum += ar;
        const int ta = (t + 1) % 3, tb = (t + 2) % 3;
        const vec3<T> da = sub(v[ta], vj), db = sub(v[tb], vj);
        s.x += cot[tb] * da.x + cot[ta] * db.x;
        s.y += cot[tb] * da.y + cot[ta] * db.y;
        s.z += cot[tb] * da.z + cot[ta] * db.z;
    }
    const T la = nrm(an);
    const vec3<T> nj = la > (T)0 ? vec3<T>{an.x / la, an.y / la, an.z / la} : vec3<T>{0, 0, 0};
    const T inv = Ssum > (T)0 ? (T)1 / Ssum : (T)0;
    const vec3<T> lv{(T)0.5 * inv * s.x, (T)0.5 * inv * s.y, (T)0.5 * inv * s.z};
    const T ldot = lv.x * nj.x + lv.y * nj.y + lv.z * nj.z;
    const T sg = ldot > (T)0 ? (T)1 : (ldot < (T)0 ? (T)-1 : (T)0);
    const T kap = (T)0.5 * nrm(lv) * sg;
    vnormal[3 * j] = nj.x;
    vnormal[3 * j + 1] = nj.y;
    vnormal[3 * j + 2] = nj.z;
    kappa[j] = kap;
    // gravity vector (0, -g, 0): -rho (gravity . v) = rho g v_y
    const T pg = rho * gy * vj.y;
    for (int64_t b = 0; b < B; ++b) {
        const int64_t o = b * V + j;
        const uint8_t fl = flags[o];
        T pd = pg + sigma * kap;
        if (pmag) pd -= pmag[o];
        p_dir[o] = fl == 1 ? (T)0 : pd;  // BoundaryType: 0 DIRICHLET, 1 NEUMANN
        const vec3<T> u = ld3(vel, o);
        const T un = u.x * nj.x + u.y * nj.y + u.z * nj.z;
        dpdn[o] = fl == 0 ? (T)0 : (rho / dt) * un;
    }
}

template <class T>
__global__ void k_pressure_update(const T* __restrict__ verts, const int32_t* __restrict__ faces, int64_t V,
                                  const int64_t* __restrict__ vptr, const int32_t* __restrict__ vidx,
                                  const T* __restrict__ vnormal, const T* __restrict__ p, const T* __restrict__ dpdn,
                                  const T* __restrict__ vel, int64_t B, T rho, T dt, T* __restrict__ new_vel)
{
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= V) return;
    const vec3<T> nj = ld3(vnormal, j);
    for (int64_t b = 0; b < B; ++b) {
        vec3<T> g{0, 0, 0};
        for (int64_t q = vptr[j]; q < vptr[j + 1]; ++q) {
            const int64_t f = vidx[q] / 3;
            vec3<T> v[3], n;
            T area;
            face_geom(verts, faces, f, v, area, n);
            vec3<T> gf{0, 0, 0};
            for (int t = 0; t < 3; ++t) {
                const T pt = p[b * V + faces[3 * f + t]];
                const vec3<T> c = crs(n, sub(v[(t + 2) % 3], v[(t + 1) % 3]));
                gf.x += pt * c.x;
                gf.y += pt * c.y;
                gf.z += pt * c.z;
            }
            const T sc = (T)1 / ((T)2 * area) / (T)3;
            g.x += gf.x * sc;
            g.y += gf.y * sc;
            g.z += gf.z * sc;
        }
        const int64_t o = b * V + j;
        const T dn = dpdn[o];
        const T c = dt / rho;
        new_vel[3 * o] = vel[3 * o] - c * (dn * nj.x + g.x);
        new_vel[3 * o + 1] = vel[3 * o + 1] - c * (dn * nj.y + g.y);
        new_vel[3 * o + 2] = vel[3 * o + 2] - c * (dn * nj.z + g.z);
    }
}

}  // namespace

extern "C" {

int mba_pressure_bc(int prec, const void* verts, const int32_t* faces, int64_t n_verts, int64_t n_faces,
                    const int64_t* vinc_ptr, const int32_t* vinc_idx, const void* vel, const uint8_t* flags,
                    const void* p_mag, int64_t batch, double density, double gravity_strength,
                    double surface_tension, double dt, void* vert_normals, void* kappa, void* p_dirichlet,
                    void* dpdn_neumann, mba_stream_t stream)
{
    if (!verts || !faces || !vinc_ptr || !vinc_idx || !vel || !flags || !vert_normals || !kappa || !p_dirichlet ||
        !dpdn_neumann)
        return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_verts <= 0 || n_faces <= 0 || batch <= 0) return MBA_ERR_SIZE;
    if (!(dt > 0.0) || !(density > 0.0)) return MBA_ERR_ARG;
    const dim3 grid((unsigned)((n_verts + 255) / 256));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_pressure_bc<float>, grid, dim3(256), 0, S(stream), (const float*)verts, faces, n_verts,
                           vinc_ptr, vinc_idx, (const float*)vel, flags, (const float*)p_mag, batch, (float)density,
                           (float)gravity_strength, (float)surface_tension, (float)dt, (float*)vert_normals,
                           (float*)kappa, (float*)p_dirichlet, (float*)dpdn_neumann);
    else
        hipLaunchKernelGGL(k_pressure_bc<double>, grid, dim3(256), 0, S(stream), (const double*)verts, faces,
                           n_verts, vinc_ptr, vinc_idx, (const double*)vel, flags, (const double*)p_mag, batch,
                           density, gravity_strength, surface_tension, dt, (double*)vert_normals, (double*)kappa,
                           (double*)p_dirichlet, (double*)dpdn_neumann);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_pressure_update(int prec, const void* verts, const int32_t* faces, int64_t n_verts, int64_t n_faces,
                        const int64_t* vinc_ptr, const int32_t* vinc_idx, const void* vert_normals, const void* p,
                        const void* dpdn, const void* vel, int64_t batch, double density, double dt, void* new_vel,
                        mba_stream_t stream)
{
    if (!verts || !faces || !vinc_ptr || !vinc_idx || !vert_normals || !p || !dpdn || !vel || !new_vel)
        return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_verts <= 0 || n_faces <= 0 || batch <= 0) return MBA_ERR_SIZE;
    if (!(dt > 0.0) || !(density > 0.0)) return MBA_ERR_ARG;
    const dim3 grid((unsigned)((n_verts + 255) / 256));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_pressure_update<float>, grid, dim3(256), 0, S(stream), (const float*)verts, faces,
                           n_verts, vinc_ptr, vinc_idx, (const float*)vert_normals, (const float*)p,
                           (const float*)dpdn, (const float*)vel, batch, (float)density, (float)dt, (float*)new_vel);
    else
        hipLaunchKernelGGL(k_pressure_update<double>, grid, dim3(256), 0, S(stream), (const double*)verts, faces,
                           n_verts, vinc_ptr, vinc_idx, (const double*)vert_normals, (const double*)p,
                           (const double*)dpdn, (const double*)vel, batch, density, dt, (double*)new_vel);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
