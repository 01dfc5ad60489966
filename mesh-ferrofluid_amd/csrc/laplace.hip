// Laplace single / double layer kernels for gfx950.
//
// Reference rule (src/mesh_BEM/utils/integrations_3d.py:252-452): for the face
// pair (fy, fx) with relation r = number of shared vertex indices, the point
// table tab[r] holds rows (r1x, r2x, r1y, r2y, W) and
//   U[fy, fx, iy, ix] = 4 A_x A_y sum_rows W phi_iy(r1y, r2y) phi_ix(r1x, r2x) f(x, y)
//   x = (1-r1x) X0 + (r1x-r2x) X1 + r2x X2 on fx, y likewise on fy,
//   single layer f = 1/(4 pi R), double layer f = (x - y).n_fy / (4 pi R^3).
// One lane per column face fx, a uniform loop over a block of row faces fy.
//
// NS rule (P0 Galerkin, 7-point rule, analytic FP64 inner integrals on
// near pairs):
//   V[i, j]  = int_Ti int_Tj 1/(4 pi R)
//   Kt[i, j] = int_Ti int_Tj d/dn_x 1/(4 pi R) = -int int (x - y).n_i / (4 pi R^3)
//   (Kt is the adjoint double layer K'; Kt[i, i] = 0 on flat triangles.)
#include <math.h>
#include <stdint.h>

#include "mba_common.hpp"

namespace {

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

constexpr int LAP_ROWS = 32;

template <class T>
__global__ __launch_bounds__(256) void k_laplace_ref(
    const T* __restrict__ verts, const int32_t* __restrict__ faces, const T* __restrict__ areas,
    const T* __restrict__ normals, int64_t F, const T* __restrict__ tab, const int64_t* __restrict__ tab_off,
    int layer, int nphi, T* __restrict__ out)
{
    __shared__ T ys[LAP_ROWS][13];  // 3 vertices, normal, area
    __shared__ int32_t yi[LAP_ROWS][3];
    const int tid = threadIdx.x;
    const int64_t fy0 = (int64_t)blockIdx.y * LAP_ROWS;
    const int64_t fx = (int64_t)blockIdx.x * blockDim.x + tid;
    for (int t = tid; t < LAP_ROWS; t += blockDim.x) {
        int64_t fy = fy0 + t;
        if (fy < F) {
            for (int v = 0; v < 3; ++v) {
                int32_t id = faces[3 * fy + v];
                yi[t][v] = id;
                for (int q = 0; q < 3; ++q) ys[t][3 * v + q] = verts[3 * (int64_t)id + q];
            }
            for (int q = 0; q < 3; ++q) ys[t][9 + q] = normals[3 * fy + q];
            ys[t][12] = areas[fy];
        }
    }
    __syncthreads();
    if (fx >= F) return;
    int32_t xi[3];
    T X[9];
    for (int v = 0; v < 3; ++v) {
        xi[v] = faces[3 * fx + v];
        for (int q = 0; q < 3; ++q) X[3 * v + q] = verts[3 * (int64_t)xi[v] + q];
    }
    const T Ax = areas[fx];
    const T inv4pi = (T)MBA_INV4PI;
    const int nrows = (int)((F - fy0) < LAP_ROWS ? (F - fy0) : LAP_ROWS);
    for (int t = 0; t < nrows; ++t) {
        int rel = 0;
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) rel += (yi[t][a] == xi[b]);
        const T* Yabs = ys[t];
        // coordinates relative to the first vertex of fy: R and (x - y).n_y of
        // near pairs lose no digits to the absolute position
        T Y[12], Xr[9];
        for (int q = 0; q < 3; ++q) {
            Y[q] = (T)0;
            Y[3 + q] = Yabs[3 + q] - Yabs[q];
            Y[6 + q] = Yabs[6 + q] - Yabs[q];
            Y[9 + q] = Yabs[9 + q];
            for (int v = 0; v < 3; ++v) Xr[3 * v + q] = X[3 * v + q] - Yabs[q];
        }
        T acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        auto accumulate = [&](int64_t g0, int64_t g1) {
            for (int64_t g = g0; g < g1; ++g) {
                const T r1x = tab[5 * g], r2x = tab[5 * g + 1], r1y = tab[5 * g + 2], r2y = tab[5 * g + 3];
                const T W = tab[5 * g + 4];
                const T px0 = (T)1 - r1x, px1 = r1x - r2x, px2 = r2x;
                const T py0 = (T)1 - r1y, py1 = r1y - r2y, py2 = r2y;
                T d[3];
                for (int q = 0; q < 3; ++q)
                    d[q] = (px0 * Xr[q] + px1 * Xr[3 + q] + px2 * Xr[6 + q]) - (py1 * Y[3 + q] + py2 * Y[6 + q]);
                const T R2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
                const T rinv = rsqrt_t(R2);
                T fv = rinv * inv4pi;
                if (layer == 2) fv *= (d[0] * Y[9] + d[1] * Y[10] + d[2] * Y[11]) * rinv * rinv;
                const T wf = W * fv;
                if (nphi == 1) {
                    acc[0] += wf;
                } else {
                    const T pyv[3] = {py0, py1, py2}, pxv[3] = {px0, px1, px2};
                    for (int iy = 0; iy < 3; ++iy)
                        for (int ix = 0; ix < 3; ++ix) acc[3 * iy + ix] += wf * pyv[iy] * pxv[ix];
                }
            }
        };
        if (__all(rel == 0)) {
            // every lane of the wave has a separate pair: one table range for
            // the wave, so the row loads are uniform (scalar cache)
            accumulate(tab_off[0], tab_off[1]);
        } else {
            // the double layer of a flat face on itself is exactly zero ((x - y)
            // lies in the plane); summing rounding noise of (x - y).n / R^3 is skipped
            const int64_t r0 = tab_off[rel], r1 = (layer == 2 && rel == 3) ? r0 : tab_off[rel + 1];
            accumulate(r0, r1);
        }
        const T s = (T)4 * Ax * Yabs[12];
        T* o = out + ((fy0 + t) * F + fx) * (nphi * nphi);
        for (int q = 0; q < nphi * nphi; ++q) o[q] = s * acc[q];
    }
}

// Deterministic packing (replaces pack_matrix_tight's scatter_add_,
// src/utils/BEM_utils.py:314-357): A[vy, vx] = sum over incidences (fy, iy) of
// vy and (fx, ix) of vx of U[fy, fx, iy, ix] (PLANAR: U[fy, fx] / 9).
template <class T>
__global__ void k_pack(const T* __restrict__ U, int64_t F, int nphi, const int64_t* __restrict__ vptr,
                       const int32_t* __restrict__ vidx, int64_t V, T* __restrict__ A)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= V * V) return;
    const int64_t vy = id / V, vx = id % V;
    T acc = 0;
    for (int64_t p = vptr[vy]; p < vptr[vy + 1]; ++p) {
        const int64_t sy = vidx[p], fy = sy / 3, iy = sy % 3;
        for (int64_t q = vptr[vx]; q < vptr[vx + 1]; ++q) {
            const int64_t sx = vidx[q], fx = sx / 3, ix = sx % 3;
            if (nphi == 1)
                acc += U[fy * F + fx] / (T)9;
            else
                acc += U[((fy * F + fx) * 3 + iy) * 3 + ix];
        }
    }
    A[id] = acc;
}

// rhs[v] = 1/2 g[v] va[v] + sum_vy Kt[vy, v] g[vy]   (rhs_layer_3d.py:122-146)
template <class T>
__global__ void k_dirichlet_rhs(const T* __restrict__ Kt, int64_t V, const T* __restrict__ g,
                                const T* __restrict__ va, T* __restrict__ rhs)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    T acc = 0;
    for (int64_t y = 0; y < V; ++y) acc += Kt[y * V + v] * g[y];
    rhs[v] = (T)0.5 * g[v] * va[v] + acc;
}

// ------------------------------------------------------------ NS P0
// Analytic inner integrals over triangle Q at point r (FP64):
//   S0 = int 1/R,  grad S0 = -sum_i u_i f_i - n sgn(d) sum_i beta_i.
__device__ void tri_S0_grad(const double Q[3][3], const double r[3], double& S0, double G[3])
{
    double e01[3], e12[3], n[3];
    for (int a = 0; a < 3; ++a) {
        e01[a] = Q[1][a] - Q[0][a];
        e12[a] = Q[2][a] - Q[1][a];
    }
    n[0] = e01[1] * e12[2] - e01[2] * e12[1];
    n[1] = e01[2] * e12[0] - e01[0] * e12[2];
    n[2] = e01[0] * e12[1] - e01[1] * e12[0];
    double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int a = 0; a < 3; ++a) n[a] /= nn;
    double d = n[0] * (r[0] - Q[0][0]) + n[1] * (r[1] - Q[0][1]) + n[2] * (r[2] - Q[0][2]);
    double ad = fabs(d);
    double rho[3];
    for (int a = 0; a < 3; ++a) rho[a] = r[a] - d * n[a];
    double scale2 = e01[0] * e01[0] + e01[1] * e01[1] + e01[2] * e01[2];
    double s0 = 0.0, bsum = 0.0, gu[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i) {
        const double* A = Q[i];
        const double* B = Q[(i + 1) % 3];
        double l[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        double ll = sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
        for (int a = 0; a < 3; ++a) l[a] /= ll;
        double u[3] = {l[1] * n[2] - l[2] * n[1], l[2] * n[0] - l[0] * n[2], l[0] * n[1] - l[1] * n[0]};
        double Ar[3] = {A[0] - rho[0], A[1] - rho[1], A[2] - rho[2]};
        double Br[3] = {B[0] - rho[0], B[1] - rho[1], B[2] - rho[2]};
        double lp = Br[0] * l[0] + Br[1] * l[1] + Br[2] * l[2];
        double lm = Ar[0] * l[0] + Ar[1] * l[1] + Ar[2] * l[2];
        double t0 = Ar[0] * u[0] + Ar[1] * u[1] + Ar[2] * u[2];
        double Rp = sqrt(Br[0] * Br[0] + Br[1] * Br[1] + Br[2] * Br[2] + d * d);
        double Rm = sqrt(Ar[0] * Ar[0] + Ar[1] * Ar[1] + Ar[2] * Ar[2] + d * d);
        double R02 = t0 * t0 + d * d;
        double f = 0.0;
        if (R02 > 1e-28 * scale2) {
            double num = (lp >= 0.0) ? (Rp + lp) : R02 / (Rp - lp);
            double den = (lm >= 0.0) ? (Rm + lm) : R02 / (Rm - lm);
            f = log(num / den);
        }
        double beta = atan(t0 * lp / (R02 + ad * Rp)) - atan(t0 * lm / (R02 + ad * Rm));
        s0 += t0 * f - ad * beta;
        bsum += beta;
        for (int a = 0; a < 3; ++a) gu[a] += u[a] * f;
    }
    S0 = s0;
    // principal value on the plane: |d| at rounding level counts as in-plane
    double sg = (ad <= 1e-12 * sqrt(scale2)) ? 0.0 : (d > 0.0 ? 1.0 : -1.0);
    for (int a = 0; a < 3; ++a) G[a] = -gu[a] - n[a] * sg * bsum;
}

template <class T>
struct P0Rec {
    T o[3];   // centroid
    T p[9];   // vertices relative to the centroid
    T n[3];
    T area;
    T pad[2];
};

template <class T>
__global__ void k_p0_records(const T* __restrict__ verts, const int32_t* __restrict__ faces, int64_t F,
                             P0Rec<T>* __restrict__ rec)
{
    int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double v[3][3], o[3];
    for (int t = 0; t < 3; ++t)
        for (int a = 0; a < 3; ++a) v[t][a] = (double)verts[3 * (int64_t)faces[3 * f + t] + a];
    for (int a = 0; a < 3; ++a) o[a] = (v[0][a] + v[1][a] + v[2][a]) / 3.0;
    double e01[3], e12[3], c[3];
    for (int a = 0; a < 3; ++a) {
        e01[a] = v[1][a] - v[0][a];
        e12[a] = v[2][a] - v[1][a];
    }
    c[0] = e01[1] * e12[2] - e01[2] * e12[1];
    c[1] = e01[2] * e12[0] - e01[0] * e12[2];
    c[2] = e01[0] * e12[1] - e01[1] * e12[0];
    double cn = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    P0Rec<T> r;
    for (int a = 0; a < 3; ++a) {
        r.o[a] = (T)o[a];
        r.n[a] = (T)(c[a] / cn);
    }
    for (int t = 0; t < 3; ++t)
        for (int a = 0; a < 3; ++a) r.p[3 * t + a] = (T)(v[t][a] - o[a]);
    r.area = (T)(0.5 * cn);
    r.pad[0] = r.pad[1] = (T)0;
    rec[f] = r;
}

// Far pairs with the 7-point rule on both triangles (near pairs are skipped and
// written by k_p0_near).  One lane per column face j, uniform loop over row
// faces i.  Outputs V and/or Kt (either may be null), row-major [F, ld].
constexpr int P0_ROWS = 64;

template <class T, int NQ>
__global__ __launch_bounds__(256) void k_p0_far(const P0Rec<T>* __restrict__ rec, int64_t F,
                                                const double* __restrict__ rule,
                                                const int64_t* __restrict__ near_ptr,
                                                const int32_t* __restrict__ near_idx,
                                                T* __restrict__ Vm, T* __restrict__ Km, int64_t ld)
{
    __shared__ T xs[P0_ROWS][NQ][4];  // test points (absolute) and weights
    __shared__ T ns[P0_ROWS][3];
    __shared__ T lam[NQ][4];
    const int tid = threadIdx.x;
    const int64_t i0 = (int64_t)blockIdx.y * P0_ROWS;
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + tid;
    if (tid < NQ * 4) lam[tid / 4][tid % 4] = (T)rule[tid];
    __syncthreads();
    for (int t = tid; t < P0_ROWS; t += blockDim.x) {
        int64_t i = i0 + t;
        if (i < F) {
            const P0Rec<T>& r = rec[i];
            for (int q = 0; q < NQ; ++q) {
                for (int a = 0; a < 3; ++a)
                    xs[t][q][a] = r.o[a] + lam[q][0] * r.p[a] + lam[q][1] * r.p[3 + a] + lam[q][2] * r.p[6 + a];
                xs[t][q][3] = lam[q][3] * (T)2 * r.area;
            }
            for (int a = 0; a < 3; ++a) ns[t][a] = r.n[a];
        }
    }
    __syncthreads();
    if (j >= F) return;
    const P0Rec<T> rj = rec[j];
    T y[NQ][3], wy[NQ];
    for (int q = 0; q < NQ; ++q) {
        for (int a = 0; a < 3; ++a)
            y[q][a] = rj.o[a] + lam[q][0] * rj.p[a] + lam[q][1] * rj.p[3 + a] + lam[q][2] * rj.p[6 + a];
        wy[q] = lam[q][3] * (T)2 * rj.area * (T)MBA_INV4PI;
    }
    int64_t np = near_ptr[j];
    const int64_t npe = near_ptr[j + 1];
    while (np < npe && near_idx[np] < i0) ++np;
    const int nrows = (int)((F - i0) < P0_ROWS ? (F - i0) : P0_ROWS);
    for (int t = 0; t < nrows; ++t) {
        const int64_t i = i0 + t;
        if (np < npe && near_idx[np] == (int32_t)i) {
            ++np;
            continue;
        }
        T sv = 0, sk = 0;
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            const T x0 = xs[t][p][0], x1 = xs[t][p][1], x2 = xs[t][p][2];
            T gv = 0, gk0 = 0, gk1 = 0, gk2 = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const T d0 = x0 - y[q][0], d1 = x1 - y[q][1], d2 = x2 - y[q][2];
                const T rinv = rsqrt_t(d0 * d0 + d1 * d1 + d2 * d2);
                const T g = wy[q] * rinv;
                const T g3 = g * rinv * rinv;
                gv += g;
                gk0 += g3 * d0;
                gk1 += g3 * d1;
                gk2 += g3 * d2;
            }
            const T wx = xs[t][p][3];
            sv += wx * gv;
            sk -= wx * (gk0 * ns[t][0] + gk1 * ns[t][1] + gk2 * ns[t][2]);
        }
        if (Vm) Vm[i * ld + j] = sv;
        if (Km) Km[i * ld + j] = sk;
    }
}

// Near pairs (CSR entry q: column face j, row face i): outer 7-point rule on
// T_i, inner integral over T_j analytic in FP64.
template <class T>
__global__ void k_p0_near(const P0Rec<T>* __restrict__ rec, int64_t F, const double* __restrict__ rule,
                          int nq, const int64_t* __restrict__ near_ptr,
                          const int32_t* __restrict__ near_idx, T* __restrict__ Vm, T* __restrict__ Km,
                          int64_t ld)
{
    const int64_t qn = near_ptr[F];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < qn; q += stride) {
        int64_t lo = 0, hi = F;
        while (hi - lo > 1) {
            int64_t mid = (lo + hi) / 2;
            if (near_ptr[mid] <= q) lo = mid; else hi = mid;
        }
        const int64_t j = lo, i = near_idx[q];
        const P0Rec<T>& ri = rec[i];
        const P0Rec<T>& rj = rec[j];
        double Qv[3][3], D[3];
        for (int a = 0; a < 3; ++a) D[a] = (double)ri.o[a] - (double)rj.o[a];
        for (int t = 0; t < 3; ++t)
            for (int a = 0; a < 3; ++a) Qv[t][a] = (double)rj.p[3 * t + a];
        double sv = 0.0, sk = 0.0;
        for (int p = 0; p < nq; ++p) {
            double x[3];
            for (int a = 0; a < 3; ++a)
                x[a] = D[a] + rule[4 * p] * (double)ri.p[a] + rule[4 * p + 1] * (double)ri.p[3 + a] +
                       rule[4 * p + 2] * (double)ri.p[6 + a];
            double wx = rule[4 * p + 3] * 2.0 * (double)ri.area * MBA_INV4PI;
            double S0, G[3];
            tri_S0_grad(Qv, x, S0, G);
            sv += wx * S0;
            sk += wx * ((double)ri.n[0] * G[0] + (double)ri.n[1] * G[1] + (double)ri.n[2] * G[2]);
        }
        if (Vm) Vm[i * ld + j] = (T)sv;
        if (Km) Km[i * ld + j] = (T)sk;
    }
}

// ------------------------------------------------------------ NS P1
// Galerkin P1 layers in the unpacked layout of the reference rule,
//   Vu[fy, fx, iy, ix] = int_Tx int_Ty phi_ix(x) phi_iy(y) / (4 pi R)
//   Ku[fy, fx, iy, ix] = int_Tx int_Ty phi_ix(x) phi_iy(y) (x - y).n_y / (4 pi R^3)
// (fy the row face carrying the density, fx the test face).  Far pairs: the
// 7-point rule on both triangles; near pairs: the rule on T_x and the inner
// integral over T_y in closed form (FP64):
//   int phi/R = phi(rho) S0 + grad phi . S1,
//   int phi (x - y).n_y / R^3 = -phi(rho) dS0/dn + h grad phi . grad_t S0,
// S1 = int (y - rho)/R, h = (x - Y0).n_y, rho the projection of x.

// S0 = int 1/R, S1 = int (y - rho)/R, G = grad_r S0 over triangle Q at r;
// also the unit normal n, the signed height d of r and rho.
__device__ void tri_S01_grad(const double Q[3][3], const double r[3], double& S0, double S1[3], double G[3],
                             double n[3], double& d, double rho[3])
{
    double e01[3], e12[3];
    for (int a = 0; a < 3; ++a) {
        e01[a] = Q[1][a] - Q[0][a];
        e12[a] = Q[2][a] - Q[1][a];
    }
    n[0] = e01[1] * e12[2] - e01[2] * e12[1];
    n[1] = e01[2] * e12[0] - e01[0] * e12[2];
    n[2] = e01[0] * e12[1] - e01[1] * e12[0];
    const double nn = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int a = 0; a < 3; ++a) n[a] /= nn;
    d = n[0] * (r[0] - Q[0][0]) + n[1] * (r[1] - Q[0][1]) + n[2] * (r[2] - Q[0][2]);
    const double ad = fabs(d);
    for (int a = 0; a < 3; ++a) rho[a] = r[a] - d * n[a];
    const double scale2 = e01[0] * e01[0] + e01[1] * e01[1] + e01[2] * e01[2];
    double s0 = 0.0, bsum = 0.0, gu[3] = {0.0, 0.0, 0.0}, s1[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < 3; ++i) {
        const double* A = Q[i];
        const double* B = Q[(i + 1) % 3];
        double l[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
        const double ll = sqrt(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
        for (int a = 0; a < 3; ++a) l[a] /= ll;
        const double u[3] = {l[1] * n[2] - l[2] * n[1], l[2] * n[0] - l[0] * n[2], l[0] * n[1] - l[1] * n[0]};
        const double Ar[3] = {A[0] - rho[0], A[1] - rho[1], A[2] - rho[2]};
        const double Br[3] = {B[0] - rho[0], B[1] - rho[1], B[2] - rho[2]};
        const double lp = Br[0] * l[0] + Br[1] * l[1] + Br[2] * l[2];
        const double lm = Ar[0] * l[0] + Ar[1] * l[1] + Ar[2] * l[2];
        const double t0 = Ar[0] * u[0] + Ar[1] * u[1] + Ar[2] * u[2];
        const double Rp = sqrt(Br[0] * Br[0] + Br[1] * Br[1] + Br[2] * Br[2] + d * d);
        const double Rm = sqrt(Ar[0] * Ar[0] + Ar[1] * Ar[1] + Ar[2] * Ar[2] + d * d);
        const double R02 = t0 * t0 + d * d;
        double f = 0.0;
        if (R02 > 1e-28 * scale2) {
            const double num = (lp >= 0.0) ? (Rp + lp) : R02 / (Rp - lp);
            const double den = (lm >= 0.0) ? (Rm + lm) : R02 / (Rm - lm);
            f = log(num / den);
        }
        const double beta = atan(t0 * lp / (R02 + ad * Rp)) - atan(t0 * lm / (R02 + ad * Rm));
        s0 += t0 * f - ad * beta;
        bsum += beta;
        const double c = 0.5 * (R02 * f + lp * Rp - lm * Rm);
        for (int a = 0; a < 3; ++a) {
            gu[a] += u[a] * f;
            s1[a] += u[a] * c;
        }
    }
    S0 = s0;
    const double sg = (ad <= 1e-12 * sqrt(scale2)) ? 0.0 : (d > 0.0 ? 1.0 : -1.0);
    for (int a = 0; a < 3; ++a) {
        G[a] = -gu[a] - n[a] * sg * bsum;
        S1[a] = s1[a];
    }
}

constexpr int P1_ROWS = 32;

template <class T, int NQ>
__global__ __launch_bounds__(256) void k_p1_far(const P0Rec<T>* __restrict__ rec, int64_t F,
                                                const double* __restrict__ rule,
                                                const int64_t* __restrict__ near_ptr,
                                                const int32_t* __restrict__ near_idx, T* __restrict__ Vu,
                                                T* __restrict__ Ku)
{
    __shared__ T ys[P1_ROWS][NQ][4];  // row-face points (absolute) and weights 2 A w / (4 pi)
    __shared__ T nys[P1_ROWS][3];
    __shared__ T lam[NQ][4];
    const int tid = threadIdx.x;
    const int64_t fy0 = (int64_t)blockIdx.y * P1_ROWS;
    const int64_t fx = (int64_t)blockIdx.x * blockDim.x + tid;
    if (tid < NQ * 4) lam[tid / 4][tid % 4] = (T)rule[tid];
    __syncthreads();
    for (int t = tid; t < P1_ROWS; t += blockDim.x) {
        const int64_t fy = fy0 + t;
        if (fy < F) {
            const P0Rec<T>& r = rec[fy];
            for (int q = 0; q < NQ; ++q) {
                for (int a = 0; a < 3; ++a)
                    ys[t][q][a] = r.o[a] + lam[q][0] * r.p[a] + lam[q][1] * r.p[3 + a] + lam[q][2] * r.p[6 + a];
                ys[t][q][3] = lam[q][3] * (T)2 * r.area * (T)MBA_INV4PI;
            }
            for (int a = 0; a < 3; ++a) nys[t][a] = r.n[a];
        }
    }
    __syncthreads();
    if (fx >= F) return;
    const P0Rec<T> rx = rec[fx];
    T x[NQ][3], wx[NQ];
    for (int p = 0; p < NQ; ++p) {
        for (int a = 0; a < 3; ++a)
            x[p][a] = rx.o[a] + lam[p][0] * rx.p[a] + lam[p][1] * rx.p[3 + a] + lam[p][2] * rx.p[6 + a];
        wx[p] = lam[p][3] * (T)2 * rx.area;
    }
    // the near relation is symmetric: the rows near fx are fx's own CSR list
    int64_t np = near_ptr[fx];
    const int64_t npe = near_ptr[fx + 1];
    while (np < npe && near_idx[np] < fy0) ++np;
    const int nrows = (int)((F - fy0) < P1_ROWS ? (F - fy0) : P1_ROWS);
    for (int t = 0; t < nrows; ++t) {
        const int64_t fy = fy0 + t;
        if (np < npe && near_idx[np] == (int32_t)fy) {
            ++np;
            continue;
        }
        const T n0 = nys[t][0], n1 = nys[t][1], n2 = nys[t][2];
        T v[9], k[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) v[q] = k[q] = (T)0;
#pragma unroll
        for (int p = 0; p < NQ; ++p) {
            T iv0 = 0, iv1 = 0, iv2 = 0, ik0 = 0, ik1 = 0, ik2 = 0;
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const T d0 = x[p][0] - ys[t][q][0], d1 = x[p][1] - ys[t][q][1], d2 = x[p][2] - ys[t][q][2];
                const T rinv = rsqrt_t(d0 * d0 + d1 * d1 + d2 * d2);
                const T g = ys[t][q][3] * rinv;
                const T gk = g * rinv * rinv * (d0 * n0 + d1 * n1 + d2 * n2);
                iv0 += g * lam[q][0];
                iv1 += g * lam[q][1];
                iv2 += g * lam[q][2];
                ik0 += gk * lam[q][0];
                ik1 += gk * lam[q][1];
                ik2 += gk * lam[q][2];
            }
            const T iv[3] = {iv0, iv1, iv2}, ik[3] = {ik0, ik1, ik2};
#pragma unroll
            for (int iy = 0; iy < 3; ++iy)
#pragma unroll
                for (int ix = 0; ix < 3; ++ix) {
                    const T wl = wx[p] * lam[p][ix];
                    v[iy * 3 + ix] += wl * iv[iy];
                    k[iy * 3 + ix] += wl * ik[iy];
                }
        }
        const int64_t o = (fy * F + fx) * 9;
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            if (Vu) Vu[o + q] = v[q];
            if (Ku) Ku[o + q] = k[q];
        }
    }
}

// Near pairs: CSR entry e of owner fx with row fy = near_idx[e] (each ordered
// near pair once, the relation being symmetric).
template <class T>
__global__ void k_p1_near(const P0Rec<T>* __restrict__ rec, int64_t F, const double* __restrict__ rule, int nq,
                          const int64_t* __restrict__ near_ptr, const int32_t* __restrict__ near_idx,
                          T* __restrict__ Vu, T* __restrict__ Ku)
{
    const int64_t qn = near_ptr[F];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < qn; e += stride) {
        int64_t lo = 0, hi = F;
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) / 2;
            if (near_ptr[mid] <= e) lo = mid; else hi = mid;
        }
        const int64_t fx = lo, fy = near_idx[e];
        const P0Rec<T>& rx = rec[fx];
        const P0Rec<T>& ry = rec[fy];
        // coordinates relative to the row face's centroid
        double Qv[3][3], D[3];
        for (int a = 0; a < 3; ++a) D[a] = (double)rx.o[a] - (double)ry.o[a];
        for (int t = 0; t < 3; ++t)
            for (int a = 0; a < 3; ++a) Qv[t][a] = (double)ry.p[3 * t + a];
        // in-plane gradients of the row face's hat functions: n x (Q_i+2 - Q_i+1) / (2A)
        double gphi[3][3];
        {
            double e01[3], e12[3], c[3];
            for (int a = 0; a < 3; ++a) {
                e01[a] = Qv[1][a] - Qv[0][a];
                e12[a] = Qv[2][a] - Qv[1][a];
            }
            c[0] = e01[1] * e12[2] - e01[2] * e12[1];
            c[1] = e01[2] * e12[0] - e01[0] * e12[2];
            c[2] = e01[0] * e12[1] - e01[1] * e12[0];
            const double a2 = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
            for (int i = 0; i < 3; ++i) {
                const double* A = Qv[(i + 1) % 3];
                const double* B = Qv[(i + 2) % 3];
                const double ed[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
                // (c / a2) x ed / a2
                gphi[i][0] = (c[1] * ed[2] - c[2] * ed[1]) / (a2 * a2);
                gphi[i][1] = (c[2] * ed[0] - c[0] * ed[2]) / (a2 * a2);
                gphi[i][2] = (c[0] * ed[1] - c[1] * ed[0]) / (a2 * a2);
            }
        }
        double v[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, k[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int p = 0; p < nq; ++p) {
            const double l0 = rule[4 * p], l1 = rule[4 * p + 1], l2 = rule[4 * p + 2];
            double xp[3];
            for (int a = 0; a < 3; ++a)
                xp[a] = D[a] + l0 * (double)rx.p[a] + l1 * (double)rx.p[3 + a] + l2 * (double)rx.p[6 + a];
            const double wx = rule[4 * p + 3] * 2.0 * (double)rx.area * MBA_INV4PI;
            double S0, S1[3], G[3], n[3], h, rho[3];
            tri_S01_grad(Qv, xp, S0, S1, G, n, h, rho);
            const double Gn = G[0] * n[0] + G[1] * n[1] + G[2] * n[2];
            const double Gt[3] = {G[0] - Gn * n[0], G[1] - Gn * n[1], G[2] - Gn * n[2]};
            const double lx[3] = {l0, l1, l2};
            for (int iy = 0; iy < 3; ++iy) {
                const double* A = Qv[(iy + 1) % 3];
                const double phr = gphi[iy][0] * (rho[0] - A[0]) + gphi[iy][1] * (rho[1] - A[1]) +
                                   gphi[iy][2] * (rho[2] - A[2]);
                const double iv = phr * S0 + gphi[iy][0] * S1[0] + gphi[iy][1] * S1[1] + gphi[iy][2] * S1[2];
                const double ik = -phr * Gn + h * (gphi[iy][0] * Gt[0] + gphi[iy][1] * Gt[1] + gphi[iy][2] * Gt[2]);
                for (int ix = 0; ix < 3; ++ix) {
                    v[iy * 3 + ix] += wx * lx[ix] * iv;
                    k[iy * 3 + ix] += wx * lx[ix] * ik;
                }
            }
        }
        const int64_t o = (fy * F + fx) * 9;
        for (int q = 0; q < 9; ++q) {
            if (Vu) Vu[o + q] = (T)v[q];
            if (Ku) Ku[o + q] = (T)k[q];
        }
    }
}

// out[v] = alpha (M g)[v] + beta out[v], M the P1 mass matrix:
// (M g)[v] = sum over faces f at v of A_f / 12 (g_v + g_a + g_b + g_c).
template <class T>
__global__ void k_mass_apply(const T* __restrict__ verts, const int32_t* __restrict__ faces,
                             const int64_t* __restrict__ vptr, const int32_t* __restrict__ vidx, int64_t V,
                             const T* __restrict__ g, T alpha, T beta, T* __restrict__ out)
{
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= V) return;
    double acc = 0.0;
    for (int64_t p = vptr[v]; p < vptr[v + 1]; ++p) {
        const int64_t f = vidx[p] / 3;
        double P[3][3], gs = (double)g[v];
        for (int t = 0; t < 3; ++t) {
            const int32_t id = faces[3 * f + t];
            gs += (double)g[id];
            for (int a = 0; a < 3; ++a) P[t][a] = (double)verts[3 * (int64_t)id + a];
        }
        double e01[3], e12[3];
        for (int a = 0; a < 3; ++a) {
            e01[a] = P[1][a] - P[0][a];
            e12[a] = P[2][a] - P[1][a];
        }
        const double c0 = e01[1] * e12[2] - e01[2] * e12[1], c1 = e01[2] * e12[0] - e01[0] * e12[2],
                     c2 = e01[0] * e12[1] - e01[1] * e12[0];
        acc += 0.5 * sqrt(c0 * c0 + c1 * c1 + c2 * c2) / 12.0 * gs;
    }
    out[v] = (T)((double)alpha * acc + (beta != (T)0 ? (double)beta * (double)out[v] : 0.0));
}

template <class T>
int p1_assemble(const void* verts, const int32_t* faces, int64_t F, const double* rule, int nq,
                const int64_t* near_ptr, const int32_t* near_idx, void* Vu, void* Ku, void* ws, hipStream_t st)
{
    P0Rec<T>* rec = (P0Rec<T>*)ws;
    hipLaunchKernelGGL(k_p0_records<T>, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, st,
                       (const T*)verts, faces, F, rec);
    MBA_LAUNCH_CHECK();
    dim3 grid((unsigned)((F + 255) / 256), (unsigned)((F + P1_ROWS - 1) / P1_ROWS));
    if (nq == 7)
        hipLaunchKernelGGL((k_p1_far<T, 7>), grid, dim3(256), 0, st, rec, F, rule, near_ptr, near_idx, (T*)Vu,
                           (T*)Ku);
    else if (nq == 3)
        hipLaunchKernelGGL((k_p1_far<T, 3>), grid, dim3(256), 0, st, rec, F, rule, near_ptr, near_idx, (T*)Vu,
                           (T*)Ku);
    else
        return MBA_ERR_ARG;
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_p1_near<T>, dim3(1024), dim3(64), 0, st, rec, F, rule, nq, near_ptr, near_idx, (T*)Vu,
                       (T*)Ku);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

template <class T>
int laplace_ref(const void* verts, const int32_t* faces, const void* areas, const void* normals,
                int64_t F, const void* tab, const int64_t* tab_off, int layer, int nphi, void* out,
                hipStream_t st)
{
    dim3 grid((unsigned)((F + 255) / 256), (unsigned)((F + LAP_ROWS - 1) / LAP_ROWS));
    hipLaunchKernelGGL(k_laplace_ref<T>, grid, dim3(256), 0, st, (const T*)verts, faces, (const T*)areas,
                       (const T*)normals, F, (const T*)tab, tab_off, layer, nphi, (T*)out);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

template <class T>
int p0_assemble(const void* verts, const int32_t* faces, int64_t F, const double* rule, int nq,
                const int64_t* near_ptr, const int32_t* near_idx, void* Vm, void* Km, int64_t ld,
                void* ws, hipStream_t st)
{
    P0Rec<T>* rec = (P0Rec<T>*)ws;
    hipLaunchKernelGGL(k_p0_records<T>, dim3((unsigned)((F + 255) / 256)), dim3(256), 0, st,
                       (const T*)verts, faces, F, rec);
    MBA_LAUNCH_CHECK();
    dim3 grid((unsigned)((F + 255) / 256), (unsigned)((F + P0_ROWS - 1) / P0_ROWS));
    if (nq == 7)
        hipLaunchKernelGGL((k_p0_far<T, 7>), grid, dim3(256), 0, st, rec, F, rule, near_ptr, near_idx,
                           (T*)Vm, (T*)Km, ld);
    else if (nq == 3)
        hipLaunchKernelGGL((k_p0_far<T, 3>), grid, dim3(256), 0, st, rec, F, rule, near_ptr, near_idx,
                           (T*)Vm, (T*)Km, ld);
    else
        return MBA_ERR_ARG;
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_p0_near<T>, dim3(1024), dim3(64), 0, st, rec, F, rule, nq, near_ptr, near_idx,
                       (T*)Vm, (T*)Km, ld);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // namespace

extern "C" {

int mba_laplace_layer_ref(int prec, const void* verts, const int32_t* faces, const void* face_areas,
                          const void* face_normals, int64_t n_verts, int64_t n_faces, const void* table,
                          const int64_t* table_offsets, int32_t layer, int32_t nphi, void* out,
                          mba_stream_t stream)
{
    if (!verts || !faces || !face_areas || !face_normals || !table || !table_offsets || !out)
        return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if ((layer != 1 && layer != 2) || (nphi != 1 && nphi != 3)) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_verts <= 0 || n_faces > (1 << 20)) return MBA_ERR_SIZE;
    if (prec == MBA_F32)
        return laplace_ref<float>(verts, faces, face_areas, face_normals, n_faces, table, table_offsets,
                                  layer, nphi, out, S(stream));
    return laplace_ref<double>(verts, faces, face_areas, face_normals, n_faces, table, table_offsets,
                               layer, nphi, out, S(stream));
}

int mba_laplace_pack(int prec, const void* unpacked, int64_t n_faces, int32_t nphi,
                     const int64_t* vinc_ptr, const int32_t* vinc_idx, int64_t n_verts, void* out,
                     mba_stream_t stream)
{
    if (!unpacked || !vinc_ptr || !vinc_idx || !out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (nphi != 1 && nphi != 3) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_verts <= 0) return MBA_ERR_SIZE;
    const int64_t tot = n_verts * n_verts;
    dim3 grid((unsigned)((tot + 255) / 256));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_pack<float>, grid, dim3(256), 0, S(stream), (const float*)unpacked, n_faces,
                           nphi, vinc_ptr, vinc_idx, n_verts, (float*)out);
    else
        hipLaunchKernelGGL(k_pack<double>, grid, dim3(256), 0, S(stream), (const double*)unpacked, n_faces,
                           nphi, vinc_ptr, vinc_idx, n_verts, (double*)out);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_laplace_dirichlet_rhs(int prec, const void* k_tight, int64_t n_verts, const void* g,
                              const void* vert_areas, void* rhs, mba_stream_t stream)
{
    if (!k_tight || !g || !vert_areas || !rhs) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_verts <= 0) return MBA_ERR_SIZE;
    dim3 grid((unsigned)((n_verts + 255) / 256));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_dirichlet_rhs<float>, grid, dim3(256), 0, S(stream), (const float*)k_tight,
                           n_verts, (const float*)g, (const float*)vert_areas, (float*)rhs);
    else
        hipLaunchKernelGGL(k_dirichlet_rhs<double>, grid, dim3(256), 0, S(stream), (const double*)k_tight,
                           n_verts, (const double*)g, (const double*)vert_areas, (double*)rhs);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

size_t mba_laplace_p0_workspace_bytes(int prec, int64_t n_faces)
{
    if (n_faces <= 0) return 0;
    return (size_t)n_faces * (prec == MBA_F32 ? sizeof(P0Rec<float>) : sizeof(P0Rec<double>));
}

int mba_laplace_p0_assemble(int prec, const void* verts, const int32_t* faces, int64_t n_verts,
                            int64_t n_faces, const double* rule, int32_t nq, const int64_t* near_ptr,
                            const int32_t* near_idx, void* v_out, void* kt_out, int64_t ld,
                            void* workspace, size_t workspace_bytes, mba_stream_t stream)
{
    if (!verts || !faces || !rule || !near_ptr || !near_idx || !workspace) return MBA_ERR_ARG;
    if (!v_out && !kt_out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_verts <= 0 || n_faces > ((int64_t)1 << 30)) return MBA_ERR_SIZE;
    if (ld < n_faces) return MBA_ERR_ARG;
    if (workspace_bytes < mba_laplace_p0_workspace_bytes(prec, n_faces)) return MBA_ERR_ALLOC;
    if (prec == MBA_F32)
        return p0_assemble<float>(verts, faces, n_faces, rule, nq, near_ptr, near_idx, v_out, kt_out, ld,
                                  workspace, S(stream));
    return p0_assemble<double>(verts, faces, n_faces, rule, nq, near_ptr, near_idx, v_out, kt_out, ld,
                               workspace, S(stream));
}

int mba_laplace_p1_assemble(int prec, const void* verts, const int32_t* faces, int64_t n_verts, int64_t n_faces,
                            const double* rule, int32_t nq, const int64_t* near_ptr, const int32_t* near_idx,
                            void* v_unpacked, void* k_unpacked, void* workspace, size_t workspace_bytes,
                            mba_stream_t stream)
{
    if (!verts || !faces || !rule || !near_ptr || !near_idx || !workspace) return MBA_ERR_ARG;
    if (!v_unpacked && !k_unpacked) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (nq != 3 && nq != 7) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_verts <= 0 || n_faces > (1 << 20)) return MBA_ERR_SIZE;
    if (workspace_bytes < mba_laplace_p0_workspace_bytes(prec, n_faces)) return MBA_ERR_ALLOC;
    if (prec == MBA_F32)
        return p1_assemble<float>(verts, faces, n_faces, rule, nq, near_ptr, near_idx, v_unpacked, k_unpacked,
                                  workspace, S(stream));
    return p1_assemble<double>(verts, faces, n_faces, rule, nq, near_ptr, near_idx, v_unpacked, k_unpacked,
                               workspace, S(stream));
}

int mba_laplace_mass_apply(int prec, const void* verts, const int32_t* faces, const int64_t* vinc_ptr,
                           const int32_t* vinc_idx, int64_t n_verts, const void* g, double alpha, double beta,
                           void* out, mba_stream_t stream)
{
    if (!verts || !faces || !vinc_ptr || !vinc_idx || !g || !out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_verts <= 0) return MBA_ERR_SIZE;
    dim3 grid((unsigned)((n_verts + 255) / 256));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_mass_apply<float>, grid, dim3(256), 0, S(stream), (const float*)verts, faces, vinc_ptr,
                           vinc_idx, n_verts, (const float*)g, (float)alpha, (float)beta, (float*)out);
    else
        hipLaunchKernelGGL(k_mass_apply<double>, grid, dim3(256), 0, S(stream), (const double*)verts, faces,
                           vinc_ptr, vinc_idx, n_verts, (const double*)g, alpha, beta, (double*)out);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // extern "C"
