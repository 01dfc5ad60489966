/*
 * CPU FP64 oracle for the mesh-BEM hot path.  TEST INFRASTRUCTURE ONLY: it is
 * loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
 * leg as the checker, never by the product path.
 *
 * Contents
 *   ref mode (literal restatement of src/mesh_BEM/cuda/mesh_MOM_3d.cu):
 *     orc_ref_zmat_faceedge   kernel_create_MOM_Zmat_3d_forward   (.cu:332-496)
 *     orc_ref_rhs_faceedge    kernel_create_MOM_rhs_3d_forward    (.cu:244-330)
 *     orc_interp_faceedge     kernel_interpolate_Icoeff_to_Js_3d_forward (.cu:498-565)
 *   ns mode (EFIE, 7-point rule, analytic 1/R on near pairs):
 *     orc_ns_zmat_faceedge, orc_ns_rhs_faceedge
 *   RWG accumulation (src/mesh_BEM/mesh_BEM_solver_3d.py:128-164):
 *     orc_gather_rwg, orc_gather_rhs
 *   analytic triangle potentials (Wilton et al. 1984, Graglia 1993):
 *     orc_tri_potentials
 *   Laplace single/double layer on a relation-typed point table
 *   (src/mesh_BEM/utils/integrations_3d.py:252-452): orc_laplace_layer
 *
 * Every complex array is interleaved (re, im) double, row-major.
 * Parity status of each function is recorded in oracle/README.md.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

typedef double complex cplx;
#define PI 3.14159265358979323846

static inline void sub3(const double *a, const double *b, double *o)
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
static inline double dot3(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
static inline void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
static inline double norm3(const double *a) { return sqrt(dot3(a, a)); }

/* RWG coefficient of local edge e of face f: sign * l / (2A), sign + iff
 * faces[f][(e+1)%3] < faces[f][(e+2)%3]  (.cu:154-206). */
static double rwg_coef(const double *verts, const int32_t *faces, const double *areas,
                       int64_t f, int e)
{
    int32_t i1 = faces[3 * f + (e + 1) % 3], i2 = faces[3 * f + (e + 2) % 3];
    double d[3];
    sub3(verts + 3 * i2, verts + 3 * i1, d);
    double c = norm3(d) / (2.0 * areas[f]);
    return (i1 < i2) ? c : -c;
}

/* ------------------------------------------------------------ ref mode Z
 * Z_loc(f1 e1, f2 e2) of .cu:332-496 with the vertex rotation of .cu:391-403:
 * the test points of local edge e1 are generated from (v_e1, v_e1+1, v_e1+2),
 * the source points likewise from e2.  out: [3F, 3F] complex, row = 3 f1 + e1. */
/* Terms of the reference's CFIE integrand (.cu:426-492), selectable so each
 * can be pinned on its own (EFIE part, MFIE part, identity term):
 *   ZT_ID   1/2 f1.f2 on f1 == f2                   (.cu:426-436)
 *   ZT_K    f1.(n1 x sum W f2 x grad_x G)           (.cu:472-473, 482-489)
 *   ZT_VEC  f1.(jk sum W f2 G)                      (.cu:475-476)
 *   ZT_SCA  div1 (-j/k) sum W div2 G                (.cu:479-480, 491-492)
 * The reference sums all four (ZT_ALL). */
enum { ZT_ID = 1, ZT_K = 2, ZT_VEC = 4, ZT_SCA = 8, ZT_ALL = 15 };
/* RHS terms: RT_E f.E_inc, RT_H f.(n x H_inc)  (.cu:316-326). */
enum { RT_E = 1, RT_H = 2, RT_ALL = 3 };

void orc_ref_zmat_faceedge_terms(const double *verts, const int32_t *faces, const double *areas,
                                 const double *normals, const double *gpx, const double *gwx, int nqx,
                                 const double *gpy, const double *gwy, int nqy, double k, int64_t F,
                                 int terms, double *out)
{
    const int64_t N = 3 * F;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t fe1 = 0; fe1 < N; ++fe1) {
        int64_t f1 = fe1 / 3;
        int e1 = (int)(fe1 % 3);
        const double *P0 = verts + 3 * faces[3 * f1 + e1];
        const double *P1 = verts + 3 * faces[3 * f1 + (e1 + 1) % 3];
        const double *P2 = verts + 3 * faces[3 * f1 + (e1 + 2) % 3];
        const double A1 = areas[f1];
        const double *n1 = normals + 3 * f1;
        const double c1 = rwg_coef(verts, faces, areas, f1, e1);
        const double div1 = 2.0 * c1;
        for (int64_t fe2 = 0; fe2 < N; ++fe2) {
            int64_t f2 = fe2 / 3;
            int e2 = (int)(fe2 % 3);
            const double *Q0 = verts + 3 * faces[3 * f2 + e2];
            const double *Q1 = verts + 3 * faces[3 * f2 + (e2 + 1) % 3];
            const double *Q2 = verts + 3 * faces[3 * f2 + (e2 + 2) % 3];
            const double A2 = areas[f2];
            const double c2 = rwg_coef(verts, faces, areas, f2, e2);
            const double div2 = 2.0 * c2;
            cplx acc = 0.0;
            for (int i = 0; i < nqx * nqx; ++i) {
                double r1 = gpx[i / nqx], r2 = gpx[i % nqx] * r1;
                double W1 = gwx[i / nqx] * gwx[i % nqx] * (2.0 * A1) * r1;
                double x[3], f1v[3];
                for (int a = 0; a < 3; ++a) {
                    x[a] = (1.0 - r1) * P0[a] + (r1 - r2) * P1[a] + r2 * P2[a];
                    f1v[a] = c1 * (x[a] - P0[a]);
                }
                if (f1 == f2 && (terms & ZT_ID)) {
                    double f2x[3];
                    for (int a = 0; a < 3; ++a) f2x[a] = c2 * (x[a] - Q0[a]);
                    acc += W1 * 0.5 * dot3(f1v, f2x);
                }
                cplx K[3] = {0, 0, 0}, L1[3] = {0, 0, 0}, L2 = 0;
                for (int j = 0; j < nqy * nqy; ++j) {
                    double r3 = gpy[j / nqy], r4 = gpy[j % nqy] * r3;
                    double W2 = gwy[j / nqy] * gwy[j % nqy] * (2.0 * A2) * r3;
                    double y[3], f2v[3], d[3];
                    for (int a = 0; a < 3; ++a) {
                        y[a] = (1.0 - r3) * Q0[a] + (r3 - r4) * Q1[a] + r4 * Q2[a];
                        f2v[a] = c2 * (y[a] - Q0[a]);
                    }
                    sub3(x, y, d);
                    double R = norm3(d);
                    cplx e = cexp(-I * k * R);
                    cplx G = e / (4.0 * PI * R);
                    /* grad_x G = -(1 + jkR) e^{-jkR} (x - y) / (4 pi R^3) */
                    cplx s = -(1.0 + I * k * R) * e / (4.0 * PI * R * R * R);
                    /* f2 x grad_x G */
                    double cr[3];
                    cross3(f2v, d, cr);
                    for (int a = 0; a < 3; ++a) {
                        K[a] += W2 * s * cr[a];
                        L1[a] += W2 * (I * k) * f2v[a] * G;
                    }
                    L2 += W2 * (-I / k) * div2 * G;
                }
                /* n1 x K */
                cplx nK[3];
                nK[0] = n1[1] * K[2] - n1[2] * K[1];
                nK[1] = n1[2] * K[0] - n1[0] * K[2];
                nK[2] = n1[0] * K[1] - n1[1] * K[0];
                for (int a = 0; a < 3; ++a)
                    acc += W1 * f1v[a] * (((terms & ZT_K) ? nK[a] : 0.0) + ((terms & ZT_VEC) ? L1[a] : 0.0));
                if (terms & ZT_SCA) acc += W1 * div1 * L2;
            }
            out[2 * (fe1 * N + fe2)] = creal(acc);
            out[2 * (fe1 * N + fe2) + 1] = cimag(acc);
        }
    }
}

void orc_ref_zmat_faceedge(const double *verts, const int32_t *faces, const double *areas,
                           const double *normals, const double *gpx, const double *gwx, int nqx,
                           const double *gpy, const double *gwy, int nqy, double k, int64_t F,
                           double *out)
{
    orc_ref_zmat_faceedge_terms(verts, faces, areas, normals, gpx, gwx, nqx, gpy, gwy, nqy, k, F, ZT_ALL, out);
}

/* ------------------------------------------------------------ ref mode RHS
 * E = z e^{-j k khat.x}, H = (cos(-60deg), sin(-60deg), 0) e^{-j k khat.x},
 * khat = (cos 30deg, sin 30deg, 0); V = sum W f.(E + n x H)  (.cu:24-70, 244-330).
 * out: [F, 3] complex. */
void orc_ref_rhs_faceedge_terms(const double *verts, const int32_t *faces, const double *areas,
                                const double *normals, const double *gp, const double *gw, int nq,
                                double k, int64_t F, int terms, double *out)
{
    const double az = 30.0 / 180.0 * PI;
    const double kh[3] = {cos(az), sin(az), 0.0};
    const double hd = az - 0.5 * PI;
    const double hp[3] = {cos(hd), sin(hd), 0.0};
#pragma omp parallel for schedule(static)
    for (int64_t fe = 0; fe < 3 * F; ++fe) {
        int64_t f = fe / 3;
        int e = (int)(fe % 3);
        const double *P0 = verts + 3 * faces[3 * f + e];
        const double *P1 = verts + 3 * faces[3 * f + (e + 1) % 3];
        const double *P2 = verts + 3 * faces[3 * f + (e + 2) % 3];
        const double *n = normals + 3 * f;
        const double c = rwg_coef(verts, faces, areas, f, e);
        cplx acc = 0.0;
        for (int i = 0; i < nq * nq; ++i) {
            double r1 = gp[i / nq], r2 = gp[i % nq] * r1;
            double W = gw[i / nq] * gw[i % nq] * (2.0 * areas[f]) * r1;
            double x[3], fv[3];
            for (int a = 0; a < 3; ++a) {
                x[a] = (1.0 - r1) * P0[a] + (r1 - r2) * P1[a] + r2 * P2[a];
                fv[a] = c * (x[a] - P0[a]);
            }
            cplx ph = cexp(-I * k * dot3(kh, x));
            double nxh[3];
            cross3(n, hp, nxh);
            double Ev[3] = {0.0, 0.0, 1.0};
            for (int a = 0; a < 3; ++a)
                acc += W * fv[a] * (((terms & RT_E) ? Ev[a] : 0.0) + ((terms & RT_H) ? nxh[a] : 0.0)) * ph;
        }
        out[2 * fe] = creal(acc);
        out[2 * fe + 1] = cimag(acc);
    }
}

void orc_ref_rhs_faceedge(const double *verts, const int32_t *faces, const double *areas,
                          const double *normals, const double *gp, const double *gw, int nq,
                          double k, int64_t F, double *out)
{
    orc_ref_rhs_faceedge_terms(verts, faces, areas, normals, gp, gw, nq, k, F, RT_ALL, out);
}

/* ------------------------------------------------------------ interpolation
 * face_vert_Js[b, 3f + t, :] = sum_{e != t} c_e (v_t - v_e) I[b, edge(f,e)]
 * (.cu:498-565).  I: [B, E] complex; out: [B, 3F, 3] complex. */
void orc_interp_faceedge(const double *verts, const int32_t *faces, const int32_t *fe_idx,
                         const double *areas, const double *Icoef, int64_t B, int64_t E,
                         int64_t F, double *out)
{
    memset(out, 0, sizeof(double) * 2 * B * 9 * F);
    for (int64_t b = 0; b < B; ++b)
        for (int64_t f = 0; f < F; ++f)
            for (int e = 0; e < 3; ++e) {
                int64_t m = fe_idx[3 * f + e];
                cplx Ie = Icoef[2 * (b * E + m)] + I * Icoef[2 * (b * E + m) + 1];
                double c = rwg_coef(verts, faces, areas, f, e);
                const double *P0 = verts + 3 * faces[3 * f + e];
                for (int s = 1; s <= 2; ++s) {
                    int t = (e + s) % 3;
                    const double *Pt = verts + 3 * faces[3 * f + t];
                    for (int a = 0; a < 3; ++a) {
                        cplx v = c * (Pt[a] - P0[a]) * Ie;
                        int64_t o = 2 * ((b * 3 * F + 3 * f + t) * 3 + a);
                        out[o] += creal(v);
                        out[o + 1] += cimag(v);
                    }
                }
            }
}

/* ------------------------------------------------------------ analytic potentials
 * For the triangle (P0, P1, P2) with unit normal n = normalise((P1-P0)x(P2-P1))
 * and observation point r:
 *   S0 = int_T 1/|r - y| dA
 *   S1 = int_T (y - rho)/|r - y| dA     (rho = projection of r on the plane)
 *   S0 = sum_i t_i f_i - |d| sum_i beta_i
 *   S1 = 1/2 sum_i u_i (R0_i^2 f_i + l+_i R+_i - l-_i R-_i)
 * with the edge quantities of Graglia (1993), eqs. 25-31.  rho is returned too. */
void orc_tri_potentials(const double *P0, const double *P1, const double *P2, const double *r,
                        double *S0, double *S1, double *rho_out)
{
    const double *P[3] = {P0, P1, P2};
    double e01[3], e12[3], n[3];
    sub3(P1, P0, e01);
    sub3(P2, P1, e12);
    cross3(e01, e12, n);
    double nn = norm3(n);
    for (int a = 0; a < 3; ++a) n[a] /= nn;
    double rp[3];
    sub3(r, P0, rp);
    double d = dot3(n, rp);
    double ad = fabs(d);
    double rho[3] = {r[0] - d * n[0], r[1] - d * n[1], r[2] - d * n[2]};
    double s0 = 0.0, s1[3] = {0.0, 0.0, 0.0};
    double scale2 = dot3(e01, e01);
    for (int i = 0; i < 3; ++i) {
        const double *A = P[i], *B = P[(i + 1) % 3];
        double l[3];
        sub3(B, A, l);
        double ll = norm3(l);
        for (int a = 0; a < 3; ++a) l[a] /= ll;
        double u[3];
        cross3(l, n, u);
        double Ar[3], Br[3];
        sub3(A, rho, Ar);
        sub3(B, rho, Br);
        double lp = dot3(Br, l), lm = dot3(Ar, l);
        double t0 = dot3(Ar, u);
        double Rp = sqrt(dot3(Br, Br) + d * d), Rm = sqrt(dot3(Ar, Ar) + d * d);
        double R02 = t0 * t0 + d * d;
        double f = 0.0;
        if (R02 > 1e-28 * scale2) {
            double num = (lp >= 0.0) ? (Rp + lp) : R02 / (Rp - lp);
            double den = (lm >= 0.0) ? (Rm + lm) : R02 / (Rm - lm);
            f = log(num / den);
        }
        double beta = atan(t0 * lp / (R02 + ad * Rp)) - atan(t0 * lm / (R02 + ad * Rm));
        s0 += t0 * f - ad * beta;
        double c = R02 * f + lp * Rp - lm * Rm;
        for (int a = 0; a < 3; ++a) s1[a] += 0.5 * u[a] * c;
    }
    *S0 = s0;
    for (int a = 0; a < 3; ++a) S1[a] = s1[a];
    if (rho_out)
        for (int a = 0; a < 3; ++a) rho_out[a] = rho[a];
}

/* ------------------------------------------------------------ ns mode Z
 * EFIE local block for test face a (rows 3a+i) and source face b (cols 3b+j):
 *   Z_ij = c_i c_j [ jk <(x - v_i).(y - v_j) G> - (4j/k) <G> ]
 * with G = e^{-jkR}/(4 pi R), the triangle rule (lam, w) on both faces
 * (physical weight w * 2A).  For near pairs G is split into
 * (e^{-jkR} - 1)/(4 pi R), integrated by the rule, plus 1/(4 pi R) whose
 * inner integral over b is analytic (orc_tri_potentials).
 * near_ptr/near_idx: CSR over source faces b listing the near test faces a.
 * out: [3F, 3F] complex. */
void orc_ns_zmat_faceedge_window(const double *verts, const int32_t *faces, const double *areas,
                                 const double *lam, const double *w, int nq, double k, int64_t F,
                                 const int64_t *near_ptr, const int32_t *near_idx, int64_t b_lo,
                                 int64_t b_hi, double *out)
{
    const int64_t N = 3 * (b_hi - b_lo);  /* leading dimension of out */
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = b_lo; b < b_hi; ++b) {
        const double *Q[3] = {verts + 3 * faces[3 * b], verts + 3 * faces[3 * b + 1],
                              verts + 3 * faces[3 * b + 2]};
        double ob[3], yp[64][3], wy[64];
        for (int a = 0; a < 3; ++a) ob[a] = (Q[0][a] + Q[1][a] + Q[2][a]) / 3.0;
        for (int j = 0; j < nq; ++j) {
            for (int a = 0; a < 3; ++a)
                yp[j][a] = lam[3 * j] * Q[0][a] + lam[3 * j + 1] * Q[1][a] + lam[3 * j + 2] * Q[2][a];
            wy[j] = w[j] * 2.0 * areas[b];
        }
        double cb[3];
        for (int e = 0; e < 3; ++e) cb[e] = rwg_coef(verts, faces, areas, b, e);
        int64_t np = near_ptr[b], npe = near_ptr[b + 1];
        for (int64_t a = 0; a < F; ++a) {
            int near = 0;
            while (np < npe && near_idx[np] < a) ++np;
            if (np < npe && near_idx[np] == a) near = 1;
            const double *P[3] = {verts + 3 * faces[3 * a], verts + 3 * faces[3 * a + 1],
                                  verts + 3 * faces[3 * a + 2]};
            double oa[3];
            for (int c = 0; c < 3; ++c) oa[c] = (P[0][c] + P[1][c] + P[2][c]) / 3.0;
            cplx I0 = 0, Ix[3] = {0, 0, 0}, Iy[3] = {0, 0, 0}, Ixy = 0;
            for (int i = 0; i < nq; ++i) {
                double x[3];
                for (int c = 0; c < 3; ++c)
                    x[c] = lam[3 * i] * P[0][c] + lam[3 * i + 1] * P[1][c] + lam[3 * i + 2] * P[2][c];
                double wx = w[i] * 2.0 * areas[a];
                cplx g = 0, gy[3] = {0, 0, 0};
                for (int j = 0; j < nq; ++j) {
                    double d[3];
                    sub3(x, yp[j], d);
                    double R = norm3(d);
                    cplx G;
                    if (near) {
                        if (R > 0.0)
                            G = (cexp(-I * k * R) - 1.0) / (4.0 * PI * R);
                        else
                            G = -I * k / (4.0 * PI);
                    } else {
                        G = cexp(-I * k * R) / (4.0 * PI * R);
                    }
                    g += wy[j] * G;
                    for (int c = 0; c < 3; ++c) gy[c] += wy[j] * G * (yp[j][c] - ob[c]);
                }
                if (near) {
                    double S0, S1[3], rho[3];
                    orc_tri_potentials(Q[0], Q[1], Q[2], x, &S0, S1, rho);
                    g += S0 / (4.0 * PI);
                    for (int c = 0; c < 3; ++c)
                        gy[c] += (S1[c] + (rho[c] - ob[c]) * S0) / (4.0 * PI);
                }
                double xl[3];
                sub3(x, oa, xl);
                I0 += wx * g;
                for (int c = 0; c < 3; ++c) {
                    Ix[c] += wx * xl[c] * g;
                    Iy[c] += wx * gy[c];
                }
                Ixy += wx * (xl[0] * gy[0] + xl[1] * gy[1] + xl[2] * gy[2]);
            }
            for (int i = 0; i < 3; ++i) {
                double ai[3];
                sub3(P[i], oa, ai);
                double ca = rwg_coef(verts, faces, areas, a, i);
                for (int j = 0; j < 3; ++j) {
                    double bj[3];
                    sub3(Q[j], ob, bj);
                    cplx dotv = Ixy - (bj[0] * Ix[0] + bj[1] * Ix[1] + bj[2] * Ix[2]) -
                                (ai[0] * Iy[0] + ai[1] * Iy[1] + ai[2] * Iy[2]) +
                                dot3(ai, bj) * I0;
                    cplx z = ca * cb[j] * (I * k * dotv - (4.0 * I / k) * I0);
                    int64_t o = 2 * ((3 * a + i) * N + 3 * (b - b_lo) + j);
                    out[o] = creal(z);
                    out[o + 1] = cimag(z);
                }
            }
        }
    }
}

void orc_ns_zmat_faceedge(const double *verts, const int32_t *faces, const double *areas,
                          const double *lam, const double *w, int nq, double k, int64_t F,
                          const int64_t *near_ptr, const int32_t *near_idx, double *out)
{
    orc_ns_zmat_faceedge_window(verts, faces, areas, lam, w, nq, k, F, near_ptr, near_idx, 0, F, out);
}

/* ------------------------------------------------------------ ns mode RHS
 * V[3f+e, d] = sum_i w_i c_e (x_i - v_e).p_d e^{-jk khat_d . x_i}
 * dirs: [ndir, 3] unit propagation directions, pols: [ndir, 3] real unit
 * polarisations.  out: [3F, ndir] complex. */
void orc_ns_rhs_faceedge(const double *verts, const int32_t *faces, const double *areas,
                         const double *lam, const double *w, int nq, double k, int64_t F,
                         const double *dirs, const double *pols, int64_t ndir, double *out)
{
#pragma omp parallel for schedule(static)
    for (int64_t f = 0; f < F; ++f) {
        const double *P[3] = {verts + 3 * faces[3 * f], verts + 3 * faces[3 * f + 1],
                              verts + 3 * faces[3 * f + 2]};
        for (int e = 0; e < 3; ++e) {
            double c = rwg_coef(verts, faces, areas, f, e);
            for (int64_t dd = 0; dd < ndir; ++dd) {
                cplx acc = 0;
                for (int i = 0; i < nq; ++i) {
                    double x[3];
                    for (int a = 0; a < 3; ++a)
                        x[a] = lam[3 * i] * P[0][a] + lam[3 * i + 1] * P[1][a] + lam[3 * i + 2] * P[2][a];
                    double wx = w[i] * 2.0 * areas[f];
                    double fp = 0.0;
                    for (int a = 0; a < 3; ++a) fp += c * (x[a] - P[e][a]) * pols[3 * dd + a];
                    acc += wx * fp * cexp(-I * k * dot3(dirs + 3 * dd, x));
                }
                out[2 * ((3 * f + e) * ndir + dd)] = creal(acc);
                out[2 * ((3 * f + e) * ndir + dd) + 1] = cimag(acc);
            }
        }
    }
}

/* ------------------------------------------------------------ RWG accumulation
 * Z[m, n] = sum_{s,t} ZF[fe_s(m), fe_t(n)], rwg_fe [E, 2] ascending; summation
 * order (ZF[0,0] + ZF[1,0]) + (ZF[0,1] + ZF[1,1]).  Row-major Z [E, E]. */
void orc_gather_rwg(const double *ZF, int64_t F, const int32_t *rwg_fe, int64_t E, double *Z)
{
    const int64_t N = 3 * F;
#pragma omp parallel for schedule(static)
    for (int64_t m = 0; m < E; ++m)
        for (int64_t n = 0; n < E; ++n) {
            double re = 0.0, im = 0.0;
            for (int t = 0; t < 2; ++t) {
                double pr = 0.0, pi = 0.0;
                for (int s = 0; s < 2; ++s) {
                    int64_t o = 2 * ((int64_t)rwg_fe[2 * m + s] * N + rwg_fe[2 * n + t]);
                    pr += ZF[o];
                    pi += ZF[o + 1];
                }
                re += pr;
                im += pi;
            }
            Z[2 * (m * E + n)] = re;
            Z[2 * (m * E + n) + 1] = im;
        }
}

/* V[m, d] = VF[fe_0(m), d] + VF[fe_1(m), d] */
void orc_gather_rhs(const double *VF, const int32_t *rwg_fe, int64_t E, int64_t ndir, double *V)
{
    for (int64_t m = 0; m < E; ++m)
        for (int64_t d = 0; d < ndir; ++d)
            for (int c = 0; c < 2; ++c)
                V[2 * (m * ndir + d) + c] = VF[2 * ((int64_t)rwg_fe[2 * m] * ndir + d) + c] +
                                            VF[2 * ((int64_t)rwg_fe[2 * m + 1] * ndir + d) + c];
}

/* ------------------------------------------------------------ Laplace layers
 * U[fy, fx, iy, ix] = 4 A_x A_y sum_{g} W_g phi_iy(y_g) phi_ix(x_g) f(x_g, y_g)
 * over the point table of the pair's relation type rel[fy*F+fx]:
 *   tab[rel]: n_tab[rel] rows of (r1x, r2x, r1y, r2y, W) with W = weight*jacobian,
 *   x = (1-r1x) X0 + (r1x-r2x) X1 + r2x X2 on face fx (stored vertex order),
 *   y likewise on face fy.
 *   single layer f = 1/(4 pi R); double layer f = (x - y).n_fy / (4 pi R^3).
 *   phi (LINEAR) = (1 - r1, r1 - r2, r2); PLANAR: 1 (nphi = 1).
 * (integrations_3d.py:345-450, BEM_utils.py:72-104, 141-172, 175-207.)
 * out: [F, F, nphi, nphi] double. */
/* Laplace single (layer 1: 1/(4 pi R)) / double (layer 2: (x - y).n_y /
 * (4 pi R^3), n_y the normal of the row face fy) layer on relation-typed
 * point tables (integrations_3d.py:252-452).  out[fy, fx, iy, ix] =
 * 4 A_x A_y sum_g W_g phi_iy(y_g) phi_ix(x_g) kernel, with x_g, y_g from the
 * table rows (r1x, r2x, r1y, r2y, W) mapped through each face's local vertex
 * order, as the reference maps them (no reordering).
 *
 * fix (bit mask; 0 reproduces the reference):
 *   LAP_REORDER   for same-vertex / same-edge pairs, map through local
 *                 vertex orders that put the shared vertex (edge) first in
 *                 both faces, as the Sauter-Schwab maps assume
 *   LAP_SYMMETRY  same-face pairs: add the transposed block, i.e. the three
 *                 sub-domains x <-> y the reference's maps leave out
 * With both bits and a proper (non-collapsed) 4-D rule in the tables, out
 * is the exact Galerkin integral to the rule's accuracy. */
enum { LAP_REORDER = 1, LAP_SYMMETRY = 2 };

/* Local orders of faces X (fx) and Y (fy) putting their shared vertices
 * first: shared vertex at position 0 (relation 1), shared edge at positions
 * (0, 1) with the same two vertices in both (relation 2); X keeps its cyclic
 * order, Y follows X's edge direction. */
static void shared_first(const int32_t *X, const int32_t *Y, int r, int px[3], int py[3])
{
    for (int i = 0; i < 3; ++i) px[i] = py[i] = i;
    if (r == 1) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                if (X[a] == Y[b])
                    for (int i = 0; i < 3; ++i) {
                        px[i] = (a + i) % 3;
                        py[i] = (b + i) % 3;
                    }
    } else if (r == 2) {
        int a0 = -1;
        for (int a = 0; a < 3 && a0 < 0; ++a) {
            int in_y0 = 0, in_y1 = 0;
            for (int b = 0; b < 3; ++b) {
                in_y0 |= X[a] == Y[b];
                in_y1 |= X[(a + 1) % 3] == Y[b];
            }
            if (in_y0 && in_y1) a0 = a;
        }
        if (a0 < 0) return;
        for (int i = 0; i < 3; ++i) px[i] = (a0 + i) % 3;
        for (int b = 0; b < 3; ++b) {
            if (Y[b] == X[px[0]]) py[0] = b;
            if (Y[b] == X[px[1]]) py[1] = b;
        }
        py[2] = 3 - py[0] - py[1];
    }
}

void orc_laplace_layer_ex(const double *verts, const int32_t *faces, const double *areas,
                          const double *normals, int64_t F, const uint8_t *rel,
                          const double *const *tab, const int64_t *n_tab, int layer, int nphi,
                          int fix, double *out)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t fy = 0; fy < F; ++fy) {
        const double *ny = normals + 3 * fy;
        for (int64_t fx = 0; fx < F; ++fx) {
            int r = rel[fy * F + fx];
            int px[3] = {0, 1, 2}, py[3] = {0, 1, 2};
            if ((fix & LAP_REORDER) && (r == 1 || r == 2)) shared_first(faces + 3 * fx, faces + 3 * fy, r, px, py);
            const double *X[3], *Y[3];
            for (int i = 0; i < 3; ++i) {
                X[i] = verts + 3 * faces[3 * fx + px[i]];
                Y[i] = verts + 3 * faces[3 * fy + py[i]];
            }
            const double *t = tab[r];
            double acc[9] = {0};
            for (int64_t g = 0; g < n_tab[r]; ++g) {
                double r1x = t[5 * g], r2x = t[5 * g + 1], r1y = t[5 * g + 2], r2y = t[5 * g + 3];
                double W = t[5 * g + 4];
                double bx[3] = {1.0 - r1x, r1x - r2x, r2x}, by[3] = {1.0 - r1y, r1y - r2y, r2y};
                double x[3], y[3], d[3];
                for (int a = 0; a < 3; ++a) {
                    x[a] = bx[0] * X[0][a] + bx[1] * X[1][a] + bx[2] * X[2][a];
                    y[a] = by[0] * Y[0][a] + by[1] * Y[1][a] + by[2] * Y[2][a];
                }
                sub3(x, y, d);
                double R = norm3(d);
                double fv = (layer == 1) ? 1.0 / (4.0 * PI * R)
                                         : dot3(d, ny) / (4.0 * PI * R * R * R);
                if (nphi == 1)
                    acc[0] += W * fv;
                else
                    for (int iy = 0; iy < 3; ++iy)
                        for (int ix = 0; ix < 3; ++ix) acc[py[iy] * 3 + px[ix]] += W * by[iy] * bx[ix] * fv;
            }
            if ((fix & LAP_SYMMETRY) && r == 3) {
                double sym[9];
                for (int iy = 0; iy < nphi; ++iy)
                    for (int ix = 0; ix < nphi; ++ix) sym[iy * nphi + ix] = acc[iy * nphi + ix] + acc[ix * nphi + iy];
                memcpy(acc, sym, sizeof(double) * nphi * nphi);
            }
            double s = 4.0 * areas[fx] * areas[fy];
            for (int q = 0; q < nphi * nphi; ++q) out[(fy * F + fx) * nphi * nphi + q] = s * acc[q];
        }
    }
}

void orc_laplace_layer(const double *verts, const int32_t *faces, const double *areas,
                       const double *normals, int64_t F, const uint8_t *rel,
                       const double *const *tab, const int64_t *n_tab, int layer, int nphi,
                       double *out)
{
    orc_laplace_layer_ex(verts, faces, areas, normals, F, rel, tab, n_tab, layer, nphi, 0, out);
}

/* ------------------------------------------------------------ P0 Laplace (ns)
 * grad_r S0 = -sum_i u_i f_i - n sgn(d) sum_i beta_i  (in-plane part from the
 * edge terms, normal part from the solid angle). */
void orc_tri_S0_grad(const double *P0, const double *P1, const double *P2, const double *r,
                     double *S0, double *G)
{
    const double *P[3] = {P0, P1, P2};
    double e01[3], e12[3], n[3];
    sub3(P1, P0, e01);
    sub3(P2, P1, e12);
    cross3(e01, e12, n);
    double nn = norm3(n);
    for (int a = 0; a < 3; ++a) n[a] /= nn;
    double rp[3];
    sub3(r, P0, rp);
    double d = dot3(n, rp), ad = fabs(d);
    double rho[3] = {r[0] - d * n[0], r[1] - d * n[1], r[2] - d * n[2]};
    double s0 = 0.0, bsum = 0.0, gu[3] = {0, 0, 0}, scale2 = dot3(e01, e01);
    for (int i = 0; i < 3; ++i) {
        const double *A = P[i], *B = P[(i + 1) % 3];
        double l[3];
        sub3(B, A, l);
        double ll = norm3(l);
        for (int a = 0; a < 3; ++a) l[a] /= ll;
        double u[3];
        cross3(l, n, u);
        double Ar[3], Br[3];
        sub3(A, rho, Ar);
        sub3(B, rho, Br);
        double lp = dot3(Br, l), lm = dot3(Ar, l), t0 = dot3(Ar, u);
        double Rp = sqrt(dot3(Br, Br) + d * d), Rm = sqrt(dot3(Ar, Ar) + d * d);
        double R02 = t0 * t0 + d * d, f = 0.0;
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
    *S0 = s0;
    /* principal value on the plane: |d| at rounding level counts as in-plane */
    double sg = (ad <= 1e-12 * sqrt(scale2)) ? 0.0 : (d > 0 ? 1.0 : -1.0);
    for (int a = 0; a < 3; ++a) G[a] = -gu[a] - n[a] * sg * bsum;
}

/* V[i,j] = int_Ti int_Tj 1/(4 pi R), Kt[i,j] = int_Ti int_Tj n_i.grad_x 1/(4 pi R):
 * the rule on both triangles for far pairs, analytic inner integral over T_j
 * for the near pairs of the CSR (columns j). out: [F, F] each (either may be NULL). */
void orc_laplace_p0(const double *verts, const int32_t *faces, int64_t F, const double *lam,
                    const double *w, int nq, const int64_t *near_ptr, const int32_t *near_idx,
                    double *Vm, double *Km)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t j = 0; j < F; ++j) {
        const double *Q[3] = {verts + 3 * faces[3 * j], verts + 3 * faces[3 * j + 1],
                              verts + 3 * faces[3 * j + 2]};
        double Aj, nj[3], c[3], e01[3], e12[3];
        sub3(Q[1], Q[0], e01);
        sub3(Q[2], Q[1], e12);
        cross3(e01, e12, c);
        Aj = 0.5 * norm3(c);
        (void)nj;
        int64_t np = near_ptr[j], npe = near_ptr[j + 1];
        for (int64_t i = 0; i < F; ++i) {
            int near = 0;
            while (np < npe && near_idx[np] < i) ++np;
            if (np < npe && near_idx[np] == i) near = 1;
            const double *P[3] = {verts + 3 * faces[3 * i], verts + 3 * faces[3 * i + 1],
                                  verts + 3 * faces[3 * i + 2]};
            double f01[3], f12[3], ni[3];
            sub3(P[1], P[0], f01);
            sub3(P[2], P[1], f12);
            cross3(f01, f12, ni);
            double Ai = 0.5 * norm3(ni);
            for (int a = 0; a < 3; ++a) ni[a] /= 2.0 * Ai;
            double sv = 0.0, sk = 0.0;
            for (int p = 0; p < nq; ++p) {
                double x[3];
                for (int a = 0; a < 3; ++a)
                    x[a] = lam[3 * p] * P[0][a] + lam[3 * p + 1] * P[1][a] + lam[3 * p + 2] * P[2][a];
                double wx = w[p] * 2.0 * Ai / (4.0 * PI);
                if (near) {
                    double S0, G[3];
                    orc_tri_S0_grad(Q[0], Q[1], Q[2], x, &S0, G);
                    sv += wx * S0;
                    sk += wx * dot3(ni, G);
                } else {
                    for (int q = 0; q < nq; ++q) {
                        double y[3], d[3];
                        for (int a = 0; a < 3; ++a)
                            y[a] = lam[3 * q] * Q[0][a] + lam[3 * q + 1] * Q[1][a] + lam[3 * q + 2] * Q[2][a];
                        sub3(x, y, d);
                        double R = norm3(d), wy = w[q] * 2.0 * Aj;
                        sv += wx * wy / R;
                        sk -= wx * wy * dot3(d, ni) / (R * R * R);
                    }
                }
            }
            if (Vm) Vm[i * F + j] = sv;
            if (Km) Km[i * F + j] = sk;
        }
    }
}

/* ------------------------------------------------------------ P1 Laplace (ns)
 * NS P1 Galerkin layers, unpacked like orc_laplace_layer:
 *   Vu[fy, fx, iy, ix] = int_Tx int_Ty phi_ix(x) phi_iy(y) / (4 pi R)
 *   Ku[fy, fx, iy, ix] = int_Tx int_Ty phi_ix(x) phi_iy(y) (x - y).n_y / (4 pi R^3)
 * Far pairs: the (lam, w) rule on both triangles.  Near pairs (fx in the
 * near list of fy): the rule on T_x, the inner integral over T_y in closed
 * form: int phi/R = phi(rho) S0 + grad phi . S1, and
 * int phi (x - y).n_y / R^3 = -phi(rho) (grad S0 . n_y) + h grad phi . (grad S0)_tangential,
 * h = (x - Y0).n_y, rho the projection of x on the plane of T_y. */
void orc_laplace_p1_ns(const double *verts, const int32_t *faces, int64_t F, const double *lam,
                       const double *w, int nq, const int64_t *near_ptr, const int32_t *near_idx,
                       double *Vu, double *Ku)
{
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t fy = 0; fy < F; ++fy) {
        const double *Y[3] = {verts + 3 * faces[3 * fy], verts + 3 * faces[3 * fy + 1],
                              verts + 3 * faces[3 * fy + 2]};
        double e01[3], e12[3], ny[3];
        sub3(Y[1], Y[0], e01);
        sub3(Y[2], Y[1], e12);
        cross3(e01, e12, ny);
        const double Ay = 0.5 * norm3(ny);
        for (int a = 0; a < 3; ++a) ny[a] /= 2.0 * Ay;
        double gphi[3][3];
        for (int i = 0; i < 3; ++i) {
            double e[3];
            sub3(Y[(i + 2) % 3], Y[(i + 1) % 3], e);
            cross3(ny, e, gphi[i]);
            for (int a = 0; a < 3; ++a) gphi[i][a] /= 2.0 * Ay;
        }
        int64_t np = near_ptr[fy];
        const int64_t npe = near_ptr[fy + 1];
        for (int64_t fx = 0; fx < F; ++fx) {
            int near = 0;
            while (np < npe && near_idx[np] < fx) ++np;
            if (np < npe && near_idx[np] == fx) near = 1;
            const double *X[3] = {verts + 3 * faces[3 * fx], verts + 3 * faces[3 * fx + 1],
                                  verts + 3 * faces[3 * fx + 2]};
            double f01[3], f12[3], c[3];
            sub3(X[1], X[0], f01);
            sub3(X[2], X[1], f12);
            cross3(f01, f12, c);
            const double Ax = 0.5 * norm3(c);
            double v[9] = {0}, k[9] = {0};
            for (int p = 0; p < nq; ++p) {
                double x[3];
                for (int a = 0; a < 3; ++a)
                    x[a] = lam[3 * p] * X[0][a] + lam[3 * p + 1] * X[1][a] + lam[3 * p + 2] * X[2][a];
                const double wx = w[p] * 2.0 * Ax / (4.0 * PI);
                double iv[3] = {0, 0, 0}, ik[3] = {0, 0, 0};
                if (near) {
                    double S0, S1[3], rho[3], S0b, G[3];
                    orc_tri_potentials(Y[0], Y[1], Y[2], x, &S0, S1, rho);
                    orc_tri_S0_grad(Y[0], Y[1], Y[2], x, &S0b, G);
                    double xr[3];
                    sub3(x, Y[0], xr);
                    const double h = dot3(xr, ny), Gn = dot3(G, ny);
                    double Gt[3] = {G[0] - Gn * ny[0], G[1] - Gn * ny[1], G[2] - Gn * ny[2]};
                    for (int i = 0; i < 3; ++i) {
                        double r1[3];
                        sub3(rho, Y[(i + 1) % 3], r1);
                        const double phr = dot3(gphi[i], r1);
                        iv[i] = phr * S0 + dot3(gphi[i], S1);
                        ik[i] = -phr * Gn + h * dot3(gphi[i], Gt);
                    }
                } else {
                    for (int q = 0; q < nq; ++q) {
                        double y[3], d[3];
                        for (int a = 0; a < 3; ++a)
                            y[a] = lam[3 * q] * Y[0][a] + lam[3 * q + 1] * Y[1][a] + lam[3 * q + 2] * Y[2][a];
                        sub3(x, y, d);
                        const double R = norm3(d), wy = w[q] * 2.0 * Ay;
                        const double gv = wy / R, gk = wy * dot3(d, ny) / (R * R * R);
                        for (int i = 0; i < 3; ++i) {
                            iv[i] += gv * lam[3 * q + i];
                            ik[i] += gk * lam[3 * q + i];
                        }
                    }
                }
                for (int iy = 0; iy < 3; ++iy)
                    for (int ix = 0; ix < 3; ++ix) {
                        v[iy * 3 + ix] += wx * lam[3 * p + ix] * iv[iy];
                        k[iy * 3 + ix] += wx * lam[3 * p + ix] * ik[iy];
                    }
            }
            for (int q = 0; q < 9; ++q) {
                if (Vu) Vu[(fy * F + fx) * 9 + q] = v[q];
                if (Ku) Ku[(fy * F + fx) * 9 + q] = k[q];
            }
        }
    }
}
