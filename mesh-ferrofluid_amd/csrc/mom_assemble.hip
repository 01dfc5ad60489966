// MoM assembly kernels for gfx950: face records, the NS EFIE face-pair
// kernel, the FP64 near-singular correction, the reference-rule CFIE kernel,
// plane-wave and reference RHS, RWG gather, vertex interpolation.
//
// Layout: face-edge blocks zf [3F, ldzf] row-major (row = test 3a+i, col =
// source 3b+j) as the reference's face_edge_Zmat (mesh_MOM_3d.cu:362), RWG
// matrix z [E, ldz] row-major (row = test RWG) as the reference's Zmn
// (mesh_BEM_solver_3d.py:167).
#include <math.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "mba_common.hpp"
#include "mba_guard.hpp"
#include "upload.hpp"

namespace {

// ----------------------------------------------------------------- records
// Per-face geometry in the assembly precision.  Positions are stored relative
// to the face centroid o (computed in double); o itself is split o = ohi + olo
// so that o_a - o_b = (ohi_a - ohi_b) + (olo_a - olo_b) keeps ~double accuracy
// in float arithmetic (the ohi difference is exact for nearby faces).
template <class T>
struct FaceRec {
    T ohi[3];
    T olo[3];
    T p[9];  // vertex t at p[3t..3t+2], relative to o
    T n[3];  // unit normal normalise((v1-v0)x(v2-v1))
    T area;
    T c[3];  // signed RWG coefficient of local edge e
    T rho;   // max |v_t - o|
    T pad;
};
static_assert(sizeof(FaceRec<float>) == 24 * 4, "record size");

template <class T>
__device__ __forceinline__ double ld(const T* p, int64_t i) { return (double)p[i]; }

template <class T>
__global__ void k_face_records(const T* __restrict__ verts, const int32_t* __restrict__ faces,
                               const T* __restrict__ areas, const T* __restrict__ normals,
                               int64_t F, FaceRec<T>* __restrict__ rec)
{
    int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    int32_t vi[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
    double v[3][3];
    for (int t = 0; t < 3; ++t)
        for (int a = 0; a < 3; ++a) v[t][a] = ld(verts, 3 * (int64_t)vi[t] + a);
    double o[3], e01[3], e12[3], cr[3];
    for (int a = 0; a < 3; ++a) {
        o[a] = (v[0][a] + v[1][a] + v[2][a]) / 3.0;
        e01[a] = v[1][a] - v[0][a];
        e12[a] = v[2][a] - v[1][a];
    }
    cr[0] = e01[1] * e12[2] - e01[2] * e12[1];
    cr[1] = e01[2] * e12[0] - e01[0] * e12[2];
    cr[2] = e01[0] * e12[1] - e01[1] * e12[0];
    double cn = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    double A = areas ? ld(areas, f) : 0.5 * cn;
    FaceRec<T> r;
    double rho = 0.0;
    for (int a = 0; a < 3; ++a) {
        T hi = (T)o[a];
        r.ohi[a] = hi;
        r.olo[a] = (T)(o[a] - (double)hi);
        r.n[a] = normals ? normals[3 * f + a] : (T)(cr[a] / cn);
    }
    for (int t = 0; t < 3; ++t) {
        double d2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            double d = v[t][a] - o[a];
            r.p[3 * t + a] = (T)d;
            d2 += d * d;
        }
        rho = fmax(rho, sqrt(d2));
    }
    for (int e = 0; e < 3; ++e) {
        int i1 = (e + 1) % 3, i2 = (e + 2) % 3;
        double l2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            double d = v[i2][a] - v[i1][a];
            l2 += d * d;
        }
        double c = sqrt(l2) / (2.0 * A);
        r.c[e] = (T)((vi[i1] < vi[i2]) ? c : -c);
    }
    r.area = (T)A;
    r.rho = (T)rho;
    r.pad = (T)0;
    rec[f] = r;
}

// ----------------------------------------------------------------- ns kernel
// One lane per source face b (columns 3b..3b+2), a uniform loop over a block of
// TA test faces staged in LDS.  Per (a, b) the 7x7 point pairs accumulate the
// moments I0 = <G>, Ix = <x G>, Iy = <y G>, Ixy = <x.y G> (x, y relative to
// their face centroids); the 9 local entries follow from
//   <(x - v_i).(y - v_j) G> = Ixy - v_j.Ix - v_i.Iy + (v_i.v_j) I0.
// Near pairs integrate (e^{-jkR} - 1)/(4 pi R); the analytic 1/(4 pi R) part is
// added by k_ns_near.
//
// The face-edge matrix is complex symmetric: the integrand of entry (a i, b j)
// is symmetric in the two faces and both sides use the same rule, so
// z(a i, b j) = z(b j, a i).  In FP64 (bound by its FP64 arithmetic) a pair
// whose test face a also lies in the window of source faces is evaluated
// once: for a < b the values go to both (a, b) and the mirrored (b, a) block,
// pairs with a > b are skipped.  This needs the near relation to be symmetric
// (a near b <=> b near a), as the centroid-distance criterion of the near-pair
// builders is.  In FP32 the hardware transcendentals make the kernel cheap
// enough that the mirrored block's column-wise stores cost more than they
// save, so every ordered pair is evaluated.
constexpr int NS_TA = 64;     // test faces per block
constexpr int NS_BLOCK = 256; // source faces per block

template <class T, int NQ>
struct NsTest {
    T x[NQ][3];  // test points relative to o_a
    T w[NQ];     // physical weight w * 2A
    T ohi[3], olo[3];
    T p[9];
    T c[3];
};

// One tile: test faces [by NS_TA, +NS_TA) x source faces b0 + [bx NS_BLOCK, +NS_BLOCK)
// of one mesh.  tst / lam are the block's LDS arrays (lam already loaded and
// synchronised); every thread of the block calls this with the same tile.
template <class T, int NQ>
__device__ __forceinline__ void ns_zf_tile(NsTest<T, NQ>* tst, const T (*lam)[4],
                                           const FaceRec<T>* __restrict__ rec, int64_t F, T kk, T k2pi,
                                           const int64_t* __restrict__ near_ptr,
                                           const int32_t* __restrict__ near_idx, int64_t b0, int64_t nb,
                                           cpx<T>* __restrict__ zf, int64_t ldzf, int64_t bx, int64_t by)
{
    const int tid = threadIdx.x;
    const int64_t a0 = by * NS_TA;
    const int64_t bl = bx * NS_BLOCK + tid;  // local source index
    const int64_t b = b0 + bl;
    constexpr bool SYM = sizeof(T) == 8;
    // whole tile below the diagonal of the window's square part: nothing to do
    // (its pairs come from the mirrored tile).  The test depends on the block
    // only, so every thread leaves before the barrier below.
    if (SYM) {
        const int64_t bend = bx * NS_BLOCK + NS_BLOCK < nb ? bx * NS_BLOCK + NS_BLOCK : nb;
        if (a0 >= b0 && a0 + NS_TA <= b0 + nb && a0 >= b0 + bend) return;
    }
    for (int t = tid; t < NS_TA; t += NS_BLOCK) {
        int64_t a = a0 + t;
        if (a < F) {
            const FaceRec<T>& r = rec[a];
            for (int i = 0; i < NQ; ++i) {
                for (int q = 0; q < 3; ++q)
                    tst[t].x[i][q] = lam[i][0] * r.p[q] + lam[i][1] * r.p[3 + q] + lam[i][2] * r.p[6 + q];
                tst[t].w[i] = lam[i][3] * (T)2 * r.area;
            }
            for (int q = 0; q < 3; ++q) {
                tst[t].ohi[q] = r.ohi[q];
                tst[t].olo[q] = r.olo[q];
                tst[t].c[q] = r.c[q];
            }
            for (int q = 0; q < 9; ++q) tst[t].p[q] = r.p[q];
        }
    }
    __syncthreads();
    if (bl >= nb || b >= F) return;

    const FaceRec<T> rb = rec[b];
    T y[NQ][3], wy[NQ];
    for (int j = 0; j < NQ; ++j) {
        for (int q = 0; q < 3; ++q)
            y[j][q] = lam[j][0] * rb.p[q] + lam[j][1] * rb.p[3 + q] + lam[j][2] * rb.p[6 + q];
        wy[j] = lam[j][3] * (T)2 * rb.area * (T)MBA_INV4PI;
    }
    int64_t np = near_ptr[b];
    const int64_t npe = near_ptr[b + 1];
    while (np < npe && near_idx[np] < a0) ++np;

    const int nt = (int)((F - a0) < NS_TA ? (F - a0) : NS_TA);
    const T kinv4 = (T)4 / kk;
    for (int t = 0; t < nt; ++t) {
        const int64_t a = a0 + t;
        T nearf = (T)0;
        if (np < npe && near_idx[np] == (int32_t)a) {
            nearf = (T)1;
            ++np;
        }
        const bool square = SYM && a >= b0 && a < b0 + nb;  // a is also a source face of this window
        if (square && a > b) continue;
        const NsTest<T, NQ>& ta = tst[t];
        T D[3];
        for (int q = 0; q < 3; ++q) D[q] = (ta.ohi[q] - rb.ohi[q]) + (ta.olo[q] - rb.olo[q]);
        T I0r = 0, I0i = 0, Ixr[3] = {0, 0, 0}, Ixi[3] = {0, 0, 0}, Iyr[3] = {0, 0, 0},
          Iyi[3] = {0, 0, 0}, Ixyr = 0, Ixyi = 0;
        // FP64: keep the test-point loop rolled (register pressure of the
        // unrolled 7 x 7 body limits FP64 occupancy to one wave per SIMD)
#pragma unroll(sizeof(T) == 4 ? NQ : 1)
        for (int i = 0; i < NQ; ++i) {
            const T xl0 = ta.x[i][0], xl1 = ta.x[i][1], xl2 = ta.x[i][2];
            const T x0 = xl0 + D[0], x1 = xl1 + D[1], x2 = xl2 + D[2];
            T gr = 0, gi = 0, gyr[3] = {0, 0, 0}, gyi[3] = {0, 0, 0};
#pragma unroll
            for (int j = 0; j < NQ; ++j) {
                const T d0 = x0 - y[j][0], d1 = x1 - y[j][1], d2 = x2 - y[j][2];
                const T R2 = d0 * d0 + d1 * d1 + d2 * d2;
                const bool zero = (R2 == (T)0);
                const T rinv = zero ? (T)0 : rsqrt_t(R2);
                const T R = R2 * rinv;
                T cs, sn;
                cis_rev(R * k2pi, cs, sn);
                const T wr = wy[j] * rinv;
                const T Gr = (cs - nearf) * wr;
                const T Gi = zero ? -kk * wy[j] : -sn * wr;
                gr += Gr;
                gi += Gi;
                gyr[0] += Gr * y[j][0];
                gyi[0] += Gi * y[j][0];
                gyr[1] += Gr * y[j][1];
                gyi[1] += Gi * y[j][1];
                gyr[2] += Gr * y[j][2];
                gyi[2] += Gi * y[j][2];
            }
            const T w = ta.w[i];
            I0r += w * gr;
            I0i += w * gi;
            Ixr[0] += w * xl0 * gr;
            Ixi[0] += w * xl0 * gi;
            Ixr[1] += w * xl1 * gr;
            Ixi[1] += w * xl1 * gi;
            Ixr[2] += w * xl2 * gr;
            Ixi[2] += w * xl2 * gi;
            Iyr[0] += w * gyr[0];
            Iyi[0] += w * gyi[0];
            Iyr[1] += w * gyr[1];
            Iyi[1] += w * gyi[1];
            Iyr[2] += w * gyr[2];
            Iyi[2] += w * gyi[2];
            Ixyr += w * (xl0 * gyr[0] + xl1 * gyr[1] + xl2 * gyr[2]);
            Ixyi += w * (xl0 * gyi[0] + xl1 * gyi[1] + xl2 * gyi[2]);
        }
        // 9 local entries: z = c_i c_j [ jk dot - (4j/k) I0 ]; mirrored to
        // rows 3b + j, columns 3(a - b0) + i when a < b lies in the window
        cpx<T>* row0 = zf + (3 * a) * ldzf + 3 * bl;
        const bool mirror = square && a < b;
        cpx<T>* rowm = mirror ? zf + (3 * b) * ldzf + 3 * (a - b0) : row0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const T ai0 = ta.p[3 * i], ai1 = ta.p[3 * i + 1], ai2 = ta.p[3 * i + 2];
            const T aIyr = ai0 * Iyr[0] + ai1 * Iyr[1] + ai2 * Iyr[2];
            const T aIyi = ai0 * Iyi[0] + ai1 * Iyi[1] + ai2 * Iyi[2];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const T bj0 = rb.p[3 * j], bj1 = rb.p[3 * j + 1], bj2 = rb.p[3 * j + 2];
                const T ab = ai0 * bj0 + ai1 * bj1 + ai2 * bj2;
                const T dr = Ixyr - (bj0 * Ixr[0] + bj1 * Ixr[1] + bj2 * Ixr[2]) - aIyr + ab * I0r;
                const T di = Ixyi - (bj0 * Ixi[0] + bj1 * Ixi[1] + bj2 * Ixi[2]) - aIyi + ab * I0i;
                const T cc = ta.c[i] * rb.c[j];
                // jk (dr + j di) = -k di + j k dr ; -(4j/k)(I0r + j I0i) = (4/k) I0i - j (4/k) I0r
                const T zr = cc * (-kk * di + kinv4 * I0i);
                const T zi = cc * (kk * dr - kinv4 * I0r);
                store_cpx(row0 + i * ldzf + j, zr, zi);
                if (mirror) store_cpx(rowm + j * ldzf + i, zr, zi);
            }
        }
    }
}

template <class T, int NQ>
__global__ __launch_bounds__(NS_BLOCK) void k_ns_zf(
    const FaceRec<T>* __restrict__ rec, int64_t F, const double* __restrict__ rule, T kk,
    T k2pi, const int64_t* __restrict__ near_ptr, const int32_t* __restrict__ near_idx,
    int64_t b0, int64_t nb, cpx<T>* __restrict__ zf, int64_t ldzf)
{
    __shared__ NsTest<T, NQ> tst[NS_TA];
    __shared__ T lam[NQ][4];
    if (threadIdx.x < NQ * 4) lam[threadIdx.x / 4][threadIdx.x % 4] = (T)rule[threadIdx.x];
    __syncthreads();
    ns_zf_tile<T, NQ>(tst, lam, rec, F, kk, k2pi, near_ptr, near_idx, b0, nb, zf, ldzf, blockIdx.x,
                      blockIdx.y);
}

// Several meshes in one launch: a 1-D grid over the tiles of all meshes
// (tile_off: prefix of the per-mesh tile counts); each block finds its mesh by
// bisection and runs the single-mesh tile on that mesh's records and output.
struct NsMeshDev {
    int64_t F;
    int64_t face_off;
    double k;
    const int64_t* near_ptr;
    const int32_t* near_idx;
    void* zf;
    int64_t ldzf;
    int64_t n_near;
};

__device__ __forceinline__ int find_segment(const int64_t* __restrict__ off, int n, int64_t t)
{
    int lo = 0, hi = n;  // off[lo] <= t < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= t) lo = mid; else hi = mid;
    }
    return lo;
}

template <class T, int NQ>
__global__ __launch_bounds__(NS_BLOCK) void k_ns_zf_batched(const FaceRec<T>* __restrict__ rec_all,
                                                            const NsMeshDev* __restrict__ meshes,
                                                            const int64_t* __restrict__ tile_off, int nmesh,
                                                            const double* __restrict__ rule)
{
    __shared__ NsTest<T, NQ> tst[NS_TA];
    __shared__ T lam[NQ][4];
    if (threadIdx.x < NQ * 4) lam[threadIdx.x / 4][threadIdx.x % 4] = (T)rule[threadIdx.x];
    __syncthreads();
    const int64_t t = blockIdx.x;
    const int m = find_segment(tile_off, nmesh, t);
    const NsMeshDev md = meshes[m];
    const int64_t lt = t - tile_off[m];
    const int64_t nbx = (md.F + NS_BLOCK - 1) / NS_BLOCK;
    ns_zf_tile<T, NQ>(tst, lam, rec_all + md.face_off, md.F, (T)md.k, (T)(md.k / (2.0 * MBA_PI)), md.near_ptr,
                      md.near_idx, 0, md.F, (cpx<T>*)md.zf, md.ldzf, lt % nbx, lt / nbx);
}

// ----------------------------------------------------------------- near (FP64)
// Analytic potentials of triangle Q (vertices relative to the origin of r):
//   S0 = int 1/|r-y|, S1 = int (y - rho)/|r-y|  (Graglia 1993, eqs. 25-31).
__device__ void tri_potentials(const double Q[3][3], const double r[3], double& S0,
                               double S1[3], double rho[3])
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
    for (int a = 0; a < 3; ++a) rho[a] = r[a] - d * n[a];
    double scale2 = e01[0] * e01[0] + e01[1] * e01[1] + e01[2] * e01[2];
    double s0 = 0.0, s1[3] = {0.0, 0.0, 0.0};
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
        double c = R02 * f + lp * Rp - lm * Rm;
        for (int a = 0; a < 3; ++a) s1[a] += 0.5 * u[a] * c;
    }
    S0 = s0;
    for (int a = 0; a < 3; ++a) S1[a] = s1[a];
}

template <class T>
__device__ void near_pair(const FaceRec<T>* __restrict__ rec, const double* __restrict__ rule,
                          int nq, double kk, const int64_t* __restrict__ near_ptr,
                          const int32_t* __restrict__ near_idx, int64_t b0, int64_t nb, int64_t q,
                          cpx<T>* __restrict__ zf, int64_t ldzf);

// One thread per near pair (CSR entry q: source b, test a).  Adds the analytic
// 1/(4 pi R) part of the inner integral, in FP64, to the 9 local entries.
template <class T>
__global__ void k_ns_near(const FaceRec<T>* __restrict__ rec, int64_t F,
                          const double* __restrict__ rule, int nq, double kk,
                          const int64_t* __restrict__ near_ptr, const int32_t* __restrict__ near_idx,
                          int64_t b0, int64_t nb, cpx<T>* __restrict__ zf, int64_t ldzf)
{
    const int64_t q0 = near_ptr[b0], q1 = near_ptr[b0 + nb];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = q0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < q1; q += stride)
        near_pair<T>(rec, rule, nq, kk, near_ptr, near_idx, b0, nb, q, zf, ldzf);
}

// Near pairs of several meshes: a grid-stride loop over all CSR entries
// (near_off: prefix of the per-mesh counts).
template <class T>
__global__ void k_ns_near_batched(const FaceRec<T>* __restrict__ rec_all, const NsMeshDev* __restrict__ meshes,
                                  const int64_t* __restrict__ near_off, int nmesh,
                                  const double* __restrict__ rule, int nq)
{
    const int64_t total = near_off[nmesh];
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total; q += stride) {
        const int m = find_segment(near_off, nmesh, q);
        const NsMeshDev& md = meshes[m];
        near_pair<T>(rec_all + md.face_off, rule, nq, md.k, md.near_ptr, md.near_idx, 0, md.F, q - near_off[m],
                     (cpx<T>*)md.zf, md.ldzf);
    }
}

template <class T>
__device__ void near_pair(const FaceRec<T>* __restrict__ rec, const double* __restrict__ rule,
                          int nq, double kk, const int64_t* __restrict__ near_ptr,
                          const int32_t* __restrict__ near_idx, int64_t b0, int64_t nb, int64_t q,
                          cpx<T>* __restrict__ zf, int64_t ldzf)
{
    // source face: last b with near_ptr[b] <= q
    int64_t lo = b0, hi = b0 + nb;  // invariant near_ptr[lo] <= q < near_ptr[hi]
    while (hi - lo > 1) {
        int64_t mid = (lo + hi) / 2;
        if (near_ptr[mid] <= q) lo = mid; else hi = mid;
    }
    const int64_t b = lo, a = near_idx[q];
    const FaceRec<T>& ra = rec[a];
    const FaceRec<T>& rb = rec[b];
    // work in the frame of o_b: Q relative to o_b, test points = o_a - o_b + x_l
    double D[3], Qv[3][3], pa[3][3];
    for (int c = 0; c < 3; ++c)
        D[c] = ((double)ra.ohi[c] - (double)rb.ohi[c]) + ((double)ra.olo[c] - (double)rb.olo[c]);
    for (int t = 0; t < 3; ++t)
        for (int c = 0; c < 3; ++c) {
            Qv[t][c] = (double)rb.p[3 * t + c];
            pa[t][c] = (double)ra.p[3 * t + c];
        }
    const double Aa = (double)ra.area;
    double I0 = 0, Ix[3] = {0, 0, 0}, Iy[3] = {0, 0, 0}, Ixy = 0;
    for (int i = 0; i < nq; ++i) {
        double xl[3], x[3];
        for (int c = 0; c < 3; ++c) {
            xl[c] = rule[4 * i] * pa[0][c] + rule[4 * i + 1] * pa[1][c] + rule[4 * i + 2] * pa[2][c];
            x[c] = xl[c] + D[c];
        }
        double w = rule[4 * i + 3] * 2.0 * Aa * MBA_INV4PI;
        double S0, S1[3], rho[3];
        tri_potentials(Qv, x, S0, S1, rho);
        double gy[3];
        for (int c = 0; c < 3; ++c) gy[c] = S1[c] + rho[c] * S0;  // int (y - o_b)/R
        I0 += w * S0;
        for (int c = 0; c < 3; ++c) {
            Ix[c] += w * xl[c] * S0;
            Iy[c] += w * gy[c];
        }
        Ixy += w * (xl[0] * gy[0] + xl[1] * gy[1] + xl[2] * gy[2]);
    }
    cpx<T>* row0 = zf + (3 * a) * ldzf + 3 * (b - b0);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double dv = Ixy - (Qv[j][0] * Ix[0] + Qv[j][1] * Ix[1] + Qv[j][2] * Ix[2]) -
                        (pa[i][0] * Iy[0] + pa[i][1] * Iy[1] + pa[i][2] * Iy[2]) +
                        (pa[i][0] * Qv[j][0] + pa[i][1] * Qv[j][1] + pa[i][2] * Qv[j][2]) * I0;
            double cc = (double)ra.c[i] * (double)rb.c[j];
            // real-valued kernel: jk dv - (4j/k) I0 is purely imaginary
            double zi = cc * (kk * dv - 4.0 / kk * I0);
            cpx<T>* p = row0 + i * ldzf + j;
            cpx<T> v = *p;
            v.im = (T)((double)v.im + zi);
            *p = v;
        }
}

// ----------------------------------------------------------------- ref kernel
// Literal CFIE of mesh_MOM_3d.cu:332-496 on an output tile: one lane per source
// face-edge column (b, e2), a uniform loop over a block of test face-edge rows
// (a, e1) staged in LDS.  Points of local edge e are generated from the
// vertices (v_e, v_e+1, v_e+2) with the collapsed rule
// x = (1 - r1) P0 + (r1 - r2) P1 + r2 P2.
template <class T>
struct RefRow {
    T n[3];
    T c1, div1;
    T P0[3];
    int32_t face;
    int32_t pad;
};

template <class T>
__global__ __launch_bounds__(256) void k_ref_zf(const FaceRec<T>* __restrict__ rec, int64_t F,
                                                const T* __restrict__ gpx, const T* __restrict__ gwx,
                                                int nqx, const T* __restrict__ gpy,
                                                const T* __restrict__ gwy, int nqy, T kk, T k2pi,
                                                int rows_per_block, cpx<T>* __restrict__ zf,
                                                int64_t ldzf)
{
    extern __shared__ unsigned char smem_raw[];
    const int NX = nqx * nqx, NY = nqy * nqy;
    RefRow<T>* rows = (RefRow<T>*)smem_raw;
    T* xs = (T*)(rows + rows_per_block);  // [rows][NX][7]: x(3), W1, f1(3)
    T* ly = xs + (size_t)rows_per_block * NX * 7;  // [NY][3]: r3, r4, wy
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_block;
    const int64_t col = (int64_t)blockIdx.x * blockDim.x + tid;
    for (int j = tid; j < NY; j += blockDim.x) {
        T r3 = gpy[j / nqy];
        ly[3 * j] = r3;
        ly[3 * j + 1] = gpy[j % nqy] * r3;
        ly[3 * j + 2] = gwy[j / nqy] * gwy[j % nqy] * r3;
    }
    for (int t = tid; t < rows_per_block; t += blockDim.x) {
        int64_t row = r0 + t;
        if (row >= 3 * F) continue;
        int64_t a = row / 3;
        int e1 = (int)(row % 3);
        const FaceRec<T>& r = rec[a];
        RefRow<T> rr;
        for (int q = 0; q < 3; ++q) {
            rr.n[q] = r.n[q];
            rr.P0[q] = r.p[3 * e1 + q];
        }
        rr.c1 = r.c[e1];
        rr.div1 = (T)2 * r.c[e1];
        rr.face = (int32_t)a;
        rr.pad = 0;
        rows[t] = rr;
        const T* P0 = r.p + 3 * e1;
        const T* P1 = r.p + 3 * ((e1 + 1) % 3);
        const T* P2 = r.p + 3 * ((e1 + 2) % 3);
        for (int i = 0; i < NX; ++i) {
            T r1 = gpx[i / nqx], r2 = gpx[i % nqx] * r1;
            T* o = xs + ((size_t)t * NX + i) * 7;
            for (int q = 0; q < 3; ++q) {
                T x = ((T)1 - r1) * P0[q] + (r1 - r2) * P1[q] + r2 * P2[q];
                o[q] = x;
                o[4 + q] = r.c[e1] * (x - P0[q]);
            }
            o[3] = gwx[i / nqx] * gwx[i % nqx] * ((T)2 * r.area) * r1;
        }
    }
    __syncthreads();
    if (col >= 3 * F) return;
    const int64_t b = col / 3;
    const int e2 = (int)(col % 3);
    const FaceRec<T> rb = rec[b];
    T Q0[3], E1[3], E2[3];
    for (int q = 0; q < 3; ++q) {
        Q0[q] = rb.p[3 * e2 + q];
        E1[q] = rb.p[3 * ((e2 + 1) % 3) + q] - Q0[q];
        E2[q] = rb.p[3 * ((e2 + 2) % 3) + q] - rb.p[3 * ((e2 + 1) % 3) + q];
    }
    const T c2 = rb.c[e2], div2 = (T)2 * rb.c[e2];
    const T twoA2 = (T)2 * rb.area;
    const T inv4pi = (T)MBA_INV4PI;
    const int nrows = (int)((3 * F - r0) < rows_per_block ? (3 * F - r0) : rows_per_block);
    for (int t = 0; t < nrows; ++t) {
        const RefRow<T> rr = rows[t];
        const FaceRec<T>& ra = rec[rr.face];
        T D[3];
        for (int q = 0; q < 3; ++q) D[q] = (ra.ohi[q] - rb.ohi[q]) + (ra.olo[q] - rb.olo[q]);
        const bool same = ((int64_t)rr.face == b);
        T accr = 0, acci = 0;
        for (int i = 0; i < NX; ++i) {
            const T* xo = xs + ((size_t)t * NX + i) * 7;
            const T xl0 = xo[0], xl1 = xo[1], xl2 = xo[2];
            const T W1 = xo[3];
            const T f10 = xo[4], f11 = xo[5], f12 = xo[6];
            const T x0 = xl0 + D[0], x1 = xl1 + D[1], x2 = xl2 + D[2];
            if (same) {
                // 1/2 f1 . f2 at the test point (.cu:426-436)
                T f20 = c2 * (xl0 - Q0[0]), f21 = c2 * (xl1 - Q0[1]), f22 = c2 * (xl2 - Q0[2]);
                accr += W1 * (T)0.5 * (f10 * f20 + f11 * f21 + f12 * f22);
            }
            T Kr[3] = {0, 0, 0}, Ki[3] = {0, 0, 0}, L1r[3] = {0, 0, 0}, L1i[3] = {0, 0, 0};
            T L2r = 0, L2i = 0;
            for (int j = 0; j < NY; ++j) {
                const T r3 = ly[3 * j], r4 = ly[3 * j + 1];
                const T W2 = ly[3 * j + 2] * twoA2;
                // y - Q0 = r3 (P1 - P0) + r4 (P2 - P1)
                const T u0 = r3 * E1[0] + r4 * E2[0];
                const T u1 = r3 * E1[1] + r4 * E2[1];
                const T u2 = r3 * E1[2] + r4 * E2[2];
                const T d0 = x0 - (Q0[0] + u0), d1 = x1 - (Q0[1] + u1), d2 = x2 - (Q0[2] + u2);
                const T R2 = d0 * d0 + d1 * d1 + d2 * d2;
                const T rinv = rsqrt_t(R2);
                const T R = R2 * rinv;
                T cs, sn;
                cis_rev(R * k2pi, cs, sn);
                // G = (cs - j sn) rinv / 4pi
                const T g = W2 * rinv * inv4pi;
                const T Gr = cs * g, Gi = -sn * g;
                // grad_x G = -(1 + jkR)(cs - j sn) rinv^3 / 4pi * d
                const T kR = kk * R;
                const T h = g * rinv * rinv;
                const T sr = -(cs + kR * sn) * h;
                const T si = -(kR * cs - sn) * h;
                const T f20 = c2 * u0, f21 = c2 * u1, f22 = c2 * u2;
                const T cr0 = f21 * d2 - f22 * d1;
                const T cr1 = f22 * d0 - f20 * d2;
                const T cr2 = f20 * d1 - f21 * d0;
                Kr[0] += sr * cr0;
                Ki[0] += si * cr0;
                Kr[1] += sr * cr1;
                Ki[1] += si * cr1;
                Kr[2] += sr * cr2;
                Ki[2] += si * cr2;
                // L1 = jk f2 G
                L1r[0] += -kk * f20 * Gi;
                L1i[0] += kk * f20 * Gr;
                L1r[1] += -kk * f21 * Gi;
                L1i[1] += kk * f21 * Gr;
                L1r[2] += -kk * f22 * Gi;
                L1i[2] += kk * f22 * Gr;
                L2r += Gr;
                L2i += Gi;
            }
            // L2 = -(j/k) div2 sum(G)
            const T l2r = div2 / kk * L2i, l2i = -div2 / kk * L2r;
            const T* n1 = rr.n;
            const T nKr0 = n1[1] * Kr[2] - n1[2] * Kr[1], nKi0 = n1[1] * Ki[2] - n1[2] * Ki[1];
            const T nKr1 = n1[2] * Kr[0] - n1[0] * Kr[2], nKi1 = n1[2] * Ki[0] - n1[0] * Ki[2];
            const T nKr2 = n1[0] * Kr[1] - n1[1] * Kr[0], nKi2 = n1[0] * Ki[1] - n1[1] * Ki[0];
            accr += W1 * (f10 * (nKr0 + L1r[0]) + f11 * (nKr1 + L1r[1]) + f12 * (nKr2 + L1r[2]) +
                          rr.div1 * l2r);
            acci += W1 * (f10 * (nKi0 + L1i[0]) + f11 * (nKi1 + L1i[1]) + f12 * (nKi2 + L1i[2]) +
                          rr.div1 * l2i);
        }
        store_cpx(zf + (r0 + t) * ldzf + col, accr, acci);
    }
}

// ----------------------------------------------------------------- RHS
// Reference incidence (mesh_MOM_3d.cu:24-70): E = z e^{-jk khat.x},
// H = (cos(-60deg), sin(-60deg), 0) e^{-jk khat.x}, khat = (cos 30deg, sin 30deg, 0).
template <class T>
__global__ void k_ref_rhs(const FaceRec<T>* __restrict__ rec, int64_t F, const T* __restrict__ gp,
                          const T* __restrict__ gw, int nq, T k2pi, cpx<T>* __restrict__ vf)
{
    int64_t fe = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (fe >= 3 * F) return;
    int64_t f = fe / 3;
    int e = (int)(fe % 3);
    const FaceRec<T> r = rec[f];
    const double az = MBA_PI / 6.0, hd = MBA_PI / 6.0 - 0.5 * MBA_PI;
    const T kh0 = (T)cos(az), kh1 = (T)sin(az);
    const T h0 = (T)cos(hd), h1 = (T)sin(hd);
    // n x H direction (H has zero z component)
    const T nh0 = -r.n[2] * h1, nh1 = r.n[2] * h0, nh2 = r.n[0] * h1 - r.n[1] * h0;
    const T* P0 = r.p + 3 * e;
    const T* P1 = r.p + 3 * ((e + 1) % 3);
    const T* P2 = r.p + 3 * ((e + 2) % 3);
    const T c = r.c[e];
    // phase of the centroid in revolutions, split off for accuracy
    const T ph0 = (kh0 * (r.ohi[0] + r.olo[0]) + kh1 * (r.ohi[1] + r.olo[1])) * k2pi;
    T accr = 0, acci = 0;
    for (int i = 0; i < nq * nq; ++i) {
        T r1 = gp[i / nq], r2 = gp[i % nq] * r1;
        T W = gw[i / nq] * gw[i % nq] * ((T)2 * r.area) * r1;
        T x[3], fv[3];
        for (int q = 0; q < 3; ++q) {
            x[q] = ((T)1 - r1) * P0[q] + (r1 - r2) * P1[q] + r2 * P2[q];
            fv[q] = c * (x[q] - P0[q]);
        }
        T cs, sn;
        cis_rev(ph0 + (kh0 * x[0] + kh1 * x[1]) * k2pi, cs, sn);
        T proj = fv[0] * nh0 + fv[1] * nh1 + fv[2] * ((T)1 + nh2);
        accr += W * proj * cs;
        acci += -W * proj * sn;
    }
    store_cpx(vf + fe, accr, acci);
}

template <class T>
__global__ void k_ns_rhs(const FaceRec<T>* __restrict__ rec, int64_t F,
                         const double* __restrict__ rule, int nq, T k2pi,
                         const double* __restrict__ dirs, const double* __restrict__ pols,
                         int64_t ndir, cpx<T>* __restrict__ vf)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= F * ndir) return;
    int64_t f = id / ndir, d = id % ndir;
    const FaceRec<T> r = rec[f];
    const T kh[3] = {(T)dirs[3 * d], (T)dirs[3 * d + 1], (T)dirs[3 * d + 2]};
    const T pl[3] = {(T)pols[3 * d], (T)pols[3 * d + 1], (T)pols[3 * d + 2]};
    const T ph0 = (kh[0] * r.ohi[0] + kh[1] * r.ohi[1] + kh[2] * r.ohi[2]) * k2pi +
                  (kh[0] * r.olo[0] + kh[1] * r.olo[1] + kh[2] * r.olo[2]) * k2pi;
    T acc[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int i = 0; i < nq; ++i) {
        T x[3];
        for (int q = 0; q < 3; ++q)
            x[q] = (T)rule[4 * i] * r.p[q] + (T)rule[4 * i + 1] * r.p[3 + q] + (T)rule[4 * i + 2] * r.p[6 + q];
        T w = (T)rule[4 * i + 3] * (T)2 * r.area;
        T cs, sn;
        cis_rev(ph0 + (kh[0] * x[0] + kh[1] * x[1] + kh[2] * x[2]) * k2pi, cs, sn);
        for (int e = 0; e < 3; ++e) {
            T fp = r.c[e] * ((x[0] - r.p[3 * e]) * pl[0] + (x[1] - r.p[3 * e + 1]) * pl[1] +
                             (x[2] - r.p[3 * e + 2]) * pl[2]);
            acc[e][0] += w * fp * cs;
            acc[e][1] += -w * fp * sn;
        }
    }
    for (int e = 0; e < 3; ++e) store_cpx(vf + (3 * f + e) * ndir + d, acc[e][0], acc[e][1]);
}

// ----------------------------------------------------------------- gathers
// Workgroup: 256 consecutive columns n x `rows` consecutive rows m (8 for
// large matrices; 1 below 8192 columns, where fewer workgroups would leave
// CUs idle).  The column's window test is done once for all its rows, and the
// workgroups of one row group take consecutive logical indices on one XCD
// (workgroup i runs on XCD i % 8), so the group's face-edge rows, read at
// scattered columns, are fetched into that XCD's L2 once.
constexpr int GATHER_XCD = 8;
inline int gather_rows(int64_t E) { return E >= 8192 ? 8 : 1; }

template <class T>
__global__ __launch_bounds__(256) void k_gather_rwg(const cpx<T>* __restrict__ zf, int64_t ldzf, int64_t col0,
                                                    int64_t ncols, const int32_t* __restrict__ rwg_fe, int64_t E,
                                                    cpx<T>* __restrict__ z, int64_t ldz, int rows)
{
    const int64_t nbx = (E + 255) / 256;
    const int64_t total = nbx * ((E + rows - 1) / rows);
    const int64_t per = (int64_t)gridDim.x / GATHER_XCD;  // gridDim.x = per * GATHER_XCD >= total
    const int64_t L = (int64_t)(blockIdx.x % GATHER_XCD) * per + blockIdx.x / GATHER_XCD;
    if (L >= total) return;
    const int64_t m0 = (L / nbx) * rows;
    const int64_t n = (L % nbx) * 256 + threadIdx.x;
    if (n >= E) return;
    const int64_t c0 = (int64_t)rwg_fe[2 * n] - col0, c1 = (int64_t)rwg_fe[2 * n + 1] - col0;
    const bool in0 = (c0 >= 0 && c0 < ncols), in1 = (c1 >= 0 && c1 < ncols);
    if (!in0 && !in1) return;
    const int64_t m1 = (m0 + rows) < E ? (m0 + rows) : E;
    for (int64_t m = m0; m < m1; ++m) {
        const cpx<T>* r0 = zf + (int64_t)rwg_fe[2 * m] * ldzf;
        const cpx<T>* r1 = zf + (int64_t)rwg_fe[2 * m + 1] * ldzf;
        // z = (zf[fe0(m), fe0(n)] + zf[fe1(m), fe0(n)]) + (zf[fe0(m), fe1(n)] + zf[fe1(m), fe1(n)]):
        // the two column pairs are summed separately, so a column split over two
        // windows rounds exactly as one gathered in a single window.
        cpx<T>* out = z + m * ldz + n;
        T re, im;
        if (in0) {
            cpx<T> a = r0[c0], b = r1[c0];
            re = a.re + b.re;
            im = a.im + b.im;
        } else {
            cpx<T> o = *out;
            re = o.re;
            im = o.im;
        }
        if (in1) {
            cpx<T> a = r0[c1], b = r1[c1];
            re = re + (a.re + b.re);
            im = im + (a.im + b.im);
        }
        store_cpx(out, re, im);
    }
}

template <class T>
__global__ void k_gather_rhs(const cpx<T>* __restrict__ vf, const int32_t* __restrict__ rwg_fe,
                             int64_t E, int64_t ndir, cpx<T>* __restrict__ v)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= E * ndir) return;
    int64_t m = id / ndir, d = id % ndir;
    cpx<T> a = vf[(int64_t)rwg_fe[2 * m] * ndir + d], b = vf[(int64_t)rwg_fe[2 * m + 1] * ndir + d];
    store_cpx(v + id, a.re + b.re, a.im + b.im);
}

// ---- batched right-hand sides and gathers (many meshes, one launch each) --
struct RhsMeshDev {
    int64_t F, face_off;
    double k;
    const double* dirs;
    const double* pols;
    int64_t ndir;
    const int32_t* rwg_fe;
    int64_t E;
    void* v;
    int64_t vf_off;  // first element of this mesh's face-edge values in the workspace
};

// thread (mesh m, face f, direction d), m found in the prefix of F ndir
template <class T>
__global__ void k_ns_rhs_batched(const FaceRec<T>* __restrict__ rec_all, const double* __restrict__ rule, int nq,
                                 const RhsMeshDev* __restrict__ meshes, int n, const int64_t* __restrict__ off,
                                 cpx<T>* __restrict__ vf_all)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= off[n]) return;
    const int m = find_segment(off, n, id);
    const RhsMeshDev M = meshes[m];
    const int64_t local = id - off[m];
    const int64_t f = local / M.ndir, d = local % M.ndir;
    const FaceRec<T> r = rec_all[M.face_off + f];
    const T k2pi = (T)(M.k / (2.0 * MBA_PI));
    const T kh[3] = {(T)M.dirs[3 * d], (T)M.dirs[3 * d + 1], (T)M.dirs[3 * d + 2]};
    const T pl[3] = {(T)M.pols[3 * d], (T)M.pols[3 * d + 1], (T)M.pols[3 * d + 2]};
    const T ph0 = (kh[0] * r.ohi[0] + kh[1] * r.ohi[1] + kh[2] * r.ohi[2]) * k2pi +
                  (kh[0] * r.olo[0] + kh[1] * r.olo[1] + kh[2] * r.olo[2]) * k2pi;
    T acc[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    for (int i = 0; i < nq; ++i) {
        T x[3];
        for (int q = 0; q < 3; ++q)
            x[q] = (T)rule[4 * i] * r.p[q] + (T)rule[4 * i + 1] * r.p[3 + q] + (T)rule[4 * i + 2] * r.p[6 + q];
        T w = (T)rule[4 * i + 3] * (T)2 * r.area;
        T cs, sn;
        cis_rev(ph0 + (kh[0] * x[0] + kh[1] * x[1] + kh[2] * x[2]) * k2pi, cs, sn);
        for (int e = 0; e < 3; ++e) {
            T fp = r.c[e] * ((x[0] - r.p[3 * e]) * pl[0] + (x[1] - r.p[3 * e + 1]) * pl[1] +
                             (x[2] - r.p[3 * e + 2]) * pl[2]);
            acc[e][0] += w * fp * cs;
            acc[e][1] += -w * fp * sn;
        }
    }
    cpx<T>* vf = vf_all + M.vf_off;
    for (int e = 0; e < 3; ++e) store_cpx(vf + (3 * f + e) * M.ndir + d, acc[e][0], acc[e][1]);
}

// thread (mesh m, RWG row, direction d), m found in the prefix of E ndir
template <class T>
__global__ void k_gather_rhs_batched(const RhsMeshDev* __restrict__ meshes, int n, const int64_t* __restrict__ off,
                                     const cpx<T>* __restrict__ vf_all)
{
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= off[n]) return;
    const int m = find_segment(off, n, id);
    const RhsMeshDev M = meshes[m];
    const int64_t local = id - off[m];
    const int64_t row = local / M.ndir, d = local % M.ndir;
    const cpx<T>* vf = vf_all + M.vf_off;
    const cpx<T> a = vf[(int64_t)M.rwg_fe[2 * row] * M.ndir + d], b = vf[(int64_t)M.rwg_fe[2 * row + 1] * M.ndir + d];
    store_cpx((cpx<T>*)M.v + local, a.re + b.re, a.im + b.im);
}

struct GatherJobDev {
    const void* zf;
    int64_t ldzf;
    const int32_t* rwg_fe;
    int64_t E;
    void* z;
    int64_t ldz;
    int rows;  // RWG rows per workgroup
};

// k_gather_rwg over the full face-edge block of many meshes: logical tile L
// (XCD-grouped as in k_gather_rwg over the whole grid) -> job by the prefix of
// tile counts, then the same sums in the same order
template <class T>
__global__ __launch_bounds__(256) void k_gather_rwg_batched(const GatherJobDev* __restrict__ jobs, int n,
                                                            const int64_t* __restrict__ tile_off)
{
    const int64_t per = (int64_t)gridDim.x / GATHER_XCD;
    const int64_t L = (int64_t)(blockIdx.x % GATHER_XCD) * per + blockIdx.x / GATHER_XCD;
    if (L >= tile_off[n]) return;
    const int q = find_segment(tile_off, n, L);
    const GatherJobDev J = jobs[q];
    const int64_t t = L - tile_off[q];
    const int64_t nbx = (J.E + 255) / 256;
    const int64_t m0 = (t / nbx) * J.rows;
    const int64_t c = (t % nbx) * 256 + threadIdx.x;
    if (c >= J.E) return;
    const cpx<T>* zf = (const cpx<T>*)J.zf;
    const int64_t c0 = J.rwg_fe[2 * c], c1 = J.rwg_fe[2 * c + 1];
    const int64_t m1 = (m0 + J.rows) < J.E ? (m0 + J.rows) : J.E;
    for (int64_t m = m0; m < m1; ++m) {
        const cpx<T>* r0 = zf + (int64_t)J.rwg_fe[2 * m] * J.ldzf;
        const cpx<T>* r1 = zf + (int64_t)J.rwg_fe[2 * m + 1] * J.ldzf;
        const cpx<T> a = r0[c0], b = r1[c0], e = r0[c1], f = r1[c1];
        const T re = (a.re + b.re) + (e.re + f.re), im = (a.im + b.im) + (e.im + f.im);
        store_cpx((cpx<T>*)J.z + m * J.ldz + c, re, im);
    }
}

// ----------------------------------------------------------------- interpolation
template <class T>
__global__ void k_interp(const T* __restrict__ verts, const int32_t* __restrict__ faces,
                         const int32_t* __restrict__ fe_idx, const T* __restrict__ areas,
                         const cpx<T>* __restrict__ icoef, int64_t B, int64_t E, int64_t F,
                         cpx<T>* __restrict__ out)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= B * F) return;
    int64_t bb = id / F, f = id % F;
    int32_t vi[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
    T v[3][3];
    for (int t = 0; t < 3; ++t)
        for (int q = 0; q < 3; ++q) v[t][q] = verts[3 * (int64_t)vi[t] + q];
    T A = areas[f];
    T accr[3][3] = {{0}}, acci[3][3] = {{0}};
    for (int e = 0; e < 3; ++e) {
        int i1 = (e + 1) % 3, i2 = (e + 2) % 3;
        T l = sqrt((v[i2][0] - v[i1][0]) * (v[i2][0] - v[i1][0]) +
                   (v[i2][1] - v[i1][1]) * (v[i2][1] - v[i1][1]) +
                   (v[i2][2] - v[i1][2]) * (v[i2][2] - v[i1][2]));
        T c = l / ((T)2 * A);
        if (!(vi[i1] < vi[i2])) c = -c;
        const int32_t ei = fe_idx[3 * f + e];  // an index outside [0, E) contributes nothing
        const cpx<T> I = (ei >= 0 && ei < E) ? icoef[bb * E + ei] : cpx<T>{(T)0, (T)0};
        for (int s = 1; s <= 2; ++s) {
            int t = (e + s) % 3;
            for (int q = 0; q < 3; ++q) {
                T rw = c * (v[t][q] - v[e][q]);
                accr[t][q] += rw * I.re;
                acci[t][q] += rw * I.im;
            }
        }
    }
    for (int t = 0; t < 3; ++t)
        for (int q = 0; q < 3; ++q) store_cpx(out + ((bb * F + f) * 3 + t) * 3 + q, accr[t][q], acci[t][q]);
}

template <class T>
__global__ void k_vertex_sum(const cpx<T>* __restrict__ fvj, const int64_t* __restrict__ vptr,
                             const int32_t* __restrict__ vidx, int64_t B, int64_t V, int64_t F,
                             cpx<T>* __restrict__ vert)
{
    int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= B * V * 3) return;
    int64_t q = id % 3, v = (id / 3) % V, bb = id / (3 * V);
    T re = 0, im = 0;
    for (int64_t p = vptr[v]; p < vptr[v + 1]; ++p) {
        cpx<T> x = fvj[(bb * 3 * F + vidx[p]) * 3 + q];
        re += x.re;
        im += x.im;
    }
    store_cpx(vert + id, re, im);
}

inline hipStream_t S(mba_stream_t s) { return (hipStream_t)s; }

template <class T>
int face_records(const void* verts, const int32_t* faces, const void* areas, const void* normals,
                 int64_t F, FaceRec<T>* rec, hipStream_t st)
{
    int threads = 256;
    int64_t blocks = (F + threads - 1) / threads;
    hipLaunchKernelGGL(k_face_records<T>, dim3((unsigned)blocks), dim3(threads), 0, st,
                       (const T*)verts, faces, (const T*)areas, (const T*)normals, F, rec);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

constexpr int64_t kMaxFaces = (int64_t)1 << 30;

size_t records_bytes(int prec, int64_t F)
{
    size_t rs = prec == MBA_F32 ? sizeof(FaceRec<float>) : sizeof(FaceRec<double>);
    return ((size_t)F * rs + 255) & ~(size_t)255;
}

// Reference-rule face-edge blocks with the face records in `rec` (at least
// records_bytes(prec, F) bytes of device memory; every record is written
// before the Z kernel reads it).
template <class T>
int zmat_ref_with(const void* verts, const int32_t* faces, const void* areas, const void* normals,
                  const void* gpx, const void* gwx, int nqx, const void* gpy, const void* gwy, int nqy,
                  double k, int64_t F, void* zf, int64_t ldzf, FaceRec<T>* rec, hipStream_t st)
{
    int rc = face_records<T>(verts, faces, areas, normals, F, rec, st);
    if (rc == MBA_OK) {
        const int NX = nqx * nqx, NY = nqy * nqy;
        const size_t row_bytes = sizeof(RefRow<T>) + (size_t)NX * 7 * sizeof(T);
        const size_t budget = 48 * 1024 - (size_t)NY * 3 * sizeof(T);
        int rpb = (int)(budget / row_bytes);
        if (rpb > 64) rpb = 64;
        if (rpb < 1) rpb = 1;
        size_t shm = (size_t)rpb * row_bytes + (size_t)NY * 3 * sizeof(T);
        dim3 grid((unsigned)((3 * F + 255) / 256), (unsigned)((3 * F + rpb - 1) / rpb));
        hipLaunchKernelGGL(k_ref_zf<T>, grid, dim3(256), shm, st, rec, F, (const T*)gpx,
                           (const T*)gwx, nqx, (const T*)gpy, (const T*)gwy, nqy, (T)k,
                           (T)(k / (2.0 * MBA_PI)), rpb, (cpx<T>*)zf, ldzf);
        if (hipGetLastError() != hipSuccess) rc = MBA_ERR_LAUNCH;
    }
    return rc;
}

template <class T>
int zmat_ref(const void* verts, const int32_t* faces, const void* areas, const void* normals,
             const void* gpx, const void* gwx, int nqx, const void* gpy, const void* gwy, int nqy,
             double k, int64_t F, void* zf, int64_t ldzf, hipStream_t st)
{
    FaceRec<T>* rec = nullptr;
    if (hipMallocAsync((void**)&rec, records_bytes(sizeof(T) == 4 ? MBA_F32 : MBA_F64, F), st) != hipSuccess)
        return MBA_ERR_ALLOC;
    const int rc = zmat_ref_with<T>(verts, faces, areas, normals, gpx, gwx, nqx, gpy, gwy, nqy, k, F, zf, ldzf,
                                    rec, st);
    (void)hipFreeAsync(rec, st);
    return rc;
}

template <class T>
int zmat_ns(const void* verts, const int32_t* faces, int64_t F, const double* rule, int nq,
            double k, const int64_t* near_ptr, const int32_t* near_idx, int64_t b0, int64_t nb,
            void* zf, int64_t ldzf, void* ws, hipStream_t st)
{
    FaceRec<T>* rec = (FaceRec<T>*)ws;
    int rc = face_records<T>(verts, faces, nullptr, nullptr, F, rec, st);
    if (rc != MBA_OK) return rc;
    dim3 grid((unsigned)((nb + NS_BLOCK - 1) / NS_BLOCK), (unsigned)((F + NS_TA - 1) / NS_TA));
    const T kk = (T)k, k2pi = (T)(k / (2.0 * MBA_PI));
    if (nq == 7) {
        hipLaunchKernelGGL((k_ns_zf<T, 7>), grid, dim3(NS_BLOCK), 0, st, rec, F, rule, kk, k2pi,
                           near_ptr, near_idx, b0, nb, (cpx<T>*)zf, ldzf);
    } else if (nq == 3) {
        hipLaunchKernelGGL((k_ns_zf<T, 3>), grid, dim3(NS_BLOCK), 0, st, rec, F, rule, kk, k2pi,
                           near_ptr, near_idx, b0, nb, (cpx<T>*)zf, ldzf);
    } else if (nq == 1) {
        hipLaunchKernelGGL((k_ns_zf<T, 1>), grid, dim3(NS_BLOCK), 0, st, rec, F, rule, kk, k2pi,
                           near_ptr, near_idx, b0, nb, (cpx<T>*)zf, ldzf);
    } else {
        return MBA_ERR_ARG;
    }
    MBA_LAUNCH_CHECK();
    // near pairs of this window: a fixed grid strides over the CSR entries
    // [near_ptr[b0], near_ptr[b0 + nb]) read on the device.
    hipLaunchKernelGGL(k_ns_near<T>, dim3(1024), dim3(64), 0, st, rec, F, rule, nq, k, near_ptr,
                       near_idx, b0, nb, (cpx<T>*)zf, ldzf);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

static_assert(sizeof(NsMeshDev) == sizeof(mba_ns_mesh), "descriptor layouts must match");

size_t batched_ws_layout(int prec, int64_t total_faces, int n, size_t* off_desc, size_t* off_tile,
                         size_t* off_near)
{
    size_t o = records_bytes(prec, total_faces);
    *off_desc = o;
    o += ((size_t)n * sizeof(NsMeshDev) + 255) & ~(size_t)255;
    *off_tile = o;
    o += ((size_t)(n + 1) * sizeof(int64_t) + 255) & ~(size_t)255;
    *off_near = o;
    o += ((size_t)(n + 1) * sizeof(int64_t) + 255) & ~(size_t)255;
    return o;
}

template <class T>
int zmat_ns_batched(const void* verts, const int32_t* faces, int64_t total_faces, const double* rule, int nq,
                    const mba_ns_mesh* meshes, int n, void* ws, hipStream_t st)
{
    size_t od, ot, on;
    batched_ws_layout(sizeof(T) == 4 ? MBA_F32 : MBA_F64, total_faces, n, &od, &ot, &on);
    FaceRec<T>* rec = (FaceRec<T>*)ws;
    int rc = face_records<T>(verts, faces, nullptr, nullptr, total_faces, rec, st);
    if (rc != MBA_OK) return rc;
    std::vector<int64_t> tile_off(n + 1, 0), near_off(n + 1, 0);
    for (int m = 0; m < n; ++m) {
        const int64_t F = meshes[m].n_faces;
        tile_off[m + 1] = tile_off[m] + ((F + NS_BLOCK - 1) / NS_BLOCK) * ((F + NS_TA - 1) / NS_TA);
        near_off[m + 1] = near_off[m] + meshes[m].n_near;
    }
    char* base = (char*)ws;
    if (mba_host::upload_tables(base + od, meshes, (size_t)n * sizeof(NsMeshDev), st) != 0 ||
        mba_host::upload_tables(base + ot, tile_off.data(), (n + 1) * sizeof(int64_t), st) != 0 ||
        mba_host::upload_tables(base + on, near_off.data(), (n + 1) * sizeof(int64_t), st) != 0)
        return MBA_ERR_LAUNCH;
    const NsMeshDev* dm = (const NsMeshDev*)(base + od);
    const int64_t* dt = (const int64_t*)(base + ot);
    const int64_t* dn = (const int64_t*)(base + on);
    const int64_t tiles = tile_off[n];
    if (tiles > 0x7fffffffLL) return MBA_ERR_SIZE;
    if (nq == 7)
        hipLaunchKernelGGL((k_ns_zf_batched<T, 7>), dim3((unsigned)tiles), dim3(NS_BLOCK), 0, st, rec, dm, dt, n,
                           rule);
    else if (nq == 3)
        hipLaunchKernelGGL((k_ns_zf_batched<T, 3>), dim3((unsigned)tiles), dim3(NS_BLOCK), 0, st, rec, dm, dt, n,
                           rule);
    else
        return MBA_ERR_ARG;
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_ns_near_batched<T>, dim3(2048), dim3(64), 0, st, rec, dm, dn, n, rule, nq);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

}  // namespace

// ======================================================= no-alloc scratch
// Device memory of mba_mom_zmat_faceedge_ref_noalloc: the per-face records of
// the reference-rule assembly, MBA_NOALLOC_SCRATCH_BYTES per GPU, held by
// every GPU the library runs on until the process ends.  Host code never
// passes `scratch` itself to a kernel (in host code the name is a host-side
// stand-in, not a device address): it uses the address hipGetSymbolAddress
// returns for the current device.
static __device__ __align__(256) unsigned char scratch[MBA_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    std::once_flag once;
    void* ptr = nullptr;  // device address of `scratch` on this GPU, or null on failure
};
DeviceScratch g_scratch[kMaxDevices];

// Address of the current device's copy of `scratch`, looked up once per
// device with that device current; null if the lookup failed.
void* scratch_address()
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    DeviceScratch& d = g_scratch[dev];
    std::call_once(d.once, [&] {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) == hipSuccess) d.ptr = p;
    });
    return d.ptr;
}

}  // namespace

// ============================================================== C ABI
// ---- batched right-hand sides / gathers: host side
struct RhsLayout {
    size_t rec, vf, mesh, offf, offe, total;
    int64_t n_vf;  // complex face-edge values of all meshes
};

inline RhsLayout rhs_batched_layout(int prec, int64_t total_faces, const mba_ns_rhs_mesh* meshes, int n)
{
    const size_t csz = prec == MBA_F32 ? sizeof(cpx<float>) : sizeof(cpx<double>);
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    RhsLayout L{};
    L.n_vf = 0;
    for (int m = 0; m < n; ++m) L.n_vf += 3 * meshes[m].n_faces * meshes[m].ndir;
    L.rec = 0;
    L.vf = al(records_bytes(prec, total_faces));
    L.mesh = L.vf + al((size_t)L.n_vf * csz);
    L.offf = L.mesh + al((size_t)n * sizeof(RhsMeshDev));
    L.offe = L.offf + al((size_t)(n + 1) * sizeof(int64_t));
    L.total = L.offe + (size_t)(n + 1) * sizeof(int64_t);
    return L;
}

template <class T>
int rhs_ns_batched(const void* verts, const int32_t* faces, int64_t total_faces, const double* rule, int nq,
                   const mba_ns_rhs_mesh* meshes, int n, void* ws, hipStream_t st)
{
    const RhsLayout L = rhs_batched_layout(sizeof(T) == 4 ? MBA_F32 : MBA_F64, total_faces, meshes, n);
    char* base = (char*)ws;
    FaceRec<T>* rec = (FaceRec<T>*)(base + L.rec);
    int rc = face_records<T>(verts, faces, nullptr, nullptr, total_faces, rec, st);
    if (rc != MBA_OK) return rc;
    std::vector<RhsMeshDev> dm(n);
    std::vector<int64_t> offf(n + 1, 0), offe(n + 1, 0);
    int64_t vf_off = 0;
    for (int m = 0; m < n; ++m) {
        const mba_ns_rhs_mesh& d = meshes[m];
        dm[m] = RhsMeshDev{d.n_faces, d.face_off, d.wavenumber, d.dirs, d.pols, d.ndir, d.rwg_fe, d.n_edges, d.v,
                           vf_off};
        vf_off += 3 * d.n_faces * d.ndir;
        offf[m + 1] = offf[m] + d.n_faces * d.ndir;
        offe[m + 1] = offe[m] + d.n_edges * d.ndir;
    }
    RhsMeshDev* d_mesh = (RhsMeshDev*)(base + L.mesh);
    int64_t* d_offf = (int64_t*)(base + L.offf);
    int64_t* d_offe = (int64_t*)(base + L.offe);
    if (mba_host::upload_tables(d_mesh, dm.data(), dm.size() * sizeof(RhsMeshDev), st) != 0 ||
        mba_host::upload_tables(d_offf, offf.data(), offf.size() * sizeof(int64_t), st) != 0 ||
        mba_host::upload_tables(d_offe, offe.data(), offe.size() * sizeof(int64_t), st) != 0)
        return MBA_ERR_LAUNCH;
    cpx<T>* vf = (cpx<T>*)(base + L.vf);
    hipLaunchKernelGGL(k_ns_rhs_batched<T>, dim3((unsigned)((offf[n] + 255) / 256)), dim3(256), 0, st, rec, rule, nq,
                       (const RhsMeshDev*)d_mesh, n, (const int64_t*)d_offf, vf);
    MBA_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_gather_rhs_batched<T>, dim3((unsigned)((offe[n] + 255) / 256)), dim3(256), 0, st,
                       (const RhsMeshDev*)d_mesh, n, (const int64_t*)d_offe, (const cpx<T>*)vf);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

inline size_t gather_batched_layout(int n, size_t* off_tiles)
{
    *off_tiles = ((size_t)n * sizeof(GatherJobDev) + 255) & ~(size_t)255;
    return *off_tiles + (size_t)(n + 1) * sizeof(int64_t);
}

extern "C" {

const char* mba_status_string(int status)
{
    switch (status) {
    case MBA_OK: return "ok";
    case MBA_ERR_ARG: return "invalid argument";
    case MBA_ERR_SIZE: return "size out of range";
    case MBA_ERR_MESH: return "malformed mesh";
    case MBA_ERR_ALLOC: return "workspace too small or allocation failed";
    case MBA_ERR_LAUNCH: return "kernel launch failed";
    case MBA_ERR_NUMERIC: return "numerical failure (zero pivot or non-finite value)";
    case MBA_ERR_TIMEOUT: return "cooperative wait timed out";
    case MBA_ERR_IO: return "file missing, unreadable or unwritable";
    default: return "unknown status";
    }
}

int mba_version(void) { return 100; }

size_t mba_mom_workspace_bytes(int prec, int64_t n_faces)
{
    if (n_faces < 0) return 0;
    return records_bytes(prec, n_faces);
}

int mba_mom_zmat_faceedge_ref(int prec, const void* verts, const int32_t* faces,
                              const void* face_areas, const void* face_normals, const void* gp_x,
                              const void* gw_x, int32_t nq_x, const void* gp_y, const void* gw_y,
                              int32_t nq_y, double wavenumber, int64_t n_verts, int64_t n_faces,
                              void* zf, int64_t ldzf, mba_stream_t stream)
{
    if (!verts || !faces || !face_areas || !face_normals || !gp_x || !gw_x || !gp_y || !gw_y || !zf)
        return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_faces > kMaxFaces || n_verts <= 0) return MBA_ERR_SIZE;
    if (nq_x < 1 || nq_x > 16 || nq_y < 1 || nq_y > 16) return MBA_ERR_SIZE;
    if (ldzf < 3 * n_faces) return MBA_ERR_ARG;
    if (wavenumber == 0.0) return MBA_ERR_ARG;
    if (prec == MBA_F32)
        return zmat_ref<float>(verts, faces, face_areas, face_normals, gp_x, gw_x, nq_x, gp_y, gw_y,
                               nq_y, wavenumber, n_faces, zf, ldzf, S(stream));
    return zmat_ref<double>(verts, faces, face_areas, face_normals, gp_x, gw_x, nq_x, gp_y, gw_y,
                            nq_y, wavenumber, n_faces, zf, ldzf, S(stream));
}

int mba_mom_zmat_faceedge_ref_noalloc(int prec, const void* verts, const int32_t* faces,
                                      const void* face_areas, const void* face_normals, const void* gp_x,
                                      const void* gw_x, int32_t nq_x, const void* gp_y, const void* gw_y,
                                      int32_t nq_y, double wavenumber, int64_t n_verts, int64_t n_faces,
                                      void* zf, int64_t ldzf, mba_stream_t stream)
{
    if (!verts || !faces || !face_areas || !face_normals || !gp_x || !gw_x || !gp_y || !gw_y || !zf)
        return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_faces > kMaxFaces || n_verts <= 0) return MBA_ERR_SIZE;
    if (nq_x < 1 || nq_x > 16 || nq_y < 1 || nq_y > 16) return MBA_ERR_SIZE;
    if (ldzf < 3 * n_faces) return MBA_ERR_ARG;
    if (wavenumber == 0.0) return MBA_ERR_ARG;
    if (records_bytes(prec, n_faces) > MBA_NOALLOC_SCRATCH_BYTES) return MBA_ERR_SIZE;
    void* rec = scratch_address();
    if (!rec) return MBA_ERR_LAUNCH;
    if (prec == MBA_F32)
        return zmat_ref_with<float>(verts, faces, face_areas, face_normals, gp_x, gw_x, nq_x, gp_y, gw_y, nq_y,
                                    wavenumber, n_faces, zf, ldzf, (FaceRec<float>*)rec, S(stream));
    return zmat_ref_with<double>(verts, faces, face_areas, face_normals, gp_x, gw_x, nq_x, gp_y, gw_y, nq_y,
                                 wavenumber, n_faces, zf, ldzf, (FaceRec<double>*)rec, S(stream));
}

int mba_mom_rhs_faceedge_ref(int prec, const void* verts, const int32_t* faces,
                             const void* face_areas, const void* face_normals, const void* gp,
                             const void* gw, int32_t nq, double wavenumber, int64_t n_verts,
                             int64_t n_faces, void* vf, mba_stream_t stream)
{
    if (!verts || !faces || !face_areas || !face_normals || !gp || !gw || !vf) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_faces > kMaxFaces || n_verts <= 0 || nq < 1 || nq > 16) return MBA_ERR_SIZE;
    hipStream_t st = S(stream);
    void* rec = nullptr;
    if (hipMallocAsync(&rec, records_bytes(prec, n_faces), st) != hipSuccess) return MBA_ERR_ALLOC;
    int rc;
    int threads = 256;
    dim3 grid((unsigned)((3 * n_faces + threads - 1) / threads));
    if (prec == MBA_F32) {
        rc = face_records<float>(verts, faces, face_areas, face_normals, n_faces, (FaceRec<float>*)rec, st);
        if (rc == MBA_OK) {
            hipLaunchKernelGGL(k_ref_rhs<float>, grid, dim3(threads), 0, st, (FaceRec<float>*)rec,
                               n_faces, (const float*)gp, (const float*)gw, nq,
                               (float)(wavenumber / (2.0 * MBA_PI)), (cpx<float>*)vf);
            if (hipGetLastError() != hipSuccess) rc = MBA_ERR_LAUNCH;
        }
    } else {
        rc = face_records<double>(verts, faces, face_areas, face_normals, n_faces, (FaceRec<double>*)rec, st);
        if (rc == MBA_OK) {
            hipLaunchKernelGGL(k_ref_rhs<double>, grid, dim3(threads), 0, st, (FaceRec<double>*)rec,
                               n_faces, (const double*)gp, (const double*)gw, nq,
                               wavenumber / (2.0 * MBA_PI), (cpx<double>*)vf);
            if (hipGetLastError() != hipSuccess) rc = MBA_ERR_LAUNCH;
        }
    }
    (void)hipFreeAsync(rec, st);
    return rc;
}

int mba_mom_interp_faceedge(int prec, const void* verts, const int32_t* faces,
                            const int32_t* face_edge_indices, const void* face_areas,
                            const void* icoef, int64_t batch, int64_t n_edges, int64_t n_verts,
                            int64_t n_faces, void* out, mba_stream_t stream)
{
    if (!verts || !faces || !face_edge_indices || !face_areas || !icoef || !out) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (batch <= 0 || n_edges <= 0 || n_faces <= 0 || n_faces > kMaxFaces || n_verts <= 0)
        return MBA_ERR_SIZE;
    int threads = 256;
    dim3 grid((unsigned)((batch * n_faces + threads - 1) / threads));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_interp<float>, grid, dim3(threads), 0, S(stream), (const float*)verts,
                           faces, face_edge_indices, (const float*)face_areas,
                           (const cpx<float>*)icoef, batch, n_edges, n_faces, (cpx<float>*)out);
    else
        hipLaunchKernelGGL(k_interp<double>, grid, dim3(threads), 0, S(stream), (const double*)verts,
                           faces, face_edge_indices, (const double*)face_areas,
                           (const cpx<double>*)icoef, batch, n_edges, n_faces, (cpx<double>*)out);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_mom_vertex_sum(int prec, const void* fvj, const int64_t* vinc_ptr, const int32_t* vinc_idx,
                       int64_t batch, int64_t n_verts, int64_t n_faces, void* vert,
                       mba_stream_t stream)
{
    if (!fvj || !vinc_ptr || !vinc_idx || !vert) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (batch <= 0 || n_verts <= 0 || n_faces <= 0) return MBA_ERR_SIZE;
    int threads = 256;
    int64_t n = batch * n_verts * 3;
    dim3 grid((unsigned)((n + threads - 1) / threads));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_vertex_sum<float>, grid, dim3(threads), 0, S(stream),
                           (const cpx<float>*)fvj, vinc_ptr, vinc_idx, batch, n_verts, n_faces,
                           (cpx<float>*)vert);
    else
        hipLaunchKernelGGL(k_vertex_sum<double>, grid, dim3(threads), 0, S(stream),
                           (const cpx<double>*)fvj, vinc_ptr, vinc_idx, batch, n_verts, n_faces,
                           (cpx<double>*)vert);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_mom_zmat_faceedge_ns(int prec, const void* verts, const int32_t* faces, int64_t n_verts,
                             int64_t n_faces, const double* rule, int32_t nq, double wavenumber,
                             const int64_t* near_ptr, const int32_t* near_idx, int64_t b0,
                             int64_t nb, void* zf, int64_t ldzf, void* workspace,
                             size_t workspace_bytes, mba_stream_t stream)
{
    if (!verts || !faces || !rule || !near_ptr || !near_idx || !zf || !workspace) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_faces > kMaxFaces || n_verts <= 0) return MBA_ERR_SIZE;
    if (nq != 1 && nq != 3 && nq != 7) return MBA_ERR_ARG;
    if (b0 < 0 || nb <= 0 || b0 + nb > n_faces) return MBA_ERR_ARG;
    if (ldzf < 3 * nb) return MBA_ERR_ARG;
    if (wavenumber <= 0.0) return MBA_ERR_ARG;
    if (workspace_bytes < mba_mom_workspace_bytes(prec, n_faces)) return MBA_ERR_ALLOC;
    if (prec == MBA_F32)
        return zmat_ns<float>(verts, faces, n_faces, rule, nq, wavenumber, near_ptr, near_idx, b0,
                              nb, zf, ldzf, workspace, S(stream));
    return zmat_ns<double>(verts, faces, n_faces, rule, nq, wavenumber, near_ptr, near_idx, b0, nb,
                           zf, ldzf, workspace, S(stream));
}

int mba_mom_rhs_faceedge_ns(int prec, const void* verts, const int32_t* faces, int64_t n_verts,
                            int64_t n_faces, const double* rule, int32_t nq, double wavenumber,
                            const double* dirs, const double* pols, int64_t ndir, void* vf,
                            mba_stream_t stream)
{
    if (!verts || !faces || !rule || !dirs || !pols || !vf) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_faces <= 0 || n_faces > kMaxFaces || n_verts <= 0 || ndir <= 0 || nq < 1 || nq > 64)
        return MBA_ERR_SIZE;
    hipStream_t st = S(stream);
    void* rec = nullptr;
    if (hipMallocAsync(&rec, records_bytes(prec, n_faces), st) != hipSuccess) return MBA_ERR_ALLOC;
    int threads = 256;
    dim3 grid((unsigned)((n_faces * ndir + threads - 1) / threads));
    int rc;
    if (prec == MBA_F32) {
        rc = face_records<float>(verts, faces, nullptr, nullptr, n_faces, (FaceRec<float>*)rec, st);
        if (rc == MBA_OK) {
            hipLaunchKernelGGL(k_ns_rhs<float>, grid, dim3(threads), 0, st, (FaceRec<float>*)rec,
                               n_faces, rule, nq, (float)(wavenumber / (2.0 * MBA_PI)), dirs, pols,
                               ndir, (cpx<float>*)vf);
            if (hipGetLastError() != hipSuccess) rc = MBA_ERR_LAUNCH;
        }
    } else {
        rc = face_records<double>(verts, faces, nullptr, nullptr, n_faces, (FaceRec<double>*)rec, st);
        if (rc == MBA_OK) {
            hipLaunchKernelGGL(k_ns_rhs<double>, grid, dim3(threads), 0, st, (FaceRec<double>*)rec,
                               n_faces, rule, nq, wavenumber / (2.0 * MBA_PI), dirs, pols, ndir,
                               (cpx<double>*)vf);
            if (hipGetLastError() != hipSuccess) rc = MBA_ERR_LAUNCH;
        }
    }
    (void)hipFreeAsync(rec, st);
    return rc;
}

int mba_mom_gather_rwg(int prec, const void* zf, int64_t ldzf, int64_t col0, int64_t ncols,
                       const int32_t* rwg_fe, int64_t n_edges, void* z, int64_t ldz,
                       mba_stream_t stream)
{
    if (!zf || !rwg_fe || !z) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_edges <= 0 || n_edges > ((int64_t)1 << 31) - 1 || ncols <= 0 || col0 < 0) return MBA_ERR_SIZE;
    if (ldz < n_edges || ldzf < ncols) return MBA_ERR_ARG;
    const int rows = gather_rows(n_edges);
    const int64_t total = ((n_edges + 255) / 256) * ((n_edges + rows - 1) / rows);
    dim3 grid((unsigned)((total + GATHER_XCD - 1) / GATHER_XCD * GATHER_XCD));
    const int threads = 256;
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_gather_rwg<float>, grid, dim3(threads), 0, S(stream),
                           (const cpx<float>*)zf, ldzf, col0, ncols, rwg_fe, n_edges,
                           (cpx<float>*)z, ldz, rows);
    else
        hipLaunchKernelGGL(k_gather_rwg<double>, grid, dim3(threads), 0, S(stream),
                           (const cpx<double>*)zf, ldzf, col0, ncols, rwg_fe, n_edges,
                           (cpx<double>*)z, ldz, rows);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

int mba_mom_gather_rhs(int prec, const void* vf, const int32_t* rwg_fe, int64_t n_edges,
                       int64_t ndir, void* v, mba_stream_t stream)
{
    if (!vf || !rwg_fe || !v) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_edges <= 0 || ndir <= 0) return MBA_ERR_SIZE;
    int threads = 256;
    dim3 grid((unsigned)((n_edges * ndir + threads - 1) / threads));
    if (prec == MBA_F32)
        hipLaunchKernelGGL(k_gather_rhs<float>, grid, dim3(threads), 0, S(stream),
                           (const cpx<float>*)vf, rwg_fe, n_edges, ndir, (cpx<float>*)v);
    else
        hipLaunchKernelGGL(k_gather_rhs<double>, grid, dim3(threads), 0, S(stream),
                           (const cpx<double>*)vf, rwg_fe, n_edges, ndir, (cpx<double>*)v);
    MBA_LAUNCH_CHECK();
    return MBA_OK;
}

size_t mba_mom_batched_workspace_bytes(int prec, int64_t total_faces, int32_t n_meshes)
{
    if (total_faces <= 0 || n_meshes <= 0 || (prec != MBA_F32 && prec != MBA_F64)) return 0;
    size_t a, b, c;
    return batched_ws_layout(prec, total_faces, n_meshes, &a, &b, &c);
}

int mba_mom_zmat_faceedge_ns_batched(int prec, const void* verts, const int32_t* faces, int64_t total_faces,
                                     const double* rule, int32_t nq, const mba_ns_mesh* meshes, int32_t n_meshes,
                                     void* workspace, size_t workspace_bytes, mba_stream_t stream)
{
    if (!verts || !faces || !rule || !meshes || !workspace) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_meshes <= 0 || total_faces <= 0 || total_faces > kMaxFaces) return MBA_ERR_SIZE;
    if (nq != 3 && nq != 7) return MBA_ERR_ARG;
    if (workspace_bytes < mba_mom_batched_workspace_bytes(prec, total_faces, n_meshes)) return MBA_ERR_ALLOC;
    int64_t sum = 0;
    for (int m = 0; m < n_meshes; ++m) {
        const mba_ns_mesh& d = meshes[m];
        if (d.n_faces <= 0 || d.face_off != sum || !d.near_ptr || !d.near_idx || !d.zf) return MBA_ERR_ARG;
        if (d.ldzf < 3 * d.n_faces || d.n_near < 0 || !(d.wavenumber > 0.0)) return MBA_ERR_ARG;
        sum += d.n_faces;
    }
    if (sum != total_faces) return MBA_ERR_ARG;
    return mba_guard([&] {
        if (prec == MBA_F32)
            return zmat_ns_batched<float>(verts, faces, total_faces, rule, nq, meshes, n_meshes, workspace,
                                          S(stream));
        return zmat_ns_batched<double>(verts, faces, total_faces, rule, nq, meshes, n_meshes, workspace, S(stream));
    });
}

size_t mba_mom_rhs_ns_batched_workspace_bytes(int prec, int64_t total_faces, const mba_ns_rhs_mesh* meshes,
                                              int32_t n_meshes)
{
    if (!meshes || total_faces <= 0 || n_meshes <= 0 || (prec != MBA_F32 && prec != MBA_F64)) return 0;
    for (int m = 0; m < n_meshes; ++m)
        if (meshes[m].n_faces <= 0 || meshes[m].ndir <= 0) return 0;
    return rhs_batched_layout(prec, total_faces, meshes, n_meshes).total;
}

int mba_mom_rhs_ns_batched(int prec, const void* verts, const int32_t* faces, int64_t total_faces,
                           const double* rule, int32_t nq, const mba_ns_rhs_mesh* meshes, int32_t n_meshes,
                           void* workspace, size_t workspace_bytes, mba_stream_t stream)
{
    if (!verts || !faces || !rule || !meshes || !workspace) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    if (n_meshes <= 0 || total_faces <= 0 || total_faces > kMaxFaces) return MBA_ERR_SIZE;
    if (nq < 1 || nq > 64) return MBA_ERR_SIZE;
    int64_t sum = 0, items_f = 0, items_e = 0;
    for (int m = 0; m < n_meshes; ++m) {
        const mba_ns_rhs_mesh& d = meshes[m];
        if (d.n_faces <= 0 || d.face_off != sum || d.ndir <= 0 || d.n_edges <= 0) return MBA_ERR_SIZE;
        if (!d.dirs || !d.pols || !d.rwg_fe || !d.v) return MBA_ERR_ARG;
        sum += d.n_faces;
        items_f += d.n_faces * d.ndir;
        items_e += d.n_edges * d.ndir;
    }
    if (sum != total_faces) return MBA_ERR_ARG;
    if ((items_f + 255) / 256 > 0x7fffffff || (items_e + 255) / 256 > 0x7fffffff) return MBA_ERR_SIZE;
    if (workspace_bytes < mba_mom_rhs_ns_batched_workspace_bytes(prec, total_faces, meshes, n_meshes))
        return MBA_ERR_ALLOC;
    return mba_guard([&] {
        if (prec == MBA_F32)
            return rhs_ns_batched<float>(verts, faces, total_faces, rule, nq, meshes, n_meshes, workspace, S(stream));
        return rhs_ns_batched<double>(verts, faces, total_faces, rule, nq, meshes, n_meshes, workspace, S(stream));
    });
}

size_t mba_mom_gather_rwg_batched_scratch_bytes(int32_t n_jobs)
{
    if (n_jobs <= 0) return 0;
    size_t o = 0;
    return gather_batched_layout(n_jobs, &o);
}

int mba_mom_gather_rwg_batched(int prec, const mba_gather_job* jobs, int32_t n_jobs, void* scratch,
                               size_t scratch_bytes, mba_stream_t stream)
{
    if (!jobs || !scratch || n_jobs <= 0) return MBA_ERR_ARG;
    if (prec != MBA_F32 && prec != MBA_F64) return MBA_ERR_ARG;
    for (int q = 0; q < n_jobs; ++q) {
        const mba_gather_job& J = jobs[q];
        if (!J.zf || !J.rwg_fe || !J.z) return MBA_ERR_ARG;
        if (J.n_edges <= 0 || J.n_edges > ((int64_t)1 << 31) - 1) return MBA_ERR_SIZE;
        if (J.ldz < J.n_edges || J.ldzf < 1) return MBA_ERR_ARG;
    }
    if (scratch_bytes < mba_mom_gather_rwg_batched_scratch_bytes(n_jobs)) return MBA_ERR_ALLOC;
    return mba_guard([&]() -> int {
        std::vector<GatherJobDev> dj(n_jobs);
        std::vector<int64_t> toff(n_jobs + 1, 0);
        for (int q = 0; q < n_jobs; ++q) {
            const mba_gather_job& J = jobs[q];
            const int rows = gather_rows(J.n_edges);
            dj[q] = GatherJobDev{J.zf, J.ldzf, J.rwg_fe, J.n_edges, J.z, J.ldz, rows};
            toff[q + 1] = toff[q] + ((J.n_edges + 255) / 256) * ((J.n_edges + rows - 1) / rows);
        }
        const int64_t grid = (toff[n_jobs] + GATHER_XCD - 1) / GATHER_XCD * GATHER_XCD;
        if (grid > 0x7fffffff) return MBA_ERR_SIZE;
        size_t o_t = 0;
        gather_batched_layout(n_jobs, &o_t);
        char* base = (char*)scratch;
        hipStream_t st = S(stream);
        if (mba_host::upload_tables(base, dj.data(), dj.size() * sizeof(GatherJobDev), st) != 0 ||
            mba_host::upload_tables(base + o_t, toff.data(), toff.size() * sizeof(int64_t), st) != 0)
            return MBA_ERR_LAUNCH;
        if (prec == MBA_F32)
            hipLaunchKernelGGL(k_gather_rwg_batched<float>, dim3((unsigned)grid), dim3(256), 0, st,
                               (const GatherJobDev*)base, n_jobs, (const int64_t*)(base + o_t));
        else
            hipLaunchKernelGGL(k_gather_rwg_batched<double>, dim3((unsigned)grid), dim3(256), 0, st,
                               (const GatherJobDev*)base, n_jobs, (const int64_t*)(base + o_t));
        MBA_LAUNCH_CHECK();
        return MBA_OK;
    });
}

}  // extern "C"
