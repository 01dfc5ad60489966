// Shared device helpers for the mesh-BEM HIP library (gfx950 / CDNA4).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mesh_bem_amd.h"

#define MBA_PI 3.14159265358979323846
#define MBA_INV4PI 0.07957747154594766788

#define MBA_LAUNCH_CHECK()                                              \
    do {                                                                \
        hipError_t _e = hipGetLastError();                              \
        if (_e != hipSuccess) return MBA_ERR_LAUNCH;                    \
    } while (0)

template <class T>
struct alignas(2 * sizeof(T)) cpx {
    T re, im;
};

template <class T>
__device__ __forceinline__ void store_cpx(cpx<T>* p, T re, T im)
{
    cpx<T> v;
    v.re = re;
    v.im = im;
    *p = v;
}

// e^{-jkR} = (c, -s) with c = cos(kR), s = sin(kR).
// float: the argument is the phase in revolutions t = kR / (2 pi); the
// fractional part is taken first so the hardware sin/cos (which take
// revolutions) work on [0, 1).  double: exact reduction to an octant and
// Taylor series (below).
__device__ __forceinline__ void cis_rev(float t, float& c, float& s)
{
    float f = __builtin_amdgcn_fractf(t);
    c = __builtin_amdgcn_cosf(f);
    s = __builtin_amdgcn_sinf(f);
}
__device__ __forceinline__ void cis_rev(double t, double& c, double& s)
{
    // t in revolutions.  q = rint(4t) and r = t - q/4 (|r| <= 1/8) are exact;
    // x = 2 pi r (|x| <= pi/4); sin x and cos x by their Taylor series through
    // x^17 and x^18 (the first omitted terms are below 1e-19), evaluated in
    // Horner form in x^2; the quadrant q mod 4 rotates (cos, sin).
    const double q = rint(4.0 * t);
    const double r = t - 0.25 * q;
    const double x = (2.0 * MBA_PI) * r;
    const double z = x * x;
    double sp = 1.0 / 355687428096000.0;
    sp = fma(sp, z, -1.0 / 1307674368000.0);
    sp = fma(sp, z, 1.0 / 6227020800.0);
    sp = fma(sp, z, -1.0 / 39916800.0);
    sp = fma(sp, z, 1.0 / 362880.0);
    sp = fma(sp, z, -1.0 / 5040.0);
    sp = fma(sp, z, 1.0 / 120.0);
    sp = fma(sp, z, -1.0 / 6.0);
    const double s0 = fma(x * z, sp, x);
    double cp = -1.0 / 6402373705728000.0;
    cp = fma(cp, z, 1.0 / 20922789888000.0);
    cp = fma(cp, z, -1.0 / 87178291200.0);
    cp = fma(cp, z, 1.0 / 479001600.0);
    cp = fma(cp, z, -1.0 / 3628800.0);
    cp = fma(cp, z, 1.0 / 40320.0);
    cp = fma(cp, z, -1.0 / 720.0);
    cp = fma(cp, z, 1.0 / 24.0);
    cp = fma(cp, z, -0.5);
    const double c0 = fma(z, cp, 1.0);
    const int n = (int)((long long)q & 3);
    c = (n == 0) ? c0 : (n == 1) ? -s0 : (n == 2) ? -c0 : s0;
    s = (n == 0) ? s0 : (n == 1) ? c0 : (n == 2) ? -s0 : -c0;
}

__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
// FP64 1/sqrt(x), x > 0: the hardware estimate refined by two Newton steps
// y <- y (1 + (1 - x y^2) / 2) (each step squares the relative error).
__device__ __forceinline__ double rsqrt_t(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-x * y, y, 1.0);
    y = fma(0.5 * y, e, y);
    e = fma(-x * y, y, 1.0);
    return fma(0.5 * y, e, y);
}

template <class T>
struct vec3 {
    T x, y, z;
};

template <class T>
__device__ __forceinline__ T dot(const vec3<T>& a, const vec3<T>& b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
