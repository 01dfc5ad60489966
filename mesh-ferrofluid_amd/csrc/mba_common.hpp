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
// revolutions) work on [0, 1).  double: the library sincos on kR.
__device__ __forceinline__ void cis_rev(float t, float& c, float& s)
{
    float f = __builtin_amdgcn_fractf(t);
    c = __builtin_amdgcn_cosf(f);
    s = __builtin_amdgcn_sinf(f);
}
__device__ __forceinline__ void cis_rev(double t, double& c, double& s)
{
    // t is in revolutions here as well; reduce exactly first (t - rint(t) is
    // exact in binary floating point), then sincos of 2 pi f with |f| <= 1/2.
    double f = t - rint(t);
    sincos(2.0 * MBA_PI * f, &s, &c);
}

__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_t(double x) { return 1.0 / sqrt(x); }

template <class T>
struct vec3 {
    T x, y, z;
};

template <class T>
__device__ __forceinline__ T dot(const vec3<T>& a, const vec3<T>& b)
{
    return a.x * b.x + a.y * b.y + a.z * b.z;
}
