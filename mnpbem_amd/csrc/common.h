// Shared device/host helpers for the MNPBEM retarded-BEM HIP library (gfx950).
//
// Complex numbers are stored interleaved (re, im), the layout of numpy /
// torch complex128 and complex64, so device buffers are torch tensors handed
// over as raw pointers.  Matrix buffers are row-major with a leading
// dimension counted in complex elements.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

#define MNP_DTYPE_C64 1   // complex64  (float pairs)
#define MNP_DTYPE_C128 2  // complex128 (double pairs)
#define MNP_DTYPE_C64X 3  // complex64 data, GEMMs on bf16 matrix cores with 3-term splitting

inline bool mnp_is_c64(int dtype) { return dtype == MNP_DTYPE_C64 || dtype == MNP_DTYPE_C64X; }

// ---------------------------------------------------------------------------
// error reporting across the C ABI: status code + thread-local message
// ---------------------------------------------------------------------------
void mnp_set_error(const std::string& msg);

#define MNP_HIP_CHECK(expr)                                                       \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) {                                                       \
      mnp_set_error(std::string(#expr) + ": " + hipGetErrorString(_e) + " (" +    \
                    __FILE__ + ":" + std::to_string(__LINE__) + ")");             \
      return 2;                                                                   \
    }                                                                             \
  } while (0)

#define MNP_REQUIRE(cond, msg)          \
  do {                                  \
    if (!(cond)) {                      \
      mnp_set_error(std::string(msg));  \
      return 1;                         \
    }                                   \
  } while (0)

#define MNP_LAUNCH_CHECK() MNP_HIP_CHECK(hipGetLastError())

// Geometry slot of batch element b.  Batched calls over several particles of
// one face count stack the per-face geometry of `slots` particles (pos, nvec:
// slots x n x 3; area, material and side flags: slots x n) and order the batch
// particle-major, gdiv wavelengths per particle: element b reads slot b / gdiv.
// gdiv <= 0: every element reads slot 0 (one particle for the whole batch).
__host__ __device__ __forceinline__ long mnp_geo_slot(long b, int gdiv) { return gdiv > 0 ? b / gdiv : 0; }

// ---------------------------------------------------------------------------
// complex arithmetic (T = float or double)
// ---------------------------------------------------------------------------
template <typename T>
struct cplx {
  T re, im;
};

template <typename T>
struct vec2_of;
template <>
struct vec2_of<double> {
  typedef double2 type;
};
template <>
struct vec2_of<float> {
  typedef float2 type;
};

template <typename T>
__device__ __forceinline__ cplx<T> cmake(T r, T i) {
  cplx<T> z;
  z.re = r;
  z.im = i;
  return z;
}
template <typename T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return cmake(a.re + b.re, a.im + b.im);
}
template <typename T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) {
  return cmake(a.re - b.re, a.im - b.im);
}
template <typename T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cmake(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
template <typename T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) {
  return cmake(a.re * s, a.im * s);
}
// fused multiply-add x*y + z, one rounding
__device__ __forceinline__ float fmadd(float x, float y, float z) { return __builtin_fmaf(x, y, z); }
__device__ __forceinline__ double fmadd(double x, double y, double z) { return __builtin_fma(x, y, z); }
// a - b*c as four fused multiply-adds per complex value (the accumulator
// enters the first one, so a chain of updates costs 4 instructions each)
template <typename T>
__device__ __forceinline__ cplx<T> cfms(cplx<T> a, cplx<T> b, cplx<T> c) {
  return cmake(fmadd(b.im, c.im, fmadd(-b.re, c.re, a.re)), fmadd(-b.im, c.re, fmadd(-b.re, c.im, a.im)));
}
// a + b*c
template <typename T>
__device__ __forceinline__ cplx<T> cfma(cplx<T> a, cplx<T> b, cplx<T> c) {
  return cmake(fmadd(-b.im, c.im, fmadd(b.re, c.re, a.re)), fmadd(b.im, c.re, fmadd(b.re, c.im, a.im)));
}
// 1 / a, scaled to avoid overflow (Smith's algorithm)
template <typename T>
__device__ __forceinline__ cplx<T> crecip(cplx<T> a) {
  T ar = a.re, ai = a.im;
  if (fabs(ar) >= fabs(ai)) {
    T r = ai / ar, d = ar + ai * r;
    return cmake(T(1) / d, -r / d);
  } else {
    T r = ar / ai, d = ai + ar * r;
    return cmake(r / d, T(-1) / d);
  }
}
template <typename T>
__device__ __forceinline__ cplx<T> cdiv(cplx<T> a, cplx<T> b) {
  return cmul(a, crecip(b));
}
template <typename T>
__device__ __forceinline__ T cabs1(cplx<T> a) {
  return fabs(a.re) + fabs(a.im);
}
template <typename T>
__device__ __forceinline__ cplx<T> cld(const T* p) {
  typedef typename vec2_of<T>::type V;
  V v = *reinterpret_cast<const V*>(p);
  return cmake(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ void cst(T* p, cplx<T> z) {
  typedef typename vec2_of<T>::type V;
  V v;
  v.x = z.re;
  v.y = z.im;
  *reinterpret_cast<V*>(p) = v;
}
// exp(i z) for complex z = zr + i zi  ->  e^{-zi} (cos zr + i sin zr)
template <typename T>
__device__ __forceinline__ cplx<T> cexpi(T zr, T zi) {
  T s, c;
  sincos(zr, &s, &c);
  T e = exp(-zi);
  return cmake(e * c, e * s);
}

inline int mnp_ceil_div(long a, long b) { return (int)((a + b - 1) / b); }
