// Multi-vector kernels of the GPU GMRES behind BEMRetIter
// (mnpbem/bem/bem_ret_iter.py:455-1877, mnpbem/bem/bem_iter.py:45-101).
//
// A "multi-vector" holds the 8N packed vectors of batch * P independent
// systems: for wavelength b it is L = n * 8P complex128 values (the (n, 8P)
// row-major block of the solver), and system (b, p) owns the elements whose
// index e satisfies e % P == p (columns c * P + p of every row).  The Arnoldi
// process of all systems runs in lockstep; these kernels compute per-system
// inner products, linear combinations and scalings of such vectors.
#include "common.h"

namespace {

constexpr int DOT_SPLIT = 64;  // partial sums per (batch, vector, system); fixed order -> deterministic
constexpr int MAXP = 16;

// partial dots: part[((b * k + j) * DOT_SPLIT + s) * P + p] = sum over chunk s of conj(V_j) W
__global__ __launch_bounds__(256) void k_mv_dot(long L, int P, const double* V, long sVk, long sVb,
                                                const double* W, long sWb, double* part) {
  const int s = blockIdx.x, j = blockIdx.y, b = blockIdx.z;
  const int T = 256 - 256 % P;  // thread stride, a multiple of P -> fixed system per thread
  const long per = ((L + DOT_SPLIT - 1) / DOT_SPLIT + T - 1) / T * T;
  const long e0 = (long)s * per;
  const long e1 = e0 + per < L ? e0 + per : L;
  const double* v = V + 2 * ((long)j * sVk + (long)b * sVb);
  const double* w = W + 2 * ((long)b * sWb);
  double ar = 0.0, ai = 0.0;
  const int tid = threadIdx.x;
  if (tid < T) {
    for (long e = e0 + tid; e < e1; e += T) {
      const double2 x = *reinterpret_cast<const double2*>(v + 2 * e);
      const double2 y = *reinterpret_cast<const double2*>(w + 2 * e);
      ar += x.x * y.x + x.y * y.y;
      ai += x.x * y.y - x.y * y.x;
    }
  }
  __shared__ double red[2][256];
  red[0][tid] = ar;
  red[1][tid] = ai;
  __syncthreads();
  if (tid < P) {
    double sr = 0.0, si = 0.0;
    for (int t = tid; t < T; t += P) {
      sr += red[0][t];
      si += red[1][t];
    }
    double* o = part + 2 * ((((long)b * gridDim.y + j) * DOT_SPLIT + s) * P + tid);
    o[0] = sr;
    o[1] = si;
  }
}

__global__ void k_mv_dot_reduce(int total, int P, const double* part, double* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // t = (b * k + j) * P + p
  if (t >= total) return;
  const int p = t % P;
  const long bj = t / P;
  double sr = 0.0, si = 0.0;
  for (int s = 0; s < DOT_SPLIT; ++s) {
    const double* q = part + 2 * ((bj * DOT_SPLIT + s) * P + p);
    sr += q[0];
    si += q[1];
  }
  out[2 * t] = sr;
  out[2 * t + 1] = si;
}

// W += sign * sum_j coef[b, j, p] V_j
__global__ __launch_bounds__(256) void k_mv_update(long L, int P, int k, const double* V, long sVk, long sVb,
                                                   const double* coef, double sign, double* W, long sWb) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= L) return;
  const int p = (int)(e % P);
  const double* c = coef + 2 * ((long)b * k * P + p);
  const double* v = V + 2 * ((long)b * sVb + e);
  double ar = 0.0, ai = 0.0;
  for (int j = 0; j < k; ++j) {
    const double cr = c[2 * (long)j * P], ci = c[2 * (long)j * P + 1];
    const double2 x = *reinterpret_cast<const double2*>(v + 2 * (long)j * sVk);
    ar += cr * x.x - ci * x.y;
    ai += cr * x.y + ci * x.x;
  }
  double* w = W + 2 * ((long)b * sWb + e);
  w[0] += sign * ar;
  w[1] += sign * ai;
}

// mode 0: out = W * s[b, p] (complex);  mode 1: out = W / sqrt(Re s[b, p]), 0 where Re s <= 0
__global__ __launch_bounds__(256) void k_mv_scale(long L, int P, int mode, const double* s, const double* W,
                                                  long sWb, double* out, long sOb) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (e >= L) return;
  const int p = (int)(e % P);
  double fr, fi;
  if (mode == 0) {
    fr = s[2 * ((long)b * P + p)];
    fi = s[2 * ((long)b * P + p) + 1];
  } else {
    const double q = s[2 * ((long)b * P + p)];
    fr = q > 0.0 ? 1.0 / sqrt(q) : 0.0;
    fi = 0.0;
  }
  const double2 x = *reinterpret_cast<const double2*>(W + 2 * ((long)b * sWb + e));
  double2 y;
  y.x = fr * x.x - fi * x.y;
  y.y = fr * x.y + fi * x.x;
  *reinterpret_cast<double2*>(out + 2 * ((long)b * sOb + e)) = y;
}

__global__ void k_widen(long count, const float* src, double* dst) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < count) dst[e] = (double)src[e];
}

}  // namespace

// out[b, j, p] = sum_{e % P == p} conj(V_j[b, e]) W[b, e] for j < k (complex128).
// V_j starts sVk complex elements after V_{j-1}; wavelength b of V_j (of W)
// starts sVb (sWb) elements after b-1.  work: >= batch * k * 64 * P complex128.
extern "C" int mnpbem_mv_dot_batched(long L, int P, int k, const void* V, long sVk, long sVb, const void* W,
                                     long sWb, void* out, void* work, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(L > 0 && P > 0 && P <= MAXP && L % P == 0, "mv_dot: bad size");
  MNP_REQUIRE(k >= 0 && k <= 65535 && batch <= 65535, "mv_dot: too many vectors");
  if (k == 0 || batch <= 0) return 0;
  hipLaunchKernelGGL(k_mv_dot, dim3(DOT_SPLIT, k, batch), dim3(256), 0, st, L, P, (const double*)V, sVk, sVb,
                     (const double*)W, sWb, (double*)work);
  MNP_LAUNCH_CHECK();
  const int total = batch * k * P;
  hipLaunchKernelGGL(k_mv_dot_reduce, dim3(mnp_ceil_div(total, 256)), dim3(256), 0, st, total, P,
                     (const double*)work, (double*)out);
  MNP_LAUNCH_CHECK();
  return 0;
}

// W[b] += sign * sum_{j<k} coef[b, j, p] V_j[b] (coef: (batch, k, P) complex128, contiguous)
extern "C" int mnpbem_mv_update_batched(long L, int P, int k, const void* V, long sVk, long sVb, const void* coef,
                                        double sign, void* W, long sWb, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(L > 0 && P > 0 && P <= MAXP && L % P == 0, "mv_update: bad size");
  if (k <= 0 || batch <= 0) return 0;
  hipLaunchKernelGGL(k_mv_update, dim3(mnp_ceil_div(L, 256), batch), dim3(256), 0, st, L, P, k,
                     (const double*)V, sVk, sVb, (const double*)coef, sign, (double*)W, sWb);
  MNP_LAUNCH_CHECK();
  return 0;
}

// out[b] = W[b] * s[b, p] (mode 0, s complex128 (batch, P)) or W[b] / sqrt(Re s[b, p])
// (mode 1; 0 where Re s <= 0).  out may alias W.
extern "C" int mnpbem_mv_scale_batched(long L, int P, int mode, const void* s, const void* W, long sWb, void* out,
                                       long sOb, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(L > 0 && P > 0 && P <= MAXP && L % P == 0 && (mode == 0 || mode == 1), "mv_scale: bad argument");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_mv_scale, dim3(mnp_ceil_div(L, 256), batch), dim3(256), 0, st, L, P, mode,
                     (const double*)s, (const double*)W, sWb, (double*)out, sOb);
  MNP_LAUNCH_CHECK();
  return 0;
}

// dst = (double) src for count real values (src complex64 or complex128)
extern "C" int mnpbem_widen(int src_dtype, long count, const void* src, double* dst, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (count <= 0) return 0;
  if (mnp_is_c64(src_dtype)) {
    hipLaunchKernelGGL(k_widen, dim3(mnp_ceil_div(count, 256)), dim3(256), 0, st, count, (const float*)src, dst);
    MNP_LAUNCH_CHECK();
    return 0;
  }
  MNP_REQUIRE(src_dtype == MNP_DTYPE_C128, "widen: unknown dtype");
  MNP_HIP_CHECK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToDevice, st));
  return 0;
}
