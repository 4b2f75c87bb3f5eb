// Per-wavelength BEM assembly of Sigma/Delta, the small-right-hand-side solve
// of GdA&H Eqs. (19)-(20), the 8N operator residual for iterative refinement,
// plane-wave excitation and the optical-theorem extinction (gfx950, batched
// over wavelengths).
//
// Storage convention.  Every N x N matrix M of the algorithm is kept as the
// row-major buffer of its transpose, buf[j*ld + i] = M[i][j]:
//   G1t, G2t  -> LU factors of G1^T, G2^T (P G^T = L U)
//   S1t       -> Sigma1^T = (G1^T)^-1 H1^T,  S2t likewise
//   Dt        -> LU of Delta^T = Sigma1^T - Sigma2^T
//   Et        -> (Delta^-1)^T
//   St        -> LU of Sigma^T = eps1 Sigma1^T - eps2 Sigma2^T + k^2 (eps1-eps2)^2 Et o (n n^T)
// (mnpbem/bem/bem_ret.py:437-523, MATLAB BEM/@bemret/private/initmat.m).
// Products M x are then "transposed" GEMVs over the buffer, and solves
// M x = v with the stored LU of M^T are transposed triangular solves.
//
// Vector layout per wavelength: (N, 8P) row-major, P = polarisations:
//   unknowns  X = [sig1 (P) | h1 (3P, index c*P+p) | sig2 (P) | h2 (3P)]
//   right side R = [phi (P) | a (3P) | De (P) | alpha (3P)]
// the packing order of mnpbem/bem/bem_ret_iter.py:1196-1232.
#include "common.h"

#include <algorithm>
#include <mutex>
#include <string>

namespace {

constexpr int MAXC = 16;  // right-hand-side columns handled per pass

// Y[b] = beta*Y[b] + alpha * M[b]^T X[b]  (M n x n buffer, X, Y n x C, C <= CM).
// A workgroup takes 64 output rows i (lane = row: every load of M is a
// coalesced 512-byte row segment) and its four waves take four quarters of
// the buffer rows j, partial sums added in wave order (deterministic).  The
// rows of X are read with wave-uniform loads, not staged through LDS: with
// LDS staging the workgroup barriers every 64 rows broke the stream of
// loads (2.9 TB/s at C = 8 over the complex64 matrices of a 2884-row batch
// of 64, against 4.7 TB/s in this form); 8 row loads in flight per lane.
template <typename T, int CM>
__global__ __launch_bounds__(256) void k_gemv_t(int n, int C, const T* __restrict__ M, long ldm, long sM,
                                                const T* __restrict__ X, long ldx, long sX, T* __restrict__ Y,
                                                long ldy, long sY, double alr, double ali, double beta) {
  constexpr int UNR = 8;
  const long b = blockIdx.y;
  M += 2 * b * sM;
  X += 2 * b * sX;
  Y += 2 * b * sY;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int i = blockIdx.x * 64 + lane;
  const int q = (((n + 3) / 4) + UNR - 1) / UNR * UNR;  // buffer rows per wave
  const int jb = wv * q, je = min(n, jb + q);
  const int ic = min(i, n - 1);
  __shared__ cplx<T> part[3][64][CM];
  cplx<T> acc[CM];
#pragma unroll
  for (int c = 0; c < CM; ++c) acc[c] = cmake<T>(0, 0);
  auto xval = [&](int j, int c) {
    const T* xr = X + 2 * ((long)j * ldx + c);
    return c < C ? cmake<T>(xr[0], xr[1]) : cmake<T>(0, 0);
  };
  int j = jb;
  for (; j + UNR <= je; j += UNR) {
    cplx<T> m[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) m[u] = cld(M + 2 * ((long)(j + u) * ldm + ic));
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
      for (int c = 0; c < CM; ++c) acc[c] = cfma(acc[c], m[u], xval(j + u, c));
  }
  for (; j < je; ++j) {
    const cplx<T> m = cld(M + 2 * ((long)j * ldm + ic));
#pragma unroll
    for (int c = 0; c < CM; ++c) acc[c] = cfma(acc[c], m, xval(j, c));
  }
  if (wv > 0)
#pragma unroll
    for (int c = 0; c < CM; ++c) part[wv - 1][lane][c] = acc[c];
  __syncthreads();
  if (wv > 0 || i >= n) return;
#pragma unroll
  for (int w = 0; w < 3; ++w)
#pragma unroll
    for (int c = 0; c < CM; ++c) acc[c] = cadd(acc[c], part[w][lane][c]);
#pragma unroll
  for (int c = 0; c < CM; ++c)
    if (c < C) {
      T* py = Y + 2 * ((long)i * ldy + c);
      const double sr = (double)acc[c].re, si = (double)acc[c].im;
      double yr = alr * sr - ali * si, yi = alr * si + ali * sr;
      if (beta != 0.0) {
        const cplx<T> y0 = cld(py);
        yr += beta * (double)y0.re;
        yi += beta * (double)y0.im;
      }
      cst(py, cmake<T>((T)yr, (T)yi));
    }
}

// complex64 form of k_gemv_t with two adjacent output rows i, i + 1 per lane:
// one 16-byte load per buffer row (a wave reads a 1 KB segment instead of
// 512 bytes, half the load instructions per byte).  Every output is the same
// sum in the same order as in k_gemv_t (bit-identical).  Needs n, ldm and sM
// even and M 16-byte aligned.
template <int CM>
__global__ __launch_bounds__(256) void k_gemv_t2(int n, int C, const float* __restrict__ M, long ldm, long sM,
                                                 const float* __restrict__ X, long ldx, long sX, float* __restrict__ Y,
                                                 long ldy, long sY, double alr, double ali, double beta) {
  constexpr int UNR = 8;
  const long b = blockIdx.y;
  M += 2 * b * sM;
  X += 2 * b * sX;
  Y += 2 * b * sY;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int i = blockIdx.x * 128 + 2 * lane;
  const int q = (((n + 3) / 4) + UNR - 1) / UNR * UNR;  // buffer rows per wave
  const int jb = wv * q, je = min(n, jb + q);
  const int ic = min(i, n - 2);
  __shared__ cplx<float> part[3][64][2 * CM];
  cplx<float> a0[CM], a1[CM];
#pragma unroll
  for (int c = 0; c < CM; ++c) a0[c] = a1[c] = cmake<float>(0, 0);
  auto xval = [&](int j, int c) {
    const float* xr = X + 2 * ((long)j * ldx + c);
    return c < C ? cmake<float>(xr[0], xr[1]) : cmake<float>(0, 0);
  };
  int j = jb;
  for (; j + UNR <= je; j += UNR) {
    float4 m[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) m[u] = *reinterpret_cast<const float4*>(M + 2 * ((long)(j + u) * ldm + ic));
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
      for (int c = 0; c < CM; ++c) {
        const cplx<float> x = xval(j + u, c);
        a0[c] = cfma(a0[c], cmake<float>(m[u].x, m[u].y), x);
        a1[c] = cfma(a1[c], cmake<float>(m[u].z, m[u].w), x);
      }
  }
  for (; j < je; ++j) {
    const float4 m = *reinterpret_cast<const float4*>(M + 2 * ((long)j * ldm + ic));
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      const cplx<float> x = xval(j, c);
      a0[c] = cfma(a0[c], cmake<float>(m.x, m.y), x);
      a1[c] = cfma(a1[c], cmake<float>(m.z, m.w), x);
    }
  }
  if (wv > 0)
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      part[wv - 1][lane][c] = a0[c];
      part[wv - 1][lane][CM + c] = a1[c];
    }
  __syncthreads();
  if (wv > 0 || i >= n) return;
#pragma unroll
  for (int w = 0; w < 3; ++w)
#pragma unroll
    for (int c = 0; c < CM; ++c) {
      a0[c] = cadd(a0[c], part[w][lane][c]);
      a1[c] = cadd(a1[c], part[w][lane][CM + c]);
    }
#pragma unroll
  for (int h = 0; h < 2; ++h)
#pragma unroll
    for (int c = 0; c < CM; ++c)
      if (c < C) {
        const cplx<float> a = h ? a1[c] : a0[c];
        float* py = Y + 2 * ((long)(i + h) * ldy + c);
        const double sr = (double)a.re, si = (double)a.im;
        double yr = alr * sr - ali * si, yi = alr * si + ali * sr;
        if (beta != 0.0) {
          const cplx<float> y0 = cld(py);
          yr += beta * (double)y0.re;
          yi += beta * (double)y0.im;
        }
        cst(py, cmake<float>((float)yr, (float)yi));
      }
}

// complex64 form of k_gemv_t for C <= 8 columns on the FP32 matrix cores
// (v_mfma_f32_16x16x4f32): the VALU form spends four FMAs per complex entry
// and column, about as many cycles as HBM needs to deliver the entry, so the
// two did not overlap fully.  Real form of the complex product with K over
// the buffer rows j: for a pair of output rows (2p, 2p + 1) a lane supplies
// A[p][q] = re / im of M(j + q, row) (one 16-byte load: both rows, both
// parts), and B[q][c] = [Xr | Xi] (columns 0-7 | 8-15) against the real
// parts and [-Xi | Xr] against the imaginary parts, so that D[p][c] holds
// Re(M X) for c < 8 and Im(M X) for c >= 8.  One accumulator per row parity.
// A workgroup of 4 waves owns 32 output rows; the waves split the buffer
// rows in quarters (multiples of 32) and their partial sums are added in
// wave order.  FP32 arithmetic like the VALU form, in another summation
// order.  Needs n, ldm and sM even and M 16-byte aligned.
typedef float gv_f32x4 __attribute__((ext_vector_type(4)));
int g_gemv_mfma = 1;  // complex64 GEMVs of <= 8 columns on the matrix cores
int g_trsv_mfma = 3;  // transposed-solve tile updates on the matrix cores: bit 0 for 5-8 columns, bit 1 for <= 4

__global__ __launch_bounds__(256) void k_gemv_t_mfma(int n, int C, const float* __restrict__ M, long ldm, long sM,
                                                     const float* __restrict__ X, long ldx, long sX,
                                                     float* __restrict__ Y, long ldy, long sY, double alr, double ali,
                                                     double beta) {
  constexpr int UNR = 8;  // 4-row groups of the buffer in flight per wave
  const long b = blockIdx.y;
  M += 2 * b * sM;
  X += 2 * b * sX;
  Y += 2 * b * sY;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int p = lane & 15, q = lane >> 4;
  const int r0 = blockIdx.x * 32;
  const int ir = min(r0 + 2 * p, n - 2);  // rows past n read row n - 2 (discarded)
  const int qs = ((n + 3) / 4 + 4 * UNR - 1) / (4 * UNR) * (4 * UNR);
  const int jb = wv * qs, je = min(n, jb + qs);
  const int c = p & 7;
  const bool cim = p >= 8, cok = c < C;
  gv_f32x4 de = {0.f, 0.f, 0.f, 0.f}, dodd = {0.f, 0.f, 0.f, 0.f};
  int j = jb;
  for (; j + 4 * UNR <= je; j += 4 * UNR) {
    float4 m[UNR];
    float2 x[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) m[u] = *reinterpret_cast<const float4*>(M + 2 * ((long)(j + 4 * u + q) * ldm + ir));
#pragma unroll
    for (int u = 0; u < UNR; ++u)
      x[u] = cok ? *reinterpret_cast<const float2*>(X + 2 * ((long)(j + 4 * u + q) * ldx + c)) : make_float2(0.f, 0.f);
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      const float b1 = cim ? x[u].y : x[u].x, b2 = cim ? x[u].x : -x[u].y;
      de = __builtin_amdgcn_mfma_f32_16x16x4f32(m[u].x, b1, de, 0, 0, 0);
      dodd = __builtin_amdgcn_mfma_f32_16x16x4f32(m[u].z, b1, dodd, 0, 0, 0);
      de = __builtin_amdgcn_mfma_f32_16x16x4f32(m[u].y, b2, de, 0, 0, 0);
      dodd = __builtin_amdgcn_mfma_f32_16x16x4f32(m[u].w, b2, dodd, 0, 0, 0);
    }
  }
  for (; j < je; j += 4) {
    const int jj = j + q;
    float4 m = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 x = make_float2(0.f, 0.f);
    if (jj < je) {
      m = *reinterpret_cast<const float4*>(M + 2 * ((long)jj * ldm + ir));
      if (cok) x = *reinterpret_cast<const float2*>(X + 2 * ((long)jj * ldx + c));
    }
    const float b1 = cim ? x.y : x.x, b2 = cim ? x.x : -x.y;
    de = __builtin_amdgcn_mfma_f32_16x16x4f32(m.x, b1, de, 0, 0, 0);
    dodd = __builtin_amdgcn_mfma_f32_16x16x4f32(m.z, b1, dodd, 0, 0, 0);
    de = __builtin_amdgcn_mfma_f32_16x16x4f32(m.y, b2, de, 0, 0, 0);
    dodd = __builtin_amdgcn_mfma_f32_16x16x4f32(m.w, b2, dodd, 0, 0, 0);
  }
  // lane holds D[4 q + v][p] for v = 0..3; parts [wave][parity][lane][v]
  __shared__ float part[4][2][64][4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    part[wv][0][lane][v] = de[v];
    part[wv][1][lane][v] = dodd[v];
  }
  __syncthreads();
  // one thread per (output row rho of the 32, column cc of the 8)
  const int rho = tid >> 3, cc = tid & 7;
  const int row = r0 + rho;
  if (row >= n || cc >= C) return;
  const int par = rho & 1, i = rho >> 1;          // D row i of the parity's accumulator
  const int lr = (i >> 2) * 16 + cc, li = lr + 8;  // lanes holding columns cc (re) and cc + 8 (im)
  const int v = i & 3;
  float sr = 0.f, si = 0.f;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    sr += part[w][par][lr][v];
    si += part[w][par][li][v];
  }
  float* py = Y + 2 * ((long)row * ldy + cc);
  double yr = alr * (double)sr - ali * (double)si, yi = alr * (double)si + ali * (double)sr;
  if (beta != 0.0) {
    const cplx<float> y0 = cld(py);
    yr += beta * (double)y0.re;
    yi += beta * (double)y0.im;
  }
  cst(py, cmake<float>((float)yr, (float)yi));
}

// The four products of the complex128 residual, Y = [G1^T.. buffers applied:
// G1 X1 | H1 X1 | G2 X2 | H2 X2], in one launch (blockIdx.z = product), with
// accumulators for exactly the CC = 4P columns.  One product per launch with
// accumulators for 16 columns kept three waves per SIMD in flight and read
// the matrices at 4.2 TB/s.  Same loop and summation order as k_gemv_t.
template <int CC>
__global__ __launch_bounds__(256) void k_gemv4_t(int n, const double* __restrict__ M0, const double* __restrict__ M1,
                                                 const double* __restrict__ M2, const double* __restrict__ M3,
                                                 long ldm, long sM, const double* __restrict__ X, long ldx, long sX,
                                                 double* __restrict__ Y, long ldy, long sY) {
  const int q = blockIdx.z;
  const double* M = q == 0 ? M0 : q == 1 ? M1 : q == 2 ? M2 : M3;
  const long b = blockIdx.y;
  M += 2 * b * sM;
  const double* Xq = X + 2 * (b * sX + (q >= 2 ? CC : 0));
  double* Yq = Y + 2 * (b * sY + (long)q * CC);
  const int tid = threadIdx.x;
  const int i = blockIdx.x * 256 + tid;
  __shared__ cplx<double> xs[64][CC];
  cplx<double> acc[CC];
#pragma unroll
  for (int c = 0; c < CC; ++c) acc[c] = cmake<double>(0, 0);
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int nj = min(64, n - j0);
    __syncthreads();
    for (int e = tid; e < nj * CC; e += 256) {
      const int jj = e / CC, c = e % CC;
      xs[jj][c] = cld(Xq + 2 * ((long)(j0 + jj) * ldx + c));
    }
    __syncthreads();
    if (i < n) {
      int jj = 0;
      for (; jj + 16 <= nj; jj += 16) {
        cplx<double> m[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) m[u] = cld(M + 2 * ((long)(j0 + jj + u) * ldm + i));
        // keep the 16 row loads issued ahead of the arithmetic (left to itself
        // the compiler interleaves them with the FMAs, fewer bytes in flight)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 16; ++u)
#pragma unroll
          for (int c = 0; c < CC; ++c) acc[c] = cfma(acc[c], m[u], xs[jj + u][c]);
      }
      for (; jj < nj; ++jj) {
        const cplx<double> m = cld(M + 2 * ((long)(j0 + jj) * ldm + i));
#pragma unroll
        for (int c = 0; c < CC; ++c) acc[c] = cfma(acc[c], m, xs[jj][c]);
      }
    }
  }
  if (i < n) {
#pragma unroll
    for (int c = 0; c < CC; ++c) cst(Yq + 2 * ((long)i * ldy + c), acc[c]);
  }
}

// One 64-row block step of the solve M x = v with the LU of M^T in the
// buffer (P M^T = L U, so x = P^T L^-T U^-T v).
//   FWD (U^T z = v, U^T lower non-unit): the block rhs is in V (already
//     reduced by earlier blocks); z_I -> Z; V rows below the block -= U_IK^T z_I.
//   BWD (L^T w = z, unit upper): the block rhs is in Z; w_I -> V; Z rows above
//     the block -= L_IK^T w_I.
// Every workgroup solves the 64 x 64 triangle (one wavefront, shuffles) and
// updates its own 256-row tile; workgroup 0 stores the block solution.
// Systems b >= half take their vectors hoff columns to the right of system
// b - half's (two factorisations stacked in one batch solving different
// column ranges of the same vectors).
template <typename T, bool FWD>
__global__ __launch_bounds__(256) void k_trsv_t_step(int n, int C, int i0, const T* __restrict__ LU, long ld,
                                                     long sLU, T* V, long ldv, long sV, T* Z, long ldz, long sZ,
                                                     const T* __restrict__ Dinv, long sD, int half, long hoff) {
  const long b = blockIdx.y;
  LU += 2 * b * sLU;
  V += 2 * ((b % half) * sV + (b / half) * hoff);
  Z += 2 * ((b % half) * sZ + (b / half) * hoff);
  const int tid = threadIdx.x, lane = tid & 63;
  const int nb = min(64, n - i0), i1 = i0 + nb;
  __shared__ cplx<T> blk[64][65];
  __shared__ cplx<T> sol[64][MAXC];
  if (Dinv) {
    // block solution = (inverse diagonal block)^T * rhs, all rows in parallel
    const T* src = FWD ? V : Z;
    const long lds = FWD ? ldv : ldz;
    const T* D = Dinv + 2 * (b * sD + ((long)(i0 / 64) * 2 + (FWD ? 0 : 1)) * 64 * 64);
    for (int e = tid; e < 64 * C; e += 256) {
      const int r = e / C, c = e % C;
      blk[r][c] = r < nb ? cld(src + 2 * ((long)(i0 + r) * lds + c)) : cmake<T>(0, 0);
    }
    __syncthreads();
    // thread (row r, column group cg): the 64 entries of column r of the
    // inverse block are loaded once and applied to the group's columns
    const int r = tid & 63, cg = tid >> 6;
    cplx<T> dcol[64];
#pragma unroll
    for (int q = 0; q < 64; ++q) dcol[q] = (q < nb && r < nb) ? cld(D + 2 * ((long)q * 64 + r)) : cmake<T>(0, 0);
    for (int c = cg; c < C; c += 4) {
      cplx<T> acc = cmake<T>(0, 0);
#pragma unroll
      for (int q = 0; q < 64; ++q) acc = cfma(acc, dcol[q], blk[q][c]);
      sol[r][c] = acc;
    }
    __syncthreads();
    if (blockIdx.x == 0) {
      T* dst = FWD ? Z : V;
      const long ldd_ = FWD ? ldz : ldv;
      for (int e = tid; e < nb * C; e += 256) {
        const int rr = e / C, c = e % C;
        cst(dst + 2 * ((long)(i0 + rr) * ldd_ + c), sol[rr][c]);
      }
    }
  } else {
  for (int e = tid; e < nb * nb; e += 256) {
    const int r = e / nb, q = e % nb;
    blk[r][q] = cld(LU + 2 * ((long)(i0 + r) * ld + i0 + q));
  }
  __syncthreads();
  if (tid < 64) {
    const T* src = FWD ? V : Z;
    const long lds = FWD ? ldv : ldz;
    cplx<T> val[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c)
      val[c] = (c < C && lane < nb) ? cld(src + 2 * ((long)(i0 + lane) * lds + c)) : cmake<T>(0, 0);
    if (FWD) {
      for (int i = 0; i < nb; ++i) {
        if (lane == i) {
          const cplx<T> inv = crecip(blk[i][i]);
#pragma unroll
          for (int c = 0; c < MAXC; ++c) val[c] = cmul(val[c], inv);
        }
        const cplx<T> u = blk[i][lane];  // U[i][r] = (U^T)[r][i]
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
          if (c < C) {
            cplx<T> zi;
            zi.re = __shfl(val[c].re, i);
            zi.im = __shfl(val[c].im, i);
            if (lane > i && lane < nb) val[c] = cfms(val[c], u, zi);
          }
      }
    } else {
      for (int i = nb - 1; i >= 0; --i) {
        const cplx<T> l = blk[i][lane];  // L[i][r] = (L^T)[r][i], r < i
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
          if (c < C) {
            cplx<T> wi;
            wi.re = __shfl(val[c].re, i);
            wi.im = __shfl(val[c].im, i);
            if (lane < i) val[c] = cfms(val[c], l, wi);
          }
      }
    }
#pragma unroll
    for (int c = 0; c < MAXC; ++c) sol[lane][c] = val[c];
    if (blockIdx.x == 0 && lane < nb) {
      T* dst = FWD ? Z : V;
      const long ldd_ = FWD ? ldz : ldv;
#pragma unroll
      for (int c = 0; c < MAXC; ++c)
        if (c < C) cst(dst + 2 * ((long)(i0 + lane) * ldd_ + c), val[c]);
    }
  }
  __syncthreads();
  }
  // tile update: FWD rows k >= i1, BWD rows k < i0
  const int k = (FWD ? i1 : 0) + blockIdx.x * 256 + tid;
  const int kend = FWD ? n : i0;
  if (k < kend) {
    T* upd = FWD ? V : Z;
    const long ldu = FWD ? ldv : ldz;
    cplx<T> acc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) acc[c] = cmake<T>(0, 0);
    int l = 0;
    for (; l + 16 <= nb; l += 16) {  // 16 row loads in flight
      cplx<T> m[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) m[u] = cld(LU + 2 * ((long)(i0 + l + u) * ld + k));
#pragma unroll
      for (int u = 0; u < 16; ++u)
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
          if (c < C) acc[c] = cfma(acc[c], m[u], sol[l + u][c]);
    }
    for (; l < nb; ++l) {
      const cplx<T> m = cld(LU + 2 * ((long)(i0 + l) * ld + k));
#pragma unroll
      for (int c = 0; c < MAXC; ++c)
        if (c < C) acc[c] = cfma(acc[c], m, sol[l][c]);
    }
#pragma unroll
    for (int c = 0; c < MAXC; ++c)
      if (c < C) {
        T* pu = upd + 2 * ((long)k * ldu + c);
        cst(pu, csub(cld(pu), acc[c]));
      }
  }
}

// Window step of the same solve with the inverse diagonal blocks: the
// window's rows [w0, w1) (up to TRSV_WIN rows, 64-row blocks) are solved
// block by block (block solution = inverse block^T * rhs, then the window's
// own off-diagonal block applied to the blocks still to solve, all in LDS);
// then each workgroup updates its tile of the rows outside the window (FWD
// rows >= w1, BWD rows < w0) with all the window's rows.  Fewer launches than
// one 64-row step each, for the same bytes of the triangle.
//   TILE 256: one row per thread, window rows summed in order.  The tiles are
//     laid out from the window outwards (FWD upwards from w1, BWD downwards
//     from w0), so workgroup 0 updates exactly the next window's rows; with
//     `ahead` it then solves that window once and writes its solution, and
//     the next launch (`have`) reads the solution instead of every workgroup
//     solving the window again (the redundant solves took ~40 % of the
//     time).  Only the first window is solved by every workgroup.
//   TILE 64: the window rows split over the four waves (four times the
//     workgroups, for few systems), wave partial sums added in wave order;
//     every workgroup solves its window.
// CM >= C: right-hand-side columns held per row.  Deterministic: the window
// solve runs the same operations in the same order whichever workgroup runs it.
constexpr int TRSV_WIN = 256;
constexpr int WSU = 32;  // loads in flight per thread in the window solve (16: 2-6 % slower solves)
template <typename T, bool FWD, int CM, int TILE, bool MF = false>
__global__ __launch_bounds__(256) void k_trsv_t_win(int n, int C, int w0, int w1, int have, int ahead,
                                                    const T* __restrict__ LU, long ld, long sLU, T* V, long ldv,
                                                    long sV, T* Z, long ldz, long sZ, const T* __restrict__ Dinv,
                                                    long sD, int half, long hoff) {
  // 256-row tiles: one-dimensional grid, workgroup 0 of every system first
  // (it solves the next window after its own tile, so starting it before the
  // other tiles lets that serial solve overlap their updates), then tiles
  // 1.. of system 0, of system 1, ...; 64-row tiles: (tile, system) grid
  long b;
  int bx;
  if constexpr (TILE == 256) {
    const int nt = max(1, FWD ? (n - w1 + 255) / 256 : (w0 + 255) / 256);
    const int nb_sys = gridDim.x / nt, L = blockIdx.x;
    if (L < nb_sys) {
      b = L;
      bx = 0;
    } else {
      b = (L - nb_sys) / (nt - 1);
      bx = 1 + (L - nb_sys) % (nt - 1);
    }
  } else {
    b = blockIdx.y;
    bx = blockIdx.x;
  }
  LU += 2 * b * sLU;
  Dinv += 2 * b * sD;
  V += 2 * ((b % half) * sV + (b / half) * hoff);
  Z += 2 * ((b % half) * sZ + (b / half) * hoff);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int W = w1 - w0;
  __shared__ cplx<T> rs[TRSV_WIN][CM];  // window right-hand side, then its solution
  __shared__ cplx<T> part[TILE == 64 ? 3 : 1][64][CM];
  T* dst = FWD ? Z : V;
  const long ldd_ = FWD ? ldz : ldv;
  // solve the window [s0, s1) whose right-hand side is in rs[0 .. s1 - s0)
  auto solve_window = [&](int s0, int s1) {
    const int nbk = (s1 - s0 + 63) / 64;
    for (int s = 0; s < nbk; ++s) {
      const int r = FWD ? s : nbk - 1 - s;
      const int i0 = s0 + 64 * r, nb = min(64, s1 - i0);
      const T* D = Dinv + 2 * (((long)(i0 / 64) * 2 + (FWD ? 0 : 1)) * 64 * 64);
      __syncthreads();  // rhs of block r complete
      // row lane of the block, columns wv, wv + 4, ...
      cplx<T> sl[CM / 4];
#pragma unroll
      for (int v = 0; v < CM / 4; ++v) sl[v] = cmake<T>(0, 0);
      for (int q0 = 0; q0 < nb; q0 += WSU) {
        cplx<T> d[WSU];
#pragma unroll
        for (int u = 0; u < WSU; ++u)
          d[u] = (q0 + u < nb && lane < nb) ? cld(D + 2 * ((long)(q0 + u) * 64 + lane)) : cmake<T>(0, 0);
#pragma unroll
        for (int u = 0; u < WSU; ++u)
#pragma unroll
          for (int v = 0; v < CM / 4; ++v) sl[v] = cfma(sl[v], d[u], rs[64 * r + q0 + u][wv + 4 * v]);
      }
      __syncthreads();
#pragma unroll
      for (int v = 0; v < CM / 4; ++v) rs[64 * r + lane][wv + 4 * v] = sl[v];
      __syncthreads();
      // the window's blocks still to solve: FWD rows [i0 + 64, s1), BWD rows [s0, i0)
      const int k0 = FWD ? i0 + 64 : s0;
      const int cnt = FWD ? s1 - k0 : i0 - s0;
      if (tid < cnt) {
        const int k = k0 + tid;
        cplx<T> acc[CM];
#pragma unroll
        for (int c = 0; c < CM; ++c) acc[c] = cmake<T>(0, 0);
        for (int l0 = 0; l0 < nb; l0 += WSU) {
          cplx<T> m[WSU];
#pragma unroll
          for (int u = 0; u < WSU; ++u)
            m[u] = l0 + u < nb ? cld(LU + 2 * ((long)(i0 + l0 + u) * ld + k)) : cmake<T>(0, 0);
#pragma unroll
          for (int u = 0; u < WSU; ++u)
#pragma unroll
            for (int c = 0; c < CM; ++c) acc[c] = cfma(acc[c], m[u], rs[64 * r + l0 + u][c]);
        }
#pragma unroll
        for (int c = 0; c < CM; ++c) rs[k - s0][c] = csub(rs[k - s0][c], acc[c]);
      }
    }
    __syncthreads();
  };
  {
    // the window's solution (have) or its right-hand side
    const T* src = have ? dst : (FWD ? V : Z);
    const long lds = have ? ldd_ : (FWD ? ldv : ldz);
    for (int e = tid; e < TRSV_WIN * CM; e += 256) {
      const int r = e / CM, c = e % CM;
      rs[r][c] = (r < W && c < C) ? cld(src + 2 * ((long)(w0 + r) * lds + c)) : cmake<T>(0, 0);
    }
  }
  if (!have) {
    solve_window(w0, w1);
    if (bx == 0)
      for (int e = tid; e < W * C; e += 256) {
        const int rr = e / C, c = e % C;
        cst(dst + 2 * ((long)(w0 + rr) * ldd_ + c), rs[rr][c]);
      }
  } else {
    __syncthreads();
  }
  // tile update with the whole window
  T* upd = FWD ? V : Z;
  const long ldu = FWD ? ldv : ldz;
  if constexpr (MF) {
    // complex64 tile update on the FP32 matrix cores (the real form of
    // k_gemv_t_mfma): wave wv owns the tile rows [64 wv, 64 wv + 64) as two
    // 32-row groups, lane (p, q) supplies rows 2p, 2p + 1 of each group (one
    // 16-byte load per group) and window row l + q; D[g][parity][i][c] holds
    // Re (c < 8) or Im (c >= 8) of column c & 7 for tile row 32 g + 2 i + parity
    static_assert(TILE == 256 && sizeof(T) == 4, "matrix-core tile update: complex64, 256-row tiles");
    const int p = lane & 15, q = lane >> 4, c = p & 7;
    const bool cim = p >= 8, cok = c < C;
    const int kt = FWD ? w1 + bx * 256 : w0 - 256 * (bx + 1);  // first row of the tile
    const int kr0 = min(max(kt + 64 * wv + 2 * p, 0), n - 2), kr1 = min(max(kt + 64 * wv + 32 + 2 * p, 0), n - 2);
    gv_f32x4 d00 = {0.f, 0.f, 0.f, 0.f}, d01 = d00, d10 = d00, d11 = d00;
    auto step = [&](const float4& m0, const float4& m1, float xr, float xi) {
      const float b1 = cim ? xi : xr, b2 = cim ? xr : -xi;
      d00 = __builtin_amdgcn_mfma_f32_16x16x4f32(m0.x, b1, d00, 0, 0, 0);
      d01 = __builtin_amdgcn_mfma_f32_16x16x4f32(m0.z, b1, d01, 0, 0, 0);
      d10 = __builtin_amdgcn_mfma_f32_16x16x4f32(m1.x, b1, d10, 0, 0, 0);
      d11 = __builtin_amdgcn_mfma_f32_16x16x4f32(m1.z, b1, d11, 0, 0, 0);
      d00 = __builtin_amdgcn_mfma_f32_16x16x4f32(m0.y, b2, d00, 0, 0, 0);
      d01 = __builtin_amdgcn_mfma_f32_16x16x4f32(m0.w, b2, d01, 0, 0, 0);
      d10 = __builtin_amdgcn_mfma_f32_16x16x4f32(m1.y, b2, d10, 0, 0, 0);
      d11 = __builtin_amdgcn_mfma_f32_16x16x4f32(m1.w, b2, d11, 0, 0, 0);
    };
    const float* LUf = reinterpret_cast<const float*>(LU);
    int l = 0;
    for (; l + 16 <= W; l += 16) {
      float4 m0[4], m1[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const long row = (long)(w0 + l + 4 * u + q) * ld;
        m0[u] = *reinterpret_cast<const float4*>(LUf + 2 * (row + kr0));
        m1[u] = *reinterpret_cast<const float4*>(LUf + 2 * (row + kr1));
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const cplx<T> x = cok ? rs[l + 4 * u + q][c] : cmake<T>(0, 0);
        step(m0[u], m1[u], x.re, x.im);
      }
    }
    for (; l < W; l += 4) {
      const int ll = l + q;
      float4 m0 = make_float4(0.f, 0.f, 0.f, 0.f), m1 = m0;
      cplx<T> x = cmake<T>(0, 0);
      if (ll < W) {
        const long row = (long)(w0 + ll) * ld;
        m0 = *reinterpret_cast<const float4*>(LUf + 2 * (row + kr0));
        m1 = *reinterpret_cast<const float4*>(LUf + 2 * (row + kr1));
        if (cok) x = rs[ll][c];
      }
      step(m0, m1, x.re, x.im);
    }
    // new right-hand side of this lane's (row, column, part) entries
    float nvs[2][2][4];
    float* updf = reinterpret_cast<float*>(upd);
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const gv_f32x4& dd = g ? (par ? d11 : d10) : (par ? d01 : d00);
          const int k = kt + 64 * wv + 32 * g + 8 * q + 2 * v + par;
          const bool lv = FWD ? k < n : k >= 0;
          float nv = 0.f;
          if (lv && cok) {
            float* pu = updf + 2 * ((long)k * ldu + c) + (cim ? 1 : 0);
            nv = *pu - dd[v];
            *pu = nv;
          }
          nvs[g][par][v] = nv;
        }
    if (ahead && bx == 0) {
      const int s0 = FWD ? w1 : max(0, w0 - 256), s1 = FWD ? min(n, w1 + 256) : w0;
      if (s1 > s0) {
        __syncthreads();  // every wave done with the current window in rs
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
          for (int par = 0; par < 2; ++par)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int k = kt + 64 * wv + 32 * g + 8 * q + 2 * v + par;
              const bool lv = FWD ? k < n : k >= 0;
              const int r = k - s0;
              if (lv && r >= 0 && r < TRSV_WIN && c < CM) {
                T* e = reinterpret_cast<T*>(&rs[r][c]);
                e[cim ? 1 : 0] = nvs[g][par][v];
              }
            }
        solve_window(s0, s1);
        for (int e = tid; e < (s1 - s0) * C; e += 256) {
          const int rr = e / C, cc = e % C;
          cst(dst + 2 * ((long)(s0 + rr) * ldd_ + cc), rs[rr][cc]);
        }
      }
    }
    return;
  }
  cplx<T> acc[CM];
#pragma unroll
  for (int c = 0; c < CM; ++c) acc[c] = cmake<T>(0, 0);
  int k;
  bool live;
  if constexpr (TILE == 256) {
    k = FWD ? w1 + bx * 256 + tid : w0 - 256 * (bx + 1) + tid;
    live = FWD ? k < n : k >= 0;
  } else {
    k = (FWD ? w1 : 0) + bx * 64 + lane;
    live = k < (FWD ? n : w0);
  }
  const int la = TILE == 64 ? 64 * wv : 0, lb = TILE == 64 ? min(W, 64 * wv + 64) : W;
  if (live) {
    int l = la;
    for (; l + 16 <= lb; l += 16) {
      cplx<T> m[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) m[u] = cld(LU + 2 * ((long)(w0 + l + u) * ld + k));
#pragma unroll
      for (int u = 0; u < 16; ++u)
#pragma unroll
        for (int c = 0; c < CM; ++c) acc[c] = cfma(acc[c], m[u], rs[l + u][c]);
    }
    for (; l < lb; ++l) {
      const cplx<T> m = cld(LU + 2 * ((long)(w0 + l) * ld + k));
#pragma unroll
      for (int c = 0; c < CM; ++c) acc[c] = cfma(acc[c], m, rs[l][c]);
    }
  }
  if constexpr (TILE == 64) {
    if (wv > 0)
#pragma unroll
      for (int c = 0; c < CM; ++c) part[wv - 1][lane][c] = acc[c];
    __syncthreads();
    if (wv > 0) return;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int c = 0; c < CM; ++c) acc[c] = cadd(acc[c], part[q][lane][c]);
  }
  cplx<T> nv[CM];
  if (live)
#pragma unroll
    for (int c = 0; c < CM; ++c)
      if (c < C) {
        T* pu = upd + 2 * ((long)k * ldu + c);
        nv[c] = csub(cld(pu), acc[c]);
        cst(pu, nv[c]);
      }
  if constexpr (TILE == 256) {
    // workgroup 0 holds the next window's updated right-hand side: solve it
    if (ahead && bx == 0) {
      const int s0 = FWD ? w1 : max(0, w0 - 256), s1 = FWD ? min(n, w1 + 256) : w0;
      if (s1 > s0) {
        __syncthreads();  // every thread done with the current window in rs
        const int r = k - s0;
#pragma unroll
        for (int c = 0; c < CM; ++c) {
          if (live && r >= 0 && r < TRSV_WIN) rs[r][c] = c < C ? nv[c] : cmake<T>(0, 0);
        }
        // rows of rs beyond the window (the last window of each direction is shorter) stay unused
        solve_window(s0, s1);
        for (int e = tid; e < (s1 - s0) * C; e += 256) {
          const int rr = e / C, c = e % C;
          cst(dst + 2 * ((long)(s0 + rr) * ldd_ + c), rs[rr][c]);
        }
      }
    }
  }
}

// Inverses of the 64 x 64 diagonal blocks of the LU factors (U_II^-1 upper,
// L_II^-1 unit lower), stored per block as [U^-1 | L^-1] (2 x 64 x 64,
// zero outside the triangle).  One workgroup of 64 threads per (block,
// matrix); thread j computes column j by substitution, touching only its own
// column of the LDS result, so no barriers are needed inside the loop.
template <typename T>
__global__ __launch_bounds__(64) void k_lu_blockinv(int n, const T* __restrict__ LU, long ld, long sLU, T* Dinv,
                                                    long sD) {
  const long b = blockIdx.y;
  const int ib = blockIdx.x, i0 = ib * 64, nb = min(64, n - i0);
  LU += 2 * b * sLU;
  T* D = Dinv + 2 * (b * sD + (long)ib * 2 * 64 * 64);
  __shared__ cplx<T> Tb[64][65];
  __shared__ cplx<T> X[64][65];
  const int j = threadIdx.x;
  for (int e = j; e < 64 * 64; e += 64) {
    const int r = e / 64, q = e % 64;
    Tb[r][q] = (r < nb && q < nb) ? cld(LU + 2 * ((long)(i0 + r) * ld + i0 + q)) : cmake<T>(0, 0);
  }
  __syncthreads();
  // U X = I (upper, back substitution)
  for (int i = 63; i >= 0; --i) {
    cplx<T> s = cmake<T>(i == j && i < nb ? T(1) : T(0), T(0));
    if (i < nb) {
      for (int l = i + 1; l < nb; ++l) s = cfms(s, Tb[i][l], X[l][j]);
      s = cdiv(s, Tb[i][i]);
    }
    X[i][j] = s;
  }
  for (int i = 0; i < 64; ++i) cst(D + 2 * ((long)i * 64 + j), X[i][j]);
  // L X = I (unit lower, forward substitution)
  for (int i = 0; i < 64; ++i) {
    cplx<T> s = cmake<T>(i == j && i < nb ? T(1) : T(0), T(0));
    if (i < nb)
      for (int l = 0; l < i; ++l) s = cfms(s, Tb[i][l], X[l][j]);
    X[i][j] = s;
  }
  for (int i = 0; i < 64; ++i) cst(D + 2 * (64 * 64 + (long)i * 64 + j), X[i][j]);
}

// x = P^T w: row interchanges ipiv[n-1], ..., ipiv[0] applied to V (n x C);
// one workgroup per matrix, one thread per column.
template <typename T>
__global__ __launch_bounds__(256) void k_perm_t(int n, int C, const int* __restrict__ ipiv, long sP, T* V, long ldv,
                                                long sV, int half, long hoff) {
  const long b = blockIdx.x;
  ipiv += b * sP;
  V += 2 * ((b % half) * sV + (b / half) * hoff);
  const int tid = threadIdx.x;
  __shared__ int sp[256];
  for (int base = n; base > 0; base -= 256) {
    const int lo = max(0, base - 256);
    __syncthreads();
    int moved = 0;
    if (tid < base - lo) {
      sp[tid] = ipiv[lo + tid];
      moved = sp[tid] != lo + tid;
    }
    // chunks without interchanges (ipiv[i] = i, the factorisations without
    // them) are skipped instead of walked pivot by pivot
    if (!__syncthreads_or(moved)) continue;
    if (tid < C) {
      for (int i = base - 1; i >= lo; --i) {
        const int p = sp[i - lo];
        if (p != i) {
          T* x = V + 2 * ((long)i * ldv + tid);
          T* y = V + 2 * ((long)p * ldv + tid);
          const cplx<T> u = cld(x), v = cld(y);
          cst(x, v);
          cst(y, u);
        }
      }
    }
  }
}

// A[b] = identity (n x n buffer, leading dimension ld)
template <typename T>
__global__ void k_identity(int n, T* A, long ld, long sA) {
  const long b = blockIdx.y;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int r = e / n, c = e % n;
  cst(A + 2 * (b * sA + (long)r * ld + c), cmake<T>(r == c ? T(1) : T(0), T(0)));
}

// D = S1 - S2;  (mode 0)
// S2 <- e1 S1 - e2 S2 + k^2 (e1-e2)^2 E o (n n^T)  (mode 1)
// per-wavelength scalars in par[b] = {e1r, e1i, e2r, e2i, k}
template <typename T>
__global__ void k_sigma(int mode, int n, const T* S1, T* S2, T* D, const T* E, long ld, long stride,
                        const double* __restrict__ nvec, const double* __restrict__ par, int gdiv) {
  const long b = blockIdx.y;
  nvec += 3L * n * mnp_geo_slot(b, gdiv);
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int r = e / n, c = e % n;
  const long o = 2 * (b * stride + (long)r * ld + c);
  const cplx<T> s1 = cld(S1 + o), s2 = cld(S2 + o);
  if (mode == 0) {
    cst(D + o, csub(s1, s2));
    return;
  }
  const double* q = par + 5 * b;
  const cplx<double> e1 = cmake(q[0], q[1]), e2 = cmake(q[2], q[3]);
  const double k = q[4];
  const cplx<double> L = csub(e1, e2);
  const cplx<double> cf = cscale(cmul(L, L), k * k);
  const double nn = nvec[3 * r] * nvec[3 * c] + nvec[3 * r + 1] * nvec[3 * c + 1] + nvec[3 * r + 2] * nvec[3 * c + 2];
  const cplx<T> ev = cld(E + o);
  const cplx<double> ed = cmake((double)ev.re, (double)ev.im);
  const cplx<double> s1d = cmake((double)s1.re, (double)s1.im), s2d = cmake((double)s2.re, (double)s2.im);
  cplx<double> v = csub(cmul(e1, s1d), cmul(e2, s2d));
  v = cadd(v, cscale(cmul(cf, ed), nn));
  cst(S2 + o, cmake<T>((T)v.re, (T)v.im));
}

// --- small elementwise steps of the direct solve (bem_ret.py:1725-1810) ---
// per-wavelength parameters par[b] = {e1r, e1i, e2r, e2i, k}
struct Par {
  cplx<double> e1, e2;
  double k;
};
__device__ __forceinline__ Par load_par(const double* par, long b) {
  const double* q = par + 5 * b;
  Par p;
  p.e1 = cmake(q[0], q[1]);
  p.e2 = cmake(q[2], q[3]);
  p.k = q[4];
  return p;
}
template <typename T>
__device__ __forceinline__ cplx<double> ldd(const T* p) {
  const cplx<T> v = cld(p);
  return cmake((double)v.re, (double)v.im);
}
template <typename T>
__device__ __forceinline__ void std_(T* p, cplx<double> v) {
  cst(p, cmake<T>((T)v.re, (T)v.im));
}
__device__ __forceinline__ cplx<double> ik_times(double k, cplx<double> z) { return cmake(-k * z.im, k * z.re); }

// step 1: W holds [Sigma1 phi | Sigma1 a] in columns [0, 4P).  Writes
//   alpha' = alpha - Sigma1 a + ik n (e1 phi)        -> W[:, 5P:8P]
//   De'    = De - e1 Sigma1 phi + ik n.(e1 a)        -> W[:, 4P:5P]
template <typename T>
__global__ void k_step_mod(int n, int P, const T* R, T* W, long ld, long stride, const double* nvec,
                           const double* par, int gdiv) {
  const long b = blockIdx.y;
  nvec += 3L * n * mnp_geo_slot(b, gdiv);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const double nx[3] = {nvec[3 * i], nvec[3 * i + 1], nvec[3 * i + 2]};
  const T* r = R + 2 * (b * stride + (long)i * ld);
  T* w = W + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    const cplx<double> e1phi = cmul(q.e1, ldd(r + 2 * p));
    cplx<double> nda = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) {
      const cplx<double> a = ldd(r + 2 * (P + c * P + p));
      nda = cadd(nda, cscale(a, nx[c]));
      cplx<double> al = ldd(r + 2 * (5 * P + c * P + p));
      al = csub(al, ldd(w + 2 * (P + c * P + p)));
      al = cadd(al, ik_times(q.k, cscale(e1phi, nx[c])));
      std_(w + 2 * (5 * P + c * P + p), al);
    }
    cplx<double> de = ldd(r + 2 * (4 * P + p));
    de = csub(de, cmul(q.e1, ldd(w + 2 * p)));
    de = cadd(de, ik_times(q.k, cmul(q.e1, nda)));
    std_(w + 2 * (4 * P + p), de);
  }
}

// step 2: U = Delta^-1 alpha' in W[:, P:4P].  rhs = De' + ik n.((e1-e2) U) -> W[:, 0:P]
template <typename T>
__global__ void k_step_rhs(int n, int P, T* W, long ld, long stride, const double* nvec, const double* par,
                           int gdiv) {
  const long b = blockIdx.y;
  nvec += 3L * n * mnp_geo_slot(b, gdiv);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const cplx<double> Ld = csub(q.e1, q.e2);
  T* w = W + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    cplx<double> s = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) s = cadd(s, cscale(ldd(w + 2 * (P + c * P + p)), nvec[3 * i + c]));
    const cplx<double> v = cadd(ldd(w + 2 * (4 * P + p)), ik_times(q.k, cmul(Ld, s)));
    std_(w + 2 * p, v);
  }
}

// step 3: sig2 = Sigma^-1 rhs in W[:, 0:P].  T = ik n ((e1-e2) sig2) + alpha' -> W[:, P:4P]
template <typename T>
__global__ void k_step_h2pre(int n, int P, T* W, long ld, long stride, const double* nvec, const double* par,
                             int gdiv) {
  const long b = blockIdx.y;
  nvec += 3L * n * mnp_geo_slot(b, gdiv);
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const cplx<double> Ld = csub(q.e1, q.e2);
  T* w = W + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    const cplx<double> ls = cmul(Ld, ldd(w + 2 * p));
    for (int c = 0; c < 3; ++c) {
      const cplx<double> v = cadd(ik_times(q.k, cscale(ls, nvec[3 * i + c])), ldd(w + 2 * (5 * P + c * P + p)));
      std_(w + 2 * (P + c * P + p), v);
    }
  }
}

// step 4: h2 = Delta^-1 T in W[:, 4P:7P] (written there by the GEMV).
//   X[:, 0:4P] = [sig2 + phi | h2 + a],  X[:, 4P:8P] = [sig2 | h2]
template <typename T>
__global__ void k_step_pack(int n, int P, const T* R, const T* W, T* X, long ld, long stride) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* r = R + 2 * (b * stride + (long)i * ld);
  const T* w = W + 2 * (b * stride + (long)i * ld);
  T* x = X + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    const cplx<T> s2 = cld(w + 2 * p);
    cst(x + 2 * p, cadd(s2, cld(r + 2 * p)));
    cst(x + 2 * (4 * P + p), s2);
    for (int c = 0; c < 3; ++c) {
      const cplx<T> h2 = cld(w + 2 * (4 * P + c * P + p));
      cst(x + 2 * (P + c * P + p), cadd(h2, cld(r + 2 * (P + c * P + p))));
      cst(x + 2 * (5 * P + c * P + p), h2);
    }
  }
}

// Dst[b][i][0..C) = Src[b][i][0..C) for n rows (row strides ld, system
// stride stride): the columns a solve with an explicit inverse wrote to its
// work space, back into the vector it solves in place
template <typename T>
__global__ void k_copy_cols(int n, int C, const T* __restrict__ Src, T* __restrict__ Dst, long ld, long stride) {
  const long b = blockIdx.y;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * C) return;
  const long i = e / C, c = e % C;
  cst(Dst + 2 * (b * stride + i * ld + c), cld(Src + 2 * (b * stride + i * ld + c)));
}

// residual of the 8N operator (GdA&H Eqs. 10, 11, 14, 17; bem_ret_iter.py:1362-1501):
// Y = [G1 X1 | H1 X1 | G2 X2 | H2 X2], each (N, 4P), in one (N, 16P) buffer.
//   phi = G1 s1 - G2 s2;  a = G1 h1 - G2 h2
//   alpha = H1 h1 - H2 h2 - ik n (e1 G1 s1 - e2 G2 s2)
//   De = e1 H1 s1 - e2 H2 s2 - ik n.(e1 G1 h1 - e2 G2 h2)
// Rout = R - A X; with R == nullptr, Rout = A X.  The norms of the refinement
// come from k_norms2 afterwards.
__global__ void k_residual(int n, int P, const double* R, const double* Y, double* Rout, long ldr, long sr,
                           long ldy, long sy, const double* nvec, const double* par, int gdiv) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  nvec += 3L * n * mnp_geo_slot(b, gdiv);
  if (i < n) {
    const Par q = load_par(par, b);
    const double* r = R ? R + 2 * (b * sr + (long)i * ldr) : nullptr;
    const double* y = Y + 2 * (b * sy + (long)i * ldy);
    double* o = Rout + 2 * (b * sr + (long)i * ldr);
    const int P4 = 4 * P;
    auto G1 = [&](int col) { return ldd(y + 2 * col); };
    auto H1 = [&](int col) { return ldd(y + 2 * (P4 + col)); };
    auto G2 = [&](int col) { return ldd(y + 2 * (2 * P4 + col)); };
    auto H2 = [&](int col) { return ldd(y + 2 * (3 * P4 + col)); };
    auto put = [&](int col, cplx<double> v) {
      if (R == nullptr) {  // plain product A X
        std_(o + 2 * col, v);
        return;
      }
      std_(o + 2 * col, csub(ldd(r + 2 * col), v));
    };
    for (int p = 0; p < P; ++p) {
      put(p, csub(G1(p), G2(p)));
      const cplx<double> t = csub(cmul(q.e1, G1(p)), cmul(q.e2, G2(p)));
      cplx<double> ndg = cmake(0.0, 0.0);
      for (int c = 0; c < 3; ++c) {
        const int hc = P + c * P + p;
        put(hc, csub(G1(hc), G2(hc)));
        const cplx<double> al = csub(csub(H1(hc), H2(hc)), ik_times(q.k, cscale(t, nvec[3 * i + c])));
        put(4 * P + hc, al);
        ndg = cadd(ndg, cscale(csub(cmul(q.e1, G1(hc)), cmul(q.e2, G2(hc))), nvec[3 * i + c]));
      }
      const cplx<double> de = csub(csub(cmul(q.e1, H1(p)), cmul(q.e2, H2(p))), ik_times(q.k, ndg));
      put(4 * P + p, de);
    }
  }
}

// nrm[b] = {sum |Out|^2, sum |Rin|^2} over the n x C block of wavelength b in a
// fixed order: one workgroup per wavelength, each thread summing a fixed
// strided set of elements, then a fixed butterfly per wave and a sum over the
// waves in index order.  The refinement's stopping decision therefore does
// not change from run to run.
__global__ __launch_bounds__(1024) void k_norms2(int n, int C, const double* __restrict__ Rin,
                                                 const double* __restrict__ Out, long ld, long stride,
                                                 double* __restrict__ nrm) {
  const long b = blockIdx.x;
  const double* r = Rin + 2 * b * stride;
  const double* o = Out + 2 * b * stride;
  double ar = 0.0, ab = 0.0;
  const long total = (long)n * C;
  for (long e = threadIdx.x; e < total; e += blockDim.x) {
    const long at = 2 * ((e / C) * ld + e % C);
    const double o0 = o[at], o1 = o[at + 1], r0 = r[at], r1 = r[at + 1];
    ar += o0 * o0 + o1 * o1;
    ab += r0 * r0 + r1 * r1;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ar += __shfl_xor(ar, off);
    ab += __shfl_xor(ab, off);
  }
  __shared__ double sr[16], sb[16];
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sr[w] = ar;
    sb[w] = ab;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0, tb = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      tr += sr[k];
      tb += sb[k];
    }
    nrm[2 * b] = tr;
    nrm[2 * b + 1] = tb;
  }
}

// X64 += (double) D ;  R32 = (float) R64   (refinement bookkeeping)
template <typename TS>
__global__ void k_axpy_cast(long count, const TS* src, double* dst) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < count) dst[e] += (double)src[e];
}
template <typename TD>
__global__ void k_copy_cast(long count, const double* src, TD* dst) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < count) dst[e] = (TD)src[e];
}
__global__ void k_fill_double(long count, double* p, double v) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e < count) p[e] = v;
}

// plane wave a = e^{i k pos.dir}/(i k0) pol, a' = i k (n.dir) a on faces whose
// side touches the embedding medium (planewave_ret.py:257-361); then the
// right side R = [phi=0 | a2-a1 | De | alpha=a2'-a1'] with
// De = -i k0 (eps2 n.a2 - eps1 n.a1) (bem_ret.py:1360-1429).
// pw[b] = {k0, kr, ki, e1r, e1i, e2r, e2i}; pol, dir (P, 3) real.
__global__ void k_planewave(int n, int P, const double* pos, const double* nvec, const int* in_med,
                            const int* out_med, const double* pol, const double* dir, const double* pw,
                            double* R, long ld, long stride, double* pot, long spot, int gdiv) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long gs = mnp_geo_slot(b, gdiv);
  pos += 3 * n * gs;
  nvec += 3 * n * gs;
  in_med += n * gs;
  out_med += n * gs;
  const double* q = pw + 7 * b;
  const double k0 = q[0], kr = q[1], ki = q[2];
  const cplx<double> e1 = cmake(q[3], q[4]), e2 = cmake(q[5], q[6]);
  double* r = R + 2 * (b * stride + (long)i * ld);
  const double px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
  const double nx[3] = {nvec[3 * i], nvec[3 * i + 1], nvec[3 * i + 2]};
  for (int p = 0; p < P; ++p) {
    const double* dp = dir + 3 * p;
    const double* pp = pol + 3 * p;
    const double kd = px * dp[0] + py * dp[1] + pz * dp[2];
    const double nd = nx[0] * dp[0] + nx[1] * dp[1] + nx[2] * dp[2];
    // phase / (i k0) = e^{i k kd} * (-i / k0)
    const cplx<double> ph = cexpi(kr * kd, ki * kd);
    const cplx<double> amp = cmake(ph.im / k0, -ph.re / k0);
    const cplx<double> ikn = cmake(-ki * nd, kr * nd);  // i k (n.dir)
    const cplx<double> z = cmake(0.0, 0.0);
    const cplx<double> a1 = in_med[i] ? amp : z, a2 = out_med[i] ? amp : z;
    const cplx<double> ap1 = in_med[i] ? cmul(ikn, amp) : z, ap2 = out_med[i] ? cmul(ikn, amp) : z;
    std_(r + 2 * p, z);
    double npol = nx[0] * pp[0] + nx[1] * pp[1] + nx[2] * pp[2];
    for (int c = 0; c < 3; ++c) {
      std_(r + 2 * (P + c * P + p), cscale(csub(a2, a1), pp[c]));
      std_(r + 2 * (5 * P + c * P + p), cscale(csub(ap2, ap1), pp[c]));
      if (pot) {
        double* o = pot + 2 * (b * spot + (long)i * 12 * P);
        std_(o + 2 * (0 * 3 * P + c * P + p), cscale(a1, pp[c]));
        std_(o + 2 * (1 * 3 * P + c * P + p), cscale(ap1, pp[c]));
        std_(o + 2 * (2 * 3 * P + c * P + p), cscale(a2, pp[c]));
        std_(o + 2 * (3 * 3 * P + c * P + p), cscale(ap2, pp[c]));
      }
    }
    const cplx<double> t = csub(cmul(e2, cscale(a2, npol)), cmul(e1, cscale(a1, npol)));
    std_(r + 2 * (4 * P + p), cmake(k0 * t.im, -k0 * t.re));  // -i k0 t
  }
}

// ext[b][p] = 4 pi / k Im(conj(pol_p) . e_p),
// e_p = sum_j e^{-i k dir_p.pos_j} area_j (i k0 h_j - i k dir_p sig_j) over
// faces whose inside (sig1, h1) or outside (sig2, h2) is the medium
// (spectrum_ret.py:231-392, planewave_ret.py:437-504).  One workgroup per (b, p).
__global__ __launch_bounds__(256) void k_extinction(int n, int P, const double* X, long ld, long stride,
                                                    const double* pos, const double* area, const int* in_med,
                                                    const int* out_med, const double* pol, const double* dir,
                                                    const double* pw, double* ext, int gdiv) {
  const long b = blockIdx.y;
  const int p = blockIdx.x;
  const long gs = mnp_geo_slot(b, gdiv);
  pos += 3 * n * gs;
  area += n * gs;
  in_med += n * gs;
  out_med += n * gs;
  const double* q = pw + 7 * b;
  const double k0 = q[0], kr = q[1], ki = q[2];
  const double* dp = dir + 3 * p;
  const double* pp = pol + 3 * p;
  cplx<double> acc = cmake(0.0, 0.0);
  for (int j = threadIdx.x; j < n; j += 256) {
    const double kd = pos[3 * j] * dp[0] + pos[3 * j + 1] * dp[1] + pos[3 * j + 2] * dp[2];
    // e^{-i k kd} area
    const cplx<double> ph = cscale(cexpi(-kr * kd, -ki * kd), area[j]);
    const double* x = X + 2 * (b * stride + (long)j * ld);
    for (int side = 0; side < 2; ++side) {
      if (!(side == 0 ? in_med[j] : out_med[j])) continue;
      const int off = side * 4 * P;
      const cplx<double> sg = ldd(x + 2 * (off + p));
      cplx<double> ph_ = cmake(0.0, 0.0);  // conj(pol).h
      for (int c = 0; c < 3; ++c) ph_ = cadd(ph_, cscale(ldd(x + 2 * (off + P + c * P + p)), pp[c]));
      const double pd = pp[0] * dp[0] + pp[1] * dp[1] + pp[2] * dp[2];
      // i k0 (pol.h) - i k (pol.dir) sig
      const cplx<double> t1 = cmake(-k0 * ph_.im, k0 * ph_.re);
      const cplx<double> ksg = cmul(cmake(kr, ki), sg);
      const cplx<double> t2 = cmake(-ksg.im * pd, ksg.re * pd);
      acc = cadd(acc, cmul(ph, csub(t1, t2)));
    }
  }
  __shared__ double sr[4], si[4];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    acc.re += __shfl_xor(acc.re, off);
    acc.im += __shfl_xor(acc.im, off);
  }
  if ((threadIdx.x & 63) == 0) {
    sr[threadIdx.x >> 6] = acc.re;
    si[threadIdx.x >> 6] = acc.im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double im = si[0] + si[1] + si[2] + si[3];
    // Re(4 pi / k) * Im(conj(pol).e); pol is real
    const double den = kr * kr + ki * ki;
    ext[b * P + p] = 4.0 * M_PI * kr / den * im;
  }
}

// Far field in the directions dirs[d] (unit vectors, the pinfty sphere of
// spectrum_ret.py:14-45) and its radiated power (spectrum_ret.py:231-459,
// planewave_ret.py:391-435): with S_h = sum_j ph_j h_j, S_s = sum_j ph_j sig_j,
// ph_j = e^{-i k dir.pos_j} area_j over the medium-side faces,
//   e = i k0 S_h - i k dir S_s,   h = i k dir x S_h,
//   dsca[b][d][p] = 0.5 Re(dir . (e x conj h)) / (0.5 nb),  nb = Re k / k0.
// One workgroup per (direction, wavelength); P <= 8.  e, h (optional):
// (batch, ndir, 3, P) complex128.
__global__ __launch_bounds__(256) void k_farfield(int n, int P, const double* X, long ld, long stride,
                                                  const double* pos, const double* area, const int* in_med,
                                                  const int* out_med, const double* dirs, const double* pw,
                                                  double* dsca, double* E, double* Hf) {
  const int d = blockIdx.x, ndir = gridDim.x;
  const long b = blockIdx.y;
  const double* q = pw + 7 * b;
  const double k0 = q[0], kr = q[1], ki = q[2];
  const double dx = dirs[3 * d], dy = dirs[3 * d + 1], dz = dirs[3 * d + 2];
  cplx<double> acc[8][4];  // [p][h_x, h_y, h_z, sig]
#pragma unroll
  for (int p = 0; p < 8; ++p)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[p][c] = cmake(0.0, 0.0);
  for (int j = threadIdx.x; j < n; j += 256) {
    const double kd = pos[3 * j] * dx + pos[3 * j + 1] * dy + pos[3 * j + 2] * dz;
    const cplx<double> ph = cscale(cexpi(-kr * kd, -ki * kd), area[j]);
    const double* x = X + 2 * (b * stride + (long)j * ld);
    for (int side = 0; side < 2; ++side) {
      if (!(side == 0 ? in_med[j] : out_med[j])) continue;
      const int off = side * 4 * P;
#pragma unroll
      for (int p = 0; p < 8; ++p) {
        if (p < P) {
#pragma unroll
          for (int c = 0; c < 3; ++c) acc[p][c] = cfma(acc[p][c], ph, ldd(x + 2 * (off + P + c * P + p)));
          acc[p][3] = cfma(acc[p][3], ph, ldd(x + 2 * (off + p)));
        }
      }
    }
  }
  __shared__ double red[4][8][4][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < 8; ++p)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      cplx<double> v = acc[p][c];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        v.re += __shfl_xor(v.re, off);
        v.im += __shfl_xor(v.im, off);
      }
      if (lane == 0) {
        red[wv][p][c][0] = v.re;
        red[wv][p][c][1] = v.im;
      }
    }
  __syncthreads();
  if (threadIdx.x < P) {
    const int p = threadIdx.x;
    cplx<double> S[4];
    for (int c = 0; c < 4; ++c)
      S[c] = cmake(red[0][p][c][0] + red[1][p][c][0] + red[2][p][c][0] + red[3][p][c][0],
                   red[0][p][c][1] + red[1][p][c][1] + red[2][p][c][1] + red[3][p][c][1]);
    const double dv[3] = {dx, dy, dz};
    const cplx<double> kc = cmake(kr, ki);
    cplx<double> e[3], h[3];
    for (int c = 0; c < 3; ++c) {
      // i k0 S_h - i k dir S_s
      const cplx<double> t = csub(cscale(S[c], k0), cscale(cmul(kc, S[3]), dv[c]));
      e[c] = cmake(-t.im, t.re);
    }
    const cplx<double> cr[3] = {csub(cscale(S[2], dv[1]), cscale(S[1], dv[2])),
                                csub(cscale(S[0], dv[2]), cscale(S[2], dv[0])),
                                csub(cscale(S[1], dv[0]), cscale(S[0], dv[1]))};
    for (int c = 0; c < 3; ++c) {
      const cplx<double> t = cmul(kc, cr[c]);
      h[c] = cmake(-t.im, t.re);
    }
    // dir . (e x conj h)
    auto cj = [](cplx<double> z) { return cmake(z.re, -z.im); };
    const cplx<double> px = csub(cmul(e[1], cj(h[2])), cmul(e[2], cj(h[1])));
    const cplx<double> py = csub(cmul(e[2], cj(h[0])), cmul(e[0], cj(h[2])));
    const cplx<double> pz = csub(cmul(e[0], cj(h[1])), cmul(e[1], cj(h[0])));
    const double nb = kr / k0;
    dsca[(b * ndir + d) * P + p] = 0.5 * (dx * px.re + dy * py.re + dz * pz.re) / (0.5 * nb);
    if (E != nullptr) {
      for (int c = 0; c < 3; ++c) {
        double* o = E + 2 * (((b * ndir + d) * 3 + c) * P + p);
        o[0] = e[c].re;
        o[1] = e[c].im;
        double* oh = Hf + 2 * (((b * ndir + d) * 3 + c) * P + p);
        oh[0] = h[c].re;
        oh[1] = h[c].im;
      }
    }
  }
}

// sca[b][p] = sum_d w_d dsca[b][d][p] in a fixed order
__global__ void k_sca_reduce(int ndir, int P, int total, const double* w, const double* dsca, double* sca) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;  // t = b * P + p
  if (t >= total) return;
  const int p = t % P;
  const long b = t / P;
  double s = 0.0;
  for (int d = 0; d < ndir; ++d) s += w[d] * dsca[(b * ndir + d) * P + p];
  sca[t] = s;
}

template <typename T>
int gemv_t(int n, int C, const T* M, long ldm, long sM, const T* X, long ldx, long sX, T* Y, long ldy, long sY,
           double alr, double ali, double beta, int batch, hipStream_t st) {
  for (int c0 = 0; c0 < C; c0 += MAXC) {
    const int cc = std::min(MAXC, C - c0);
    if constexpr (sizeof(T) == 4) {
      if (n % 2 == 0 && ldm % 2 == 0 && sM % 2 == 0 && ((uintptr_t)M & 15) == 0) {
        if (cc > 4 && cc <= 8 && g_gemv_mfma) {  // at <= 4 columns the VALU form keeps up with HBM
          hipLaunchKernelGGL(k_gemv_t_mfma, dim3(mnp_ceil_div(n, 32), batch), dim3(256), 0, st, n, cc,
                             (const float*)M, ldm, sM, (const float*)(X + 2 * c0), ldx, sX, (float*)(Y + 2 * c0), ldy,
                             sY, alr, ali, beta);
          MNP_LAUNCH_CHECK();
          continue;
        }
        const dim3 grid2(mnp_ceil_div(n, 128), batch);
#define MNP_GEMV2(CM)                                                                                             \
  hipLaunchKernelGGL((k_gemv_t2<CM>), grid2, dim3(256), 0, st, n, cc, (const float*)M, ldm, sM,                   \
                     (const float*)(X + 2 * c0), ldx, sX, (float*)(Y + 2 * c0), ldy, sY, alr, ali, beta)
        if (cc <= 4)
          MNP_GEMV2(4);
        else if (cc <= 8)
          MNP_GEMV2(8);
        else
          MNP_GEMV2(16);
#undef MNP_GEMV2
        MNP_LAUNCH_CHECK();
        continue;
      }
    }
    const dim3 grid(mnp_ceil_div(n, 64), batch);
#define MNP_GEMV(CM)                                                                                            \
  hipLaunchKernelGGL((k_gemv_t<T, CM>), grid, dim3(256), 0, st, n, cc, M, ldm, sM, X + 2 * c0, ldx, sX, Y + 2 * c0, \
                     ldy, sY, alr, ali, beta)
    if (cc <= 4)
      MNP_GEMV(4);
    else if (cc <= 8)
      MNP_GEMV(8);
    else
      MNP_GEMV(16);
#undef MNP_GEMV
    MNP_LAUNCH_CHECK();
  }
  return 0;
}

// Both triangular solves of lu_solve_t by window steps (TRSV_WIN rows per
// launch; tiles of 64 rows with the window split over the waves when the
// 256-row tiles would give fewer than 512 workgroups).
template <typename T, int CM, int TILE>
int trsv_t_windows_cm(int n, int C, const T* LU, long ld, long sLU, T* V, long ldv, long sV, T* Z, long ldz,
                      long sZ, const T* Dinv, long sD, int half, long hoff, int batch, hipStream_t st) {
  const int nblk = (n + 63) / 64, R = TRSV_WIN / 64;
  // with 256-row tiles every launch after the first reads its window's
  // solution, written by workgroup 0 of the launch before
  const int la = TILE == 256 ? 1 : 0;
  bool first = true;
  // complex64 tile updates on the matrix cores: row pairs by 16-byte loads
  constexpr bool can_mf = TILE == 256 && sizeof(T) == 4;
  const bool mf = can_mf && (g_trsv_mfma & (CM > 4 ? 1 : 2)) && n % 2 == 0 && ld % 2 == 0 && sLU % 2 == 0 &&
                  ((uintptr_t)LU & 15) == 0;
  for (int b0 = 0; b0 < nblk; b0 += R) {
    const int w0 = 64 * b0, w1 = std::min(n, 64 * (b0 + R));
    if (!first && la && w1 >= n) break;  // solved by the previous launch, nothing left to update
    const int ntf = std::max(1, mnp_ceil_div(n - w1, TILE));
    const dim3 grid = TILE == 256 ? dim3(ntf * batch) : dim3(ntf, batch);
    if (mf)
      hipLaunchKernelGGL((k_trsv_t_win<T, true, CM, TILE, can_mf>), grid, dim3(256), 0, st, n, C, w0, w1,
                         first ? 0 : la, la, LU, ld, sLU, V, ldv, sV, Z, ldz, sZ, Dinv, sD, half, hoff);
    else
      hipLaunchKernelGGL((k_trsv_t_win<T, true, CM, TILE>), grid, dim3(256), 0, st, n, C, w0, w1, first ? 0 : la,
                         la, LU, ld, sLU, V, ldv, sV, Z, ldz, sZ, Dinv, sD, half, hoff);
    MNP_LAUNCH_CHECK();
    first = false;
  }
  first = true;
  for (int b1 = nblk; b1 > 0; b1 -= R) {
    const int b0 = std::max(0, b1 - R);
    const int w0 = 64 * b0, w1 = std::min(n, 64 * b1);
    if (!first && la && w0 <= 0) break;
    const int ntb = std::max(1, mnp_ceil_div(w0, TILE));
    const dim3 grid = TILE == 256 ? dim3(ntb * batch) : dim3(ntb, batch);
    if (mf)
      hipLaunchKernelGGL((k_trsv_t_win<T, false, CM, TILE, can_mf>), grid, dim3(256), 0, st, n, C, w0, w1,
                         first ? 0 : la, la, LU, ld, sLU, V, ldv, sV, Z, ldz, sZ, Dinv, sD, half, hoff);
    else
      hipLaunchKernelGGL((k_trsv_t_win<T, false, CM, TILE>), grid, dim3(256), 0, st, n, C, w0, w1, first ? 0 : la,
                         la, LU, ld, sLU, V, ldv, sV, Z, ldz, sZ, Dinv, sD, half, hoff);
    MNP_LAUNCH_CHECK();
    first = false;
  }
  return 0;
}

template <typename T>
int trsv_t_windows(int n, int C, const T* LU, long ld, long sLU, T* V, long ldv, long sV, T* Z, long ldz, long sZ,
                   const T* Dinv, long sD, int half, long hoff, int batch, hipStream_t st) {
  const bool wide = (long)mnp_ceil_div(n, 256) * batch >= 512;
#define MNP_WIN(CM)                                                                                             \
  return wide ? trsv_t_windows_cm<T, CM, 256>(n, C, LU, ld, sLU, V, ldv, sV, Z, ldz, sZ, Dinv, sD, half, hoff, \
                                              batch, st)                                                     \
              : trsv_t_windows_cm<T, CM, 64>(n, C, LU, ld, sLU, V, ldv, sV, Z, ldz, sZ, Dinv, sD, half, hoff,  \
                                             batch, st);
  if (C <= 4) MNP_WIN(4)
  if (C <= 8) MNP_WIN(8)
  MNP_WIN(16)
#undef MNP_WIN
}

// V <- M^-1 V (n x C) with the LU of M^T; Z (n x C) is work space.  Systems
// b >= half solve the columns hoff to the right in the vectors of b - half.
template <typename T>
int lu_solve_t(int n, int C, const T* LU, long ld, long sLU, const int* ipiv, long sP, T* V, long ldv, long sV,
               T* Z, long ldz, long sZ, int batch, hipStream_t st, const T* Dinv = nullptr, long sD = 0,
               int half = 0, long hoff = 0) {
  if (half <= 0) half = batch;
  const int nblk = (n + 63) / 64;
  for (int c0 = 0; c0 < C; c0 += MAXC) {
    const int cc = std::min(MAXC, C - c0);
    T* v = V + 2 * c0;
    T* z = Z + 2 * c0;
    if (Dinv) {
      int rc = trsv_t_windows<T>(n, cc, LU, ld, sLU, v, ldv, sV, z, ldz, sZ, Dinv, sD, half, hoff, batch, st);
      if (rc) return rc;
      hipLaunchKernelGGL(k_perm_t<T>, dim3(batch), dim3(256), 0, st, n, cc, ipiv, sP, v, ldv, sV, half, hoff);
      MNP_LAUNCH_CHECK();
      continue;
    }
    for (int ib = 0; ib < nblk; ++ib) {
      const int i0 = ib * 64, i1 = std::min(n, i0 + 64);
      dim3 grid(std::max(1, mnp_ceil_div(n - i1, 256)), batch);
      hipLaunchKernelGGL((k_trsv_t_step<T, true>), grid, dim3(256), 0, st, n, cc, i0, LU, ld, sLU, v, ldv, sV, z,
                         ldz, sZ, Dinv, sD, half, hoff);
      MNP_LAUNCH_CHECK();
    }
    for (int ib = nblk - 1; ib >= 0; --ib) {
      const int i0 = ib * 64;
      dim3 grid(std::max(1, mnp_ceil_div(i0, 256)), batch);
      hipLaunchKernelGGL((k_trsv_t_step<T, false>), grid, dim3(256), 0, st, n, cc, i0, LU, ld, sLU, v, ldv, sV, z,
                         ldz, sZ, Dinv, sD, half, hoff);
      MNP_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_perm_t<T>, dim3(batch), dim3(256), 0, st, n, cc, ipiv, sP, v, ldv, sV, half, hoff);
    MNP_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace

extern "C" int mnpbem_gemv_mfma(int on) {
  const int prev = g_gemv_mfma;
  g_gemv_mfma = on != 0;
  return prev;
}

extern "C" int mnpbem_trsv_mfma(int mask) {
  const int prev = g_trsv_mfma;
  g_trsv_mfma = mask < 0 ? 3 : (mask & 3);
  return prev;
}

extern "C" int mnpbem_gemv_t_batched(int dtype, int n, int C, const void* M, long ldm, long sM, const void* X,
                                     long ldx, long sX, void* Y, long ldy, long sY, double alr, double ali,
                                     double beta, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || C <= 0 || batch <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    return gemv_t<double>(n, C, (const double*)M, ldm, sM, (const double*)X, ldx, sX, (double*)Y, ldy, sY, alr,
                          ali, beta, batch, st);
  if (mnp_is_c64(dtype))
    return gemv_t<float>(n, C, (const float*)M, ldm, sM, (const float*)X, ldx, sX, (float*)Y, ldy, sY, alr, ali,
                         beta, batch, st);
  mnp_set_error("gemv_t: unknown dtype");
  return 1;
}

extern "C" int mnpbem_lu_solve_t_batched(int dtype, int n, int C, const void* LU, long ld, long sLU,
                                         const int* ipiv, long sP, void* V, long ldv, long sV, void* Z, long ldz,
                                         long sZ, const void* Dinv, long sD, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || C <= 0 || batch <= 0) return 0;
  MNP_REQUIRE(ldz >= C, "lu_solve_t: work space narrower than the right-hand side");
  if (dtype == MNP_DTYPE_C128)
    return lu_solve_t<double>(n, C, (const double*)LU, ld, sLU, ipiv, sP, (double*)V, ldv, sV, (double*)Z, ldz,
                              sZ, batch, st, (const double*)Dinv, sD);
  if (mnp_is_c64(dtype))
    return lu_solve_t<float>(n, C, (const float*)LU, ld, sLU, ipiv, sP, (float*)V, ldv, sV, (float*)Z, ldz, sZ,
                             batch, st, (const float*)Dinv, sD);
  mnp_set_error("lu_solve_t: unknown dtype");
  return 1;
}

extern "C" int mnpbem_identity_batched(int dtype, int n, void* A, long ld, long sA, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || batch <= 0) return 0;
  dim3 grid(mnp_ceil_div((long)n * n, 256), batch);
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_identity<double>, grid, dim3(256), 0, st, n, (double*)A, ld, sA);
  else if (mnp_is_c64(dtype))
    hipLaunchKernelGGL(k_identity<float>, grid, dim3(256), 0, st, n, (float*)A, ld, sA);
  else {
    mnp_set_error("identity: unknown dtype");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// mode 0: D = S1 - S2 ; mode 1: S2 <- Sigma (needs E)
extern "C" int mnpbem_sigma_batched(int dtype, int mode, int n, const void* S1, void* S2, void* D, const void* E,
                                    long ld, long stride, const double* nvec, const double* par, int gdiv,
                                    int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || batch <= 0) return 0;
  dim3 grid(mnp_ceil_div((long)n * n, 256), batch);
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_sigma<double>, grid, dim3(256), 0, st, mode, n, (const double*)S1, (double*)S2,
                       (double*)D, (const double*)E, ld, stride, nvec, par, gdiv);
  else if (mnp_is_c64(dtype))
    hipLaunchKernelGGL(k_sigma<float>, grid, dim3(256), 0, st, mode, n, (const float*)S1, (float*)S2, (float*)D,
                       (const float*)E, ld, stride, nvec, par, gdiv);
  else {
    mnp_set_error("sigma: unknown dtype");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// Direct solve X = A^-1 R for ``batch`` wavelengths (bem_ret.py:1694-1810).
// R, W, X: (batch, n, 8P) buffers of the factorisation's dtype, stride in
// complex elements; W is work space.
extern "C" int mnpbem_bem_solve_batched(int dtype, int n, int P, const void* G1t, const int* piv1, const void* G2t,
                                        const int* piv2, const void* S1t, const void* Et, const void* St,
                                        const int* pivS, long ldm, long smat, long spiv, const void* D1,
                                        const void* D2, const void* DS, long sD, const void* R, void* W, void* X,
                                        long ldv, long sv, const double* nvec, const double* par, int gdiv,
                                        int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldv >= 8 * P, "bem_solve: bad size (P <= 4)");
  if (batch <= 0) return 0;
  dim3 eg(mnp_ceil_div(n, 256), batch);
  // the factors of G1 and G2 as one stack (the solver's layout): both
  // triangular solves of the last step run as one batch of 2 batch systems
  const long esz = dtype == MNP_DTYPE_C128 ? 16 : 8;
  const bool stacked = (const char*)G2t == (const char*)G1t + (long)batch * smat * esz &&
                       piv2 == piv1 + (long)batch * spiv &&
                       ((D1 == nullptr && D2 == nullptr) ||
                        (D1 != nullptr && (const char*)D2 == (const char*)D1 + (long)batch * sD * esz));
#define MNP_SOLVE(T)                                                                                               \
  {                                                                                                                \
    const T* r = (const T*)R;                                                                                      \
    T* w = (T*)W;                                                                                                  \
    T* x = (T*)X;                                                                                                  \
    int rc;                                                                                                        \
    /* W[:,0:4P] = Sigma1 [phi | a] */                                                                             \
    if ((rc = gemv_t<T>(n, 4 * P, (const T*)S1t, ldm, smat, r, ldv, sv, w, ldv, sv, 1.0, 0.0, 0.0, batch, st)))   \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_step_mod<T>, eg, dim3(256), 0, st, n, P, r, w, ldv, sv, nvec, par, gdiv);                 \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* W[:,P:4P] = Delta^-1 alpha' */                                                                              \
    if ((rc = gemv_t<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * 5 * P, ldv, sv, w + 2 * P, ldv, sv, 1.0, 0.0, \
                        0.0, batch, st)))                                                                          \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_step_rhs<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par, gdiv);                    \
    MNP_LAUNCH_CHECK();                                                                                            \
    if ((rc = lu_solve_t<T>(n, P, (const T*)St, ldm, smat, pivS, spiv, w, ldv, sv, w + 2 * P, ldv, sv, batch,    \
                            st, (const T*)DS, sD)))                                                                \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_step_h2pre<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par, gdiv);                  \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* W[:,4P:7P] = Delta^-1 T */                                                                                  \
    if ((rc = gemv_t<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * P, ldv, sv, w + 2 * 4 * P, ldv, sv, 1.0, 0.0, \
                        0.0, batch, st)))                                                                          \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_step_pack<T>, eg, dim3(256), 0, st, n, P, r, w, x, ldv, sv);                              \
    MNP_LAUNCH_CHECK();                                                                                            \
    if (stacked) {                                                                                                 \
      /* G1 and G2 factors stacked: one solve over 2 batch systems */                                            \
      if ((rc = lu_solve_t<T>(n, 4 * P, (const T*)G1t, ldm, smat, piv1, spiv, x, ldv, sv, w, ldv, sv, 2 * batch,  \
                              st, (const T*)D1, sD, batch, 4 * P)))                                                \
        return rc;                                                                                                 \
      return 0;                                                                                                    \
    }                                                                                                              \
    if ((rc = lu_solve_t<T>(n, 4 * P, (const T*)G1t, ldm, smat, piv1, spiv, x, ldv, sv, w, ldv, sv, batch, st,    \
                            (const T*)D1, sD)))                                                                    \
      return rc;                                                                                                   \
    if ((rc = lu_solve_t<T>(n, 4 * P, (const T*)G2t, ldm, smat, piv2, spiv, x + 2 * 4 * P, ldv, sv,                \
                            w + 2 * 4 * P, ldv, sv, batch, st, (const T*)D2, sD)))                                 \
      return rc;                                                                                                   \
    return 0;                                                                                                      \
  }
  if (dtype == MNP_DTYPE_C128) MNP_SOLVE(double)
  if (mnp_is_c64(dtype)) MNP_SOLVE(float)
#undef MNP_SOLVE
  mnp_set_error("bem_solve: unknown dtype");
  return 1;
}

// The BEM solve of mnpbem_bem_solve_batched with explicit inverses in place
// of the LU factors of G1, G2 and Sigma (G1i, G2i, Si: the transposed
// buffers of G1^-1, G2^-1, Sigma^-1, as getri leaves them for the factors of
// the transposed buffers): every solve is a matrix-vector product at full
// bandwidth instead of a chain of triangular steps.  For callers that apply
// one factorisation many times with a few right-hand sides (a preconditioner
// inside GMRES), where the triangular steps of a few systems are
// latency-bound.
extern "C" int mnpbem_bem_solve_inv_batched(int dtype, int n, int P, const void* G1i, const void* G2i,
                                            const void* S1t, const void* Et, const void* Si, long ldm, long smat,
                                            const void* R, void* W, void* X, long ldv, long sv, const double* nvec,
                                            const double* par, int gdiv, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldv >= 8 * P, "bem_solve_inv: bad size (P <= 4)");
  if (batch <= 0) return 0;
  dim3 eg(mnp_ceil_div(n, 256), batch);
#define MNP_SOLVE_INV(T)                                                                                           \
  {                                                                                                                \
    const T* r = (const T*)R;                                                                                      \
    T* w = (T*)W;                                                                                                  \
    T* x = (T*)X;                                                                                                  \
    int rc;                                                                                                        \
    if ((rc = gemv_t<T>(n, 4 * P, (const T*)S1t, ldm, smat, r, ldv, sv, w, ldv, sv, 1.0, 0.0, 0.0, batch, st)))   \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_step_mod<T>, eg, dim3(256), 0, st, n, P, r, w, ldv, sv, nvec, par, gdiv);                 \
    MNP_LAUNCH_CHECK();                                                                                            \
    if ((rc = gemv_t<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * 5 * P, ldv, sv, w + 2 * P, ldv, sv, 1.0, 0.0, \
                        0.0, batch, st)))                                                                          \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_step_rhs<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par, gdiv);                    \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* Sigma^-1 on columns 0..P-1, through columns P..2P-1 (free here) */                                         \
    if ((rc = gemv_t<T>(n, P, (const T*)Si, ldm, smat, w, ldv, sv, w + 2 * P, ldv, sv, 1.0, 0.0, 0.0, batch, st))) \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_copy_cols<T>, dim3(mnp_ceil_div((long)n * P, 256), batch), dim3(256), 0, st, n, P,      \
                       (const T*)(w + 2 * P), w, ldv, sv);                                                         \
    MNP_LAUNCH_CHECK();                                                                                            \
    hipLaunchKernelGGL(k_step_h2pre<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par, gdiv);                  \
    MNP_LAUNCH_CHECK();                                                                                            \
    if ((rc = gemv_t<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * P, ldv, sv, w + 2 * 4 * P, ldv, sv, 1.0, 0.0, \
                        0.0, batch, st)))                                                                          \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_step_pack<T>, eg, dim3(256), 0, st, n, P, r, w, x, ldv, sv);                              \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* G1^-1, G2^-1 on the two halves of x, into w, then back into x */                                          \
    if ((rc = gemv_t<T>(n, 4 * P, (const T*)G1i, ldm, smat, x, ldv, sv, w, ldv, sv, 1.0, 0.0, 0.0, batch, st)))   \
      return rc;                                                                                                   \
    if ((rc = gemv_t<T>(n, 4 * P, (const T*)G2i, ldm, smat, x + 2 * 4 * P, ldv, sv, w + 2 * 4 * P, ldv, sv, 1.0,  \
                        0.0, 0.0, batch, st)))                                                                     \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_copy_cols<T>, dim3(mnp_ceil_div((long)n * 8 * P, 256), batch), dim3(256), 0, st, n,     \
                       8 * P, (const T*)w, x, ldv, sv);                                                            \
    MNP_LAUNCH_CHECK();                                                                                            \
    return 0;                                                                                                      \
  }
  if (dtype == MNP_DTYPE_C128) MNP_SOLVE_INV(double)
  if (mnp_is_c64(dtype)) MNP_SOLVE_INV(float)
#undef MNP_SOLVE_INV
  mnp_set_error("bem_solve_inv: unknown dtype");
  return 1;
}

// ---------------------------------------------------------------------------
// The same solve with its work space W in a static device array instead of a
// caller buffer, for callers that cannot allocate device memory (code under
// HIP graph capture that must not reach an allocator, serving loops,
// harnesses that forbid hipMalloc).  Each device holds its own copy of
// `scratch` (64 MiB) from the library's load until the process ends; its
// address is looked up once per device, with that device current, on the
// first call there, and only that address is handed to kernels.  Every call
// on one device shares the array: calls that overlap on one device (two
// streams, overlapping graph replays) race.  Nothing in it has to start at
// zero: every element the solve reads from W it has written first in the same
// call.
// ---------------------------------------------------------------------------
constexpr size_t kSolveScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kSolveScratchBytes];

namespace {
constexpr int kScratchDevices = 64;
std::once_flag g_scratch_once[kScratchDevices];
void* g_scratch_ptr[kScratchDevices];
hipError_t g_scratch_rc[kScratchDevices];

// device address of `scratch` on the current device
int scratch_address(void** out) {
  int dev = 0;
  MNP_HIP_CHECK(hipGetDevice(&dev));
  MNP_REQUIRE(dev >= 0 && dev < kScratchDevices, "bem_solve_noalloc: device ordinal above 63");
  std::call_once(g_scratch_once[dev], [dev]() {
    void* p = nullptr;
    g_scratch_rc[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_ptr[dev] = g_scratch_rc[dev] == hipSuccess ? p : nullptr;
  });
  if (g_scratch_rc[dev] != hipSuccess || g_scratch_ptr[dev] == nullptr) {
    mnp_set_error(std::string("bem_solve_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                  hipGetErrorString(g_scratch_rc[dev]));
    return 2;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" int mnpbem_bem_solve_batched_noalloc(int dtype, int n, int P, const void* G1t, const int* piv1,
                                                const void* G2t, const int* piv2, const void* S1t, const void* Et,
                                                const void* St, const int* pivS, long ldm, long smat, long spiv,
                                                const void* D1, const void* D2, const void* DS, long sD,
                                                const void* R, void* X, long ldv, long sv, const double* nvec,
                                                const double* par, int gdiv, int batch, void* stream) {
  MNP_REQUIRE(dtype == MNP_DTYPE_C128 || mnp_is_c64(dtype), "bem_solve_noalloc: unknown dtype");
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldv >= 8 * P && sv >= 0, "bem_solve_noalloc: bad size (P <= 4)");
  if (batch <= 0) return 0;
  // W takes R's layout: element (b, i, c) at b sv + i ldv + c
  const size_t esz = dtype == MNP_DTYPE_C128 ? 16 : 8;
  const size_t need = ((size_t)(batch - 1) * (size_t)sv + (size_t)(n - 1) * (size_t)ldv + 8 * (size_t)P) * esz;
  if (need > kSolveScratchBytes) {
    mnp_set_error("bem_solve_noalloc: work space of " + std::to_string(need) + " bytes (batch " +
                  std::to_string(batch) + ", n " + std::to_string(n) + ", P " + std::to_string(P) +
                  ") exceeds the static " + std::to_string(kSolveScratchBytes) +
                  "-byte scratch; split the batch or use mnpbem_bem_solve_batched");
    return 1;
  }
  void* w = nullptr;
  int rc = scratch_address(&w);
  if (rc) return rc;
  return mnpbem_bem_solve_batched(dtype, n, P, G1t, piv1, G2t, piv2, S1t, Et, St, pivS, ldm, smat, spiv, D1, D2,
                                  DS, sD, R, w, X, ldv, sv, nvec, par, gdiv, batch, stream);
}

extern "C" int mnpbem_bem_combine_batched(int n, int P, const void* Y, long ldy, long sy, const void* R, void* Out,
                                          long ldv, long sv, const double* nvec, const double* par, double* nrm,
                                          int gdiv, int batch, void* stream);

// Rout = R - A X (complex128), Y: (batch, n, 16P) work buffer; nrm: (batch, 2)
// receives ||R - A X||^2 and ||R||^2 (nrm may be NULL; with R NULL, Rout =
// A X).  G1t.. are complex128 transposed Green matrices (not factored).
extern "C" int mnpbem_bem_residual_batched(int n, int P, const void* G1t, const void* H1t, const void* G2t,
                                           const void* H2t, long ldm, long smat, const void* R, const void* X,
                                           void* Rout, long ldv, long sv, void* Y, long ldy, long sy,
                                           const double* nvec, const double* par, double* nrm, int gdiv, int batch,
                                           void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldy >= 16 * P, "bem_residual: bad size");
  if (batch <= 0) return 0;
  const double* x = (const double*)X;
  double* y = (double*)Y;
  const double *g1 = (const double*)G1t, *h1 = (const double*)H1t, *g2 = (const double*)G2t,
               *h2 = (const double*)H2t;
  dim3 grid(mnp_ceil_div(n, 256), batch, 4);
  switch (P) {
    case 1:
      hipLaunchKernelGGL(k_gemv4_t<4>, grid, dim3(256), 0, st, n, g1, h1, g2, h2, ldm, smat, x, ldv, sv, y, ldy, sy);
      break;
    case 2:
      hipLaunchKernelGGL(k_gemv4_t<8>, grid, dim3(256), 0, st, n, g1, h1, g2, h2, ldm, smat, x, ldv, sv, y, ldy, sy);
      break;
    case 3:
      hipLaunchKernelGGL(k_gemv4_t<12>, grid, dim3(256), 0, st, n, g1, h1, g2, h2, ldm, smat, x, ldv, sv, y, ldy,
                         sy);
      break;
    default:
      hipLaunchKernelGGL(k_gemv4_t<16>, grid, dim3(256), 0, st, n, g1, h1, g2, h2, ldm, smat, x, ldv, sv, y, ldy,
                         sy);
      break;
  }
  MNP_LAUNCH_CHECK();
  return mnpbem_bem_combine_batched(n, P, Y, ldy, sy, R, Rout, ldv, sv, nvec, par, nrm, gdiv, batch, stream);
}

// Combination step of the 8N operator from precomputed products
// Y = [G1 X1 | H1 X1 | G2 X2 | H2 X2] (batch, n, 16P) complex128:
// Out = R - A X when R is given (nrm, if given, receives ||R - A X||^2 and
// ||R||^2 per wavelength), Out = A X when R is NULL
// (mnpbem/bem/bem_ret_iter.py:1362-1501 after its four Green matvecs).
extern "C" int mnpbem_bem_combine_batched(int n, int P, const void* Y, long ldy, long sy, const void* R, void* Out,
                                          long ldv, long sv, const double* nvec, const double* par, double* nrm,
                                          int gdiv, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldy >= 16 * P && ldv >= 8 * P, "bem_combine: bad size");
  MNP_REQUIRE(nrm == nullptr || (R != nullptr && R != Out), "bem_combine: norms need R, distinct from Out");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_residual, dim3(mnp_ceil_div(n, 256), batch), dim3(256), 0, st, n, P, (const double*)R,
                     (const double*)Y, (double*)Out, ldv, sv, ldy, sy, nvec, par, gdiv);
  MNP_LAUNCH_CHECK();
  if (nrm != nullptr) {
    hipLaunchKernelGGL(k_norms2, dim3(batch), dim3(1024), 0, st, n, 8 * P, (const double*)R, (const double*)Out, ldv,
                       sv, nrm);
    MNP_LAUNCH_CHECK();
  }
  return 0;
}

// Right side of the permittivity sensitivity solve A x' = -(dA/deps1) x for
// the inside permittivity eps1 (k1 = k0 sqrt(eps1), dk1/deps1 = kap):
// with Y = [G1 X1 | H1 X1 | G1' X1 | H1' X1] (n x 16P, ' = d/dk1, X1 =
// [sig1 | h1]) and the operator of GdA&H Eqs. (10, 11, 14, 17),
//   dA x = explicit:  alpha -= ik0 n (G1 s1),  De += H1 s1 - ik0 n.(G1 h1)
//        + kap * (phi: G1' s1, a: G1' h1, alpha: H1' h1 - ik0 n eps1 G1' s1,
//                 De: eps1 H1' s1 - ik0 n.(eps1 G1' h1)),
// R = -dA x in the right-side layout [phi | a | De | alpha].
__global__ void k_deps_rhs(int n, int P, const double* Y, long ldy, long sy, double* R, long ldv, long sv,
                           const double* nvec, const double* par, const double* kap) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const cplx<double> kp = cmake(kap[2 * b], kap[2 * b + 1]);
  const double* y = Y + 2 * (b * sy + (long)i * ldy);
  double* r = R + 2 * (b * sv + (long)i * ldv);
  const int P4 = 4 * P;
  auto G = [&](int col) { return ldd(y + 2 * col); };
  auto H = [&](int col) { return ldd(y + 2 * (P4 + col)); };
  auto Gd = [&](int col) { return ldd(y + 2 * (2 * P4 + col)); };
  auto Hd = [&](int col) { return ldd(y + 2 * (3 * P4 + col)); };
  auto put = [&](int col, cplx<double> v) { std_(r + 2 * col, cmake(-v.re, -v.im)); };
  for (int p = 0; p < P; ++p) {
    put(p, cmul(kp, Gd(p)));
    cplx<double> ndg = cmake(0.0, 0.0), nddg = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) {
      const int hc = P + c * P + p;
      const double nc = nvec[3 * i + c];
      put(hc, cmul(kp, Gd(hc)));
      // alpha_c
      const cplx<double> ex = ik_times(q.k, cscale(G(p), -nc));
      const cplx<double> im = csub(Hd(hc), ik_times(q.k, cscale(cmul(q.e1, Gd(p)), nc)));
      put(4 * P + hc, cadd(ex, cmul(kp, im)));
      ndg = cadd(ndg, cscale(G(hc), nc));
      nddg = cadd(nddg, cscale(Gd(hc), nc));
    }
    const cplx<double> ex = csub(H(p), ik_times(q.k, ndg));
    const cplx<double> im = csub(cmul(q.e1, Hd(p)), ik_times(q.k, cmul(q.e1, nddg)));
    put(4 * P + p, cadd(ex, cmul(kp, im)));
  }
}

extern "C" int mnpbem_deps_rhs(int n, int P, const void* Y, long ldy, long sy, void* R, long ldv, long sv,
                               const double* nvec, const double* par, const double* kap, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && ldy >= 16 * P && ldv >= 8 * P, "deps_rhs: bad size");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_deps_rhs, dim3(mnp_ceil_div(n, 256), batch), dim3(256), 0, st, n, P, (const double*)Y, ldy,
                     sy, (double*)R, ldv, sv, nvec, par, kap);
  MNP_LAUNCH_CHECK();
  return 0;
}

// dst(double) += src (float or double), count real values
extern "C" int mnpbem_accumulate(int src_dtype, long count, const void* src, double* dst, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (count <= 0) return 0;
  dim3 grid(mnp_ceil_div(count, 256));
  if (mnp_is_c64(src_dtype))
    hipLaunchKernelGGL(k_axpy_cast<float>, grid, dim3(256), 0, st, count, (const float*)src, dst);
  else
    hipLaunchKernelGGL(k_axpy_cast<double>, grid, dim3(256), 0, st, count, (const double*)src, dst);
  MNP_LAUNCH_CHECK();
  return 0;
}

// dst = (dst type) src(double), count real values
extern "C" int mnpbem_convert(int dst_dtype, long count, const double* src, void* dst, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (count <= 0) return 0;
  dim3 grid(mnp_ceil_div(count, 256));
  if (mnp_is_c64(dst_dtype))
    hipLaunchKernelGGL(k_copy_cast<float>, grid, dim3(256), 0, st, count, src, (float*)dst);
  else
    hipLaunchKernelGGL(k_copy_cast<double>, grid, dim3(256), 0, st, count, src, (double*)dst);
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_planewave_batched(int n, int P, const double* pos, const double* nvec, const int* in_med,
                                        const int* out_med, const double* pol, const double* dir, const double* pw,
                                        double* R, long ldv, long sv, double* pot, long spot, int gdiv, int batch,
                                        void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && ldv >= 8 * P, "planewave: bad size");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_planewave, dim3(mnp_ceil_div(n, 256), batch), dim3(256), 0, st, n, P, pos, nvec, in_med,
                     out_med, pol, dir, pw, R, ldv, sv, pot, spot, gdiv);
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_extinction_batched(int n, int P, const double* X, long ldv, long sv, const double* pos,
                                         const double* area, const int* in_med, const int* out_med,
                                         const double* pol, const double* dir, const double* pw, double* ext,
                                         int gdiv, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0, "extinction: bad size");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_extinction, dim3(P, batch), dim3(256), 0, st, n, P, X, ldv, sv, pos, area, in_med, out_med,
                     pol, dir, pw, ext, gdiv);
  MNP_LAUNCH_CHECK();
  return 0;
}

// Scattering cross section of the plane wave (planewave_ret.py:391-435 over
// spectrum_ret.py:395-459): dirs / w are the far-field sphere's directions and
// weights (ndir); dsca: (batch, ndir, P) doubles (differential, normalised by
// the incoming 0.5 nb); sca: (batch, P); E, H: NULL or (batch, ndir, 3, P)
// complex128 far fields.  P <= 8.
extern "C" int mnpbem_scattering_batched(int n, int P, const double* X, long ldv, long sv, const double* pos,
                                         const double* area, const int* in_med, const int* out_med, int ndir,
                                         const double* dirs, const double* w, const double* pw, double* dsca,
                                         double* sca, double* E, double* H, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P <= 8 && ndir > 0 && ndir <= 2147483647, "scattering: bad size");
  MNP_REQUIRE((E == nullptr) == (H == nullptr), "scattering: give both E and H or neither");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_farfield, dim3(ndir, batch), dim3(256), 0, st, n, P, X, ldv, sv, pos, area, in_med, out_med,
                     dirs, pw, dsca, E, H);
  MNP_LAUNCH_CHECK();
  const int total = batch * P;
  hipLaunchKernelGGL(k_sca_reduce, dim3(mnp_ceil_div(total, 256)), dim3(256), 0, st, ndir, P, total, w, dsca, sca);
  MNP_LAUNCH_CHECK();
  return 0;
}

// Inverses of the 64 x 64 diagonal blocks of LU factors from
// mnpbem_getrf_batched, for mnpbem_lu_solve_t_batched / mnpbem_bem_solve_batched:
// Dinv[b] holds ceil(n/64) blocks of [U_II^-1 | L_II^-1] (2 x 64 x 64 each);
// sD >= ceil(n/64) * 2 * 64 * 64 complex elements.
extern "C" int mnpbem_lu_blockinv_batched(int dtype, int n, const void* LU, long ld, long sLU, void* Dinv, long sD,
                                          int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || batch <= 0) return 0;
  const int nblk = (n + 63) / 64;
  MNP_REQUIRE(sD >= (long)nblk * 2 * 64 * 64, "lu_blockinv: Dinv stride too small");
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_lu_blockinv<double>, dim3(nblk, batch), dim3(64), 0, st, n, (const double*)LU, ld, sLU,
                       (double*)Dinv, sD);
  else if (mnp_is_c64(dtype))
    hipLaunchKernelGGL(k_lu_blockinv<float>, dim3(nblk, batch), dim3(64), 0, st, n, (const float*)LU, ld, sLU,
                       (float*)Dinv, sD);
  else {
    mnp_set_error("lu_blockinv: unknown dtype");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------
// Adjoint (transposed) solve A^T y = c with the factors of the direct solve.
//
// From the operator of GdA&H Eqs. (10, 11, 14, 17) (see k_residual),
//   (A^T y)_s1 =  G1^T (y_phi - ik e1 n.y_al) + e1 H1^T y_De
//   (A^T y)_h1 =  G1^T (y_a - ik e1 n y_De)   +    H1^T y_al
//   (A^T y)_s2 = -G2^T (y_phi - ik e2 n.y_al) - e2 H2^T y_De
//   (A^T y)_h2 = -G2^T (y_a - ik e2 n y_De)   -    H2^T y_al
// Multiplying by G^-T and eliminating as in the forward solve
// (bem_ret.py:1694-1810 transposed) gives, with g1 = G1^-T c1, g2 = G2^-T c2,
// gs = g1_s + g2_s, gh = g1_h + g2_h and L = e1 - e2:
//   y_De = Sigma^-T (gs + ik L n.(Delta^-T gh))
//   y_al = Delta^-T (gh + ik L n y_De)
//   y_phi = g1_s + ik e1 n.y_al - e1 (Sigma1^T y_De)
//   y_a   = g1_h + ik e1 n y_De - Sigma1^T y_al
// The transposed products are plain row-major GEMVs over the stored buffers
// (buf = M^T), and M^-T v is the ordinary LU solve with the stored LU of M^T.
// y is returned in the right-side layout [phi | a | De | alpha], c is given in
// the unknowns' layout [sig1 | h1 | sig2 | h2].
// ---------------------------------------------------------------------------
namespace {

// Y[b] = beta Y[b] + buf[b] X[b], y_i = sum_j buf[i*ld + j] x_j (the product
// with M^T for the stored buffer of M).  One wavefront per row with the lanes
// along the row (coalesced); X chunks of 64 rows staged column-major in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_gemv_n(int n, int C, const T* __restrict__ M, long ldm, long sM,
                                                const T* __restrict__ X, long ldx, long sX, T* __restrict__ Y,
                                                long ldy, long sY, double beta) {
  const long b = blockIdx.y;
  M += 2 * b * sM;
  X += 2 * b * sX;
  Y += 2 * b * sY;
  const int tid = threadIdx.x, lane = tid & 63;
  const int i = blockIdx.x * 4 + (tid >> 6);
  __shared__ cplx<T> xs[MAXC][64];
  cplx<T> acc[MAXC];
#pragma unroll
  for (int c = 0; c < MAXC; ++c) acc[c] = cmake<T>(0, 0);
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int nj = min(64, n - j0);
    __syncthreads();
    for (int e = tid; e < nj * C; e += 256) {
      const int jj = e / C, c = e % C;
      xs[c][jj] = cld(X + 2 * ((long)(j0 + jj) * ldx + c));
    }
    __syncthreads();
    if (i < n && lane < nj) {
      const cplx<T> m = cld(M + 2 * ((long)i * ldm + j0 + lane));
#pragma unroll
      for (int c = 0; c < MAXC; ++c)
        if (c < C) acc[c] = cfma(acc[c], m, xs[c][lane]);
    }
  }
#pragma unroll
  for (int c = 0; c < MAXC; ++c)
    if (c < C) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        acc[c].re += __shfl_xor(acc[c].re, off);
        acc[c].im += __shfl_xor(acc[c].im, off);
      }
    }
  if (lane == 0 && i < n) {
#pragma unroll
    for (int c = 0; c < MAXC; ++c)
      if (c < C) {
        T* py = Y + 2 * ((long)i * ldy + c);
        double yr = (double)acc[c].re, yi = (double)acc[c].im;
        if (beta != 0.0) {
          const cplx<T> y0 = cld(py);
          yr += beta * (double)y0.re;
          yi += beta * (double)y0.im;
        }
        cst(py, cmake<T>((T)yr, (T)yi));
      }
  }
}

// One 64-row block step of the solve B x = v with the row-major LU of the
// buffer B (P B = L U, x = U^-1 L^-1 P v; P applied beforehand by k_perm_n).
//   FWD (L z = v, unit lower): block rhs in V; z_I -> Z; V rows below -= L_KI z_I.
//   BWD (U x = z, upper): block rhs in Z; x_I -> V; Z rows above -= U_KI x_I.
// Dinv, when given, holds the inverse diagonal blocks [U^-1 | L^-1].
template <typename T, bool FWD>
__global__ __launch_bounds__(256) void k_trsv_n_step(int n, int C, int i0, const T* __restrict__ LU, long ld,
                                                     long sLU, T* V, long ldv, long sV, T* Z, long ldz, long sZ,
                                                     const T* __restrict__ Dinv, long sD) {
  const long b = blockIdx.y;
  LU += 2 * b * sLU;
  V += 2 * b * sV;
  Z += 2 * b * sZ;
  const int tid = threadIdx.x, lane = tid & 63;
  const int nb = min(64, n - i0), i1 = i0 + nb;
  __shared__ cplx<T> blk[64][65];
  __shared__ cplx<T> sol[64][MAXC];
  const T* src = FWD ? V : Z;
  const long lds = FWD ? ldv : ldz;
  if (Dinv) {
    const T* D = Dinv + 2 * (b * sD + ((long)(i0 / 64) * 2 + (FWD ? 1 : 0)) * 64 * 64);
    for (int e = tid; e < 64 * C; e += 256) {
      const int r = e / C, c = e % C;
      blk[r][c] = r < nb ? cld(src + 2 * ((long)(i0 + r) * lds + c)) : cmake<T>(0, 0);
    }
    __syncthreads();
    // thread (row r, column group cg): row r of the inverse block
    const int r = tid & 63, cg = tid >> 6;
    cplx<T> drow[64];
#pragma unroll
    for (int q = 0; q < 64; ++q) drow[q] = (q < nb && r < nb) ? cld(D + 2 * ((long)r * 64 + q)) : cmake<T>(0, 0);
    for (int c = cg; c < C; c += 4) {
      cplx<T> acc = cmake<T>(0, 0);
#pragma unroll
      for (int q = 0; q < 64; ++q) acc = cfma(acc, drow[q], blk[q][c]);
      sol[r][c] = acc;
    }
    __syncthreads();
  } else {
    for (int e = tid; e < nb * nb; e += 256) {
      const int r = e / nb, q = e % nb;
      blk[r][q] = cld(LU + 2 * ((long)(i0 + r) * ld + i0 + q));
    }
    __syncthreads();
    if (tid < 64) {
      cplx<T> val[MAXC];
#pragma unroll
      for (int c = 0; c < MAXC; ++c)
        val[c] = (c < C && lane < nb) ? cld(src + 2 * ((long)(i0 + lane) * lds + c)) : cmake<T>(0, 0);
      if (FWD) {
        for (int i = 0; i < nb; ++i) {
          const cplx<T> l = blk[lane < nb ? lane : 0][i];  // L[r][i]
#pragma unroll
          for (int c = 0; c < MAXC; ++c)
            if (c < C) {
              cplx<T> zi;
              zi.re = __shfl(val[c].re, i);
              zi.im = __shfl(val[c].im, i);
              if (lane > i && lane < nb) val[c] = cfms(val[c], l, zi);
            }
        }
      } else {
        for (int i = nb - 1; i >= 0; --i) {
          if (lane == i) {
            const cplx<T> inv = crecip(blk[i][i]);
#pragma unroll
            for (int c = 0; c < MAXC; ++c) val[c] = cmul(val[c], inv);
          }
          const cplx<T> u = blk[lane < nb ? lane : 0][i];  // U[r][i], r < i
#pragma unroll
          for (int c = 0; c < MAXC; ++c)
            if (c < C) {
              cplx<T> xi;
              xi.re = __shfl(val[c].re, i);
              xi.im = __shfl(val[c].im, i);
              if (lane < i) val[c] = cfms(val[c], u, xi);
            }
        }
      }
#pragma unroll
      for (int c = 0; c < MAXC; ++c) sol[lane][c] = val[c];
    }
    __syncthreads();
  }
  if (blockIdx.x == 0) {
    T* dst = FWD ? Z : V;
    const long ldd_ = FWD ? ldz : ldv;
    for (int e = tid; e < nb * C; e += 256) {
      const int rr = e / C, c = e % C;
      cst(dst + 2 * ((long)(i0 + rr) * ldd_ + c), sol[rr][c]);
    }
  }
  // tile update: FWD rows k >= i1, BWD rows k < i0; row k reads its own
  // 64-entry segment LU[k][i0:i1]
  const int k = (FWD ? i1 : 0) + blockIdx.x * 256 + tid;
  const int kend = FWD ? n : i0;
  if (k < kend) {
    T* upd = FWD ? V : Z;
    const long ldu = FWD ? ldv : ldz;
    const T* row = LU + 2 * ((long)k * ld + i0);
    cplx<T> acc[MAXC];
#pragma unroll
    for (int c = 0; c < MAXC; ++c) acc[c] = cmake<T>(0, 0);
    int l = 0;
    for (; l + 16 <= nb; l += 16) {
      cplx<T> m[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) m[u] = cld(row + 2 * (l + u));
#pragma unroll
      for (int u = 0; u < 16; ++u)
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
          if (c < C) acc[c] = cfma(acc[c], m[u], sol[l + u][c]);
    }
    for (; l < nb; ++l) {
      const cplx<T> m = cld(row + 2 * l);
#pragma unroll
      for (int c = 0; c < MAXC; ++c)
        if (c < C) acc[c] = cfma(acc[c], m, sol[l][c]);
    }
#pragma unroll
    for (int c = 0; c < MAXC; ++c)
      if (c < C) {
        T* pu = upd + 2 * ((long)k * ldu + c);
        cst(pu, csub(cld(pu), acc[c]));
      }
  }
}

// v = P v: row interchanges ipiv[0], ..., ipiv[n-1] in order; one workgroup
// per matrix, one thread per column.
template <typename T>
__global__ __launch_bounds__(256) void k_perm_n(int n, int C, const int* __restrict__ ipiv, long sP, T* V, long ldv,
                                                long sV) {
  const long b = blockIdx.x;
  ipiv += b * sP;
  V += 2 * b * sV;
  const int tid = threadIdx.x;
  __shared__ int sp[256];
  for (int lo = 0; lo < n; lo += 256) {
    const int hi = min(n, lo + 256);
    __syncthreads();
    int moved = 0;
    if (tid < hi - lo) {
      sp[tid] = ipiv[lo + tid];
      moved = sp[tid] != lo + tid;
    }
    // chunks without interchanges are skipped (as in k_perm_t)
    if (!__syncthreads_or(moved)) continue;
    if (tid < C) {
      for (int i = lo; i < hi; ++i) {
        const int p = sp[i - lo];
        if (p != i) {
          T* x = V + 2 * ((long)i * ldv + tid);
          T* y = V + 2 * ((long)p * ldv + tid);
          const cplx<T> u = cld(x), v = cld(y);
          cst(x, v);
          cst(y, u);
        }
      }
    }
  }
}

template <typename T>
int gemv_n(int n, int C, const T* M, long ldm, long sM, const T* X, long ldx, long sX, T* Y, long ldy, long sY,
           double beta, int batch, hipStream_t st) {
  for (int c0 = 0; c0 < C; c0 += MAXC) {
    const int cc = std::min(MAXC, C - c0);
    hipLaunchKernelGGL(k_gemv_n<T>, dim3(mnp_ceil_div(n, 4), batch), dim3(256), 0, st, n, cc, M, ldm, sM,
                       X + 2 * c0, ldx, sX, Y + 2 * c0, ldy, sY, beta);
    MNP_LAUNCH_CHECK();
  }
  return 0;
}

// V <- M^-T V (n x C) with the LU of M^T; Z (n x C) is work space.
template <typename T>
int lu_solve_n(int n, int C, const T* LU, long ld, long sLU, const int* ipiv, long sP, T* V, long ldv, long sV,
               T* Z, long ldz, long sZ, int batch, hipStream_t st, const T* Dinv = nullptr, long sD = 0) {
  const int nblk = (n + 63) / 64;
  for (int c0 = 0; c0 < C; c0 += MAXC) {
    const int cc = std::min(MAXC, C - c0);
    T* v = V + 2 * c0;
    T* z = Z + 2 * c0;
    hipLaunchKernelGGL(k_perm_n<T>, dim3(batch), dim3(256), 0, st, n, cc, ipiv, sP, v, ldv, sV);
    MNP_LAUNCH_CHECK();
    for (int ib = 0; ib < nblk; ++ib) {
      const int i0 = ib * 64, i1 = std::min(n, i0 + 64);
      dim3 grid(std::max(1, mnp_ceil_div(n - i1, 256)), batch);
      hipLaunchKernelGGL((k_trsv_n_step<T, true>), grid, dim3(256), 0, st, n, cc, i0, LU, ld, sLU, v, ldv, sV, z,
                         ldz, sZ, Dinv, sD);
      MNP_LAUNCH_CHECK();
    }
    for (int ib = nblk - 1; ib >= 0; --ib) {
      const int i0 = ib * 64;
      dim3 grid(std::max(1, mnp_ceil_div(i0, 256)), batch);
      hipLaunchKernelGGL((k_trsv_n_step<T, false>), grid, dim3(256), 0, st, n, cc, i0, LU, ld, sLU, v, ldv, sV, z,
                         ldz, sZ, Dinv, sD);
      MNP_LAUNCH_CHECK();
    }
  }
  return 0;
}

// Y = C (n x cols per wavelength, same layout)
template <typename T>
__global__ void k_adj_copy(int n, int cols, const T* C, T* Y, long ld, long stride) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const long o = 2 * (b * stride + (long)i * ld);
  for (int c = 0; c < cols; ++c) cst(Y + o + 2 * c, cld(C + o + 2 * c));
}

// a2: W[:, 0:P] = g1_s + g2_s, W[:, P:4P] = g1_h + g2_h  (Y holds [g1 | g2])
template <typename T>
__global__ void k_adj_sum(int n, int P, const T* Y, T* W, long ld, long stride) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T* y = Y + 2 * (b * stride + (long)i * ld);
  T* w = W + 2 * (b * stride + (long)i * ld);
  for (int c = 0; c < 4 * P; ++c) cst(w + 2 * c, cadd(cld(y + 2 * c), cld(y + 2 * (4 * P + c))));
}

// a4: W[:, 4P:5P] = gs + ik L n.u,  u = W[:, 5P:8P]
template <typename T>
__global__ void k_adj_rhs(int n, int P, T* W, long ld, long stride, const double* nvec, const double* par) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const cplx<double> L = csub(q.e1, q.e2);
  T* w = W + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    cplx<double> s = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) s = cadd(s, cscale(ldd(w + 2 * (5 * P + c * P + p)), nvec[3 * i + c]));
    std_(w + 2 * (4 * P + p), cadd(ldd(w + 2 * p), ik_times(q.k, cmul(L, s))));
  }
}

// a6: W[:, P:4P] = gh + ik L n y_De,  y_De = W[:, 4P:5P]
template <typename T>
__global__ void k_adj_t(int n, int P, T* W, long ld, long stride, const double* nvec, const double* par) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const cplx<double> L = csub(q.e1, q.e2);
  T* w = W + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    const cplx<double> ly = cmul(L, ldd(w + 2 * (4 * P + p)));
    for (int c = 0; c < 3; ++c) {
      T* o = w + 2 * (P + c * P + p);
      std_(o, cadd(ldd(o), ik_times(q.k, cscale(ly, nvec[3 * i + c]))));
    }
  }
}

// a8: Y = [g1_s + ik e1 n.y_al - e1 w_De | g1_h + ik e1 n y_De - w_al | y_De | y_al]
// with W = [w_De | w_al | y_De | y_al]
template <typename T>
__global__ void k_adj_final(int n, int P, T* Y, const T* W, long ld, long stride, const double* nvec,
                            const double* par) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  T* y = Y + 2 * (b * stride + (long)i * ld);
  const T* w = W + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    const cplx<double> yde = ldd(w + 2 * (4 * P + p));
    cplx<double> nya = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) {
      const double nc = nvec[3 * i + c];
      const cplx<double> yal = ldd(w + 2 * (5 * P + c * P + p));
      nya = cadd(nya, cscale(yal, nc));
      T* o = y + 2 * (P + c * P + p);
      std_(o, csub(cadd(ldd(o), ik_times(q.k, cmul(q.e1, cscale(yde, nc)))), ldd(w + 2 * (P + c * P + p))));
      std_(y + 2 * (5 * P + c * P + p), yal);
    }
    std_(y + 2 * p, csub(cadd(ldd(y + 2 * p), ik_times(q.k, cmul(q.e1, nya))), cmul(q.e1, ldd(w + 2 * p))));
    std_(y + 2 * (4 * P + p), yde);
  }
}

// Z = [y_phi - ik e1 n.y_al | y_a - ik e1 n y_De | y_phi - ik e2 n.y_al | y_a - ik e2 n y_De]
__global__ void k_adj_z(int n, int P, const double* Yv, double* Z, long ld, long stride, const double* nvec,
                        const double* par) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const double* y = Yv + 2 * (b * stride + (long)i * ld);
  double* z = Z + 2 * (b * stride + (long)i * ld);
  for (int p = 0; p < P; ++p) {
    const cplx<double> yphi = ldd(y + 2 * p), yde = ldd(y + 2 * (4 * P + p));
    cplx<double> nya = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) nya = cadd(nya, cscale(ldd(y + 2 * (5 * P + c * P + p)), nvec[3 * i + c]));
    for (int s = 0; s < 2; ++s) {
      const cplx<double> e = s == 0 ? q.e1 : q.e2;
      const int off = 4 * P * s;
      std_(z + 2 * (off + p), csub(yphi, ik_times(q.k, cmul(e, nya))));
      for (int c = 0; c < 3; ++c)
        std_(z + 2 * (off + P + c * P + p),
             csub(ldd(y + 2 * (P + c * P + p)), ik_times(q.k, cmul(e, cscale(yde, nvec[3 * i + c])))));
    }
  }
}

// T = [G1^T Z1 | H1^T Wy | G2^T Z2 | H2^T Wy] (n x 16P), Wy = [y_De | y_al];
// Out = C - A^T y (norms as in k_residual), or A^T y when C is null.
__global__ void k_adj_combine(int n, int P, const double* C, const double* T, double* Out, long ldv, long sv,
                              long ldt, long stt, const double* par) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const Par q = load_par(par, b);
    const double* t = T + 2 * (b * stt + (long)i * ldt);
    const double* c = C ? C + 2 * (b * sv + (long)i * ldv) : nullptr;
    double* o = Out + 2 * (b * sv + (long)i * ldv);
    const int P4 = 4 * P;
    auto put = [&](int col, cplx<double> v) {
      if (c == nullptr) {
        std_(o + 2 * col, v);
        return;
      }
      std_(o + 2 * col, csub(ldd(c + 2 * col), v));
    };
    for (int col = 0; col < P4; ++col) {
      const cplx<double> g1 = ldd(t + 2 * col), h1 = ldd(t + 2 * (P4 + col));
      const cplx<double> g2 = ldd(t + 2 * (2 * P4 + col)), h2 = ldd(t + 2 * (3 * P4 + col));
      if (col < P) {
        put(col, cadd(g1, cmul(q.e1, h1)));
        put(P4 + col, cmake(-(g2.re + cmul(q.e2, h2).re), -(g2.im + cmul(q.e2, h2).im)));
      } else {
        put(col, cadd(g1, h1));
        put(P4 + col, cmake(-(g2.re + h2.re), -(g2.im + h2.im)));
      }
    }
  }
}

}  // namespace

// Adjoint solve Y = A^-T C for ``batch`` wavelengths with the factors of
// mnpbem_bem_solve_batched (same arguments; C in the unknowns' layout, Y in
// the right-side layout).  W is work space of the shape of C.
extern "C" int mnpbem_bem_solve_adjoint_batched(int dtype, int n, int P, const void* G1t, const int* piv1,
                                                const void* G2t, const int* piv2, const void* S1t, const void* Et,
                                                const void* St, const int* pivS, long ldm, long smat, long spiv,
                                                const void* D1, const void* D2, const void* DS, long sD,
                                                const void* C, void* W, void* Y, long ldv, long sv,
                                                const double* nvec, const double* par, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldv >= 8 * P, "bem_solve_adjoint: bad size (P <= 4)");
  if (batch <= 0) return 0;
  dim3 eg(mnp_ceil_div(n, 256), batch);
#define MNP_ADJ(T)                                                                                                 \
  {                                                                                                                \
    T* y = (T*)Y;                                                                                                  \
    T* w = (T*)W;                                                                                                  \
    int rc;                                                                                                        \
    hipLaunchKernelGGL(k_adj_copy<T>, eg, dim3(256), 0, st, n, 8 * P, (const T*)C, y, ldv, sv);                    \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* a1: Y = [G1^-T c1 | G2^-T c2] */                                                                            \
    if ((rc = lu_solve_n<T>(n, 4 * P, (const T*)G1t, ldm, smat, piv1, spiv, y, ldv, sv, w, ldv, sv, batch, st,    \
                            (const T*)D1, sD)))                                                                    \
      return rc;                                                                                                   \
    if ((rc = lu_solve_n<T>(n, 4 * P, (const T*)G2t, ldm, smat, piv2, spiv, y + 2 * 4 * P, ldv, sv,                \
                            w + 2 * 4 * P, ldv, sv, batch, st, (const T*)D2, sD)))                                 \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_adj_sum<T>, eg, dim3(256), 0, st, n, P, y, w, ldv, sv);                                   \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* a3: u = Delta^-T gh -> W[:, 5P:8P] */                                                                       \
    if ((rc = gemv_n<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * P, ldv, sv, w + 2 * 5 * P, ldv, sv, 0.0,      \
                        batch, st)))                                                                               \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_adj_rhs<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par);                           \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* a5: y_De = Sigma^-T rhs in W[:, 4P:5P] */                                                                   \
    if ((rc = lu_solve_n<T>(n, P, (const T*)St, ldm, smat, pivS, spiv, w + 2 * 4 * P, ldv, sv, w, ldv, sv, batch, \
                            st, (const T*)DS, sD)))                                                                \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_adj_t<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par);                             \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* a6: y_al = Delta^-T t -> W[:, 5P:8P] */                                                                     \
    if ((rc = gemv_n<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * P, ldv, sv, w + 2 * 5 * P, ldv, sv, 0.0,      \
                        batch, st)))                                                                               \
      return rc;                                                                                                   \
    /* a7: W[:, 0:4P] = Sigma1^T [y_De | y_al] */                                                                  \
    if ((rc = gemv_n<T>(n, 4 * P, (const T*)S1t, ldm, smat, w + 2 * 4 * P, ldv, sv, w, ldv, sv, 0.0, batch, st))) \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_adj_final<T>, eg, dim3(256), 0, st, n, P, y, w, ldv, sv, nvec, par);                      \
    MNP_LAUNCH_CHECK();                                                                                            \
    return 0;                                                                                                      \
  }
  if (dtype == MNP_DTYPE_C128) MNP_ADJ(double)
  if (mnp_is_c64(dtype)) MNP_ADJ(float)
#undef MNP_ADJ
  mnp_set_error("bem_solve_adjoint: unknown dtype");
  return 1;
}

// The adjoint solve with explicit inverses (the buffers that getri leaves in
// place of the factors of mnpbem_bem_solve_adjoint_batched: G1i, G2i, Si;
// S1t and Et as there): each triangular pair becomes one matrix product in
// the same orientation.  C and Y must not overlap; W is work space of the
// shape of C.
extern "C" int mnpbem_bem_solve_adjoint_inv_batched(int dtype, int n, int P, const void* G1i, const void* G2i,
                                                    const void* S1t, const void* Et, const void* Si, long ldm,
                                                    long smat, const void* C, void* W, void* Y, long ldv, long sv,
                                                    const double* nvec, const double* par, int batch,
                                                    void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldv >= 8 * P, "bem_solve_adjoint_inv: bad size (P <= 4)");
  MNP_REQUIRE(C != Y, "bem_solve_adjoint_inv: C and Y must be different buffers");
  if (batch <= 0) return 0;
  dim3 eg(mnp_ceil_div(n, 256), batch);
#define MNP_ADJ_INV(T)                                                                                             \
  {                                                                                                                \
    const T* c = (const T*)C;                                                                                      \
    T* y = (T*)Y;                                                                                                  \
    T* w = (T*)W;                                                                                                  \
    int rc;                                                                                                        \
    /* a1: Y = [G1^-T c1 | G2^-T c2] */                                                                            \
    if ((rc = gemv_n<T>(n, 4 * P, (const T*)G1i, ldm, smat, c, ldv, sv, y, ldv, sv, 0.0, batch, st))) return rc;  \
    if ((rc = gemv_n<T>(n, 4 * P, (const T*)G2i, ldm, smat, c + 2 * 4 * P, ldv, sv, y + 2 * 4 * P, ldv, sv, 0.0,   \
                        batch, st)))                                                                               \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_adj_sum<T>, eg, dim3(256), 0, st, n, P, y, w, ldv, sv);                                   \
    MNP_LAUNCH_CHECK();                                                                                            \
    if ((rc = gemv_n<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * P, ldv, sv, w + 2 * 5 * P, ldv, sv, 0.0,      \
                        batch, st)))                                                                               \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_adj_rhs<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par);                           \
    MNP_LAUNCH_CHECK();                                                                                            \
    /* a5: Sigma^-T rhs through W[:, 0:P] (free here), back into W[:, 4P:5P] */                                   \
    if ((rc = gemv_n<T>(n, P, (const T*)Si, ldm, smat, w + 2 * 4 * P, ldv, sv, w, ldv, sv, 0.0, batch, st)))      \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_copy_cols<T>, dim3(mnp_ceil_div((long)n * P, 256), batch), dim3(256), 0, st, n, P,      \
                       (const T*)w, w + 2 * 4 * P, ldv, sv);                                                       \
    MNP_LAUNCH_CHECK();                                                                                            \
    hipLaunchKernelGGL(k_adj_t<T>, eg, dim3(256), 0, st, n, P, w, ldv, sv, nvec, par);                             \
    MNP_LAUNCH_CHECK();                                                                                            \
    if ((rc = gemv_n<T>(n, 3 * P, (const T*)Et, ldm, smat, w + 2 * P, ldv, sv, w + 2 * 5 * P, ldv, sv, 0.0,      \
                        batch, st)))                                                                               \
      return rc;                                                                                                   \
    if ((rc = gemv_n<T>(n, 4 * P, (const T*)S1t, ldm, smat, w + 2 * 4 * P, ldv, sv, w, ldv, sv, 0.0, batch, st))) \
      return rc;                                                                                                   \
    hipLaunchKernelGGL(k_adj_final<T>, eg, dim3(256), 0, st, n, P, y, w, ldv, sv, nvec, par);                      \
    MNP_LAUNCH_CHECK();                                                                                            \
    return 0;                                                                                                      \
  }
  if (dtype == MNP_DTYPE_C128) MNP_ADJ_INV(double)
  if (mnp_is_c64(dtype)) MNP_ADJ_INV(float)
#undef MNP_ADJ_INV
  mnp_set_error("bem_solve_adjoint_inv: unknown dtype");
  return 1;
}

// Out = C - A^T Y in complex128 (nrm: ||C - A^T Y||^2 and ||C||^2 per
// wavelength, may be NULL; with C NULL, Out = A^T Y).  G1t.. are the
// complex128 transposed Green matrices; Z (batch, n, 8P) and T (batch, n,
// 16P) are work space.
extern "C" int mnpbem_bem_residual_adjoint_batched(int n, int P, const void* G1t, const void* H1t, const void* G2t,
                                                   const void* H2t, long ldm, long smat, const void* C,
                                                   const void* Yv, void* Out, long ldv, long sv, void* Z, void* T,
                                                   long ldt, long stt, const double* nvec, const double* par,
                                                   double* nrm, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P * 4 <= MAXC && ldt >= 16 * P && ldv >= 8 * P, "bem_residual_adjoint: bad size");
  if (batch <= 0) return 0;
  MNP_REQUIRE(nrm == nullptr || (C != nullptr && C != Out), "bem_residual_adjoint: norms need C, distinct from Out");
  const double* y = (const double*)Yv;
  double* z = (double*)Z;
  double* t = (double*)T;
  const int P4 = 4 * P;
  dim3 eg(mnp_ceil_div(n, 256), batch);
  hipLaunchKernelGGL(k_adj_z, eg, dim3(256), 0, st, n, P, y, z, ldv, sv, nvec, par);
  MNP_LAUNCH_CHECK();
  int rc;
  if ((rc = gemv_n<double>(n, P4, (const double*)G1t, ldm, smat, z, ldv, sv, t, ldt, stt, 0.0, batch, st)))
    return rc;
  if ((rc = gemv_n<double>(n, P4, (const double*)H1t, ldm, smat, y + 2 * P4, ldv, sv, t + 2 * P4, ldt, stt, 0.0,
                           batch, st)))
    return rc;
  if ((rc = gemv_n<double>(n, P4, (const double*)G2t, ldm, smat, z + 2 * P4, ldv, sv, t + 2 * 2 * P4, ldt, stt,
                           0.0, batch, st)))
    return rc;
  if ((rc = gemv_n<double>(n, P4, (const double*)H2t, ldm, smat, y + 2 * P4, ldv, sv, t + 2 * 3 * P4, ldt, stt,
                           0.0, batch, st)))
    return rc;
  hipLaunchKernelGGL(k_adj_combine, eg, dim3(256), 0, st, n, P, (const double*)C, t, (double*)Out, ldv, sv, ldt,
                     stt, par);
  MNP_LAUNCH_CHECK();
  if (nrm != nullptr) {
    hipLaunchKernelGGL(k_norms2, dim3(batch), dim3(1024), 0, st, n, 8 * P, (const double*)C, (const double*)Out, ldv,
                       sv, nrm);
    MNP_LAUNCH_CHECK();
  }
  return 0;
}

// Y[b] = buf[b] X[b] (the product with M^T for the stored buffer of M), beta
// as in mnpbem_gemv_t_batched.
extern "C" int mnpbem_gemv_n_batched(int dtype, int n, int C, const void* M, long ldm, long sM, const void* X,
                                     long ldx, long sX, void* Y, long ldy, long sY, double beta, int batch,
                                     void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || C <= 0 || batch <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    return gemv_n<double>(n, C, (const double*)M, ldm, sM, (const double*)X, ldx, sX, (double*)Y, ldy, sY, beta,
                          batch, st);
  if (mnp_is_c64(dtype))
    return gemv_n<float>(n, C, (const float*)M, ldm, sM, (const float*)X, ldx, sX, (float*)Y, ldy, sY, beta, batch,
                         st);
  mnp_set_error("gemv_n: unknown dtype");
  return 1;
}

// V <- M^-T V with the LU of M^T (mnpbem_lu_solve_t_batched's transpose).
extern "C" int mnpbem_lu_solve_n_batched(int dtype, int n, int C, const void* LU, long ld, long sLU,
                                         const int* ipiv, long sP, void* V, long ldv, long sV, void* Z, long ldz,
                                         long sZ, const void* Dinv, long sD, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n <= 0 || C <= 0 || batch <= 0) return 0;
  MNP_REQUIRE(ldz >= C, "lu_solve_n: work space narrower than the right-hand side");
  if (dtype == MNP_DTYPE_C128)
    return lu_solve_n<double>(n, C, (const double*)LU, ld, sLU, ipiv, sP, (double*)V, ldv, sV, (double*)Z, ldz,
                              sZ, batch, st, (const double*)Dinv, sD);
  if (mnp_is_c64(dtype))
    return lu_solve_n<float>(n, C, (const float*)LU, ld, sLU, ipiv, sP, (float*)V, ldv, sV, (float*)Z, ldz, sZ,
                             batch, st, (const float*)Dinv, sD);
  mnp_set_error("lu_solve_n: unknown dtype");
  return 1;
}

// ---------------------------------------------------------------------------
// Right side of the BEM equations from an excitation's potentials (GdA&H
// Eqs. 10, 11, 15, 18; bem_ret.py:1360-1429):
//   phi = phi2 - phi1,  a = a2 - a1,
//   alpha = a2' - a1' - i k0 n (eps2 phi2 - eps1 phi1),
//   De = eps2 phi2' - eps1 phi1' - i k0 (eps2 n.a2 - eps1 n.a1),
// written as R = [phi | a | De | alpha] (n x 8P).  phi*, phi*p: n x P, a*, a*p:
// n x 3 x P complex128 (row-major, index c P + p), any of them NULL for zero;
// eps1, eps2: per-face permittivities inside / outside (complex128).
// ---------------------------------------------------------------------------
namespace {
__global__ void k_excitation(int n, int P, const double* phi1, const double* phi1p, const double* a1,
                             const double* a1p, const double* phi2, const double* phi2p, const double* a2,
                             const double* a2p, const double* nvec, const double* eps1, const double* eps2, double k0,
                             double* R, long ldv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  auto ld = [&](const double* q, long e) { return q ? cmake(q[2 * e], q[2 * e + 1]) : cmake(0.0, 0.0); };
  const cplx<double> e1 = cmake(eps1[2 * i], eps1[2 * i + 1]), e2 = cmake(eps2[2 * i], eps2[2 * i + 1]);
  const double nx[3] = {nvec[3 * i], nvec[3 * i + 1], nvec[3 * i + 2]};
  double* r = R + 2 * (long)i * ldv;
  for (int p = 0; p < P; ++p) {
    const long sp = (long)i * P + p;
    const cplx<double> f1 = ld(phi1, sp), f2 = ld(phi2, sp);
    std_(r + 2 * p, csub(f2, f1));
    const cplx<double> t = csub(cmul(e2, f2), cmul(e1, f1));   // eps2 phi2 - eps1 phi1
    cplx<double> na1 = cmake(0.0, 0.0), na2 = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) {
      const long sa = (long)i * 3 * P + c * P + p;
      const cplx<double> va1 = ld(a1, sa), va2 = ld(a2, sa);
      std_(r + 2 * (P + c * P + p), csub(va2, va1));
      na1 = cadd(na1, cscale(va1, nx[c]));
      na2 = cadd(na2, cscale(va2, nx[c]));
      // alpha_c = a2p - a1p - i k0 n_c t
      const cplx<double> al = csub(csub(ld(a2p, sa), ld(a1p, sa)), ik_times(k0, cscale(t, nx[c])));
      std_(r + 2 * (5 * P + c * P + p), al);
    }
    const cplx<double> dp = csub(cmul(e2, ld(phi2p, sp)), cmul(e1, ld(phi1p, sp)));
    const cplx<double> na = csub(cmul(e2, na2), cmul(e1, na1));
    std_(r + 2 * (4 * P + p), csub(dp, ik_times(k0, na)));
  }
}
}  // namespace

extern "C" int mnpbem_excitation(int n, int P, const double* phi1, const double* phi1p, const double* a1,
                                 const double* a1p, const double* phi2, const double* phi2p, const double* a2,
                                 const double* a2p, const double* nvec, const double* eps1, const double* eps2,
                                 double k0, double* R, long ldv, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && ldv >= 8 * P && k0 > 0, "excitation: bad size");
  hipLaunchKernelGGL(k_excitation, dim3(mnp_ceil_div(n, 256)), dim3(256), 0, st, n, P, phi1, phi1p, a1, a1p, phi2,
                     phi2p, a2, a2p, nvec, eps1, eps2, k0, R, ldv);
  MNP_LAUNCH_CHECK();
  return 0;
}
