// Batched dense complex linear algebra on CDNA4 (gfx950): MFMA complex GEMM,
// recursive LU with partial pivoting, recursive triangular solves, and the
// LU-solve with many right-hand sides.
//
// These replace scipy.linalg.lu_factor / lu_solve (LAPACK zgetrf / zgetrs) as
// called by BEMRet.init (mnpbem/bem/bem_ret.py:437-523, via
// mnpbem/utils/gpu.py:145-279).  Matrices are row-major, complex interleaved,
// ``batch`` matrices of one size spaced ``stride`` complex elements apart.
// Pivot choice follows LAPACK: largest |re|+|im| in the column, first index
// on ties, no scaling of the column when the pivot is exactly zero (the
// 1-based column is then reported in ``info``).
#include "common.h"

#include <cstdlib>
#include <type_traits>

#include <vector>

namespace {

// ---------------------------------------------------------------------------
// MFMA traits: 16x16x4 single-block matrix instructions.
// Operand layout: lane l supplies A[l%16][l/16] and B[l/16][l%16].  The four
// accumulator values of lane l (r = 0..3) sit in column l%16 and row
//   f32: 4*(l/16) + r        f64: l/16 + 4*r
// (the f64 row interleave was measured on gfx950 with a lane-index probe).
// ---------------------------------------------------------------------------
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// ---------------------------------------------------------------------------
// C[b] -= A[b] * B[b]   (complex, row-major; M x K times K x N)
// Workgroup: 256 threads = 4 waves, 64x64 output tile, each wave 32x32 as
// 2x2 blocks of 16x16.  K is staged through LDS in slices of 16 with the real
// and imaginary parts split into separate planes; a complex product is four
// real MFMAs (re*re - im*im, re*im + im*re) accumulated in FP32/FP64.
// ---------------------------------------------------------------------------
constexpr int GT = 64;   // output tile edge
constexpr int GK = 16;   // K slice
constexpr int GP = 1;    // LDS row padding (elements)

template <typename T>
__global__ __launch_bounds__(256) void k_cgemm_sub(int M, int N, int K, const T* __restrict__ A,
                                                   long lda, long sA, const T* __restrict__ B, long ldb,
                                                   long sB, T* __restrict__ C, long ldc, long sC) {
  typedef typename Mfma<T>::v4 v4;
  const long bz = blockIdx.z;
  A += 2 * bz * sA;
  B += 2 * bz * sB;
  C += 2 * bz * sC;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  __shared__ T Ar[GK][GT + GP], Ai[GK][GT + GP], Br[GK][GT + GP], Bi[GK][GT + GP];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // global load assignment: A tile 64 x 16 -> row tid/4, k (tid%4)*4 .. +3
  const int la_r = tid >> 2, la_k = (tid & 3) * 4;
  // B tile 16 x 64 -> k tid/16, col (tid%16)*4 .. +3
  const int lb_k = tid >> 4, lb_n = (tid & 15) * 4;

  cplx<T> ra[4], rb[4];
  auto gload = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int gr = m0 + la_r, gk = k0 + la_k + q;
      ra[q] = (gr < M && gk < K) ? cld(A + 2 * ((long)gr * lda + gk)) : cmake<T>(0, 0);
      int gk2 = k0 + lb_k, gc = n0 + lb_n + q;
      rb[q] = (gk2 < K && gc < N) ? cld(B + 2 * ((long)gk2 * ldb + gc)) : cmake<T>(0, 0);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      Ar[la_k + q][la_r] = ra[q].re;
      Ai[la_k + q][la_r] = ra[q].im;
      Br[lb_k][lb_n + q] = rb[q].re;
      Bi[lb_k][lb_n + q] = rb[q].im;
    }
  };

  v4 accR[2][2], accI[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      accR[i][j] = v4{0, 0, 0, 0};
      accI[i][j] = v4{0, 0, 0, 0};
    }

  gload(0);
  sstore();
  __syncthreads();
  for (int k0 = 0; k0 < K; k0 += GK) {
    const bool more = (k0 + GK) < K;
    if (more) gload(k0 + GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      const int kl = kk + (lane >> 4);
      T ar[2], ai[2], br[2], bi[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        ar[i] = Ar[kl][wm + 16 * i + (lane & 15)];
        ai[i] = Ai[kl][wm + 16 * i + (lane & 15)];
        br[i] = Br[kl][wn + 16 * i + (lane & 15)];
        bi[i] = Bi[kl][wn + 16 * i + (lane & 15)];
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          accR[i][j] = Mfma<T>::mma(ar[i], br[j], accR[i][j]);
          accR[i][j] = Mfma<T>::mma(-ai[i], bi[j], accR[i][j]);
          accI[i][j] = Mfma<T>::mma(ar[i], bi[j], accI[i][j]);
          accI[i][j] = Mfma<T>::mma(ai[i], br[j], accI[i][j]);
        }
    }
    __syncthreads();
    if (more) {
      sstore();
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int gc = n0 + wn + 16 * j + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = m0 + wm + 16 * i + Mfma<T>::row(lane, r);
        if (gr < M && gc < N) {
          T* p = C + 2 * ((long)gr * ldc + gc);
          cplx<T> c = cld(p);
          c.re -= accR[i][j][r];
          c.im -= accI[i][j][r];
          cst(p, c);
        }
      }
    }
}

// ---------------------------------------------------------------------------
// C[b] -= A[b] * B[b] for complex64 data on bf16 matrix cores.  Every float
// x is split into x_hi = bf16(x) and x_lo = bf16(x - x_hi); a real product is
// x_hi y_hi + x_hi y_lo + x_lo y_hi (the lo*lo term, ~2^-16 relative, is
// dropped), accumulated in FP32 by v_mfma_f32_16x16x32_bf16.  A complex
// product takes 12 such MFMAs.  Accuracy is ~1e-5 relative per product; the
// mixed-precision BEM solve recovers FP64 accuracy by iterative refinement.
// Tile shapes (template BM x BN, waves of WM x 32): 128 x 128 with 8 waves of
// 64 x 32 (4 x 2 blocks of 16 x 16, 128 KB LDS, one workgroup per CU) for
// the bulk of the factorisation; 64 x 64 with 4 waves of 32 x 32 (64 KB LDS,
// two workgroups per CU) for the thin products deep in the recursion, where
// the per-workgroup latency dominates.  A shape-class table chooses between
// them (GemmTuning below).  Both accumulate every output element over K in
// the same order, so the choice does not change the result.  K slices of 32
// are staged in LDS as four bf16 planes (re_hi, re_lo, im_hi, im_lo), B
// stored transposed so that every MFMA operand is one 16-byte LDS read per
// lane.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

constexpr int XK = 32;   // K slice

__device__ __forceinline__ bf16x8 bneg(bf16x8 v) {
  u32x4 u = __builtin_bit_cast(u32x4, v);
  u ^= 0x80008000u;
  return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ f32x4 mfma_bf16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// LDS layout: rows of XK = 32 bf16 (four 16-byte chunks, 64 B), chunk c of
// row r stored at chunk c ^ ((r >> 1) & 3): the 16-byte stores and the
// fragment reads of eight consecutive rows then fall in distinct banks.
__device__ __forceinline__ int xsw(int r, int c) { return r * XK + ((c ^ ((r >> 1) & 3)) << 3); }

// split 8 floats (stride 2 in v: real or imaginary parts of 8 complex) into
// bf16 hi and lo vectors
__device__ __forceinline__ void split8(const float* v, bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const float x = v[2 * t];
    const __bf16 h = (__bf16)x;
    hi[t] = h;
    lo[t] = (__bf16)(x - (float)h);
  }
}

// Output tile of a workgroup.  The dispatcher hands consecutive workgroups
// of the linearised grid to the 8 XCDs in turn, and each XCD has its own L2;
// with remap, XCD x instead gets one contiguous range of the tiles, ordered
// matrix by matrix and, inside a matrix, in groups of 4 tile rows walked
// column by column, so the workgroups resident on one XCD share A row panels
// and B column panels in its L2.  A bijection on the tiles: results do not
// depend on it.
struct Tile {
  int x, y;
  long z;
};
__device__ __forceinline__ Tile gemm_tile(int remap) {
  Tile t;
  if (!remap) {
    t.x = blockIdx.x;
    t.y = blockIdx.y;
    t.z = blockIdx.z;
    return t;
  }
  const long gx = gridDim.x, gy = gridDim.y;
  const long T = gx * gy * gridDim.z;
  const long L = blockIdx.x + gx * (blockIdx.y + gy * (long)blockIdx.z);
  const long q = T / 8, r = T % 8, xcd = L % 8;
  const long id = xcd * q + (xcd < r ? xcd : r) + L / 8;
  const long per = gx * gy;
  t.z = id / per;
  const long rem = id - t.z * per;
  constexpr long GROUP = 4;
  const long g = rem / (GROUP * gx);
  const long y0 = g * GROUP;
  const long gs = (gy - y0) < GROUP ? (gy - y0) : GROUP;
  const long w = rem - g * GROUP * gx;
  t.y = (int)(y0 + w % gs);
  t.x = (int)(w / gs);
  return t;
}

// ---------------------------------------------------------------------------
// Software pipeline of the K loop.  A loop that runs the MFMAs of a K slice,
// then (after waiting for the global loads) the bf16 split and LDS stores of
// the next slice, then the barrier, sends both waves of a SIMD into the split
// together, and the matrix pipe idles through its ~150 vector instructions
// (measured: 283-307 TFLOP/s on large shapes, against 331-347 for this form).
// Here slice k+1 is already in registers when the MFMAs of slice k start: its
// split and LDS stores sit between the MFMAs of the first column block, the
// global loads of slice k+2 after them.  While slice k+2 lies inside K the
// body is one basic block with compile-time stages (the loop is unrolled by
// two), so the compiler may interleave the split with the MFMAs and move LDS
// writes of one stage past reads of the other: rows of A beyond M and
// columns of B beyond N are clamped to the last row / column instead of
// zero-filled (they only feed outputs that are not stored).  The slices near
// the end of K go through a generic loop with zero fill.  Explicit
// interleaving with sched_group_barrier (split among the MFMAs, loads
// after them or right after the split of their registers) was measured
// no faster than the compiler's own order.
// ---------------------------------------------------------------------------
template <int BM, int BN, int WM, bool CPRE = (BM == 64), bool LEAF = false>
__global__ __launch_bounds__((BM / WM) * (BN / 32) * 64) void k_cgemm_sub_bf16x3(
    int M, int N, int K, const float* __restrict__ A, long lda, long sA, const float* __restrict__ B, long ldb,
    long sB, float* __restrict__ C, long ldc, long sC, int remap, int assign, const float* __restrict__ Dl,
    long sDl) {
  const Tile tl = gemm_tile(remap);
  const long bz = tl.z;
  A += 2 * bz * sA;
  B += 2 * bz * sB;
  C += 2 * bz * sC;
  constexpr int WN = 32, FI = WM / 16, FJ = WN / 16, NWN = BN / WN;
  static_assert(FJ == 2, "two column blocks per wave");
  const int m0 = tl.y * BM, n0 = tl.x * BN;
  __shared__ __attribute__((aligned(16))) __bf16 As[2][4][BM * XK];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][4][BN * XK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = (wave / NWN) * WM, wn = (wave % NWN) * WN;
  const int la_r = tid >> 2, la_c = tid & 3;
  const int lb_n = tid % BN, lb_c = tid / BN;
  const int ar = min(m0 + la_r, M - 1), bc = min(n0 + lb_n, N - 1);
  const float* Ap = A + 2 * ((long)ar * lda + la_c * 8);
  const float* Bp = B + 2 * ((long)(lb_c * 8) * ldb + bc);
  float ra[16], rb[16];
  // slice k0 .. k0 + 31 inside K
  auto gload_full = [&](int k0) {
    const float4* src = reinterpret_cast<const float4*>(Ap + 2 * k0);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 v = src[q];
      ra[4 * q] = v.x;
      ra[4 * q + 1] = v.y;
      ra[4 * q + 2] = v.z;
      ra[4 * q + 3] = v.w;
    }
    const float* b = Bp + 2 * (long)k0 * ldb;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const float2 v = *reinterpret_cast<const float2*>(b + 2 * (long)q * ldb);
      rb[2 * q] = v.x;
      rb[2 * q + 1] = v.y;
    }
  };
  // slice at k0 with k beyond K zero-filled
  auto gload_edge = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int gk = k0 + la_c * 8 + q;
      const float2 v = gk < K ? *reinterpret_cast<const float2*>(A + 2 * ((long)ar * lda + gk)) : make_float2(0.f, 0.f);
      ra[2 * q] = v.x;
      ra[2 * q + 1] = v.y;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int gk = k0 + lb_c * 8 + q;
      const float2 v = gk < K ? *reinterpret_cast<const float2*>(B + 2 * ((long)gk * ldb + bc)) : make_float2(0.f, 0.f);
      rb[2 * q] = v.x;
      rb[2 * q + 1] = v.y;
    }
  };
  const int oa = xsw(la_r, la_c), ob = xsw(lb_n, lb_c);
  auto sstore = [&](int stg) {
    bf16x8 h, l;
    split8(ra, h, l);
    *reinterpret_cast<bf16x8*>(&As[stg][0][oa]) = h;
    *reinterpret_cast<bf16x8*>(&As[stg][1][oa]) = l;
    split8(ra + 1, h, l);
    *reinterpret_cast<bf16x8*>(&As[stg][2][oa]) = h;
    *reinterpret_cast<bf16x8*>(&As[stg][3][oa]) = l;
    split8(rb, h, l);
    *reinterpret_cast<bf16x8*>(&Bs[stg][0][ob]) = h;
    *reinterpret_cast<bf16x8*>(&Bs[stg][1][ob]) = l;
    split8(rb + 1, h, l);
    *reinterpret_cast<bf16x8*>(&Bs[stg][2][ob]) = h;
    *reinterpret_cast<bf16x8*>(&Bs[stg][3][ob]) = l;
  };
  f32x4 accR[FI][FJ], accI[FI][FJ];
#pragma unroll
  for (int i = 0; i < FI; ++i)
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      accR[i][j] = f32x4{0, 0, 0, 0};
      accI[i][j] = f32x4{0, 0, 0, 0};
    }
  const int fr = lane & 15, fc = lane >> 4;
  int oaf[FI], obf[FJ];
#pragma unroll
  for (int i = 0; i < FI; ++i) oaf[i] = xsw(wm + 16 * i + fr, fc);
#pragma unroll
  for (int j = 0; j < FJ; ++j) obf[j] = xsw(wn + 16 * j + fr, fc);
  // the 12 MFMAs of column block j, in the order of k_cgemm_sub_bf16x3
  auto mma = [&](int j, const bf16x8 (&a)[FI][4], const bf16x8 (&b)[4]) {
    const bf16x8 brh = b[0], brl = b[1], bih = b[2], bil = b[3];
    const bf16x8 nbih = bneg(bih), nbil = bneg(bil);
#pragma unroll
    for (int i = 0; i < FI; ++i) accR[i][j] = mfma_bf16(a[i][0], brh, accR[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accI[i][j] = mfma_bf16(a[i][0], bih, accI[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accR[i][j] = mfma_bf16(a[i][0], brl, accR[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accI[i][j] = mfma_bf16(a[i][0], bil, accI[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accR[i][j] = mfma_bf16(a[i][1], brh, accR[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accI[i][j] = mfma_bf16(a[i][1], bih, accI[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accR[i][j] = mfma_bf16(a[i][2], nbih, accR[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accI[i][j] = mfma_bf16(a[i][2], brh, accI[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accR[i][j] = mfma_bf16(a[i][2], nbil, accR[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accI[i][j] = mfma_bf16(a[i][2], brl, accI[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accR[i][j] = mfma_bf16(a[i][3], nbih, accR[i][j]);
#pragma unroll
    for (int i = 0; i < FI; ++i) accI[i][j] = mfma_bf16(a[i][3], brh, accI[i][j]);
  };
  auto frags = [&](int stg, bf16x8 (&a)[FI][4], bf16x8 (&b)[FJ][4]) {
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) a[i][pl] = *reinterpret_cast<const bf16x8*>(&As[stg][pl][oaf[i]]);
#pragma unroll
    for (int j = 0; j < FJ; ++j)
#pragma unroll
      for (int pl = 0; pl < 4; ++pl) b[j][pl] = *reinterpret_cast<const bf16x8*>(&Bs[stg][pl][obf[j]]);
  };
  // steady state: slice s from stage S, slice s+1 (registers) to stage S^1,
  // slice s+2 (inside K) loaded into the registers
  auto body = [&](auto stc, int s) {
    constexpr int S = decltype(stc)::value;
    bf16x8 a[FI][4], b[FJ][4];
    frags(S, a, b);
    mma(0, a, b[0]);
    sstore(S ^ 1);
    mma(1, a, b[1]);
    gload_full((s + 2) * XK);
    __syncthreads();
  };
  // CPRE: the C tile is read into registers before the K loop.  The 64 x 64
  // tile always does it: its launches are the few-row, small-K products of
  // the recursions, where the epilogue's wait for C was a large part of the
  // launch (13-33 % faster at K <= 256).  The 128 x 128 tile has no registers
  // to spare for it on long K loops; its prefetching instance (tile variant
  // 2) is for the short ones the shape table assigns to it.
  float2 cpre[CPRE ? FI * FJ * 4 : 1];
  if (CPRE && !assign) {
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) {
        const int gc = min(n0 + wn + 16 * j + (lane & 15), N - 1);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gr = min(m0 + wm + 16 * i + 4 * (lane >> 4) + r, M - 1);
          cpre[(i * FJ + j) * 4 + r] = *reinterpret_cast<const float2*>(C + 2 * ((long)gr * ldc + gc));
        }
      }
  }
  const int nk = (K + XK - 1) / XK, nfull = K / XK;
  if (nfull >= 1)
    gload_full(0);
  else
    gload_edge(0);
  sstore(0);
  if (nk > 1) {
    if (nfull >= 2)
      gload_full(XK);
    else
      gload_edge(XK);
  }
  __syncthreads();
  int s = 0;
  for (; s + 3 < nfull; s += 2) {
    body(std::integral_constant<int, 0>(), s);
    body(std::integral_constant<int, 1>(), s + 1);
  }
  if (s + 2 < nfull) {
    body(std::integral_constant<int, 0>(), s);
    ++s;
  }
  for (; s < nk; ++s) {
    const int stg = s & 1;
    bf16x8 a[FI][4], b[FJ][4];
    frags(stg, a, b);
    mma(0, a, b[0]);
    mma(1, a, b[1]);
    if (s + 1 < nk) sstore(stg ^ 1);
    if (s + 2 < nk) gload_edge((s + 2) * XK);
    __syncthreads();
  }
  if constexpr (LEAF) {
    // Block-row leaf in the epilogue (one row tile, M <= BM = 128): C_new =
    // C - A B stays on chip as the B operand of X = D C_new, D the block's
    // 128 x 128 inverse (rows and columns past M zero), and X is stored in
    // place of C.  C_new goes to LDS as bf16 planes in the B layout, its four
    // K slices of 32 rows in the two stages of both operand arrays; D's
    // fragments come straight from memory with the same split.  The splits
    // and the MFMA order are those of the stand-alone in-place product X <- D
    // X on this tile, so the result is bit-identical to it.
    typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
    static_assert(BM == 128 && BN == 128, "leaf epilogue on the 128 x 128 tile");
    __syncthreads();  // every wave done with the last slice's fragments
    auto bpl = [&](int sl, int pl) -> __bf16* { return sl < 2 ? &Bs[sl][pl][0] : &As[sl - 2][pl][0]; };
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) {
        const int col = wn + 16 * j + fr, gc = n0 + col;
        const int k0 = wm + 16 * i + 4 * (lane >> 4);
        bf16x4 rh, rl, ih, il;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gr = m0 + k0 + r;
          const float2 c = (gr < M && gc < N) ? *reinterpret_cast<const float2*>(C + 2 * ((long)gr * ldc + gc))
                                              : make_float2(0.f, 0.f);
          const float xr = c.x - accR[i][j][r], xi = c.y - accI[i][j][r];
          const __bf16 hr = (__bf16)xr, hi = (__bf16)xi;
          rh[r] = hr;
          rl[r] = (__bf16)(xr - (float)hr);
          ih[r] = hi;
          il[r] = (__bf16)(xi - (float)hi);
        }
        const int sl = k0 >> 5, kk = k0 & 31;
        const int off = xsw(col, kk >> 3) + (kk & 7);
        *reinterpret_cast<bf16x4*>(bpl(sl, 0) + off) = rh;
        *reinterpret_cast<bf16x4*>(bpl(sl, 1) + off) = rl;
        *reinterpret_cast<bf16x4*>(bpl(sl, 2) + off) = ih;
        *reinterpret_cast<bf16x4*>(bpl(sl, 3) + off) = il;
      }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) {
        accR[i][j] = f32x4{0, 0, 0, 0};
        accI[i][j] = f32x4{0, 0, 0, 0};
      }
    const float* Dp = Dl + 2 * bz * sDl;
#pragma unroll 1
    for (int sl = 0; sl < 4; ++sl) {
      bf16x8 a[FI][4], b[FJ][4];
#pragma unroll
      for (int i = 0; i < FI; ++i) {
        const float4* src = reinterpret_cast<const float4*>(Dp + 2 * ((long)(wm + 16 * i + fr) * BM + XK * sl + 8 * fc));
        float v[16];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float4 t = src[q];
          v[4 * q] = t.x;
          v[4 * q + 1] = t.y;
          v[4 * q + 2] = t.z;
          v[4 * q + 3] = t.w;
        }
        split8(v, a[i][0], a[i][1]);
        split8(v + 1, a[i][2], a[i][3]);
      }
#pragma unroll
      for (int j = 0; j < FJ; ++j)
#pragma unroll
        for (int pl = 0; pl < 4; ++pl) b[j][pl] = *reinterpret_cast<const bf16x8*>(bpl(sl, pl) + obf[j]);
      mma(0, a, b[0]);
      mma(1, a, b[1]);
    }
#pragma unroll
    for (int i = 0; i < FI; ++i)
#pragma unroll
      for (int j = 0; j < FJ; ++j) {
        const int gc = n0 + wn + 16 * j + fr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int gr = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
          if (gr < M && gc < N)
            *reinterpret_cast<float2*>(C + 2 * ((long)gr * ldc + gc)) = make_float2(accR[i][j][r], accI[i][j][r]);
        }
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < FI; ++i)
#pragma unroll
    for (int j = 0; j < FJ; ++j) {
      const int gc = n0 + wn + 16 * j + fr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gr = m0 + wm + 16 * i + 4 * (lane >> 4) + r;
        if (gr < M && gc < N) {
          float* p = C + 2 * ((long)gr * ldc + gc);
          float2 c;
          if (assign)
            c = make_float2(0.f, 0.f);
          else if constexpr (CPRE)
            c = cpre[(i * FJ + j) * 4 + r];
          else
            c = *reinterpret_cast<float2*>(p);
          c.x = assign ? accR[i][j][r] : c.x - accR[i][j][r];
          c.y = assign ? accI[i][j][r] : c.y - accI[i][j][r];
          *reinterpret_cast<float2*>(p) = c;
        }
      }
    }
}

// ---------------------------------------------------------------------------
// LU of the panel A[c0:N, c0:c0+w] (w <= W) with partial pivoting, one
// workgroup per matrix.  Thread t holds rows t, t+NT, ... (S rows) of the
// panel in registers; per column: workgroup argmax of |re|+|im|, row
// exchange through LDS, scaling and rank-1 update of the remaining columns.
// ipiv receives global (0-based) row indices.  Swaps are applied inside the
// panel columns; with fullswap the leaf also exchanges the two rows in every
// column outside the panel as soon as the pivot is known (the interchanges
// LAPACK's recursive getrf applies later with laswp: the same permutation,
// applied before any later use of those rows), otherwise the callers apply
// them to the other columns.
// ---------------------------------------------------------------------------
template <typename T, int W, int S, int NT>
__global__ __launch_bounds__(NT) void k_lu_leaf(int N, int c0, int w, T* A, long lda, long sA,
                                                int* ipiv, long sP, int* info, int fullswap) {
  const long b = blockIdx.x;
  A += 2 * b * sA;
  ipiv += b * sP;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  constexpr int NW = NT / 64;
  const int m = N - c0;
  cplx<T> a[S][W];
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int r = tid + s * NT;
#pragma unroll
    for (int c = 0; c < W; ++c)
      a[s][c] = (r < m && c < w) ? cld(A + 2 * ((long)(c0 + r) * lda + c0 + c)) : cmake<T>(0, 0);
  }
  __shared__ T sval[2 * NW];
  __shared__ int sidx[2 * NW];
  __shared__ cplx<T> rowc[W], rowp[W];

#pragma unroll
  for (int c = 0; c < W; ++c) {
    if (c < w) {
    T best = T(-1);
    int bi = 0x7fffffff;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int r = tid + s * NT;
      if (r >= c && r < m) {
        const T v = cabs1(a[s][c]);
        if (v > best) {
          best = v;
          bi = r;
        }
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const T ov = __shfl_xor(best, off);
      const int oi = __shfl_xor(bi, off);
      if (ov > best || (ov == best && oi < bi)) {
        best = ov;
        bi = oi;
      }
    }
    if (lane == 0) {
      sval[(c & 1) * NW + wv] = best;
      sidx[(c & 1) * NW + wv] = bi;
    }
    __syncthreads();
    // every thread reduces the per-wave candidates (double-buffered by column
    // parity, so the next column's writes cannot race these reads)
    int p;
    {
      T bb = sval[(c & 1) * NW];
      int ii = sidx[(c & 1) * NW];
      for (int q = 1; q < NW; ++q) {
        const T v = sval[(c & 1) * NW + q];
        const int iq = sidx[(c & 1) * NW + q];
        if (v > bb || (v == bb && iq < ii)) {
          bb = v;
          ii = iq;
        }
      }
      if (ii < c || ii >= m) ii = c;  // only NaN entries left: keep the diagonal
      p = ii;
    }
    if (tid == 0) ipiv[c0 + c] = c0 + p;
    if (fullswap && p != c) {
      // columns outside the panel: memory that no other part of this launch touches
      T* rc_ = A + 2 * ((long)(c0 + c) * lda);
      T* rp_ = A + 2 * ((long)(c0 + p) * lda);
      for (int j = tid; j < N; j += NT) {
        if (j >= c0 && j < c0 + w) continue;
        const cplx<T> x = cld(rc_ + 2 * j), y = cld(rp_ + 2 * j);
        cst(rc_ + 2 * j, y);
        cst(rp_ + 2 * j, x);
      }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int r = tid + s * NT;
      if (r == p)
#pragma unroll
        for (int cc = 0; cc < W; ++cc) rowp[cc] = a[s][cc];
      if (r == c)
#pragma unroll
        for (int cc = 0; cc < W; ++cc) rowc[cc] = a[s][cc];
    }
    __syncthreads();
    if (p != c) {
#pragma unroll
      for (int s = 0; s < S; ++s) {
        const int r = tid + s * NT;
        if (r == c)
#pragma unroll
          for (int cc = 0; cc < W; ++cc) a[s][cc] = rowp[cc];
        if (r == p)
#pragma unroll
          for (int cc = 0; cc < W; ++cc) a[s][cc] = rowc[cc];
      }
    }
    const cplx<T> pv = rowp[c];
    const bool zero = (pv.re == T(0) && pv.im == T(0));
    if (zero && tid == 0 && info[b] == 0) info[b] = c0 + c + 1;
    const cplx<T> inv = zero ? cmake<T>(1, 0) : crecip(pv);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int r = tid + s * NT;
      if (r > c && r < m) {
        const cplx<T> l = cmul(a[s][c], inv);
        a[s][c] = l;
#pragma unroll
        for (int cc = c + 1; cc < W; ++cc) a[s][cc] = cfms(a[s][cc], l, rowp[cc]);
      }
    }
    // no barrier here: the next column writes rowp / rowc only after its
    // first barrier, which every thread reaches after finishing this column,
    // and its candidates go to the other parity of sval / sidx
    }
  }
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int r = tid + s * NT;
    if (r < m)
#pragma unroll
      for (int c = 0; c < W; ++c)
        if (c < w) cst(A + 2 * ((long)(c0 + r) * lda + c0 + c), a[s][c]);
  }
}

// ---------------------------------------------------------------------------
// Apply row interchanges ipiv[k0..k1) in order (forward) or reverse order to
// columns [clo, chi).  Pivots are staged through LDS; interchanges with
// ipiv[i] == i touch no memory, and chunks of 1024 pivots without any are
// skipped whole.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_laswp(int k0, int k1, int clo, int chi, T* A, long lda,
                                               long sA, const int* ipiv, long sP, int reverse) {
  const long b = blockIdx.y;
  A += 2 * b * sA;
  ipiv += b * sP;
  const int col = clo + blockIdx.x * 256 + threadIdx.x;
  __shared__ int sp[1024];
  const int n = k1 - k0;
  for (int base = 0; base < n; base += 1024) {
    const int cnt = min(1024, n - base);
    __syncthreads();
    int moved = 0;
    for (int q = threadIdx.x; q < cnt; q += 256) {
      const int i = reverse ? (k1 - 1 - base - q) : (k0 + base + q);
      sp[q] = ipiv[i];
      moved |= sp[q] != i;
    }
    // a chunk without interchanges (the factorisations without them record
    // ipiv[i] = i) is skipped instead of walked pivot by pivot
    if (!__syncthreads_or(moved)) continue;
    for (int q = 0; q < cnt; ++q) {
      const int i = reverse ? (k1 - 1 - base - q) : (k0 + base + q);
      const int p = sp[q];
      if (p != i && col < chi) {
        T* x = A + 2 * ((long)i * lda + col);
        T* y = A + 2 * ((long)p * lda + col);
        const cplx<T> u = cld(x), v = cld(y);
        cst(x, v);
        cst(y, u);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Small triangular solve X = T^-1 B for an n x n (n <= TLN) block T: lower
// with unit diagonal (LOWER) or upper with its diagonal.  T sits in LDS
// (zero-padded to TLN, so no bounds tests in the loops); every thread owns
// one right-hand-side column in registers and runs the substitution without
// further synchronisation.
// ---------------------------------------------------------------------------
template <typename T>
struct TLeaf;
template <>
struct TLeaf<double> {
  static constexpr int n = 32;
};
template <>
struct TLeaf<float> {
  static constexpr int n = 64;
};

template <typename T, bool LOWER, int TLN>
__global__ __launch_bounds__(256) void k_trsm_leaf(int n, int nrhs, const T* Tm, long ldt, long sT, T* B,
                                                   long ldb, long sB) {
  const long b = blockIdx.y;
  Tm += 2 * b * sT;
  B += 2 * b * sB;
  // The substitution always runs in upper-triangular form; a lower system is
  // mapped onto it by reversing the row and column order (index n-1-i).
  __shared__ cplx<T> Ts[TLN][TLN];
  __shared__ cplx<T> rd[TLN];
  const int tid = threadIdx.x;
  for (int e = tid; e < TLN * TLN; e += 256) {
    const int i = e / TLN, j = e % TLN;
    const int si = LOWER ? n - 1 - i : i, sj = LOWER ? n - 1 - j : j;
    Ts[i][j] = (i < n && j < n) ? cld(Tm + 2 * ((long)si * ldt + sj)) : cmake<T>(0, 0);
  }
  __syncthreads();
  if (tid < TLN) rd[tid] = (tid >= n) ? cmake<T>(0, 0) : (LOWER ? cmake<T>(1, 0) : crecip(Ts[tid][tid]));
  __syncthreads();
  const int gc = blockIdx.x * 256 + tid;
  if (gc >= nrhs) return;
  cplx<T> x[TLN];
#pragma unroll
  for (int i = 0; i < TLN; ++i) {
    const int si = LOWER ? n - 1 - i : i;
    x[i] = (i < n) ? cld(B + 2 * ((long)si * ldb + gc)) : cmake<T>(0, 0);
  }
#pragma unroll
  for (int i = TLN - 1; i >= 0; --i) {
#pragma unroll
    for (int l = i + 1; l < TLN; ++l) x[i] = cfms(x[i], Ts[i][l], x[l]);
    x[i] = cmul(x[i], rd[i]);
  }
#pragma unroll
  for (int i = 0; i < TLN; ++i) {
    const int si = LOWER ? n - 1 - i : i;
    if (i < n) cst(B + 2 * ((long)si * ldb + gc), x[i]);
  }
}

// The same leaf for few right-hand sides, where one thread per column leaves
// the substitution chain of TLN steps as the whole launch time.  A workgroup
// takes 64 columns of one matrix (lane = column) and its four waves take four
// blocks of RB = TLN / 4 rows (upper-triangular form as above).  Block k (k =
// 3 .. 0): wave k solves its diagonal block in registers, writes the solved
// rows to LDS; after the barrier, waves q < k subtract T[block q][block k]
// times them from their rows.  Four barriers per solve; the result equals
// the substitution up to rounding (the sums run in another order).
template <typename T, bool LOWER, int TLN>
__global__ __launch_bounds__(256) void k_trsm_leaf_rows(int n, int nrhs, const T* Tm, long ldt, long sT, T* B,
                                                        long ldb, long sB) {
  constexpr int RB = TLN / 4;
  const long b = blockIdx.y;
  Tm += 2 * b * sT;
  B += 2 * b * sB;
  __shared__ cplx<T> Ts[TLN][TLN];
  __shared__ cplx<T> rd[TLN];
  __shared__ cplx<T> xs[TLN][64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int e = tid; e < TLN * TLN; e += 256) {
    const int i = e / TLN, j = e % TLN;
    const int si = LOWER ? n - 1 - i : i, sj = LOWER ? n - 1 - j : j;
    Ts[i][j] = (i < n && j < n) ? cld(Tm + 2 * ((long)si * ldt + sj)) : cmake<T>(0, 0);
  }
  __syncthreads();
  if (tid < TLN) rd[tid] = (tid >= n) ? cmake<T>(0, 0) : (LOWER ? cmake<T>(1, 0) : crecip(Ts[tid][tid]));
  const int gc = blockIdx.x * 64 + lane;
  const bool live = gc < nrhs;
  cplx<T> x[RB];
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int r = wv * RB + i;
    const int si = LOWER ? n - 1 - r : r;
    x[i] = (live && r < n) ? cld(B + 2 * ((long)si * ldb + gc)) : cmake<T>(0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int k = 3; k >= 0; --k) {
    if (wv == k) {
#pragma unroll
      for (int i = RB - 1; i >= 0; --i) {
#pragma unroll
        for (int l = i + 1; l < RB; ++l) x[i] = cfms(x[i], Ts[k * RB + i][k * RB + l], x[l]);
        x[i] = cmul(x[i], rd[k * RB + i]);
      }
#pragma unroll
      for (int i = 0; i < RB; ++i) xs[k * RB + i][lane] = x[i];
    }
    __syncthreads();
    if (wv < k) {
#pragma unroll
      for (int l = 0; l < RB; ++l) {
        const cplx<T> xl = xs[k * RB + l][lane];
#pragma unroll
        for (int i = 0; i < RB; ++i) x[i] = cfms(x[i], Ts[wv * RB + i][k * RB + l], xl);
      }
    }
  }
  if (live) {
#pragma unroll
    for (int i = 0; i < RB; ++i) {
      const int r = wv * RB + i;
      const int si = LOWER ? n - 1 - r : r;
      if (r < n) cst(B + 2 * ((long)si * ldb + gc), x[i]);
    }
  }
}

__global__ void k_fill_int(int* p, long n, int v) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// C[b] -= A[b] B[b] for thin complex64 products (N <= 32, K <= 64: the
// updates at the bottom of the LU recursion, M up to the full height, N = K
// = 8, 16 or 32) in FP32 on the vector units.  On the 64 x 64 bf16x3 tile
// such a product is mostly padding and its launch time is latency; here one
// thread computes one output element (K fused complex multiply-adds in order
// of k, then C - sum), rows of 256 / N per workgroup, B staged in LDS.  More
// accurate than the split-bf16 product.  Measured on the 2884-face sweep:
// 149.5 wavelengths/s with every product on the bf16x3 tiles, 152.5 with this
// kernel up to N = 16, 154.5 up to N = 32, 148.9 up to N = 64.  Since the
// 64 x 64 tile reads its C tile ahead of its K loop, it takes the N = 17..32
// products (M ~2800, K = 64) faster: one of them ran ~300 us here at batch
// 128; sweep 295-300 -> 290-292 ms with this kernel up to N = 16 only.
constexpr int THIN_N = 32, THIN_K = 64;
constexpr int THIN_N_DEFAULT = 16;
int g_thin_n = THIN_N_DEFAULT;  // products with N <= g_thin_n (and K <= THIN_K) use k_cgemm_thin
__global__ __launch_bounds__(256) void k_cgemm_thin(int M, int N, int K, const float* __restrict__ A, long lda,
                                                    long sA, const float* __restrict__ B, long ldb, long sB,
                                                    float* __restrict__ C, long ldc, long sC) {
  const long b = blockIdx.y;
  A += 2 * b * sA;
  B += 2 * b * sB;
  C += 2 * b * sC;
  __shared__ cplx<float> Bs[THIN_K][THIN_N];
  const int tid = threadIdx.x;
  for (int e = tid; e < K * N; e += 256) {
    const int k = e / N, c = e % N;
    Bs[k][c] = cld(B + 2 * ((long)k * ldb + c));
  }
  __syncthreads();
  const int rpb = 256 / N;
  const int rr = tid / N, c = tid % N;
  const int r = blockIdx.x * rpb + rr;
  if (rr >= rpb || r >= M) return;
  const float* a = A + 2 * (long)r * lda;
  cplx<float> acc = cmake<float>(0.f, 0.f);
  for (int k = 0; k < K; ++k) acc = cfma(acc, cld(a + 2 * k), Bs[k][c]);
  float* p = C + 2 * ((long)r * ldc + c);
  cst(p, csub(cld(p), acc));
}

// ---------------------------------------------------------------------------
// host-side recursion
// ---------------------------------------------------------------------------
template <typename T>
struct Bat {
  T* A;
  long lda, sA;
  int batch;
  hipStream_t st;
  T* at(long r, long c) const { return A + 2 * (r * lda + c); }
};

// Optional timing of every complex-GEMM launch with a pair of HIP events on
// the launching stream (enabled by mnpbem_prof_begin), for the roofline of
// the dominant kernel: algorithmic flops 8*M*N*K*batch per launch.
struct GemmProf {
  bool on = false;
  std::vector<hipEvent_t> ev;
  std::vector<double> flops;
  double bytes = 0.0;   // compulsory bytes: A and B read once, C read and written once
  double cbytes = 0.0;  // of which the writes of C
  std::vector<int> shape;  // per timed launch: M, N, K, batch, kernel (0-2: bf16x3 tile, 3: generic, 4: assign128, 5: block-row leaf)
  size_t used = 0;
};
GemmProf g_prof;

// Tile choice of the bf16x3 GEMM per shape class: class of a dimension x is
// 0 for x <= 64, 1 for <= 128, ... 5 for <= 2048, 6 above; cell (cm, cn, ck)
// holds the variant (0: 128 x 128, 1: 64 x 64).  The library starts from the
// measured default (64 x 64 for the products with min(M, N) <= 64 and
// K <= 64) and takes a table set by the autotuner.  It also records, per
// cell, one shape that was launched, so the autotuner can time the shapes a
// workload actually uses.
constexpr int GC = 7;
struct GemmTuning {
  unsigned char var[GC * GC * GC];
  int seen[GC * GC * GC][4];  // M, N, K, batch of one launch (M = 0: none)
  GemmTuning() {
    for (int cm = 0; cm < GC; ++cm)
      for (int cn = 0; cn < GC; ++cn)
        for (int ck = 0; ck < GC; ++ck) var[(cm * GC + cn) * GC + ck] = ((cm == 0 || cn == 0) && ck == 0) ? 1 : 0;
    for (auto& q : seen) q[0] = q[1] = q[2] = q[3] = 0;
  }
};
GemmTuning g_tune;

inline int gclass(int x) {
  int c = 0;
  for (int t = 64; x > t && c < GC - 1; t <<= 1) ++c;
  return c;
}
inline int gcell(int M, int N, int K) { return (gclass(M) * GC + gclass(N)) * GC + gclass(K); }

// XCD-aware tile order of the bf16x3 GEMM (each XCD takes a contiguous range
// of tiles; see gemm_tile)
inline int gemm_xcd_remap() { return 1; }

// assign = 1: C = A B instead of C -= A B
int launch_bf16x3(int variant, int M, int N, int K, const float* A, long lda, long sA, const float* B, long ldb,
                  long sB, float* C, long ldc, long sC, int batch, hipStream_t st, int assign = 0) {
  if (variant == 1) {
    dim3 grid(mnp_ceil_div(N, 64), mnp_ceil_div(M, 64), batch);
    hipLaunchKernelGGL((k_cgemm_sub_bf16x3<64, 64, 32>), grid, dim3(256), 0, st, M, N, K, A, lda, sA, B, ldb, sB,
                       C, ldc, sC, gemm_xcd_remap(), assign, nullptr, 0L);
  } else if (variant == 2) {
    dim3 grid(mnp_ceil_div(N, 128), mnp_ceil_div(M, 128), batch);
    hipLaunchKernelGGL((k_cgemm_sub_bf16x3<128, 128, 64, true>), grid, dim3(512), 0, st, M, N, K, A, lda, sA, B, ldb,
                       sB, C, ldc, sC, gemm_xcd_remap(), assign, nullptr, 0L);
  } else {
    dim3 grid(mnp_ceil_div(N, 128), mnp_ceil_div(M, 128), batch);
    hipLaunchKernelGGL((k_cgemm_sub_bf16x3<128, 128, 64>), grid, dim3(512), 0, st, M, N, K, A, lda, sA, B, ldb, sB,
                       C, ldc, sC, gemm_xcd_remap(), assign, nullptr, 0L);
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// X <- D X in place for n <= 128 rows on the 128 x 128 tile (one row tile), timed like gemm_sub
int launch_bf16x3_assign128(int n, int nrhs, const float* D, long ldd, long sD, float* X, long ldx, long sX,
                            int batch, hipStream_t st);

// gmode 1 (complex64 only): bf16 split-product GEMM
// assign = 1 (bf16x3 kernel only): C = A B
template <typename T>
int gemm_sub(int M, int N, int K, const T* A, long lda, long sA, const T* B, long ldb, long sB, T* C, long ldc,
             long sC, int batch, hipStream_t st, int gmode, int assign = 0) {
  if (M <= 0 || N <= 0 || K <= 0 || batch <= 0) return 0;
  if constexpr (sizeof(T) == 4) {
    if (gmode == 1 && !assign && N <= g_thin_n && K <= THIN_K) {
      dim3 grid(mnp_ceil_div(M, 256 / N), batch);
      hipLaunchKernelGGL(k_cgemm_thin, grid, dim3(256), 0, st, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC);
      MNP_LAUNCH_CHECK();
      return 0;
    }
  }
  const bool timed = g_prof.on && g_prof.used + 2 <= g_prof.ev.size();
  if (timed) MNP_HIP_CHECK(hipEventRecord(g_prof.ev[g_prof.used], st));
  if constexpr (sizeof(T) == 4) {
    if (gmode == 1) {
      const int cell = gcell(M, N, K);
      if (g_tune.seen[cell][0] == 0) {
        g_tune.seen[cell][0] = M;
        g_tune.seen[cell][1] = N;
        g_tune.seen[cell][2] = K;
        g_tune.seen[cell][3] = batch;
      }
      int rc = launch_bf16x3(g_tune.var[cell], M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC, batch, st, assign);
      if (rc) return rc;
    } else {
      dim3 grid(mnp_ceil_div(N, GT), mnp_ceil_div(M, GT), batch);
      hipLaunchKernelGGL(k_cgemm_sub<T>, grid, dim3(256), 0, st, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC);
    }
  } else {
    dim3 grid(mnp_ceil_div(N, GT), mnp_ceil_div(M, GT), batch);
    hipLaunchKernelGGL(k_cgemm_sub<T>, grid, dim3(256), 0, st, M, N, K, A, lda, sA, B, ldb, sB, C, ldc, sC);
  }
  MNP_LAUNCH_CHECK();
  if (timed) {
    MNP_HIP_CHECK(hipEventRecord(g_prof.ev[g_prof.used + 1], st));
    g_prof.used += 2;
    g_prof.flops.push_back(8.0 * M * (double)N * K * batch);
    int kern = 3;
    if constexpr (sizeof(T) == 4) kern = gmode == 1 ? g_tune.var[gcell(M, N, K)] : 3;
    g_prof.shape.insert(g_prof.shape.end(), {M, N, K, batch, kern});
    g_prof.bytes += (double)batch * 2.0 * sizeof(T) * ((double)M * K + (double)K * N + 2.0 * M * N);
    g_prof.cbytes += (double)batch * 2.0 * sizeof(T) * ((double)M * N);
  }
  return 0;
}

int launch_bf16x3_assign128(int n, int nrhs, const float* D, long ldd, long sD, float* X, long ldx, long sX,
                            int batch, hipStream_t st) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return 0;
  const bool timed = g_prof.on && g_prof.used + 2 <= g_prof.ev.size();
  if (timed) MNP_HIP_CHECK(hipEventRecord(g_prof.ev[g_prof.used], st));
  int rc = launch_bf16x3(0, n, nrhs, n, D, ldd, sD, X, ldx, sX, X, ldx, sX, batch, st, 1);
  if (rc) return rc;
  if (timed) {
    MNP_HIP_CHECK(hipEventRecord(g_prof.ev[g_prof.used + 1], st));
    g_prof.used += 2;
    g_prof.flops.push_back(8.0 * n * (double)nrhs * n * batch);
    g_prof.shape.insert(g_prof.shape.end(), {n, nrhs, n, batch, 4});
    g_prof.bytes += (double)batch * 8.0 * ((double)n * n + 2.0 * n * nrhs);
    g_prof.cbytes += (double)batch * 8.0 * ((double)n * nrhs);
  }
  return 0;
}

// One block row of a left-looking solve in one launch: C <- D (C - A B) with
// M <= 128 rows (one row tile) and D the block's 128 x 128 inverse (per
// matrix stride sD); timed like gemm_sub (update and leaf flops).
int launch_bf16x3_leaf(int M, int N, int K, const float* A, long lda, long sA, const float* B, long ldb, long sB,
                       float* C, long ldc, long sC, const float* D, long sD, int batch, hipStream_t st) {
  if (M <= 0 || N <= 0 || batch <= 0) return 0;
  MNP_REQUIRE(M <= 128 && K > 0, "leaf product: M above 128 or K empty");
  const bool timed = g_prof.on && g_prof.used + 2 <= g_prof.ev.size();
  if (timed) MNP_HIP_CHECK(hipEventRecord(g_prof.ev[g_prof.used], st));
  dim3 grid(mnp_ceil_div(N, 128), 1, batch);
  hipLaunchKernelGGL((k_cgemm_sub_bf16x3<128, 128, 64, false, true>), grid, dim3(512), 0, st, M, N, K, A, lda, sA, B,
                     ldb, sB, C, ldc, sC, gemm_xcd_remap(), 0, D, sD);
  MNP_LAUNCH_CHECK();
  if (timed) {
    MNP_HIP_CHECK(hipEventRecord(g_prof.ev[g_prof.used + 1], st));
    g_prof.used += 2;
    g_prof.flops.push_back(8.0 * M * (double)N * (K + M) * batch);
    g_prof.shape.insert(g_prof.shape.end(), {M, N, K, batch, 5});
    g_prof.bytes += (double)batch * 8.0 * ((double)M * K + (double)K * N + 2.0 * M * N + 128.0 * 128.0);
    g_prof.cbytes += (double)batch * 8.0 * ((double)M * N);
  }
  return 0;
}

// Columns (right-hand sides x batch) up to which a triangular leaf of more
// than half the leaf size runs k_trsm_leaf_rows instead of one thread per
// column (set by mnpbem_trsm_rows_max; 0 disables).  Measured on the leaves
// of 64 rows: 50 against 69-82 us per getrs (both leaves) up to 500 x 64
// columns, 178 against 102 us at 760 x 128; the row-blocked form has the
// shorter chain but less throughput.
constexpr int TRSM_ROWS_DEFAULT = 32768;
long g_trsm_rows_max = TRSM_ROWS_DEFAULT;

// Inverse diagonal blocks of a triangular factor for the leaves of trsm_rec:
// per matrix (stride sD) and 64-row block, [U^-1 | L^-1] as two 64 x 64
// row-major blocks (the layout of k_lu_blockinv in solve.hip); blk is the
// block index of the node's first row.
template <typename T>
struct LeafInv {
  const T* D = nullptr;
  long sD = 0;
  int blk = 0;
  int bs = 64;  // block rows: 64 (k_lu_blockinv layout) or 128 (k_blockinv128)
  LeafInv at(int rows) const {
    LeafInv r = *this;
    r.blk += rows / bs;
    return r;
  }
};

// X <- T^-1 X, T n x n at Tm (lower unit or upper), X n x nrhs.  With the
// inverse diagonal blocks (complex64 on the bf16x3 kernel), a leaf of up to
// 64 rows is the product X <- D X with its block's inverse D instead of a
// substitution: one row tile covers the leaf, so every workgroup reads its
// whole column strip of X before writing it back.  The recursion splits at
// multiples of 64 rows from the first row, so leaves start on block
// boundaries and only the last leaf of the matrix is shorter than 64.
template <typename T>
int trsm_rec(bool lower, int n, int nrhs, const T* Tm, long ldt, long sT, T* X, long ldx, long sX, int batch,
             hipStream_t st, int gmode, LeafInv<T> inv = LeafInv<T>()) {
  if (n <= 0 || nrhs <= 0) return 0;
  constexpr int TL = TLeaf<T>::n;
  const bool prod = sizeof(T) == 4 && inv.D != nullptr && gmode == 1;
  if constexpr (sizeof(T) == 4) {
    if (prod && n <= inv.bs) {
      const int bs = inv.bs;
      const T* D = inv.D + 2 * ((long)inv.blk * 2 * bs * bs + (lower ? bs * bs : 0));
      // in place: the tile must cover all n rows (the 128-row tile above 64)
      if (n > 64) return launch_bf16x3_assign128(n, nrhs, D, bs, inv.sD, X, ldx, sX, batch, st);
      return gemm_sub<T>(n, nrhs, n, D, bs, inv.sD, X, ldx, sX, X, ldx, sX, batch, st, gmode, 1);
    }
  }
  if (n <= TL) {
    if (n > TL / 2 && (long)nrhs * batch <= g_trsm_rows_max) {
      dim3 grid(mnp_ceil_div(nrhs, 64), batch);
      if (lower)
        hipLaunchKernelGGL((k_trsm_leaf_rows<T, true, TL>), grid, dim3(256), 0, st, n, nrhs, Tm, ldt, sT, X, ldx, sX);
      else
        hipLaunchKernelGGL((k_trsm_leaf_rows<T, false, TL>), grid, dim3(256), 0, st, n, nrhs, Tm, ldt, sT, X, ldx,
                           sX);
      MNP_LAUNCH_CHECK();
      return 0;
    }
    if constexpr (TL > 32) {
      // 17 .. 32 rows (the 32-column level of the LU recursion)
      if (n > 16 && n <= 32 && (long)nrhs * batch <= g_trsm_rows_max) {
        dim3 grid(mnp_ceil_div(nrhs, 64), batch);
        if (lower)
          hipLaunchKernelGGL((k_trsm_leaf_rows<T, true, 32>), grid, dim3(256), 0, st, n, nrhs, Tm, ldt, sT, X, ldx,
                             sX);
        else
          hipLaunchKernelGGL((k_trsm_leaf_rows<T, false, 32>), grid, dim3(256), 0, st, n, nrhs, Tm, ldt, sT, X, ldx,
                             sX);
        MNP_LAUNCH_CHECK();
        return 0;
      }
    }
    dim3 grid(mnp_ceil_div(nrhs, 256), batch);
    // smallest unrolled size that holds the triangle
#define MNP_TRSM_LEAF(SZ)                                                                                    \
  if (n <= (SZ)) {                                                                                           \
    if (lower)                                                                                               \
      hipLaunchKernelGGL((k_trsm_leaf<T, true, SZ>), grid, dim3(256), 0, st, n, nrhs, Tm, ldt, sT, X, ldx, sX); \
    else                                                                                                     \
      hipLaunchKernelGGL((k_trsm_leaf<T, false, SZ>), grid, dim3(256), 0, st, n, nrhs, Tm, ldt, sT, X, ldx, sX); \
    MNP_LAUNCH_CHECK();                                                                                      \
    return 0;                                                                                                \
  }
    MNP_TRSM_LEAF(8)
    MNP_TRSM_LEAF(16)
    MNP_TRSM_LEAF(32)
    if constexpr (TL >= 64) {
      MNP_TRSM_LEAF(64)
    }
#undef MNP_TRSM_LEAF
    return 0;
  }
  // split at a multiple of the leaf size from the node's first row
  const int SB = prod ? inv.bs : TL;
  int n1 = ((n / 2 + SB - 1) / SB) * SB;
  if (n1 >= n) n1 = n - SB;
  const int n2 = n - n1;
  int rc;
  if (lower) {
    if ((rc = trsm_rec(true, n1, nrhs, Tm, ldt, sT, X, ldx, sX, batch, st, gmode, inv))) return rc;
    if ((rc = gemm_sub(n2, nrhs, n1, Tm + 2 * ((long)n1 * ldt), ldt, sT, X, ldx, sX, X + 2 * ((long)n1 * ldx), ldx,
                       sX, batch, st, gmode)))
      return rc;
    return trsm_rec(true, n2, nrhs, Tm + 2 * ((long)n1 * ldt + n1), ldt, sT, X + 2 * ((long)n1 * ldx), ldx, sX,
                    batch, st, gmode, inv.at(n1));
  } else {
    if ((rc = trsm_rec(false, n2, nrhs, Tm + 2 * ((long)n1 * ldt + n1), ldt, sT, X + 2 * ((long)n1 * ldx), ldx, sX,
                       batch, st, gmode, inv.at(n1))))
      return rc;
    if ((rc = gemm_sub(n1, nrhs, n2, Tm + 2 * (long)n1, ldt, sT, X + 2 * ((long)n1 * ldx), ldx, sX, X, ldx, sX,
                       batch, st, gmode)))
      return rc;
    return trsm_rec(false, n1, nrhs, Tm, ldt, sT, X, ldx, sX, batch, st, gmode, inv);
  }
}

// Right-hand sides from which the many-column triangular solves with 128-row
// inverse diagonal blocks run block row by block row (trsm_left) instead of
// recursively (set by mnpbem_trsm_left_rhs; 0 disables).
constexpr int TRSM_LEFT_DEFAULT = 512;
int g_trsm_left_rhs = TRSM_LEFT_DEFAULT;
// 1: the block rows after the first fold their leaf into the update's
// epilogue (launch_bf16x3_leaf); 0: update and leaf as two launches
// (mnpbem_trsm_leaf_fused).
int g_trsm_leaf_fused = 1;

// X <- T^-1 X block row by block row (left-looking), with the inverse
// diagonal blocks: for block row k of inv.bs rows (lower: first to last,
// upper: last to first), X_k -= T_k,s X_s over all rows s solved so far (one
// product whose K grows with k, so nearly all flops run on long-K products
// that keep their C tile in registers), then X_k <- D_k X_k.  The recursion
// of trsm_rec spends its lowest levels on products of K = 128..384 whose C
// read and write dominate.  Same operation count as trsm_rec with
// inverse-block leaves.
template <typename T>
int trsm_left(bool lower, int n, int nrhs, const T* Tm, long ldt, long sT, T* X, long ldx, long sX, int batch,
              hipStream_t st, int gmode, LeafInv<T> inv) {
  const int bs = inv.bs;
  const int nb = (n + bs - 1) / bs;
  for (int q = 0; q < nb; ++q) {
    const int k = lower ? q : nb - 1 - q;
    const int r0 = k * bs, m = min(bs, n - r0), r1 = r0 + m;
    int rc;
    if constexpr (sizeof(T) == 4) {
      if (g_trsm_leaf_fused && bs == 128 && (lower ? r0 > 0 : r1 < n)) {
        const float* D = (const float*)inv.D + 2 * ((long)(inv.blk + k) * 2 * bs * bs + (lower ? bs * bs : 0));
        if (lower)
          rc = launch_bf16x3_leaf(m, nrhs, r0, (const float*)Tm + 2 * ((long)r0 * ldt), ldt, sT, (const float*)X, ldx,
                                  sX, (float*)X + 2 * ((long)r0 * ldx), ldx, sX, D, inv.sD, batch, st);
        else
          rc = launch_bf16x3_leaf(m, nrhs, n - r1, (const float*)Tm + 2 * ((long)r0 * ldt + r1), ldt, sT,
                                  (const float*)X + 2 * ((long)r1 * ldx), ldx, sX, (float*)X + 2 * ((long)r0 * ldx),
                                  ldx, sX, D, inv.sD, batch, st);
        if (rc) return rc;
        continue;
      }
    }
    if (lower && r0 > 0) {
      if ((rc = gemm_sub<T>(m, nrhs, r0, Tm + 2 * ((long)r0 * ldt), ldt, sT, X, ldx, sX, X + 2 * ((long)r0 * ldx),
                            ldx, sX, batch, st, gmode)))
        return rc;
    } else if (!lower && r1 < n) {
      if ((rc = gemm_sub<T>(m, nrhs, n - r1, Tm + 2 * ((long)r0 * ldt + r1), ldt, sT, X + 2 * ((long)r1 * ldx), ldx,
                            sX, X + 2 * ((long)r0 * ldx), ldx, sX, batch, st, gmode)))
        return rc;
    }
    if ((rc = trsm_rec<T>(lower, m, nrhs, Tm + 2 * ((long)r0 * ldt + r0), ldt, sT, X + 2 * ((long)r0 * ldx), ldx, sX,
                          batch, st, gmode, inv.at(r0))))
      return rc;
  }
  return 0;
}

// the many-column solves: left-looking with 128-row inverse blocks on the
// bf16x3 products, else the recursion
template <typename T>
int trsm_many(bool lower, int n, int nrhs, const T* Tm, long ldt, long sT, T* X, long ldx, long sX, int batch,
              hipStream_t st, int gmode, LeafInv<T> inv) {
  if (sizeof(T) == 4 && gmode == 1 && inv.D != nullptr && inv.bs == 128 && g_trsm_left_rhs > 0 &&
      nrhs >= g_trsm_left_rhs && n > 2 * inv.bs)
    return trsm_left<T>(lower, n, nrhs, Tm, ldt, sT, X, ldx, sX, batch, st, gmode, inv);
  return trsm_rec<T>(lower, n, nrhs, Tm, ldt, sT, X, ldx, sX, batch, st, gmode, inv);
}

template <typename T>
int laswp(int k0, int k1, int clo, int chi, T* A, long lda, long sA, const int* ipiv, long sP, int batch, bool rev,
          hipStream_t st) {
  if (k1 <= k0 || chi <= clo) return 0;
  dim3 grid(mnp_ceil_div(chi - clo, 256), batch);
  hipLaunchKernelGGL(k_laswp<T>, grid, dim3(256), 0, st, k0, k1, clo, chi, A, lda, sA, ipiv, sP, rev ? 1 : 0);
  MNP_LAUNCH_CHECK();
  return 0;
}

// leaf configurations: W panel columns, S rows per thread, NT threads.  The
// panel (S*W complex per thread) stays in registers without spilling, so the
// panel narrows as the remaining rows m = N - c0 grow.  complex64: W = 8 up
// to 3072 rows, 4 up to 6144, 2 up to 12288 (512 threads), then one column
// with 1024 threads up to 32768 rows.  complex128: 8 up to 3072, 4 up to
// 6144, then one column with 512 threads up to 20480 rows.
template <typename T>
int leaf_width(int m) {
  if (m <= 3072) return 8;
  if (m <= 6144) return 4;
  if (sizeof(T) == 4) {
    if (m <= 12288) return 2;
    return m <= 32768 ? 1 : 0;
  }
  return m <= 20480 ? 1 : 0;
}

template <typename T, int W, int SMAX, int NT>
int leaf_launch_w(int N, int c0, int w, T* A, long lda, long sA, int* ipiv, long sP, int* info, int batch,
                  hipStream_t st, int fullswap) {
  const int m = N - c0;
#define MNP_LEAF(S_)                                                                                       \
  if (m <= (S_)*NT) {                                                                                      \
    hipLaunchKernelGGL((k_lu_leaf<T, W, S_, NT>), dim3(batch), dim3(NT), 0, st, N, c0, w, A, lda, sA, ipiv,  \
                       sP, info, fullswap);                                                                \
    MNP_LAUNCH_CHECK();                                                                                    \
    return 0;                                                                                              \
  }
  MNP_LEAF(1)
  MNP_LEAF(2)
  MNP_LEAF(3)
  MNP_LEAF(6)
  if constexpr (SMAX >= 12) {
    MNP_LEAF(12)
  }
  if constexpr (SMAX >= 24) {
    MNP_LEAF(24)
  }
  if constexpr (SMAX >= 32) {
    MNP_LEAF(32)
  }
  if constexpr (SMAX >= 40) {
    MNP_LEAF(40)
  }
#undef MNP_LEAF
  mnp_set_error("getrf: matrix too tall for the register-resident panel");
  return 1;
}

template <typename T>
int leaf_launch(int W, int N, int c0, int w, T* A, long lda, long sA, int* ipiv, long sP, int* info, int batch,
                hipStream_t st, int fullswap) {
  if (W == 8) return leaf_launch_w<T, 8, 6, 512>(N, c0, w, A, lda, sA, ipiv, sP, info, batch, st, fullswap);
  if (W == 4) return leaf_launch_w<T, 4, 12, 512>(N, c0, w, A, lda, sA, ipiv, sP, info, batch, st, fullswap);
  if constexpr (sizeof(T) == 4) {
    if (W == 2) return leaf_launch_w<T, 2, 24, 512>(N, c0, w, A, lda, sA, ipiv, sP, info, batch, st, fullswap);
    if (W == 1) return leaf_launch_w<T, 1, 32, 1024>(N, c0, w, A, lda, sA, ipiv, sP, info, batch, st, fullswap);
  } else {
    if (W == 1) return leaf_launch_w<T, 1, 40, 512>(N, c0, w, A, lda, sA, ipiv, sP, info, batch, st, fullswap);
  }
  mnp_set_error("getrf: matrix too tall for the register-resident panel");
  return 1;
}

// ---------------------------------------------------------------------------
// Tall panels (more rows than the register-resident leaf holds with 8
// columns): the panel of width w <= TALL_W is factored column by column by
// many workgroups at once, two launches per column and no grid-wide barrier:
//   k_tall_pivot (one workgroup per matrix): reduce the per-chunk candidates
//     of column c to the pivot row p (largest |re|+|im|, lowest row on ties,
//     the rule of k_lu_leaf), exchange rows c and p inside the panel, record
//     ipiv and 1 / pivot;
//   k_tall_update (row chunks x matrices): scale column c below the diagonal,
//     rank-1 update of the panel columns right of c, and the per-chunk
//     candidates of column c+1.
// The same pivots and the same arithmetic per entry as the leaf, so the
// factors do not depend on which path ran.  Candidates live in a small static
// device buffer (no allocation; concurrent tall factorisations on one device
// would share it).
constexpr int TALL_W = 32;       // panel width
constexpr int TALL_CH = 512;     // rows per workgroup chunk (256 threads x 2)
constexpr int TALL_MAXC = 64;    // chunks per matrix (32768 rows)
constexpr int TALL_MAXB = 512;   // matrices per launch
__device__ double g_tall_val[TALL_MAXB * TALL_MAXC];
__device__ int g_tall_idx[TALL_MAXB * TALL_MAXC];
__device__ double g_tall_inv[2 * TALL_MAXB];

// per-chunk argmax of |re|+|im| of column col over the rows >= col of chunk
// blockIdx.x (rows [r0 + k TALL_CH, r0 + (k+1) TALL_CH) of [r0, N)); each
// thread reads the rows it updated itself in k_tall_update
template <typename T>
__device__ void tall_candidates(int N, int r0, int col, const T* A, long lda, long b, int nch) {
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int lo = r0 + blockIdx.x * TALL_CH;
  T best = T(-1);
  int bi = 0x7fffffff;
  for (int r = lo + tid; r < min(N, lo + TALL_CH); r += 256) {
    if (r < col) continue;
    const T v = cabs1(cld(A + 2 * ((long)r * lda + col)));
    if (v > best) {
      best = v;
      bi = r;
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const T ov = __shfl_xor(best, off);
    const int oi = __shfl_xor(bi, off);
    if (ov > best || (ov == best && oi < bi)) {
      best = ov;
      bi = oi;
    }
  }
  __shared__ T sv[4];
  __shared__ int si[4];
  if (lane == 0) {
    sv[wv] = best;
    si[wv] = bi;
  }
  __syncthreads();
  if (tid == 0) {
    for (int q = 1; q < 4; ++q)
      if (sv[q] > best || (sv[q] == best && si[q] < bi)) {
        best = sv[q];
        bi = si[q];
      }
    g_tall_val[b * TALL_MAXC + blockIdx.x] = (double)best;
    g_tall_idx[b * TALL_MAXC + blockIdx.x] = bi;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tall_first(int N, int c0, T* A, long lda, long sA, int nch) {
  const long b = blockIdx.y;
  tall_candidates<T>(N, c0, c0, A + 2 * b * sA, lda, b, nch);
}

template <typename T>
__global__ __launch_bounds__(64) void k_tall_pivot(int N, int c0, int w, int c, T* A, long lda, long sA, int* ipiv,
                                                   long sP, int* info, int nch) {
  const long b = blockIdx.x;
  A += 2 * b * sA;
  const int tid = threadIdx.x;
  const int r = c0 + c;
  __shared__ int sp;
  if (tid == 0) {
    double best = -1.0;
    int bi = 0x7fffffff;
    for (int q = 0; q < nch; ++q) {
      const double v = g_tall_val[b * TALL_MAXC + q];
      const int iq = g_tall_idx[b * TALL_MAXC + q];
      if (v > best || (v == best && iq < bi)) {
        best = v;
        bi = iq;
      }
    }
    if (bi < r || bi >= N) bi = r;  // only NaN entries left: keep the diagonal
    sp = bi;
    ipiv[b * sP + r] = bi;
  }
  __syncthreads();
  const int p = sp;
  if (p != r)
    for (int t = tid; t < w; t += 64) {
      T* x = A + 2 * ((long)r * lda + c0 + t);
      T* y = A + 2 * ((long)p * lda + c0 + t);
      const cplx<T> u = cld(x), v = cld(y);
      cst(x, v);
      cst(y, u);
    }
  __syncthreads();
  if (tid == 0) {
    const cplx<T> pv = cld(A + 2 * ((long)r * lda + r));
    const bool zero = (pv.re == T(0) && pv.im == T(0));
    if (zero && info[b] == 0) info[b] = r + 1;
    const cplx<T> inv = zero ? cmake<T>(1, 0) : crecip(pv);
    g_tall_inv[2 * b] = (double)inv.re;
    g_tall_inv[2 * b + 1] = (double)inv.im;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_tall_update(int N, int c0, int w, int c, T* A, long lda, long sA, int nch) {
  const long b = blockIdx.y;
  A += 2 * b * sA;
  const int r0 = c0 + c;  // pivot row
  __shared__ cplx<T> prow[TALL_W];
  const int tid = threadIdx.x;
  if (tid < w) prow[tid] = cld(A + 2 * ((long)r0 * lda + c0 + tid));
  __syncthreads();
  const cplx<T> inv = cmake<T>((T)g_tall_inv[2 * b], (T)g_tall_inv[2 * b + 1]);
  // chunks cover rows [c0, N) in TALL_CH blocks; only rows below the pivot change
  const int lo = c0 + blockIdx.x * TALL_CH;
  for (int r = lo + tid; r < min(N, lo + TALL_CH); r += 256) {
    if (r <= r0) continue;
    T* row = A + 2 * ((long)r * lda + c0);
    const cplx<T> l = cmul(cld(row + 2 * c), inv);
    cst(row + 2 * c, l);
    for (int cc = c + 1; cc < w; ++cc) cst(row + 2 * cc, cfms(cld(row + 2 * cc), l, prow[cc]));
  }
  if (c + 1 < w) {
    __syncthreads();
    tall_candidates<T>(N, c0, c0 + c + 1, A, lda, b, nch);
  }
}

// rows per matrix above which the panel is factored by the multi-workgroup
// column kernels: where the register leaf is down to one column (measured on
// gfx950, complex64 LU: 0.82 s instead of 1.22 s at 25000 rows, 0.33 instead
// of 0.36 at 12500, but slower than the 2- and 4-column leaves below);
// 0 = the built-in thresholds (set by mnpbem_getrf_tall_rows)
int g_tall_rows = 0;

template <typename T>
bool tall_ok(int m, int batch) {
  const int thr = g_tall_rows > 0 ? g_tall_rows : (sizeof(T) == 4 ? 12288 : 6144);
  return m > thr && batch <= TALL_MAXB && (m + TALL_CH - 1) / TALL_CH <= TALL_MAXC;
}

template <typename T>
int tall_panel(int N, int c0, int w, T* A, long lda, long sA, int* ipiv, long sP, int* info, int batch,
               hipStream_t st) {
  const int m = N - c0;
  // chunks of the rows [c0, N): the candidates of column c are searched in
  // rows >= c0 + c, so a chunk may hold none (its candidate is then empty)
  const int nch = (m + TALL_CH - 1) / TALL_CH;
  hipLaunchKernelGGL(k_tall_first<T>, dim3(nch, batch), dim3(256), 0, st, N, c0, A, lda, sA, nch);
  MNP_LAUNCH_CHECK();
  for (int c = 0; c < w; ++c) {
    hipLaunchKernelGGL(k_tall_pivot<T>, dim3(batch), dim3(64), 0, st, N, c0, w, c, A, lda, sA, ipiv, sP, info, nch);
    MNP_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tall_update<T>, dim3(nch, batch), dim3(256), 0, st, N, c0, w, c, A, lda, sA, nch);
    MNP_LAUNCH_CHECK();
  }
  return 0;
}

// recursive LU of columns [c0, c0+n) (rows c0..N-1).  fullswap: the leaves
// exchange whole rows (every factorisation whose first panel is not a tall
// one, so that no tall panel follows either), and the interchanges need no
// laswp launches.
template <typename T>
int getrf_rec(int N, int c0, int n, T* A, long lda, long sA, int* ipiv, long sP, int* info, int batch,
              hipStream_t st, int gmode, bool fullswap) {
  const bool tall = tall_ok<T>(N - c0, batch);
  const int W = tall ? TALL_W : leaf_width<T>(N - c0);
  if (W == 0) {
    mnp_set_error("getrf: matrix too tall for the register-resident panel");
    return 1;
  }
  if (n <= W) {
    if (tall) return tall_panel<T>(N, c0, n, A, lda, sA, ipiv, sP, info, batch, st);
    return leaf_launch<T>(W, N, c0, n, A, lda, sA, ipiv, sP, info, batch, st, fullswap ? 1 : 0);
  }
  int n1 = ((n / 2 + W - 1) / W) * W;
  if (n1 >= n) n1 = n - W;
  const int n2 = n - n1;
  int rc;
  if ((rc = getrf_rec(N, c0, n1, A, lda, sA, ipiv, sP, info, batch, st, gmode, fullswap))) return rc;
  if (!fullswap && (rc = laswp<T>(c0, c0 + n1, c0 + n1, c0 + n, A, lda, sA, ipiv, sP, batch, false, st))) return rc;
  T* A11 = A + 2 * ((long)c0 * lda + c0);
  T* A12 = A + 2 * ((long)c0 * lda + c0 + n1);
  T* A21 = A + 2 * ((long)(c0 + n1) * lda + c0);
  T* A22 = A + 2 * ((long)(c0 + n1) * lda + c0 + n1);
  if ((rc = trsm_rec<T>(true, n1, n2, A11, lda, sA, A12, lda, sA, batch, st, gmode))) return rc;
  if ((rc = gemm_sub<T>(N - c0 - n1, n2, n1, A21, lda, sA, A12, lda, sA, A22, lda, sA, batch, st, gmode))) return rc;
  if ((rc = getrf_rec(N, c0 + n1, n2, A, lda, sA, ipiv, sP, info, batch, st, gmode, fullswap))) return rc;
  if (fullswap) return 0;
  return laswp<T>(c0 + n1, c0 + n, c0, c0 + n1, A, lda, sA, ipiv, sP, batch, false, st);
}

// ---------------------------------------------------------------------------
// LU without row interchanges, for matrices whose partial pivoting makes none
// (the Green matrices G1, G2 and Delta of the BEM system are diagonally
// dominant in this sense: every multiplier has |re| + |im| <= 0.51 on the
// test geometries).  Partial pivoting takes the entry of largest |re| + |im|
// of the remaining column (lowest row on ties); it keeps the diagonal at
// every step exactly when no multiplier exceeds 1 in |re| + |im|, and then
// its factors are these.  The factorisation records that test per matrix,
// flag[b] = 1 when some multiplier exceeds 1 and 2 when one exceeds
// NP_GUARD (a growth guard for callers that accept a factorisation without
// the interchanges, e.g. as the inner solve of iterative refinement), so
// that the caller can factor again with interchanges.
//
// Panel of w <= W columns [c0, c0 + w): k_lu_np_diag (one workgroup per
// matrix) factors the w x w diagonal block in LDS, stores it and leaves
// U11^-1 in a static device buffer; k_lu_np_rows then forms the multipliers
// below as a product, L21 = A21 U11^-1, 64 rows per workgroup, every output
// a sum over k <= j in order.  No per-column reduction or row exchange: two
// launches per W columns against W / 8 for the pivoting leaf.  The buffer
// holds NP_MAXB matrices (the entry point runs larger batches in chunks);
// concurrent factorisations on one device would share it.
//
// At the bottom nodes of the recursion (W < n <= 2W columns: a panel of W and
// one of nx = n - W) the left panel's kernels also do the node's triangular
// solve and update: k_lu_np_diag solves U12 = L11^-1 A12 (the nx columns to
// the right of its block, in place) and k_lu_np_rows, after forming its rows
// of L21, subtracts L21 U12 from the same rows' nx columns.  That replaces a
// triangular leaf launch and a thin product launch per node, and L21 is used
// from LDS instead of being read back.
// ---------------------------------------------------------------------------
template <typename T>
struct NpW;
template <>
struct NpW<float> {
  static constexpr int w = 32;
};
template <>
struct NpW<double> {
  static constexpr int w = 16;
};
constexpr int NP_MAXB = 512;
constexpr float NP_GUARD = 1024.0f;

template <typename T>
__device__ __forceinline__ int np_level(cplx<T> l) {
  const T a = cabs1(l);
  return a > T(NP_GUARD) ? 2 : (a > T(1) ? 1 : 0);
}
__device__ float g_np_uinv_f[NP_MAXB * 64 * 64 * 2];
__device__ double g_np_uinv_d[NP_MAXB * 32 * 32 * 2];
template <typename T>
__device__ __forceinline__ T* np_uinv();
template <>
__device__ __forceinline__ float* np_uinv<float>() {
  return g_np_uinv_f;
}
template <>
__device__ __forceinline__ double* np_uinv<double>() {
  return g_np_uinv_d;
}

// Diagonal block: thread (row i = tid / 4, column group cg = tid % 4) holds
// row i's columns cg, cg + 4, ... in registers; per column k the pivot row
// and the column go through LDS (double-buffered, one barrier per column)
// and every row below updates its own entries.  Then X = U11^-1 column by
// column, four lanes per column splitting each sum (adjacent lanes, added
// by shuffles in a fixed order).
// 1 / a with one real division (|a|^2 well inside the range: pivots of
// diagonally dominant, O(1)-scaled blocks)
template <typename T>
__device__ __forceinline__ cplx<T> crecip_np(cplx<T> a) {
  const T sc = T(1) / (a.re * a.re + a.im * a.im);
  return cmake(a.re * sc, -a.im * sc);
}

template <typename T, int W>
__global__ __launch_bounds__(256) void k_lu_np_diag(int c0, int w, T* A, long lda, long sA, int* info, int* flag,
                                                    int nx) {
  constexpr int CPT = W / 4;
  const long b = blockIdx.x;
  A += 2 * b * sA;
  const int tid = threadIdx.x;
  const int i = tid >> 2, cg = tid & 3;
  const bool own = i < w;
  __shared__ cplx<T> rowb[2][W], colb[2][W];
  __shared__ cplx<T> U[W][W + 1];
  __shared__ cplx<T> dinv[W];
  cplx<T> a[CPT];
#pragma unroll
  for (int jj = 0; jj < CPT; ++jj) {
    const int j = cg + 4 * jj;
    a[jj] = (own && j < w) ? cld(A + 2 * ((long)(c0 + i) * lda + c0 + j)) : cmake<T>(0, 0);
  }
  int lev = 0;
  for (int k = 0; k < w; ++k) {
    const int pb = k & 1;
    // branch-free selects of the thread's entry in column k
    cplx<T> ck = a[0];
#pragma unroll
    for (int jj = 1; jj < CPT; ++jj) {
      const bool hit = cg + 4 * jj == k;
      ck.re = hit ? a[jj].re : ck.re;
      ck.im = hit ? a[jj].im : ck.im;
    }
    if (i == k)
#pragma unroll
      for (int jj = 0; jj < CPT; ++jj) rowb[pb][cg + 4 * jj] = a[jj];
    if (own && i > k && cg == (k & 3)) colb[pb][i] = ck;
    __syncthreads();
    const cplx<T> pv = rowb[pb][k];
    const bool zero = pv.re == T(0) && pv.im == T(0);
    if (tid == 0 && zero && info[b] == 0) info[b] = c0 + k + 1;
    if (own && i > k) {
      const cplx<T> l = cmul(colb[pb][i], zero ? cmake<T>(1, 0) : crecip_np(pv));
      lev = max(lev, np_level(l));
#pragma unroll
      for (int jj = 0; jj < CPT; ++jj) {
        const int j = cg + 4 * jj;
        const cplx<T> nv = cfms(a[jj], l, rowb[pb][j]);
        const bool upd = j > k, put = j == k;
        a[jj].re = put ? l.re : (upd ? nv.re : a[jj].re);
        a[jj].im = put ? l.im : (upd ? nv.im : a[jj].im);
      }
    }
  }
#pragma unroll
  for (int jj = 0; jj < CPT; ++jj) {
    const int j = cg + 4 * jj;
    if (own && j < w) cst(A + 2 * ((long)(c0 + i) * lda + c0 + j), a[jj]);
    if (i < W) U[i][j] = a[jj];
  }
  __syncthreads();
  if (tid < W) {
    const cplx<T> pv = U[tid][tid];
    dinv[tid] = (tid >= w || (pv.re == T(0) && pv.im == T(0))) ? cmake<T>(1, 0) : crecip_np(pv);
  }
  __syncthreads();
  // X = U11^-1: thread (column jc, part q) keeps rows q, q + 4, ... of column jc
  {
    const int jc = tid >> 2, q = tid & 3;
    cplx<T> xr[CPT];
#pragma unroll
    for (int m = 0; m < CPT; ++m) xr[m] = cmake<T>(0, 0);
#pragma unroll
    for (int ii = W - 1; ii >= 0; --ii) {
      cplx<T> sm = cmake<T>(0, 0);
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        const int l = q + 4 * m;
        if (l > ii) {  // compile-time test; xr[m] is zero for rows beyond jc
          const cplx<T> t = cfma(sm, U[ii][l], xr[m]);
          sm = t;
        }
      }
      sm.re += __shfl_xor(sm.re, 1);
      sm.im += __shfl_xor(sm.im, 1);
      sm.re += __shfl_xor(sm.re, 2);
      sm.im += __shfl_xor(sm.im, 2);
      const cplx<T> d = cmake<T>(ii == jc ? T(1) : T(0), T(0));
      cplx<T> xv = cmul(csub(d, sm), dinv[ii]);
      const bool live = ii <= jc && jc < w;
      xv.re = live ? xv.re : T(0);
      xv.im = live ? xv.im : T(0);
#pragma unroll
      for (int m = 0; m < CPT; ++m) {
        const bool hit = q + 4 * m == ii;
        xr[m].re = hit ? xv.re : xr[m].re;
        xr[m].im = hit ? xv.im : xr[m].im;
      }
    }
    T* Xg = np_uinv<T>() + 2 * b * W * W;
    if (jc < W)
#pragma unroll
      for (int m = 0; m < CPT; ++m) cst(Xg + 2 * ((long)(q + 4 * m) * W + jc), xr[m]);
  }
  // U12 = L11^-1 A12 for the nx columns right of the block (unit lower L11 in
  // U below the diagonal): the columns in LDS, one thread per column
  if (nx > 0) {
    __shared__ cplx<T> X12[W][W + 1];
    for (int e = tid; e < W * W; e += 256) {
      const int r = e / W, c = e % W;
      X12[r][c] = (r < w && c < nx) ? cld(A + 2 * ((long)(c0 + r) * lda + c0 + w + c)) : cmake<T>(0, 0);
    }
    __syncthreads();
    if (tid < nx) {
#pragma unroll 1
      for (int r = 1; r < w; ++r) {
        cplx<T> s = X12[r][tid];
        for (int k = 0; k < r; ++k) s = cfms(s, U[r][k], X12[k][tid]);
        X12[r][tid] = s;
      }
      for (int r = 0; r < w; ++r) cst(A + 2 * ((long)(c0 + r) * lda + c0 + w + tid), X12[r][tid]);
    }
  }
  if (lev) atomicMax(flag + b, lev);
}

// 64 rows per workgroup; wave v forms columns [v W/4, (v+1) W/4) of
// L21 = A21 U11^-1 (sums over k below the wave's last column: U11^-1 is
// upper), lane (row group, column group) 4 rows x W/16 columns, every output
// summed over k in order
template <typename T, int W>
__global__ __launch_bounds__(256) void k_lu_np_rows(int N, int c0, int w, T* A, long lda, long sA, int* flag,
                                                    int nx) {
  constexpr int CPW = W / 4, CW = CPW / 4, RW = 4;
  const long b = blockIdx.y;
  A += 2 * b * sA;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  __shared__ cplx<T> As[64][W + 1];
  __shared__ cplx<T> Xs[W][W + 1];
  const int r0 = c0 + w + blockIdx.x * 64;
  const int nr = min(64, N - r0);
  for (int e = tid; e < 64 * W; e += 256) {
    const int ii = e / W, j = e % W;
    As[ii][j] = (ii < nr && j < w) ? cld(A + 2 * ((long)(r0 + ii) * lda + c0 + j)) : cmake<T>(0, 0);
  }
  const T* Xg = np_uinv<T>() + 2 * b * W * W;
  for (int e = tid; e < W * W; e += 256) Xs[e / W][e % W] = cld(Xg + 2 * e);
  __syncthreads();
  const int rg = lane >> 2, cs = lane & 3;
  const int jb = wv * CPW + cs * CW;
  const int kend = min(w, (wv + 1) * CPW);
  cplx<T> acc[RW][CW];
#pragma unroll
  for (int ii = 0; ii < RW; ++ii)
#pragma unroll
    for (int j = 0; j < CW; ++j) acc[ii][j] = cmake<T>(0, 0);
  for (int k = 0; k < kend; ++k) {
    cplx<T> av[RW], xv[CW];
#pragma unroll
    for (int ii = 0; ii < RW; ++ii) av[ii] = As[RW * rg + ii][k];
#pragma unroll
    for (int j = 0; j < CW; ++j) xv[j] = Xs[k][jb + j];
#pragma unroll
    for (int ii = 0; ii < RW; ++ii)
#pragma unroll
      for (int j = 0; j < CW; ++j) acc[ii][j] = cfma(acc[ii][j], av[ii], xv[j]);
  }
  int lev = 0;
#pragma unroll
  for (int ii = 0; ii < RW; ++ii) {
    const int r = RW * rg + ii;
    if (r < nr)
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const int c = jb + j;
        if (c < w) {
          cst(A + 2 * ((long)(r0 + r) * lda + c0 + c), acc[ii][j]);
          lev = max(lev, np_level(acc[ii][j]));
        }
      }
  }
  if (nx > 0) {
    // A22 -= L21 U12 on this workgroup's rows and the nx columns right of the
    // panel (U12 solved in place by k_lu_np_diag); every output summed over k
    // in order
    __syncthreads();  // every wave done with As and Xs
#pragma unroll
    for (int ii = 0; ii < RW; ++ii)
#pragma unroll
      for (int j = 0; j < CW; ++j) As[RW * rg + ii][jb + j] = acc[ii][j];
    for (int e = tid; e < W * W; e += 256) {
      const int k = e / W, c = e % W;
      Xs[k][c] = (k < w && c < nx) ? cld(A + 2 * ((long)(c0 + k) * lda + c0 + w + c)) : cmake<T>(0, 0);
    }
    __syncthreads();
    cplx<T> up[RW][CW];
#pragma unroll
    for (int ii = 0; ii < RW; ++ii)
#pragma unroll
      for (int j = 0; j < CW; ++j) up[ii][j] = cmake<T>(0, 0);
    for (int k = 0; k < w; ++k) {
      cplx<T> av[RW], xv[CW];
#pragma unroll
      for (int ii = 0; ii < RW; ++ii) av[ii] = As[RW * rg + ii][k];
#pragma unroll
      for (int j = 0; j < CW; ++j) xv[j] = Xs[k][jb + j];
#pragma unroll
      for (int ii = 0; ii < RW; ++ii)
#pragma unroll
        for (int j = 0; j < CW; ++j) up[ii][j] = cfma(up[ii][j], av[ii], xv[j]);
    }
#pragma unroll
    for (int ii = 0; ii < RW; ++ii) {
      const int r = RW * rg + ii;
      if (r < nr)
#pragma unroll
        for (int j = 0; j < CW; ++j) {
          const int c = jb + j;
          if (c < nx) {
            T* pa = A + 2 * ((long)(r0 + r) * lda + c0 + w + c);
            cst(pa, csub(cld(pa), up[ii][j]));
          }
        }
    }
  }
  if (lev) atomicMax(flag + b, lev);
}

__global__ void k_iota_rows(int n, int* ipiv, long sP) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) ipiv[b * sP + i] = i;
}

// recursive LU without interchanges of columns [c0, c0 + n) (rows c0..N-1)
template <typename T>
int getrf_np_rec(int N, int c0, int n, T* A, long lda, long sA, int* info, int* flag, int batch, hipStream_t st,
                 int gmode) {
  constexpr int W = NpW<T>::w;
  if (n <= 2 * W) {
    // a panel of up to W columns; with n > W the left panel of W columns
    // also solves and updates the nx = n - W columns of the right one
    const int w = n <= W ? n : W, nx = n - w;
    hipLaunchKernelGGL((k_lu_np_diag<T, W>), dim3(batch), dim3(256), 0, st, c0, w, A, lda, sA, info, flag, nx);
    MNP_LAUNCH_CHECK();
    if (N - c0 - w > 0) {
      dim3 grid(mnp_ceil_div(N - c0 - w, 64), batch);
      hipLaunchKernelGGL((k_lu_np_rows<T, W>), grid, dim3(256), 0, st, N, c0, w, A, lda, sA, flag, nx);
      MNP_LAUNCH_CHECK();
    }
    if (nx > 0) return getrf_np_rec(N, c0 + w, nx, A, lda, sA, info, flag, batch, st, gmode);
    return 0;
  }
  int n1 = ((n / 2 + W - 1) / W) * W;
  if (n1 >= n) n1 = n - W;
  const int n2 = n - n1;
  int rc;
  if ((rc = getrf_np_rec(N, c0, n1, A, lda, sA, info, flag, batch, st, gmode))) return rc;
  T* A11 = A + 2 * ((long)c0 * lda + c0);
  T* A12 = A + 2 * ((long)c0 * lda + c0 + n1);
  T* A21 = A + 2 * ((long)(c0 + n1) * lda + c0);
  T* A22 = A + 2 * ((long)(c0 + n1) * lda + c0 + n1);
  if ((rc = trsm_rec<T>(true, n1, n2, A11, lda, sA, A12, lda, sA, batch, st, gmode))) return rc;
  if ((rc = gemm_sub<T>(N - c0 - n1, n2, n1, A21, lda, sA, A12, lda, sA, A22, lda, sA, batch, st, gmode))) return rc;
  return getrf_np_rec(N, c0 + n1, n2, A, lda, sA, info, flag, batch, st, gmode);
}

// ---------------------------------------------------------------------------
// Inverse from the LU factors (LAPACK getri's operation count): with P A = L U,
// A^-1 = U^-1 L^-1 P.  L^-1 is formed by a recursion that keeps the zero
// upper triangle of the identity out of the work (n^3/6 complex MACs instead
// of the n^3/2 of a triangular solve against a full identity), the upper
// solve against it is dense (n^3/2), and P becomes a column scatter.
// ---------------------------------------------------------------------------
// X (n x n, identity on entry) <- L^-1, L unit lower triangular
template <typename T>
int trsm_lower_ident(int n, const T* L, long ldl, long sL, T* X, long ldx, long sX, int batch, hipStream_t st,
                     int gmode, LeafInv<T> inv = LeafInv<T>()) {
  constexpr int TL = TLeaf<T>::n;
  const int SB = (sizeof(T) == 4 && inv.D != nullptr && gmode == 1) ? inv.bs : TL;
  if (n <= 2 * SB) return trsm_rec<T>(true, n, n, L, ldl, sL, X, ldx, sX, batch, st, gmode, inv);
  int n1 = ((n / 2 + SB - 1) / SB) * SB;
  if (n1 >= n) n1 = n - SB;
  const int n2 = n - n1;
  int rc;
  if ((rc = trsm_lower_ident<T>(n1, L, ldl, sL, X, ldx, sX, batch, st, gmode, inv))) return rc;
  // X21 = -L22^-1 L21 X11 (X21 is zero on entry)
  if ((rc = gemm_sub<T>(n2, n1, n1, L + 2 * ((long)n1 * ldl), ldl, sL, X, ldx, sX, X + 2 * ((long)n1 * ldx), ldx,
                        sX, batch, st, gmode)))
    return rc;
  if ((rc = trsm_rec<T>(true, n2, n1, L + 2 * ((long)n1 * ldl + n1), ldl, sL, X + 2 * ((long)n1 * ldx), ldx, sX,
                        batch, st, gmode, inv.at(n1))))
    return rc;
  return trsm_lower_ident<T>(n2, L + 2 * ((long)n1 * ldl + n1), ldl, sL, X + 2 * ((long)n1 * ldx + n1), ldx, sX,
                             batch, st, gmode, inv.at(n1));
}

// W[b] = identity (n x n, leading dimension ld)
template <typename T>
__global__ void k_eye(int n, T* W, long ld, long sW) {
  const long b = blockIdx.y;
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= (long)n * n) return;
  const int r = (int)(e / n), c = (int)(e % n);
  cst(W + 2 * (b * sW + (long)r * ld + c), cmake<T>(r == c ? T(1) : T(0), T(0)));
}

// perm[b][j]: the row interchanges ipiv[0..n) applied in order to 0..n-1
// (P I = I[perm]); one workgroup per matrix, the swaps run in LDS (128 KB of
// the 160 KB per CU: n up to GETRI_MAXN, the row limit of the complex64 getrf).
constexpr int GETRI_MAXN = 32768;
__global__ __launch_bounds__(256) void k_perm_array(int n, const int* __restrict__ ipiv, long sP, int* perm,
                                                    long sQ) {
  __shared__ int arr[GETRI_MAXN];
  const long b = blockIdx.x;
  ipiv += b * sP;
  perm += b * sQ;
  for (int j = threadIdx.x; j < n; j += blockDim.x) arr[j] = j;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int i = 0; i < n; ++i) {
      const int p = ipiv[i];
      if (p != i) {
        const int t = arr[i];
        arr[i] = arr[p];
        arr[p] = t;
      }
    }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) perm[j] = arr[j];
}

// Out[r][perm[i]] = W[r][i]  (A^-1 = W P: column i of W lands in column perm[i])
template <typename T>
__global__ __launch_bounds__(256) void k_scatter_cols(int n, const T* __restrict__ W, long ldw, long sW,
                                                      const int* __restrict__ perm, long sQ, T* Out, long ldo,
                                                      long sO) {
  const long b = blockIdx.z;
  const int r = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int c = perm[b * sQ + i];
  cst(Out + 2 * (b * sO + (long)r * ldo + c), cld(W + 2 * (b * sW + (long)r * ldw + i)));
}

// A (holding its LU and pivots from getrf) <- A^-1; W: n x n work per matrix
// (the identity fill, L^-1 and U^-1 L^-1 live there); Q: n ints per matrix.
template <typename T>
int getri(int n, T* A, long lda, long sA, const int* ipiv, long sP, T* W, long ldw, long sW, int* Q, long sQ,
          int batch, hipStream_t st, int gmode, LeafInv<T> inv = LeafInv<T>()) {
  dim3 gi(mnp_ceil_div((long)n * n, 256), batch);
  hipLaunchKernelGGL(k_eye<T>, gi, dim3(256), 0, st, n, W, ldw, sW);
  MNP_LAUNCH_CHECK();
  int rc;
  if ((rc = trsm_lower_ident<T>(n, A, lda, sA, W, ldw, sW, batch, st, gmode, inv))) return rc;
  if ((rc = trsm_many<T>(false, n, n, A, lda, sA, W, ldw, sW, batch, st, gmode, inv))) return rc;
  hipLaunchKernelGGL(k_perm_array, dim3(batch), dim3(256), 0, st, n, ipiv, sP, Q, sQ);
  MNP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_scatter_cols<T>, dim3(mnp_ceil_div(n, 256), n, batch), dim3(256), 0, st, n, W, ldw, sW, Q,
                     sQ, A, lda, sA);
  MNP_LAUNCH_CHECK();
  return 0;
}

// W <- A^-1 from LU factors made without row interchanges (P = I): the
// inverse U^-1 L^-1 is formed in W and stays there (no permutation array,
// no scatter back into A); A is only read.
template <typename T>
int getri_np(int n, const T* A, long lda, long sA, T* W, long ldw, long sW, int batch, hipStream_t st, int gmode,
             LeafInv<T> inv = LeafInv<T>()) {
  dim3 gi(mnp_ceil_div((long)n * n, 256), batch);
  hipLaunchKernelGGL(k_eye<T>, gi, dim3(256), 0, st, n, W, ldw, sW);
  MNP_LAUNCH_CHECK();
  int rc;
  if ((rc = trsm_lower_ident<T>(n, A, lda, sA, W, ldw, sW, batch, st, gmode, inv))) return rc;
  return trsm_many<T>(false, n, n, A, lda, sA, W, ldw, sW, batch, st, gmode, inv);
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int mnpbem_cgemm_sub_batched(int dtype, int M, int N, int K, const void* A, long lda, long sA,
                                        const void* B, long ldb, long sB, void* C, long ldc, long sC, int batch,
                                        void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "cgemm: negative size");
  if (dtype == MNP_DTYPE_C128)
    return gemm_sub<double>(M, N, K, (const double*)A, lda, sA, (const double*)B, ldb, sB, (double*)C, ldc, sC,
                            batch, st, 0);
  if (mnp_is_c64(dtype))
    return gemm_sub<float>(M, N, K, (const float*)A, lda, sA, (const float*)B, ldb, sB, (float*)C, ldc, sC, batch,
                           st, dtype == MNP_DTYPE_C64X ? 1 : 0);
  mnp_set_error("cgemm: unknown dtype");
  return 1;
}

// C[b] -= A[b] B[b] for complex64 data on the bf16x3 kernel with a given tile
// variant (0: 128 x 128, 1: 64 x 64), bypassing the shape-class table: the
// autotuner times the variants with it.
extern "C" int mnpbem_cgemm_sub_variant(int variant, int M, int N, int K, const void* A, long lda, long sA,
                                        const void* B, long ldb, long sB, void* C, long ldc, long sC, int batch,
                                        void* stream) {
  MNP_REQUIRE(variant >= 0 && variant <= 2, "cgemm_variant: variant must be 0, 1 or 2");
  MNP_REQUIRE(M >= 0 && N >= 0 && K >= 0 && batch >= 0, "cgemm_variant: negative size");
  if (M == 0 || N == 0 || K == 0 || batch == 0) return 0;
  return launch_bf16x3(variant, M, N, K, (const float*)A, lda, sA, (const float*)B, ldb, sB, (float*)C, ldc, sC,
                       batch, (hipStream_t)stream);
}

// Rows above which getrf factors its panels with the multi-workgroup column
// kernels instead of the register-resident leaf (rows > 0), or the built-in
// thresholds (rows = 0: 12288 for complex64, 6144 for complex128).  Returns
// the previous setting.
// Columns (right-hand sides x batch) up to which the triangular-solve leaves
// (17 to 64 rows complex64, 17 to 32 complex128) use the row-blocked kernel; cols < 0
// restores the built-in 32768, 0 disables it.  Returns the previous setting.
extern "C" int mnpbem_trsm_rows_max(int cols) {
  const int prev = (int)g_trsm_rows_max;
  g_trsm_rows_max = cols < 0 ? TRSM_ROWS_DEFAULT : cols;
  return prev;
}

// Right-hand sides from which getrs / getri with 128-row inverse blocks solve
// block row by block row (left-looking); rhs < 0 restores the built-in 512, 0
// disables it (recursive solves only).  Returns the previous setting.
// 1 (built-in): the block rows of the left-looking solves after the first
// fold their inverse-block leaf into the update product's epilogue; 0: two
// launches per block row.  Returns the previous setting.
extern "C" int mnpbem_trsm_leaf_fused(int on) {
  const int prev = g_trsm_leaf_fused;
  g_trsm_leaf_fused = on ? 1 : 0;
  return prev;
}

// Widest product (N columns, K <= 64) that goes to the FP32 vector kernel
// k_cgemm_thin instead of the bf16x3 tiles: n in 0..32 (0: none), n < 0
// restores the built-in 16.  Returns the previous setting.
extern "C" int mnpbem_gemm_thin_max(int n) {
  const int prev = g_thin_n;
  g_thin_n = n < 0 ? THIN_N_DEFAULT : (n > THIN_N ? THIN_N : n);
  return prev;
}

extern "C" int mnpbem_trsm_left_rhs(int rhs) {
  const int prev = g_trsm_left_rhs;
  g_trsm_left_rhs = rhs < 0 ? TRSM_LEFT_DEFAULT : rhs;
  return prev;
}

extern "C" int mnpbem_getrf_tall_rows(int rows) {
  const int prev = g_tall_rows;
  g_tall_rows = rows > 0 ? rows : 0;
  return prev;
}

// Shape-class table of the bf16x3 GEMM (7 x 7 x 7 cells, index (cm * 7 + cn)
// * 7 + ck, class c of a dimension x: x <= 64 * 2^c, last class unbounded).
// set: table[343] of variants (0, 1 or 2); get: copies it out.
extern "C" int mnpbem_gemm_tuning_set(const unsigned char* table, int ncells) {
  MNP_REQUIRE(ncells == GC * GC * GC && table != nullptr, "gemm_tuning: expected 343 cells");
  for (int c = 0; c < ncells; ++c) MNP_REQUIRE(table[c] <= 2, "gemm_tuning: variants are 0, 1 or 2");
  for (int c = 0; c < ncells; ++c) g_tune.var[c] = table[c];
  return 0;
}
extern "C" int mnpbem_gemm_tuning_get(unsigned char* table, int ncells) {
  MNP_REQUIRE(ncells == GC * GC * GC && table != nullptr, "gemm_tuning: expected 343 cells");
  for (int c = 0; c < ncells; ++c) table[c] = g_tune.var[c];
  return 0;
}
// One launched shape per cell: shapes[4 c .. 4 c + 3] = M, N, K, batch (M = 0:
// no launch in that class since the last clear); clear != 0 resets the record.
extern "C" int mnpbem_gemm_tuning_seen(int* shapes, int ncells, int clear) {
  MNP_REQUIRE(ncells == GC * GC * GC, "gemm_tuning: expected 343 cells");
  for (int c = 0; c < ncells; ++c)
    for (int q = 0; q < 4; ++q) {
      if (shapes) shapes[4 * c + q] = g_tune.seen[c][q];
      if (clear) g_tune.seen[c][q] = 0;
    }
  return 0;
}

extern "C" int mnpbem_getrf_batched(int dtype, int n, void* A, long lda, long sA, int* ipiv, long sP, int* info,
                                    int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && batch >= 0 && lda >= n, "getrf: bad size");
  MNP_REQUIRE(dtype == MNP_DTYPE_C128 || mnp_is_c64(dtype), "getrf: unknown dtype");
  MNP_REQUIRE(dtype != MNP_DTYPE_C128 || leaf_width<double>(n) > 0,
              "getrf: complex128 matrices are limited to 20480 rows");
  MNP_REQUIRE(dtype == MNP_DTYPE_C128 || leaf_width<float>(n) > 0,
              "getrf: complex64 matrices are limited to 32768 rows");
  if (n == 0 || batch == 0) return 0;
  hipLaunchKernelGGL(k_fill_int, dim3(mnp_ceil_div(batch, 256)), dim3(256), 0, st, info, (long)batch, 0);
  MNP_LAUNCH_CHECK();
  if (dtype == MNP_DTYPE_C128)
    return getrf_rec<double>(n, 0, n, (double*)A, lda, sA, ipiv, sP, info, batch, st, 0, !tall_ok<double>(n, batch));
  if (mnp_is_c64(dtype))
    return getrf_rec<float>(n, 0, n, (float*)A, lda, sA, ipiv, sP, info, batch, st, dtype == MNP_DTYPE_C64X ? 1 : 0,
                            !tall_ok<float>(n, batch));
  mnp_set_error("getrf: unknown dtype");
  return 1;
}

// LU without row interchanges (ipiv = identity); flag[b] = 1 when partial
// pivoting would have interchanged rows of matrix b, 2 when a multiplier
// exceeds NP_GUARD in |re| + |im|
extern "C" int mnpbem_getrf_nopiv_batched(int dtype, int n, void* A, long lda, long sA, int* ipiv, long sP,
                                          int* info, int* flag, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && batch >= 0 && lda >= n && sP >= n, "getrf_nopiv: bad size");
  MNP_REQUIRE(dtype == MNP_DTYPE_C128 || mnp_is_c64(dtype), "getrf_nopiv: unknown dtype");
  if (n == 0 || batch == 0) return 0;
  hipLaunchKernelGGL(k_fill_int, dim3(mnp_ceil_div(batch, 256)), dim3(256), 0, st, info, (long)batch, 0);
  MNP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_fill_int, dim3(mnp_ceil_div(batch, 256)), dim3(256), 0, st, flag, (long)batch, 0);
  MNP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_iota_rows, dim3(mnp_ceil_div(n, 256), batch), dim3(256), 0, st, n, ipiv, sP);
  MNP_LAUNCH_CHECK();
  const long esz = dtype == MNP_DTYPE_C128 ? 16 : 8;
  for (int b0 = 0; b0 < batch; b0 += NP_MAXB) {
    const int nb = std::min(NP_MAXB, batch - b0);
    void* Ab = (char*)A + (long)b0 * sA * esz;
    int rc = dtype == MNP_DTYPE_C128
                 ? getrf_np_rec<double>(n, 0, n, (double*)Ab, lda, sA, info + b0, flag + b0, nb, st, 0)
                 : getrf_np_rec<float>(n, 0, n, (float*)Ab, lda, sA, info + b0, flag + b0, nb, st,
                                       dtype == MNP_DTYPE_C64X ? 1 : 0);
    if (rc) return rc;
  }
  return 0;
}

// B <- A^-1 B with A = P^T L U from mnpbem_getrf_batched (no transpose)
extern "C" int mnpbem_getrs_batched(int dtype, int n, int nrhs, const void* A, long lda, long sA, const int* ipiv,
                                    long sP, void* B, long ldb, long sB, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && nrhs >= 0 && batch >= 0 && lda >= n && ldb >= nrhs, "getrs: bad size");
  if (n == 0 || nrhs == 0 || batch == 0) return 0;
  int rc;
  if (dtype == MNP_DTYPE_C128) {
    if ((rc = laswp<double>(0, n, 0, nrhs, (double*)B, ldb, sB, ipiv, sP, batch, false, st))) return rc;
    if ((rc = trsm_rec<double>(true, n, nrhs, (const double*)A, lda, sA, (double*)B, ldb, sB, batch, st, 0)))
      return rc;
    return trsm_rec<double>(false, n, nrhs, (const double*)A, lda, sA, (double*)B, ldb, sB, batch, st, 0);
  }
  if (mnp_is_c64(dtype)) {
    const int gm = dtype == MNP_DTYPE_C64X ? 1 : 0;
    if ((rc = laswp<float>(0, n, 0, nrhs, (float*)B, ldb, sB, ipiv, sP, batch, false, st))) return rc;
    if ((rc = trsm_rec<float>(true, n, nrhs, (const float*)A, lda, sA, (float*)B, ldb, sB, batch, st, gm)))
      return rc;
    return trsm_rec<float>(false, n, nrhs, (const float*)A, lda, sA, (float*)B, ldb, sB, batch, st, gm);
  }
  mnp_set_error("getrs: unknown dtype");
  return 1;
}

// A <- A^-1 from its getrf factors (LU in A, pivots ipiv), i.e. LAPACK getri;
// W (n x n per matrix, leading dimension ldw) and Q (n ints per matrix) are
// work space.  Replaces the inverse that bem_ret.py forms with a solve against
// the identity.
extern "C" int mnpbem_getri_batched(int dtype, int n, void* A, long lda, long sA, const int* ipiv, long sP,
                                    void* W, long ldw, long sW, int* Q, long sQ, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && batch >= 0 && lda >= n && ldw >= n && sQ >= n, "getri: bad size");
  MNP_REQUIRE(n <= GETRI_MAXN, "getri: n above 32768");
  if (n == 0 || batch == 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    return getri<double>(n, (double*)A, lda, sA, ipiv, sP, (double*)W, ldw, sW, Q, sQ, batch, st, 0);
  if (mnp_is_c64(dtype))
    return getri<float>(n, (float*)A, lda, sA, ipiv, sP, (float*)W, ldw, sW, Q, sQ, batch, st,
                        dtype == MNP_DTYPE_C64X ? 1 : 0);
  mnp_set_error("getri: unknown dtype");
  return 1;
}

// getrs and getri with the inverse diagonal blocks of the factors (Dinv from
// mnpbem_lu_blockinv_batched on the same factors, per-matrix stride sD): for
// dtype MNPBEM_DTYPE_C64X the 64-row leaves of the triangular solves are
// products with those inverses on the bf16x3 kernel; other dtypes ignore Dinv.
extern "C" int mnpbem_getrs_blockinv_batched(int dtype, int n, int nrhs, const void* A, long lda, long sA,
                                             const int* ipiv, long sP, const void* Dinv, long sD, int bs, void* B,
                                             long ldb, long sB, int batch, void* stream) {
  if (dtype != MNP_DTYPE_C64X || Dinv == nullptr)
    return mnpbem_getrs_batched(dtype, n, nrhs, A, lda, sA, ipiv, sP, B, ldb, sB, batch, stream);
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && nrhs >= 0 && batch >= 0 && lda >= n && ldb >= nrhs, "getrs: bad size");
  MNP_REQUIRE(bs == 64 || bs == 128, "getrs: block size must be 64 or 128");
  MNP_REQUIRE(sD >= (long)((n + bs - 1) / bs) * 2 * bs * bs, "getrs: Dinv stride too small");
  if (n == 0 || nrhs == 0 || batch == 0) return 0;
  LeafInv<float> inv;
  inv.D = (const float*)Dinv;
  inv.sD = sD;
  inv.bs = bs;
  int rc;
  if ((rc = laswp<float>(0, n, 0, nrhs, (float*)B, ldb, sB, ipiv, sP, batch, false, st))) return rc;
  if ((rc = trsm_many<float>(true, n, nrhs, (const float*)A, lda, sA, (float*)B, ldb, sB, batch, st, 1, inv)))
    return rc;
  return trsm_many<float>(false, n, nrhs, (const float*)A, lda, sA, (float*)B, ldb, sB, batch, st, 1, inv);
}

extern "C" int mnpbem_getri_blockinv_batched(int dtype, int n, void* A, long lda, long sA, const int* ipiv, long sP,
                                             const void* Dinv, long sD, int bs, void* W, long ldw, long sW, int* Q,
                                             long sQ, int batch, void* stream) {
  if (dtype != MNP_DTYPE_C64X || Dinv == nullptr)
    return mnpbem_getri_batched(dtype, n, A, lda, sA, ipiv, sP, W, ldw, sW, Q, sQ, batch, stream);
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && batch >= 0 && lda >= n && ldw >= n && sQ >= n, "getri: bad size");
  MNP_REQUIRE(n <= GETRI_MAXN, "getri: n above 32768");
  MNP_REQUIRE(bs == 64 || bs == 128, "getri: block size must be 64 or 128");
  MNP_REQUIRE(sD >= (long)((n + bs - 1) / bs) * 2 * bs * bs, "getri: Dinv stride too small");
  if (n == 0 || batch == 0) return 0;
  LeafInv<float> inv;
  inv.D = (const float*)Dinv;
  inv.sD = sD;
  inv.bs = bs;
  return getri<float>(n, (float*)A, lda, sA, ipiv, sP, (float*)W, ldw, sW, Q, sQ, batch, st, 1, inv);
}

extern "C" int mnpbem_getri_nopiv_blockinv_batched(int dtype, int n, const void* A, long lda, long sA,
                                                   const void* Dinv, long sD, int bs, void* W, long ldw, long sW,
                                                   int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && batch >= 0 && lda >= n && ldw >= n, "getri_nopiv: bad size");
  if (n == 0 || batch == 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    return getri_np<double>(n, (const double*)A, lda, sA, (double*)W, ldw, sW, batch, st, 0);
  if (!mnp_is_c64(dtype)) {
    mnp_set_error("getri_nopiv: unknown dtype");
    return 1;
  }
  LeafInv<float> inv;
  if (dtype == MNP_DTYPE_C64X && Dinv != nullptr) {
    MNP_REQUIRE(bs == 64 || bs == 128, "getri_nopiv: block size must be 64 or 128");
    MNP_REQUIRE(sD >= (long)((n + bs - 1) / bs) * 2 * bs * bs, "getri_nopiv: Dinv stride too small");
    inv.D = (const float*)Dinv;
    inv.sD = sD;
    inv.bs = bs;
  }
  return getri_np<float>(n, (const float*)A, lda, sA, (float*)W, ldw, sW, batch, st,
                         dtype == MNP_DTYPE_C64X ? 1 : 0, inv);
}

// Inverses of the 128 x 128 diagonal blocks of LU factors, from those of their
// 64 x 64 halves (k_lu_blockinv in solve.hip) and the off-diagonal 64 x 64
// blocks of the factors:
//   L^-1 = [[L0^-1, 0], [-L1^-1 L10 L0^-1, L1^-1]],
//   U^-1 = [[U0^-1, -U0^-1 U01 U1^-1], [0, U1^-1]].
// Output per matrix and 128-row block: [U^-1 | L^-1], two 128 x 128 row-major
// blocks, zero outside the triangle (and beyond n).  One workgroup per
// (block, matrix); the products in FP32 through LDS, each output summed over
// k in order.
namespace {
__device__ __forceinline__ void mm64(const cplx<float> (&S1)[64][65], const cplx<float> (&S2)[64][65],
                                     cplx<float> (&acc)[16], int r, int cq) {
#pragma unroll
  for (int m = 0; m < 16; ++m) acc[m] = cmake<float>(0.f, 0.f);
  for (int k = 0; k < 64; ++k) {
    const cplx<float> a = S1[r][k];
#pragma unroll
    for (int m = 0; m < 16; ++m) acc[m] = cfma(acc[m], a, S2[k][cq + 4 * m]);
  }
}

__global__ __launch_bounds__(256) void k_blockinv128(int n, const float* __restrict__ LU, long ld, long sLU,
                                                     const float* __restrict__ D64, long sD64, float* D128,
                                                     long sD128) {
  const long b = blockIdx.y;
  const int p = blockIdx.x, i0 = 128 * p;
  const int nb = min(128, n - i0), n1 = max(0, nb - 64);  // rows of the block and of its second half
  LU += 2 * b * sLU;
  const float* d0 = D64 + 2 * (b * sD64 + (long)(2 * p) * 2 * 64 * 64);  // [U0^-1 | L0^-1]
  const float* d1 = d0 + 2 * 2 * 64 * 64;                                // [U1^-1 | L1^-1] (if n1 > 0)
  float* o = D128 + 2 * (b * sD128 + (long)p * 2 * 128 * 128);
  __shared__ cplx<float> S1[64][65], S2[64][65], S3[64][65];
  const int tid = threadIdx.x, r = tid >> 2, cq = tid & 3;
  // diagonal halves and the zero blocks (U: lower left, L: upper right)
  for (int tri = 0; tri < 2; ++tri)
    for (int e = tid; e < 128 * 128; e += 256) {
      const int rr = e / 128, cc = e % 128;
      const bool lo = rr >= 64, le = cc < 64;
      if ((tri == 0 && !lo && !le) || (tri == 1 && lo && le)) continue;  // computed below
      cplx<float> v = cmake<float>(0.f, 0.f);
      if (!lo && le)
        v = cld(d0 + 2 * ((long)tri * 64 * 64 + rr * 64 + cc));
      else if (lo && !le && n1 > 0)
        v = cld(d1 + 2 * ((long)tri * 64 * 64 + (rr - 64) * 64 + (cc - 64)));
      cst(o + 2 * ((long)tri * 128 * 128 + rr * 128 + cc), v);
    }
  cplx<float> acc[16];
  for (int tri = 0; tri < 2; ++tri) {
    // tri 1 (L): X = -L1^-1 (L10 L0^-1); tri 0 (U): Y = -U0^-1 (U01 U1^-1)
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
      const int rr = e / 64, cc = e % 64;
      if (tri == 1) {
        S1[rr][cc] = rr < n1 ? cld(LU + 2 * ((long)(i0 + 64 + rr) * ld + i0 + cc)) : cmake<float>(0.f, 0.f);
        S2[rr][cc] = cld(d0 + 2 * ((long)64 * 64 + rr * 64 + cc));
      } else {
        S1[rr][cc] = cc < n1 ? cld(LU + 2 * ((long)(i0 + rr) * ld + i0 + 64 + cc)) : cmake<float>(0.f, 0.f);
        S2[rr][cc] = n1 > 0 ? cld(d1 + 2 * ((long)rr * 64 + cc)) : cmake<float>(0.f, 0.f);
      }
    }
    __syncthreads();
    mm64(S1, S2, acc, r, cq);
#pragma unroll
    for (int m = 0; m < 16; ++m) S3[r][cq + 4 * m] = acc[m];
    __syncthreads();
    for (int e = tid; e < 64 * 64; e += 256) {
      const int rr = e / 64, cc = e % 64;
      if (tri == 1)
        S1[rr][cc] = n1 > 0 ? cld(d1 + 2 * ((long)64 * 64 + rr * 64 + cc)) : cmake<float>(0.f, 0.f);
      else
        S1[rr][cc] = cld(d0 + 2 * ((long)rr * 64 + cc));
    }
    __syncthreads();
    mm64(S1, S3, acc, r, cq);
    const int ro = tri == 1 ? 64 : 0, co = tri == 1 ? 0 : 64;
#pragma unroll
    for (int m = 0; m < 16; ++m)
      cst(o + 2 * ((long)tri * 128 * 128 + (ro + r) * 128 + co + cq + 4 * m),
          cmake<float>(-acc[m].re, -acc[m].im));
  }
}
}  // namespace

extern "C" int mnpbem_lu_blockinv128_batched(int dtype, int n, const void* LU, long ld, long sLU, const void* D64,
                                             long sD64, void* D128, long sD128, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(mnp_is_c64(dtype), "lu_blockinv128: complex64 factors only");
  if (n <= 0 || batch <= 0) return 0;
  MNP_REQUIRE(sD64 >= (long)((n + 63) / 64) * 2 * 64 * 64, "lu_blockinv128: D64 stride too small");
  MNP_REQUIRE(sD128 >= (long)((n + 127) / 128) * 2 * 128 * 128, "lu_blockinv128: D128 stride too small");
  hipLaunchKernelGGL(k_blockinv128, dim3((n + 127) / 128, batch), dim3(256), 0, st, n, (const float*)LU, ld, sLU,
                     (const float*)D64, sD64, (float*)D128, sD128);
  MNP_LAUNCH_CHECK();
  return 0;
}

// Start timing complex-GEMM launches (up to ``capacity`` launches).
extern "C" int mnpbem_prof_begin(int capacity) {
  MNP_REQUIRE(capacity > 0, "prof_begin: capacity must be positive");
  if ((int)g_prof.ev.size() < 2 * capacity) {
    for (size_t i = g_prof.ev.size(); i < (size_t)2 * capacity; ++i) {
      hipEvent_t e;
      MNP_HIP_CHECK(hipEventCreate(&e));
      g_prof.ev.push_back(e);
    }
  }
  g_prof.used = 0;
  g_prof.flops.clear();
  g_prof.shape.clear();
  g_prof.bytes = 0.0;
  g_prof.cbytes = 0.0;
  g_prof.on = true;
  return 0;
}

// Stop timing; after the stream work has completed, report the summed
// kernel time (ms), summed algorithmic flops and the number of launches.
extern "C" int mnpbem_prof_end(double* total_ms, double* total_flops, long* launches) {
  g_prof.on = false;
  double ms = 0.0, fl = 0.0;
  for (size_t i = 0; i + 1 < g_prof.used; i += 2) {
    MNP_HIP_CHECK(hipEventSynchronize(g_prof.ev[i + 1]));
    float t = 0.f;
    MNP_HIP_CHECK(hipEventElapsedTime(&t, g_prof.ev[i], g_prof.ev[i + 1]));
    ms += t;
  }
  for (double f : g_prof.flops) fl += f;
  *total_ms = ms;
  *total_flops = fl;
  *launches = (long)(g_prof.used / 2);
  return 0;
}

// Compulsory HBM bytes of the launches timed since mnpbem_prof_begin (each
// operand read once, C read and written once).
extern "C" int mnpbem_prof_bytes(double* total_bytes) {
  *total_bytes = g_prof.bytes;
  return 0;
}

// Per-launch record of the launches timed since mnpbem_prof_begin (after
// mnpbem_prof_end): shapes[5 i ..] = M, N, K, batch, kernel (bf16x3 tiles 0:
// 128 x 128, 1: 64 x 64, 2: 128 x 128 with C prefetch; 3: the FP32/FP64
// kernel; 4: in-place 128-row product; 5: block-row update with its leaf)
// and ms[i] its HIP-event time, for up
// to cap launches;
// returns the number of launches recorded.
extern "C" int mnpbem_prof_launches(int cap, int* shapes, float* ms) {
  MNP_REQUIRE(cap >= 0, "prof_launches: negative capacity");
  const int nl = (int)(g_prof.used / 2);
  const int m = nl < cap ? nl : cap;
  for (int i = 0; i < m; ++i) {
    MNP_HIP_CHECK(hipEventSynchronize(g_prof.ev[2 * i + 1]));
    float t = 0.f;
    MNP_HIP_CHECK(hipEventElapsedTime(&t, g_prof.ev[2 * i], g_prof.ev[2 * i + 1]));
    ms[i] = t;
    for (int q = 0; q < 5; ++q) shapes[5 * i + q] = g_prof.shape[5 * i + q];
  }
  return m;
}

// The part of those bytes that writes C (each output element once).
extern "C" int mnpbem_prof_cbytes(double* c_write_bytes) {
  *c_write_bytes = g_prof.cbytes;
  return 0;
}
