// Hierarchical (H-) matrices of the Green functions: adaptive cross
// approximation of the admissible blocks, dense leaf blocks, block matvec.
//
// Reference: mnpbem/greenfun/hmatrix.py (HMatrix.aca :159-193, _aca_block
// :195-336, mtimes_vec :450-541, full :338-448), mnpbem/greenfun/
// aca_compgreen_ret.py:96-208 (ACA entries are entries of the assembled
// Green matrix), mnpbem/greenfun/clustertree.py (cluster ordering).
//
// Layout.  The cluster tree orders the faces: cluster position c holds face
// perm[c].  A block covers cluster rows [r0, r0 + m) and columns [c0, c0 + n).
// Block table `blk` (int64, 8 per block):
//   {r0, m, c0, n, rank, off_a, off_b, off_t}
// rank < 0 marks a dense block whose values sit column-major at val + off_a
// (val[off_a + j m + i] = M[perm[r0+i], perm[c0+j]]); rank >= 0 marks a
// low-rank block M ~ U V^T with U (m x rank) at U + off_a and V (n x rank) at
// V + off_b, both column-major; off_t is the running sum of ranks (the block's
// row offset in the matvec workspace).  Leaf table `leafs` (int64, 4 per leaf
// cluster): {r0, m, first, last} -> lst[first:last] lists the blocks whose
// row range contains the leaf (dense blocks of the leaf row and every
// low-rank block over it).  Values are complex64 (dtype 1) or complex128
// (dtype 2); vectors and all arithmetic are complex128.
#include "common.h"
#include "green_entry.h"

namespace {

constexpr int ACA_KMAX = 512;      // largest rank bound an ACA launch accepts
constexpr int ACA_MAXMN = 65536;   // largest block side (bitmask capacity)

struct Blk {
  long r0, m, c0, n, rank, oa, ob, ot;
};
__device__ __forceinline__ Blk load_blk(const long* blk, long b) {
  const long* q = blk + 8 * b;
  Blk x;
  x.r0 = q[0];
  x.m = q[1];
  x.c0 = q[2];
  x.n = q[3];
  x.rank = q[4];
  x.oa = q[5];
  x.ob = q[6];
  x.ot = q[7];
  return x;
}

__device__ __forceinline__ cplx<double> ld2(const double* p) {
  const double2 v = *reinterpret_cast<const double2*>(p);
  return cmake(v.x, v.y);
}
template <typename T>
__device__ __forceinline__ cplx<double> ldw(const T* p) {
  const cplx<T> v = cld(p);
  return cmake((double)v.re, (double)v.im);
}

// Entry sources of the ACA and dense-block kernels: M[i][j] for face indices
// (i, j).  DenseSrc reads an assembled complex128 matrix with element strides
// (si, sj); GreenSrc evaluates the refined Green function of one side in
// place (G for kind 0, F + hsign 2 pi [i == j] for kind 1), so no n x n
// matrix exists (the approach of mex/acagreen/acagreen.cpp:79-127, which
// fills ACA rows and columns from positions, normals and areas).
struct DenseSrc {
  const double* m;
  long si, sj;
  __device__ __forceinline__ cplx<double> operator()(int i, int j) const {
    const double2 v = *reinterpret_cast<const double2*>(m + 2 * ((long)i * si + (long)j * sj));
    return cmake(v.x, v.y);
  }
};
struct GreenSrc {
  int kind;
  double hsign;
  const double *pos, *nvec, *area;
  const int* mat;
  const double* kmat;
  const long* rowptr;
  const int* pcol;
  const double *g, *f;
  int ord1;
  __device__ __forceinline__ cplx<double> operator()(int i, int j) const {
    return green_value(kind, hsign, i, j, pos, nvec, area, mat, kmat, rowptr, pcol, g, f, ord1);
  }
};

// (|v|, index, v) candidate for the pivot searches; ties keep the lower index
struct Cand {
  double a;
  int idx;
  double re, im;
};
__device__ __forceinline__ bool better(double a, int i, double b, int j) { return a > b || (a == b && i < j); }

__device__ Cand block_argmax(Cand c, Cand* sh) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double oa = __shfl_xor(c.a, off);
    const int oi = __shfl_xor(c.idx, off);
    const double ore = __shfl_xor(c.re, off);
    const double oim = __shfl_xor(c.im, off);
    if (better(oa, oi, c.a, c.idx)) {
      c.a = oa;
      c.idx = oi;
      c.re = ore;
      c.im = oim;
    }
  }
  if (lane == 0) sh[w] = c;
  __syncthreads();
  Cand r = sh[0];
  for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
    if (better(sh[q].a, sh[q].idx, r.a, r.idx)) r = sh[q];
  __syncthreads();
  return r;
}

__device__ __forceinline__ cplx<double> wave_sum(cplx<double> v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v.re += __shfl_xor(v.re, off);
    v.im += __shfl_xor(v.im, off);
  }
  return v;
}

// Partially pivoted ACA of one admissible block per workgroup, FP64
// (hmatrix.py:195-336): pivot row p (first 0), residual row, pivot column =
// argmax |row| with used columns read as 0, stop when |pivot| < 1e-15 or
// ||u_k|| ||v_k|| < htol ||sum_r u_r v_r^T||_F (norm updated with the cross
// terms 2 Re sum_s (u_s^H u_k)(v_s^H v_k)), next row = argmax |u_k| with
// used rows read as 0.  Entries M[i][j] come from the source functor (an
// assembled matrix or the Green function evaluated in place).  U, V go to a
// workspace with per-block offsets woff[2 b], woff[2 b + 1] and column
// stride m (n).
template <class SRC>
__global__ __launch_bounds__(256) void k_aca(const long* blk, const int* bids, const int* perm, SRC src, double* Uw,
                                             double* Vw, const long* woff, int kmax, double htol, int* rank_out) {
  const int tid = threadIdx.x;
  const int lb = blockIdx.x;
  const Blk B = load_blk(blk, bids[lb]);
  if (B.m > ACA_MAXMN || B.n > ACA_MAXMN || B.m < 1 || B.n < 1) {  // outside the bitmask capacity
    if (threadIdx.x == 0) rank_out[lb] = -1;
    return;
  }
  const int m = (int)B.m, n = (int)B.n;
  double* U = Uw + 2 * woff[2 * lb];
  double* V = Vw + 2 * woff[2 * lb + 1];
  const int* prow = perm + B.r0;
  const int* pcol = perm + B.c0;
  int maxr = m < n ? m : n;
  if (kmax < maxr) maxr = kmax;

  __shared__ unsigned usedr[ACA_MAXMN / 32], usedc[ACA_MAXMN / 32];
  __shared__ double uk[2 * ACA_KMAX], vq[2 * ACA_KMAX];
  __shared__ double du[2 * ACA_KMAX], dv[2 * ACA_KMAX];
  __shared__ Cand sh[4];
  __shared__ int s_stop;
  for (int w = tid; w < (m + 31) / 32; w += blockDim.x) usedr[w] = 0u;
  for (int w = tid; w < (n + 31) / 32; w += blockDim.x) usedc[w] = 0u;
  double frob = 0.0;  // meaningful in thread 0
  int p = 0, rank = 0;
  __syncthreads();

  for (int k = 0; k < maxr; ++k) {
    for (int r = tid; r < k; r += blockDim.x) {
      const cplx<double> z = ld2(U + 2 * ((long)r * m + p));
      uk[2 * r] = z.re;
      uk[2 * r + 1] = z.im;
    }
    __syncthreads();
    // residual row p -> V[:, k]
    const int fi = prow[p];
    Cand c;
    c.a = -1.0;
    c.idx = n;
    c.re = c.im = 0.0;
    for (int j = tid; j < n; j += blockDim.x) {
      cplx<double> v = src(fi, pcol[j]);
      for (int r = 0; r < k; ++r) v = cfms(v, cmake(uk[2 * r], uk[2 * r + 1]), ld2(V + 2 * ((long)r * n + j)));
      double* o = V + 2 * ((long)k * n + j);
      o[0] = v.re;
      o[1] = v.im;
      const double a = ((usedc[j >> 5] >> (j & 31)) & 1u) ? 0.0 : hypot(v.re, v.im);
      if (better(a, j, c.a, c.idx)) {
        c.a = a;
        c.idx = j;
        c.re = v.re;
        c.im = v.im;
      }
    }
    const Cand pc = block_argmax(c, sh);
    if (hypot(pc.re, pc.im) < 1e-15) break;  // uniform across the workgroup
    const int q = pc.idx;
    for (int r = tid; r < k; r += blockDim.x) {
      const cplx<double> z = ld2(V + 2 * ((long)r * n + q));
      vq[2 * r] = z.re;
      vq[2 * r + 1] = z.im;
    }
    if (tid == 0) {
      usedr[p >> 5] |= 1u << (p & 31);
      usedc[q >> 5] |= 1u << (q & 31);
    }
    __syncthreads();
    // residual column q / pivot -> U[:, k]
    const cplx<double> inv = crecip(cmake(pc.re, pc.im));
    const int fj = pcol[q];
    Cand cr;
    cr.a = -1.0;
    cr.idx = m;
    cr.re = cr.im = 0.0;
    for (int i = tid; i < m; i += blockDim.x) {
      cplx<double> v = src(prow[i], fj);
      for (int r = 0; r < k; ++r) v = cfms(v, ld2(U + 2 * ((long)r * m + i)), cmake(vq[2 * r], vq[2 * r + 1]));
      const cplx<double> u = cmul(v, inv);
      double* o = U + 2 * ((long)k * m + i);
      o[0] = u.re;
      o[1] = u.im;
      const double a = ((usedr[i >> 5] >> (i & 31)) & 1u) ? 0.0 : hypot(u.re, u.im);
      if (better(a, i, cr.a, cr.idx)) {
        cr.a = a;
        cr.idx = i;
      }
    }
    const Cand pr = block_argmax(cr, sh);  // its barrier orders the U[:, k] writes before the dots
    rank = k + 1;
    // dots u_s^H u_k and v_s^H v_k for s <= k, one wave per s
    const int lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
    for (int s = wv; s <= k; s += nw) {
      cplx<double> a = cmake(0.0, 0.0), b = cmake(0.0, 0.0);
      for (int i = lane; i < m; i += 64) {
        const cplx<double> x = ld2(U + 2 * ((long)s * m + i)), y = ld2(U + 2 * ((long)k * m + i));
        a.re += x.re * y.re + x.im * y.im;
        a.im += x.re * y.im - x.im * y.re;
      }
      for (int j = lane; j < n; j += 64) {
        const cplx<double> x = ld2(V + 2 * ((long)s * n + j)), y = ld2(V + 2 * ((long)k * n + j));
        b.re += x.re * y.re + x.im * y.im;
        b.im += x.re * y.im - x.im * y.re;
      }
      a = wave_sum(a);
      b = wave_sum(b);
      if (lane == 0) {
        du[2 * s] = a.re;
        du[2 * s + 1] = a.im;
        dv[2 * s] = b.re;
        dv[2 * s + 1] = b.im;
      }
    }
    __syncthreads();
    if (tid == 0) {
      const double nt = du[2 * k] * dv[2 * k];
      double cross = 0.0;
      for (int s = 0; s < k; ++s) cross += du[2 * s] * dv[2 * s] - du[2 * s + 1] * dv[2 * s + 1];
      frob += nt + 2.0 * cross;
      s_stop = (frob > 0.0 && sqrt(nt) < htol * sqrt(fabs(frob))) ? 1 : 0;
    }
    __syncthreads();
    if (s_stop) break;
    p = pr.idx;
  }
  if (tid == 0) rank_out[lb] = rank;
}

// dense blocks: val[off_a + j m + i] = M[perm[r0+i]][perm[c0+j]]
template <typename TS, class SRC>
__global__ __launch_bounds__(256) void k_hmat_dense(const long* blk, const int* bids, const int* perm, SRC src,
                                                    TS* val) {
  const Blk B = load_blk(blk, bids[blockIdx.x]);
  const long mn = B.m * B.n;
  for (long e = threadIdx.x; e < mn; e += blockDim.x) {
    const long j = e / B.m, i = e - j * B.m;
    const cplx<double> v = src(perm[B.r0 + i], perm[B.c0 + j]);
    cst(val + 2 * (B.oa + e), cmake<TS>((TS)v.re, (TS)v.im));
  }
}

// low-rank blocks: copy the first `rank` columns of the ACA workspace into
// the packed factors (rank recorded in blk)
template <typename TS>
__global__ __launch_bounds__(256) void k_hmat_compact(const long* blk, const int* bids, const double* Uw,
                                                      const double* Vw, const long* woff, TS* U, TS* V) {
  const int lb = blockIdx.x;
  const Blk B = load_blk(blk, bids[lb]);
  const double* u = Uw + 2 * woff[2 * lb];
  const double* v = Vw + 2 * woff[2 * lb + 1];
  const long nu = B.m * B.rank, nv = B.n * B.rank;
  for (long e = threadIdx.x; e < nu; e += blockDim.x) cst(U + 2 * (B.oa + e), cmake<TS>((TS)u[2 * e], (TS)u[2 * e + 1]));
  for (long e = threadIdx.x; e < nv; e += blockDim.x) cst(V + 2 * (B.ob + e), cmake<TS>((TS)v[2 * e], (TS)v[2 * e + 1]));
}

// matvec phase 0: XW = X in cluster order (row c of XW = row perm[c] of X),
// so that the block products below read their columns contiguously
__global__ __launch_bounds__(256) void k_hmat_xperm(int n, int ncol, const int* perm, const double* X, long ldx,
                                                    double* XW) {
  const long e = (long)blockIdx.x * 256 + threadIdx.x;
  if (e >= (long)n * ncol) return;
  const long c = e / ncol;
  const int q = (int)(e % ncol);
  const double* x = X + 2 * ((long)perm[c] * ldx + q);
  XW[2 * e] = x[0];
  XW[2 * e + 1] = x[1];
}

// matvec phase 1: t_b = V_b^T x[cols of b] for every low-rank block (one
// wave per rank column, four column loads in flight per lane; the x rows of
// the block are contiguous in XW and come from L2 for the later columns)
template <typename TS, int NC>
__global__ __launch_bounds__(256) void k_hmat_vt(const long* blk, const TS* Vv, const double* XW, int ncol,
                                                 double* T) {
  const Blk B = load_blk(blk, blockIdx.x);
  if (B.rank <= 0) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double* xb = XW + 2 * (B.c0 * (long)ncol);
  for (long r = wv; r < B.rank; r += 4) {
    cplx<double> acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = cmake(0.0, 0.0);
    const TS* v = Vv + 2 * (B.ob + r * B.n);
    long j = lane;
    for (; j + 192 < B.n; j += 256) {
      cplx<double> a[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] = ldw(v + 2 * (j + 64 * u));
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const double* x = xb + 2 * ((j + 64 * u) * ncol);
#pragma unroll
        for (int c = 0; c < NC; ++c)
          if (c < ncol) acc[c] = cfma(acc[c], a[u], ld2(x + 2 * c));
      }
    }
    for (; j < B.n; j += 64) {
      const cplx<double> a = ldw(v + 2 * j);
      const double* x = xb + 2 * (j * ncol);
#pragma unroll
      for (int c = 0; c < NC; ++c)
        if (c < ncol) acc[c] = cfma(acc[c], a, ld2(x + 2 * c));
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      if (c < ncol) {
        const cplx<double> s = wave_sum(acc[c]);
        if (lane == 0) {
          double* o = T + 2 * ((B.ot + r) * ncol + c);
          o[0] = s.re;
          o[1] = s.im;
        }
      }
    }
  }
}

// matvec phase 2: one workgroup per (leaf row cluster, 256-row chunk); the
// waves share the leaf's rows in groups of 64 and split its block list; the
// group's partial sums meet in LDS and row i of the leaf is written once, to
// Y[perm[r0 + i]] (deterministic, no atomics).  Per block a wave stages the
// right-hand side it multiplies (the block's x rows, or its t slice) in LDS,
// 64 rows at a time, and keeps 16 matrix loads per lane in flight.
template <int NC>
struct RowsCfg {
  static constexpr int nw = 4;  // waves per workgroup (8 measured no faster)
};

template <typename TS, int NC>
__global__ __launch_bounds__(64 * RowsCfg<NC>::nw) void k_hmat_rows(const long* leafs, const int* lst,
                                                                    const long* blk, const int* perm, const TS* val,
                                                                    const TS* Uv, const double* T, const double* XW,
                                                                    int ncol, double* Y, long ldy, int accumulate) {
  constexpr int NW = RowsCfg<NC>::nw;
  const long* lf = leafs + 4 * blockIdx.x;
  const long lr0 = lf[0], lm = lf[1], first = lf[2], last = lf[3];
  const long i0 = (long)blockIdx.y * 256;
  if (i0 >= lm) return;  // uniform
  const int mc = (int)(lm - i0 < 256 ? lm - i0 : 256);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int RG = (mc + 63) / 64;  // row groups of 64
  const int WPG = NW / RG;        // waves per row group
  const int g = wv / WPG, sub = wv % WPG;
  const bool active = g < RG;
  const long li = i0 + (long)g * 64 + lane;  // row within the leaf
  const bool rowok = active && (g * 64 + lane) < mc;
  __shared__ cplx<double> rhs[NW][64][NC];  // per wave: the right-hand side rows of the current block
  cplx<double> acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = cmake(0.0, 0.0);
  if (active) {
    for (long e = first + sub; e < last; e += WPG) {
      const Blk B = load_blk(blk, lst[e]);
      const bool dense = B.rank < 0;
      // matrix column q of the block: dense values val[oa + q m + row], or U[oa + q m + row]
      const long row = dense ? li : lr0 + li - B.r0;
      const TS* a = (dense ? val : Uv) + 2 * (B.oa + row);
      const long nq = dense ? B.n : B.rank;
      for (long q0 = 0; q0 < nq; q0 += 64) {
        const int nb = (int)(nq - q0 < 64 ? nq - q0 : 64);
        // stage: x rows of columns c0 + q0.. (dense) or t rows ot + q0.. (low rank)
        const double* src = dense ? XW + 2 * ((B.c0 + q0) * (long)ncol) : T + 2 * ((B.ot + q0) * (long)ncol);
        if (lane < nb) {
#pragma unroll
          for (int c = 0; c < NC; ++c)
            rhs[wv][lane][c] = c < ncol ? ld2(src + 2 * ((long)lane * ncol + c)) : cmake(0.0, 0.0);
        }
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        int q = 0;
        for (; q + 16 <= nb; q += 16) {
          cplx<double> v[16];
#pragma unroll
          for (int u = 0; u < 16; ++u) v[u] = rowok ? ldw(a + 2 * (q0 + q + u) * B.m) : cmake(0.0, 0.0);
#pragma unroll
          for (int u = 0; u < 16; ++u)
#pragma unroll
            for (int c = 0; c < NC; ++c)
              if (c < ncol) acc[c] = cfma(acc[c], v[u], rhs[wv][q + u][c]);
        }
        for (; q < nb; ++q) {
          const cplx<double> v = rowok ? ldw(a + 2 * (q0 + q) * B.m) : cmake(0.0, 0.0);
#pragma unroll
          for (int c = 0; c < NC; ++c)
            if (c < ncol) acc[c] = cfma(acc[c], v, rhs[wv][q][c]);
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
    }
  }
  __shared__ double red[NW][64][NC][2];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    red[wv][lane][c][0] = acc[c].re;
    red[wv][lane][c][1] = acc[c].im;
  }
  __syncthreads();
  if (rowok && sub == 0) {
    double* y = Y + 2 * ((long)perm[lr0 + li] * ldy);
    for (int c = 0; c < ncol; ++c) {
      double sr = 0.0, si = 0.0;
      for (int q = 0; q < WPG; ++q) {
        sr += red[wv + q][lane][c][0];
        si += red[wv + q][lane][c][1];
      }
      if (accumulate) {
        sr += y[2 * c];
        si += y[2 * c + 1];
      }
      y[2 * c] = sr;
      y[2 * c + 1] = si;
    }
  }
}

// Packed product layouts.  The block storage above serves the block
// operations (densify, recompression, sums); for products the blocks are
// also copied once into two packed layouts, so that each phase of a product
// streams long contiguous panels instead of walking 50-100 blocks per leaf
// with a descriptor, a right-hand-side fetch and a few column loads each:
//  * column panels (phase 1): the low-rank blocks that share a column
//    cluster [c0, c0 + n) with their V factors side by side, transposed:
//    VS[vs + j K + o] = V_b[j][r] for the panel's output o = (block b, rank
//    column r), o < K; output o is row tix[og + o] of the workspace R.
//  * row panels (phase 2): per leaf row cluster and 64-row group, the rows of
//    every block over it side by side, column-major with leading dimension
//    mi: P[p + q mi + i]; column q multiplies row gidx[g + q] of R.
// R: (n + sum of ranks) x ncol complex128 rows: X in cluster order in rows
// [0, n), the t = V^T x values of block b in rows n + off_t + r; allocated
// with 16 columns per row, so that the kernels may read a full bucket of
// columns past the last row.
//
// Pack pieces.  Column piece (one per low-rank block, 8 int64):
//   {dst, K, n, ob, rank, tdst, trow, 0}:
//   VS[dst + j K + r] = V[ob + r n + j], tix[tdst + r] = trow + r.
// Row piece (one per row group and block over it, 8 int64):
//   {dst, mi, ncols, src, ld, dense, gdst, gstart}:
//   P[dst + q mi + i] = (dense ? val : U)[src + q ld + i],
//   gidx[gdst + q] = gstart + q.
template <typename TS>
__global__ __launch_bounds__(256) void k_hmat_pack_cols(const long* desc, const TS* V, TS* VS, int* tix) {
  const long* d = desc + 8 * (long)blockIdx.x;
  const long dst = d[0], K = d[1], n = d[2], ob = d[3], rank = d[4], tdst = d[5], trow = d[6];
  for (long e = threadIdx.x; e < n * rank; e += blockDim.x) {
    const long r = e / n, j = e - r * n;
    cst(VS + 2 * (dst + j * K + r), cld(V + 2 * (ob + e)));
  }
  for (long r = threadIdx.x; r < rank; r += blockDim.x) tix[tdst + r] = (int)(trow + r);
}

template <typename TS>
__global__ __launch_bounds__(256) void k_hmat_pack_rows(const long* desc, const TS* val, const TS* U, TS* P,
                                                        int* gidx) {
  const long* d = desc + 8 * (long)blockIdx.x;
  const long dst = d[0], mi = d[1], nc = d[2], src = d[3], ld = d[4], dense = d[5], gdst = d[6], gstart = d[7];
  const TS* S = dense ? val : U;
  for (long e = threadIdx.x; e < nc * mi; e += blockDim.x) {
    const long q = e / mi, i = e - q * mi;
    cst(P + 2 * (dst + e), cld(S + 2 * (src + q * ld + i)));
  }
  for (long q = threadIdx.x; q < nc; q += blockDim.x) gidx[gdst + q] = (int)(gstart + q);
}

// Sum of the NW waves' accumulators for column c, in wave order, into
// wave 0 (deterministic); red is NW x 64 complex128 of LDS.
template <int NW>
__device__ __forceinline__ cplx<double> wave_order_sum(cplx<double> (*red)[64], cplx<double> a, int wv, int lane) {
  red[wv][lane] = a;
  __syncthreads();
  cplx<double> s = red[0][lane];
#pragma unroll
  for (int w = 1; w < NW; ++w) s = cadd(s, red[w][lane]);
  __syncthreads();
  return s;
}

// phase 1 on column panels: one workgroup per (panel, 64 outputs); lane =
// output (every load of VS is a coalesced row segment), the NW waves take
// consecutive parts of the panel's n rows, x rows of R by wave-uniform loads,
// 8 panel loads in flight per lane; writes the outputs' t rows of R.
template <typename TS, int NC, int NW, int UNR>
__global__ __launch_bounds__(64 * NW) void k_hmat_vt_pk(const long* items, const TS* __restrict__ VS,
                                                        const int* __restrict__ tix, const double* __restrict__ RX,
                                                        int ncol, double* __restrict__ RT) {
  const long* it = items + 6 * (long)blockIdx.x;
  const long vs = it[0], K = it[1], o0 = it[2], n = it[3], c0 = it[4], og = it[5];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const long o = o0 + lane;
  const long oc = o < K ? o : K - 1;
  const long q = ((n + NW - 1) / NW + UNR - 1) / UNR * UNR;
  const long jb = wv * q, je = jb + q < n ? jb + q : n;
  const double* xb = RX + 2 * c0 * ncol;
  cplx<double> acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = cmake(0.0, 0.0);
  // columns c >= ncol read the next row's values (in bounds: R holds 16
  // columns per row) into accumulators that are never written
  auto xval = [&](long j, int c) { return ld2(xb + 2 * (j * ncol + c)); };
  long j = jb;
  for (; j + UNR <= je; j += UNR) {
    cplx<double> m[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) m[u] = ldw(VS + 2 * (vs + (j + u) * K + oc));
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = cfma(acc[c], m[u], xval(j + u, c));
  }
  for (; j < je; ++j) {
    const cplx<double> m = ldw(VS + 2 * (vs + j * K + oc));
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = cfma(acc[c], m, xval(j, c));
  }
  __shared__ cplx<double> red[NW][64];
  const long row = o < K ? tix[og + o] : 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c < ncol) {
      const cplx<double> s = wave_order_sum<NW>(red, acc[c], wv, lane);
      if (wv == 0 && o < K) {
        double* t = RT + 2 * (row * ncol + c);
        t[0] = s.re;
        t[1] = s.im;
      }
    }
  }
}

// phase 2 on row panels: one workgroup per (leaf row cluster, 64-row
// group); lane = row, the NW waves take consecutive parts of the panel's K
// columns; the right-hand-side row of column q (gidx, then the row of R) by
// wave-uniform loads, 8 panel loads in flight per lane.  Row i of the group
// is written once, to Y[perm[r0 + i]] (no atomics, deterministic).
template <typename TS, int NC, int NW, int UNR>
__global__ __launch_bounds__(64 * NW) void k_hmat_rows_pk(const long* items, const TS* __restrict__ P,
                                                          const int* __restrict__ gidx, const int* perm,
                                                          const double* __restrict__ R, int ncol,
                                                          double* __restrict__ Y, long ldy, int accumulate) {
  const long* it = items + 6 * (long)blockIdx.x;
  const long p = it[0], mi = it[1], K = it[2], g = it[3], r0 = it[4];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
  const long ic = lane < mi ? lane : mi - 1;
  const long q = ((K + NW - 1) / NW + UNR - 1) / UNR * UNR;
  const long qb = wv * q, qe = qb + q < K ? qb + q : K;
  const int* gx = gidx + g;
  cplx<double> acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = cmake(0.0, 0.0);
  // unconditional loads as in k_hmat_vt_pk (R holds 16 columns per row)
  auto xval = [&](long qq, int c) { return ld2(R + 2 * ((long)gx[qq] * ncol + c)); };
  long k = qb;
  for (; k + UNR <= qe; k += UNR) {
    cplx<double> m[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) m[u] = ldw(P + 2 * (p + (k + u) * mi + ic));
#pragma unroll
    for (int u = 0; u < UNR; ++u)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] = cfma(acc[c], m[u], xval(k + u, c));
  }
  for (; k < qe; ++k) {
    const cplx<double> m = ldw(P + 2 * (p + k * mi + ic));
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = cfma(acc[c], m, xval(k, c));
  }
  __shared__ cplx<double> red[NW][64];
  double* y = Y + 2 * ((long)perm[r0 + ic] * ldy);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    if (c < ncol) {
      cplx<double> s = wave_order_sum<NW>(red, acc[c], wv, lane);
      if (wv == 0 && lane < mi) {
        if (accumulate) s = cadd(s, ld2(y + 2 * c));
        y[2 * c] = s.re;
        y[2 * c + 1] = s.im;
      }
    }
  }
}

// waves per workgroup and panel loads in flight per lane of the two product
// kernels (measured on the 25 000-face dimer's H-matrices: 8 / 8 for both;
// 16 loads in flight or 16 waves per row-panel workgroup were slower)
#ifndef MNP_HM_NW1
#define MNP_HM_NW1 8
#endif
#ifndef MNP_HM_NW2
#define MNP_HM_NW2 8
#endif
#ifndef MNP_HM_UNR1
#define MNP_HM_UNR1 8
#endif
#ifndef MNP_HM_UNR2
#define MNP_HM_UNR2 8
#endif
template <typename TS, int NC>
int matvec_packed_launch(int n, int ncitem, const long* citems, const void* VS, const int* tix, int nritem,
                         const long* ritems, const void* P, const int* gidx, const int* perm, int ncol,
                         const double* X, long ldx, double* Y, long ldy, int accumulate, double* R, hipStream_t st) {
  constexpr int NW1 = MNP_HM_NW1;
  constexpr int NW2 = MNP_HM_NW2;
  constexpr int U1 = MNP_HM_UNR1, U2 = MNP_HM_UNR2;
  hipLaunchKernelGGL(k_hmat_xperm, dim3(mnp_ceil_div((long)n * ncol, 256)), dim3(256), 0, st, n, ncol, perm, X, ldx,
                     R);
  MNP_LAUNCH_CHECK();
  if (ncitem > 0) {
    hipLaunchKernelGGL((k_hmat_vt_pk<TS, NC, NW1, U1>), dim3(ncitem), dim3(64 * NW1), 0, st, citems, (const TS*)VS, tix,
                       (const double*)R, ncol, R);
    MNP_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_hmat_rows_pk<TS, NC, NW2, U2>), dim3(nritem), dim3(64 * NW2), 0, st, ritems, (const TS*)P, gidx,
                     perm, (const double*)R, ncol, Y, ldy, accumulate);
  MNP_LAUNCH_CHECK();
  return 0;
}

// densify: Out[perm[r0+i]][perm[c0+j]] (row-major, ldo) for every block
template <typename TS>
__global__ __launch_bounds__(256) void k_hmat_full(const long* blk, const int* perm, const TS* val, const TS* Uv,
                                                   const TS* Vv, double* Out, long ldo) {
  const Blk B = load_blk(blk, blockIdx.x);
  const long mn = B.m * B.n;
  for (long e = threadIdx.x; e < mn; e += blockDim.x) {
    const long j = e / B.m, i = e - j * B.m;
    cplx<double> s;
    if (B.rank < 0) {
      s = ldw(val + 2 * (B.oa + e));
    } else {
      s = cmake(0.0, 0.0);
      for (long r = 0; r < B.rank; ++r)
        s = cfma(s, ldw(Uv + 2 * (B.oa + r * B.m + i)), ldw(Vv + 2 * (B.ob + r * B.n + j)));
    }
    double* o = Out + 2 * ((long)perm[B.r0 + i] * ldo + perm[B.c0 + j]);
    o[0] = s.re;
    o[1] = s.im;
  }
}

// ---------------------------------------------------------------------------
// Recompression of the low-rank blocks (hmatrix.py:672-737): QR of both
// factors U = Q1 R1, V = Q2 R2; SVD of X = R1 R2^T = W S Z^H; keep the
// leading r singular values, r = #{cumsum(s) < (1 - htol) sum(s)} (at least
// 1); new factors U' = Q1 (W_r S_r) = Q1 (X Z_r), V' = Q2 conj(Z_r).
//   k_trunc_qr   one workgroup per (block, factor): classical Gram-Schmidt
//                with one reorthogonalisation pass and deflation of
//                dependent columns, Q into a complex128 workspace, R (k x k)
//                beside it
//   k_trunc_svd  one wave per block: one-sided (Hestenes) Jacobi on X, lane
//                i owning rows i and i + 64 of X and of Z, so the rotations
//                need no barriers; the column norms give S.  X and Z sit in
//                LDS for ranks <= 64, in a global workspace up to 128
//   k_trunc_form U', V' = Q (X Z_r), Q conj(Z_r) in the storage type
// Singular vectors are fixed up to one phase per pair, so the new factors may
// differ from a LAPACK SVD's by column phases; their product does not.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum1(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

constexpr double QR_DEFLATE = 1e-13;  // residual / original norm below which a column is dependent
constexpr int TR_KMAX = 128;   // largest rank the recompression accepts
constexpr int TR_KLDS = 64;    // ranks up to this keep the SVD work in LDS

// Q, R of factor `side` (0: U, m rows; 1: V, n rows) of block bids[lb].
// Q at Q + qoff[2 lb + side] (column-major, rows x k); R column-major k x k at
// R + (2 lb + side) kst^2.
template <typename TS>
__global__ __launch_bounds__(256) void k_trunc_qr(const long* blk, const int* bids, const TS* Uv, const TS* Vv,
                                                  const long* qoff, double* Q, double* R, int kst) {
  const int lb = blockIdx.x, side = blockIdx.y, tid = threadIdx.x;
  const Blk B = load_blk(blk, bids[lb]);
  const int k = (int)B.rank;
  if (k < 1 || k > kst) return;  // k_trunc_svd reports the block
  const long rows = side ? B.n : B.m;
  const TS* src = side ? Vv + 2 * B.ob : Uv + 2 * B.oa;
  double* q = Q + 2 * qoff[2 * lb + side];
  double* r = R + 2 * (long)(2 * lb + side) * kst * kst;
  __shared__ cplx<double> d[TR_KMAX];
  __shared__ double part[4];
  for (long e = tid; e < rows * k; e += blockDim.x) {
    const cplx<double> v = ldw(src + 2 * e);
    q[2 * e] = v.re;
    q[2 * e + 1] = v.im;
  }
  for (int e = tid; e < k * k; e += blockDim.x) r[2 * e] = r[2 * e + 1] = 0.0;
  __syncthreads();
  const int lane = tid & 63, wv = tid >> 6;
  for (int j = 0; j < k; ++j) {
    double* qj = q + 2 * (long)j * rows;
    double p0 = 0.0;  // ||f_j||^2 before the projections
    for (long i = tid; i < rows; i += blockDim.x) {
      const cplx<double> v = ld2(qj + 2 * i);
      p0 += v.re * v.re + v.im * v.im;
    }
    p0 = wave_sum1(p0);
    if (lane == 0) part[wv] = p0;
    __syncthreads();
    const double n0 = sqrt(part[0] + part[1] + part[2] + part[3]);
    for (int pass = 0; pass < 2; ++pass) {
      for (int c = wv; c < j; c += 4) {  // d_c = q_c^H q_j
        const double* qc = q + 2 * (long)c * rows;
        cplx<double> s = cmake(0.0, 0.0);
        for (long i = lane; i < rows; i += 64) {
          const cplx<double> x = ld2(qc + 2 * i), y = ld2(qj + 2 * i);
          s.re += x.re * y.re + x.im * y.im;
          s.im += x.re * y.im - x.im * y.re;
        }
        s = wave_sum(s);
        if (lane == 0) {
          d[c] = s;
          r[2 * (j * k + c)] += s.re;
          r[2 * (j * k + c) + 1] += s.im;
        }
      }
      __syncthreads();
      for (long i = tid; i < rows; i += blockDim.x) {
        cplx<double> acc = ld2(qj + 2 * i);
        for (int c = 0; c < j; ++c) acc = csub(acc, cmul(ld2(q + 2 * ((long)c * rows + i)), d[c]));
        qj[2 * i] = acc.re;
        qj[2 * i + 1] = acc.im;
      }
      __syncthreads();
    }
    double ps = 0.0;
    for (long i = tid; i < rows; i += blockDim.x) {
      const cplx<double> v = ld2(qj + 2 * i);
      ps += v.re * v.re + v.im * v.im;
    }
    ps = wave_sum1(ps);
    if (lane == 0) part[wv] = ps;
    __syncthreads();
    double nrm = sqrt(part[0] + part[1] + part[2] + part[3]);
    // a column left with rounding noise only (in the span of the previous
    // ones, e.g. the repeated factors of a sum) is deflated: zero Q column,
    // zero diagonal.  Normalising the noise would feed non-orthogonal
    // directions into every later projection.
    const bool dependent = !(nrm > QR_DEFLATE * n0);
    if (dependent) nrm = 0.0;
    if (tid == 0) {
      r[2 * (j * k + j)] = nrm;
      r[2 * (j * k + j) + 1] = 0.0;
    }
    const double inv = dependent ? 0.0 : 1.0 / nrm;
    for (long i = tid; i < rows; i += blockDim.x) {
      qj[2 * i] *= inv;
      qj[2 * i + 1] *= inv;
    }
    __syncthreads();
  }
}

// SVD of X = R1 R2^T and the truncation rule.  RT: on entry R1, R2 (column-
// major k x k at RT + (2 lb + side) kst^2); on exit T1 = X Z_r and
// T2 = conj(Z_r) (k x r, column-major, same places).  rank_out receives r.
// Factors whose Frobenius norm is below machine epsilon are kept as they are
// (T = R, r = k); a block whose singular values sum below epsilon becomes a
// zero rank-1 block (hmatrix.py:703-721).  X and Z (k x k each) live in
// dynamic LDS when XZ is null (kst <= 64), else in XZ + 2 lb kst^2 (complex).
__global__ __launch_bounds__(64) void k_trunc_svd(const long* blk, const int* bids, double* RT, int kst,
                                                  double htol, double* XZ, int* rank_out) {
  extern __shared__ double dyn[];
  const int lb = blockIdx.x, i = threadIdx.x;
  const Blk B = load_blk(blk, bids[lb]);
  const int k = (int)B.rank;
  if (k < 1 || k > kst) {
    if (i == 0) rank_out[lb] = -1;
    return;
  }
  cplx<double>* X = XZ ? reinterpret_cast<cplx<double>*>(XZ) + 2 * (long)lb * kst * kst
                       : reinterpret_cast<cplx<double>*>(dyn);
  cplx<double>* Z = X + k * k;
  __shared__ double sig[TR_KMAX];
  __shared__ int ord[TR_KMAX];
  double* r1 = RT + 2 * (long)(2 * lb) * kst * kst;
  double* r2 = RT + 2 * (long)(2 * lb + 1) * kst * kst;
  const double eps = 2.220446049250313e-16;
  // ||U||_F = ||R1||_F, ||V||_F = ||R2||_F
  double n1 = 0.0, n2 = 0.0;
  for (int ii = i; ii < k; ii += 64)
    for (int l = ii; l < k; ++l) {
      const cplx<double> a = ld2(r1 + 2 * (l * k + ii)), b = ld2(r2 + 2 * (l * k + ii));
      n1 += a.re * a.re + a.im * a.im;
      n2 += b.re * b.re + b.im * b.im;
    }
  n1 = sqrt(wave_sum1(n1));
  n2 = sqrt(wave_sum1(n2));
  if (n1 < eps || n2 < eps) {  // T = R already in place
    if (i == 0) rank_out[lb] = k;
    return;
  }
  // X[i][j] = sum_{l >= max(i, j)} R1[i][l] R2[j][l]; Z = I
  for (int ii = i; ii < k; ii += 64)
    for (int j = 0; j < k; ++j) {
      cplx<double> s = cmake(0.0, 0.0);
      for (int l = ii > j ? ii : j; l < k; ++l) s = cfma(s, ld2(r1 + 2 * (l * k + ii)), ld2(r2 + 2 * (l * k + j)));
      X[j * k + ii] = s;
      Z[j * k + ii] = cmake(ii == j ? 1.0 : 0.0, 0.0);
    }
  // one-sided Jacobi: rotate column pairs until all are orthogonal.  Lane i
  // owns rows i, i + 64 of X and Z; only the sums cross lanes.
  for (int sweep = 0; sweep < 60; ++sweep) {
    int rotated = 0;
    for (int p = 0; p < k - 1; ++p)
      for (int q = p + 1; q < k; ++q) {
        double a = 0.0, b = 0.0, gr = 0.0, gi = 0.0;  // g = x_p^H x_q
        for (int ii = i; ii < k; ii += 64) {
          const cplx<double> xp = X[p * k + ii], xq = X[q * k + ii];
          a += xp.re * xp.re + xp.im * xp.im;
          b += xq.re * xq.re + xq.im * xq.im;
          gr += xp.re * xq.re + xp.im * xq.im;
          gi += xp.re * xq.im - xp.im * xq.re;
        }
        a = wave_sum1(a);
        b = wave_sum1(b);
        gr = wave_sum1(gr);
        gi = wave_sum1(gi);
        const double g = hypot(gr, gi);
        if (!(g > eps * sqrt(a * b)) || g == 0.0) continue;
        rotated = 1;
        // columns [x_p, e^{-i phi} x_q] have real Gram [[a, g], [g, b]]
        const double zeta = (b - a) / (2.0 * g);
        const double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        const cplx<double> w = cmake(gr / g, -gi / g);
        for (int ii = i; ii < k; ii += 64) {
          const cplx<double> xp = X[p * k + ii], yq = cmul(X[q * k + ii], w);
          X[p * k + ii] = csub(cscale(xp, c), cscale(yq, s));
          X[q * k + ii] = cadd(cscale(xp, s), cscale(yq, c));
          const cplx<double> zp = Z[p * k + ii], zq = cmul(Z[q * k + ii], w);
          Z[p * k + ii] = csub(cscale(zp, c), cscale(zq, s));
          Z[q * k + ii] = cadd(cscale(zp, s), cscale(zq, c));
        }
      }
    if (!rotated) break;
  }
  // singular values = column norms of X Z; order by descending value, ties by index
  for (int p = 0; p < k; ++p) {
    double a = 0.0;
    for (int ii = i; ii < k; ii += 64) {
      const cplx<double> x = X[p * k + ii];
      a += x.re * x.re + x.im * x.im;
    }
    a = sqrt(wave_sum1(a));
    if (i == 0) sig[p] = a;
  }
  __syncthreads();
  for (int ii = i; ii < k; ii += 64) {
    const double me = sig[ii];
    int pos = 0;
    for (int p = 0; p < k; ++p) pos += (sig[p] > me || (sig[p] == me && p < ii)) ? 1 : 0;
    ord[pos] = ii;
  }
  __syncthreads();
  double tot = 0.0;
  for (int p = 0; p < k; ++p) tot += sig[ord[p]];
  if (tot < eps) {  // zero rank-1 block
    for (int ii = i; ii < k; ii += 64) {
      r1[2 * ii] = r1[2 * ii + 1] = 0.0;
      r2[2 * ii] = r2[2 * ii + 1] = 0.0;
    }
    if (i == 0) rank_out[lb] = 1;
    return;
  }
  int rr = 0;
  double cum = 0.0;
  const double thr = (1.0 - htol) * tot;
  for (int p = 0; p < k; ++p) {
    cum += sig[ord[p]];
    rr += cum < thr ? 1 : 0;
  }
  if (rr == 0) rr = 1;
  for (int ii = i; ii < k; ii += 64)
    for (int c = 0; c < rr; ++c) {
      const int zc = ord[c];
      const cplx<double> x = X[zc * k + ii], z = Z[zc * k + ii];
      r1[2 * (c * k + ii)] = x.re;
      r1[2 * (c * k + ii) + 1] = x.im;
      r2[2 * (c * k + ii)] = z.re;
      r2[2 * (c * k + ii) + 1] = -z.im;
    }
  if (i == 0) rank_out[lb] = rr;
}

// new factors: U' = Q1 T1 (m x r) at the new table's off_a, V' = Q2 T2
// (n x r) at off_b, in the storage type
template <typename TS>
__global__ __launch_bounds__(256) void k_trunc_form(const long* blk_old, const long* blk_new, const int* bids,
                                                    const long* qoff, const double* Q, const double* T, int kst,
                                                    TS* Un, TS* Vn) {
  const int lb = blockIdx.x, side = blockIdx.y;
  const Blk Bo = load_blk(blk_old, bids[lb]);
  const Blk Bn = load_blk(blk_new, bids[lb]);
  const int k = (int)Bo.rank, r = (int)Bn.rank;
  if (k < 1 || k > kst || r < 0 || r > k) return;
  const long rows = side ? Bo.n : Bo.m;
  const double* q = Q + 2 * qoff[2 * lb + side];
  const double* t = T + 2 * (long)(2 * lb + side) * kst * kst;
  TS* out = side ? Vn + 2 * Bn.ob : Un + 2 * Bn.oa;
  for (long e = (long)blockIdx.z * blockDim.x + threadIdx.x; e < rows * r; e += (long)gridDim.z * blockDim.x) {
    const long i = e % rows, c = e / rows;
    cplx<double> s = cmake(0.0, 0.0);
    for (int a = 0; a < k; ++a) s = cfma(s, ld2(q + 2 * ((long)a * rows + i)), ld2(t + 2 * (c * k + a)));
    cst(out + 2 * e, cmake<TS>((TS)s.re, (TS)s.im));
  }
}

// Out = alpha X + beta Y (Y may be null: Out = alpha X), elementwise over
// `count` complex values of the storage type; Out may alias X or Y
template <typename TS>
__global__ __launch_bounds__(256) void k_hmat_axpby(long count, double ar, double ai, const TS* X, double br,
                                                    double bi, const TS* Y, TS* Out) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (long)gridDim.x * blockDim.x) {
    const cplx<double> x = ldw(X + 2 * e);
    cplx<double> o = cmul(cmake(ar, ai), x);
    if (Y) o = cfma(o, cmake(br, bi), ldw(Y + 2 * e));
    cst(Out + 2 * e, cmake<TS>((TS)o.re, (TS)o.im));
  }
}

// low-rank blocks of a sum: U = [Ua, Ub], V = [Va, Vb] (columns of a first);
// the three tables share the block geometry, ranks add up
template <typename TS>
__global__ __launch_bounds__(256) void k_hmat_concat(const long* blk_a, const long* blk_b, const long* blk_n,
                                                     const int* bids, const TS* Ua, const TS* Va, const TS* Ub,
                                                     const TS* Vb, TS* Un, TS* Vn) {
  const long b = bids[blockIdx.x];
  const Blk A = load_blk(blk_a, b), Bb = load_blk(blk_b, b), N = load_blk(blk_n, b);
  const long ka = A.rank > 0 ? A.rank : 0, kb = Bb.rank > 0 ? Bb.rank : 0;
  for (long e = threadIdx.x; e < N.m * (ka + kb); e += blockDim.x) {
    const TS* s = e < N.m * ka ? Ua + 2 * (A.oa + e) : Ub + 2 * (Bb.oa + e - N.m * ka);
    cst(Un + 2 * (N.oa + e), cld(s));
  }
  for (long e = threadIdx.x; e < N.n * (ka + kb); e += blockDim.x) {
    const TS* s = e < N.n * ka ? Va + 2 * (A.ob + e) : Vb + 2 * (Bb.ob + e - N.n * ka);
    cst(Vn + 2 * (N.ob + e), cld(s));
  }
}

// diagonal of the diagonal dense blocks (r0 == c0, m == n) into out (complex128,
// face order), or (mode 1) set those blocks to the identity
template <typename TS>
__global__ __launch_bounds__(256) void k_hmat_diag(const long* blk, const int* bids, const int* perm, TS* val,
                                                   double* out, int mode) {
  const Blk B = load_blk(blk, bids[blockIdx.x]);
  for (long i = threadIdx.x; i < B.m; i += blockDim.x) {
    TS* v = val + 2 * (B.oa + i * B.m + i);
    if (mode == 0) {
      const cplx<double> d = ldw(v);
      out[2 * (long)perm[B.r0 + i]] = d.re;
      out[2 * (long)perm[B.r0 + i] + 1] = d.im;
    } else {
      cst(v, cmake<TS>((TS)1, (TS)0));
    }
  }
}

template <typename TS, int NC>
int matvec_launch(int n, int nblk, int nleaf, int maxchunks, const long* blk, const long* leafs, const int* lst,
                  const int* perm, const void* val, const void* U, const void* V, int ncol, const double* X,
                  long ldx, double* Y, long ldy, int accumulate, double* T, double* XW, hipStream_t st) {
  hipLaunchKernelGGL(k_hmat_xperm, dim3(mnp_ceil_div((long)n * ncol, 256)), dim3(256), 0, st, n, ncol, perm, X, ldx,
                     XW);
  MNP_LAUNCH_CHECK();
  if (nblk > 0) {
    hipLaunchKernelGGL((k_hmat_vt<TS, NC>), dim3(nblk), dim3(256), 0, st, blk, (const TS*)V, (const double*)XW, ncol,
                       T);
    MNP_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_hmat_rows<TS, NC>), dim3(nleaf, maxchunks), dim3(64 * RowsCfg<NC>::nw), 0, st, leafs, lst,
                     blk, perm,
                     (const TS*)val, (const TS*)U, (const double*)T, (const double*)XW, ncol, Y, ldy, accumulate);
  MNP_LAUNCH_CHECK();
  return 0;
}

}  // namespace

// ACA of the low-rank blocks bids[0:nb] of the block table, entries from the
// complex128 matrix M[i][j] = src[i si + j sj] (element strides).  Uw / Vw:
// complex128 workspaces, block l's factors at woff[2 l] / woff[2 l + 1] with
// room for min(m, n, kmax) columns.  rank_out[l] receives the rank.
extern "C" int mnpbem_hmat_aca(int nb, const long* blk, const int* bids, const int* perm, const void* src, long si,
                               long sj, void* Uw, void* Vw, const long* woff, int kmax, double htol, int* rank_out,
                               void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(kmax >= 1 && kmax <= ACA_KMAX, "hmat_aca: kmax must be in [1, 512]");
  MNP_REQUIRE(htol >= 0.0, "hmat_aca: htol < 0");
  if (nb <= 0) return 0;
  const DenseSrc s{(const double*)src, si, sj};
  hipLaunchKernelGGL(k_aca<DenseSrc>, dim3(nb), dim3(256), 0, st, blk, bids, perm, s, (double*)Uw, (double*)Vw,
                     woff, kmax, htol, rank_out);
  MNP_LAUNCH_CHECK();
  return 0;
}

// dense blocks bids[0:nb]: copy from the complex128 matrix M[i][j] = src[i si + j sj] into val (dtype)
extern "C" int mnpbem_hmat_dense(int dtype, int nb, const long* blk, const int* bids, const int* perm,
                                 const void* src, long si, long sj, void* val, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (nb <= 0) return 0;
  const DenseSrc s{(const double*)src, si, sj};
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL((k_hmat_dense<double, DenseSrc>), dim3(nb), dim3(256), 0, st, blk, bids, perm, s,
                       (double*)val);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL((k_hmat_dense<float, DenseSrc>), dim3(nb), dim3(256), 0, st, blk, bids, perm, s,
                       (float*)val);
  else {
    mnp_set_error("hmat_dense: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// The two fills above with entries generated in place from the geometry and
// the refinement tables of one side, never from an assembled matrix: kind 0
// G, kind 1 F + hsign 2 pi [i == j] (H1 with hsign +1, H2 with -1).  kmat:
// the nmat complex wavenumbers of one wavelength (float64 pairs); mat: the
// side's material index per face (pairs of different materials are 0);
// rowptr (int64, n + 1), pcol, g, f: the (row, col)-sorted refined pairs and
// their order + 1 expansion coefficients, as for mnpbem_green_select.
extern "C" int mnpbem_hmat_aca_green(int nb, const long* blk, const int* bids, const int* perm, int kind,
                                     double hsign, const double* pos, const double* nvec, const double* area,
                                     const int* mat, const double* kmat, int order, const long* rowptr,
                                     const int* pcol, const double* g, const double* f, void* Uw, void* Vw,
                                     const long* woff, int kmax, double htol, int* rank_out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(kmax >= 1 && kmax <= ACA_KMAX, "hmat_aca_green: kmax must be in [1, 512]");
  MNP_REQUIRE(htol >= 0.0, "hmat_aca_green: htol < 0");
  MNP_REQUIRE(kind == 0 || kind == 1, "hmat_aca_green: kind must be 0 (G) or 1 (F / H)");
  MNP_REQUIRE(order >= 0, "hmat_aca_green: order < 0");
  if (nb <= 0) return 0;
  const GreenSrc s{kind, hsign, pos, nvec, area, mat, kmat, rowptr, pcol, g, f, order + 1};
  hipLaunchKernelGGL(k_aca<GreenSrc>, dim3(nb), dim3(256), 0, st, blk, bids, perm, s, (double*)Uw, (double*)Vw,
                     woff, kmax, htol, rank_out);
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_hmat_dense_green(int dtype, int nb, const long* blk, const int* bids, const int* perm,
                                       int kind, double hsign, const double* pos, const double* nvec,
                                       const double* area, const int* mat, const double* kmat, int order,
                                       const long* rowptr, const int* pcol, const double* g, const double* f,
                                       void* val, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(kind == 0 || kind == 1, "hmat_dense_green: kind must be 0 (G) or 1 (F / H)");
  MNP_REQUIRE(order >= 0, "hmat_dense_green: order < 0");
  if (nb <= 0) return 0;
  const GreenSrc s{kind, hsign, pos, nvec, area, mat, kmat, rowptr, pcol, g, f, order + 1};
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL((k_hmat_dense<double, GreenSrc>), dim3(nb), dim3(256), 0, st, blk, bids, perm, s,
                       (double*)val);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL((k_hmat_dense<float, GreenSrc>), dim3(nb), dim3(256), 0, st, blk, bids, perm, s,
                       (float*)val);
  else {
    mnp_set_error("hmat_dense_green: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// low-rank blocks bids[0:nb]: packed U, V (dtype) from the ACA workspaces
extern "C" int mnpbem_hmat_compact(int dtype, int nb, const long* blk, const int* bids, const void* Uw,
                                   const void* Vw, const long* woff, void* U, void* V, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (nb <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_hmat_compact<double>, dim3(nb), dim3(256), 0, st, blk, bids, (const double*)Uw,
                       (const double*)Vw, woff, (double*)U, (double*)V);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL(k_hmat_compact<float>, dim3(nb), dim3(256), 0, st, blk, bids, (const double*)Uw,
                       (const double*)Vw, woff, (float*)U, (float*)V);
  else {
    mnp_set_error("hmat_compact: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// Y (+)= M X for the H-matrix (hmatrix.py:450-541) of n rows.  X, Y: (n, ldx
// / ldy) complex128 row-major in face order, ncol <= 16 columns used.  T:
// complex128 workspace of (sum of ranks) * ncol elements; XW: complex128
// workspace of n * ncol elements (X in cluster order).  maxchunks =
// ceil(max leaf size / 256).
extern "C" int mnpbem_hmat_matvec(int dtype, int n, int nblk, int nleaf, int maxchunks, const long* blk,
                                  const long* leafs, const int* lst, const int* perm, const void* val, const void* U,
                                  const void* V, int ncol, const void* X, long ldx, void* Y, long ldy, int accumulate,
                                  void* T, void* XW, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(ncol >= 1 && ncol <= 16 && ldx >= ncol && ldy >= ncol, "hmat_matvec: 1 <= ncol <= 16");
  MNP_REQUIRE(n >= 1 && nleaf >= 1 && maxchunks >= 1 && maxchunks <= 65535, "hmat_matvec: bad leaf table");
  MNP_REQUIRE(XW != nullptr, "hmat_matvec: XW workspace required");
  const double* x = (const double*)X;
  double* y = (double*)Y;
  double* t = (double*)T;
  double* xw = (double*)XW;
#define MNP_MV(TS, NC)                                                                                          \
  return matvec_launch<TS, NC>(n, nblk, nleaf, maxchunks, blk, leafs, lst, perm, val, U, V, ncol, x, ldx, y, ldy, \
                               accumulate, t, xw, st)
  if (dtype == MNP_DTYPE_C128) {
    if (ncol <= 4) MNP_MV(double, 4);
    if (ncol <= 8) MNP_MV(double, 8);
    MNP_MV(double, 16);
  }
  if (dtype == MNP_DTYPE_C64) {
    if (ncol <= 4) MNP_MV(float, 4);
    if (ncol <= 8) MNP_MV(float, 8);
    MNP_MV(float, 16);
  }
#undef MNP_MV
  mnp_set_error("hmat_matvec: dtype must be 1 or 2");
  return 1;
}

// Copy the blocks into the packed product layouts (the layouts and piece
// descriptors are described above k_hmat_pack_cols): ncp column pieces
// building VS and tix, nrp row pieces building P and gidx.
extern "C" int mnpbem_hmat_pack(int dtype, int ncp, const long* cdesc, int nrp, const long* rdesc, const void* val,
                                const void* U, const void* V, void* VS, int* tix, void* P, int* gidx, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(ncp >= 0 && nrp >= 0, "hmat_pack: negative piece count");
  if (dtype != MNP_DTYPE_C128 && dtype != MNP_DTYPE_C64) {
    mnp_set_error("hmat_pack: dtype must be 1 or 2");
    return 1;
  }
#define MNP_PK(TS)                                                                                               \
  if (ncp > 0) {                                                                                                 \
    hipLaunchKernelGGL(k_hmat_pack_cols<TS>, dim3(ncp), dim3(256), 0, st, cdesc, (const TS*)V, (TS*)VS, tix);    \
    MNP_LAUNCH_CHECK();                                                                                          \
  }                                                                                                              \
  if (nrp > 0) {                                                                                                 \
    hipLaunchKernelGGL(k_hmat_pack_rows<TS>, dim3(nrp), dim3(256), 0, st, rdesc, (const TS*)val, (const TS*)U,   \
                       (TS*)P, gidx);                                                                            \
    MNP_LAUNCH_CHECK();                                                                                          \
  }
  if (dtype == MNP_DTYPE_C128) {
    MNP_PK(double)
  } else {
    MNP_PK(float)
  }
#undef MNP_PK
  return 0;
}

// Y (+)= M X on the packed layouts: phase 0 writes X in cluster order into
// rows [0, n) of the workspace R (rows of ncol values; (n + sum of ranks)
// * 16 complex128 elements allocated),
// phase 1 the t rows of R from the ncitem column-panel items (6 int64:
// {vs, K, o0, n, c0, og}), phase 2 Y from the nritem row-panel items
// (6 int64: {p, mi, K, g, r0, 0}).  X, Y: (n, ldx / ldy) complex128
// row-major in face order, 1 <= ncol <= 16.
extern "C" int mnpbem_hmat_matvec_packed(int dtype, int n, int ncitem, const long* citems, const void* VS,
                                         const int* tix, int nritem, const long* ritems, const void* P,
                                         const int* gidx, const int* perm, int ncol, const void* X, long ldx, void* Y,
                                         long ldy, int accumulate, void* R, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(ncol >= 1 && ncol <= 16 && ldx >= ncol && ldy >= ncol, "hmat_matvec_packed: 1 <= ncol <= 16");
  MNP_REQUIRE(n >= 1 && nritem >= 1 && ncitem >= 0, "hmat_matvec_packed: bad item tables");
  MNP_REQUIRE(R != nullptr, "hmat_matvec_packed: R workspace required");
  const double* x = (const double*)X;
  double* y = (double*)Y;
  double* r = (double*)R;
#define MNP_MV(TS, NC)                                                                                            \
  return matvec_packed_launch<TS, NC>(n, ncitem, citems, VS, tix, nritem, ritems, P, gidx, perm, ncol, x, ldx, y, \
                                      ldy, accumulate, r, st)
  if (dtype == MNP_DTYPE_C128) {
    if (ncol <= 4) MNP_MV(double, 4);
    if (ncol <= 8) MNP_MV(double, 8);
    MNP_MV(double, 16);
  }
  if (dtype == MNP_DTYPE_C64) {
    if (ncol <= 4) MNP_MV(float, 4);
    if (ncol <= 8) MNP_MV(float, 8);
    MNP_MV(float, 16);
  }
#undef MNP_MV
  mnp_set_error("hmat_matvec_packed: dtype must be 1 or 2");
  return 1;
}

// Out (n x n complex128, row-major, face order, leading dimension ldo) = the
// H-matrix (hmatrix.py:338-448)
extern "C" int mnpbem_hmat_full(int dtype, int nblk, const long* blk, const int* perm, const void* val,
                                const void* U, const void* V, void* Out, long ldo, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (nblk <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_hmat_full<double>, dim3(nblk), dim3(256), 0, st, blk, perm, (const double*)val,
                       (const double*)U, (const double*)V, (double*)Out, ldo);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL(k_hmat_full<float>, dim3(nblk), dim3(256), 0, st, blk, perm, (const float*)val,
                       (const float*)U, (const float*)V, (double*)Out, ldo);
  else {
    mnp_set_error("hmat_full: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// Recompression, step 1 (hmatrix.py:672-737): for the low-rank blocks
// bids[0:nb] (each of rank 1 <= k <= kst <= 128), QR of both factors into Q
// (complex128, factor `side` of block l at qoff[2 l + side], column-major),
// SVD of R1 R2^T and the truncation rule at htol.  RT: complex128 workspace of
// 2 nb kst^2 elements; on return it holds the k x r transforms that
// mnpbem_hmat_trunc_form applies.  XZ: complex128 workspace of 2 nb kst^2
// elements, required for kst > 64 (null otherwise).  rank_out[l] receives the
// new rank (-1: block outside the supported ranks).
extern "C" int mnpbem_hmat_truncate(int dtype, int nb, const long* blk, const int* bids, const void* U, const void* V,
                                    const long* qoff, int kst, double htol, void* Q, void* RT, void* XZ,
                                    int* rank_out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(kst >= 1 && kst <= TR_KMAX, "hmat_truncate: ranks above 128 are not supported");
  MNP_REQUIRE(kst <= TR_KLDS || XZ != nullptr, "hmat_truncate: ranks above 64 need the XZ workspace");
  MNP_REQUIRE(htol >= 0.0 && htol < 1.0, "hmat_truncate: htol must lie in [0, 1)");
  if (nb <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_trunc_qr<double>, dim3(nb, 2), dim3(256), 0, st, blk, bids, (const double*)U,
                       (const double*)V, qoff, (double*)Q, (double*)RT, kst);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL(k_trunc_qr<float>, dim3(nb, 2), dim3(256), 0, st, blk, bids, (const float*)U,
                       (const float*)V, qoff, (double*)Q, (double*)RT, kst);
  else {
    mnp_set_error("hmat_truncate: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  const bool lds = kst <= TR_KLDS;
  const size_t bytes = lds ? (size_t)2 * kst * kst * 2 * sizeof(double) : 0;
  if (bytes > 65536)
    MNP_HIP_CHECK(
        hipFuncSetAttribute((const void*)k_trunc_svd, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  hipLaunchKernelGGL(k_trunc_svd, dim3(nb), dim3(64), bytes, st, blk, bids, (double*)RT, kst, htol,
                     lds ? nullptr : (double*)XZ, rank_out);
  MNP_LAUNCH_CHECK();
  return 0;
}

// Recompression, step 2: U' = Q1 T1, V' = Q2 T2 into the packed factors of
// the new block table (same geometry, ranks from mnpbem_hmat_truncate; T is
// its RT workspace).  zsplit: workgroups per (block, factor).
extern "C" int mnpbem_hmat_trunc_form(int dtype, int nb, const long* blk_old, const long* blk_new, const int* bids,
                                      const long* qoff, const void* Q, const void* T, int kst, int zsplit, void* Un,
                                      void* Vn, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(kst >= 1 && kst <= TR_KMAX, "hmat_trunc_form: ranks above 128 are not supported");
  MNP_REQUIRE(zsplit >= 1 && zsplit <= 65535, "hmat_trunc_form: bad zsplit");
  if (nb <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_trunc_form<double>, dim3(nb, 2, zsplit), dim3(256), 0, st, blk_old, blk_new, bids, qoff,
                       (const double*)Q, (const double*)T, kst, (double*)Un, (double*)Vn);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL(k_trunc_form<float>, dim3(nb, 2, zsplit), dim3(256), 0, st, blk_old, blk_new, bids, qoff,
                       (const double*)Q, (const double*)T, kst, (float*)Un, (float*)Vn);
  else {
    mnp_set_error("hmat_trunc_form: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// Out = alpha X + beta Y over `count` complex values of the storage type
// (scalar multiple, negation and the dense part of a sum, hmatrix.py:553-670);
// Y may be null.
extern "C" int mnpbem_hmat_axpby(int dtype, long count, double ar, double ai, const void* X, double br, double bi,
                                 const void* Y, void* Out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(count >= 0, "hmat_axpby: negative count");
  if (count == 0) return 0;
  const long want = (count + 255) / 256;
  const int grid = (int)(want < 8192 ? want : 8192);
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_hmat_axpby<double>, dim3(grid), dim3(256), 0, st, count, ar, ai, (const double*)X, br, bi,
                       (const double*)Y, (double*)Out);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL(k_hmat_axpby<float>, dim3(grid), dim3(256), 0, st, count, ar, ai, (const float*)X, br, bi,
                       (const float*)Y, (float*)Out);
  else {
    mnp_set_error("hmat_axpby: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// Low-rank blocks of a sum of two H-matrices on one tree (hmatrix.py:592-670):
// factors of blk_n's blocks bids[0:nb] become [Ua, Ub], [Va, Vb].
extern "C" int mnpbem_hmat_concat(int dtype, int nb, const long* blk_a, const long* blk_b, const long* blk_n,
                                  const int* bids, const void* Ua, const void* Va, const void* Ub, const void* Vb,
                                  void* Un, void* Vn, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (nb <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_hmat_concat<double>, dim3(nb), dim3(256), 0, st, blk_a, blk_b, blk_n, bids,
                       (const double*)Ua, (const double*)Va, (const double*)Ub, (const double*)Vb, (double*)Un,
                       (double*)Vn);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL(k_hmat_concat<float>, dim3(nb), dim3(256), 0, st, blk_a, blk_b, blk_n, bids,
                       (const float*)Ua, (const float*)Va, (const float*)Ub, (const float*)Vb, (float*)Un,
                       (float*)Vn);
  else {
    mnp_set_error("hmat_concat: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// Diagonal dense blocks bids[0:nb] (r0 == c0, m == n): mode 0 writes their
// diagonal to out (complex128, face order; hmatrix.py:758-778), mode 1 sets
// their diagonal to one (identity, hmatrix.py:780-798, after a zero fill).
extern "C" int mnpbem_hmat_diag(int dtype, int nb, const long* blk, const int* bids, const int* perm, void* val,
                                void* out, int mode, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(mode == 0 || mode == 1, "hmat_diag: mode must be 0 or 1");
  MNP_REQUIRE(mode == 1 || out != nullptr, "hmat_diag: output required");
  if (nb <= 0) return 0;
  if (dtype == MNP_DTYPE_C128)
    hipLaunchKernelGGL(k_hmat_diag<double>, dim3(nb), dim3(256), 0, st, blk, bids, perm, (double*)val, (double*)out,
                       mode);
  else if (dtype == MNP_DTYPE_C64)
    hipLaunchKernelGGL(k_hmat_diag<float>, dim3(nb), dim3(256), 0, st, blk, bids, perm, (float*)val, (double*)out,
                       mode);
  else {
    mnp_set_error("hmat_diag: dtype must be 1 or 2");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}
