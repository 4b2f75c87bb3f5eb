// Geometry derivatives of the extinction cross section (gfx950): the
// discrete adjoint of the retarded BEM solve, contracted with the derivatives
// of every geometric ingredient of the operator, the plane-wave right side and
// the optical-theorem functional.  No counterpart in the reference (it has no
// derivatives); the formulas differentiate, term by term, what the library
// assembles after greenret_refined.py:122-810, compgreen_ret.py:192-400,
// planewave_ret.py:257-504 and spectrum_ret.py:231-392.
//
// ext_bp = Im(c . x) with A x = b (c scaled by 4 pi Re k / |k|^2), y = A^-T c:
//   d ext = Im(dc . x + y . db - y . dA x).
// The per-face geometric inputs are pos (centroid), nvec, tvec1, tvec2, area
// and the 6 nodes of each face; the derivative with respect to each is
// accumulated in Gr[b][p][face][31] = [pos 3 | nvec 3 | tvec1 3 | tvec2 3 |
// area 1 | nodes 18], which the caller chains to the mesh vertices.
//
// y . A x splits over the four Green matrices: with
//   LG1 = -[y_phi - ik0 e1 n.y_al, y_a - ik0 e1 n y_De], LH1 = -[e1 y_De, y_al],
//   LG2 = +[y_phi - ik0 e2 n.y_al, y_a - ik0 e2 n y_De], LH2 = +[e2 y_De, y_al],
//   R1 = [s1, h1], R2 = [s2, h2]   (4 components each, per face and pol),
// -y . A x = sum_s sum_ij (LG_s(i).R_s(j)) G_s[i][j] + (LH_s(i).R_s(j)) H_s[i][j]
// plus the normals inside LG (the explicit n of GdA&H Eqs. 14, 17).
#include "common.h"

namespace {

constexpr int NG = 31;  // per-face gradient slots
constexpr int S_POS = 0, S_NVEC = 3, S_T1 = 6, S_T2 = 9, S_AREA = 12, S_NODES = 13;
constexpr int QSTRIDE = 19;  // quadrature table row: shape[6], d/dx[6], d/dy[6], weight
constexpr double TINY = 2.220446049250313e-16;

__device__ __forceinline__ cplx<double> ldc(const double* p) { return cmake(p[0], p[1]); }
__device__ __forceinline__ cplx<double> ik_times(double k, cplx<double> z) { return cmake(-k * z.im, k * z.re); }
__device__ __forceinline__ double im_mul(cplx<double> a, cplx<double> b) { return a.re * b.im + a.im * b.re; }

struct Par {
  cplx<double> e1, e2;
  double k0;
};
__device__ __forceinline__ Par load_par(const double* par, long b) {
  const double* q = par + 5 * b;
  Par p;
  p.e1 = cmake(q[0], q[1]);
  p.e2 = cmake(q[2], q[3]);
  p.k0 = q[4];
  return p;
}

// ---------------------------------------------------------------------------
// extinction functional C = kappa [c_s1 | c_h1 | c_s2 | c_h2] with
// c_sig = -ik (pol.dir) ph, c_h = ik0 pol ph, ph = e^{-ik dir.pos} area, on the
// sides whose medium is the embedding one (k_extinction's sum written as a
// vector), kappa = 4 pi Re k / |k|^2.
__global__ void k_ext_functional(int n, int P, const double* pos, const double* area, const int* in_med,
                                 const int* out_med, const double* pol, const double* dir, const double* pw,
                                 double* C, long ld, long stride) {
  const long b = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double* q = pw + 7 * b;
  const double k0 = q[0], kr = q[1], ki = q[2];
  const double kap = 4.0 * M_PI * kr / (kr * kr + ki * ki);
  double* c = C + 2 * (b * stride + (long)j * ld);
  for (int p = 0; p < P; ++p) {
    const double* dp = dir + 3 * p;
    const double* pp = pol + 3 * p;
    const double kd = pos[3 * j] * dp[0] + pos[3 * j + 1] * dp[1] + pos[3 * j + 2] * dp[2];
    const cplx<double> ph = cscale(cexpi(-kr * kd, -ki * kd), area[j] * kap);
    const double pd = pp[0] * dp[0] + pp[1] * dp[1] + pp[2] * dp[2];
    const cplx<double> csig = cscale(cmul(cmake(ki, -kr), ph), pd);  // -ik pd ph
    const cplx<double> ch = ik_times(k0, ph);
    for (int side = 0; side < 2; ++side) {
      const bool on = side == 0 ? in_med[j] : out_med[j];
      const int off = 4 * P * side;
      const cplx<double> z = cmake(0.0, 0.0);
      cst(c + 2 * (off + p), on ? csig : z);
      for (int cc = 0; cc < 3; ++cc) cst(c + 2 * (off + P + cc * P + p), on ? cscale(ch, pp[cc]) : z);
    }
  }
}

// ---------------------------------------------------------------------------
// L (b, i, side, p, 8) = [LG | LH], R (b, j, side, p, 4), Wc (b, j, p) = the
// weight of H_s[j][j] summed over the sides (it carries the closed-surface
// correction of the self term).
__global__ void k_grad_prep(int n, int P, const double* X, const double* Y, long ld, long stride, const double* nvec,
                            const double* par, double* L, double* R, double* Wc) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const Par q = load_par(par, b);
  const double* x = X + 2 * (b * stride + (long)i * ld);
  const double* y = Y + 2 * (b * stride + (long)i * ld);
  const double nv[3] = {nvec[3 * i], nvec[3 * i + 1], nvec[3 * i + 2]};
  for (int p = 0; p < P; ++p) {
    const cplx<double> yphi = ldc(y + 2 * p), yde = ldc(y + 2 * (4 * P + p));
    cplx<double> ya[3], yal[3], nya = cmake(0.0, 0.0);
    for (int c = 0; c < 3; ++c) {
      ya[c] = ldc(y + 2 * (P + c * P + p));
      yal[c] = ldc(y + 2 * (5 * P + c * P + p));
      nya = cadd(nya, cscale(yal[c], nv[c]));
    }
    cplx<double> wc = cmake(0.0, 0.0);
    for (int s = 0; s < 2; ++s) {
      const cplx<double> e = s == 0 ? q.e1 : q.e2;
      const double sg = s == 0 ? -1.0 : 1.0;
      double* l = L + 2 * ((((b * n + i) * 2 + s) * P + p) * 8);
      double* r = R + 2 * ((((b * n + i) * 2 + s) * P + p) * 4);
      cplx<double> lv[8], rv[4];
      lv[0] = cscale(csub(yphi, ik_times(q.k0, cmul(e, nya))), sg);
      lv[4] = cscale(cmul(e, yde), sg);
      for (int c = 0; c < 3; ++c) {
        lv[1 + c] = cscale(csub(ya[c], ik_times(q.k0, cmul(e, cscale(yde, nv[c])))), sg);
        lv[5 + c] = cscale(yal[c], sg);
      }
      const int xo = 4 * P * s;
      rv[0] = ldc(x + 2 * (xo + p));
      for (int c = 0; c < 3; ++c) rv[1 + c] = ldc(x + 2 * (xo + P + c * P + p));
      for (int t = 0; t < 8; ++t) cst(l + 2 * t, lv[t]);
      for (int t = 0; t < 4; ++t) {
        cst(r + 2 * t, rv[t]);
        wc = cadd(wc, cmul(lv[4 + t], rv[t]));
      }
    }
    cst(Wc + 2 * ((b * n + i) * P + p), wc);
  }
}

__device__ __forceinline__ void weights(const double* L, const double* R, long b, int n, int P, int i, int j, int s,
                                        int p, cplx<double>& wg, cplx<double>& wh) {
  const double* l = L + 2 * ((((b * n + i) * 2 + s) * P + p) * 8);
  const double* r = R + 2 * ((((b * n + j) * 2 + s) * P + p) * 4);
  wg = cmake(0.0, 0.0);
  wh = cmake(0.0, 0.0);
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const cplx<double> rv = ldc(r + 2 * t);
    wg = cadd(wg, cmul(ldc(l + 2 * t), rv));
    wh = cadd(wh, cmul(ldc(l + 2 * (4 + t)), rv));
  }
}

// ---------------------------------------------------------------------------
// Far-field pairs (every (i, j), i != j, not refined), both sides:
//   G = a_j e/d, F = a_j (n_i.r) h, r = pos_i - pos_j, e = e^{ikd},
//   h = (ikd - 1) e / d^3, h' = dh/dd = e (3 - 3ikd - k^2 d^2) / d^4,
//   dG/dr = a_j h r, dF/dr = a_j (h n_i + (n_i.r) h' r / d), dF/dn_i = a_j h r,
//   dG/da_j = e/d, dF/da_j = (n_i.r) h,
// and the closed-surface correction corr_j = -2 pi - sum_i term_ij with the
// far terms term_ij = -area_i (n_i.r) / d^3 (i in j's closed surface).
// PASS 0: lane = row i (pos_i, nvec_i, area_i); PASS 1: lane = column j
// (pos_j, area_j).  grid (ceil(n/64), 1, batch): 64 own faces per workgroup,
// its four waves take four quarters of the other index, the quarters' sums
// are added in wave order through LDS and each own face's slots are written
// by one thread (no atomics: the same sums in the same order on every run).
template <int PASS>
__global__ __launch_bounds__(256) void k_grad_far(int n, int P, const double* __restrict__ pos,
                                                  const double* __restrict__ nvec, const double* __restrict__ area,
                                                  const int* __restrict__ mat0, const int* __restrict__ mat1,
                                                  const int* __restrict__ surf, const double* __restrict__ kmat,
                                                  int nmat, const long* __restrict__ rowptr,
                                                  const int* __restrict__ pcol, const int* __restrict__ colptr,
                                                  const int* __restrict__ crow, const double* __restrict__ L,
                                                  const double* __restrict__ R, const double* __restrict__ Wc,
                                                  double* Gr) {
  const long b = blockIdx.z;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const bool act = blockIdx.x * 64 + lane < n;
  const int own = act ? blockIdx.x * 64 + lane : n - 1;
  const int quarter = (n + 3) / 4;
  const int o0 = min(n, wv * quarter), o1 = act ? min(n, o0 + quarter) : o0;
  // sorted refined partners of the own face
  int q, qend;
  const int* part;
  if (PASS == 0) {
    q = (int)rowptr[own];
    qend = (int)rowptr[own + 1];
    part = pcol;
  } else {
    q = colptr[own];
    qend = colptr[own + 1];
    part = crow;
  }
  {
    int lo = q, hi = qend;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (part[m] < o0)
        lo = m + 1;
      else
        hi = m;
    }
    q = lo;
  }
  constexpr int PM = 4;
  double acc[PM][7];
#pragma unroll
  for (int p = 0; p < PM; ++p)
#pragma unroll
    for (int t = 0; t < 7; ++t) acc[p][t] = 0.0;
  const double po[3] = {pos[3 * own], pos[3 * own + 1], pos[3 * own + 2]};
  const int m0o = mat0[own], m1o = mat1[own];
  for (int other = o0; other < o1; ++other) {
    while (q < qend && part[q] < other) ++q;
    if (other == own || (q < qend && part[q] == other)) continue;
    const int i = PASS == 0 ? own : other, j = PASS == 0 ? other : own;
    const double rx = pos[3 * i] - pos[3 * j], ry = pos[3 * i + 1] - pos[3 * j + 1],
                 rz = pos[3 * i + 2] - pos[3 * j + 2];
    const double d = fmax(sqrt(rx * rx + ry * ry + rz * rz), TINY);
    const double id = 1.0 / d;
    const double ni[3] = {nvec[3 * i], nvec[3 * i + 1], nvec[3 * i + 2]};
    const double nr = ni[0] * rx + ni[1] * ry + ni[2] * rz;
    const double aj = area[j];
    const double r3[3] = {rx, ry, rz};
    for (int s = 0; s < 2; ++s) {
      const int mi = s == 0 ? (PASS == 0 ? m0o : mat0[i]) : (PASS == 0 ? m1o : mat1[i]);
      const int mj = s == 0 ? (PASS == 0 ? mat0[j] : m0o) : (PASS == 0 ? mat1[j] : m1o);
      if (mi != mj) continue;
      const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
      double sn, cs;
      sincos(kr * d, &sn, &cs);
      const double ex = exp(-ki * d);
      const cplx<double> e = cmake(ex * cs, ex * sn);
      const cplx<double> ikd = cmake(-ki * d, kr * d);
      const cplx<double> g0 = cscale(e, id);
      const cplx<double> h = cscale(cmul(cadd(ikd, cmake(-1.0, 0.0)), e), id * id * id);
      // 3 - 3ikd - (kd)^2 with (kd)^2 = (kr^2 - ki^2 + 2i kr ki) d^2
      const cplx<double> poly = cmake(3.0 - 3.0 * ikd.re - (kr * kr - ki * ki) * d * d,
                                      -3.0 * ikd.im - 2.0 * kr * ki * d * d);
      const cplx<double> hp = cscale(cmul(poly, e), id * id * id * id);
      for (int p = 0; p < P && p < PM; ++p) {
        cplx<double> wg, wh;
        weights(L, R, b, n, P, i, j, s, p, wg, wh);
        const double gh = im_mul(wg, h) * aj;              // Im(WG a_j h)
        const double hh = im_mul(wh, h) * aj;              // Im(WH a_j h)
        const double hhp = im_mul(wh, hp) * aj * nr * id;  // Im(WH a_j h') (n.r)/d
        for (int c = 0; c < 3; ++c) {
          const double gp = gh * r3[c] + hh * ni[c] + hhp * r3[c];
          if (PASS == 0) {
            acc[p][c] += gp;
            acc[p][3 + c] += hh * r3[c];
          } else {
            acc[p][c] -= gp;
          }
        }
        if (PASS == 1) acc[p][3] += im_mul(wg, g0) + im_mul(wh, h) * nr;
      }
    }
    if (surf[i] == surf[j]) {
      const double ai = area[i];
      const double id3 = id * id * id, id5 = id3 * id * id;
      for (int p = 0; p < P && p < PM; ++p) {
        const double wc = Wc[2 * ((b * n + j) * P + p) + 1];
        for (int c = 0; c < 3; ++c) {
          const double gp = wc * ai * (ni[c] * id3 - 3.0 * nr * r3[c] * id5);
          if (PASS == 0) {
            acc[p][c] += gp;
            acc[p][3 + c] += wc * ai * r3[c] * id3;
          } else {
            acc[p][c] -= gp;
          }
        }
        if (PASS == 0) acc[p][6] += wc * nr * id3;
      }
    }
  }
  __shared__ double wsum[3][64][PM * 7];
  if (wv > 0)
#pragma unroll
    for (int p = 0; p < PM; ++p)
#pragma unroll
      for (int t = 0; t < 7; ++t) wsum[wv - 1][lane][p * 7 + t] = acc[p][t];
  __syncthreads();
  if (wv > 0 || !act) return;
  for (int w = 0; w < 3; ++w)
#pragma unroll
    for (int p = 0; p < PM; ++p)
#pragma unroll
      for (int t = 0; t < 7; ++t) acc[p][t] += wsum[w][lane][p * 7 + t];
  for (int p = 0; p < P && p < PM; ++p) {
    double* gr = Gr + ((b * P + p) * n + own) * (long)NG;
    for (int c = 0; c < 3; ++c) gr[S_POS + c] += acc[p][c];
    if (PASS == 0) {
      for (int c = 0; c < 3; ++c) gr[S_NVEC + c] += acc[p][3 + c];
      gr[S_AREA] += acc[p][6];
    } else {
      gr[S_AREA] += acc[p][3];
    }
  }
}

// ---------------------------------------------------------------------------
// forward-mode dual numbers with 3 tangent directions
struct D3 {
  double v, d[3];
};
__device__ __forceinline__ D3 dc(double v) { return D3{v, {0.0, 0.0, 0.0}}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) {
  return D3{a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
__device__ __forceinline__ D3 operator-(D3 a, D3 b) {
  return D3{a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
__device__ __forceinline__ D3 operator*(D3 a, D3 b) {
  return D3{a.v * b.v,
            {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
__device__ __forceinline__ D3 operator*(double s, D3 a) { return D3{s * a.v, {s * a.d[0], s * a.d[1], s * a.d[2]}}; }
__device__ __forceinline__ D3 operator*(D3 a, double s) { return s * a; }
__device__ __forceinline__ D3 drecip(D3 a) {
  const double r = 1.0 / a.v, r2 = -r * r;
  return D3{r, {a.d[0] * r2, a.d[1] * r2, a.d[2] * r2}};
}
__device__ __forceinline__ D3 dsqrt(D3 a) {
  const double s = sqrt(a.v), h = 0.5 / s;
  return D3{s, {a.d[0] * h, a.d[1] * h, a.d[2] * h}};
}
__device__ __forceinline__ D3 dmax(D3 a, D3 b) { return a.v >= b.v ? a : b; }
__device__ __forceinline__ D3 dmaxc(D3 a, double b) { return a.v >= b ? a : dc(b); }

struct QP3 {
  D3 p[3];
  D3 w;
};

// point and weight of quadrature node tq on the 6-node face X, with tangents
// along the 3 coordinates of node c0 (c0 < 0: no tangent)
__device__ __forceinline__ QP3 face_point_d3(const double* X, const double* tq, int c0) {
  double p[3] = {0, 0, 0}, tx[3] = {0, 0, 0}, ty[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = X[3 * c + d];
      p[d] += tq[c] * x;
      tx[d] += tq[6 + c] * x;
      ty[d] += tq[12 + c] * x;
    }
  }
  const double sp = c0 >= 0 ? tq[c0] : 0.0, sx = c0 >= 0 ? tq[6 + c0] : 0.0, sy = c0 >= 0 ? tq[12 + c0] : 0.0;
  QP3 q;
  D3 TX[3], TY[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    q.p[d] = dc(p[d]);
    TX[d] = dc(tx[d]);
    TY[d] = dc(ty[d]);
    q.p[d].d[d] = sp;
    TX[d].d[d] = sx;
    TY[d].d[d] = sy;
  }
  const D3 nx = TX[1] * TY[2] - TX[2] * TY[1];
  const D3 ny = TX[2] * TY[0] - TX[0] * TY[2];
  const D3 nz = TX[0] * TY[1] - TX[1] * TY[0];
  q.w = (0.5 * tq[18]) * dsqrt(nx * nx + ny * ny + nz * nz);
  return q;
}

// Refined pairs: the coefficients of k_refine_coeffs (greenret_refined.py:
// 122-516) evaluated with dual numbers along one group of 3 inputs, then
//   G = (sum_n g_n (ik)^n) e^{ikd},  H = (sum_n n_i.f_n (ik)^n) e^{ikd},
// differentiated and contracted with the weights, plus the closed-surface
// term area_i fstat / a_j of corr_j.  grid (ceil(npairs/256), 11 groups):
// group 0 pos_i, 1 nvec_i, 2 tvec1_i, 3 tvec2_i, 4..9 node (g-4) of face j,
// 10 pos_j.  Each (pair, group) writes its 3 derivative values (and, group 0,
// the two area terms of the closed-surface correction for faces i and j) to
// the work space Ws[((b P + p) 11 + group) npairs + pair][5]; k_grad_refined_sum
// adds them per face in a fixed order (no atomics).
template <int ORD>
__global__ __launch_bounds__(256) void k_grad_refined(
    int n, int P, int batch, int npairs, const int* __restrict__ prow, const int* __restrict__ pcol,
    const int* __restrict__ pkind, const double* __restrict__ pos, const double* __restrict__ nvec,
    const double* __restrict__ tvec1, const double* __restrict__ tvec2, const double* __restrict__ area,
    const double* __restrict__ nodes, const int* __restrict__ mat0, const int* __restrict__ mat1,
    const int* __restrict__ surf, const double* __restrict__ kmat, int nmat, const double* __restrict__ qnear,
    int nq_near, const double* __restrict__ qself, int nq_self, const double* __restrict__ L,
    const double* __restrict__ R, const double* __restrict__ Wc, double* __restrict__ Ws) {
  extern __shared__ double qs[];
  double* qn = qs;
  double* qp = qs + nq_near * QSTRIDE;
  for (int e = threadIdx.x; e < nq_near * QSTRIDE; e += blockDim.x) qn[e] = qnear[e];
  for (int e = threadIdx.x; e < nq_self * QSTRIDE; e += blockDim.x) qp[e] = qself[e];
  __syncthreads();
  const int pr = blockIdx.x * blockDim.x + threadIdx.x;
  const int grp = blockIdx.y;
  if (pr >= npairs) return;
  const int i = prow[pr], j = pcol[pr], kind = pkind[pr];
  double X[18];
#pragma unroll
  for (int e = 0; e < 18; ++e) X[e] = nodes[18 * (long)j + e];
  const int c0 = (grp >= 4 && grp <= 9) ? grp - 4 : -1;
  D3 pi_[3], ni[3], t1[3], t2[3], pj[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    pi_[c] = dc(pos[3 * i + c]);
    ni[c] = dc(nvec[3 * i + c]);
    t1[c] = dc(tvec1[3 * i + c]);
    t2[c] = dc(tvec2[3 * i + c]);
    pj[c] = dc(pos[3 * j + c]);
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (grp == 0) pi_[c].d[c] = 1.0;
    if (grp == 1) ni[c].d[c] = 1.0;
    if (grp == 2) t1[c].d[c] = 1.0;
    if (grp == 3) t2[c].d[c] = 1.0;
    if (grp == 10) pj[c].d[c] = 1.0;
  }
  double fact[ORD + 1];
  fact[0] = 1.0;
#pragma unroll
  for (int m = 1; m <= ORD; ++m) fact[m] = fact[m - 1] * m;
  D3 gs[ORD + 1], fx[ORD + 1], fy[ORD + 1], fz[ORD + 1];
#pragma unroll
  for (int m = 0; m <= ORD; ++m) gs[m] = fx[m] = fy[m] = fz[m] = dc(0.0);
  D3 fq = dc(0.0);
  if (kind == 2) {
    D3 rmax = dc(0.0);
    for (int qq = 0; qq < nq_self; ++qq) {
      const QP3 Q = face_point_d3(X, qp + qq * QSTRIDE, c0);
      const D3 vx = pi_[0] - Q.p[0], vy = pi_[1] - Q.p[1], vz = pi_[2] - Q.p[2];
      rmax = dmax(rmax, dmaxc(dsqrt(vx * vx + vy * vy + vz * vz), TINY));
    }
    D3 f1[ORD + 1], f2[ORD + 1], f3[ORD + 1];
#pragma unroll
    for (int m = 0; m <= ORD; ++m) f1[m] = f2[m] = f3[m] = dc(0.0);
    for (int qq = 0; qq < nq_self; ++qq) {
      const QP3 Q = face_point_d3(X, qp + qq * QSTRIDE, c0);
      const D3 vx = pi_[0] - Q.p[0], vy = pi_[1] - Q.p[1], vz = pi_[2] - Q.p[2];
      const D3 r = dmaxc(dsqrt(vx * vx + vy * vy + vz * vz), TINY);
      const D3 rr = dmax(r, 1e-4 * rmax);
      const D3 ndr = ni[0] * vx + ni[1] * vy + ni[2] * vz;
      const D3 in1 = t1[0] * vx + t1[1] * vy + t1[2] * vz;
      const D3 in2 = t2[0] * vx + t2[1] * vy + t2[2] * vz;
      const D3 w = Q.w;
      const D3 ir = drecip(r);
      const D3 irr3 = drecip(rr * rr * rr);
      D3 rnm1 = ir, rnm3 = ir * ir * ir, rn = dc(1.0);
#pragma unroll
      for (int m = 0; m <= ORD; ++m) {
        gs[m] = gs[m] + w * rnm1;
        f1[m] = f1[m] + (double)(m - 1) * (w * ndr * rnm3);
        f2[m] = f2[m] + (double)(m - 1) * (w * in1 * rn * irr3);
        f3[m] = f3[m] + (double)(m - 1) * (w * in2 * rn * irr3);
        rnm1 = rnm1 * r;
        rnm3 = rnm3 * r;
        rn = rn * r;
      }
      fq = fq - w * ndr * ir * ir * ir;
    }
    // The static normal part of the self term, n.f_0 = (n.n) f1[0] with
    // f1[0] = fstat, cancels against the self term of the closed-surface
    // correction in H_ii = n.(f_0 + corr_i n): H_ii(order 0) = (n.t1) f2[0] +
    // (n.t2) f3[0] + (n.n)(-2 pi - sum_{k != i} term_ki).  Both parts are of
    // order w / r^3 at the polar nodes next to the centroid, so they are
    // dropped here together (the k != i sum is differentiated elsewhere).
    f1[0] = dc(0.0);
#pragma unroll
    for (int m = 0; m <= ORD; ++m) {
      const double fi = 1.0 / fact[m];
      gs[m] = fi * gs[m];
      const D3 a = fi * f1[m], b2 = fi * f2[m], c2 = fi * f3[m];
      fx[m] = ni[0] * a + t1[0] * b2 + t2[0] * c2;
      fy[m] = ni[1] * a + t1[1] * b2 + t2[1] * c2;
      fz[m] = ni[2] * a + t1[2] * b2 + t2[2] * c2;
    }
  } else {
    const D3 dx0 = pj[0] - pi_[0], dy0 = pj[1] - pi_[1], dz0 = pj[2] - pi_[2];
    const D3 r0 = dsqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
    for (int qq = 0; qq < nq_near; ++qq) {
      const QP3 Q = face_point_d3(X, qn + qq * QSTRIDE, c0);
      const D3 vx = pi_[0] - Q.p[0], vy = pi_[1] - Q.p[1], vz = pi_[2] - Q.p[2];
      const D3 r = dmaxc(dsqrt(vx * vx + vy * vy + vz * vz), TINY);
      const D3 dl = r - r0;
      const D3 w = Q.w;
      const D3 ir = drecip(r), ir2 = ir * ir, ir3 = ir2 * ir;
      D3 dn = dc(1.0), dnm1 = dc(0.0);
#pragma unroll
      for (int m = 0; m <= ORD; ++m) {
        gs[m] = gs[m] + (1.0 / fact[m]) * (w * dn * ir);
        const double fnm1 = fact[m > 0 ? m - 1 : 0];
        const D3 fsc = (m == 0) ? dc(0.0) - ir3 : (dc(0.0) - (1.0 / fact[m]) * (dn * ir3)) + (1.0 / fnm1) * (dnm1 * ir2);
        fx[m] = fx[m] + w * vx * fsc;
        fy[m] = fy[m] + w * vy * fsc;
        fz[m] = fz[m] + w * vz * fsc;
        dnm1 = dn;
        dn = dn * dl;
      }
      fq = fq - w * (ni[0] * vx + ni[1] * vy + ni[2] * vz) * ir3;
    }
  }
  // phase distance (centroids) and its tangent
  const double rx = pos[3 * i] - pos[3 * j], ry = pos[3 * i + 1] - pos[3 * j + 1], rz = pos[3 * i + 2] - pos[3 * j + 2];
  const double d = fmax(sqrt(rx * rx + ry * ry + rz * rz), TINY);
  double dd[3] = {0.0, 0.0, 0.0};
  if (i != j && (grp == 0 || grp == 10)) {
    const double sg = grp == 0 ? 1.0 / d : -1.0 / d;
    dd[0] = sg * rx;
    dd[1] = sg * ry;
    dd[2] = sg * rz;
  }
  const double nv[3] = {ni[0].v, ni[1].v, ni[2].v};
  // n_i.f_n and its tangents (including the explicit n_i of H = n_i.F)
  double nf[ORD + 1], dnf[ORD + 1][3];
#pragma unroll
  for (int m = 0; m <= ORD; ++m) {
    nf[m] = nv[0] * fx[m].v + nv[1] * fy[m].v + nv[2] * fz[m].v;
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      dnf[m][t] = nv[0] * fx[m].d[t] + nv[1] * fy[m].d[t] + nv[2] * fz[m].d[t];
      if (grp == 1) dnf[m][t] += (t == 0 ? fx[m].v : t == 1 ? fy[m].v : fz[m].v);
    }
  }
  const bool same_surf = surf[i] == surf[j];
  const double ai = area[i], aj = area[j];
  for (long b = 0; b < batch; ++b) {
    cplx<double> ag[2], af[2], dag[2][3], daf[2][3], E[2];
    bool on[2];
    for (int s = 0; s < 2; ++s) {
      const int mi = s == 0 ? mat0[i] : mat1[i];
      on[s] = mi == (s == 0 ? mat0[j] : mat1[j]);
      if (!on[s]) continue;
      const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
      const cplx<double> ik = cmake(-ki, kr);
      double sn, cs;
      sincos(kr * d, &sn, &cs);
      const double ex = exp(-ki * d);
      E[s] = cmake(ex * cs, ex * sn);
      cplx<double> pw = cmake(1.0, 0.0);
      ag[s] = af[s] = cmake(0.0, 0.0);
      for (int t = 0; t < 3; ++t) dag[s][t] = daf[s][t] = cmake(0.0, 0.0);
#pragma unroll
      for (int m = 0; m <= ORD; ++m) {
        ag[s] = cadd(ag[s], cscale(pw, gs[m].v));
        af[s] = cadd(af[s], cscale(pw, nf[m]));
#pragma unroll
        for (int t = 0; t < 3; ++t) {
          dag[s][t] = cadd(dag[s][t], cscale(pw, gs[m].d[t]));
          daf[s][t] = cadd(daf[s][t], cscale(pw, dnf[m][t]));
        }
        pw = cmul(pw, ik);
      }
      // d/dt [A e^{ikd}] = (dA + ik A dd) e^{ikd}
      for (int t = 0; t < 3; ++t) {
        dag[s][t] = cmul(cadd(dag[s][t], cscale(cmul(ik, ag[s]), dd[t])), E[s]);
        daf[s][t] = cmul(cadd(daf[s][t], cscale(cmul(ik, af[s]), dd[t])), E[s]);
      }
    }
    for (int p = 0; p < P; ++p) {
      double out[3] = {0.0, 0.0, 0.0};
      for (int s = 0; s < 2; ++s) {
        if (!on[s]) continue;
        cplx<double> wg, wh;
        weights(L, R, b, n, P, i, j, s, p, wg, wh);
        for (int t = 0; t < 3; ++t) out[t] += im_mul(wg, dag[s][t]) + im_mul(wh, daf[s][t]);
      }
      double* ws = Ws + ((((b * P + p) * 11 + grp) * (long)npairs) + pr) * 5;
      double ar_i = 0.0, ar_j = 0.0;
      if (i == j && grp == 1) {
        // d(n.n)/dn * fstat of the cancelled static part (n.n) (fstat + corr_i);
        // the corr_i half is added with the per-face terms
        const double wc = Wc[2 * ((b * n + j) * P + p) + 1];
        for (int t = 0; t < 3; ++t) out[t] += 2.0 * nv[t] * fq.v * wc;
      }
      if (same_surf && i != j) {
        const double wc = Wc[2 * ((b * n + j) * P + p) + 1];
        for (int t = 0; t < 3; ++t) out[t] -= wc * ai * fq.d[t] / aj;
        if (grp == 0) {
          ar_i = -wc * fq.v / aj;
          ar_j = wc * ai * fq.v / (aj * aj);
        }
      }
      for (int t = 0; t < 3; ++t) ws[t] = out[t];
      ws[3] = ar_i;
      ws[4] = ar_j;
    }
  }
}

// Per face and group, the sums of k_grad_refined's work space into Gr in a
// fixed order: groups 0..3 (derivatives at face i) over the face's pairs by
// row, groups 4..9 (nodes of face j) over its pairs by column; the thread of
// group 0 also adds group 10 (pos_j, by column) to the same position slots and
// both area terms of the closed-surface correction (by row, then by column).
// grid (ceil(n/256), 10, batch P).
__global__ __launch_bounds__(256) void k_grad_refined_sum(int n, int npairs, const long* __restrict__ rowptr,
                                                          const int* __restrict__ colptr,
                                                          const int* __restrict__ cpair,
                                                          const double* __restrict__ Ws, double* Gr) {
  const int f = blockIdx.x * 256 + threadIdx.x;
  const int grp = blockIdx.y;
  const long bp = blockIdx.z;
  if (f >= n) return;
  const double* wg = Ws + (bp * 11 + grp) * (long)npairs * 5;
  const int slot = grp == 0 ? S_POS : grp == 1 ? S_NVEC : grp == 2 ? S_T1 : grp == 3 ? S_T2
                 : S_NODES + 3 * (grp - 4);
  double s[3] = {0.0, 0.0, 0.0}, sa = 0.0;
  if (grp <= 3) {
    for (long q = rowptr[f]; q < rowptr[f + 1]; ++q) {
      const double* w = wg + q * 5;
      s[0] += w[0];
      s[1] += w[1];
      s[2] += w[2];
      sa += w[3];
    }
  } else {
    for (int q = colptr[f]; q < colptr[f + 1]; ++q) {
      const double* w = wg + (long)cpair[q] * 5;
      s[0] += w[0];
      s[1] += w[1];
      s[2] += w[2];
    }
  }
  if (grp == 0) {
    const double* w10 = Ws + (bp * 11 + 10) * (long)npairs * 5;
    for (int q = colptr[f]; q < colptr[f + 1]; ++q) {
      const double* w = w10 + (long)cpair[q] * 5;
      s[0] += w[0];
      s[1] += w[1];
      s[2] += w[2];
      sa += wg[(long)cpair[q] * 5 + 4];
    }
  }
  double* gr = Gr + (bp * n + f) * (long)NG;
  for (int t = 0; t < 3; ++t) gr[slot + t] += s[t];
  if (grp == 0) gr[S_AREA] += sa;
}

// ---------------------------------------------------------------------------
// Per-face terms: the explicit normals of A (Q = [G1 X1 | G2 X2]), the plane
// wave right side (all of b_i is proportional to amp = e^{ik dir.pos}, so
// db_i/dpos = ik dir b_i) and the extinction functional (ph = e^{-ik dir.pos}
// area).  Rb: the right side R; X, Y, Q: (batch, n, 8P).
__global__ void k_grad_local(int n, int P, const double* X, const double* Y, const double* Q, const double* Rb,
                             long ld, long stride, const double* pos, const double* nvec, const double* area,
                             const int* in_med, const int* out_med, const double* pol, const double* dir,
                             const double* pw, const double* corr, const double* Wc, double* Gr) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* q = pw + 7 * b;
  const double k0 = q[0], kr = q[1], ki = q[2];
  const cplx<double> e1 = cmake(q[3], q[4]), e2 = cmake(q[5], q[6]);
  const cplx<double> ik = cmake(-ki, kr);
  const double kap = 4.0 * M_PI * kr / (kr * kr + ki * ki);
  const long o = 2 * (b * stride + (long)i * ld);
  const double* x = X + o;
  const double* y = Y + o;
  const double* qv = Q + o;
  const double* rb = Rb + o;
  const double nv[3] = {nvec[3 * i], nvec[3 * i + 1], nvec[3 * i + 2]};
  const int fin = in_med[i], fout = out_med[i];
  const double sig = (double)(fout - fin);
  for (int p = 0; p < P; ++p) {
    double* gr = Gr + ((b * P + p) * n + i) * (long)NG;
    const double* dp = dir + 3 * p;
    const double* pp = pol + 3 * p;
    double gpos[3] = {0, 0, 0}, gnv[3] = {0, 0, 0}, garea = 0.0;
    // explicit normals of A
    const cplx<double> yde = ldc(y + 2 * (4 * P + p));
    const cplx<double> q1s = ldc(qv + 2 * p), q2s = ldc(qv + 2 * (4 * P + p));
    for (int c = 0; c < 3; ++c) {
      const cplx<double> yal = ldc(y + 2 * (5 * P + c * P + p));
      const cplx<double> q1h = ldc(qv + 2 * (P + c * P + p)), q2h = ldc(qv + 2 * (5 * P + c * P + p));
      const cplx<double> t = cadd(cmul(yal, csub(cmul(e1, q1s), cmul(e2, q2s))),
                                  cmul(yde, csub(cmul(e1, q1h), cmul(e2, q2h))));
      gnv[c] += ik_times(k0, t).im;
      // the self term's closed-surface correction enters H_ii as corr_i (n_i.n_i)
      gnv[c] += 2.0 * corr[i] * nv[c] * Wc[2 * ((b * n + i) * P + p) + 1];
    }
    // right side
    const int cols[8] = {p, P + p, 2 * P + p, 3 * P + p, 4 * P + p, 5 * P + p, 6 * P + p, 7 * P + p};
    cplx<double> yb = cmake(0.0, 0.0);
    for (int u = 0; u < 8; ++u) yb = cadd(yb, cmul(ldc(y + 2 * cols[u]), ldc(rb + 2 * cols[u])));
    const double iyb = im_mul(ik, yb);
    const double kd = pos[3 * i] * dp[0] + pos[3 * i + 1] * dp[1] + pos[3 * i + 2] * dp[2];
    const cplx<double> ph = cexpi(kr * kd, ki * kd);
    const cplx<double> amp = cmake(ph.im / k0, -ph.re / k0);
    cplx<double> ypa = cmake(0.0, 0.0);  // y_al . pol
    for (int c = 0; c < 3; ++c) ypa = cadd(ypa, cscale(ldc(y + 2 * (5 * P + c * P + p)), pp[c]));
    const double t_dir = im_mul(cscale(cmul(ik, amp), sig), ypa);
    const cplx<double> ede = csub(cscale(e2, (double)fout), cscale(e1, (double)fin));
    // -ik0 (e2 out - e1 in) amp y_De
    const double t_pol = ik_times(-k0, cmul(cmul(ede, amp), yde)).im;
    for (int c = 0; c < 3; ++c) {
      gpos[c] += iyb * dp[c];
      gnv[c] += t_dir * dp[c] + t_pol * pp[c];
    }
    // extinction functional
    const cplx<double> phc = cscale(cexpi(-kr * kd, -ki * kd), kap);
    const double pd = pp[0] * dp[0] + pp[1] * dp[1] + pp[2] * dp[2];
    cplx<double> T = cmake(0.0, 0.0);
    for (int side = 0; side < 2; ++side) {
      if (!(side == 0 ? fin : fout)) continue;
      const int off = 4 * P * side;
      cplx<double> phh = cmake(0.0, 0.0);
      for (int c = 0; c < 3; ++c) phh = cadd(phh, cscale(ldc(x + 2 * (off + P + c * P + p)), pp[c]));
      const cplx<double> sg = ldc(x + 2 * (off + p));
      T = cadd(T, csub(ik_times(k0, phh), cscale(cmul(ik, sg), pd)));
    }
    T = cmul(T, phc);  // per unit area
    const double ikT = im_mul(cmake(ki, -kr), T) * area[i];  // Im(-ik T area)
    for (int c = 0; c < 3; ++c) gpos[c] += ikT * dp[c];
    garea += T.im;
    for (int c = 0; c < 3; ++c) {
      gr[S_POS + c] += gpos[c];
      gr[S_NVEC + c] += gnv[c];
    }
    gr[S_AREA] += garea;
  }
}

}  // namespace

extern "C" int mnpbem_ext_functional(int n, int P, const double* pos, const double* area, const int* in_med,
                                     const int* out_med, const double* pol, const double* dir, const double* pw,
                                     void* C, long ldv, long sv, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && ldv >= 8 * P, "ext_functional: bad size");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_ext_functional, dim3(mnp_ceil_div(n, 256), batch), dim3(256), 0, st, n, P, pos, area, in_med,
                     out_med, pol, dir, pw, (double*)C, ldv, sv);
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_grad_prep(int n, int P, const void* X, const void* Y, long ldv, long sv, const double* nvec,
                                const double* par, void* L, void* R, void* Wc, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P <= 4 && ldv >= 8 * P, "grad_prep: bad size (P <= 4)");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_grad_prep, dim3(mnp_ceil_div(n, 256), batch), dim3(256), 0, st, n, P, (const double*)X,
                     (const double*)Y, ldv, sv, nvec, par, (double*)L, (double*)R, (double*)Wc);
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_grad_far(int n, int P, const double* pos, const double* nvec, const double* area,
                               const int* mat0, const int* mat1, const int* surf, const double* kmat, int nmat,
                               const long* rowptr, const int* pcol, const int* colptr, const int* crow,
                               const void* L, const void* R, const void* Wc, double* Gr, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && P <= 4, "grad_far: bad size (P <= 4)");
  if (batch <= 0) return 0;
  dim3 grid(mnp_ceil_div(n, 64), 1, batch);
  hipLaunchKernelGGL(k_grad_far<0>, grid, dim3(256), 0, st, n, P, pos, nvec, area, mat0, mat1, surf, kmat, nmat,
                     rowptr, pcol, colptr, crow, (const double*)L, (const double*)R, (const double*)Wc, Gr);
  MNP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_grad_far<1>, grid, dim3(256), 0, st, n, P, pos, nvec, area, mat0, mat1, surf, kmat, nmat,
                     rowptr, pcol, colptr, crow, (const double*)L, (const double*)R, (const double*)Wc, Gr);
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_grad_refined(int order, int n, int P, int npairs, const int* prow, const int* pcol,
                                   const int* pkind, const double* pos, const double* nvec, const double* tvec1,
                                   const double* tvec2, const double* area, const double* nodes, const int* mat0,
                                   const int* mat1, const int* surf, const double* kmat, int nmat,
                                   const double* qnear, int nq_near, const double* qself, int nq_self,
                                   const void* L, const void* R, const void* Wc, const long* rowptr,
                                   const int* colptr, const int* cpair, double* Ws, long ws_len, double* Gr,
                                   int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(order >= 0 && order <= 7, "grad_refined: order must be in 0..7");
  MNP_REQUIRE(n > 0 && P > 0 && P <= 4, "grad_refined: bad size (P <= 4)");
  MNP_REQUIRE(nq_near > 0 && nq_self > 0 && (nq_near + nq_self) * QSTRIDE * 8 <= 60000,
              "grad_refined: quadrature tables too large");
  if (npairs == 0 || batch <= 0) return 0;
  MNP_REQUIRE(ws_len >= (long)batch * P * 11 * npairs * 5,
              "grad_refined: work space below batch * P * 11 * npairs * 5 doubles");
  const size_t shm = (size_t)(nq_near + nq_self) * QSTRIDE * sizeof(double);
  dim3 grid(mnp_ceil_div(npairs, 256), 11);
#define MNP_GR(O)                                                                                                  \
  case O:                                                                                                          \
    hipLaunchKernelGGL(k_grad_refined<O>, grid, dim3(256), shm, st, n, P, batch, npairs, prow, pcol, pkind, pos,   \
                       nvec, tvec1, tvec2, area, nodes, mat0, mat1, surf, kmat, nmat, qnear, nq_near, qself,       \
                       nq_self, (const double*)L, (const double*)R, (const double*)Wc, Ws);                        \
    break;
  switch (order) {
    MNP_GR(0) MNP_GR(1) MNP_GR(2) MNP_GR(3) MNP_GR(4) MNP_GR(5) MNP_GR(6) MNP_GR(7)
  }
#undef MNP_GR
  MNP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_grad_refined_sum, dim3(mnp_ceil_div(n, 256), 10, batch * P), dim3(256), 0, st, n, npairs,
                     rowptr, colptr, cpair, (const double*)Ws, Gr);
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_grad_local(int n, int P, const void* X, const void* Y, const void* Q, const void* Rb, long ldv,
                                 long sv, const double* pos, const double* nvec, const double* area,
                                 const int* in_med, const int* out_med, const double* pol, const double* dir,
                                 const double* pw, const double* corr, const void* Wc, double* Gr, int batch,
                                 void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P > 0 && ldv >= 8 * P, "grad_local: bad size");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_grad_local, dim3(mnp_ceil_div(n, 256), batch), dim3(256), 0, st, n, P, (const double*)X,
                     (const double*)Y, (const double*)Q, (const double*)Rb, ldv, sv, pos, nvec, area, in_med,
                     out_med, pol, dir, pw, corr, (const double*)Wc, Gr);
  MNP_LAUNCH_CHECK();
  return 0;
}
