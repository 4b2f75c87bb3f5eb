// Retarded Green functions on curved boundary elements (gfx950).
//
// Set-up (once per geometry):
//   mnpbem_refine_coeffs    - order-n (ik)^n expansion coefficients of the
//                             refined (self and near) entries, deriv='cart'
//                             (mnpbem/greenfun/greenret_refined.py:122-516,
//                             MATLAB Greenfun/@greenret/private/init.m:42-177)
//   mnpbem_closed_correction - Fuchs-Liu closed-surface correction added to
//                             the order-0 coefficient of the self terms
//                             (mnpbem/greenfun/compgreen_ret.py:192-400,
//                             compgreen_stat.py:219-465)
// Per wavelength batch:
//   mnpbem_green_assemble   - G = area_j/d e^{ikd}, F = n_i.r (ik-1/d)/d^2
//                             area_j e^{ikd} for all face pairs of one side
//                             (greenret_refined.py:654-800), then the refined
//                             entries sum_n c_n (ik)^n e^{ikd}, and +/-2 pi on
//                             the self diagonal for H1/H2 (:800-810).
// Matrices are written transposed: entry (i, j) lands at out[j*ld + i], so a
// row-major LU of the buffer factors the transpose (the LU solve then yields
// H G^-1 directly; see mnpbem_getrs_batched).
#include "common.h"
#include "green_entry.h"

namespace {

constexpr int QSTRIDE = 19;  // quadrature table row: shape[6], d/dx[6], d/dy[6], weight

struct QPoint {
  double p[3];
  double w;
};

// point and weight of quadrature node q on the curved 6-node face X[6][3]:
// p = sum N_c X_c,  w = 0.5 * w_q * |dX/dx x dX/dy|  (particle.py:1364-1429,
// 1520-1575)
__device__ __forceinline__ QPoint face_point(const double* __restrict__ X, const double* __restrict__ tq) {
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
  const double nx = tx[1] * ty[2] - tx[2] * ty[1];
  const double ny = tx[2] * ty[0] - tx[0] * ty[2];
  const double nz = tx[0] * ty[1] - tx[1] * ty[0];
  QPoint q;
  q.p[0] = p[0];
  q.p[1] = p[1];
  q.p[2] = p[2];
  q.w = 0.5 * tq[18] * sqrt(nx * nx + ny * ny + nz * nz);
  return q;
}

constexpr int QSTRIDE9 = 28;  // quadrilateral table row: shape[9], d/dx[9], d/dy[9], weight

// point and weight of quadrature node q on the curved 9-node quadrilateral
// X[9][3] over [-1, 1]^2: w = w_q |dX/dx x dX/dy| (particle.py:1405-1429)
__device__ __forceinline__ QPoint face_point9(const double* __restrict__ X, const double* __restrict__ tq) {
  double p[3] = {0, 0, 0}, tx[3] = {0, 0, 0}, ty[3] = {0, 0, 0};
#pragma unroll
  for (int c = 0; c < 9; ++c) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double x = X[3 * c + d];
      p[d] += tq[c] * x;
      tx[d] += tq[9 + c] * x;
      ty[d] += tq[18 + c] * x;
    }
  }
  const double nx = tx[1] * ty[2] - tx[2] * ty[1];
  const double ny = tx[2] * ty[0] - tx[0] * ty[2];
  const double nz = tx[0] * ty[1] - tx[1] * ty[0];
  QPoint q;
  q.p[0] = p[0];
  q.p[1] = p[1];
  q.p[2] = p[2];
  q.w = tq[27] * sqrt(nx * nx + ny * ny + nz * nz);
  return q;
}

// the 6 nodes of sub-triangle `sub` of a face in the 9-node layout: a
// triangle is (v0, v1, v2, m01, m12, m20); a quadrilateral integrates as
// (v0, v1, v2, m01, m12, c) and (v2, v3, v0, m23, m30, c)
// (particle.py:695-720)
__device__ __forceinline__ void sub_triangle(const double* __restrict__ X9, int nv, int sub, double* X) {
  const int tri[6] = {0, 1, 2, 4, 5, 6}, q1[6] = {0, 1, 2, 4, 5, 8}, q2[6] = {2, 3, 0, 6, 7, 8};
  const int* ix = nv == 3 ? tri : (sub == 0 ? q1 : q2);
#pragma unroll
  for (int c = 0; c < 6; ++c)
#pragma unroll
    for (int d = 0; d < 3; ++d) X[3 * c + d] = X9[3 * ix[c] + d];
}

template <int ORD>
__global__ __launch_bounds__(256) void k_refine_coeffs(
    int npairs, const int* __restrict__ prow, const int* __restrict__ pcol, const int* __restrict__ pkind,
    const double* __restrict__ pos1, const double* __restrict__ nvec1, const double* __restrict__ tvec1,
    const double* __restrict__ tvec2, const double* __restrict__ pos2, const double* __restrict__ nodes9,
    const int* __restrict__ nvert2, const double* __restrict__ qnear, int nq_near,
    const double* __restrict__ qself, int nq_self, const double* __restrict__ qselfq, int nq_selfq,
    double* __restrict__ g, double* __restrict__ f, double* __restrict__ fstat) {
  extern __shared__ double qs[];
  double* qn = qs;
  double* qp = qs + nq_near * QSTRIDE;
  for (int e = threadIdx.x; e < nq_near * QSTRIDE; e += blockDim.x) qn[e] = qnear[e];
  for (int e = threadIdx.x; e < nq_self * QSTRIDE; e += blockDim.x) qp[e] = qself[e];
  __syncthreads();
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int i = prow[p], j = pcol[p], kind = pkind[p];
  double X9[27], X[18];
#pragma unroll
  for (int e = 0; e < 27; ++e) X9[e] = nodes9[27 * (long)j + e];
  const int nv = nvert2[j];
  sub_triangle(X9, nv, 0, X);
  // self terms of a quadrilateral: the 9-node shape functions with the
  // quadrilateral polar rule (read from global memory); of a triangle: the
  // 6-node shape functions with the triangle polar rule
  const bool quad = nv == 4;
  const int nqs = quad ? nq_selfq : nq_self;
  auto self_point = [&](int q) { return quad ? face_point9(X9, qselfq + q * QSTRIDE9) : face_point(X, qp + q * QSTRIDE); };
  const double pi_[3] = {pos1[3 * i], pos1[3 * i + 1], pos1[3 * i + 2]};
  const double ni[3] = {nvec1[3 * i], nvec1[3 * i + 1], nvec1[3 * i + 2]};
  const double tiny = 2.220446049250313e-16;
  double fact[ORD + 1];
  fact[0] = 1.0;
#pragma unroll
  for (int n = 1; n <= ORD; ++n) fact[n] = fact[n - 1] * n;
  double gs[ORD + 1], fx[ORD + 1], fy[ORD + 1], fz[ORD + 1];
#pragma unroll
  for (int n = 0; n <= ORD; ++n) gs[n] = fx[n] = fy[n] = fz[n] = 0.0;
  double fq = 0.0;

  if (kind == 2) {
    // self term: polar rule around the centroid; tangential parts use
    // rr = max(r, 1e-4 max r)  (greenret_refined.py:208-258)
    const double t1[3] = {tvec1[3 * i], tvec1[3 * i + 1], tvec1[3 * i + 2]};
    const double t2[3] = {tvec2[3 * i], tvec2[3 * i + 1], tvec2[3 * i + 2]};
    double rmax = 0.0;
    for (int q = 0; q < nqs; ++q) {
      const QPoint Q = self_point(q);
      const double vx = pi_[0] - Q.p[0], vy = pi_[1] - Q.p[1], vz = pi_[2] - Q.p[2];
      rmax = fmax(rmax, fmax(sqrt(vx * vx + vy * vy + vz * vz), tiny));
    }
    double f1[ORD + 1], f2[ORD + 1], f3[ORD + 1];
#pragma unroll
    for (int n = 0; n <= ORD; ++n) f1[n] = f2[n] = f3[n] = 0.0;
    for (int q = 0; q < nqs; ++q) {
      const QPoint Q = self_point(q);
      const double vx = pi_[0] - Q.p[0], vy = pi_[1] - Q.p[1], vz = pi_[2] - Q.p[2];
      const double r = fmax(sqrt(vx * vx + vy * vy + vz * vz), tiny);
      const double rr = fmax(r, 1e-4 * rmax);
      const double ndr = ni[0] * vx + ni[1] * vy + ni[2] * vz;
      const double in1 = t1[0] * vx + t1[1] * vy + t1[2] * vz;
      const double in2 = t2[0] * vx + t2[1] * vy + t2[2] * vz;
      const double w = Q.w;
      const double ir = 1.0 / r, irr3 = 1.0 / (rr * rr * rr);
      double rnm1 = ir;               // r^(n-1)
      double rnm3 = ir * ir * ir;     // r^(n-3)
      double rn = 1.0;                // r^n
#pragma unroll
      for (int n = 0; n <= ORD; ++n) {
        gs[n] += w * rnm1;
        f1[n] += w * (n - 1) * ndr * rnm3;
        f2[n] += w * (n - 1) * in1 * rn * irr3;
        f3[n] += w * (n - 1) * in2 * rn * irr3;
        rnm1 *= r;
        rnm3 *= r;
        rn *= r;
      }
      fq -= w * ndr * ir * ir * ir;
    }
#pragma unroll
    for (int n = 0; n <= ORD; ++n) {
      gs[n] /= fact[n];
      const double a = f1[n] / fact[n], b = f2[n] / fact[n], c = f3[n] / fact[n];
      fx[n] = ni[0] * a + t1[0] * b + t2[0] * c;
      fy[n] = ni[1] * a + t1[1] * b + t2[1] * c;
      fz[n] = ni[2] * a + t1[2] * b + t2[2] * c;
    }
  } else {
    // near pair: 28-point rule on the source face, expanded about the
    // centroid distance r0 (greenret_refined.py:357-449)
    const double dx0 = pos2[3 * j] - pi_[0], dy0 = pos2[3 * j + 1] - pi_[1], dz0 = pos2[3 * j + 2] - pi_[2];
    const double r0 = sqrt(dx0 * dx0 + dy0 * dy0 + dz0 * dz0);
    for (int qq = 0; qq < (quad ? 2 : 1) * nq_near; ++qq) {
      const int q = qq % nq_near;
      if (quad && qq == nq_near) sub_triangle(X9, nv, 1, X);   // second triangle of the quadrilateral
      const QPoint Q = face_point(X, qn + q * QSTRIDE);
      const double vx = pi_[0] - Q.p[0], vy = pi_[1] - Q.p[1], vz = pi_[2] - Q.p[2];
      const double r = fmax(sqrt(vx * vx + vy * vy + vz * vz), tiny);
      const double dl = r - r0;
      const double w = Q.w;
      const double ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
      double dn = 1.0, dnm1 = 0.0;  // dl^n, dl^(n-1)
#pragma unroll
      for (int n = 0; n <= ORD; ++n) {
        gs[n] += w * dn * ir / fact[n];
        const double fnm1 = fact[n > 0 ? n - 1 : 0];
        const double fsc = (n == 0) ? -ir3 : (-dn * ir3 / fact[n] + dnm1 * ir2 / fnm1);
        fx[n] += w * vx * fsc;
        fy[n] += w * vy * fsc;
        fz[n] += w * vz * fsc;
        dnm1 = dn;
        dn *= dl;
      }
      fq -= w * (ni[0] * vx + ni[1] * vy + ni[2] * vz) * ir3;
    }
  }
#pragma unroll
  for (int n = 0; n <= ORD; ++n) {
    g[(long)p * (ORD + 1) + n] = gs[n];
    f[((long)p * 3 + 0) * (ORD + 1) + n] = fx[n];
    f[((long)p * 3 + 1) * (ORD + 1) + n] = fy[n];
    f[((long)p * 3 + 2) * (ORD + 1) + n] = fz[n];
  }
  fstat[p] = fq;
}

// one wavefront per column j: f_j = sum_i a_i F_ij / a_j over the rows of
// j's closed surface, F the quasistatic normal derivative with the refined
// entries (fstat) in place of -n_i.r_ij/d^3 a_j; corr_j = -2 pi dir - f_j is
// added to the order-0 Cartesian coefficient of the self pair as corr_j n_j.
__global__ __launch_bounds__(256) void k_closed_correction(
    int n, const double* __restrict__ pos, const double* __restrict__ nvec, const double* __restrict__ area,
    const int* __restrict__ lo, const int* __restrict__ hi, const int* __restrict__ colptr,
    const int* __restrict__ crow, const int* __restrict__ cpair, const double* __restrict__ fstat, double dir,
    int ord1, double* __restrict__ f, double* __restrict__ corr) {
  const int lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  const double pj[3] = {pos[3 * j], pos[3 * j + 1], pos[3 * j + 2]};
  const double aj = area[j];
  const int c0 = colptr[j], c1 = colptr[j + 1];
  double s = 0.0;
  for (int i = lo[j] + lane; i < hi[j]; i += 64) {
    // binary search i among the refined rows of column j (sorted)
    int a = c0, b = c1;
    while (a < b) {
      const int m = (a + b) >> 1;
      if (crow[m] < i)
        a = m + 1;
      else
        b = m;
    }
    double Fij;
    if (a < c1 && crow[a] == i) {
      Fij = fstat[cpair[a]];
    } else {
      const double x = pos[3 * i] - pj[0], y = pos[3 * i + 1] - pj[1], z = pos[3 * i + 2] - pj[2];
      const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
      Fij = -(nvec[3 * i] * x + nvec[3 * i + 1] * y + nvec[3 * i + 2] * z) / (d * d * d) * aj;
    }
    s += area[i] * Fij / aj;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) {
    const double cj = -2.0 * M_PI * dir - s;
    corr[j] = cj;
    for (int q = c0; q < c1; ++q)
      if (crow[q] == j) {
        const long pr = cpair[q];
        f[(pr * 3 + 0) * ord1] += cj * nvec[3 * j];
        f[(pr * 3 + 1) * ord1] += cj * nvec[3 * j + 1];
        f[(pr * 3 + 2) * ord1] += cj * nvec[3 * j + 2];
      }
  }
}

template <typename TO>
__device__ __forceinline__ void store_c(TO* p, double re, double im) {
  typedef typename vec2_of<TO>::type V;
  V v;
  v.x = (TO)re;
  v.y = (TO)im;
  *reinterpret_cast<V*>(p) = v;
}

constexpr int AS_I = 256;  // rows (field faces) per workgroup
constexpr int AS_J = 8;    // columns (source faces) per workgroup

// dense far-field pass over all (i, j) of one side, batch = wavelength.
// Unconnected pairs (different materials on this side) are written as 0.
template <typename TO, bool DUAL>
__global__ __launch_bounds__(AS_I) void k_green_assemble(
    int n, const double* __restrict__ pos, const double* __restrict__ nvec, const double* __restrict__ area,
    const int* __restrict__ mat, const double* __restrict__ kmat, int nmat, TO* __restrict__ G, TO* __restrict__ H,
    long ld, long stride, float* __restrict__ G2, float* __restrict__ H2, int gdiv) {
  const long b = blockIdx.z;
  const int i = blockIdx.x * AS_I + threadIdx.x;
  const int j0 = blockIdx.y * AS_J;
  if (i >= n) return;
  const long gs = mnp_geo_slot(b, gdiv);
  pos += 3 * n * gs;
  nvec += 3 * n * gs;
  area += n * gs;
  mat += n * gs;
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const double nxi = nvec[3 * i], nyi = nvec[3 * i + 1], nzi = nvec[3 * i + 2];
  const int mi = mat[i];
  const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
  G += 2 * b * stride;
  H += 2 * b * stride;
  if (DUAL) {
    G2 += 2 * b * stride;
    H2 += 2 * b * stride;
  }
  const int jn = min(AS_J, n - j0);
  for (int jj = 0; jj < jn; ++jj) {
    const int j = j0 + jj;
    double gr = 0, gi = 0, fr = 0, fi = 0;
    if (mat[j] == mi) {
      const double x = xi - pos[3 * j], y = yi - pos[3 * j + 1], z = zi - pos[3 * j + 2];
      const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
      const double id = 1.0 / d, aj = area[j];
      // phase e^{ikd} = e^{-ki d} (cos kr d + i sin kr d)
      double s, c;
      sincos(kr * d, &s, &c);
      const double e = exp(-ki * d);
      const double phr = e * c, phi = e * s;
      const double ga = aj * id;
      gr = ga * phr;
      gi = ga * phi;
      // (ik - 1/d)/d^2 * n.r * a_j
      const double ndr = nxi * x + nyi * y + nzi * z;
      const double sc = ndr * id * id * aj;
      const double ar = (-ki - id) * sc, ai = kr * sc;
      fr = ar * phr - ai * phi;
      fi = ar * phi + ai * phr;
    }
    const long o = 2 * ((long)j * ld + i);
    store_c(G + o, gr, gi);
    store_c(H + o, fr, fi);
    if (DUAL) {
      store_c(G2 + o, gr, gi);
      store_c(H2 + o, fr, fi);
    }
  }
}

// refined entries: value = (sum_n c_n (ik)^n) e^{ikd}; H gets +/-2 pi on the
// self diagonal after the phase (greenret_refined.py:698-810).  With several
// geometry slots (gdiv > 0) the pair lists of the slots are concatenated:
// slot s owns pairs [poff[s], poff[s+1]) with face indices local to the slot,
// and npairs is the longest list (the grid's extent).
template <typename TO, bool DUAL>
__global__ __launch_bounds__(256) void k_green_overlay(
    int n, int npairs, int ord1, const int* __restrict__ prow, const int* __restrict__ pcol,
    const double* __restrict__ g, const double* __restrict__ f, const double* __restrict__ pos,
    const double* __restrict__ nvec, const int* __restrict__ mat, const double* __restrict__ kmat, int nmat,
    double hsign, TO* __restrict__ G, TO* __restrict__ H, long ld, long stride, float* __restrict__ G2,
    float* __restrict__ H2, int gdiv, const long* __restrict__ poff) {
  const long b = blockIdx.y;
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  if (gdiv > 0) {
    const long gs = mnp_geo_slot(b, gdiv);
    const long q = poff[gs] + p;
    if (q >= poff[gs + 1]) return;
    p = (int)q;
    pos += 3 * n * gs;
    nvec += 3 * n * gs;
    mat += n * gs;
  }
  const int i = prow[p], j = pcol[p];
  const int mi = mat[i];
  if (mat[j] != mi) return;
  const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
  const double x = pos[3 * i] - pos[3 * j], y = pos[3 * i + 1] - pos[3 * j + 1], z = pos[3 * i + 2] - pos[3 * j + 2];
  const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
  const double nx = nvec[3 * i], ny = nvec[3 * i + 1], nz = nvec[3 * i + 2];
  // (ik)^n with ik = -ki + i kr
  double pr = 1.0, pim = 0.0;
  double sgr = 0, sgi = 0, sfr = 0, sfi = 0;
  for (int q = 0; q < ord1; ++q) {
    const double gq = g[(long)p * ord1 + q];
    const double fq = nx * f[((long)p * 3 + 0) * ord1 + q] + ny * f[((long)p * 3 + 1) * ord1 + q] +
                      nz * f[((long)p * 3 + 2) * ord1 + q];
    sgr += gq * pr;
    sgi += gq * pim;
    sfr += fq * pr;
    sfi += fq * pim;
    const double tr = pr * (-ki) - pim * kr;
    pim = pr * kr + pim * (-ki);
    pr = tr;
  }
  double s, c;
  sincos(kr * d, &s, &c);
  const double e = exp(-ki * d);
  const double phr = e * c, phi = e * s;
  const double gr = sgr * phr - sgi * phi, gi = sgr * phi + sgi * phr;
  double hr = sfr * phr - sfi * phi, hi = sfr * phi + sfi * phr;
  if (i == j) hr += hsign * 2.0 * M_PI;
  const long o = 2 * (b * stride + (long)j * ld + i);
  store_c(G + o, gr, gi);
  store_c(H + o, hr, hi);
  if (DUAL) {
    store_c(G2 + o, gr, gi);
    store_c(H2 + o, hr, hi);
  }
}

}  // namespace

// Refinement classification (refine_utils.py:29-158): pair (i, j) is refined
// with kind 2 where the centroid distance is zero and kind 1 where
// (d - r_j) / r_i < rel or d - r_j < abs.  The distance follows
// refine_utils.py:12-26: d^2 = (|a|^2 + |b|^2) - 2 a.b with squares below
// 1e-10 set to zero; |a|^2 is summed x, y, z in order and a.b accumulated with
// fused multiply-adds, as numpy's row sums and its BLAS product do, so pairs
// on the cutoff boundary fall on the same side.  One wave per row i scans all
// columns in chunks of 64; pass 0 counts (cnt[i]), pass 1 writes the row's
// pairs in column order at off[i] (row-major, columns ascending).
__device__ __forceinline__ double sq3(const double* p) {
  return __dadd_rn(__dadd_rn(__dmul_rn(p[0], p[0]), __dmul_rn(p[1], p[1])), __dmul_rn(p[2], p[2]));
}

__global__ __launch_bounds__(256) void k_refine_classify(int n, const double* __restrict__ pos,
                                                         const double* __restrict__ brad, double rel, double abs_,
                                                         int pass, int* cnt, const long* off, int* prow, int* pcol,
                                                         int* pkind) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;  // whole wave exits together
  const double* a = pos + 3 * (long)i;
  const double sa = sq3(a), ri = brad[i];
  long base = pass ? off[i] : 0;
  int count = 0;
  for (int j0 = 0; j0 < n; j0 += 64) {
    const int j = j0 + lane;
    int kind = 0;
    if (j < n) {
      const double* b = pos + 3 * (long)j;
      const double dot = __fma_rn(a[2], b[2], __fma_rn(a[1], b[1], __dmul_rn(a[0], b[0])));
      double dsq = __dsub_rn(__dadd_rn(sa, sq3(b)), __dmul_rn(2.0, dot));
      if (dsq < 1e-10) dsq = 0.0;
      const double d = sqrt(dsq);
      const double d2 = __dsub_rn(d, brad[j]);
      if (d == 0.0)
        kind = 2;
      else if (d2 < abs_ || d2 / ri < rel)
        kind = 1;
    }
    const unsigned long long m = __ballot(kind != 0);
    if (pass && kind) {
      const long at = base + count + __popcll(m & ((1ull << lane) - 1ull));
      prow[at] = i;
      pcol[at] = j;
      pkind[at] = kind;
    }
    count += __popcll(m);
  }
  if (!pass && lane == 0) cnt[i] = count;
}

extern "C" int mnpbem_refine_coeffs(int order, int npairs, const int* prow, const int* pcol, const int* pkind,
                                    const double* pos1, const double* nvec1, const double* tvec1,
                                    const double* tvec2, const double* pos2, const double* nodes9,
                                    const int* nvert2, const double* qnear, int nq_near, const double* qself,
                                    int nq_self, const double* qselfq, int nq_selfq, double* g, double* f,
                                    double* fstat, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(order >= 0 && order <= 7, "refine_coeffs: order must be in 0..7");
  MNP_REQUIRE(nq_near > 0 && nq_self > 0 && (nq_near + nq_self) * QSTRIDE * 8 <= 60000,
              "refine_coeffs: quadrature tables too large");
  MNP_REQUIRE(nq_selfq >= 0 && (nq_selfq == 0 || qselfq != nullptr), "refine_coeffs: bad quadrilateral table");
  if (npairs == 0) return 0;
  const size_t shm = (size_t)(nq_near + nq_self) * QSTRIDE * sizeof(double);
  dim3 grid(mnp_ceil_div(npairs, 256));
#define MNP_RC(O)                                                                                             \
  case O:                                                                                                     \
    hipLaunchKernelGGL(k_refine_coeffs<O>, grid, dim3(256), shm, st, npairs, prow, pcol, pkind, pos1, nvec1, \
                       tvec1, tvec2, pos2, nodes9, nvert2, qnear, nq_near, qself, nq_self, qselfq, nq_selfq, \
                       g, f, fstat);                                                                          \
    break;
  switch (order) {
    MNP_RC(0) MNP_RC(1) MNP_RC(2) MNP_RC(3) MNP_RC(4) MNP_RC(5) MNP_RC(6) MNP_RC(7)
  }
#undef MNP_RC
  MNP_LAUNCH_CHECK();
  return 0;
}

extern "C" int mnpbem_closed_correction(int n, const double* pos, const double* nvec, const double* area,
                                        const int* lo, const int* hi, const int* colptr, const int* crow,
                                        const int* cpair, const double* fstat, double dir, int order, double* f,
                                        double* corr, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_closed_correction, dim3(mnp_ceil_div(n, 4)), dim3(256), 0, st, n, pos, nvec, area, lo, hi,
                     colptr, crow, cpair, fstat, dir, order + 1, f, corr);
  MNP_LAUNCH_CHECK();
  return 0;
}

// G, H of one side for ``batch`` wavelengths.  kmat: (batch, nmat) complex
// wavenumbers; mat: per-face material index (0-based) on this side;
// hsign = +1 (H1, inside) or -1 (H2, outside).  dtype selects the outputs:
// C128, C64, or C128 plus a complex64 copy in G32/H32 (dual = 1).
extern "C" int mnpbem_green_assemble(int dtype, int dual, int n, const double* pos, const double* nvec,
                                     const double* area, const int* mat, const double* kmat, int nmat,
                                     int order, int npairs, const int* prow, const int* pcol, const double* g,
                                     const double* f, double hsign, void* G, void* H, long ld, long stride,
                                     void* G32, void* H32, int gdiv, const long* poff, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && ld >= n && batch >= 0 && nmat >= 1, "green_assemble: bad size");
  MNP_REQUIRE(gdiv <= 0 || poff != nullptr, "green_assemble: several geometry slots need the pair offsets poff");
  if (gdiv <= 0) poff = nullptr;
  MNP_REQUIRE(!(dual && dtype != MNP_DTYPE_C128), "green_assemble: dual output needs complex128 primary");
  if (n == 0 || batch == 0) return 0;
  dim3 g1(mnp_ceil_div(n, AS_I), mnp_ceil_div(n, AS_J), batch);
  dim3 g2(mnp_ceil_div(npairs, 256), batch);
  const int ord1 = order + 1;
  if (dtype == MNP_DTYPE_C128 && dual) {
    hipLaunchKernelGGL((k_green_assemble<double, true>), g1, dim3(AS_I), 0, st, n, pos, nvec, area, mat, kmat, nmat,
                       (double*)G, (double*)H, ld, stride, (float*)G32, (float*)H32, gdiv);
    MNP_LAUNCH_CHECK();
    if (npairs > 0)
      hipLaunchKernelGGL((k_green_overlay<double, true>), g2, dim3(256), 0, st, n, npairs, ord1, prow, pcol, g, f,
                         pos, nvec, mat, kmat, nmat, hsign, (double*)G, (double*)H, ld, stride, (float*)G32,
                         (float*)H32, gdiv, poff);
  } else if (dtype == MNP_DTYPE_C128) {
    hipLaunchKernelGGL((k_green_assemble<double, false>), g1, dim3(AS_I), 0, st, n, pos, nvec, area, mat, kmat,
                       nmat, (double*)G, (double*)H, ld, stride, (float*)nullptr, (float*)nullptr, gdiv);
    MNP_LAUNCH_CHECK();
    if (npairs > 0)
      hipLaunchKernelGGL((k_green_overlay<double, false>), g2, dim3(256), 0, st, n, npairs, ord1, prow, pcol, g, f,
                         pos, nvec, mat, kmat, nmat, hsign, (double*)G, (double*)H, ld, stride, (float*)nullptr,
                         (float*)nullptr, gdiv, poff);
  } else if (mnp_is_c64(dtype)) {
    hipLaunchKernelGGL((k_green_assemble<float, false>), g1, dim3(AS_I), 0, st, n, pos, nvec, area, mat, kmat,
                       nmat, (float*)G, (float*)H, ld, stride, (float*)nullptr, (float*)nullptr, gdiv);
    MNP_LAUNCH_CHECK();
    if (npairs > 0)
      hipLaunchKernelGGL((k_green_overlay<float, false>), g2, dim3(256), 0, st, n, npairs, ord1, prow, pcol, g, f,
                         pos, nvec, mat, kmat, nmat, hsign, (float*)G, (float*)H, ld, stride, (float*)nullptr,
                         (float*)nullptr, gdiv, poff);
  } else {
    mnp_set_error("green_assemble: unknown dtype");
    return 1;
  }
  MNP_LAUNCH_CHECK();
  return 0;
}

// Refined pairs (see k_refine_classify).  pass 0: cnt[i] = pairs of row i
// (prow/pcol/pkind unused); pass 1: pairs written at off[i] (exclusive prefix
// sums of cnt, int64) into prow, pcol, pkind (int32).
extern "C" int mnpbem_refine_classify(int n, const double* pos, const double* brad, double rel_cutoff,
                                      double abs_cutoff, int pass, int* cnt, const long* off, int* prow, int* pcol,
                                      int* pkind, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && (pass == 0 || pass == 1), "refine_classify: bad argument");
  MNP_REQUIRE(pass == 1 || cnt != nullptr, "refine_classify: pass 0 needs cnt");
  MNP_REQUIRE(pass == 0 || (off != nullptr && prow != nullptr && pcol != nullptr && pkind != nullptr),
              "refine_classify: pass 1 needs off, prow, pcol, pkind");
  hipLaunchKernelGGL(k_refine_classify, dim3(mnp_ceil_div(n, 4)), dim3(256), 0, st, n, pos, brad, rel_cutoff,
                     abs_cutoff, pass, cnt, off, prow, pcol, pkind);
  MNP_LAUNCH_CHECK();
  return 0;
}

// ---------------------------------------------------------------------------
// Fields at the faces from the surface charges and currents of one side
// (compgreen_ret.py:912-969): with Gp = grad G (deriv='cart'),
//   e_i = i k0 sum_j G_ij h_j - sum_j Hp_ij sig_j,   h_i = sum_j Hp_ij x h_j,
// Hp = Gp + hsign 2 pi n_i on the diagonal (H1p inside, H2p outside).
// G_ij = a_j e^{ikd}/d, Gp_ij = (ik - 1/d)/d^2 (r_i - r_j) a_j e^{ikd}, and
// for refined pairs sum_n (ik)^n {g_n, f_n} e^{ikd}; only faces whose side
// material matches the field face's contribute, with that material's k.
// One workgroup per (field face, wavelength); the refined pairs of row i are
// found by binary search in its (column-sorted) slice of the pair list.
// X: (batch, n, 8P) densities [sig1 | h1 | sig2 | h2]; E, H: (batch, n, 3P).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_surface_field(
    int n, int P, int side, double hsign, const double* __restrict__ pos, const double* __restrict__ nvec,
    const double* __restrict__ area, const int* __restrict__ mat, const double* __restrict__ kmat, int nmat,
    const double* __restrict__ k0v, const long* __restrict__ rowptr, const int* __restrict__ pcol,
    const double* __restrict__ g, const double* __restrict__ f, int ord1, const double* __restrict__ X, long ldx,
    long sx, double* __restrict__ E, double* __restrict__ Hf, long lde, long se) {
  const int i = blockIdx.x;
  const long b = blockIdx.y;
  const int mi = mat[i];
  const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
  const double k0 = k0v[b];
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const long p0 = rowptr[i], p1 = rowptr[i + 1];
  const int off = side * 4 * P;
  const double* xb = X + 2 * b * sx;
  cplx<double> ea[4][3], ha[4][3];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int c = 0; c < 3; ++c) ea[p][c] = ha[p][c] = cmake(0.0, 0.0);
  for (int j = threadIdx.x; j < n; j += 256) {
    if (mat[j] != mi) continue;
    const double x = xi - pos[3 * j], y = yi - pos[3 * j + 1], z = zi - pos[3 * j + 2];
    const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
    double s, c;
    sincos(kr * d, &s, &c);
    const double ex = exp(-ki * d);
    const cplx<double> ph = cmake(ex * c, ex * s);
    // refined pair?
    long lo = p0, hi = p1;
    while (lo < hi) {
      const long mid = (lo + hi) >> 1;
      if (pcol[mid] < j) lo = mid + 1; else hi = mid;
    }
    cplx<double> G, Gp[3];
    if (lo < p1 && pcol[lo] == j) {
      double pr = 1.0, pim = 0.0;  // (ik)^q, ik = -ki + i kr
      cplx<double> sg = cmake(0.0, 0.0), sf[3] = {cmake(0.0, 0.0), cmake(0.0, 0.0), cmake(0.0, 0.0)};
      for (int q = 0; q < ord1; ++q) {
        const double gq = g[lo * ord1 + q];
        sg = cadd(sg, cmake(gq * pr, gq * pim));
        for (int cc = 0; cc < 3; ++cc) {
          const double fq = f[(lo * 3 + cc) * ord1 + q];
          sf[cc] = cadd(sf[cc], cmake(fq * pr, fq * pim));
        }
        const double tr = pr * (-ki) - pim * kr;
        pim = pr * kr + pim * (-ki);
        pr = tr;
      }
      G = cmul(sg, ph);
      for (int cc = 0; cc < 3; ++cc) Gp[cc] = cmul(sf[cc], ph);
    } else {
      const double id = 1.0 / d, aj = area[j];
      G = cscale(ph, aj * id);
      // (ik - 1/d)/d^2 a_j e^{ikd}
      const cplx<double> fa = cmul(cmake((-ki - id) * id * id * aj, kr * id * id * aj), ph);
      Gp[0] = cscale(fa, x);
      Gp[1] = cscale(fa, y);
      Gp[2] = cscale(fa, z);
    }
    if (j == i) {
      for (int cc = 0; cc < 3; ++cc) Gp[cc].re += hsign * 2.0 * M_PI * nvec[3 * i + cc];
    }
    const double* xj = xb + 2 * ((long)j * ldx + off);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      if (p < P) {
        const cplx<double> sg = cld(xj + 2 * p);
        const cplx<double> h0 = cld(xj + 2 * (P + p)), h1 = cld(xj + 2 * (2 * P + p)), h2 = cld(xj + 2 * (3 * P + p));
        const cplx<double> hv[3] = {h0, h1, h2};
        for (int cc = 0; cc < 3; ++cc) {
          // i k0 G h_c - Gp_c sig
          const cplx<double> gh = cmul(G, hv[cc]);
          ea[p][cc] = cadd(ea[p][cc], csub(cmake(-k0 * gh.im, k0 * gh.re), cmul(Gp[cc], sg)));
        }
        ha[p][0] = cadd(ha[p][0], csub(cmul(Gp[1], h2), cmul(Gp[2], h1)));
        ha[p][1] = cadd(ha[p][1], csub(cmul(Gp[2], h0), cmul(Gp[0], h2)));
        ha[p][2] = cadd(ha[p][2], csub(cmul(Gp[0], h1), cmul(Gp[1], h0)));
      }
    }
  }
  __shared__ double red[4][4][6][2];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      cplx<double> v = q < 3 ? ea[p][q] : ha[p][q - 3];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        v.re += __shfl_xor(v.re, o);
        v.im += __shfl_xor(v.im, o);
      }
      if (lane == 0) {
        red[wv][p][q][0] = v.re;
        red[wv][p][q][1] = v.im;
      }
    }
  __syncthreads();
  if (threadIdx.x < 6 * P) {
    const int p = threadIdx.x % P, q = threadIdx.x / P;
    const double vr = red[0][p][q][0] + red[1][p][q][0] + red[2][p][q][0] + red[3][p][q][0];
    const double vi = red[0][p][q][1] + red[1][p][q][1] + red[2][p][q][1] + red[3][p][q][1];
    double* out = (q < 3 ? E : Hf) + 2 * (b * se + (long)i * lde + (q % 3) * P + p);
    out[0] = vr;
    out[1] = vi;
  }
}

// Surface fields e, h on side `side` (0 inside, 1 outside) of every face from
// the densities of that side (compgreen_ret.py:912-969; hsign = +1 for H1p,
// -1 for H2p).  rowptr: int64 (n + 1) offsets of each row's refined pairs in
// the (row, col)-sorted pair list (pcol, g, f as for mnpbem_green_assemble);
// k0: vacuum wavenumber per wavelength.  P <= 4.
extern "C" int mnpbem_surface_field(int n, int P, int side, double hsign, const double* pos, const double* nvec,
                                    const double* area, const int* mat, const double* kmat, int nmat,
                                    const double* k0, int order, const long* rowptr, const int* pcol,
                                    const double* g, const double* f, const void* X, long ldx, long sx, void* E,
                                    void* H, long lde, long se, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && P >= 1 && P <= 4 && (side == 0 || side == 1) && ldx >= 8 * P && lde >= 3 * P,
              "surface_field: bad size");
  if (batch <= 0) return 0;
  hipLaunchKernelGGL(k_surface_field, dim3(n, batch), dim3(256), 0, st, n, P, side, hsign, pos, nvec, area, mat,
                     kmat, nmat, k0, rowptr, pcol, g, f, order + 1, (const double*)X, ldx, sx, (double*)E,
                     (double*)H, lde, se);
  MNP_LAUNCH_CHECK();
  return 0;
}

// Gp (hsign 0), H1p = Gp + 2 pi n_i on the diagonal (hsign +1) or H2p (-1) of
// one side as a (batch, n, 3, n) array: out[((b n + i) 3 + c) n + j]
// (compgreen_ret.py:780-851, greenret_refined.py:812-864).  Pairs whose side
// materials differ are 0.  One thread per (i, j), threads along j.
__global__ __launch_bounds__(256) void k_green_deriv(int n, double hsign, const double* __restrict__ pos,
                                                     const double* __restrict__ nvec, const double* __restrict__ area,
                                                     const int* __restrict__ mat, const double* __restrict__ kmat,
                                                     int nmat, const long* __restrict__ rowptr,
                                                     const int* __restrict__ pcol, const double* __restrict__ g,
                                                     const double* __restrict__ f, int ord1,
                                                     double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  const long b = blockIdx.z;
  if (j >= n) return;
  const int mi = mat[i];
  cplx<double> v[3] = {cmake(0.0, 0.0), cmake(0.0, 0.0), cmake(0.0, 0.0)};
  if (mat[j] == mi) {
    const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
    const GreenEntry e = green_entry(i, j, kr, ki, pos, area, rowptr, pcol, g, f, ord1);
    for (int cc = 0; cc < 3; ++cc) v[cc] = e.Gp[cc];
    if (i == j)
      for (int cc = 0; cc < 3; ++cc) v[cc].re += hsign * 2.0 * M_PI * nvec[3 * i + cc];
  }
  double* o = out + 2 * (((b * n + i) * 3) * (long)n + j);
  for (int cc = 0; cc < 3; ++cc) store_c(o + 2 * (long)cc * n, v[cc].re, v[cc].im);
}

// Selected entries (i_e, j_e) of one side at one wavelength: kind 0 G; kind 1
// F + hsign 2 pi [i == j] (F, H1, H2); kind 2 Gp + hsign 2 pi n_i [i == j]
// (Gp, H1p, H2p), three values per entry (compgreen_ret.py:860-908, MATLAB
// Greenfun/@greenret/private/eval2.m).
__global__ __launch_bounds__(256) void k_green_select(long count, int kind, double hsign,
                                                      const int* __restrict__ rows, const int* __restrict__ cols,
                                                      const double* __restrict__ pos, const double* __restrict__ nvec,
                                                      const double* __restrict__ area, const int* __restrict__ mat,
                                                      const double* __restrict__ kmat, const long* __restrict__ rowptr,
                                                      const int* __restrict__ pcol, const double* __restrict__ g,
                                                      const double* __restrict__ f, int ord1,
                                                      double* __restrict__ out) {
  const long e = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (e >= count) return;
  const int i = rows[e], j = cols[e];
  const int mi = mat[i];
  cplx<double> v[3] = {cmake(0.0, 0.0), cmake(0.0, 0.0), cmake(0.0, 0.0)};
  if (mat[j] == mi) {
    const GreenEntry ge = green_entry(i, j, kmat[2 * mi], kmat[2 * mi + 1], pos, area, rowptr, pcol, g, f, ord1);
    if (kind == 0) {
      v[0] = ge.G;
    } else if (kind == 1) {
      v[0] = cadd(cadd(cscale(ge.Gp[0], nvec[3 * i]), cscale(ge.Gp[1], nvec[3 * i + 1])),
                  cscale(ge.Gp[2], nvec[3 * i + 2]));
      if (i == j) v[0].re += hsign * 2.0 * M_PI;
    } else {
      for (int cc = 0; cc < 3; ++cc) v[cc] = ge.Gp[cc];
      if (i == j)
        for (int cc = 0; cc < 3; ++cc) v[cc].re += hsign * 2.0 * M_PI * nvec[3 * i + cc];
    }
  }
  const int w = kind == 2 ? 3 : 1;
  for (int cc = 0; cc < w; ++cc) store_c(out + 2 * (e * w + cc), v[cc].re, v[cc].im);
}

// ---------------------------------------------------------------------------
// d/dk of the Green matrices of one side (for derivatives with respect to
// the material's permittivity, k = k0 sqrt(eps)): from G = a_j e^{ikd}/d and
// F = (n.r) a_j (ik - 1/d)/d^2 e^{ikd},
//   dG/dk = i a_j e^{ikd},   dF/dk = -k (n.r) a_j e^{ikd} / d,
// and for refined pairs (G, F) = S(k) e^{ikd} with S = sum_q c_q (ik)^q,
//   d/dk = (S'(k) + i d S(k)) e^{ikd},  S' = sum_q q i c_q (ik)^(q-1).
// The +-2 pi diagonal does not depend on k.  Layout as the assembly
// (transposed buffers, complex128).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(AS_I) void k_green_dk(int n, const double* __restrict__ pos,
                                                    const double* __restrict__ nvec, const double* __restrict__ area,
                                                    const int* __restrict__ mat, const double* __restrict__ kmat,
                                                    int nmat, double* __restrict__ dG, double* __restrict__ dH,
                                                    long ld, long stride) {
  const long b = blockIdx.z;
  const int i = blockIdx.x * AS_I + threadIdx.x;
  const int j0 = blockIdx.y * AS_J;
  if (i >= n) return;
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const double nxi = nvec[3 * i], nyi = nvec[3 * i + 1], nzi = nvec[3 * i + 2];
  const int mi = mat[i];
  const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
  dG += 2 * b * stride;
  dH += 2 * b * stride;
  const int jn = min(AS_J, n - j0);
  for (int jj = 0; jj < jn; ++jj) {
    const int j = j0 + jj;
    double gr = 0, gi = 0, fr = 0, fi = 0;
    if (mat[j] == mi && j != i) {
      const double x = xi - pos[3 * j], y = yi - pos[3 * j + 1], z = zi - pos[3 * j + 2];
      const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
      const double aj = area[j];
      double s, c;
      sincos(kr * d, &s, &c);
      const double e = exp(-ki * d);
      const double phr = e * c, phi = e * s;
      // i a e^{ikd}
      gr = -aj * phi;
      gi = aj * phr;
      // -k (n.r) a e^{ikd} / d
      const double sc = -(nxi * x + nyi * y + nzi * z) * aj / d;
      const double tr = sc * kr, ti = sc * ki;
      fr = tr * phr - ti * phi;
      fi = tr * phi + ti * phr;
    }
    const long o = 2 * ((long)j * ld + i);
    dG[o] = gr;
    dG[o + 1] = gi;
    dH[o] = fr;
    dH[o + 1] = fi;
  }
}

__global__ __launch_bounds__(256) void k_green_dk_overlay(int npairs, int ord1, const int* __restrict__ prow,
                                                           const int* __restrict__ pcol, const double* __restrict__ g,
                                                           const double* __restrict__ f,
                                                           const double* __restrict__ pos,
                                                           const double* __restrict__ nvec,
                                                           const int* __restrict__ mat,
                                                           const double* __restrict__ kmat, int nmat, double* dG,
                                                           double* dH, long ld, long stride) {
  const long b = blockIdx.y;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int i = prow[p], j = pcol[p];
  const int mi = mat[i];
  if (mat[j] != mi) return;
  const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
  const double x = pos[3 * i] - pos[3 * j], y = pos[3 * i + 1] - pos[3 * j + 1], z = pos[3 * i + 2] - pos[3 * j + 2];
  const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
  const double nx = nvec[3 * i], ny = nvec[3 * i + 1], nz = nvec[3 * i + 2];
  // S = sum c_q (ik)^q and S' = sum q i c_q (ik)^(q-1), ik = -ki + i kr
  cplx<double> ikp = cmake(1.0, 0.0), ikm = cmake(0.0, 0.0);  // (ik)^q, (ik)^(q-1)
  cplx<double> sg = cmake(0.0, 0.0), sf = cmake(0.0, 0.0), dsg = cmake(0.0, 0.0), dsf = cmake(0.0, 0.0);
  const cplx<double> ik = cmake(-ki, kr);
  for (int q = 0; q < ord1; ++q) {
    const double gq = g[(long)p * ord1 + q];
    const double fq = nx * f[((long)p * 3 + 0) * ord1 + q] + ny * f[((long)p * 3 + 1) * ord1 + q] +
                      nz * f[((long)p * 3 + 2) * ord1 + q];
    sg = cadd(sg, cscale(ikp, gq));
    sf = cadd(sf, cscale(ikp, fq));
    if (q > 0) {
      // q i (ik)^(q-1)
      const cplx<double> t = cmake(-q * ikm.im, q * ikm.re);
      dsg = cadd(dsg, cscale(t, gq));
      dsf = cadd(dsf, cscale(t, fq));
    }
    ikm = ikp;
    ikp = cmul(ikp, ik);
  }
  double s, c;
  sincos(kr * d, &s, &c);
  const double e = exp(-ki * d);
  const cplx<double> ph = cmake(e * c, e * s);
  const cplx<double> id = cmake(0.0, d);  // i d
  const cplx<double> vg = cmul(cadd(dsg, cmul(id, sg)), ph);
  const cplx<double> vf = cmul(cadd(dsf, cmul(id, sf)), ph);
  const long o = 2 * (b * stride + (long)j * ld + i);
  dG[o] = vg.re;
  dG[o + 1] = vg.im;
  dH[o] = vf.re;
  dH[o + 1] = vf.im;
}

// dG/dk, dH/dk of side `mat` (complex128, transposed buffers) for a batch of
// wavelengths; arguments as mnpbem_green_assemble.
extern "C" int mnpbem_green_dk(int n, const double* pos, const double* nvec, const double* area, const int* mat,
                               const double* kmat, int nmat, int order, int npairs, const int* prow, const int* pcol,
                               const double* g, const double* f, void* dG, void* dH, long ld, long stride,
                               int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n >= 0 && ld >= n && batch >= 0 && nmat >= 1, "green_dk: bad size");
  if (n == 0 || batch == 0) return 0;
  dim3 g1(mnp_ceil_div(n, AS_I), mnp_ceil_div(n, AS_J), batch);
  hipLaunchKernelGGL(k_green_dk, g1, dim3(AS_I), 0, st, n, pos, nvec, area, mat, kmat, nmat, (double*)dG,
                     (double*)dH, ld, stride);
  MNP_LAUNCH_CHECK();
  if (npairs > 0) {
    dim3 g2(mnp_ceil_div(npairs, 256), batch);
    hipLaunchKernelGGL(k_green_dk_overlay, g2, dim3(256), 0, st, npairs, order + 1, prow, pcol, g, f, pos, nvec,
                       mat, kmat, nmat, (double*)dG, (double*)dH, ld, stride);
    MNP_LAUNCH_CHECK();
  }
  return 0;
}


// ---------------------------------------------------------------------------
// Matrix-free Green products for the FP64 refinement residual of the mixed
// path: Y[:, yg + c] = (G X)[:, c], Y[:, yh + c] = (H X)[:, c] for the
// columns X[:, xoff + c], c < cols, of one side and every wavelength of the
// batch.  The entries are recomputed with the assembly's formulas (same
// operations in the same order as k_green_assemble / k_green_overlay) instead
// of being read from stored complex128 matrices: 16 bytes per entry and per
// matrix are traded for one sincos and one exp per pair.
//   k_green_apply      plain entries of every pair j != i: thread per row i,
//                      faces j staged through LDS in tiles of GA_T
//   k_green_apply_ref  the refined pairs of a row (one wave per row, fixed
//                      order wave reduction): refined - plain off the
//                      diagonal, refined (+ hsign 2 pi for H) on it
// ---------------------------------------------------------------------------
constexpr int GA_T = 64;

template <int C>
__global__ __launch_bounds__(256) void k_green_apply(int n, int cols, const double* __restrict__ pos,
                                                     const double* __restrict__ nvec, const double* __restrict__ area,
                                                     const int* __restrict__ mat, const double* __restrict__ kmat,
                                                     int nmat, const double* __restrict__ X, long ldx, long sx,
                                                     int xoff, double* __restrict__ Y, long ldy, long sy, int yg,
                                                     int yh) {
  const long b = blockIdx.y;
  const int i = blockIdx.x * 256 + threadIdx.x;
  __shared__ double sp[GA_T][4];  // x, y, z, area
  __shared__ int sm[GA_T];
  __shared__ cplx<double> sxv[GA_T][C];
  const bool act = i < n;
  const int ii = act ? i : 0;
  const double xi = pos[3 * ii], yi = pos[3 * ii + 1], zi = pos[3 * ii + 2];
  const double nxi = nvec[3 * ii], nyi = nvec[3 * ii + 1], nzi = nvec[3 * ii + 2];
  const int mi = mat[ii];
  const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
  const double* xb = X + 2 * (b * sx + xoff);
  cplx<double> ag[C], ah[C];
#pragma unroll
  for (int c = 0; c < C; ++c) ag[c] = ah[c] = cmake(0.0, 0.0);
  for (int j0 = 0; j0 < n; j0 += GA_T) {
    const int nj = min(GA_T, n - j0);
    __syncthreads();
    for (int e = threadIdx.x; e < nj; e += 256) {
      const int j = j0 + e;
      sp[e][0] = pos[3 * j];
      sp[e][1] = pos[3 * j + 1];
      sp[e][2] = pos[3 * j + 2];
      sp[e][3] = area[j];
      sm[e] = mat[j];
    }
    for (int e = threadIdx.x; e < nj * C; e += 256) {
      const int jj = e / C, c = e % C;
      sxv[jj][c] = c < cols ? cld(xb + 2 * ((long)(j0 + jj) * ldx + c)) : cmake(0.0, 0.0);
    }
    __syncthreads();
    if (!act) continue;
    for (int jj = 0; jj < nj; ++jj) {
      if (j0 + jj == i || sm[jj] != mi) continue;
      const double x = xi - sp[jj][0], y = yi - sp[jj][1], z = zi - sp[jj][2];
      const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
      const double id = 1.0 / d, aj = sp[jj][3];
      double s, c;
      sincos(kr * d, &s, &c);
      const double e = exp(-ki * d);
      const double phr = e * c, phi = e * s;
      const double ga = aj * id;
      const cplx<double> G = cmake(ga * phr, ga * phi);
      const double ndr = nxi * x + nyi * y + nzi * z;
      const double sc = ndr * id * id * aj;
      const double ar = (-ki - id) * sc, ai = kr * sc;
      const cplx<double> H = cmake(ar * phr - ai * phi, ar * phi + ai * phr);
#pragma unroll
      for (int q = 0; q < C; ++q) {
        ag[q] = cfma(ag[q], G, sxv[jj][q]);
        ah[q] = cfma(ah[q], H, sxv[jj][q]);
      }
    }
  }
  if (!act) return;
  double* yb = Y + 2 * (b * sy + (long)i * ldy);
  for (int c = 0; c < cols; ++c) {
    yb[2 * (yg + c)] = ag[c].re;
    yb[2 * (yg + c) + 1] = ag[c].im;
    yb[2 * (yh + c)] = ah[c].re;
    yb[2 * (yh + c) + 1] = ah[c].im;
  }
}

template <int C>
__global__ __launch_bounds__(256) void k_green_apply_ref(
    int n, int cols, double hsign, const double* __restrict__ pos, const double* __restrict__ nvec,
    const double* __restrict__ area, const int* __restrict__ mat, const double* __restrict__ kmat, int nmat,
    int ord1, const long* __restrict__ rowptr, const int* __restrict__ pcol, const double* __restrict__ g,
    const double* __restrict__ f, const double* __restrict__ X, long ldx, long sx, int xoff, double* __restrict__ Y,
    long ldy, long sy, int yg, int yh) {
  const long b = blockIdx.y;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;  // whole wave
  const double xi = pos[3 * i], yi = pos[3 * i + 1], zi = pos[3 * i + 2];
  const double nx = nvec[3 * i], ny = nvec[3 * i + 1], nz = nvec[3 * i + 2];
  const int mi = mat[i];
  const double kr = kmat[2 * (b * nmat + mi)], ki = kmat[2 * (b * nmat + mi) + 1];
  const double* xb = X + 2 * (b * sx + xoff);
  cplx<double> ag[C], ah[C];
#pragma unroll
  for (int c = 0; c < C; ++c) ag[c] = ah[c] = cmake(0.0, 0.0);
  for (long p = rowptr[i] + lane; p < rowptr[i + 1]; p += 64) {
    const int j = pcol[p];
    if (mat[j] != mi) continue;
    const double x = xi - pos[3 * j], y = yi - pos[3 * j + 1], z = zi - pos[3 * j + 2];
    const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
    double s, c;
    sincos(kr * d, &s, &c);
    const double e = exp(-ki * d);
    const double phr = e * c, phi = e * s;
    // refined value (k_green_overlay)
    double pr = 1.0, pim = 0.0;
    double sgr = 0, sgi = 0, sfr = 0, sfi = 0;
    for (int q = 0; q < ord1; ++q) {
      const double gq = g[p * ord1 + q];
      const double fq = nx * f[(p * 3 + 0) * ord1 + q] + ny * f[(p * 3 + 1) * ord1 + q] + nz * f[(p * 3 + 2) * ord1 + q];
      sgr += gq * pr;
      sgi += gq * pim;
      sfr += fq * pr;
      sfi += fq * pim;
      const double tr = pr * (-ki) - pim * kr;
      pim = pr * kr + pim * (-ki);
      pr = tr;
    }
    double gr = sgr * phr - sgi * phi, gi = sgr * phi + sgi * phr;
    double hr = sfr * phr - sfi * phi, hi = sfr * phi + sfi * phr;
    if (j == i) {
      hr += hsign * 2.0 * M_PI;
    } else {  // minus the plain entry k_green_apply added (k_green_assemble)
      const double id = 1.0 / d, aj = area[j];
      const double ga = aj * id;
      gr -= ga * phr;
      gi -= ga * phi;
      const double ndr = nx * x + ny * y + nz * z;
      const double sc = ndr * id * id * aj;
      const double ar = (-ki - id) * sc, ai = kr * sc;
      hr -= ar * phr - ai * phi;
      hi -= ar * phi + ai * phr;
    }
    const double* xj = xb + 2 * ((long)j * ldx);
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (q < cols) {
        const cplx<double> v = cld(xj + 2 * q);
        ag[q] = cfma(ag[q], cmake(gr, gi), v);
        ah[q] = cfma(ah[q], cmake(hr, hi), v);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < C; ++q) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ag[q].re += __shfl_xor(ag[q].re, o);
      ag[q].im += __shfl_xor(ag[q].im, o);
      ah[q].re += __shfl_xor(ah[q].re, o);
      ah[q].im += __shfl_xor(ah[q].im, o);
    }
  }
  if (lane == 0) {
    double* yb = Y + 2 * (b * sy + (long)i * ldy);
    for (int q = 0; q < cols; ++q) {
      yb[2 * (yg + q)] += ag[q].re;
      yb[2 * (yg + q) + 1] += ag[q].im;
      yb[2 * (yh + q)] += ah[q].re;
      yb[2 * (yh + q) + 1] += ah[q].im;
    }
  }
}

// Y[:, yg:yg+cols] = G X[:, xoff:xoff+cols], Y[:, yh:yh+cols] = H X[...] for
// the Green functions of the side whose per-face material indices are `mat`
// (hsign +1: H1 = F + 2 pi, -1: H2 = F - 2 pi), complex128, batch of
// wavelengths (kmat as for mnpbem_green_assemble; rowptr, pcol, g, f as for
// mnpbem_surface_field).  Same values as mnpbem_green_assemble followed by a
// product with the assembled matrices, without storing them.  cols <= 16.
extern "C" int mnpbem_green_apply(int n, int cols, double hsign, const double* pos, const double* nvec,
                                  const double* area, const int* mat, const double* kmat, int nmat, int order,
                                  const long* rowptr, const int* pcol, const double* g, const double* f,
                                  const void* X, long ldx, long sx, int xoff, void* Y, long ldy, long sy, int yg,
                                  int yh, int batch, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && cols >= 1 && cols <= 16 && xoff >= 0 && xoff + cols <= ldx && yg >= 0 && yh >= 0 &&
                  yg + cols <= ldy && yh + cols <= ldy,
              "green_apply: bad size");
  if (batch <= 0) return 0;
  dim3 g1(mnp_ceil_div(n, 256), batch), g2(mnp_ceil_div(n, 4), batch);
  const double* x = (const double*)X;
  double* y = (double*)Y;
#define MNP_GA(CC)                                                                                                  \
  {                                                                                                                 \
    hipLaunchKernelGGL(k_green_apply<CC>, g1, dim3(256), 0, st, n, cols, pos, nvec, area, mat, kmat, nmat, x, ldx,  \
                       sx, xoff, y, ldy, sy, yg, yh);                                                               \
    MNP_LAUNCH_CHECK();                                                                                             \
    hipLaunchKernelGGL(k_green_apply_ref<CC>, g2, dim3(256), 0, st, n, cols, hsign, pos, nvec, area, mat, kmat,     \
                       nmat, order + 1, rowptr, pcol, g, f, x, ldx, sx, xoff, y, ldy, sy, yg, yh);                  \
    MNP_LAUNCH_CHECK();                                                                                             \
    return 0;                                                                                                       \
  }
  if (cols <= 4) MNP_GA(4)
  if (cols <= 8) MNP_GA(8)
  MNP_GA(16)
#undef MNP_GA
}

// Gp / H1p / H2p of one side (hsign 0 / +1 / -1) for `batch` wavelengths as
// (batch, n, 3, n) complex128, the layout of CompGreenRet.eval(i, i, 'Gp' /
// 'H1p' / 'H2p') (compgreen_ret.py:780-851).  rowptr (int64, n + 1): offsets
// of each row's refined pairs in the (row, col)-sorted list pcol / f.
extern "C" int mnpbem_green_deriv(int n, double hsign, const double* pos, const double* nvec, const double* area,
                                  const int* mat, const double* kmat, int nmat, int order, const long* rowptr,
                                  const int* pcol, const double* g, const double* f, void* out, int batch,
                                  void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && n <= 65535 && nmat >= 1 && batch >= 0, "green_deriv: bad size (n <= 65535)");
  if (batch == 0) return 0;
  hipLaunchKernelGGL(k_green_deriv, dim3(mnp_ceil_div(n, 256), n, batch), dim3(256), 0, st, n, hsign, pos, nvec,
                     area, mat, kmat, nmat, rowptr, pcol, g, f, order + 1, (double*)out);
  MNP_LAUNCH_CHECK();
  return 0;
}

// Selected entries of one side at one wavelength (kmat: nmat complex
// wavenumbers): kind 0 G, 1 F / H (hsign), 2 Gp / Hp (three values per
// entry); rows, cols: count int32 face indices (< n, checked by the caller);
// out: count (kinds 0, 1) or count x 3 (kind 2) complex128.
extern "C" int mnpbem_green_select(int n, int kind, double hsign, const double* pos, const double* nvec,
                                   const double* area, const int* mat, const double* kmat, int nmat, int order,
                                   const long* rowptr, const int* pcol, const double* g, const double* f, long count,
                                   const int* rows, const int* cols, void* out, void* stream) {
  hipStream_t st = (hipStream_t)stream;
  MNP_REQUIRE(n > 0 && nmat >= 1 && count >= 0 && kind >= 0 && kind <= 2, "green_select: bad arguments");
  if (count == 0) return 0;
  hipLaunchKernelGGL(k_green_select, dim3(mnp_ceil_div(count, 256)), dim3(256), 0, st, count, kind, hsign, rows,
                     cols, pos, nvec, area, mat, kmat, rowptr, pcol, g, f, order + 1, (double*)out);
  MNP_LAUNCH_CHECK();
  return 0;
}
