// Single entries of the refined retarded Green functions of one side,
// shared by the selection / derivative kernels and the H-matrix kernels that
// generate their entries in place (no assembled matrix).
#pragma once

#include "common.h"

// ---------------------------------------------------------------------------
// Single Green entries of one side with refinement: G_ij and Gp_ij = grad_i G
// (deriv='cart'), e^{ikd} included; refined pairs (binary search in row i's
// column-sorted slice of the pair list) take sum_n (ik)^n {g_n, f_n} before
// the phase (greenret_refined.py:698-844, MATLAB Greenfun/@greenret/private/
// eval2.m).  F_ij = n_i . Gp_ij.  The diagonal terms of H and Hp are added by
// the callers.
// ---------------------------------------------------------------------------
struct GreenEntry {
  cplx<double> G, Gp[3];
};

__device__ __forceinline__ GreenEntry green_entry(int i, int j, double kr, double ki, const double* __restrict__ pos,
                                                  const double* __restrict__ area, const long* __restrict__ rowptr,
                                                  const int* __restrict__ pcol, const double* __restrict__ g,
                                                  const double* __restrict__ f, int ord1) {
  GreenEntry e;
  const double x = pos[3 * i] - pos[3 * j], y = pos[3 * i + 1] - pos[3 * j + 1], z = pos[3 * i + 2] - pos[3 * j + 2];
  const double d = fmax(sqrt(x * x + y * y + z * z), 2.220446049250313e-16);
  double s, c;
  sincos(kr * d, &s, &c);
  const double ex = exp(-ki * d);
  const cplx<double> ph = cmake(ex * c, ex * s);
  const long p1 = rowptr[i + 1];
  long lo = rowptr[i], hi = p1;
  while (lo < hi) {
    const long mid = (lo + hi) >> 1;
    if (pcol[mid] < j) lo = mid + 1; else hi = mid;
  }
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
    e.G = cmul(sg, ph);
    for (int cc = 0; cc < 3; ++cc) e.Gp[cc] = cmul(sf[cc], ph);
  } else {
    const double id = 1.0 / d, aj = area[j];
    e.G = cscale(ph, aj * id);
    const cplx<double> fa = cmul(cmake((-ki - id) * id * id * aj, kr * id * id * aj), ph);
    e.Gp[0] = cscale(fa, x);
    e.Gp[1] = cscale(fa, y);
    e.Gp[2] = cscale(fa, z);
  }
  return e;
}

// One scalar entry of side matrix G (kind 0) or F + hsign 2 pi [i == j]
// (kind 1: F, H1, H2) between field face i and source face j; 0 when the two
// faces hold different materials on this side.  kmat: the side's nmat
// complex wavenumbers of one wavelength (compgreen_ret.py:507-553).
__device__ __forceinline__ cplx<double> green_value(int kind, double hsign, int i, int j,
                                                    const double* __restrict__ pos, const double* __restrict__ nvec,
                                                    const double* __restrict__ area, const int* __restrict__ mat,
                                                    const double* __restrict__ kmat, const long* __restrict__ rowptr,
                                                    const int* __restrict__ pcol, const double* __restrict__ g,
                                                    const double* __restrict__ f, int ord1) {
  const int mi = mat[i];
  if (mat[j] != mi) return cmake(0.0, 0.0);
  const GreenEntry e = green_entry(i, j, kmat[2 * mi], kmat[2 * mi + 1], pos, area, rowptr, pcol, g, f, ord1);
  if (kind == 0) return e.G;
  cplx<double> v = cadd(cadd(cscale(e.Gp[0], nvec[3 * i]), cscale(e.Gp[1], nvec[3 * i + 1])),
                        cscale(e.Gp[2], nvec[3 * i + 2]));
  if (i == j) v.re += hsign * 2.0 * M_PI;
  return v;
}
