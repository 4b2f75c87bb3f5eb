/* C ABI of libmnpbem_amd.so: MNPBEM's retarded BEM hot path on MI355X (gfx950).
 *
 * The reference (rberwickt/MNPBEM) has no native ABI for this path: its
 * surface is the Python API (ComParticle, CompGreenRet.eval, BEMRet.init /
 * solve, PlaneWaveRet.potential / extinction).  These entry points are the
 * device-side operations that API is built from; each comment names the
 * reference code it replaces (paths relative to the reference root).
 *
 * Conventions
 *   - all pointers are device pointers unless stated; sizes are counts of
 *     elements; complex arrays are interleaved (re, im) like numpy/torch
 *     complex64 / complex128;
 *   - dtype: 1 = complex64, 2 = complex128, 3 = complex64 data whose GEMMs run
 *     on bf16 matrix cores with 3-term split products (~1e-5 relative per
 *     product; used under FP64 iterative refinement);
 *   - matrices are row-major with leading dimension ld (complex elements);
 *     "batched" calls process `batch` matrices spaced `stride` elements
 *     apart; pivot arrays are int32, 0-based, spaced `spiv` apart;
 *   - every call is asynchronous on `stream` (a hipStream_t, NULL = default)
 *     and returns 0 on success, 1 for invalid arguments, 2 for a HIP error;
 *     mnpbem_last_error() returns the message of the last failure on the
 *     calling thread;
 *   - nothing here allocates device memory except mnpbem_prof_begin (HIP
 *     events); calls on one stream are ordered, calls that share output
 *     buffers on different streams race;
 *   - geometry slots (gdiv): calls with a gdiv argument can batch several
 *     particles of one face count n.  Their per-face inputs are then stacked
 *     per particle (pos, nvec: slots x n x 3; area, mat, in_med, out_med:
 *     slots x n) and the batch is ordered particle-major with gdiv wavelengths
 *     per particle, so element b reads slot b / gdiv.  gdiv = 0: one geometry
 *     for the whole batch.  (No counterpart in the reference, which loops over
 *     particles and wavelengths in Python.)
 */
#ifndef MNPBEM_AMD_H
#define MNPBEM_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define MNPBEM_C64 1
#define MNPBEM_C128 2
#define MNPBEM_C64_BF16X3 3

int mnpbem_abi_version(void);
const char* mnpbem_last_error(void);
int mnpbem_device_sync(void);

/* ---- dense complex linear algebra (replaces scipy.linalg.lu_factor /
 *      lu_solve, i.e. LAPACK zgetrf / zgetrs, as dispatched by
 *      mnpbem/utils/gpu.py:145-279 from mnpbem/bem/bem_ret.py:437-523) ---- */

/* C[b] -= A[b] B[b]  (M x K times K x N), MFMA complex GEMM. */
int mnpbem_cgemm_sub_batched(int dtype, int M, int N, int K, const void* A, long lda, long sA, const void* B,
                             long ldb, long sB, void* C, long ldc, long sC, int batch, void* stream);

/* Tile selection of the complex64 bf16x3 GEMM (no counterpart in the
 * reference, whose LU runs in LAPACK / cuSOLVER; this is the tile autotuner
 * of the mixed-precision path).  A 7 x 7 x 7 table over shape classes of
 * (M, N, K) (class c: x <= 64 * 2^c, class 6 unbounded; cell (cm*7+cn)*7+ck)
 * picks variant 0 (128 x 128 tile), 1 (64 x 64 tile) or 2 (128 x 128 tile that
 * reads its C tile before the K loop) per launch; all give identical
 * results.  _variant forces a variant (complex64 data). */
int mnpbem_cgemm_sub_variant(int variant, int M, int N, int K, const void* A, long lda, long sA, const void* B,
                             long ldb, long sB, void* C, long ldc, long sC, int batch, void* stream);
int mnpbem_gemm_tuning_set(const unsigned char* table, int ncells);
int mnpbem_gemm_tuning_get(unsigned char* table, int ncells);
/* shapes[4c..4c+3] = M, N, K, batch of one launch per class since the last
 * clear (M = 0: none); clear != 0 resets the record. */
int mnpbem_gemm_tuning_seen(int* shapes, int ncells, int clear);

/* In-place LU with partial pivoting, P A = L U (largest |re|+|im| pivot,
 * LAPACK convention); info[b] = 1-based column of the first exactly-zero
 * pivot or 0.  Replaces lu_factor_dispatch (mnpbem/utils/gpu.py:145-191).
 * Matrices above the tall-panel threshold (mnpbem_getrf_tall_rows) keep their
 * pivot candidates in one device-wide buffer of this library: two such
 * factorisations must not run concurrently on one device (different streams
 * or overlapping graph replays); smaller matrices have no such restriction. */
int mnpbem_getrf_batched(int dtype, int n, void* A, long lda, long sA, int* ipiv, long spiv, int* info, int batch,
                         void* stream);

/* The same factorisation without row interchanges (ipiv = identity), for
 * matrices whose partial pivoting makes none; flag[b] = 1 when it would have
 * made one (some multiplier has |re|+|im| > 1), and then the factors differ
 * from mnpbem_getrf_batched's, 2 when a multiplier exceeds 1024 (growth guard);
 * with flag[b] = 0 they are the partial-pivoting factors.  info as above.  Panels of 64 (complex64) or 32
 * (complex128) columns: the diagonal block factored in LDS, the multipliers
 * below as the product with its inverse.  No counterpart in the reference
 * (an alternative path to lu_factor_dispatch, mnpbem/utils/gpu.py:145-191,
 * for the Green matrices).  The inverse diagonal blocks live in one
 * device-wide buffer of this library: two such factorisations must not run
 * concurrently on one device. */
int mnpbem_getrf_nopiv_batched(int dtype, int n, void* A, long lda, long sA, int* ipiv, long spiv, int* info,
                               int* flag, int batch, void* stream);

/* B <- A^-1 B for A factored by mnpbem_getrf_batched, nrhs columns.
 * Replaces lu_solve_dispatch / lu_solve_native (mnpbem/utils/gpu.py:194-279). */
int mnpbem_getrs_batched(int dtype, int n, int nrhs, const void* A, long lda, long sA, const int* ipiv, long spiv,
                         void* B, long ldb, long sB, int batch, void* stream);

/* A <- A^-1 from its mnpbem_getrf_batched factors (LAPACK getri): A^-1 =
 * U^-1 L^-1 P with L^-1 formed by a recursion that skips the zero triangle of
 * the identity; W (n x n per matrix) and Q (n ints per matrix, n <= 32768) are
 * work space.  Replaces the explicit inverse that bem_ret.py:437-523 forms by
 * a solve against the identity. */
int mnpbem_getri_batched(int dtype, int n, void* A, long lda, long sA, const int* ipiv, long spiv, void* W, long ldw,
                         long sW, int* Q, long sQ, int batch, void* stream);

/* mnpbem_getrs_batched and mnpbem_getri_batched given the inverse diagonal
 * blocks of the factors, bs = 64 rows (Dinv from mnpbem_lu_blockinv_batched)
 * or 128 rows (from mnpbem_lu_blockinv128_batched), per-matrix stride sD.  For
 * MNPBEM_DTYPE_C64X the bs-row leaves of the triangular solves become products
 * with those inverses on the bf16x3 kernel instead of substitutions; every
 * other dtype, or Dinv = NULL, runs the plain entry point.  Same contract as
 * those (bem_ret.py:437-523). */
int mnpbem_getrs_blockinv_batched(int dtype, int n, int nrhs, const void* A, long lda, long sA, const int* ipiv,
                                  long spiv, const void* Dinv, long sD, int bs, void* B, long ldb, long sB, int batch,
                                  void* stream);
int mnpbem_getri_blockinv_batched(int dtype, int n, void* A, long lda, long sA, const int* ipiv, long spiv,
                                  const void* Dinv, long sD, int bs, void* W, long ldw, long sW, int* Q, long sQ,
                                  int batch, void* stream);

/* W <- A^-1 for A factored without row interchanges (mnpbem_getrf_nopiv_batched
 * with no interchange recorded, so P = I): the inverse U^-1 L^-1 is built in
 * W and left there (A is only read; no permutation work, no copy back into
 * A).  Dinv / bs as for mnpbem_getri_blockinv_batched (NULL: substitution
 * leaves).  Replaces, like mnpbem_getri_batched, the explicit inverse of
 * bem_ret.py:437-523. */
int mnpbem_getri_nopiv_blockinv_batched(int dtype, int n, const void* A, long lda, long sA, const void* Dinv, long sD,
                                        int bs, void* W, long ldw, long sW, int batch, void* stream);

/* Inverses of the 128 x 128 diagonal blocks of complex64 LU factors (as
 * [U^-1 | L^-1] per block, two 128 x 128 row-major blocks) from the 64 x 64
 * ones of mnpbem_lu_blockinv_batched (D64) and the factors' off-diagonal
 * 64 x 64 blocks; for the 128-row leaves of the blockinv solves above.
 * Replaces nothing in the reference (its inverses are whole matrices,
 * bem_ret.py:437-523). */
int mnpbem_lu_blockinv128_batched(int dtype, int n, const void* LU, long ld, long sLU, const void* D64, long sD64,
                                  void* D128, long sD128, int batch, void* stream);

/* Y[b] = beta Y[b] + alpha M[b]^T X[b] for n x C blocks X, Y (any C). */
int mnpbem_gemv_t_batched(int dtype, int n, int C, const void* M, long ldm, long sM, const void* X, long ldx,
                          long sX, void* Y, long ldy, long sY, double alpha_re, double alpha_im, double beta,
                          int batch, void* stream);

/* Solve M x = v in place for the n x C block V given the LU of M^T
 * (the transposed-storage convention of BEMRet's factors); Z (n x C,
 * leading dimension ldz >= C) is work space. */
int mnpbem_lu_solve_t_batched(int dtype, int n, int C, const void* LU, long ld, long sLU, const int* ipiv,
                              long spiv, void* V, long ldv, long sV, void* Z, long ldz, long sZ, const void* Dinv,
                              long sD, int batch, void* stream);

/* Rows above which mnpbem_getrf_batched factors its panels with the
 * multi-workgroup column kernels (k_tall_pivot / k_tall_update) instead of
 * the register-resident leaf; rows = 0 restores the built-in thresholds (12288
 * complex64, 6144 complex128).  Returns the previous setting.  A tuning knob
 * of this library (the reference's LAPACK/cuSOLVER getrf has none). */
int mnpbem_getrf_tall_rows(int rows);

/* Columns (right-hand sides x batch) up to which the leaves of the triangular
 * solves (the diagonal blocks of 17..64 rows complex64, 17..32 complex128
 * inside getrf, getrs and getri) run a row-blocked kernel (four waves over
 * four row blocks of 64 columns) instead of one thread per column; cols < 0
 * restores the built-in 32768, 0 disables it.  Returns the previous setting.
 * A tuning knob of this library (no reference counterpart). */
int mnpbem_trsm_rows_max(int cols);

/* Right-hand sides from which the many-column triangular solves of
 * mnpbem_getrs_blockinv_batched and mnpbem_getri_blockinv_batched (complex64
 * on the bf16x3 products, bs = 128, n > 256) run block row by block row
 * (left-looking: one product over all solved rows, then the inverse-block
 * leaf) instead of recursively; rhs < 0 restores the built-in 512, 0 disables
 * it.  Returns the previous setting.  A tuning knob of this library (no
 * reference counterpart). */
int mnpbem_trsm_left_rhs(int rhs);

/* Widest complex64 bf16x3-path product (N columns, with K <= 64) that runs on
 * the FP32 vector kernel for thin products instead of the bf16x3 tiles: 0..32
 * (0: none), n < 0 restores the built-in 16.  Returns the previous setting.
 * A tuning knob of this library (no reference counterpart). */
int mnpbem_gemm_thin_max(int n);

/* 1 (built-in): complex64 matrix-vector products of at most 8 columns
 * (mnpbem_gemv_t_batched and the solves built on it, n even, 16-byte aligned
 * buffer) run on the FP32 matrix cores; 0: on the vector ALUs (two output
 * rows per lane).  Same FP32 arithmetic in another summation order.  Returns
 * the previous setting.  A tuning knob of this library (no reference
 * counterpart). */
int mnpbem_gemv_mfma(int on);

/* Tile updates of the windowed transposed triangular solves (complex64,
 * many systems, n even, 16-byte aligned factors) on the FP32 matrix cores:
 * bit 0 for 5-8 right-hand-side columns, bit 1 for at most 4; mask < 0
 * restores the built-in 3 (both).  Same FP32 arithmetic in another summation
 * order.  Returns the previous mask.  A tuning knob of this library (no
 * reference counterpart). */
int mnpbem_trsv_mfma(int mask);

/* 1 (built-in): in those block-row solves every block row after the first is
 * one launch, C <- D (C - A B), the inverse-block leaf folded into the update
 * product's epilogue (bit-identical to the two-launch form); 0: update and
 * leaf as two launches.  Returns the previous setting.  A tuning knob of this
 * library (no reference counterpart). */
int mnpbem_trsm_leaf_fused(int on);

/* Inverses of the 64 x 64 diagonal blocks of getrf factors ([U_II^-1 | L_II^-1]
 * per block, sD >= ceil(n/64)*2*64*64 elements) so that the block steps of
 * mnpbem_lu_solve_t_batched are parallel matrix-vector products (Dinv = NULL
 * there falls back to per-block substitution). */
int mnpbem_lu_blockinv_batched(int dtype, int n, const void* LU, long ld, long sLU, void* Dinv, long sD, int batch,
                               void* stream);

int mnpbem_identity_batched(int dtype, int n, void* A, long ld, long sA, int batch, void* stream);

/* ---- Green functions (replaces mnpbem/greenfun/greenret_refined.py:78-867,
 *      compgreen_ret.py:192-505 and compgreen_stat.py:219-465) ---- */

/* Refinement classification (replaces refinematrix, mnpbem/greenfun/
 * refine_utils.py:12-158): pair (i, j) gets kind 2 when the centroid distance
 * is zero, kind 1 when (d - r_j) / r_i < rel_cutoff or d - r_j < abs_cutoff,
 * d^2 = |a|^2 + |b|^2 - 2 a.b with squares < 1e-10 set to 0.  pos: n x 3,
 * brad: element radii.  pass 0 writes cnt[i] (pairs per row); pass 1 writes
 * the pairs of row i, columns ascending, from off[i] (int64 exclusive prefix
 * sums of cnt) into prow / pcol / pkind (int32). */
int mnpbem_refine_classify(int n, const double* pos, const double* brad, double rel_cutoff, double abs_cutoff,
                           int pass, int* cnt, const long* off, int* prow, int* pcol, int* pkind, void* stream);

/* Order-`order` (<= 7) expansion coefficients of refined face pairs
 * (deriv='cart'): for pair p = (row i, col j, kind 2 = self, 1 = near)
 *   g[p][n], f[p][c][n] (c = x,y,z) and the quasistatic normal derivative
 *   fstat[p] = -sum_q w_q n_i.(r_i - x_q)/|r_i - x_q|^3.
 * Field faces: pos1, nvec1, tvec1, tvec2 (n1 x 3 doubles); source faces:
 * pos2 (n2 x 3), nodes9 (n2 x 9 x 3: corners 0-3, edge midpoints 4-7, centre
 * 8 of a quadrilateral; corners 0-2 and midpoints 4-6 of a triangle) and
 * nvert2 (n2 ints, 3 or 4); qnear / qself: host-built tables of nq rows
 * [shape(6), d/dx(6), d/dy(6), weight] for the 28-point rule and the
 * triangle polar rule; qselfq: nq_selfq rows [shape(9), d/dx(9), d/dy(9),
 * weight] of the quadrilateral polar rule (may be NULL with 0 rows when no
 * face is a quadrilateral).  Quadrilaterals integrate near pairs as their two
 * 6-node triangles.  (greenret_refined.py:122-516, particle.py:1364-1575) */
int mnpbem_refine_coeffs(int order, int npairs, const int* prow, const int* pcol, const int* pkind,
                         const double* pos1, const double* nvec1, const double* tvec1, const double* tvec2,
                         const double* pos2, const double* nodes9, const int* nvert2, const double* qnear,
                         int nq_near, const double* qself, int nq_self, const double* qselfq, int nq_selfq,
                         double* g, double* f, double* fstat, void* stream);

/* Fuchs-Liu closed-surface correction: corr_j = -2 pi dir - sum_i a_i F_ij / a_j
 * over rows lo[j] <= i < hi[j] (j's closed surface), with the refined pairs of
 * column j listed (sorted by row) in crow[colptr[j]..colptr[j+1]) / cpair;
 * adds corr_j n_j to f[self pair of j][:][0].  (compgreen_ret.py:192-400) */
int mnpbem_closed_correction(int n, const double* pos, const double* nvec, const double* area, const int* lo,
                             const int* hi, const int* colptr, const int* crow, const int* cpair,
                             const double* fstat, double dir, int order, double* f, double* corr, void* stream);

/* G and H = F + hsign 2 pi I of one side (inside: mat = inside material,
 * hsign = +1; outside: -1) for `batch` wavelengths with wavenumbers
 * kmat[b][m] (complex, doubles), written transposed: entry (i, j) at
 * out[b*stride + j*ld + i].  dual = 1 (dtype complex128) also writes a
 * complex64 copy to G32/H32.  With geometry slots (gdiv > 0) the refined pairs
 * and coefficients of the slots are concatenated, slot s owning pairs
 * [poff[s], poff[s+1]) (int64, device) with face indices local to the slot,
 * and npairs is the longest slot's count.  Replaces CompGreenRet.eval(i, i,
 * 'G'/'H1'/'H2') (compgreen_ret.py:507-858, greenret_refined.py:654-810). */
int mnpbem_green_assemble(int dtype, int dual, int n, const double* pos, const double* nvec, const double* area,
                          const int* mat, const double* kmat, int nmat, int order, int npairs, const int* prow,
                          const int* pcol, const double* g, const double* f, double hsign, void* G, void* H,
                          long ld, long stride, void* G32, void* H32, int gdiv, const long* poff, int batch,
                          void* stream);

/* ---- BEM solver (replaces mnpbem/bem/bem_ret.py:293-525 and 1360-1831) ----
 * Vectors per wavelength are n x 8P row-major (P polarisations):
 *   unknowns X = [sig1 | h1 (3P) | sig2 | h2 (3P)],
 *   right side R = [phi | a (3P) | De | alpha (3P)]  (bem_ret_iter.py:1196-1232).
 * par[b] = {Re eps1, Im eps1, Re eps2, Im eps2, k0} per wavelength. */

/* mode 0: D = S1 - S2 (Delta^T); mode 1: S2 <- eps1 S1 - eps2 S2
 * + k0^2 (eps1 - eps2)^2 E o (n n^T)  (Sigma^T, bem_ret.py:476-495). */
int mnpbem_sigma_batched(int dtype, int mode, int n, const void* S1, void* S2, void* D, const void* E, long ld,
                         long stride, const double* nvec, const double* par, int gdiv, int batch, void* stream);

/* X = A^-1 R with the factors of BEMRet.init: LU(G1^T), LU(G2^T), Sigma1^T,
 * (Delta^-1)^T, LU(Sigma^T) (Eqs. 19-20 of Garcia de Abajo & Howie,
 * bem_ret.py:1694-1810); D1, D2, DS: their inverse diagonal blocks from
 * mnpbem_lu_blockinv_batched (or NULL).  W is work space of R's shape; P <= 4. */
int mnpbem_bem_solve_batched(int dtype, int n, int P, const void* G1t, const int* piv1, const void* G2t,
                             const int* piv2, const void* S1t, const void* Et, const void* St, const int* pivS,
                             long ldm, long smat, long spiv, const void* D1, const void* D2, const void* DS, long sD,
                             const void* R, void* W, void* X, long ldv, long sv, const double* nvec,
                             const double* par, int gdiv, int batch, void* stream);

/* mnpbem_bem_solve_batched with explicit inverses instead of the LU factors of
 * G1, G2 and Sigma: G1i, G2i, Si are the transposed buffers of G1^-1, G2^-1,
 * Sigma^-1 (getri on the factors of the transposed buffers), so every solve is
 * a matrix-vector product.  For a factorisation applied many times to a few
 * right-hand sides (the block preconditioner inside GMRES), where the
 * triangular steps of a few systems are latency-bound.  Same vectors, work
 * space and parameters as mnpbem_bem_solve_batched; replaces, like it, the
 * solve of bem_ret.py:1569-1831 (with the reference's own explicit G1^-1,
 * G2^-1 of bem_ret.py:437-523). */
int mnpbem_bem_solve_inv_batched(int dtype, int n, int P, const void* G1i, const void* G2i, const void* S1t,
                                 const void* Et, const void* Si, long ldm, long smat, const void* R, void* W, void* X,
                                 long ldv, long sv, const double* nvec, const double* par, int gdiv, int batch,
                                 void* stream);

/* mnpbem_bem_solve_batched without the work space argument W: no counterpart
 * in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); it allocates nothing.  W lives in a static 64 MiB device array
 * of this library (one per device, held until the process ends), so the
 * batch must satisfy ((batch-1) sv + (n-1) ldv + 8P) x element size <= 64 MiB
 * (complex64 c. 23.6 MB for 64 wavelengths of N = 2884, P = 2), else status 1.
 * Calls that overlap on one device (several streams, overlapping graph
 * replays) share that array and race. */
int mnpbem_bem_solve_batched_noalloc(int dtype, int n, int P, const void* G1t, const int* piv1, const void* G2t,
                                     const int* piv2, const void* S1t, const void* Et, const void* St,
                                     const int* pivS, long ldm, long smat, long spiv, const void* D1, const void* D2,
                                     const void* DS, long sD, const void* R, void* X, long ldv, long sv,
                                     const double* nvec, const double* par, int gdiv, int batch, void* stream);

/* Rout = R - A X for the complex128 8N operator (bem_ret_iter.py:1362-1501)
 * with unfactored transposed G1, H1, G2, H2; Y is n x 16P work space;
 * nrm[b] = {||R - A X||^2, ||R||^2} (NULL to skip; Rout must not alias R
 * then), summed in a fixed order so the values repeat bit for bit. */
int mnpbem_bem_residual_batched(int n, int P, const void* G1t, const void* H1t, const void* G2t, const void* H2t,
                                long ldm, long smat, const void* R, const void* X, void* Rout, long ldv, long sv,
                                void* Y, long ldy, long sy, const double* nvec, const double* par, double* nrm,
                                int gdiv, int batch, void* stream);

/* ---- adjoint solve for the extinction gradients (no counterpart in the
 *      reference, which has no derivatives; it transposes the solve of
 *      bem_ret.py:1694-1810) ----
 * Y = A^-T C with the factors of mnpbem_bem_solve_batched (same arguments;
 * C in the unknowns' layout [sig1 | h1 | sig2 | h2], Y in the right-side
 * layout [phi | a | De | alpha]); W is work space shaped like C. */
int mnpbem_bem_solve_adjoint_batched(int dtype, int n, int P, const void* G1t, const int* piv1, const void* G2t,
                                     const int* piv2, const void* S1t, const void* Et, const void* St,
                                     const int* pivS, long ldm, long smat, long spiv, const void* D1, const void* D2,
                                     const void* DS, long sD, const void* C, void* W, void* Y, long ldv, long sv,
                                     const double* nvec, const double* par, int batch, void* stream);

/* The same adjoint solve with explicit inverses: G1i, G2i, Si are the
 * buffers that mnpbem_getri_blockinv_batched leaves in place of the factors
 * of G1t, G2t, St (S1t, Et as above); each pair of triangular solves becomes
 * one matrix product.  C and Y must be different buffers.  No counterpart in
 * the reference (see above). */
int mnpbem_bem_solve_adjoint_inv_batched(int dtype, int n, int P, const void* G1i, const void* G2i,
                                         const void* S1t, const void* Et, const void* Si, long ldm, long smat,
                                         const void* C, void* W, void* Y, long ldv, long sv, const double* nvec,
                                         const double* par, int batch, void* stream);

/* Out = C - A^T Y (complex128; C NULL: Out = A^T Y) with unfactored
 * transposed G1, H1, G2, H2; Z (n x 8P) and T (n x 16P, ldt >= 16P) are work
 * space; nrm[b] = {||C - A^T Y||^2, ||C||^2} (may be NULL). */
int mnpbem_bem_residual_adjoint_batched(int n, int P, const void* G1t, const void* H1t, const void* G2t,
                                        const void* H2t, long ldm, long smat, const void* C, const void* Y, void* Out,
                                        long ldv, long sv, void* Z, void* T, long ldt, long stt, const double* nvec,
                                        const double* par, double* nrm, int batch, void* stream);

/* ---- geometry derivatives of the extinction (no counterpart in the
 *      reference; they differentiate greenret_refined.py:122-810,
 *      compgreen_ret.py:192-400, planewave_ret.py:257-504 and
 *      spectrum_ret.py:231-392 term by term) ----
 * Per-face gradient buffer Gr (batch, P, n, 31) float64, zeroed by the caller:
 * [pos 3 | nvec 3 | tvec1 3 | tvec2 3 | area 1 | nodes 18].  L (batch, n, 2, P,
 * 8), R (batch, n, 2, P, 4), Wc (batch, n, P) complex128 weights from
 * mnpbem_grad_prep.  P <= 4. */
/* C = kappa c, the optical-theorem functional of mnpbem_extinction_batched
 * written as a vector in the unknowns' layout (ext = Im(C . X)). */
int mnpbem_ext_functional(int n, int P, const double* pos, const double* area, const int* in_med,
                          const int* out_med, const double* pol, const double* dir, const double* pw, void* C,
                          long ldv, long sv, int batch, void* stream);
/* weights of the four Green matrices in -Y . A X from the solution X and the adjoint Y */
int mnpbem_grad_prep(int n, int P, const void* X, const void* Y, long ldv, long sv, const double* nvec,
                     const double* par, void* L, void* R, void* Wc, int batch, void* stream);
/* far-field pairs and the far terms of the closed-surface correction; rowptr
 * (int64, n+1) / pcol and colptr (n+1) / crow: refined pairs by row and by column.
 * Each face's sums are formed by one workgroup in a fixed order (no atomics). */
int mnpbem_grad_far(int n, int P, const double* pos, const double* nvec, const double* area, const int* mat0,
                    const int* mat1, const int* surf, const double* kmat, int nmat, const long* rowptr,
                    const int* pcol, const int* colptr, const int* crow, const void* L, const void* R,
                    const void* Wc, double* Gr, int batch, void* stream);
/* refined (near and self) pairs: the refinement quadrature in dual numbers.
 * Every (pair, derivative group) writes its terms to the work space Ws
 * (ws_len >= batch * P * 11 * npairs * 5 doubles), which a second kernel adds
 * per face in a fixed order: by row (rowptr, int64, n+1) and by column
 * (colptr, n+1, and cpair: the pair indices sorted by column).  Gr receives
 * the same sums in the same order on every run (no atomics). */
int mnpbem_grad_refined(int order, int n, int P, int npairs, const int* prow, const int* pcol, const int* pkind,
                        const double* pos, const double* nvec, const double* tvec1, const double* tvec2,
                        const double* area, const double* nodes, const int* mat0, const int* mat1, const int* surf,
                        const double* kmat, int nmat, const double* qnear, int nq_near, const double* qself,
                        int nq_self, const void* L, const void* R, const void* Wc, const long* rowptr,
                        const int* colptr, const int* cpair, double* Ws, long ws_len, double* Gr, int batch,
                        void* stream);
/* per-face terms: explicit normals of A (Q = [G1 X1 | G2 X2]), the plane-wave
 * right side Rb and the functional; corr: closed-surface correction per face */
int mnpbem_grad_local(int n, int P, const void* X, const void* Y, const void* Q, const void* Rb, long ldv, long sv,
                      const double* pos, const double* nvec, const double* area, const int* in_med,
                      const int* out_med, const double* pol, const double* dir, const double* pw,
                      const double* corr, const void* Wc, double* Gr, int batch, void* stream);

/* Y = beta Y + buf X, y_i = sum_j buf[i*ldm + j] x_j: the product with M^T
 * for the transposed buffer of M (the transpose of mnpbem_gemv_t_batched). */
int mnpbem_gemv_n_batched(int dtype, int n, int C, const void* M, long ldm, long sM, const void* X, long ldx,
                          long sX, void* Y, long ldy, long sY, double beta, int batch, void* stream);

/* V <- M^-T V with the LU of M^T (the transpose of mnpbem_lu_solve_t_batched). */
int mnpbem_lu_solve_n_batched(int dtype, int n, int C, const void* LU, long ld, long sLU, const int* ipiv,
                              long spiv, void* V, long ldv, long sV, void* Z, long ldz, long sZ, const void* Dinv,
                              long sD, int batch, void* stream);

/* Right side R = [phi | a | De | alpha] (n x 8P, leading dimension ldv) of the
 * BEM equations for one wavelength from an excitation's potentials: phi1,
 * phi1p, phi2, phi2p (n x P), a1, a1p, a2, a2p (n x 3 x P, index c P + p),
 * complex128, any NULL for zero; eps1, eps2: per-face complex permittivities
 * inside / outside; k0 = 2 pi / lambda.  Replaces BEMRet._excitation
 * (bem_ret.py:1360-1429; GdA&H Eqs. 10, 11, 15, 18). */
int mnpbem_excitation(int n, int P, const double* phi1, const double* phi1p, const double* a1, const double* a1p,
                      const double* phi2, const double* phi2p, const double* a2, const double* a2p,
                      const double* nvec, const double* eps1, const double* eps2, double k0, double* R, long ldv,
                      void* stream);

/* dst (double) += src (complex64 or complex128 as reals); dst = (dtype) src. */
int mnpbem_accumulate(int src_dtype, long count, const void* src, double* dst, void* stream);
int mnpbem_convert(int dst_dtype, long count, const double* src, void* dst, void* stream);

/* ---- plane-wave excitation and extinction (replaces
 *      mnpbem/simulation/planewave_ret.py:257-504, bem_ret.py:1360-1429 and
 *      mnpbem/spectrum/spectrum_ret.py:231-392) ----
 * pw[b] = {k0, Re k, Im k, Re eps1, Im eps1, Re eps2, Im eps2}; pol, dir:
 * P x 3 doubles; in_med / out_med: per face, 1 when the inside / outside of
 * the face is the embedding medium.  pot (optional, NULL to skip) receives
 * [a1 | a1' | a2 | a2'] as n x 12P per wavelength. */
int mnpbem_planewave_batched(int n, int P, const double* pos, const double* nvec, const int* in_med,
                             const int* out_med, const double* pol, const double* dir, const double* pw, double* R,
                             long ldv, long sv, double* pot, long spot, int gdiv, int batch, void* stream);

/* ext[b][p] = 4 pi / k Im(pol_p . e_p(dir_p)) from the solution X (complex128). */
int mnpbem_extinction_batched(int n, int P, const double* X, long ldv, long sv, const double* pos,
                              const double* area, const int* in_med, const int* out_med, const double* pol,
                              const double* dir, const double* pw, double* ext, int gdiv, int batch, void* stream);

/* Surface fields (replaces CompGreenRet.field, compgreen_ret.py:912-969):
 * e_i = i k0 sum_j G_ij h_j - sum_j Hp_ij sig_j, h_i = sum_j Hp_ij x h_j on
 * side `side` (0 inside with H1p = Gp + 2 pi n, hsign = +1; 1 outside with
 * H2p = Gp - 2 pi n, hsign = -1), Gp = grad G (deriv 'cart', refined pairs
 * from the coefficients g, f of mnpbem_refine_coeffs).  rowptr (int64, n + 1):
 * offsets of each row's pairs in the (row, col)-sorted list; k0: vacuum
 * wavenumber per wavelength; X: batch x n x 8P densities; E, H: batch x n x 3P
 * (index c P + p) complex128.  P <= 4. */
int mnpbem_surface_field(int n, int P, int side, double hsign, const double* pos, const double* nvec,
                         const double* area, const int* mat, const double* kmat, int nmat, const double* k0,
                         int order, const long* rowptr, const int* pcol, const double* g, const double* f,
                         const void* X, long ldx, long sx, void* E, void* H, long lde, long se, int batch,
                         void* stream);

/* Matrix-free Green products of one side for the FP64 residual of the mixed
 * solve (replaces the four complex128 matvecs of the BEM operator,
 * bem_ret_iter.py:1362-1501 afun, with the entries of CompGreenRet.eval
 * recomputed instead of stored): Y[:, yg + c] = (G X)[:, xoff + c],
 * Y[:, yh + c] = (H X)[:, xoff + c], c < cols <= 16, complex128; mat, kmat,
 * rowptr, pcol, g, f as for mnpbem_surface_field; hsign +1 (H1) / -1 (H2). */
int mnpbem_green_apply(int n, int cols, double hsign, const double* pos, const double* nvec, const double* area,
                       const int* mat, const double* kmat, int nmat, int order, const long* rowptr,
                       const int* pcol, const double* g, const double* f, const void* X, long ldx, long sx,
                       int xoff, void* Y, long ldy, long sy, int yg, int yh, int batch, void* stream);

/* Cartesian derivative of the Green function of one side, Gp (hsign 0),
 * H1p = Gp + 2 pi n_i on the diagonal (hsign +1) or H2p (hsign -1), for
 * `batch` wavelengths as (batch, n, 3, n) complex128 in the layout of
 * CompGreenRet.eval(i, i, 'Gp' / 'H1p' / 'H2p') (replaces compgreen_ret.py:
 * 780-851 and greenret_refined.py:812-864); pairs whose materials on this side
 * differ are 0.  mat, kmat, rowptr, pcol, g, f as for mnpbem_surface_field;
 * n <= 65535. */
int mnpbem_green_deriv(int n, double hsign, const double* pos, const double* nvec, const double* area,
                       const int* mat, const double* kmat, int nmat, int order, const long* rowptr, const int* pcol,
                       const double* g, const double* f, void* out, int batch, void* stream);

/* Selected entries (rows[e], cols[e]) of one side at one wavelength (kmat: nmat
 * complex wavenumbers): kind 0 G; kind 1 F + hsign 2 pi on the diagonal (F, H1,
 * H2); kind 2 Gp + hsign 2 pi n_i on the diagonal (Gp, H1p, H2p), three values
 * per entry; out: count (x 3) complex128.  Replaces CompGreenRet.eval(i, j,
 * key, enei, ind) (compgreen_ret.py:860-908, greenret_refined.py via
 * compgreen_ret.py:1770-1857; MATLAB Greenfun/@greenret/private/eval2.m). */
int mnpbem_green_select(int n, int kind, double hsign, const double* pos, const double* nvec, const double* area,
                        const int* mat, const double* kmat, int nmat, int order, const long* rowptr, const int* pcol,
                        const double* g, const double* f, long count, const int* rows, const int* cols, void* out,
                        void* stream);

/* ---- permittivity derivatives (the gradient path; no counterpart in the
 *      reference, which has no derivatives) ----
 * d/dk of the Green matrices of one side (dG = i a e^{ikd}, dF = -k (n.r) a
 * e^{ikd} / d, refined pairs (S'(k) + i d S(k)) e^{ikd}); arguments as
 * mnpbem_green_assemble, complex128 transposed outputs. */
int mnpbem_green_dk(int n, const double* pos, const double* nvec, const double* area, const int* mat,
                    const double* kmat, int nmat, int order, int npairs, const int* prow, const int* pcol,
                    const double* g, const double* f, void* dG, void* dH, long ld, long stride, int batch,
                    void* stream);
/* R = -(dA/deps1) x from Y = [G1 X1 | H1 X1 | G1' X1 | H1' X1] (n x 16P,
 * ' = d/dk1, X1 = [sig1 | h1]) and kap = dk1/deps1 per wavelength (complex). */
int mnpbem_deps_rhs(int n, int P, const void* Y, long ldy, long sy, void* R, long ldv, long sv, const double* nvec,
                    const double* par, const double* kap, int batch, void* stream);

/* Combination step of the 8N operator from precomputed products
 * Y = [G1 X1 | H1 X1 | G2 X2 | H2 X2] (n x 16P per wavelength, complex128):
 * Out = R - A X, or Out = A X when R is NULL; nrm as for
 * mnpbem_bem_residual_batched or NULL (the tail of BEMRetIter._afun,
 * bem_ret_iter.py:1440-1501, after its four Green matvecs). */
int mnpbem_bem_combine_batched(int n, int P, const void* Y, long ldy, long sy, const void* R, void* Out, long ldv,
                               long sv, const double* nvec, const double* par, double* nrm, int gdiv, int batch,
                               void* stream);

/* ---- GMRES vector kernels (replace the Arnoldi / Givens vector algebra
 *      of scipy.sparse.linalg.gmres called at mnpbem/bem/bem_iter.py:79-89) ----
 * A multi-vector holds batch x P systems: L = n * 8P complex128 values per
 * wavelength, system p owning the elements e with e % P == p.  Vector j of a
 * basis starts sVk elements after vector j-1; wavelength b starts sVb (sWb,
 * sOb) elements after b-1.  P <= 16. */
/* out[b][j][p] = sum conj(V_j) W over system (b, p), j < k; work holds
 * >= batch * k * 64 * P complex128 (fixed-order partial sums). */
int mnpbem_mv_dot_batched(long L, int P, int k, const void* V, long sVk, long sVb, const void* W, long sWb,
                          void* out, void* work, int batch, void* stream);
/* W += sign * sum_{j<k} coef[b][j][p] V_j  (coef: batch x k x P complex128) */
int mnpbem_mv_update_batched(long L, int P, int k, const void* V, long sVk, long sVb, const void* coef, double sign,
                             void* W, long sWb, int batch, void* stream);
/* mode 0: out = W * s[b][p] (complex); mode 1: out = W / sqrt(Re s[b][p]),
 * 0 where Re s <= 0.  out may alias W. */
int mnpbem_mv_scale_batched(long L, int P, int mode, const void* s, const void* W, long sWb, void* out, long sOb,
                            int batch, void* stream);
/* dst = (double) src, count reals (src complex64 or complex128) */
int mnpbem_widen(int src_dtype, long count, const void* src, double* dst, void* stream);

/* ---- H-matrices (replace HMatrix.aca / _aca_block / mtimes_vec / full,
 *      mnpbem/greenfun/hmatrix.py:159-541, with entries taken from the
 *      assembled Green matrix as in aca_compgreen_ret.py:164-193) ----
 * perm: cluster position -> face (ClusterTree.ind[:, 0], int32).  Block
 * table blk: int64 x 8 per block {r0, m, c0, n, rank, off_a, off_b, off_t}
 * (cluster rows [r0, r0+m), columns [c0, c0+n); rank < 0: dense block,
 * column-major values at val + off_a; rank >= 0: factors U (m x rank, at
 * U + off_a) and V (n x rank, at V + off_b), column-major, block = U V^T;
 * off_t = sum of the ranks of the preceding low-rank blocks).  Leaf table
 * leafs: int64 x 4 per leaf cluster {r0, m, first, last}; lst[first:last]
 * (int32) lists the blocks whose rows contain the leaf.  src is a complex128
 * matrix with element strides: M[i][j] = src[i * si + j * sj].  Block values
 * are complex64 (dtype 1) or complex128 (dtype 2). */
/* partially pivoted ACA (FP64) of blocks bids[0:nb]; workspace factors of
 * block l at Uw + woff[2l], Vw + woff[2l+1] with min(m, n, kmax) columns;
 * rank_out[l] = rank (-1 if a side exceeds 65536).  kmax <= 512. */
int mnpbem_hmat_aca(int nb, const long* blk, const int* bids, const int* perm, const void* src, long si, long sj,
                    void* Uw, void* Vw, const long* woff, int kmax, double htol, int* rank_out, void* stream);
/* dense blocks bids[0:nb] copied from src into val */
int mnpbem_hmat_dense(int dtype, int nb, const long* blk, const int* bids, const int* perm, const void* src, long si,
                      long sj, void* val, void* stream);
/* The same two fills with entries generated in place from one side's
 * geometry and refinement tables (no assembled matrix; replaces the
 * entry functions greenret::getrow / getcol of
 * mex/acagreen/acagreen.cpp:79-127, here with the refined near-field
 * entries of CompGreenRet.eval, compgreen_ret.py:507-553): kind 0 G, kind 1
 * F + hsign 2 pi [i == j] (H1: hsign +1, H2: -1).  pos, nvec (n x 3), area
 * (n), mat (int32 n, the side's material per face; faces of different
 * materials give 0), kmat (nmat complex wavenumbers of one wavelength as
 * float64 pairs), rowptr (int64 n + 1), pcol, g, f: the (row, col)-sorted
 * refined pairs and their order + 1 coefficients, as for
 * mnpbem_green_select. */
int mnpbem_hmat_aca_green(int nb, const long* blk, const int* bids, const int* perm, int kind, double hsign,
                          const double* pos, const double* nvec, const double* area, const int* mat,
                          const double* kmat, int order, const long* rowptr, const int* pcol, const double* g,
                          const double* f, void* Uw, void* Vw, const long* woff, int kmax, double htol,
                          int* rank_out, void* stream);
int mnpbem_hmat_dense_green(int dtype, int nb, const long* blk, const int* bids, const int* perm, int kind,
                            double hsign, const double* pos, const double* nvec, const double* area, const int* mat,
                            const double* kmat, int order, const long* rowptr, const int* pcol, const double* g,
                            const double* f, void* val, void* stream);
/* packed U, V of the low-rank blocks bids[0:nb] (ranks in blk) from the ACA workspaces */
int mnpbem_hmat_compact(int dtype, int nb, const long* blk, const int* bids, const void* Uw, const void* Vw,
                        const long* woff, void* U, void* V, void* stream);
/* Y (+)= M X for the n x n H-matrix, X and Y complex128 (n x ldx / ldy, face
 * order), ncol <= 16; T: complex128 work space of (sum of ranks) * ncol; XW:
 * complex128 work space of n * ncol (X in cluster order); maxchunks =
 * ceil(largest leaf / 256). */
int mnpbem_hmat_matvec(int dtype, int n, int nblk, int nleaf, int maxchunks, const long* blk, const long* leafs,
                       const int* lst, const int* perm, const void* val, const void* U, const void* V, int ncol,
                       const void* X, long ldx, void* Y, long ldy, int accumulate, void* T, void* XW,
                       void* stream);
/* Packed product layouts (also replaces HMatrix.mtimes_vec, hmatrix.py:450-541,
 * through mnpbem_hmat_matvec_packed): copies the blocks into column panels
 * (VS, tix: the V factors of the low-rank blocks sharing a column cluster,
 * side by side and transposed) and row panels (P, gidx: per leaf row cluster
 * and 64-row group, the rows of every block over it side by side); ncp / nrp
 * pieces of 8 int64 each, described in csrc/hmat.hip above k_hmat_pack_cols. */
int mnpbem_hmat_pack(int dtype, int ncp, const long* cdesc, int nrp, const long* rdesc, const void* val,
                     const void* U, const void* V, void* VS, int* tix, void* P, int* gidx, void* stream);
/* Y (+)= M X on the packed layouts (replaces HMatrix.mtimes_vec,
 * hmatrix.py:450-541): ncitem column-panel items {vs, K, o0, n, c0, og} and
 * nritem row-panel items {p, mi, K, g, r0, 0} (int64); X, Y complex128
 * (n x ldx / ldy, face order), 1 <= ncol <= 16; R: complex128 work space of
 * (n + sum of ranks) * 16 elements. */
int mnpbem_hmat_matvec_packed(int dtype, int n, int ncitem, const long* citems, const void* VS, const int* tix,
                              int nritem, const long* ritems, const void* P, const int* gidx, const int* perm,
                              int ncol, const void* X, long ldx, void* Y, long ldy, int accumulate, void* R,
                              void* stream);
/* Out = M as an n x n complex128 row-major matrix (face order, leading dimension ldo) */
int mnpbem_hmat_full(int dtype, int nblk, const long* blk, const int* perm, const void* val, const void* U,
                     const void* V, void* Out, long ldo, void* stream);
/* Recompression of the low-rank blocks bids[0:nb] (replaces HMatrix.truncate /
 * _truncate_block, hmatrix.py:672-737), step 1: QR of both factors (Q into the
 * complex128 workspace Q, factor side s of block l at qoff[2 l + s]), SVD of
 * R1 R2^T, rank r = #{cumsum(s) < (1 - htol) sum(s)} (at least 1).  Ranks
 * 1 <= k <= kst <= 128.  RT: complex128 workspace of 2 nb kst^2 elements that
 * returns the k x r transforms; XZ: the same size, needed only for kst > 64
 * (else null).  rank_out[l] = r, or -1 for a block outside these limits. */
int mnpbem_hmat_truncate(int dtype, int nb, const long* blk, const int* bids, const void* U, const void* V,
                         const long* qoff, int kst, double htol, void* Q, void* RT, void* XZ, int* rank_out,
                         void* stream);
/* Step 2: U' = Q1 T1, V' = Q2 T2 into the packed factors of blk_new (the
 * block table with the new ranks and offsets); zsplit workgroups per factor. */
int mnpbem_hmat_trunc_form(int dtype, int nb, const long* blk_old, const long* blk_new, const int* bids,
                           const long* qoff, const void* Q, const void* T, int kst, int zsplit, void* Un, void* Vn,
                           void* stream);
/* Out = alpha X + beta Y over count complex values of the storage type; Y may
 * be null (replaces the scalar multiple, negation and dense part of the sum,
 * hmatrix.py:553-670). */
int mnpbem_hmat_axpby(int dtype, long count, double ar, double ai, const void* X, double br, double bi,
                      const void* Y, void* Out, void* stream);
/* Low-rank factors of a sum on one partition: blk_n's blocks bids[0:nb] get
 * U = [Ua, Ub], V = [Va, Vb] (replaces the stacking in _plus_hmat,
 * hmatrix.py:631-666; truncate follows). */
int mnpbem_hmat_concat(int dtype, int nb, const long* blk_a, const long* blk_b, const long* blk_n, const int* bids,
                       const void* Ua, const void* Va, const void* Ub, const void* Vb, void* Un, void* Vn,
                       void* stream);
/* Diagonal dense blocks bids[0:nb]: mode 0 writes their diagonal to out
 * (complex128, face order; replaces HMatrix.diag, hmatrix.py:758-778), mode 1
 * sets it to one (HMatrix.eye_hmat, hmatrix.py:780-798). */
int mnpbem_hmat_diag(int dtype, int nb, const long* blk, const int* bids, const int* perm, void* val, void* out,
                     int mode, void* stream);

/* Scattering cross section (planewave_ret.py:391-435 over the far field of
 * spectrum_ret.py:231-459): far fields e, h in the ndir directions dirs
 * (ndir x 3, unit vectors of the far-field sphere, weights w), dsca[b][d][p] =
 * 0.5 Re(dir . (e x conj h)) / (0.5 nb), sca[b][p] = sum_d w_d dsca[b][d][p];
 * E, H: NULL or (batch, ndir, 3, P) complex128 far fields.  P <= 8. */
int mnpbem_scattering_batched(int n, int P, const double* X, long ldv, long sv, const double* pos,
                              const double* area, const int* in_med, const int* out_med, int ndir,
                              const double* dirs, const double* w, const double* pw, double* dsca, double* sca,
                              double* E, double* H, int batch, void* stream);

/* ---- measurement: HIP-event timing of every complex-GEMM launch ----
 * (host pointers for the results of mnpbem_prof_end) */
int mnpbem_prof_begin(int capacity);
int mnpbem_prof_end(double* total_ms, double* total_flops, long* launches);
/* compulsory bytes of the timed launches: A, B read once, C read and written once */
int mnpbem_prof_bytes(double* total_bytes);
/* the part of those bytes that writes C (each output element once) */
int mnpbem_prof_cbytes(double* c_write_bytes);
/* per timed launch (after mnpbem_prof_end): shapes[5 i ..] = M, N, K, batch,
 * kernel (bf16x3 tiles 0: 128 x 128, 1: 64 x 64, 2: 128 x 128 with C
 * prefetch; 3: FP32/FP64 kernel; 4: in-place 128-row product; 5: block-row
 * update with its leaf), ms[i] = its
 * HIP-event time; up to cap
 * launches; returns the number recorded.  No reference counterpart
 * (measurement only). */
int mnpbem_prof_launches(int cap, int* shapes, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* MNPBEM_AMD_H */
