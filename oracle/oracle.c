/*
 * oracle.c -- CPU restatement of the reference's contraction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: the parity tests, the smoke check and
 * the CPU-baseline leg of the benchmark load it.  The product (cytnx_amd/) never links, loads
 * or calls it, and must fail loudly when its own HIP library is missing.
 *
 * Every function cites the reference code it restates (paths relative to the Cytnx tree).
 * Each function is pinned by an independent check in tests/test_oracle.py:
 *   orc_permute        -- numpy.transpose closed form and the reference's permute fixtures
 *   orc_matmul         -- numpy.matmul closed form (F64 / C128)
 *   orc_tensordot      -- numpy.tensordot closed form and the densecontres fixtures
 *   orc_block_contract -- the contract_{L,R,ans}{1,2,3} fixtures and the dense closed form
 *   orc_block_fullcontract -- the dense closed form (sum of products)
 *   orc_tridiag_eigh   -- numpy.linalg.eigh on the assembled tridiagonal matrix
 *   orc_lanczos_gnd    -- numpy.linalg.eigh of the explicit operator, and the reference's
 *                         27x27 known eigenvalue (-1628.9964650426593)
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_F64 3
#define ORC_C128 1

/* ------------------------------------------------------------------------------------------ */
/* a2: physical permute.  Restates the odometer loop of                                       */
/* src/backend/utils_internal_cpu/Movemem_cpu.cpp:152-191: walk the source in memory order,   */
/* carry the multi-index, and recompute the destination offset only when a carry happened;    */
/* otherwise step it by the destination stride of the source's last axis.                     */
/* ------------------------------------------------------------------------------------------ */
int orc_permute(int elem_size, int rank, const uint64_t *old_shape, const uint64_t *mapper,
                const void *src, void *dst) {
  if (rank <= 0) {
    memcpy(dst, src, (size_t)elem_size);
    return 0;
  }
  uint64_t newshape[16], dst_stride[16], idx[16], inv[16];
  if (rank > 16) return -1;
  uint64_t total = 1;
  for (int i = 0; i < rank; i++) {
    newshape[i] = old_shape[mapper[i]];
    inv[mapper[i]] = (uint64_t)i;
    total *= old_shape[i];
    idx[i] = 0;
  }
  uint64_t acc = 1;
  for (int i = rank - 1; i >= 0; i--) {
    dst_stride[i] = acc;
    acc *= newshape[i];
  }
  const char *s = (const char *)src;
  char *d = (char *)dst;
  /* destination stride of the source's fastest axis: shifter_new[invmapper[rank-1]] */
  const uint64_t step = dst_stride[inv[rank - 1]];
  uint64_t loc = 0;
  for (uint64_t n = 0; n < total; n++) {
    int carried = 0;
    for (int i = rank - 1; i > 0; i--) {
      if (idx[i] >= old_shape[i]) {
        carried = 1;
        idx[i] = 0;
        idx[i - 1]++;
      } else {
        break;
      }
    }
    if (carried) {
      loc = 0;
      for (int j = 0; j < rank; j++) loc += dst_stride[j] * idx[mapper[j]];
    } else if (n != 0) {
      loc += step;
    }
    memcpy(d + loc * (uint64_t)elem_size, s + n * (uint64_t)elem_size, (size_t)elem_size);
    idx[rank - 1]++;
  }
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* a5: row-major C(MxN) = A(MxK) . B(KxN), caller-allocated output.                           */
/* The reference calls BLAS (src/backend/linalg_internal_cpu/Matmul_internal.cpp:49-61) with   */
/* beta=0; the restatement is the defining triple sum.  An optional external row-major dgemm  */
/* (cblas signature, 64-bit ints) can be installed for timing the CPU baseline.               */
/* ------------------------------------------------------------------------------------------ */
typedef void (*orc_cblas_dgemm_t)(int order, int ta, int tb, int64_t m, int64_t n, int64_t k,
                                  double alpha, const double *a, int64_t lda, const double *b,
                                  int64_t ldb, double beta, double *c, int64_t ldc);
static orc_cblas_dgemm_t g_dgemm = NULL;

void orc_set_dgemm(void *fn) { g_dgemm = (orc_cblas_dgemm_t)fn; }

static void matmul_d(double *C, const double *A, const double *B, int64_t M, int64_t K,
                     int64_t N, double beta) {
  if (g_dgemm && M > 0 && N > 0 && K > 0) {
    g_dgemm(101 /*row major*/, 111, 111, M, N, K, 1.0, A, K, B, N, beta, C, N);
    return;
  }
  for (int64_t i = 0; i < M; i++) {
    double *c = C + i * N;
    if (beta == 0.0)
      for (int64_t j = 0; j < N; j++) c[j] = 0.0;
    for (int64_t k = 0; k < K; k++) {
      const double a = A[i * K + k];
      const double *b = B + k * N;
      for (int64_t j = 0; j < N; j++) c[j] += a * b[j];
    }
  }
}

static void matmul_z(double complex *C, const double complex *A, const double complex *B,
                     int64_t M, int64_t K, int64_t N, double beta) {
  for (int64_t i = 0; i < M; i++) {
    double complex *c = C + i * N;
    if (beta == 0.0)
      for (int64_t j = 0; j < N; j++) c[j] = 0.0;
    for (int64_t k = 0; k < K; k++) {
      const double complex a = A[i * K + k];
      const double complex *b = B + k * N;
      for (int64_t j = 0; j < N; j++) c[j] += a * b[j];
    }
  }
}

int orc_matmul(int dtype, void *out, const void *A, const void *B, int64_t M, int64_t K,
               int64_t N) {
  if (dtype == ORC_F64)
    matmul_d((double *)out, (const double *)A, (const double *)B, M, K, N, 0.0);
  else if (dtype == ORC_C128)
    matmul_z((double complex *)out, (const double complex *)A, (const double complex *)B, M, K,
             N, 0.0);
  else
    return -1;
  return 0;
}

static int elem_size_of(int dtype) { return dtype == ORC_C128 ? 16 : (dtype == ORC_F64 ? 8 : 0); }

/* ------------------------------------------------------------------------------------------ */
/* a4: dense Tensordot.  Restates _Tensordot_generic (src/linalg/Tensordot.cpp:29-110):        */
/* mapperL = non-contracted L axes ++ idxl, mapperR = idxr ++ non-contracted R axes; permute   */
/* both to matrices, one GEMM, output axes = L free (in order) then R free (in order).         */
/* ------------------------------------------------------------------------------------------ */
int orc_tensordot(int dtype, int rankL, const uint64_t *shapeL, const void *L, int rankR,
                  const uint64_t *shapeR, const void *R, int ncomm, const uint64_t *idxl,
                  const uint64_t *idxr, void *out) {
  const int es = elem_size_of(dtype);
  if (!es || rankL > 16 || rankR > 16) return -1;
  uint64_t mapL[16], mapR[16];
  int isc[16];
  int nl = 0, nr = 0;
  int64_t Ml = 1, Comm = 1, Nr = 1;
  for (int i = 0; i < rankL; i++) isc[i] = 0;
  for (int c = 0; c < ncomm; c++) {
    if (shapeL[idxl[c]] != shapeR[idxr[c]]) return -2;
    isc[idxl[c]] = 1;
  }
  for (int i = 0; i < rankL; i++)
    if (!isc[i]) {
      mapL[nl++] = (uint64_t)i;
      Ml *= (int64_t)shapeL[i];
    }
  for (int c = 0; c < ncomm; c++) {
    mapL[nl++] = idxl[c];
    Comm *= (int64_t)shapeL[idxl[c]];
  }
  for (int i = 0; i < rankR; i++) isc[i] = 0;
  for (int c = 0; c < ncomm; c++) {
    mapR[nr++] = idxr[c];
    isc[idxr[c]] = 1;
  }
  for (int i = 0; i < rankR; i++)
    if (!isc[i]) {
      mapR[nr++] = (uint64_t)i;
      Nr *= (int64_t)shapeR[i];
    }
  void *Lm = malloc((size_t)(Ml * Comm) * es + 16);
  void *Rm = malloc((size_t)(Comm * Nr) * es + 16);
  if (!Lm || !Rm) {
    free(Lm);
    free(Rm);
    return -3;
  }
  orc_permute(es, rankL, shapeL, mapL, L, Lm);
  orc_permute(es, rankR, shapeR, mapR, R, Rm);
  orc_matmul(dtype, out, Lm, Rm, Ml, Comm, Nr);
  free(Lm);
  free(Rm);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* a9: BlockUniTensor::contract, general case (src/BlockUniTensor.cpp:948-1228, the non-MKL    */
/* branch :1182-1214).  For each left block a (ascending), Lmat = block_a permuted by          */
/* mapperL = non_comm1 ++ comm1 and viewed as (M x comm_dim); for each right block b whose     */
/* common-qindex tuple equals a's (ascending b, the order mp[] lists them in), Rmat_b = block_b */
/* permuted by mapperR = comm2 ++ non_comm2; the target block is the output block whose itoi  */
/* equals (itoiL[a][non_comm1], itoiR[b][non_comm2]); out[targ] += Lmat . Rmat_b.  Output      */
/* blocks start at zero (Init zero-fills, :983).                                               */
/*                                                                                            */
/* Block metadata is flat: itoi[nb*rank], shape[nb*rank], data[nb].                           */
/* ------------------------------------------------------------------------------------------ */
static int tuple_eq_sel(const uint64_t *a, const uint64_t *sa, const uint64_t *b,
                        const uint64_t *sb, int n) {
  for (int i = 0; i < n; i++)
    if (a[sa[i]] != b[sb[i]]) return 0;
  return 1;
}

int orc_block_contract(int dtype, int rankL, int nbL, const uint64_t *itoiL,
                       const uint64_t *shapeL, const void *const *dataL, int rankR, int nbR,
                       const uint64_t *itoiR, const uint64_t *shapeR, const void *const *dataR,
                       int ncomm, const uint64_t *comm1, const uint64_t *comm2, int nbC,
                       const uint64_t *itoiC, const uint64_t *shapeC, void *const *dataC) {
  const int es = elem_size_of(dtype);
  if (!es || rankL > 16 || rankR > 16) return -1;
  uint64_t nc1[16], nc2[16], mapL[16], mapR[16];
  int n1 = 0, n2 = 0, isc[16];
  for (int i = 0; i < rankL; i++) isc[i] = 0;
  for (int c = 0; c < ncomm; c++) isc[comm1[c]] = 1;
  for (int i = 0; i < rankL; i++)
    if (!isc[i]) nc1[n1++] = (uint64_t)i;
  for (int i = 0; i < rankR; i++) isc[i] = 0;
  for (int c = 0; c < ncomm; c++) isc[comm2[c]] = 1;
  for (int i = 0; i < rankR; i++)
    if (!isc[i]) nc2[n2++] = (uint64_t)i;
  for (int i = 0; i < n1; i++) mapL[i] = nc1[i];
  for (int c = 0; c < ncomm; c++) mapL[n1 + c] = comm1[c];
  for (int c = 0; c < ncomm; c++) mapR[c] = comm2[c];
  for (int i = 0; i < n2; i++) mapR[ncomm + i] = nc2[i];
  const int rankC = n1 + n2;

  /* output blocks start at zero */
  for (int t = 0; t < nbC; t++) {
    uint64_t sz = 1;
    for (int i = 0; i < rankC; i++) sz *= shapeC[(uint64_t)t * rankC + i];
    memset(dataC[t], 0, (size_t)sz * es);
  }
  void **Rmat = (void **)calloc((size_t)nbR + 1, sizeof(void *));
  uint64_t key[32];
  for (int a = 0; a < nbL; a++) {
    const uint64_t *ia = itoiL + (uint64_t)a * rankL;
    const uint64_t *sa = shapeL + (uint64_t)a * rankL;
    int64_t comm_dim = 1, M = 1;
    for (int c = 0; c < ncomm; c++) comm_dim *= (int64_t)sa[comm1[c]];
    for (int i = 0; i < n1; i++) M *= (int64_t)sa[nc1[i]];
    void *Lmat = NULL;
    for (int b = 0; b < nbR; b++) {
      const uint64_t *ib = itoiR + (uint64_t)b * rankR;
      const uint64_t *sb = shapeR + (uint64_t)b * rankR;
      if (!tuple_eq_sel(ia, comm1, ib, comm2, ncomm)) continue;
      if (!Lmat) {
        Lmat = malloc((size_t)(M * comm_dim) * es + 16);
        orc_permute(es, rankL, sa, mapL, dataL[a], Lmat);
      }
      int64_t N = 1;
      for (int i = 0; i < n2; i++) N *= (int64_t)sb[nc2[i]];
      if (!Rmat[b]) {
        Rmat[b] = malloc((size_t)(comm_dim * N) * es + 16);
        orc_permute(es, rankR, sb, mapR, dataR[b], Rmat[b]);
      }
      for (int i = 0; i < n1; i++) key[i] = ia[nc1[i]];
      for (int i = 0; i < n2; i++) key[n1 + i] = ib[nc2[i]];
      int targ = -1;
      for (int t = 0; t < nbC && targ < 0; t++) {
        const uint64_t *it = itoiC + (uint64_t)t * rankC;
        int eq = 1;
        for (int i = 0; i < rankC; i++)
          if (it[i] != key[i]) {
            eq = 0;
            break;
          }
        if (eq) targ = t;
      }
      if (targ < 0) {
        free(Lmat);
        for (int r = 0; r < nbR; r++) free(Rmat[r]);
        free(Rmat);
        return -4; /* the target block must exist in a zero-flux output */
      }
      /* out[targ] += Matmul(Lmat, Rmat[b]) : a fresh product, then an add pass */
      void *prod = malloc((size_t)(M * N) * es + 16);
      orc_matmul(dtype, prod, Lmat, Rmat[b], M, comm_dim, N);
      if (dtype == ORC_F64) {
        double *o = (double *)dataC[targ];
        const double *p = (const double *)prod;
        for (int64_t e = 0; e < M * N; e++) o[e] += p[e];
      } else {
        double complex *o = (double complex *)dataC[targ];
        const double complex *p = (const double complex *)prod;
        for (int64_t e = 0; e < M * N; e++) o[e] += p[e];
      }
      free(prod);
    }
    free(Lmat);
  }
  for (int r = 0; r < nbR; r++) free(Rmat[r]);
  free(Rmat);
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* a9 special case: every leg contracted -> scalar (src/BlockUniTensor.cpp:886-947).  Both     */
/* sides are permuted to the common-label order; each pair of blocks with equal permuted itoi */
/* contributes Vectordot(flatten, flatten) (plain, unconjugated dot).                          */
/* ------------------------------------------------------------------------------------------ */
int orc_block_fullcontract(int dtype, int rank, int nbL, const uint64_t *itoiL,
                           const uint64_t *shapeL, const void *const *dataL, int nbR,
                           const uint64_t *itoiR, const uint64_t *shapeR,
                           const void *const *dataR, const uint64_t *comm1,
                           const uint64_t *comm2, void *out) {
  const int es = elem_size_of(dtype);
  if (!es || rank > 16) return -1;
  double complex acc = 0.0;
  for (int b = 0; b < nbL; b++) {
    const uint64_t *ib = itoiL + (uint64_t)b * rank;
    for (int a = 0; a < nbR; a++) {
      const uint64_t *ia = itoiR + (uint64_t)a * rank;
      if (!tuple_eq_sel(ib, comm1, ia, comm2, rank)) continue;
      uint64_t n = 1;
      for (int i = 0; i < rank; i++) n *= shapeL[(uint64_t)b * rank + i];
      void *lp = malloc((size_t)n * es + 16), *rp = malloc((size_t)n * es + 16);
      orc_permute(es, rank, shapeL + (uint64_t)b * rank, comm1, dataL[b], lp);
      orc_permute(es, rank, shapeR + (uint64_t)a * rank, comm2, dataR[a], rp);
      if (dtype == ORC_F64) {
        double s = 0.0;
        for (uint64_t e = 0; e < n; e++) s += ((double *)lp)[e] * ((double *)rp)[e];
        acc += s;
      } else {
        double complex s = 0.0;
        for (uint64_t e = 0; e < n; e++)
          s += ((double complex *)lp)[e] * ((double complex *)rp)[e];
        acc += s;
      }
      free(lp);
      free(rp);
    }
  }
  if (dtype == ORC_F64)
    *(double *)out = creal(acc);
  else
    *(double complex *)out = acc;
  return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* a15: symmetric tridiagonal eigensolver.  The reference calls LAPACKE_dstev(jobz='V')        */
/* (src/backend/linalg_internal_cpu/Tridiag_internal.cpp:9-42) through linalg::Tridiag        */
/* (src/linalg/Tridiag.cpp:10-104).  Restated as the implicit symmetric QR algorithm with a    */
/* Wilkinson shift (Golub & Van Loan, Matrix Computations, Alg. 8.3.2 and 8.3.3): bulge-chasing */
/* Givens steps on the unreduced trailing block, deflation when |e_i| <= eps(|d_i|+|d_i+1|).   */
/* Output: ascending eigenvalues w[n]; Z is row-major n x n with ROW k = eigenvector k, the    */
/* layout Tridiag(..., is_row=true) hands to Lanczos.                                          */
/* ------------------------------------------------------------------------------------------ */
int orc_tridiag_eigh(int n, const double *diag, const double *sub, double *w, double *Z) {
  if (n <= 0) return 0;
  double *a = (double *)malloc(sizeof(double) * n);
  double *b = (double *)malloc(sizeof(double) * (n > 1 ? n - 1 : 1));
  double *Q = (double *)malloc(sizeof(double) * n * n); /* Q[i*n+k]: column k = vector k */
  memcpy(a, diag, sizeof(double) * n);
  if (n > 1) memcpy(b, sub, sizeof(double) * (n - 1));
  for (int i = 0; i < n * n; i++) Q[i] = 0.0;
  for (int i = 0; i < n; i++) Q[i * n + i] = 1.0;
  const double eps = 2.220446049250313e-16;
  int hi = n - 1, iters = 0, status = 0;
  while (hi > 0) {
    for (int i = 0; i < hi; i++)
      if (fabs(b[i]) <= eps * (fabs(a[i]) + fabs(a[i + 1]))) b[i] = 0.0;
    while (hi > 0 && b[hi - 1] == 0.0) hi--;
    if (hi == 0) break;
    int lo = hi - 1;
    while (lo > 0 && b[lo - 1] != 0.0) lo--;
    if (++iters > 60 * n) {
      status = 1;
      break;
    }
    /* Wilkinson shift from the trailing 2x2 of the unreduced block [lo, hi] */
    const double dd = 0.5 * (a[hi - 1] - a[hi]);
    const double bb = b[hi - 1];
    const double den = dd + (dd >= 0 ? 1.0 : -1.0) * hypot(dd, bb);
    const double mu = (den != 0.0) ? a[hi] - bb * bb / den : a[hi];
    double x = a[lo] - mu, z = b[lo];
    for (int k = lo; k < hi; k++) {
      const double r = hypot(x, z);
      double c = 1.0, s = 0.0;
      if (r != 0.0) {
        c = x / r;
        s = z / r;
      }
      if (k > lo) b[k - 1] = r; /* the bulge at (k-1, k+1) is annihilated */
      const double ak = a[k], ak1 = a[k + 1], bk = b[k];
      a[k] = c * c * ak + 2.0 * c * s * bk + s * s * ak1;
      a[k + 1] = s * s * ak - 2.0 * c * s * bk + c * c * ak1;
      b[k] = c * s * (ak1 - ak) + (c * c - s * s) * bk;
      if (k + 1 < hi) {
        x = b[k];
        z = s * b[k + 1]; /* new bulge at (k, k+2) */
        b[k + 1] = c * b[k + 1];
      }
      for (int i = 0; i < n; i++) {
        const double qk = Q[i * n + k], qk1 = Q[i * n + k + 1];
        Q[i * n + k] = c * qk + s * qk1;
        Q[i * n + k + 1] = -s * qk + c * qk1;
      }
    }
  }
  /* ascending order (selection sort; n is at most a few hundred) */
  int *ord = (int *)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 0; i < n; i++) {
    int m = i;
    for (int j = i + 1; j < n; j++)
      if (a[ord[j]] < a[ord[m]]) m = j;
    int t = ord[i];
    ord[i] = ord[m];
    ord[m] = t;
  }
  for (int k = 0; k < n; k++) {
    w[k] = a[ord[k]];
    if (Z)
      for (int i = 0; i < n; i++) Z[k * n + i] = Q[i * n + ord[k]];
  }
  free(ord);
  free(a);
  free(b);
  free(Q);
  return status;
}

/* ------------------------------------------------------------------------------------------ */
/* a14: Lanczos ground state, restating _Lanczos_Gnd_general_Ut                               */
/* (src/linalg/Lanczos_Gnd_Ut.cpp:29-162) on flat real vectors: plain three-term recurrence,   */
/* no re-orthogonalisation, tridiagonal eigensolve every iteration, stop when |E - E_prev| <   */
/* CvgCrit or beta == 0, eigenvector rebuilt from the stored Krylov vectors.                   */
/* matvec(x, y, ctx) must write y = H x.                                                       */
/* Returns the number of tridiagonal entries used (>= 1) or a negative error.                  */
/* ------------------------------------------------------------------------------------------ */
typedef void (*orc_matvec_t)(const double *x, double *y, void *ctx);

static double dot_d(const double *x, const double *y, int64_t n) {
  double s = 0.0;
  for (int64_t i = 0; i < n; i++) s += x[i] * y[i];
  return s;
}

int orc_lanczos_gnd(orc_matvec_t matvec, void *ctx, int64_t n, const double *Tin,
                    double CvgCrit, int Maxiter, double *E_out, double *evec_out) {
  if (Maxiter < 2 || CvgCrit <= 0 || n <= 0) return -1;
  double *V = (double *)malloc(sizeof(double) * n * (size_t)(Maxiter + 1)); /* Krylov basis */
  double *nw = (double *)malloc(sizeof(double) * n);
  double *tmp = (double *)malloc(sizeof(double) * n);
  double *As = (double *)malloc(sizeof(double) * (Maxiter + 1));
  double *Bs = (double *)malloc(sizeof(double) * (Maxiter + 1));
  double *w = (double *)malloc(sizeof(double) * (Maxiter + 1));
  double *Z = (double *)malloc(sizeof(double) * (size_t)(Maxiter + 1) * (Maxiter + 1));
  const double nrm = sqrt(dot_d(Tin, Tin, n));
  double *psi1 = V, *psi0 = NULL;
  for (int64_t i = 0; i < n; i++) psi1[i] = Tin[i] / nrm;
  int nv = 1;

  matvec(psi1, nw, ctx);
  double alpha = dot_d(nw, psi1, n);
  As[0] = alpha;
  for (int64_t i = 0; i < n; i++) nw[i] -= alpha * psi1[i];
  double beta = sqrt(dot_d(nw, nw, n));
  Bs[0] = beta;
  psi0 = psi1;
  psi1 = V + (size_t)nv * n;
  for (int64_t i = 0; i < n; i++) psi1[i] = nw[i] / beta;
  nv++;
  double E = alpha;
  int na = 1, have_eig = 0;

  for (int it = 1; it < Maxiter; it++) {
    matvec(psi1, nw, ctx);
    alpha = dot_d(nw, psi1, n);
    As[na++] = alpha;
    for (int64_t i = 0; i < n; i++) tmp[i] = alpha * psi1[i] + beta * psi0[i];
    for (int64_t i = 0; i < n; i++) nw[i] -= tmp[i];
    if (orc_tridiag_eigh(na, As, Bs, w, Z) != 0) break;
    have_eig = 1;
    beta = sqrt(dot_d(nw, nw, n));
    Bs[na - 1] = beta;
    if (beta == 0) break;
    psi0 = psi1;
    psi1 = V + (size_t)nv * n;
    for (int64_t i = 0; i < n; i++) psi1[i] = nw[i] / beta;
    nv++;
    const double Ediff = fabs(E - w[0]);
    if (Ediff < CvgCrit) break;
    E = w[0];
  }
  if (!have_eig) { /* Maxiter >= 2 always reaches one solve unless it failed */
    orc_tridiag_eigh(na, As, Bs, w, Z);
  }
  *E_out = w[0];
  if (evec_out) {
    for (int64_t i = 0; i < n; i++) evec_out[i] = 0.0;
    for (int k = 0; k < na; k++) {
      const double c = Z[k];
      const double *v = V + (size_t)k * n;
      for (int64_t i = 0; i < n; i++) evec_out[i] += c * v[i];
    }
  }
  free(V);
  free(nw);
  free(tmp);
  free(As);
  free(Bs);
  free(w);
  free(Z);
  return na;
}
