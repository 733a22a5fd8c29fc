/*
 * imsrg_oracle.c -- CPU FP64 restatement of the reference IMSRG(2) Magnus
 * commutator engine (n-parzuchowski/EOM-IMSRG_FULL, directory im-SRG/).
 *
 * TEST INFRASTRUCTURE ONLY: this file is the oracle.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker.  The product HIP library never links it.
 *
 * Every function restates one reference routine, in the reference's own
 * storage layout, and cites the file:line it follows.  Differences from the
 * reference that are deliberate:
 *   - orbits are 0-based;
 *   - pair/CC-row lookups use dense tables instead of xmap/qmap searches
 *     (same values, basic_IMSRG.f90:508-630, cross_coupled.f90:61-177);
 *   - 6j symbols come from an own Racah-formula evaluation in long double
 *     instead of anglib's log-factorial d6ji (anglib.f:285-377); pinned
 *     against sympy in tests/test_oracle_angular.py;
 *   - v_same(L) is evaluated as v_elem(...,L): on the scalar Magnus path
 *     pphh_ph is never set (only EOM_IMSRG.f90/response.f90 set it), so the
 *     COMMON-block cache of basic_IMSRG.f90:2007-2028 always equals that;
 *   - dgemm is a plain loop nest (same alpha/beta semantics as BLAS).
 */
#include "imsrg_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SQRT2 1.4142135623730951 /* sqrt(2.d0) */

/* ------------------------------------------------------------------ */
/* reference layout tables, basic_IMSRG.f90:118-134 (0-based gam index) */
static const int SEA1[6] = {1, 1, 1, 2, 3, 2};
static const int SEA2[6] = {1, 2, 3, 2, 3, 3};
static const int SQS[6] = {1, 0, 0, 1, 1, 0};

struct or_basis {
  int nsp, nhole, npart;
  int *n, *l, *jj, *tz, *con;
  int *holes, *parts, *hb4, *pb4;
  int jtot_max, lmax;
  /* single-particle (l,j,tz) groups: members in orbit order */
  int *grp_of, *grp_start, *grp_cnt, *grp_list, ngroups;
  /* two-body (J,P,Tz) blocks, basic_IMSRG.f90:477-641 */
  int nblocks;
  int *lamJ, *lamP, *lamTz;
  int *nsub;  /* nsub[3*q + c-1] */
  int **qn;   /* qn[3*q + c-1] -> 2*n orbit ints */
  int *pidx;  /* pidx[(J/2)*nsp*nsp + i*nsp + j], i<=j: index in class list */
  /* Pandya blocks, cross_coupled.f90:61-177 */
  int ncc, *ccJ, *ccnph, *ccrlen;
  int *ccrow, *cccol; /* [(q*nsp + a)*nsp + b] */
  /* 6j table {j1 j2 J3; j4 j5 J6}: j half-integer, J integer */
  int nhalf, nJ;
  double *sixj;
  int64_t neq;
};

struct or_op {
  const or_basis *B;
  int herm;
  double E0;
  double *fhh, *fph, *fpp; /* fhh(A,A) fph(P,A) fpp(P,P), column-major */
  double *buf;             /* all gam arrays */
  double **gam;            /* gam[6*q + g] */
  size_t nbuf;
};

struct or_cc {
  const or_basis *B;
  int herm;
  double **X; /* X[q] rlen x nph column-major */
  double *buf;
  double **W; /* scratch W[q] rlen x rlen */
  double *wbuf;
};

static int g_threads = 0;
void or_set_threads(int n) { g_threads = n; }
static int nthreads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

static void *xcalloc(size_t n, size_t s) {
  void *p = calloc(n ? n : 1, s);
  if (!p) {
    fprintf(stderr, "oracle: out of memory\n");
    abort();
  }
  return p;
}

static inline int triangle(int j1, int j2, int J) { return J <= j1 + j2 && J >= abs(j1 - j2); }
static inline int ipow_m1(int e) { return (e % 2 == 0) ? 1 : -1; } /* (-1)**e, any sign */

/* ------------------------------------------------------------------ */
/* 6j / 9j (own Racah evaluation; reference anglib.f:285-377, 546-566) */
static long double g_fact[256];
static int g_fact_ready = 0;
static void init_fact(void) {
  if (g_fact_ready) return;
  g_fact[0] = 1.0L;
  for (int i = 1; i < 256; i++) g_fact[i] = g_fact[i - 1] * (long double)i;
  g_fact_ready = 1;
}
static long double tri_delta(int a, int b, int c) {
  return sqrtl(g_fact[(a + b - c) / 2] * g_fact[(a - b + c) / 2] * g_fact[(-a + b + c) / 2] /
               g_fact[(a + b + c) / 2 + 1]);
}
static int tri_ok(int a, int b, int c) {
  return a >= 0 && b >= 0 && c >= 0 && c <= a + b && c >= abs(a - b) && ((a + b + c) & 1) == 0;
}
/* {j1 j2 j3; j4 j5 j6}, all doubled */
static double racah6j(int j1, int j2, int j3, int j4, int j5, int j6) {
  init_fact();
  if (!tri_ok(j1, j2, j3) || !tri_ok(j1, j5, j6) || !tri_ok(j4, j2, j6) || !tri_ok(j4, j5, j3))
    return 0.0;
  int a1 = (j1 + j2 + j3) / 2, a2 = (j1 + j5 + j6) / 2, a3 = (j4 + j2 + j6) / 2,
      a4 = (j4 + j5 + j3) / 2;
  int b1 = (j1 + j2 + j4 + j5) / 2, b2 = (j2 + j3 + j5 + j6) / 2, b3 = (j3 + j1 + j6 + j4) / 2;
  int tmin = a1;
  if (a2 > tmin) tmin = a2;
  if (a3 > tmin) tmin = a3;
  if (a4 > tmin) tmin = a4;
  int tmax = b1;
  if (b2 < tmax) tmax = b2;
  if (b3 < tmax) tmax = b3;
  long double s = 0.0L;
  for (int t = tmin; t <= tmax; t++) {
    long double den = g_fact[t - a1] * g_fact[t - a2] * g_fact[t - a3] * g_fact[t - a4] *
                      g_fact[b1 - t] * g_fact[b2 - t] * g_fact[b3 - t];
    long double term = g_fact[t + 1] / den;
    s += (t & 1) ? -term : term;
  }
  s *= tri_delta(j1, j2, j3) * tri_delta(j1, j5, j6) * tri_delta(j4, j2, j6) *
       tri_delta(j4, j5, j3);
  return (double)s;
}
double or_ninej(int j1, int j2, int j3, int j4, int j5, int j6, int j7, int j8, int j9) {
  /* sum over x of (-1)^{2x}(2x+1){j1 j4 j7; j8 j9 x}{j2 j5 j8; j4 x j6}{j3 j6 j9; x j1 j2} */
  int xmin = abs(j1 - j9), t;
  t = abs(j4 - j8);
  if (t > xmin) xmin = t;
  t = abs(j2 - j6);
  if (t > xmin) xmin = t;
  int xmax = j1 + j9;
  if (j4 + j8 < xmax) xmax = j4 + j8;
  if (j2 + j6 < xmax) xmax = j2 + j6;
  long double s = 0.0L;
  for (int x = xmin; x <= xmax; x += 2) {
    long double v = (long double)(x + 1) * racah6j(j1, j4, j7, j8, j9, x) *
                    racah6j(j2, j5, j8, j4, x, j6) * racah6j(j3, j6, j9, x, j1, j2);
    s += (x & 1) ? -v : v;
  }
  return (double)s;
}

/* sixj(j1,j2,j3,j4,j5,j6) of basic_IMSRG.f90:2509-2555: j1,j2,j4,j5 half-integer,
 * j3,j6 integer, triangle-checked table lookup */
double or_sixj(const or_basis *B, int j1, int j2, int j3, int j4, int j5, int j6) {
  if (!(triangle(j1, j2, j3) && triangle(j4, j5, j3) && triangle(j1, j5, j6) && triangle(j4, j2, j6)))
    return 0.0;
  int h = B->nhalf, nJ = B->nJ;
  int i1 = (j1 - 1) / 2, i2 = (j2 - 1) / 2, i4 = (j4 - 1) / 2, i5 = (j5 - 1) / 2;
  if (i1 >= h || i2 >= h || i4 >= h || i5 >= h || j3 / 2 >= nJ || j6 / 2 >= nJ)
    return racah6j(j1, j2, j3, j4, j5, j6);
  size_t idx = ((((size_t)i1 * h + i2) * h + i4) * h + i5) * nJ * nJ + (size_t)(j3 / 2) * nJ + j6 / 2;
  return B->sixj[idx];
}

/* ------------------------------------------------------------------ */
/* basis: read_sp_basis side effects (basic_IMSRG.f90:152-342), find_holes
 * (:345-400), allocate_blocks (:477-641), allocate_CCMAT (cross_coupled.f90:61-177),
 * store_6j (:2183-2288) */
or_basis *or_basis_create(int nsp, const int *n, const int *l, const int *jj, const int *tz,
                          const int *con) {
  or_basis *B = (or_basis *)xcalloc(1, sizeof(or_basis));
  B->nsp = nsp;
  B->n = (int *)xcalloc(nsp, sizeof(int));
  B->l = (int *)xcalloc(nsp, sizeof(int));
  B->jj = (int *)xcalloc(nsp, sizeof(int));
  B->tz = (int *)xcalloc(nsp, sizeof(int));
  B->con = (int *)xcalloc(nsp, sizeof(int));
  memcpy(B->n, n, nsp * sizeof(int));
  memcpy(B->l, l, nsp * sizeof(int));
  memcpy(B->jj, jj, nsp * sizeof(int));
  memcpy(B->tz, tz, nsp * sizeof(int));
  memcpy(B->con, con, nsp * sizeof(int));
  int A = 0;
  for (int i = 0; i < nsp; i++) A += con[i];
  B->nhole = A;
  B->npart = nsp - A;
  B->holes = (int *)xcalloc(A, sizeof(int));
  B->parts = (int *)xcalloc(nsp - A, sizeof(int));
  B->hb4 = (int *)xcalloc(nsp, sizeof(int));
  B->pb4 = (int *)xcalloc(nsp, sizeof(int));
  int r1 = 0, r2 = 0, hsum = 0;
  for (int i = 0; i < nsp; i++) {
    B->hb4[i] = hsum;          /* holes before i */
    B->pb4[i] = i - hsum;      /* particles before i */
    if (con[i]) B->holes[r1++] = i;
    else B->parts[r2++] = i;
    hsum += con[i];
  }
  B->jtot_max = 0;
  B->lmax = 0;
  for (int i = 0; i < nsp; i++) {
    if (jj[i] > B->jtot_max) B->jtot_max = jj[i];
    if (l[i] > B->lmax) B->lmax = l[i];
  }
  /* (l,j,tz) groups (basic_IMSRG.f90:318-340): members listed in orbit order */
  int ng = 0;
  B->grp_of = (int *)xcalloc(nsp, sizeof(int));
  B->grp_start = (int *)xcalloc(nsp, sizeof(int));
  B->grp_cnt = (int *)xcalloc(nsp, sizeof(int));
  B->grp_list = (int *)xcalloc(nsp, sizeof(int));
  for (int i = 0; i < nsp; i++) B->grp_of[i] = -1;
  int pos = 0;
  for (int i = 0; i < nsp; i++) {
    if (B->grp_of[i] >= 0) continue;
    B->grp_start[ng] = pos;
    for (int k = i; k < nsp; k++)
      if (l[k] == l[i] && jj[k] == jj[i] && tz[k] == tz[i]) {
        B->grp_of[k] = ng;
        B->grp_list[pos++] = k;
        B->grp_cnt[ng]++;
      }
    ng++;
  }
  B->ngroups = ng;
  /* blocks: J slowest, Tz, P fastest (basic_IMSRG.f90:528-531) */
  int JTM = B->jtot_max;
  B->nblocks = (JTM + 1) * 6;
  B->lamJ = (int *)xcalloc(B->nblocks, sizeof(int));
  B->lamP = (int *)xcalloc(B->nblocks, sizeof(int));
  B->lamTz = (int *)xcalloc(B->nblocks, sizeof(int));
  B->nsub = (int *)xcalloc(3 * B->nblocks, sizeof(int));
  B->qn = (int **)xcalloc(3 * B->nblocks, sizeof(int *));
  size_t nn2 = (size_t)nsp * nsp;
  B->pidx = (int *)xcalloc((size_t)(JTM + 1) * nn2, sizeof(int));
  for (size_t i = 0; i < (size_t)(JTM + 1) * nn2; i++) B->pidx[i] = -1;
  int q = 0;
  for (int Jt = 0; Jt <= 2 * JTM; Jt += 2)
    for (int Tz = -1; Tz <= 1; Tz++)
      for (int P = 0; P <= 1; P++) {
        B->lamJ[q] = Jt;
        B->lamP[q] = P;
        B->lamTz[q] = Tz;
        int cnt[3] = {0, 0, 0};
        for (int pass = 0; pass < 2; pass++) {
          cnt[0] = cnt[1] = cnt[2] = 0;
          for (int i = 0; i < nsp; i++)
            for (int j = i; j < nsp; j++) {
              if (Tz != (tz[i] + tz[j]) / 2) continue;
              if (P != (l[i] + l[j]) % 2) continue;
              if (!triangle(jj[i], jj[j], Jt)) continue;
              int cx = con[i] + con[j];
              int k = cnt[cx]++;
              if (pass == 1) {
                B->qn[3 * q + cx][2 * k] = i;
                B->qn[3 * q + cx][2 * k + 1] = j;
                B->pidx[(size_t)(Jt / 2) * nn2 + (size_t)i * nsp + j] = k;
              }
            }
          if (pass == 0)
            for (int c = 0; c < 3; c++) {
              B->nsub[3 * q + c] = cnt[c];
              B->qn[3 * q + c] = (int *)xcalloc(2 * (cnt[c] ? cnt[c] : 1), sizeof(int));
            }
        }
        q++;
      }
  /* Pandya blocks: JC fastest, |Tz|, P slowest (cross_coupled.f90:108-114) */
  B->ncc = (JTM + 1) * 4;
  B->ccJ = (int *)xcalloc(B->ncc, sizeof(int));
  B->ccnph = (int *)xcalloc(B->ncc, sizeof(int));
  B->ccrlen = (int *)xcalloc(B->ncc, sizeof(int));
  B->ccrow = (int *)xcalloc((size_t)B->ncc * nn2, sizeof(int));
  B->cccol = (int *)xcalloc((size_t)B->ncc * nn2, sizeof(int));
  for (size_t i = 0; i < (size_t)B->ncc * nn2; i++) B->ccrow[i] = B->cccol[i] = -1;
  for (int q1 = 0; q1 < B->ncc; q1++) {
    int JC = (q1 % (JTM + 1)) * 2;
    int PAR = q1 / (2 * JTM + 2);
    int TZ = (q1 % (2 * JTM + 2)) / (JTM + 1);
    int nb = 0, r = 0;
    for (int i = 0; i < nsp; i++)
      for (int j = 0; j < nsp; j++) {
        if (!triangle(jj[i], jj[j], JC)) continue;
        if ((l[i] + l[j]) % 2 != PAR) continue;
        if (abs(tz[i] - tz[j]) / 2 != TZ) continue;
        if (con[i] == 0 && con[j] == 1) B->cccol[((size_t)q1 * nsp + i) * nsp + j] = nb++;
        B->ccrow[((size_t)q1 * nsp + i) * nsp + j] = r++;
      }
    B->ccJ[q1] = JC;
    B->ccnph[q1] = nb;
    B->ccrlen[q1] = r;
  }
  /* neq (IMSRG_ODE.f90:36-44, the length of vectorize) */
  int64_t neq = 1 + (int64_t)A * A + (int64_t)A * (nsp - A) + (int64_t)(nsp - A) * (nsp - A);
  for (q = 0; q < B->nblocks; q++) {
    int64_t np = B->nsub[3 * q], nb = B->nsub[3 * q + 1], nh = B->nsub[3 * q + 2];
    neq += (nh * nh + nh + nb * nb + nb + np * np + np) / 2 + nb * np + nh * np + nh * nb;
  }
  B->neq = neq;
  /* 6j table */
  B->nhalf = (JTM + 1) / 2;
  B->nJ = JTM + 1;
  size_t h = B->nhalf, nJ = B->nJ;
  B->sixj = (double *)xcalloc(h * h * h * h * nJ * nJ, sizeof(double));
  for (size_t i1 = 0; i1 < h; i1++)
    for (size_t i2 = 0; i2 < h; i2++)
      for (size_t i4 = 0; i4 < h; i4++)
        for (size_t i5 = 0; i5 < h; i5++)
          for (size_t a = 0; a < nJ; a++)
            for (size_t b = 0; b < nJ; b++)
              B->sixj[((((i1 * h + i2) * h + i4) * h + i5) * nJ + a) * nJ + b] =
                  racah6j(2 * i1 + 1, 2 * i2 + 1, 2 * a, 2 * i4 + 1, 2 * i5 + 1, 2 * b);
  return B;
}

void or_basis_free(or_basis *B) {
  if (!B) return;
  free(B->n); free(B->l); free(B->jj); free(B->tz); free(B->con);
  free(B->holes); free(B->parts); free(B->hb4); free(B->pb4);
  free(B->grp_of); free(B->grp_start); free(B->grp_cnt); free(B->grp_list);
  for (int i = 0; i < 3 * B->nblocks; i++) free(B->qn[i]);
  free(B->qn); free(B->nsub); free(B->lamJ); free(B->lamP); free(B->lamTz); free(B->pidx);
  free(B->ccJ); free(B->ccnph); free(B->ccrlen); free(B->ccrow); free(B->cccol);
  free(B->sixj);
  free(B);
}
int64_t or_basis_neq(const or_basis *B) { return B->neq; }
int or_basis_nblocks(const or_basis *B) { return B->nblocks; }
void or_basis_block_info(const or_basis *B, int q, int *info) {
  info[0] = B->lamJ[q]; info[1] = B->lamP[q]; info[2] = B->lamTz[q];
  info[3] = B->nsub[3 * q]; info[4] = B->nsub[3 * q + 1]; info[5] = B->nsub[3 * q + 2];
}
const int *or_basis_block_pairs(const or_basis *B, int q, int c) { return B->qn[3 * q + c - 1]; }
int or_basis_ncc(const or_basis *B) { return B->ncc; }
void or_basis_cc_info(const or_basis *B, int q, int *info) {
  info[0] = B->ccJ[q]; info[1] = B->ccrlen[q]; info[2] = B->ccnph[q];
}

static inline int nclass(const or_basis *B, int q, int c) { return B->nsub[3 * q + c - 1]; }
static inline int grows(const or_basis *B, int q, int g) { return nclass(B, q, SEA1[g]); }
static inline int gcols(const or_basis *B, int q, int g) { return nclass(B, q, SEA2[g]); }
static inline int pair_index(const or_basis *B, int J, int i, int j) {
  return B->pidx[(size_t)(J / 2) * B->nsp * B->nsp + (size_t)i * B->nsp + j];
}
static inline int block_index(int J, int T, int P) { return 3 * J + 2 * (T + 1) + P; } /* :1064-1070 */

/* ------------------------------------------------------------------ */
/* operators: allocate_blocks / duplicate_sq_op (basic_IMSRG.f90:477-641, 2736-2918) */
or_op *or_op_create(const or_basis *B, int herm) {
  or_op *op = (or_op *)xcalloc(1, sizeof(or_op));
  op->B = B;
  op->herm = herm;
  op->E0 = 0.0;
  int A = B->nhole, P = B->npart;
  op->fhh = (double *)xcalloc((size_t)A * A, sizeof(double));
  op->fph = (double *)xcalloc((size_t)P * A, sizeof(double));
  op->fpp = (double *)xcalloc((size_t)P * P, sizeof(double));
  size_t tot = 0;
  for (int q = 0; q < B->nblocks; q++)
    for (int g = 0; g < 6; g++) tot += (size_t)grows(B, q, g) * gcols(B, q, g);
  op->nbuf = tot;
  op->buf = (double *)xcalloc(tot, sizeof(double));
  op->gam = (double **)xcalloc(6 * B->nblocks, sizeof(double *));
  size_t off = 0;
  for (int q = 0; q < B->nblocks; q++)
    for (int g = 0; g < 6; g++) {
      op->gam[6 * q + g] = op->buf + off;
      off += (size_t)grows(B, q, g) * gcols(B, q, g);
    }
  return op;
}
void or_op_free(or_op *op) {
  if (!op) return;
  free(op->fhh); free(op->fph); free(op->fpp); free(op->buf); free(op->gam);
  free(op);
}
void or_op_set_herm(or_op *op, int herm) { op->herm = herm; }
int or_op_get_herm(const or_op *op) { return op->herm; }
double or_op_get_e0(const or_op *op) { return op->E0; }
void or_op_set_e0(or_op *op, double e0) { op->E0 = e0; }

#define GX(op, q, g, i, j) ((op)->gam[6 * (q) + (g)][(size_t)(i) + (size_t)(j) * grows((op)->B, q, g)])

/* repackage, basic_IMSRG.f90:3857-3918 */
void or_op_from_vec(or_op *op, const double *v) {
  const or_basis *B = op->B;
  int A = B->nhole, P = B->npart;
  size_t k = 0;
  op->E0 = v[k++];
  for (size_t i = 0; i < (size_t)A * A; i++) op->fhh[i] = v[k++];
  for (size_t i = 0; i < (size_t)P * P; i++) op->fpp[i] = v[k++];
  for (size_t i = 0; i < (size_t)P * A; i++) op->fph[i] = v[k++];
  for (int q = 0; q < B->nblocks; q++)
    for (int g = 0; g < 6; g++) {
      int nr = grows(B, q, g), nc = gcols(B, q, g);
      for (int i = 0; i < nc; i++)
        for (int j = SQS[g] ? i : 0; j < nr; j++) {
          GX(op, q, g, j, i) = v[k];
          if (SQS[g]) GX(op, q, g, i, j) = v[k] * op->herm;
          k++;
        }
    }
}
/* vectorize, basic_IMSRG.f90:3794-3854 */
void or_op_to_vec(const or_op *op, double *v) {
  const or_basis *B = op->B;
  int A = B->nhole, P = B->npart;
  size_t k = 0;
  v[k++] = op->E0;
  for (size_t i = 0; i < (size_t)A * A; i++) v[k++] = op->fhh[i];
  for (size_t i = 0; i < (size_t)P * P; i++) v[k++] = op->fpp[i];
  for (size_t i = 0; i < (size_t)P * A; i++) v[k++] = op->fph[i];
  for (int q = 0; q < B->nblocks; q++)
    for (int g = 0; g < 6; g++) {
      int nr = grows(B, q, g), nc = gcols(B, q, g);
      for (int i = 0; i < nc; i++)
        for (int j = SQS[g] ? i : 0; j < nr; j++) v[k++] = GX(op, q, g, j, i);
    }
}

/* f_elem, basic_IMSRG.f90:1147-1177 */
double or_f_elem(int a, int b, const or_op *op) {
  const or_basis *B = op->B;
  int c1 = B->con[a], c2 = B->con[b], A = B->nhole, P = B->npart;
  (void)A;
  switch (c1 + c2) {
    case 0: return op->fpp[(size_t)(a - B->hb4[a]) + (size_t)(b - B->hb4[b]) * P];
    case 1:
      if (c1 > c2) return op->fph[(size_t)(b - B->hb4[b]) + (size_t)(a - B->pb4[a]) * P] * op->herm;
      return op->fph[(size_t)(a - B->hb4[a]) + (size_t)(b - B->pb4[b]) * P];
    default: return op->fhh[(size_t)(a - B->pb4[a]) + (size_t)(b - B->pb4[b]) * B->nhole];
  }
}

/* v_elem, basic_IMSRG.f90:1270-1385 (pphh_ph branch is EOM-only) */
double or_v_elem(int a, int b, int c, int d, int J, const or_op *op) {
  const or_basis *B = op->B;
  int ja = B->jj[a], jb = B->jj[b], jc = B->jj[c], jd = B->jj[d];
  if (!(triangle(ja, jb, J) && triangle(jc, jd, J))) return 0.0;
  int P = (B->l[a] + B->l[b]) % 2;
  if ((B->l[c] + B->l[d]) % 2 != P) return 0.0;
  int T = (B->tz[a] + B->tz[b]) / 2;
  if (B->tz[c] + B->tz[d] != 2 * T) return 0.0;
  int q = block_index(J, T, P);
  int C1 = B->con[a] + B->con[b] + 1, C2 = B->con[c] + B->con[d] + 1;
  int qx = C1 * C2;
  int g = (qx == 9) ? 4 : (qx == 6 ? 5 : qx - 1);
  double pre = 1.0;
  int i1, i2;
  if (a > b) {
    i1 = pair_index(B, J, b, a);
    pre = ipow_m1(1 + (ja + jb - J) / 2);
  } else {
    if (a == b) pre = pre * SQRT2;
    i1 = pair_index(B, J, a, b);
  }
  if (c > d) {
    i2 = pair_index(B, J, d, c);
    pre = pre * ipow_m1(1 + (jc + jd - J) / 2);
  } else {
    if (c == d) pre = pre * SQRT2;
    i2 = pair_index(B, J, c, d);
  }
  if (C1 > C2) return GX(op, q, g, i2, i1) * op->herm * pre;
  return GX(op, q, g, i1, i2) * pre;
}

/* twobody_monopole, basic_IMSRG.f90:2031-2046 */
static double twobody_monopole(int a, int b, int ja, int jb, const or_op *H) {
  double sm = 0.0;
  for (int JT = abs(ja - jb); JT <= ja + jb; JT += 2) sm = sm + (JT + 1) * or_v_elem(a, b, a, b, JT, H);
  return sm / (ja + 1.0) / (jb + 1.0);
}

/* ------------------------------------------------------------------ */
/* column-major dgemm with BLAS alpha/beta semantics */
static void or_dgemm(char ta, char tb, int m, int n, int k, double alpha, const double *A, int lda,
                     const double *Bm, int ldb, double beta, double *C, int ldc) {
  if (m <= 0 || n <= 0) return;
  for (int j = 0; j < n; j++) {
    double *cj = C + (size_t)j * ldc;
    if (beta == 0.0)
      for (int i = 0; i < m; i++) cj[i] = 0.0;
    else if (beta != 1.0)
      for (int i = 0; i < m; i++) cj[i] *= beta;
    if (alpha == 0.0 || k <= 0) continue;
    if (ta == 'N') {
      for (int l = 0; l < k; l++) {
        double b = (tb == 'N') ? Bm[(size_t)l + (size_t)j * ldb] : Bm[(size_t)j + (size_t)l * ldb];
        b *= alpha;
        const double *al = A + (size_t)l * lda;
        for (int i = 0; i < m; i++) cj[i] += al[i] * b;
      }
    } else {
      for (int i = 0; i < m; i++) {
        const double *ai = A + (size_t)i * lda;
        double s = 0.0;
        if (tb == 'N') {
          const double *bj = Bm + (size_t)j * ldb;
          for (int l = 0; l < k; l++) s += ai[l] * bj[l];
        } else {
          for (int l = 0; l < k; l++) s += ai[l] * Bm[(size_t)j + (size_t)l * ldb];
        }
        cj[i] += alpha * s;
      }
    }
  }
}

/* ------------------------------------------------------------------ */
/* commutator_110, commutators.f90:10-33 */
double or_comm110(const or_op *L, const or_op *R) {
  const or_basis *B = L->B;
  int A = B->nhole, P = B->npart;
  double sm = 0.0;
  for (int i = 0; i < A; i++) {
    int ji = B->jj[B->holes[i]];
    for (int a = 0; a < P; a++) {
      double lf = L->fph[a + (size_t)i * P], rf = R->fph[a + (size_t)i * P];
      sm = sm + (lf * rf * L->herm - rf * lf * R->herm) * (ji + 1);
    }
  }
  return sm;
}

/* commutator_111, commutators.f90:36-80 */
void or_comm111(const or_op *L, const or_op *R, or_op *RES) {
  const or_basis *B = L->B;
  int hol = B->nhole, par = B->npart;
  double hL = L->herm, hR = R->herm;
  double *th1 = xcalloc((size_t)hol * hol, 8), *th2 = xcalloc((size_t)hol * hol, 8);
  double *tb1 = xcalloc((size_t)par * hol, 8), *tb2 = xcalloc((size_t)par * hol, 8);
  double *tp1 = xcalloc((size_t)par * par, 8), *tp2 = xcalloc((size_t)par * par, 8);
  or_dgemm('N', 'N', hol, hol, hol, 1.0, L->fhh, hol, R->fhh, hol, 0.0, th1, hol);
  or_dgemm('T', 'N', hol, hol, par, 1.0, L->fph, par, R->fph, par, 0.0, th2, hol);
  for (int j = 0; j < hol; j++)
    for (int i = 0; i < hol; i++) {
      size_t ij = i + (size_t)j * hol, ji = j + (size_t)i * hol;
      RES->fhh[ij] = th1[ij] + th2[ij] * hL - (th1[ji] * hL + th2[ji]) * hR;
    }
  or_dgemm('N', 'N', par, hol, hol, 1.0, L->fph, par, R->fhh, hol, 0.0, tb1, par);
  or_dgemm('N', 'N', par, hol, par, 1.0, L->fpp, par, R->fph, par, 0.0, tb2, par);
  for (size_t i = 0; i < (size_t)par * hol; i++) RES->fph[i] = tb1[i] + tb2[i];
  or_dgemm('N', 'N', par, hol, hol, 1.0, R->fph, par, L->fhh, hol, 0.0, tb1, par);
  or_dgemm('N', 'N', par, hol, par, 1.0, R->fpp, par, L->fph, par, 0.0, tb2, par);
  for (size_t i = 0; i < (size_t)par * hol; i++) RES->fph[i] = RES->fph[i] - tb1[i] - tb2[i];
  or_dgemm('N', 'T', par, par, hol, 1.0, L->fph, par, R->fph, par, 0.0, tp1, par);
  or_dgemm('N', 'N', par, par, par, 1.0, L->fpp, par, R->fpp, par, 0.0, tp2, par);
  for (int j = 0; j < par; j++)
    for (int i = 0; i < par; i++) {
      size_t ij = i + (size_t)j * par, ji = j + (size_t)i * par;
      RES->fpp[ij] = tp1[ij] * hR + tp2[ij] - (tp1[ji] + tp2[ji] * hR) * hL;
    }
  free(th1); free(th2); free(tb1); free(tb2); free(tp1); free(tp2);
}

/* shared inner sum of commutator_121 (commutators.f90:83-280) for orbits pk,qk */
static double c121_sum(const or_op *L, const or_op *R, int pk, int qk) {
  const or_basis *B = L->B;
  int A = B->nhole, P = B->npart, jq = B->jj[qk];
  double sm = 0.0;
  for (int a = 0; a < P; a++) {
    int ak = B->parts[a];
    for (int i = 0; i < A; i++) {
      int ik = B->holes[i];
      if (B->jj[ik] != B->jj[ak] || B->l[ik] != B->l[ak] || B->tz[ik] != B->tz[ak]) continue;
      int ji = B->jj[ik];
      double smx = 0, smy = 0, smx2 = 0, smy2 = 0;
      for (int JT = abs(ji - jq); JT <= ji + jq; JT += 2) {
        smx = smx + or_v_elem(ak, pk, ik, qk, JT, R) * (JT + 1);
        smx2 = smx2 + or_v_elem(ak, pk, ik, qk, JT, L) * (JT + 1);
        smy = smy + or_v_elem(ik, pk, ak, qk, JT, R) * (JT + 1);
        smy2 = smy2 + or_v_elem(ik, pk, ak, qk, JT, L) * (JT + 1);
      }
      sm = sm + L->fph[a + (size_t)i * P] * (L->herm * smx - smy) -
           R->fph[a + (size_t)i * P] * (R->herm * smx2 - smy2);
    }
  }
  return sm;
}
static inline int same_ljt(const or_basis *B, int a, int b) {
  return B->jj[a] == B->jj[b] && B->l[a] == B->l[b] && B->tz[a] == B->tz[b];
}

/* commutator_121, commutators.f90:83-280 */
void or_comm121(const or_op *L, const or_op *R, or_op *RES) {
  const or_basis *B = L->B;
  int A = B->nhole, P = B->npart;
  for (int p = 0; p < A; p++)
    for (int q = p; q < A; q++) {
      int pk = B->holes[p], qk = B->holes[q];
      if (!same_ljt(B, pk, qk)) continue;
      double sm = c121_sum(L, R, pk, qk);
      RES->fhh[p + (size_t)q * A] = RES->fhh[p + (size_t)q * A] + sm / (B->jj[pk] + 1.0);
      RES->fhh[q + (size_t)p * A] = RES->fhh[p + (size_t)q * A] * RES->herm;
    }
  for (int p = 0; p < P; p++)
    for (int q = p; q < P; q++) {
      int pk = B->parts[p], qk = B->parts[q];
      if (!same_ljt(B, pk, qk)) continue;
      double sm = c121_sum(L, R, pk, qk);
      RES->fpp[p + (size_t)q * P] = RES->fpp[p + (size_t)q * P] + sm / (B->jj[pk] + 1.0);
      RES->fpp[q + (size_t)p * P] = RES->fpp[p + (size_t)q * P] * RES->herm;
    }
  for (int p = 0; p < P; p++)
    for (int q = 0; q < A; q++) {
      int pk = B->parts[p], qk = B->holes[q];
      if (!same_ljt(B, pk, qk)) continue;
      double sm = c121_sum(L, R, pk, qk);
      RES->fph[p + (size_t)q * P] = RES->fph[p + (size_t)q * P] + sm / (B->jj[pk] + 1.0);
    }
}

/* commutator_122, commutators.f90:283-396 (overwrites every stored element) */
void or_comm122(const or_op *L, const or_op *R, or_op *RES) {
  const or_basis *B = L->B;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
  for (int q = 0; q < B->nblocks; q++) {
    int JT = B->lamJ[q];
    for (int g = 0; g < 6; g++) {
      int n1 = grows(B, q, g), n2 = gcols(B, q, g);
      if (n1 * n2 == 0) continue;
      const int *qa = B->qn[3 * q + SEA1[g] - 1], *qb = B->qn[3 * q + SEA2[g] - 1];
      for (int IX = 0; IX < n1; IX++) {
        int a = qa[2 * IX], b = qa[2 * IX + 1];
        for (int JX = SQS[g] ? IX : 0; JX < n2; JX++) {
          int c = qb[2 * JX], d = qb[2 * JX + 1];
          double sm = 0.0;
          int gs, cnt;
          gs = B->grp_start[B->grp_of[a]]; cnt = B->grp_cnt[B->grp_of[a]];
          for (int t = 0; t < cnt; t++) {
            int i = B->grp_list[gs + t];
            sm = sm + or_f_elem(a, i, L) * or_v_elem(i, b, c, d, JT, R) - or_f_elem(a, i, R) * or_v_elem(i, b, c, d, JT, L);
          }
          gs = B->grp_start[B->grp_of[b]]; cnt = B->grp_cnt[B->grp_of[b]];
          for (int t = 0; t < cnt; t++) {
            int i = B->grp_list[gs + t];
            sm = sm + or_f_elem(b, i, L) * or_v_elem(a, i, c, d, JT, R) - or_f_elem(b, i, R) * or_v_elem(a, i, c, d, JT, L);
          }
          gs = B->grp_start[B->grp_of[c]]; cnt = B->grp_cnt[B->grp_of[c]];
          for (int t = 0; t < cnt; t++) {
            int i = B->grp_list[gs + t];
            sm = sm - or_f_elem(i, c, L) * or_v_elem(a, b, i, d, JT, R) + or_f_elem(i, c, R) * or_v_elem(a, b, i, d, JT, L);
          }
          gs = B->grp_start[B->grp_of[d]]; cnt = B->grp_cnt[B->grp_of[d]];
          for (int t = 0; t < cnt; t++) {
            int i = B->grp_list[gs + t];
            sm = sm - or_f_elem(i, d, L) * or_v_elem(a, b, c, i, JT, R) + or_f_elem(i, d, R) * or_v_elem(a, b, c, i, JT, L);
          }
          sm = sm / sqrt(1.0 + (a == b)) / sqrt(1.0 + (c == d));
          GX(RES, q, g, IX, JX) = sm;
          if (SQS[g]) GX(RES, q, g, JX, IX) = sm * RES->herm;
        }
      }
    }
  }
}

/* commutator_220, commutators.f90:399-432 */
double or_comm220(const or_op *L, const or_op *R) {
  const or_basis *B = L->B;
  if (R->herm * L->herm != -1) return 0.0;
  double sm = 0.0;
  for (int q = 0; q < B->nblocks; q++) {
    int nh = B->nsub[3 * q + 2], np = B->nsub[3 * q];
    double smx = 0.0;
    for (int IX = 0; IX < np; IX++)
      for (int JX = 0; JX < nh; JX++)
        smx = smx - GX(L, q, 2, IX, JX) * GX(R, q, 2, IX, JX) * R->herm;
    sm = sm + smx * (B->lamJ[q] + 1);
  }
  return sm * 2.0;
}

/* commutator_221, commutators.f90:435-578 (must follow 222_pp_hh) */
static double c221_sum(const or_op *w1, const or_op *w2, int ik, int jk, int pm, int anti) {
  const or_basis *B = w1->B;
  int A = B->nhole, P = B->npart, ji = B->jj[ik];
  double sm = 0.0;
  for (int c = 0; c < A; c++) {
    int ck = B->holes[c], jc = B->jj[ck];
    for (int JT = abs(jc - ji); JT <= jc + ji; JT += 2) {
      if (anti) sm = sm + (or_v_elem(ck, ik, ck, jk, JT, w1) - pm * or_v_elem(ck, jk, ck, ik, JT, w1)) * (JT + 1);
      else sm = sm + or_v_elem(ck, ik, ck, jk, JT, w1) * (JT + 1);
    }
  }
  for (int c = 0; c < P; c++) {
    int ck = B->parts[c], jc = B->jj[ck];
    for (int JT = abs(jc - ji); JT <= jc + ji; JT += 2) {
      if (anti) sm = sm + (or_v_elem(ck, ik, ck, jk, JT, w2) - pm * or_v_elem(ck, jk, ck, ik, JT, w2)) * (JT + 1);
      else sm = sm + or_v_elem(ck, ik, ck, jk, JT, w2) * (JT + 1);
    }
  }
  return sm;
}
void or_comm221(const or_op *L, const or_op *R, or_op *RES, const or_op *w1, const or_op *w2) {
  const or_basis *B = L->B;
  int A = B->nhole, P = B->npart, pm = R->herm * L->herm;
  for (int i = 0; i < A; i++)
    for (int j = i; j < A; j++) {
      int ik = B->holes[i], jk = B->holes[j];
      if (!same_ljt(B, ik, jk)) continue;
      double sm = c221_sum(w1, w2, ik, jk, pm, 1);
      RES->fhh[i + (size_t)j * A] = RES->fhh[i + (size_t)j * A] + sm / (B->jj[ik] + 1.0);
      RES->fhh[j + (size_t)i * A] = RES->fhh[i + (size_t)j * A] * RES->herm;
    }
  for (int i = 0; i < P; i++)
    for (int j = i; j < P; j++) {
      int ik = B->parts[i], jk = B->parts[j];
      if (!same_ljt(B, ik, jk)) continue;
      double sm = c221_sum(w1, w2, ik, jk, pm, 1);
      RES->fpp[i + (size_t)j * P] = RES->fpp[i + (size_t)j * P] + sm / (B->jj[ik] + 1.0);
      RES->fpp[j + (size_t)i * P] = RES->fpp[i + (size_t)j * P] * RES->herm;
    }
  for (int i = 0; i < P; i++)
    for (int j = 0; j < A; j++) {
      int ik = B->parts[i], jk = B->holes[j];
      if (!same_ljt(B, ik, jk)) continue;
      double sm = c221_sum(w1, w2, ik, jk, pm, 0);
      RES->fph[i + (size_t)j * P] = RES->fph[i + (size_t)j * P] + sm / (B->jj[ik] + 1.0);
    }
}

/* commutator_222_pp_hh, commutators.f90:581-768 (L%pphh_ph is false on this path) */
void or_comm222_pp_hh(const or_op *L, const or_op *R, or_op *RES, or_op *w1, or_op *w2) {
  const or_basis *B = L->B;
  int pm = R->herm * L->herm;
  double hL = L->herm, hR = R->herm;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
  for (int q = 0; q < B->nblocks; q++) {
    int nh = B->nsub[3 * q + 2], np = B->nsub[3 * q], nb = B->nsub[3 * q + 1];
    double *L1 = L->gam[6 * q], *L2 = L->gam[6 * q + 1], *L3 = L->gam[6 * q + 2];
    double *L4 = L->gam[6 * q + 3], *L5 = L->gam[6 * q + 4], *L6 = L->gam[6 * q + 5];
    double *R1 = R->gam[6 * q], *R2 = R->gam[6 * q + 1], *R3 = R->gam[6 * q + 2];
    double *R5 = R->gam[6 * q + 4], *R6 = R->gam[6 * q + 5];
    double *a1 = w1->gam[6 * q], *a2 = w1->gam[6 * q + 1], *a3 = w1->gam[6 * q + 2];
    double *a4 = w1->gam[6 * q + 3], *a5 = w1->gam[6 * q + 4], *a6 = w1->gam[6 * q + 5];
    double *b1 = w2->gam[6 * q], *b2 = w2->gam[6 * q + 1], *b3 = w2->gam[6 * q + 2];
    double *b4 = w2->gam[6 * q + 3], *b5 = w2->gam[6 * q + 4], *b6 = w2->gam[6 * q + 5];
    (void)L4;
    for (int g = 0; g < 6; g++) {
      size_t sz = (size_t)grows(B, q, g) * gcols(B, q, g);
      memset(w1->gam[6 * q + g], 0, sz * 8);
      memset(w2->gam[6 * q + g], 0, sz * 8);
    }
    if (np != 0) or_dgemm('N', 'N', np, np, np, 1.0, L1, np, R1, np, 0.0, a1, np);
    if (nh != 0) or_dgemm('N', 'N', nh, nh, nh, 1.0, L5, nh, R5, nh, 0.0, b5, nh);
    if (np * nh != 0) {
      or_dgemm('T', 'N', nh, nh, np, 1.0, L3, np, R3, np, 0.0, a5, nh);
      for (size_t i = 0; i < (size_t)nh * nh; i++) a5[i] = a5[i] * hL;
      or_dgemm('N', 'T', np, np, nh, 1.0, L3, np, R3, np, 0.0, b1, np);
      for (size_t i = 0; i < (size_t)np * np; i++) b1[i] = b1[i] * hR;
      or_dgemm('N', 'N', np, nh, np, -1.0, R1, np, L3, np, 0.0, a3, np);
      or_dgemm('N', 'N', np, nh, np, 1.0, L1, np, R3, np, 1.0, a3, np);
      or_dgemm('N', 'N', np, nh, nh, 1.0, R3, np, L5, nh, 0.0, b3, np);
      or_dgemm('N', 'N', np, nh, nh, 1.0, L3, np, R5, nh, -1.0, b3, np);
    }
    /* Vpppp, Vhhhh, Vpphh */
    double *X1 = RES->gam[6 * q], *X5 = RES->gam[6 * q + 4], *X3 = RES->gam[6 * q + 2];
    for (int j = 0; j < np; j++)
      for (int i = 0; i < np; i++) {
        size_t ij = i + (size_t)j * np, ji = j + (size_t)i * np;
        X1[ij] = X1[ij] + a1[ij] - b1[ij] - (a1[ji] - b1[ji]) * pm;
      }
    for (int j = 0; j < nh; j++)
      for (int i = 0; i < nh; i++) {
        size_t ij = i + (size_t)j * nh, ji = j + (size_t)i * nh;
        X5[ij] = X5[ij] + a5[ij] - b5[ij] - (a5[ji] - b5[ji]) * pm;
      }
    for (size_t i = 0; i < (size_t)np * nh; i++) X3[i] = X3[i] + a3[i] - b3[i];
    if (nb * nh != 0) {
      or_dgemm('N', 'T', nb, nb, nh, 1.0, L6, nb, R6, nb, 0.0, b4, nb);
      for (size_t i = 0; i < (size_t)nb * nb; i++) b4[i] = b4[i] * hR;
    }
    if (nb * np != 0) {
      or_dgemm('T', 'N', nb, nb, np, 1.0, L2, np, R2, np, 0.0, a4, nb);
      for (size_t i = 0; i < (size_t)nb * nb; i++) a4[i] = a4[i] * hL;
    }
    double *X4 = RES->gam[6 * q + 3];
    for (int j = 0; j < nb; j++)
      for (int i = 0; i < nb; i++) {
        size_t ij = i + (size_t)j * nb, ji = j + (size_t)i * nb;
        X4[ij] = X4[ij] + a4[ij] - b4[ij] - (a4[ji] - b4[ji]) * pm;
      }
    if (nb * nh != 0) {
      or_dgemm('N', 'N', nb, nh, nh, 1.0, R6, nb, L5, nh, 0.0, b6, nb);
      or_dgemm('N', 'N', nb, nh, nh, 1.0, L6, nb, R5, nh, -1.0, b6, nb);
    }
    if (nb * np != 0) {
      or_dgemm('N', 'N', np, nb, np, 1.0, R1, np, L2, np, 0.0, a2, np);
      or_dgemm('N', 'N', np, nb, np, 1.0, L1, np, R2, np, -1.0, a2, np);
    }
    if (nb * np * nh != 0) {
      or_dgemm('T', 'N', nb, nh, np, -1.0 * hR, R2, np, L3, np, 0.0, a6, nb);
      or_dgemm('T', 'N', nb, nh, np, hL, L2, np, R3, np, 1.0, a6, nb);
      or_dgemm('N', 'T', np, nb, nh, 1.0, R3, np, L6, nb, 0.0, b2, np);
      or_dgemm('N', 'T', np, nb, nh, hR, L3, np, R6, nb, -1.0 * hL, b2, np);
    }
    double *X6 = RES->gam[6 * q + 5], *X2 = RES->gam[6 * q + 1];
    for (size_t i = 0; i < (size_t)nb * nh; i++) X6[i] = X6[i] + a6[i] - b6[i];
    for (size_t i = 0; i < (size_t)np * nb; i++) X2[i] = X2[i] + a2[i] - b2[i];
  }
}

/* ------------------------------------------------------------------ */
/* Pandya storage: allocate_CCMAT/dup_CC/allocate_CC_wkspc
 * (cross_coupled.f90:61-177, 863-914, 1075-1096) */
or_cc *or_cc_create(const or_basis *B) {
  or_cc *cc = (or_cc *)xcalloc(1, sizeof(or_cc));
  cc->B = B;
  size_t tot = 0, wtot = 0;
  for (int q = 0; q < B->ncc; q++) {
    tot += (size_t)B->ccrlen[q] * B->ccnph[q];
    wtot += (size_t)B->ccrlen[q] * B->ccrlen[q];
  }
  cc->buf = (double *)xcalloc(tot, 8);
  cc->X = (double **)xcalloc(B->ncc, sizeof(double *));
  cc->W = (double **)xcalloc(B->ncc, sizeof(double *));
  cc->wbuf = NULL;
  size_t off = 0;
  for (int q = 0; q < B->ncc; q++) {
    cc->X[q] = cc->buf + off;
    off += (size_t)B->ccrlen[q] * B->ccnph[q];
  }
  (void)wtot;
  return cc;
}
void or_cc_free(or_cc *cc) {
  if (!cc) return;
  free(cc->buf); free(cc->X); free(cc->W); free(cc->wbuf);
  free(cc);
}
const double *or_cc_block(const or_cc *cc, int q) { return cc->X[q]; }

/* calculate_cross_coupled, cross_coupled.f90:1190-1300 */
void or_calc_cross_coupled(const or_op *HS, or_cc *CC) {
  const or_basis *B = HS->B;
  int A = B->nhole, N = B->nsp, JTM = B->jtot_max;
  CC->herm = HS->herm;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
  for (int q1 = 0; q1 < B->ncc; q1++) {
    int JC = (q1 % (JTM + 1)) * 2;
    int PAR = q1 / (2 * JTM + 2);
    int TZ = (q1 % (2 * JTM + 2)) / (JTM + 1);
    int rlen = B->ccrlen[q1];
    memset(CC->X[q1], 0, (size_t)rlen * B->ccnph[q1] * 8);
    for (int hg = 0; hg < A; hg++)
      for (int pg = 0; pg < N - A; pg++) {
        int h = B->holes[hg], p = B->parts[pg];
        int jp = B->jj[p], jh = B->jj[h], lp = B->l[p], lh = B->l[h], tp = B->tz[p], th = B->tz[h];
        if (!triangle(jp, jh, JC)) continue;
        if ((lp + lh) % 2 != PAR) continue;
        if (abs(tp - th) / 2 != TZ) continue;
        int NB = B->cccol[((size_t)q1 * N + p) * N + h];
        for (int a = 0; a < N; a++)
          for (int b = 0; b < N; b++) {
            int ja = B->jj[a], jb = B->jj[b], la = B->l[a], lb = B->l[b], ta = B->tz[a], tb = B->tz[b];
            if (!triangle(ja, jb, JC)) continue;
            if ((la + lb) % 2 != PAR) continue;
            if (abs(ta - tb) / 2 != TZ) continue;
            int RI = B->ccrow[((size_t)q1 * N + a) * N + b];
            if (((la + lh) % 2 == (lb + lp) % 2) && (ta + th == tb + tp)) {
              int Jmin = abs(jp - jb), t2 = abs(ja - jh);
              if (t2 > Jmin) Jmin = t2;
              int Jmax = jp + jb;
              if (ja + jh < Jmax) Jmax = ja + jh;
              double sm = 0.0;
              for (int JT = Jmin; JT <= Jmax; JT += 2)
                sm = sm + ipow_m1(JT / 2) * (JT + 1) * or_sixj(B, jp, jh, JC, ja, jb, JT) *
                              or_v_elem(h, a, p, b, JT, HS);
              CC->X[q1][(size_t)RI + (size_t)NB * rlen] =
                  sm * ipow_m1((jh + jb + JC) / 2) * 1.0 * sqrt(JC + 1.0);
            }
          }
      }
  }
}

/* commutator_222_ph, commutators.f90:771-925 */
void or_comm222_ph(const or_cc *LCC, const or_cc *RCC, or_op *RES, or_cc *WCC) {
  const or_basis *B = RES->B;
  int N = B->nsp, JTM = B->jtot_max;
  if (!WCC->W[0]) {
    size_t wtot = 0;
    for (int q = 0; q < B->ncc; q++) wtot += (size_t)B->ccrlen[q] * B->ccrlen[q];
    WCC->wbuf = (double *)xcalloc(wtot, 8);
    size_t off = 0;
    for (int q = 0; q < B->ncc; q++) {
      WCC->W[q] = WCC->wbuf + off;
      off += (size_t)B->ccrlen[q] * B->ccrlen[q];
    }
  }
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
  for (int q = 0; q < B->ncc; q++) {
    int nb = B->ccnph[q], r = B->ccrlen[q];
    if (nb * r == 0) continue;
    or_dgemm('N', 'T', r, r, nb, 1.0, LCC->X[q], r, RCC->X[q], r, 0.0, WCC->W[q], r);
  }
  double hL = LCC->herm, hR = RCC->herm;
#define WX(qx, i, j) WCC->W[qx][(size_t)(i) + (size_t)(j) * B->ccrlen[qx]]
#define RV(i, j, qx) B->ccrow[((size_t)(qx) * N + (i)) * N + (j)]
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads())
  for (int q = 0; q < B->nblocks; q++) {
    int Jtot = B->lamJ[q];
    for (int g = 0; g < 6; g++) {
      int n1 = grows(B, q, g), n2 = gcols(B, q, g);
      if (n1 * n2 == 0) continue;
      const int *qa = B->qn[3 * q + SEA1[g] - 1], *qb = B->qn[3 * q + SEA2[g] - 1];
      for (int IX = 0; IX < n1; IX++) {
        double pre = 1.0;
        int i = qa[2 * IX], j = qa[2 * IX + 1];
        if (i == j) pre = .70710678118;
        int ji = B->jj[i], jj = B->jj[j], li = B->l[i], lj = B->l[j], ti = B->tz[i], tj = B->tz[j];
        for (int JX = SQS[g] ? IX : 0; JX < n2; JX++) {
          double pre2 = 1.0;
          int k = qb[2 * JX], l = qb[2 * JX + 1];
          if (k == l) pre2 = .70710678118;
          int jk = B->jj[k], jl = B->jj[l], lk = B->l[k], ll = B->l[l], tk = B->tz[k], tl = B->tz[l];
          double sm = 0.0;
          int jmin = abs(jj - jl), t2 = abs(ji - jk);
          if (t2 > jmin) jmin = t2;
          int jmax = jj + jl;
          if (ji + jk < jmax) jmax = ji + jk;
          int Tz = abs(ti - tk) / 2;
          if (abs(tl - tj) != Tz * 2) continue;
          int PAR = (li + lk) % 2;
          if ((ll + lj) % 2 != PAR) continue;
          int phase = ipow_m1((ji + jj + jk + jl) / 2);
          for (int JP = jmin; JP <= jmax; JP += 2) {
            int qx = JP / 2 + Tz * (JTM + 1) + 2 * PAR * (JTM + 1);
            int rjl = RV(j, l, qx), rik = RV(i, k, qx), gjl = RV(l, j, qx), gik = RV(k, i, qx);
            sm = sm - ipow_m1((jl + jk + Jtot) / 2) *
                          ((WX(qx, gik, rjl) + phase * WX(qx, gjl, rik)) * hL -
                           (phase * WX(qx, rik, gjl) + WX(qx, rjl, gik)) * hR) *
                          or_sixj(B, ji, jj, Jtot, jl, jk, JP);
          }
          Tz = abs(ti - tl) / 2;
          if (abs(tk - tj) != Tz * 2) continue;
          PAR = (li + ll) % 2;
          if ((lk + lj) % 2 != PAR) continue;
          jmin = abs(ji - jl);
          t2 = abs(jj - jk);
          if (t2 > jmin) jmin = t2;
          jmax = ji + jl;
          if (jj + jk < jmax) jmax = jj + jk;
          phase = ipow_m1((ji + jj + jk + jl) / 2);
          for (int JP = jmin; JP <= jmax; JP += 2) {
            int qx = JP / 2 + Tz * (JTM + 1) + 2 * PAR * (JTM + 1);
            int ril = RV(i, l, qx), rjk = RV(j, k, qx), gil = RV(l, i, qx), gjk = RV(k, j, qx);
            sm = sm + ((WX(qx, gil, rjk) + phase * WX(qx, gjk, ril)) * hL -
                       (phase * WX(qx, ril, gjk) + WX(qx, rjk, gil)) * hR) *
                          or_sixj(B, ji, jj, Jtot, jk, jl, JP);
          }
          GX(RES, q, g, IX, JX) = GX(RES, q, g, IX, JX) + sm * pre * pre2;
          if (SQS[g]) GX(RES, q, g, JX, IX) = GX(RES, q, g, IX, JX) * RES->herm;
        }
      }
    }
  }
#undef WX
#undef RV
}

/* the reference call sequence of one commutator (IMSRG_MAGNUS.f90:207-219) */
void or_commutator(const or_op *L, const or_op *R, or_op *RES, int with_e0) {
  const or_basis *B = L->B;
  or_cc *LCC = or_cc_create(B), *RCC = or_cc_create(B), *WCC = or_cc_create(B);
  or_op *w1 = or_op_create(B, 1), *w2 = or_op_create(B, 1);
  or_calc_cross_coupled(R, RCC);
  or_calc_cross_coupled(L, LCC);
  if (with_e0) RES->E0 = or_comm110(L, R) + or_comm220(L, R);
  or_comm111(L, R, RES);
  or_comm121(L, R, RES);
  or_comm122(L, R, RES);
  or_comm222_pp_hh(L, R, RES, w1, w2);
  or_comm221(L, R, RES, w1, w2);
  or_comm222_ph(LCC, RCC, RES, WCC);
  or_cc_free(LCC); or_cc_free(RCC); or_cc_free(WCC);
  or_op_free(w1); or_op_free(w2);
}

/* ------------------------------------------------------------------ */
/* operator algebra, basic_IMSRG.f90:2942-2969, 3450-3562, 3921-3952, 4030-4064 */
void or_add(const or_op *A, double ax, const or_op *Bo, double bx, or_op *C) {
  const or_basis *B = A->B;
  size_t A2 = (size_t)B->nhole * B->nhole, P2 = (size_t)B->npart * B->npart,
         PA = (size_t)B->npart * B->nhole;
  C->E0 = A->E0 * ax + Bo->E0 * bx;
  for (size_t i = 0; i < A2; i++) C->fhh[i] = A->fhh[i] * ax + Bo->fhh[i] * bx;
  for (size_t i = 0; i < P2; i++) C->fpp[i] = A->fpp[i] * ax + Bo->fpp[i] * bx;
  for (size_t i = 0; i < PA; i++) C->fph[i] = A->fph[i] * ax + Bo->fph[i] * bx;
  for (size_t i = 0; i < A->nbuf; i++) C->buf[i] = A->buf[i] * ax + Bo->buf[i] * bx;
}
void or_copy(const or_op *H, or_op *op) {
  const or_basis *B = H->B;
  op->herm = H->herm;
  op->E0 = H->E0;
  memcpy(op->fhh, H->fhh, (size_t)B->nhole * B->nhole * 8);
  memcpy(op->fpp, H->fpp, (size_t)B->npart * B->npart * 8);
  memcpy(op->fph, H->fph, (size_t)B->npart * B->nhole * 8);
  memcpy(op->buf, H->buf, H->nbuf * 8);
}
void or_clear(or_op *C) {
  const or_basis *B = C->B;
  C->E0 = 0.0;
  memset(C->fhh, 0, (size_t)B->nhole * B->nhole * 8);
  memset(C->fpp, 0, (size_t)B->npart * B->npart * 8);
  memset(C->fph, 0, (size_t)B->npart * B->nhole * 8);
  memset(C->buf, 0, C->nbuf * 8);
}
void or_scale(or_op *A, double ax) {
  const or_basis *B = A->B;
  A->E0 = A->E0 * ax;
  for (size_t i = 0; i < (size_t)B->nhole * B->nhole; i++) A->fhh[i] *= ax;
  for (size_t i = 0; i < (size_t)B->npart * B->npart; i++) A->fpp[i] *= ax;
  for (size_t i = 0; i < (size_t)B->npart * B->nhole; i++) A->fph[i] *= ax;
  for (size_t i = 0; i < A->nbuf; i++) A->buf[i] *= ax;
}
static double sumsq(const double *x, size_t n) {
  double s = 0.0;
  for (size_t i = 0; i < n; i++) s += x[i] * x[i];
  return s;
}
double or_frob_norm(const or_op *op) {
  const or_basis *B = op->B;
  double sm = sumsq(op->fhh, (size_t)B->nhole * B->nhole) +
              2 * sumsq(op->fph, (size_t)B->npart * B->nhole) +
              sumsq(op->fpp, (size_t)B->npart * B->npart);
  for (int q = 0; q < B->nblocks; q++)
    for (int g = 0; g < 6; g++) {
      size_t n = (size_t)grows(B, q, g) * gcols(B, q, g);
      double s = sumsq(op->gam[6 * q + g], n);
      sm = sm + s;
      if (!SQS[g]) sm = sm + s;
    }
  return sqrt(sm);
}
double or_mbpt2(const or_op *H) {
  const or_basis *B = H->B;
  double sm = 0.0;
  for (int q = 0; q < B->nblocks; q++) {
    double sj = 0.0;
    int np = B->nsub[3 * q], nh = B->nsub[3 * q + 2];
    const int *pp = B->qn[3 * q], *hh = B->qn[3 * q + 2];
    for (int II = 0; II < np; II++) {
      int i = pp[2 * II], j = pp[2 * II + 1];
      double fi = or_f_elem(i, i, H), fj = or_f_elem(j, j, H);
      for (int JJ = 0; JJ < nh; JJ++) {
        int k = hh[2 * JJ], l = hh[2 * JJ + 1];
        double fk = or_f_elem(k, k, H), fl = or_f_elem(l, l, H);
        double eden = fk + fl - fi - fj;
        double x = GX(H, q, 2, II, JJ);
        sj = sj + x * x / eden;
      }
    }
    sm = sm + (B->lamJ[q] + 1.0) * sj;
  }
  return sm;
}

/* normal_order, basic_IMSRG.f90:2375-2506 */
void or_normal_order(or_op *H) {
  const or_basis *B = H->B;
  int A = B->nhole, P = B->npart, N = B->nsp;
  or_op *T = or_op_create(B, H->herm);
  or_copy(H, T);
  double sm = 0.0;
  for (int i = 0; i < N; i++) sm = sm + or_f_elem(i, i, T) * B->con[i] * (B->jj[i] + 1);
  sm = sm * 2.0;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++) {
      if (B->con[i] * B->con[j] == 0) continue;
      for (int JT = abs(B->jj[i] - B->jj[j]); JT <= B->jj[i] + B->jj[j]; JT += 2)
        sm = sm + or_v_elem(i, j, i, j, JT, T) * (JT + 1);
    }
  H->E0 = sm * 0.5;
#define NO_SUM(ax, bx)                                                                  \
  double s2 = 0.0;                                                                      \
  for (int ii = 0; ii < A; ii++) {                                                      \
    int ix = B->holes[ii];                                                              \
    for (int JT = abs(B->jj[ix] - B->jj[ax]); JT <= B->jj[ix] + B->jj[ax]; JT += 2)     \
      s2 = s2 + (JT + 1.0) / (B->jj[ax] + 1) * or_v_elem(ax, ix, bx, ix, JT, T);        \
  }
  for (int a = 0; a < A; a++)
    for (int b = a; b < A; b++) {
      int ax = B->holes[a], bx = B->holes[b];
      if (!same_ljt(B, ax, bx)) continue;
      NO_SUM(ax, bx)
      H->fhh[a + (size_t)b * A] = or_f_elem(ax, bx, T) + s2;
      H->fhh[b + (size_t)a * A] = H->fhh[a + (size_t)b * A];
    }
  for (int a = 0; a < P; a++)
    for (int b = 0; b < A; b++) {
      int ax = B->parts[a], bx = B->holes[b];
      if (!same_ljt(B, ax, bx)) continue;
      NO_SUM(ax, bx)
      H->fph[a + (size_t)b * P] = or_f_elem(ax, bx, T) + s2;
    }
  for (int a = 0; a < P; a++)
    for (int b = a; b < P; b++) {
      int ax = B->parts[a], bx = B->parts[b];
      if (!same_ljt(B, ax, bx)) continue;
      NO_SUM(ax, bx)
      H->fpp[a + (size_t)b * P] = or_f_elem(ax, bx, T) + s2;
      H->fpp[b + (size_t)a * P] = H->fpp[a + (size_t)b * P];
    }
#undef NO_SUM
  or_op_free(T);
}

/* ------------------------------------------------------------------ */
/* generators, generators.f90:4-374 */
enum { GEN_WHITE = 0, GEN_ATAN = 1 };
static void gen_common(const or_op *Hden, const or_op *Hnum, or_op *ETA, int kind) {
  const or_basis *B = Hden->B;
  int A = B->nhole, P = B->npart;
  ETA->herm = -1;
  memset(ETA->fpp, 0, (size_t)P * P * 8);
  memset(ETA->fhh, 0, (size_t)A * A * 8);
  for (int a = 0; a < P; a++) {
    int ak = B->parts[a];
    for (int i = 0; i < A; i++) {
      int ik = B->holes[i];
      if (!same_ljt(B, ak, ik)) continue;
      int ji = B->jj[ik];
      double Eden = 0.0;
      for (int JT = 0; JT <= 2 * ji; JT += 2) Eden = Eden - (JT + 1) * or_v_elem(ak, ik, ak, ik, JT, Hden);
      Eden = Eden / (ji + 1.0) / (ji + 1.0);
      Eden = Eden + Hden->fpp[a + (size_t)a * P] - Hden->fhh[i + (size_t)i * A];
      double h = Hnum->fph[a + (size_t)i * P];
      ETA->fph[a + (size_t)i * P] = (kind == GEN_WHITE) ? h / Eden : atan(2 * h / Eden) * 0.5;
    }
  }
  memset(ETA->buf, 0, ETA->nbuf * 8);
  for (int q = 0; q < B->nblocks; q++) {
    int np = B->nsub[3 * q], nh = B->nsub[3 * q + 2];
    const int *pp = B->qn[3 * q], *hh = B->qn[3 * q + 2];
    for (int IX = 0; IX < np; IX++) {
      int a = pp[2 * IX], b = pp[2 * IX + 1], ja = B->jj[a], jb = B->jj[b];
      for (int JX = 0; JX < nh; JX++) {
        int i = hh[2 * JX], j = hh[2 * JX + 1], ji = B->jj[i], jj = B->jj[j];
        double Eden = 0.0;
        Eden = Eden + twobody_monopole(a, b, ja, jb, Hden);
        Eden = Eden + twobody_monopole(i, j, ji, jj, Hden);
        Eden = Eden - twobody_monopole(a, i, ja, ji, Hden);
        Eden = Eden - twobody_monopole(a, j, ja, jj, Hden);
        Eden = Eden - twobody_monopole(i, b, ji, jb, Hden);
        Eden = Eden - twobody_monopole(j, b, jj, jb, Hden);
        Eden = Eden + or_f_elem(a, a, Hden) + or_f_elem(b, b, Hden) - or_f_elem(i, i, Hden) -
               or_f_elem(j, j, Hden);
        double h = GX(Hnum, q, 2, IX, JX);
        GX(ETA, q, 2, IX, JX) = (kind == GEN_WHITE) ? h / Eden : 0.5 * atan(2 * h / Eden);
      }
    }
  }
}
void or_build_white(const or_op *H, or_op *ETA) { gen_common(H, H, ETA, GEN_WHITE); }
void or_build_atan(const or_op *H, or_op *ETA) { gen_common(H, H, ETA, GEN_ATAN); }
void or_build_atanxx(const or_op *Hold, const or_op *H, or_op *ETA) { gen_common(Hold, H, ETA, GEN_ATAN); }
/* build_gs_w2, generators.f90:109-160 */
void or_build_w2(const or_op *H, or_op *ETA) {
  const or_basis *B = H->B;
  ETA->herm = -1;
  or_op *ETA1 = or_op_create(B, -1), *H2 = or_op_create(B, 1), *ETA2 = or_op_create(B, -1);
  or_build_atan(H, ETA1);
  or_commutator(ETA1, H, H2, 1);
  or_build_atanxx(H, H2, ETA2);
  or_add(ETA1, 1.5, ETA2, 0.5, ETA);
  or_op_free(ETA1); or_op_free(H2); or_op_free(ETA2);
}
void or_build_generator(int kind, const or_op *H, or_op *ETA) {
  if (kind == 0) or_build_white(H, ETA);
  else if (kind == 1) or_build_atan(H, ETA);
  else or_build_w2(H, ETA);
}

/* ------------------------------------------------------------------ */
/* BCH_EXPAND, IMSRG_MAGNUS.f90:149-234 */
int or_bch_expand(or_op *HS, const or_op *G, const or_op *H) {
  const or_basis *B = H->B;
  const double conv = (double)1e-6f;
  or_op *INT1 = or_op_create(B, 1), *INT2 = or_op_create(B, 1), *AD = or_op_create(B, 1);
  or_op *w1 = or_op_create(B, 1), *w2 = or_op_create(B, 1);
  or_cc *ADCC = or_cc_create(B), *GCC = or_cc_create(B), *WCC = or_cc_create(B);
  double coef = 1.0;
  or_copy(H, HS);
  or_copy(HS, INT2);
  int iw, ncomm = 0;
  for (iw = 2; iw <= 40; iw++) {
    coef = coef / (iw - 1.0);
    or_copy(HS, INT1);
    or_copy(INT2, AD);
    or_clear(INT2);
    or_calc_cross_coupled(AD, ADCC);
    or_calc_cross_coupled(G, GCC);
    INT2->E0 = or_comm110(G, AD) + or_comm220(G, AD);
    or_comm111(G, AD, INT2);
    or_comm121(G, AD, INT2);
    or_comm122(G, AD, INT2);
    or_comm222_pp_hh(G, AD, INT2, w1, w2);
    or_comm221(G, AD, INT2, w1, w2);
    or_comm222_ph(GCC, ADCC, INT2, WCC);
    ncomm++;
    or_add(INT1, 1.0, INT2, coef, HS);
    double adv = or_frob_norm(INT2) * coef;
    if (adv < conv) break;
  }
  or_op_free(INT1); or_op_free(INT2); or_op_free(AD); or_op_free(w1); or_op_free(w2);
  or_cc_free(ADCC); or_cc_free(GCC); or_cc_free(WCC);
  if (iw > 40) return -1; /* STOP 'clipping BCH' */
  return ncomm;
}

/* MAGNUS_EXPAND, IMSRG_MAGNUS.f90:622-693 */
int or_magnus_expand(or_op *DG, const or_op *G) {
  const or_basis *B = G->B;
  const double conv = (double)1e-6f;
  const double cof[7] = {1.0, -0.5, .0833333333, 0.0, -0.00138888888, 0.0, 3.306878e-5};
  double advals[7] = {0, 0, 0, 0, 0, 0, 0};
  or_op *INT1 = or_op_create(B, -1), *INT2 = or_op_create(B, -1), *AD = or_op_create(B, -1);
  or_op *w1 = or_op_create(B, 1), *w2 = or_op_create(B, 1);
  or_cc *ADCC = or_cc_create(B), *GCC = or_cc_create(B), *WCC = or_cc_create(B);
  int ncomm = 0;
  or_copy(DG, INT2);
  advals[0] = or_frob_norm(INT2);
  double fullnorm = or_frob_norm(G);
  if (!(fullnorm < (double)1e-9f)) {
    int q = 1;
    for (int i = 2; i <= 7; i++) {
      or_copy(DG, INT1);
      or_copy(INT2, AD);
      double adnorm = advals[i - q - 1];
      if (fabs(cof[i - 1]) > (double)1e-6f) q = 1;
      else q = 2;
      if (fabs(adnorm / fullnorm) < conv) break;
      or_calc_cross_coupled(AD, ADCC);
      or_calc_cross_coupled(G, GCC);
      or_comm111(G, AD, INT2);
      or_comm121(G, AD, INT2);
      or_comm122(G, AD, INT2);
      or_comm222_pp_hh(G, AD, INT2, w1, w2);
      or_comm221(G, AD, INT2, w1, w2);
      or_comm222_ph(GCC, ADCC, INT2, WCC);
      ncomm++;
      or_add(INT1, 1.0, INT2, cof[i - 1], DG);
      advals[i - 1] = or_frob_norm(INT2) * fabs(cof[i - 1]);
    }
  }
  or_op_free(INT1); or_op_free(INT2); or_op_free(AD); or_op_free(w1); or_op_free(w2);
  or_cc_free(ADCC); or_cc_free(GCC); or_cc_free(WCC);
  return ncomm;
}

/* magnus_decouple, IMSRG_MAGNUS.f90:15-128 (no checkpointing) */
int or_magnus_flow(or_op *HS, or_op *G, int gen_kind, double ds, int max_steps, double *s_out,
                   double *e0_out, double *emp2_out, double *crit_out, int *nmag_out, int *nbch_out) {
  const or_basis *B = HS->B;
  or_op *H = or_op_create(B, 1), *DG = or_op_create(B, -1);
  or_clear(G);
  G->herm = -1;
  DG->herm = -1;
  or_build_generator(gen_kind, HS, DG);
  or_copy(HS, H);
  double s = 0.0;
  double E_mbpt2 = or_mbpt2(HS);
  double crit = fabs(E_mbpt2);
  int steps = 0;
  while (crit > (double)1e-6f) {
    if (steps >= max_steps) break;
    steps++;
    int nm = or_magnus_expand(DG, G);
    or_add(G, 1.0, DG, ds, G); /* euler_step :756-766 */
    s = s + ds;
    int nb = or_bch_expand(HS, G, H);
    if (nb < 0) {
      or_op_free(H); or_op_free(DG);
      return -steps;
    }
    or_build_generator(gen_kind, HS, DG);
    E_mbpt2 = or_mbpt2(HS);
    crit = fabs(E_mbpt2);
    int k = steps - 1;
    if (s_out) s_out[k] = s;
    if (e0_out) e0_out[k] = HS->E0;
    if (emp2_out) emp2_out[k] = HS->E0 + E_mbpt2;
    if (crit_out) crit_out[k] = crit;
    if (nmag_out) nmag_out[k] = nm;
    if (nbch_out) nbch_out[k] = nb;
  }
  or_op_free(H); or_op_free(DG);
  return steps;
}

/* ------------------------------------------------------------------ */
/* brute-force J-coupled sums, brute_force_testing.f90:1765-2076 */
double or_brute_1body(const or_op *AA, const or_op *BB, int a, int b) {
  const or_basis *B = AA->B;
  int N = B->nsp, JTM = B->jtot_max * 2, ja = B->jj[a], jb = B->jj[b];
  double sm = 0.0;
  for (int i = 0; i < N; i++)
    sm = sm + or_f_elem(a, i, AA) * or_f_elem(i, b, BB) - or_f_elem(a, i, BB) * or_f_elem(i, b, AA);
  if (ja == jb) {
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int Jt = 0; Jt <= JTM; Jt += 2)
          sm = sm + (B->con[i] - B->con[j]) * (Jt + 1.0) / (ja + 1.0) *
                        (or_f_elem(i, j, AA) * or_v_elem(j, a, i, b, Jt, BB) -
                         or_f_elem(i, j, BB) * or_v_elem(j, a, i, b, Jt, AA));
    for (int i = 0; i < N; i++)
      for (int j = 0; j < N; j++)
        for (int k = 0; k < N; k++) {
          int occ = B->con[i] * B->con[j] * (1 - B->con[k]) + (1 - B->con[i]) * (1 - B->con[j]) * B->con[k];
          if (!occ) continue;
          for (int Jt = 0; Jt <= JTM; Jt += 2)
            sm = sm + occ * (Jt + 1.0) / (ja + 1.0) *
                          (or_v_elem(a, k, i, j, Jt, AA) * or_v_elem(i, j, b, k, Jt, BB) -
                           or_v_elem(a, k, i, j, Jt, BB) * or_v_elem(i, j, b, k, Jt, AA)) / 2.0;
        }
  }
  return sm;
}
double or_brute_0body(const or_op *AA, const or_op *BB) {
  const or_basis *B = AA->B;
  int N = B->nsp, JTM = B->jtot_max * 2;
  double sm = 0.0;
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      sm = sm + (B->con[i] - B->con[j]) * (B->jj[i] + 1.0) * or_f_elem(i, j, AA) * or_f_elem(j, i, BB);
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < N; k++)
        for (int l = 0; l < N; l++) {
          int occ = B->con[i] * B->con[j] * (1 - B->con[k]) * (1 - B->con[l]) -
                    B->con[k] * B->con[l] * (1 - B->con[i]) * (1 - B->con[j]);
          if (!occ) continue;
          for (int Jt = 0; Jt <= JTM; Jt += 2)
            sm = sm + occ * (Jt + 1.0) * or_v_elem(i, j, k, l, Jt, AA) * or_v_elem(k, l, i, j, Jt, BB) / 4.0;
        }
  return sm;
}
double or_brute_2body(const or_op *AA, const or_op *BB, int a, int b, int c, int d, int Jtot) {
  const or_basis *B = AA->B;
  int N = B->nsp, JTM = B->jtot_max * 2;
  int ja = B->jj[a], jb = B->jj[b], jc = B->jj[c], jd = B->jj[d];
  double sm = 0.0;
  for (int i = 0; i < N; i++) {
    sm = sm + or_f_elem(a, i, AA) * or_v_elem(i, b, c, d, Jtot, BB) +
         or_f_elem(b, i, AA) * or_v_elem(a, i, c, d, Jtot, BB) -
         or_f_elem(i, c, AA) * or_v_elem(a, b, i, d, Jtot, BB) -
         or_f_elem(i, d, AA) * or_v_elem(a, b, c, i, Jtot, BB);
    sm = sm - or_f_elem(a, i, BB) * or_v_elem(i, b, c, d, Jtot, AA) -
         or_f_elem(b, i, BB) * or_v_elem(a, i, c, d, Jtot, AA) +
         or_f_elem(i, c, BB) * or_v_elem(a, b, i, d, Jtot, AA) +
         or_f_elem(i, d, BB) * or_v_elem(a, b, c, i, Jtot, AA);
  }
  for (int i = 0; i < N; i++)
    for (int j = 0; j < N; j++)
      sm = sm + 0.5 * (1 - B->con[i] - B->con[j]) *
                    (or_v_elem(a, b, i, j, Jtot, AA) * or_v_elem(i, j, c, d, Jtot, BB) -
                     or_v_elem(a, b, i, j, Jtot, BB) * or_v_elem(i, j, c, d, Jtot, AA));
  for (int i = 0; i < N; i++) {
    int ji = B->jj[i];
    for (int j = 0; j < N; j++) {
      int jj = B->jj[j];
      int dn = B->con[i] - B->con[j];
      if (dn == 0) continue;
      for (int J1 = 0; J1 <= JTM; J1 += 2)
        for (int J2 = 0; J2 <= JTM; J2 += 2) {
          sm = sm + dn * (ipow_m1((J1 + J2 + ja - jc) / 2) * (J1 + 1.0) * (J2 + 1.0) *
                              or_ninej(jj, J1, ja, J2, ji, jb, jc, jd, Jtot) *
                              or_v_elem(j, a, i, d, J1, AA) * or_v_elem(i, b, j, c, J2, BB) -
                          ipow_m1((J1 + J2 + jb - jc) / 2) * (J1 + 1.0) * (J2 + 1.0) *
                              or_ninej(jj, J1, jb, J2, ji, ja, jc, jd, Jtot) *
                              or_v_elem(j, b, i, d, J1, AA) * or_v_elem(i, a, j, c, J2, BB) *
                              ipow_m1((ja + jb - Jtot) / 2) -
                          ipow_m1((J1 + J2 + ja - jd) / 2) * (J1 + 1.0) * (J2 + 1.0) *
                              or_ninej(jj, J1, ja, J2, ji, jb, jd, jc, Jtot) *
                              or_v_elem(j, a, i, c, J1, AA) * or_v_elem(i, b, j, d, J2, BB) *
                              ipow_m1((jc + jd - Jtot) / 2) +
                          ipow_m1((J1 + J2 + jb - jd) / 2) * (J1 + 1.0) * (J2 + 1.0) *
                              or_ninej(jj, J1, jb, J2, ji, ja, jd, jc, Jtot) *
                              or_v_elem(j, b, i, c, J1, AA) * or_v_elem(i, a, j, d, J2, BB) *
                              ipow_m1((ja + jb + jc + jd) / 2));
        }
    }
  }
  return sm;
}
