/*
 * dbs_oracle.c -- CPU FP64 restatement of the Dirac B-spline (DBS) radial
 * core: the P- and Q-function bases, their moments, the diagonal-cell
 * quadrature and the R^k bands of the four integral families PPPP, QQQQ,
 * PQPQ and QPQP.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle); the product never links it.
 *
 * Reference routines followed (paths relative to the reference checkout):
 *   gauleg                 src/libs/ZCOM/gauss_points.f90:2-51
 *                          (the Numerical Recipes Gauss-Legendre routine)
 *   alloc_DBS_gauss        src/libs/DBS/DBS_gauss.f90:72-150 (gr, grm, grw)
 *   Def_pbsp / Def_qbsp    src/libs/DBS/DBS_gauss.f90:153-252 (tp, tq, pbsp, qbsp)
 *   moment_pp / moment_qq  src/libs/DBS/DBS_moments.f90:166-244
 *   triang_pppp/qqqq       src/libs/DBS/mrk_pppp.f90:95-179, mrk_qqqq.f90
 *   triang_pqpq/qpqp       src/libs/DBS/mrk_pqpq.f90:105-193, mrk_qpqp.f90
 *   mrk_pppp/qqqq/pqpq/qpqp src/libs/DBS/mrk_pppp.f90:2-77 and siblings
 *
 * A family f has the basis b1 on the r1 pair and b2 on the r2 pair:
 *   0 PPPP (P,P)   1 QQQQ (Q,Q)   2 PQPQ (P,Q)   3 QPQP (Q,P)
 * Layouts are the reference's (Fortran column-major):
 *   t(ns+ks); gr,grm,grw(nv,ks); pbsp(nv+1,ks,ksp), qbsp(nv+1,ks,ksq);
 *   moments rkm(ks*ks,nv) (first kb(kb+1)/2 rows: pairs i<=j packed);
 *   rkd(ks*ks,ks*ks,nv) (pairs of b1 x pairs of b2);
 *   rkb(ns,ns,ks,ks), rkb(i,j,ip,jp) = R^k(B1_i B1_i+ip-1 ; B2_j B2_j+jp-1).
 * No scaling law: every cell is computed (the DBS routines do not use it).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "dbs_oracle.h"

/* integer power by repeated squaring (Fortran real**integer) */
double orc_dpow(double x, int n) {
  if (n < 0) return 1.0 / orc_dpow(x, -n);
  double r = 1.0, b = x;
  while (n) {
    if (n & 1) r *= b;
    b *= b;
    n >>= 1;
  }
  return r;
}

/* Gauss-Legendre points and weights on [x1,x2] by Newton's method from the
 * Chebyshev-like first guess, converged to |dz| <= 1e-14 (gauleg). */
void orc_gauleg(double x1, double x2, double* x, double* w, int n) {
  const double eps = 1.0e-14;
  const int m = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; i++) {
    double z = cos(3.141592654 * (i - 0.25) / (n + 0.5)), z1, pp, p1, p2, p3;
    do {
      p1 = 1.0;
      p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (fabs(z - z1) > eps);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

/* de Boor recurrence: the kb B-splines of order kb on knots tb that are
 * nonzero at x, left = 1-based index of the knot interval's left end
 * (zbsplvb, src/libs/DBS/zbsplvd.f90:160-271). */
static void dbs_bvalues(const double* tb, int kb, int left, double x, double* b) {
  double dr[32], dl[32];
  b[0] = 1.0;
  for (int j = 1; j < kb; j++) {
    dr[j - 1] = tb[left + j - 1] - x;
    dl[j - 1] = x - tb[left - j];
    double saved = 0.0;
    for (int r = 0; r < j; r++) {
      const double term = b[r] / (dr[r] + dl[j - 1 - r]);
      b[r] = saved + dr[r] * term;
      saved = dl[j - 1 - r] * term;
    }
    b[j] = saved;
  }
}


void orc_dbs_free(OrcDbs* d) {
  if (!d) return;
  free(d->t); free(d->tp); free(d->tq); free(d->gx); free(d->gw);
  free(d->gr); free(d->grm); free(d->grw); free(d->pbsp); free(d->qbsp);
  free(d);
}

static double* knots_for(const double* t, int ks, int ns, int kb, int nsb) {
  /* tb(1:kb) = t(1); tb(kb+1:nsb) = t(ks+1:ns); tb(nsb+1:nsb+kb) = t(ns+1:ns+kb) */
  double* tb = (double*)malloc(sizeof(double) * (nsb + kb));
  for (int i = 0; i < kb; i++) tb[i] = t[0];
  for (int i = kb; i < nsb; i++) tb[i] = t[i - kb + ks];
  for (int i = 0; i < kb; i++) tb[nsb + i] = t[ns + i];
  return tb;
}

static double* values_for(const OrcDbs* d, const double* tb, int kb) {
  const int nv = d->nv, ks = d->ks;
  double* v = (double*)calloc((size_t)(nv + 1) * ks * kb, sizeof(double));
  double b[32];
  for (int m = 0; m < ks; m++)
    for (int iv = 0; iv < nv; iv++) {
      dbs_bvalues(tb, kb, iv + kb, d->gr[iv + nv * m], b);
      for (int ith = 0; ith < kb; ith++) v[iv + (nv + 1) * (m + ks * ith)] = b[ith];
    }
  return v;
}

OrcDbs* orc_dbs_new(int ks, int ns, int ksp, int ksq, const double* t) {
  OrcDbs* d = (OrcDbs*)calloc(1, sizeof(OrcDbs));
  const int nv = ns - ks + 1;
  d->ks = ks; d->ns = ns; d->nv = nv; d->ksp = ksp; d->ksq = ksq;
  d->nsp = nv + ksp - 1; d->nsq = nv + ksq - 1;
  d->t = (double*)malloc(sizeof(double) * (ns + ks));
  memcpy(d->t, t, sizeof(double) * (ns + ks));
  d->gx = (double*)malloc(sizeof(double) * ks);
  d->gw = (double*)malloc(sizeof(double) * ks);
  orc_gauleg(0.0, 1.0, d->gx, d->gw, ks);
  d->gr = (double*)malloc(sizeof(double) * nv * ks);
  d->grm = (double*)malloc(sizeof(double) * nv * ks);
  d->grw = (double*)malloc(sizeof(double) * nv * ks);
  for (int m = 0; m < ks; m++)
    for (int iv = 0; iv < nv; iv++) {
      const double h = t[iv + ks] - t[iv + ks - 1];  /* t(iv+ks) - t(iv+ks-1), 1-based iv */
      d->gr[iv + nv * m] = h * d->gx[m] + t[iv + ks - 1];
      d->grm[iv + nv * m] = 1.0 / d->gr[iv + nv * m];
      d->grw[iv + nv * m] = h * d->gw[m];
    }
  d->tp = knots_for(t, ks, ns, ksp, d->nsp);
  d->tq = knots_for(t, ks, ns, ksq, d->nsq);
  d->pbsp = values_for(d, d->tp, ksp);
  d->qbsp = values_for(d, d->tq, ksq);
  return d;
}

void orc_dbs_tables(const OrcDbs* d, double* gr, double* grw, double* pbsp, double* qbsp, double* tp,
                    double* tq) {
  const int nv = d->nv, ks = d->ks;
  if (gr) memcpy(gr, d->gr, sizeof(double) * nv * ks);
  if (grw) memcpy(grw, d->grw, sizeof(double) * nv * ks);
  if (pbsp) memcpy(pbsp, d->pbsp, sizeof(double) * (nv + 1) * ks * d->ksp);
  if (qbsp) memcpy(qbsp, d->qbsp, sizeof(double) * (nv + 1) * ks * d->ksq);
  if (tp) memcpy(tp, d->tp, sizeof(double) * (d->nsp + d->ksp));
  if (tq) memcpy(tq, d->tq, sizeof(double) * (d->nsq + d->ksq));
}

static void basis_of(const OrcDbs* d, int q, const double** tb, const double** vals, int* kb) {
  *tb = q ? d->tq : d->tp;
  *vals = q ? d->qbsp : d->pbsp;
  *kb = q ? d->ksq : d->ksp;
}

/* moments of basis q (0 P, 1 Q): rkm(ii,iv) = sum_m B_i B_j r^k w over cell
 * iv, pairs i<=j packed (moment_pp, DBS_moments.f90:166-190). */
void orc_dbs_moments(const OrcDbs* d, int q, int k, double* rkm) {
  const int nv = d->nv, ks = d->ks, kk = ks * ks;
  const double* tb; const double* v; int kb;
  basis_of(d, q, &tb, &v, &kb);
  memset(rkm, 0, sizeof(double) * kk * nv);
  double gw[32], b1[32][32], b2[32][32];
  for (int iv = 0; iv < nv; iv++) {
    for (int m = 0; m < ks; m++) {
      const int c = iv + nv * m;
      gw[m] = d->grw[c];
      if (k > 0) gw[m] = d->grw[c] * orc_dpow(d->gr[c], k);
      if (k < 0) gw[m] = d->grw[c] * orc_dpow(d->grm[c], -k);
    }
    for (int j = 0; j < kb; j++)
      for (int m = 0; m < ks; m++) {
        b1[j][m] = v[iv + (nv + 1) * (m + ks * j)];
        b2[j][m] = b1[j][m] * gw[m];
      }
    int ii = 0;
    for (int i = 0; i < kb; i++)
      for (int j = i; j < kb; j++) {
        double s = 0.0;
        for (int m = 0; m < ks; m++) s += b1[i][m] * b2[j][m];
        rkm[ii++ + kk * iv] = s;
      }
  }
}

static void family_bases(int family, int* q1, int* q2) {
  static const int b1[4] = {0, 1, 0, 1}, b2[4] = {0, 1, 1, 0};
  *q1 = b1[family];
  *q2 = b2[family];
}

/* inner integrals INT(pair, m) = int_{t_iv}^{gr(iv,m)} B_j B_jp r^k dr of
 * basis q at the inner Gauss points of cell iv (first integration of the
 * triang_* routines) */
static void inner_ints(const OrcDbs* d, int q, int k, int iv, const double* x, const double* w,
                       double ints[][32]) {
  const int ks = d->ks, nv = d->nv;
  const double* tb; const double* v; int kb;
  basis_of(d, q, &tb, &v, &kb);
  const int left = iv + kb;  /* 1-based knot index: iv+kb-1 with 1-based iv */
  const double xbase = tb[left - 1];
  double bt[32][32], gx[32], gwv[32];
  for (int m = 0; m < ks; m++) {
    const double r = d->gr[iv + nv * m];
    for (int i = 0; i < ks; i++) {
      gx[i] = (r - xbase) * x[i] + xbase;
      dbs_bvalues(tb, kb, left, gx[i], bt[i]);
      gwv[i] = (r - xbase) * w[i];
      gx[i] = gwv[i] * orc_dpow(gx[i], k);
    }
    int jj = 0;
    for (int j = 0; j < kb; j++) {
      double g[32];
      for (int i = 0; i < ks; i++) g[i] = gx[i] * bt[i][j];
      for (int jp = j; jp < kb; jp++) {
        double s = 0.0;
        for (int i = 0; i < ks; i++) s += g[i] * bt[i][jp];
        ints[jj++][m] = s;
      }
    }
  }
}

/* outer pair values P(pair, m) = B_i B_ip (grw grm^(k+1)) of basis q */
static void outer_pairs(const OrcDbs* d, int q, int k, int iv, double outer[][32]) {
  const int ks = d->ks, nv = d->nv;
  const double* tb; const double* v; int kb;
  basis_of(d, q, &tb, &v, &kb);
  double g[32];
  for (int m = 0; m < ks; m++) g[m] = d->grw[iv + nv * m] * orc_dpow(d->grm[iv + nv * m], k + 1);
  int ii = 0;
  for (int i = 0; i < kb; i++)
    for (int ip = i; ip < kb; ip++) {
      for (int m = 0; m < ks; m++)
        outer[ii][m] = v[iv + (nv + 1) * (m + ks * i)] * v[iv + (nv + 1) * (m + ks * ip)] * g[m];
      ii++;
    }
}

/* Diagonal cell iv of family f: rkd(ii,jj) over pairs ii of b1 (r1) and jj
 * of b2 (r2) = [r1 pair outer, r2 pair inner] + [r2 pair outer, r1 pair
 * inner], summed in the order of the reference routine of the family
 * (triang_pppp: a + a^T; triang_pqpq / triang_qpqp: the P-outer term first). */
void orc_dbs_triang(const OrcDbs* d, int family, int k, int iv, double* rkd) {
  const int ks = d->ks, kk = ks * ks;
  int q1, q2;
  family_bases(family, &q1, &q2);
  double x[32], w[32];
  orc_gauleg(0.0, 1.0, x, w, ks);
  static double int1[528][32], int2[528][32], out1[528][32], out2[528][32];
  const int kb1 = q1 ? d->ksq : d->ksp, kb2 = q2 ? d->ksq : d->ksp;
  const int ik1 = kb1 * (kb1 + 1) / 2, ik2 = kb2 * (kb2 + 1) / 2;
  inner_ints(d, q1, k, iv, x, w, int1);
  inner_ints(d, q2, k, iv, x, w, int2);
  outer_pairs(d, q1, k, iv, out1);
  outer_pairs(d, q2, k, iv, out2);
  /* r1-outer first unless the family's reference routine adds the P-outer
   * term first and r1 is Q (QPQP) */
  const int r2_first = (family == 3);
  for (int ii = 0; ii < ik1; ii++)
    for (int jj = 0; jj < ik2; jj++) {
      double s1 = 0.0, s2 = 0.0;
      for (int m = 0; m < ks; m++) s1 += out1[ii][m] * int2[jj][m];  /* r1 pair outer */
      for (int m = 0; m < ks; m++) s2 += out2[jj][m] * int1[ii][m];  /* r2 pair outer */
      rkd[ii + kk * jj] = r2_first ? (0.0 + s2) + s1 : (0.0 + s1) + s2;
    }
}

/* R^k band of family f (mrk_pppp and siblings): cell pairs (iv, jv) in
 * the reference's loop order, jv outer. */
void orc_dbs_mrk(const OrcDbs* d, int family, int k, double* rkb) {
  const int ks = d->ks, ns = d->ns, nv = d->nv, kk = ks * ks;
  int q1, q2;
  family_bases(family, &q1, &q2);
  const int kb1 = q1 ? d->ksq : d->ksp, kb2 = q2 ? d->ksq : d->ksp;
  double* m1k = (double*)malloc(sizeof(double) * kk * nv);
  double* m1m = (double*)malloc(sizeof(double) * kk * nv);
  double* m2k = (double*)malloc(sizeof(double) * kk * nv);
  double* m2m = (double*)malloc(sizeof(double) * kk * nv);
  double* rkd = (double*)malloc(sizeof(double) * kk * kk * nv);
  orc_dbs_moments(d, q1, k, m1k);
  orc_dbs_moments(d, q1, -(k + 1), m1m);
  orc_dbs_moments(d, q2, k, m2k);
  orc_dbs_moments(d, q2, -(k + 1), m2m);
  for (int iv = 0; iv < nv; iv++) orc_dbs_triang(d, family, k, iv, rkd + (size_t)kk * kk * iv);
  memset(rkb, 0, sizeof(double) * (size_t)ns * ns * ks * ks);
  for (int jv = 0; jv < nv; jv++) {
    int jj = 0;
    for (int jh = 0; jh < kb2; jh++)
      for (int jhp = jh; jhp < kb2; jhp++, jj++) {
        const int j = jv + jh, jp = jhp - jh;
        for (int iv = 0; iv < nv; iv++) {
          int ii = 0;
          for (int ih = 0; ih < kb1; ih++)
            for (int ihp = ih; ihp < kb1; ihp++, ii++) {
              const int i = iv + ih, ip = ihp - ih;
              double c;
              if (iv < jv) c = m1k[ii + kk * iv] * m2m[jj + kk * jv];
              else if (iv > jv) c = m2k[jj + kk * jv] * m1m[ii + kk * iv];
              else c = rkd[ii + (size_t)kk * (jj + (size_t)kk * iv)];
              rkb[i + (size_t)ns * (j + (size_t)ns * (ip + (size_t)ks * jp))] += c;
            }
        }
      }
  }
  free(m1k); free(m1m); free(m2k); free(m2m); free(rkd);
}
