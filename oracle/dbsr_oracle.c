/*
 * dbsr_oracle.c -- CPU FP64 restatement of the Dirac R^k assembly: R_data
 * over the four integral families and the accumulators it updates.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle); the product never links it.
 *
 * Reference routines followed (paths relative to the reference checkout):
 *   Gen_matrix (icase 2)  src/dbsr/DBSR_MAT3_MPI/gen_matrix.f90:11-14
 *                         (R_data for btype 0..3 on one list)
 *   R_data                src/dbsr/DBSR_MAT3_MPI/R_data.f90:2-290
 *                         itype 1 :92-165, 2 :167-212, 3 :214-248, 4 :250-281
 *   alloc_dbsr_matrix     updates.f90:2-76 (block numbering icc, icb, ibb)
 *   UPDATE_HS             updates.f90:109-177
 *   UPDATE_HW / HV / HB   updates.f90:181-261
 *   UPDATE_ACF            updates.f90:265-289
 *   Target_h              target_subs.f90:2-31
 *   KBORT                 src/libs/DBS/DBS_orbitals_pq.f90:248-277
 *   quadr_pq              src/libs/DBS/quadr_pq.f90:2-38
 * and the leaf routines density / convol / SUM_AmB / bav of bsr_oracle.c.
 *
 * State (Fortran column-major, 1-based indices in the lists):
 *   pq(ns,2,nbf) large (1) and small (2) components; ipbs(nbf) channel of
 *   an orbital (0 bound); V_ch(ms,nv_ch) the <.|p> vectors with keys
 *   (i_ch, j_ch) sorted lexicographically; iptar(nch) target of a channel.
 * Accumulators: hch(ms,ms,iicc), hcp(ms,iicb), hp(iibb), acf(iicc,0:mk),
 *   htarg(ntarg(ntarg+1)/2), numbered as alloc_dbsr_matrix does on one
 *   process (diagonal channel blocks first).  iitar = 0 (the default): the
 *   target-structure records only update htarg.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dbs_oracle.h"

#define IBO 32768

const char* orc_last_error(void);
static char g_msg[160];
static int fail(int code, const char* msg) {
  snprintf(g_msg, sizeof g_msg, "%s", msg);
  return code;
}
const char* orc_dbsr_last_error(void) { return g_msg; }

typedef struct {
  const OrcDbs* g;
  int ns, ks, ms, nbf, nch, npert, mk, ntarg, nvch;
  double *pq, *vch;
  int *ipbs, *ich, *jch, *iptar;
  int iicc, iicb, iibb;
  int* icc;
  double *hch, *hcp, *hp, *acf, *htarg;
  double* rkb[4];
  int rkb_k[4];
} OrcDbsr;

OrcDbsr* orc_dbsr_new(const OrcDbs* g, int nbf, const double* pq, const int* ipbs, int nvch, const int* ich,
                      const int* jch, const double* vch, int nch, int npert, int mk, int ntarg, const int* iptar) {
  OrcDbsr* w = (OrcDbsr*)calloc(1, sizeof(OrcDbsr));
  const int ns = g->ns, ms = 2 * ns;
  w->g = g; w->ns = ns; w->ks = g->ks; w->ms = ms; w->nbf = nbf; w->nch = nch; w->npert = npert;
  w->mk = mk; w->ntarg = ntarg; w->nvch = nvch;
  w->pq = (double*)malloc(sizeof(double) * (size_t)ns * 2 * nbf);
  memcpy(w->pq, pq, sizeof(double) * (size_t)ns * 2 * nbf);
  w->ipbs = (int*)malloc(sizeof(int) * nbf);
  memcpy(w->ipbs, ipbs, sizeof(int) * nbf);
  w->ich = (int*)malloc(sizeof(int) * (nvch + 1));
  w->jch = (int*)malloc(sizeof(int) * (nvch + 1));
  w->vch = (double*)malloc(sizeof(double) * (size_t)ms * (nvch + 1));
  if (nvch) {
    memcpy(w->ich, ich, sizeof(int) * nvch);
    memcpy(w->jch, jch, sizeof(int) * nvch);
    memcpy(w->vch, vch, sizeof(double) * (size_t)ms * nvch);
  }
  w->iptar = (int*)malloc(sizeof(int) * (nch + 1));
  memcpy(w->iptar, iptar, sizeof(int) * nch);
  /* alloc_dbsr_matrix on one process: diagonal blocks 1..nch, then the
   * lower triangle row by row */
  w->icc = (int*)calloc((size_t)nch * nch + 1, sizeof(int));
  int i = 0;
  for (int c = 1; c <= nch; c++) w->icc[(c - 1) + nch * (c - 1)] = ++i;
  for (int c = 2; c <= nch; c++)
    for (int d = 1; d < c; d++) {
      ++i;
      w->icc[(c - 1) + nch * (d - 1)] = i;
      w->icc[(d - 1) + nch * (c - 1)] = i;
    }
  w->iicc = i;
  w->iicb = nch * npert;
  w->iibb = npert * (npert + 1) / 2;
  w->hch = (double*)calloc((size_t)ms * ms * w->iicc + 1, sizeof(double));
  w->hcp = (double*)calloc((size_t)ms * w->iicb + 1, sizeof(double));
  w->hp = (double*)calloc((size_t)w->iibb + 1, sizeof(double));
  w->acf = (double*)calloc((size_t)w->iicc * (mk + 1) + 1, sizeof(double));
  w->htarg = (double*)calloc((size_t)ntarg * (ntarg + 1) / 2 + 1, sizeof(double));
  for (int f = 0; f < 4; f++) w->rkb_k[f] = -1;
  return w;
}

void orc_dbsr_free(OrcDbsr* w) {
  if (!w) return;
  free(w->pq); free(w->ipbs); free(w->ich); free(w->jch); free(w->vch); free(w->iptar); free(w->icc);
  free(w->hch); free(w->hcp); free(w->hp); free(w->acf); free(w->htarg);
  for (int f = 0; f < 4; f++) free(w->rkb[f]);
  free(w);
}

void orc_dbsr_sizes(const OrcDbsr* w, int* sizes) {
  sizes[0] = w->iicc; sizes[1] = w->iicb; sizes[2] = w->iibb; sizes[3] = w->ms;
}

void orc_dbsr_get(const OrcDbsr* w, double* hch, double* hcp, double* hp, double* acf, double* htarg) {
  const size_t ms = (size_t)w->ms;
  if (hch) memcpy(hch, w->hch, sizeof(double) * ms * ms * w->iicc);
  if (hcp) memcpy(hcp, w->hcp, sizeof(double) * ms * w->iicb);
  if (hp) memcpy(hp, w->hp, sizeof(double) * w->iibb);
  if (acf) memcpy(acf, w->acf, sizeof(double) * (size_t)w->iicc * (w->mk + 1));
  if (htarg) memcpy(htarg, w->htarg, sizeof(double) * (size_t)w->ntarg * (w->ntarg + 1) / 2);
}

static const double* comp(const OrcDbsr* w, int o, int ip) {
  return w->pq + (size_t)w->ns * ((ip - 1) + 2 * (size_t)(o - 1));
}

/* KBORT: binary search of (i,j) in the sorted keys; 0 when absent */
static int kbort(const OrcDbsr* w, int i, int j) {
  int k = 1, l = w->nvch;
  while (k <= l) {
    const int m = (k + l) / 2;
    if (i < w->ich[m - 1]) l = m - 1;
    else if (i > w->ich[m - 1]) k = m + 1;
    else if (j < w->jch[m - 1]) l = m - 1;
    else if (j > w->jch[m - 1]) k = m + 1;
    else return m;
  }
  return 0;
}

static int chk_ch(const OrcDbsr* w, int ich, int jch, const char* who) {
  if (ich < 0 || jch < 0 || ich > w->nch || jch > w->nch) return fail(-10, who);
  return 0;
}

/* UPDATE_HS: d (ns,ks) 's' band or (ns,ns) 'x' matrix into quadrant
 * (ii,jj) of block icc(ich,jch) */
static int update_hs(OrcDbsr* w, int ich, int jch, const double* d, char sym, int ii, int jj) {
  const int ns = w->ns, ks = w->ks, ms = w->ms;
  if (ich < 1 || jch < 1 || ich > w->nch || jch > w->nch) return fail(-10, "UPDATE_HS: channel index out of range");
  double* y = (double*)calloc((size_t)ns * ns, sizeof(double));
  if (sym == 's') {
    for (int jp = 0; jp < ks; jp++)
      for (int i = 0; i < ns - jp; i++) {
        const int j = i + jp;
        y[i + (size_t)ns * j] = d[i + (size_t)ns * jp];
        y[j + (size_t)ns * i] = d[i + (size_t)ns * jp];
      }
  } else {
    memcpy(y, d, sizeof(double) * (size_t)ns * ns);
  }
  const int b = w->icc[(ich - 1) + w->nch * (jch - 1)];
  double* h = w->hch + (size_t)ms * ms * (b - 1);
  const int i1 = (ii - 1) * ns, j1 = (jj - 1) * ns;
  if (ich > jch) {
    for (int c = 0; c < ns; c++)
      for (int r = 0; r < ns; r++) h[(i1 + r) + (size_t)ms * (j1 + c)] += y[r + (size_t)ns * c];
  } else if (jch > ich) {
    for (int c = 0; c < ns; c++)
      for (int r = 0; r < ns; r++) h[(i1 + r) + (size_t)ms * (j1 + c)] += y[c + (size_t)ns * r];
  } else {
    for (size_t q = 0; q < (size_t)ns * ns; q++) y[q] = y[q] / 2.0;
    for (int c = 0; c < ns; c++)
      for (int r = 0; r < ns; r++) h[(i1 + r) + (size_t)ms * (j1 + c)] += y[r + (size_t)ns * c];
    for (int c = 0; c < ns; c++)
      for (int r = 0; r < ns; r++) h[(j1 + r) + (size_t)ms * (i1 + c)] += y[c + (size_t)ns * r];
  }
  free(y);
  return 0;
}

/* UPDATE_HW: C v w^T (transposed when ich < jch, symmetrised when equal) */
static int update_hw(OrcDbsr* w, int ich, int jch, const double* v, const double* u, double C) {
  const int ms = w->ms;
  if (ich < 0 || jch < 0 || ich > w->nch || jch > w->nch) return fail(-10, "UPDATE_HW: channel index out of range");
  const int b = w->icc[(ich - 1) + w->nch * (jch - 1)];
  double* h = w->hch + (size_t)ms * ms * (b - 1);
  for (int j = 0; j < ms; j++)
    for (int i = 0; i < ms; i++) {
      double x;
      if (ich > jch) x = C * (v[i] * u[j]);
      else if (jch > ich) x = C * (v[j] * u[i]);
      else x = (C / 2.0) * (v[i] * u[j] + v[j] * u[i]);
      h[i + (size_t)ms * j] += x;
    }
  return 0;
}

static int update_hv(OrcDbsr* w, int ich, int ip, const double* v, double C) {
  if (ich < 1 || ich > w->nch) return fail(-13, "UPDATE_HV: channel index out is of range");
  if (ip < 1 || ip > w->npert) return fail(-14, "UPDATE_HV: bound index out is of range");
  double* h = w->hcp + (size_t)w->ms * ((ich - 1) * w->npert + (ip - 1));
  for (int i = 0; i < w->ms; i++) h[i] += C * v[i];
  return 0;
}

static int update_hb(OrcDbsr* w, int ic, int jc, double c) {
  if (ic > w->npert || jc > w->npert || ic <= 0 || jc <= 0) return fail(-12, "UPDATE_HB: index is out of range");
  const int i = ic > jc ? ic : jc, j = ic > jc ? jc : ic;
  w->hp[i * (i - 1) / 2 + j - 1] += c;
  return 0;
}

static int update_acf(OrcDbsr* w, int k, int ich, int jch, double C) {
  if (k > w->mk || k < 0) return fail(-17, "Update_cf: k out of range");
  if (ich < 0 || jch < 0 || ich > w->nch || jch > w->nch) return fail(-10, "Update_cf: channel index out of range");
  const int b = w->icc[(ich - 1) + w->nch * (jch - 1)];
  w->acf[(b - 1) + (size_t)w->iicc * k] += C;
  return 0;
}

static int target_h(OrcDbsr* w, int ich, int jch, double C) {
  if (ich < 0 || jch < 0 || ich > w->nch || jch > w->nch) return fail(-10, "target_h: channel index out of range");
  if (ich == 0 || jch == 0) return fail(-10, "target_h: bound orbital in a target-structure record");
  const int it = w->iptar[ich - 1], jt = w->iptar[jch - 1];
  const int i = it > jt ? it : jt, j = it > jt ? jt : it;
  w->htarg[(i - 1) * i / 2 + j - 1] += C;
  return 0;
}

/* quadr_pq: <p_io | r^m | p_jo> over both components */
static double quadr_pq(const OrcDbsr* w, int io, int jo, int m) {
  const OrcDbs* g = w->g;
  const int nv = g->nv, ks = g->ks;
  double s = 0.0, gx[32], gw[32];
  for (int pass = 0; pass < 2; pass++) {
    const int kb = pass ? g->ksq : g->ksp;
    const double* b = pass ? g->qbsp : g->pbsp;
    const double* pi_ = comp(w, io, pass + 1);
    const double* pj_ = comp(w, jo, pass + 1);
    for (int iv = 0; iv < nv; iv++) {
      for (int q = 0; q < ks; q++) gx[q] = orc_dpow(g->gr[iv + nv * q], m) * g->grw[iv + nv * q];
      for (int ith = 0; ith < kb; ith++) {
        const double pi = pi_[iv + ith];
        for (int q = 0; q < ks; q++) gw[q] = gx[q] * b[iv + (nv + 1) * (q + ks * ith)];
        for (int jth = 0; jth < kb; jth++) {
          const double pj = pj_[iv + jth];
          double t = 0.0;
          for (int q = 0; q < ks; q++) t += b[iv + (nv + 1) * (q + ks * jth)] * gw[q];
          s = s + t * pi * pj;
        }
      }
    }
  }
  return s;
}

double orc_dbsr_quadr_pq(const OrcDbsr* w, int io, int jo, int m) { return quadr_pq(w, io, jo, m); }

static const double* band(OrcDbsr* w, int f, int k) {
  if (w->rkb_k[f] != k) {
    const size_t n = (size_t)w->ns * w->ns * w->ks * w->ks;
    if (!w->rkb[f]) w->rkb[f] = (double*)malloc(sizeof(double) * n);
    orc_dbs_mrk(w->g, f, k, w->rkb[f]);
    w->rkb_k[f] = k;
  }
  return w->rkb[f];
}

static int bad_orb(const OrcDbsr* w, int o) { return o < 1 || o > w->nbf; }

/* R_data(itype, kpol, btype) for one list, records taken in list order. */
static int rdata_one(OrcDbsr* w, int itype, int kpol, int atype, int n, const int* K1, const int* K2, const int* K3,
                     const int* K4, const double* cd) {
  const int ns = w->ns, ks = w->ks, ms = w->ms;
  int ip1, ip2, jp1, jp2;
  switch (atype) {
    case 0: ip1 = 1; ip2 = 1; jp1 = 1; jp2 = 1; break;
    case 1: ip1 = 2; ip2 = 2; jp1 = 2; jp2 = 2; break;
    case 2:
      ip1 = 1; ip2 = 1; jp1 = 2; jp2 = 2;
      if (itype == 4) { ip1 = 1; ip2 = 2; jp1 = 1; jp2 = 2; }
      break;
    default:
      ip1 = 2; ip2 = 2; jp1 = 1; jp2 = 1;
      if (itype == 4) { ip1 = 2; ip2 = 1; jp1 = 2; jp2 = 1; }
      break;
  }
  const double* rkb = band(w, atype, kpol);
  const size_t nn = (size_t)ns * ns;
  double* d = (double*)calloc(nn, sizeof(double));
  double* dd = (double*)calloc(nn, sizeof(double));
  double* xx = (double*)calloc(nn, sizeof(double));
  double* v = (double*)calloc(ns, sizeof(double));
  double* vv = (double*)calloc(ms, sizeof(double));
  int rc = 0;
  if (itype == 1) {
    int iii = 0, jjj = 0;
    double S = 0.0;
    for (int q = 0; q < n && !rc; q++) {
      if (K1[q] != iii) {
        const int i1 = K1[q] / IBO, i2 = K1[q] % IBO;
        if (bad_orb(w, i1) || bad_orb(w, i2)) { rc = fail(-30, "orbital index out of range"); break; }
        orc_density(ns, ks, d, comp(w, i1, ip1), comp(w, i2, ip2), 's');
        orc_convol(ns, ks, xx, d, 2, 's', 's', rkb);
      }
      if (K2[q] != jjj) {
        const int j1 = K2[q] / IBO, j2 = K2[q] % IBO;
        if (bad_orb(w, j1) || bad_orb(w, j2)) { rc = fail(-30, "orbital index out of range"); break; }
        orc_density(ns, ks, dd, comp(w, j1, jp1), comp(w, j2, jp2), 's');
      }
      if (K1[q] != iii || K2[q] != jjj) S = orc_sum_amb(ns, ks, xx, dd, 's');
      iii = K1[q]; jjj = K2[q];
      const double C = S * cd[q];
      const int ic = -K3[q], jc = -K4[q], io = -ic, jo = -jc;
      if (ic > 0 && jc > 0) {
        rc = update_hb(w, ic, jc, C);
      } else if (ic < 0 && jc > 0) {
        const int io1 = io / IBO, io2 = io % IBO, k = kbort(w, io1, io2);
        if (k == 0) { rc = fail(-19, "RK: KBORT =0"); break; }
        rc = update_hv(w, w->ipbs[io1 - 1], jc, w->vch + (size_t)ms * (k - 1), C);
      } else if (ic < 0 && jc < 0) {
        const int io1 = io / IBO, io2 = io % IBO, jo1 = jo / IBO, jo2 = jo % IBO;
        const int k = kbort(w, io1, io2), l = kbort(w, jo1, jo2);
        if (k == 0 || l == 0) { rc = fail(-19, "RK: KBORT =0"); break; }
        rc = update_hw(w, w->ipbs[io1 - 1], w->ipbs[jo1 - 1], w->vch + (size_t)ms * (k - 1),
                       w->vch + (size_t)ms * (l - 1), C);
      } else if (ic < 0 && jc == 0) {
        const int io1 = io / IBO, io2 = io % IBO;
        if (bad_orb(w, io1) || bad_orb(w, io2)) { rc = fail(-30, "orbital index out of range"); break; }
        rc = target_h(w, w->ipbs[io1 - 1], w->ipbs[io2 - 1], C);
      } else {
        rc = fail(-22, " R_data: unknown structure for itype=1");
      }
    }
  } else if (itype == 2) {
    int jjj = 0, iii = 0;
    for (int q = 0; q < n && !rc; q++) {
      const int jj = K1[q], ii = K2[q], ich = K3[q], io = K4[q];
      if (jj != jjj || ii != iii) {
        if (jj != jjj) {
          const int j1 = jj / IBO, j2 = jj % IBO;
          if (bad_orb(w, j1) || bad_orb(w, j2)) { rc = fail(-30, "orbital index out of range"); break; }
          orc_density(ns, ks, dd, comp(w, j1, jp1), comp(w, j2, jp2), 's');
          orc_convol(ns, ks, xx, dd, 1, 's', 's', rkb);
          jjj = jj;
        }
        if (bad_orb(w, ii)) { rc = fail(-30, "orbital index out of range"); break; }
        orc_bav(ns, ks, xx, comp(w, ii, ip1), v, 's', 'r');
        iii = ii;
      }
      memset(vv, 0, sizeof(double) * ms);
      memcpy(vv + (ip2 == 1 ? 0 : ns), v, sizeof(double) * ns);
      if (io > 0) {
        const int io1 = io / IBO, io2 = io % IBO, k = kbort(w, io1, io2);
        if (k == 0) { rc = fail(-19, "RK: KBORT = 0"); break; }
        rc = update_hw(w, ich, w->ipbs[io1 - 1], vv, w->vch + (size_t)ms * (k - 1), cd[q]);
      } else if (io < 0) {
        rc = update_hv(w, ich, -io, vv, cd[q]);
      } else {
        rc = fail(-23, " R_data: uknown structure for itype=2 ");
      }
    }
  } else if (itype == 3 || itype == 4) {
    const char sym = itype == 3 ? 's' : 'x';
    const int icase = itype == 3 ? 1 : 3;
    const size_t nd = itype == 3 ? (size_t)ns * ks : nn;
    double CA = 0.0;
    memset(xx, 0, sizeof(double) * nn);
    for (int q = 0; q < n && !rc; q++) {
      const int j1 = K2[q], j2 = K3[q];
      if (bad_orb(w, j1) || bad_orb(w, j2)) { rc = fail(-30, "orbital index out of range"); break; }
      orc_density(ns, ks, dd, comp(w, j1, jp1), comp(w, j2, jp2), sym);
      for (size_t e = 0; e < nd; e++) xx[e] = xx[e] + cd[q] * dd[e];
      if (itype == 3 && atype <= 1) CA = CA + cd[q] * quadr_pq(w, j1, j2, kpol);
      if (q < n - 1 && K1[q] == K1[q + 1]) continue;
      orc_convol(ns, ks, dd, xx, icase, 's', 's', rkb);
      const int ich = K1[q] / IBO, jch = K1[q] % IBO;
      rc = update_hs(w, ich, jch, dd, sym, ip1, ip2);
      if (!rc && itype == 3 && atype <= 1) rc = update_acf(w, kpol, ich, jch, CA);
      memset(xx, 0, sizeof(double) * nn);
      CA = 0.0;
    }
  } else {
    rc = fail(-21, " R_data: unknown itype ");
  }
  free(d); free(dd); free(xx); free(v); free(vv);
  return rc;
}

/* Gen_matrix for icase 2: R_data(itype, kpol, 0..3) on one list */
int orc_dbsr_rdata(OrcDbsr* w, int itype, int kpol, int n, const int* K1, const int* K2, const int* K3,
                   const int* K4, const double* cd) {
  if (n == 0) return 0;
  for (int a = 0; a < 4; a++) {
    const int rc = rdata_one(w, itype, kpol, a, n, K1, K2, K3, K4, cd);
    if (rc) return rc;
  }
  return 0;
}
