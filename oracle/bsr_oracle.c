/*
 * bsr_oracle.c -- CPU FP64 restatement of the radial core of bsr_mat.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: tests/, the
 * smoke check and the CPU-baseline leg of bench.py may load it as the
 * checker.  The product (bsr_amd/) never links or calls it.
 *
 * Every routine follows the loop structure and quadrature of the reference
 * routine it cites (paths relative to the reference checkout).  Arrays that
 * cross the reference's Fortran interfaces keep Fortran column-major layout
 * so that fixtures, the reference build and the GPU library exchange the
 * same bytes:
 *   t(ns+ks)            knots
 *   rkb(ns,ns,ks,ks)    R^k band, rkb(i,j,ip,jp) = R^k(B_i B_i+ip-1 ; B_j B_j+jp-1)
 *   d(ns,ks) / d(ns,ns) densities (upper band 's' / full 'x')
 *   sb(ns,ks)           Gram band, lower storage: sb(i,jp) = <B_i|B_i+jp-ks>
 *   hcc(ns,ns,nch(nch+1)/2), hcb(ns,nch,ncp), hbb(ncp(ncp+1)/2),
 *   acf(nch,nch,0:mk), htarg/otarg(nch(nch+1)/2)
 * Internal C indices are 0-based; comments give the 1-based Fortran names.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define IBI 32768 /* packing base for orbital pairs, mod_cmdata.f90:49 */
#define IBO 32768 /* packing base for overlaps, mod_param.f90:77 */

static char g_err[256];
const char* orc_last_error(void) { return g_err; }
static int fail(int code, const char* msg) {
  snprintf(g_err, sizeof g_err, "%s", msg);
  return code;
}

/* integer power by repeated squaring (Fortran real**integer) */
static double powi(double x, int n) {
  if (n < 0) return 1.0 / powi(x, -n);
  double r = 1.0, b = x;
  while (n) {
    if (n & 1) r *= b;
    b *= b;
    n >>= 1;
  }
  return r;
}

/* ------------------------------------------------------------------ */
/* a1: knot sequences                                                  */
/* ------------------------------------------------------------------ */

/* Semi-exponential grid: ml steps of h in z*r up to 1, then geometric
 * growth by (1+h) while the step stays <= hmax, then steps of hmax up to
 * rmax; ks-fold end knots.  Follows mkgrid1, src/libs/BS/define_grid.f90:213-276.
 * Outputs t (1-based t(1..ns+ks) in t[0..ns+ks-1]), ns, ml, me and the
 * adjusted h (=1/ml).  Returns 0 or a negative code. */
int orc_mkgrid1(int ks, double z, double h, double hmax, double rmax, int cap,
                double* t, int* ns_out, int* ml_out, int* me_out, double* h_out) {
  double zr = rmax * z, zh = hmax * z;
  int nt = (int)((zr / h + 1 + ks) * 2);
  double* tt = (double*)calloc((size_t)nt + 2 * ks + 4, sizeof(double));
  int ns = ks; /* tt is 1-based: tt[n] == t(n) */
  int ml = (int)(1.0 / h + 0.5);
  h = 1.0 / ml;
  for (int i = 1; i <= ml; i++) { ns++; tt[ns] = tt[ns - 1] + h; }
  if (zh < h) zh = h;
  int me = 0;
  for (;;) {
    tt[ns + 1] = tt[ns] * (1.0 + h);
    if (tt[ns + 1] - tt[ns] > zh) break;
    ns++;
    if (tt[ns] > zr) break;
    me++;
  }
  if (tt[ns] < zr) {
    for (;;) {
      tt[ns + 1] = tt[ns] + zh;
      ns++;
      if (tt[ns] >= zr) break;
    }
  }
  if (tt[ns] > zr) {
    ns--; me--;
    for (int i = ns + 1; i <= ns + ks; i++) tt[i] = zr;
    tt[ns] = (tt[ns + 1] + tt[ns - 1]) / 2.0;
  }
  if (ns + ks > cap) { free(tt); return fail(-1, "orc_mkgrid1: knot buffer too small"); }
  for (int i = 1; i <= ns + ks; i++) t[i - 1] = tt[i] / z;
  *ns_out = ns; *ml_out = ml; *me_out = me; *h_out = h;
  free(tt);
  return 0;
}

/* Original grid (grid_type 0): ml = 1/h intervals, me geometric, m linear
 * steps h2; follows mkgrid, src/libs/BS/define_grid.f90:86-153. */
int orc_mkgrid0(int ks, double z, double h, double hmax, double rmax, int cap,
                double* t, int* ns_out, int* ml_out, int* me_out, double* h_out) {
  double h0 = h;
  int ml = (int)(1.0 / h + 0.1);
  h = 1.0 / ml;
  double hp1 = 1.0 + h;
  double tmax = z * rmax, zh = z * hmax;
  double me1d = log(zh / h) / log(hp1) + 1.0;
  int me1 = (int)(me1d > 0.0 ? me1d : 0.0);
  int me2 = (int)(log(tmax) / log(hp1) + 1);
  int me, m;
  double h2 = 0.0;
  if (me2 <= me1) { me = me2; m = 0; }
  else {
    me = me1;
    double tx = powi(hp1, me);
    h2 = h * tx / hp1;
    m = (int)lround((tmax - tx) / h2);
  }
  int n = ml + me + m + ks - 1;
  int ns = n, nt = ns + ks;
  if (nt > cap) return fail(-1, "orc_mkgrid0: knot buffer too small");
  double* tt = (double*)calloc((size_t)nt + 2, sizeof(double));
  for (int i = ks + 1; i <= ks + ml; i++) tt[i] = tt[i - 1] + h0;
  for (int i = ks + ml + 1; i <= ks + me + ml; i++) tt[i] = tt[i - 1] * hp1;
  for (int i = ks + me + ml + 1; i <= n + 1; i++) tt[i] = tt[i - 1] + h2;
  for (int i = n + 2; i <= nt; i++) tt[i] = tt[n + 1];
  for (int i = 1; i <= nt; i++) t[i - 1] = tt[i] / z;
  *ns_out = ns; *ml_out = ml; *me_out = me; *h_out = h;
  free(tt);
  return 0;
}

/* ------------------------------------------------------------------ */
/* a2: Gauss-Legendre on [0,1]                                          */
/* ------------------------------------------------------------------ */

/* k-point Gauss-Legendre nodes (ascending) and weights on [0,1].  The
 * reference looks them up in a 15-digit table (src/libs/BS/gauss.f90:2-283);
 * here they are computed by Newton iteration on P_k, and the tests compare
 * them with that table. */
void orc_gauss(int k, double* x, double* w) {
  const double pi = 3.14159265358979323846;
  for (int i = 1; i <= k; i++) {
    double z = cos(pi * (i - 0.25) / (k + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; it++) {
      double p0 = 1.0, p1 = z;
      for (int n = 2; n <= k; n++) {
        double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
        p0 = p1; p1 = p2;
      }
      if (k == 1) { p1 = z; p0 = 1.0; }
      dp = k * (z * p1 - p0) / (z * z - 1.0);
      double dz = p1 / dp;
      z -= dz;
      if (fabs(dz) < 1e-16) break;
    }
    {
      double p0 = 1.0, p1 = z;
      for (int n = 2; n <= k; n++) {
        double p2 = ((2.0 * n - 1.0) * z * p1 - (n - 1.0) * p0) / n;
        p0 = p1; p1 = p2;
      }
      dp = k * (z * p1 - p0) / (z * z - 1.0);
    }
    x[i - 1] = (1.0 - z) / 2.0;
    w[i - 1] = 1.0 / ((1.0 - z * z) * dp * dp);
  }
}

/* ------------------------------------------------------------------ */
/* a3: B-spline values (de Boor recurrence)                            */
/* ------------------------------------------------------------------ */

/* Values of the ks B-splines B_{left-ks+1..left} (1-based) at x, with
 * t(left) <= x <= t(left+1).  The recurrence of vbsplvb,
 * src/libs/BS/vbsplvd.f90:173-286 (values only, nderiv=1). */
static void bvalues(const double* t, int ks, int left, double x, double* b) {
  double dr[32], dl[32];
  b[0] = 1.0;
  for (int j = 1; j < ks; j++) {
    dr[j - 1] = t[left + j - 1] - x;    /* t(left+j) - x */
    dl[j - 1] = x - t[left - j];        /* x - t(left+1-j) */
    double saved = 0.0;
    for (int r = 0; r < j; r++) {
      double term = b[r] / (dr[r] + dl[j - 1 - r]);
      b[r] = saved + dr[r] * term;
      saved = dl[j - 1 - r] * term;
    }
    b[j] = saved;
  }
}

typedef struct {
  int ks, ns, nv, ml, me;
  double h;
  double* t;   /* ns+ks */
  double *gx, *gw; /* ks */
  double *gr, *grm, *grw; /* [iv*ks+m] */
  double* bsp; /* [(iv*ks+m)*ks+ith] = B_{iv+ith}(gr(iv,m)), 0-based spline */
  double* sb;  /* (ns,ks) lower band */
  /* cached R^k band */
  int krk;
  double* rkb; /* ns*ns*ks*ks */
  double *rkd1, *rkd2, *rkd; /* ik*nv, ik*nv, ik*ik*nv */
} orc_grid;

void orc_mrm(const orc_grid* g, int mm, double* rm);

/* Gaussian points, weights and B-spline values per interval, as initvb
 * (src/libs/BS/define_spline.f90:67-129), and the Gram band sb = mrm(0)
 * (define_spline.f90:162, mrm.f90:2-91). */
orc_grid* orc_grid_new(int ks, int ns, int ml, int me, double h, const double* t) {
  orc_grid* g = (orc_grid*)calloc(1, sizeof(orc_grid));
  g->ks = ks; g->ns = ns; g->nv = ns - ks + 1; g->ml = ml; g->me = me; g->h = h;
  int nv = g->nv, ik = ks * (ks + 1) / 2;
  g->t = (double*)malloc(sizeof(double) * (ns + ks));
  memcpy(g->t, t, sizeof(double) * (ns + ks));
  g->gx = (double*)malloc(sizeof(double) * ks);
  g->gw = (double*)malloc(sizeof(double) * ks);
  orc_gauss(ks, g->gx, g->gw);
  g->gr = (double*)malloc(sizeof(double) * nv * ks);
  g->grm = (double*)malloc(sizeof(double) * nv * ks);
  g->grw = (double*)malloc(sizeof(double) * nv * ks);
  g->bsp = (double*)calloc((size_t)(nv + 1) * ks * ks, sizeof(double));
  for (int m = 0; m < ks; m++)
    for (int iv = 0; iv < nv; iv++) {
      double tl = t[iv + ks - 1], tr = t[iv + ks];   /* t(iv+ks-1), t(iv+ks) 1-based iv */
      double x = (tr - tl) * g->gx[m] + tl;
      g->gr[iv * ks + m] = x;
      g->grm[iv * ks + m] = 1.0 / x;
      g->grw[iv * ks + m] = (tr - tl) * g->gw[m];
      double b[32];
      bvalues(t, ks, iv + ks, x, b);
      for (int ith = 0; ith < ks; ith++) g->bsp[(iv * ks + m) * ks + ith] = b[ith];
    }
  g->sb = (double*)calloc((size_t)ns * ks, sizeof(double));
  orc_mrm(g, 0, g->sb);
  g->krk = -100;
  g->rkb = (double*)calloc((size_t)ns * ns * ks * ks, sizeof(double));
  g->rkd1 = (double*)calloc((size_t)ik * nv, sizeof(double));
  g->rkd2 = (double*)calloc((size_t)ik * nv, sizeof(double));
  g->rkd = (double*)calloc((size_t)ik * ik * nv, sizeof(double));
  return g;
}

void orc_grid_free(orc_grid* g) {
  if (!g) return;
  free(g->t); free(g->gx); free(g->gw); free(g->gr); free(g->grm); free(g->grw);
  free(g->bsp); free(g->sb); free(g->rkb); free(g->rkd1); free(g->rkd2); free(g->rkd);
  free(g);
}

int orc_grid_nv(const orc_grid* g) { return g->nv; }

/* copy out tables: gr,grw (nv*ks, [iv*ks+m]), bsp ((nv)*ks*ks), sb (ns*ks) */
void orc_grid_tables(const orc_grid* g, double* gr, double* grw, double* bsp, double* sb) {
  int n = g->nv * g->ks;
  if (gr) memcpy(gr, g->gr, sizeof(double) * n);
  if (grw) memcpy(grw, g->grw, sizeof(double) * n);
  if (bsp) memcpy(bsp, g->bsp, sizeof(double) * n * g->ks);
  if (sb) memcpy(sb, g->sb, sizeof(double) * g->ns * g->ks);
}

/* B-spline values at an arbitrary point r (all ns), for building test
 * orbitals; zero outside [t(ks), t(ns+1)]. */
void orc_bspline_at(const orc_grid* g, double r, double* out) {
  int ks = g->ks, ns = g->ns;
  memset(out, 0, sizeof(double) * ns);
  const double* t = g->t;
  if (r < t[ks - 1] || r > t[ns]) return;
  int left = ks;
  while (left < ns && r >= t[left]) left++;
  double b[32];
  bvalues(t, ks, left, r, b);
  for (int i = 0; i < ks; i++) out[left - ks + i] = b[i];
}

/* ------------------------------------------------------------------ */
/* a4: <B_i | r^mm | B_j> band, lower storage; mrm.f90:2-91             */
/* ------------------------------------------------------------------ */
void orc_mrm(const orc_grid* g, int mm, double* rm) {
  int ks = g->ks, ns = g->ns, nv = g->nv;
  memset(rm, 0, sizeof(double) * ns * ks);
  for (int ith = 0; ith < ks; ith++)
    for (int jth = 0; jth <= ith; jth++) {
      int jcol = jth - ith + ks - 1;
      for (int iv = 0; iv < nv; iv++) {
        int irow = iv + ith;
        double s = 0.0;
        for (int m = 0; m < ks; m++) {
          double bi = g->grw[iv * ks + m] * g->bsp[(iv * ks + m) * ks + ith];
          double bj = g->bsp[(iv * ks + m) * ks + jth];
          if (mm == 0) s += bi * bj;
          else if (mm == -1) s += bi * g->grm[iv * ks + m] * bj;
          else if (mm <= -2) s += bi * powi(g->grm[iv * ks + m], -mm) * bj;
          else if (mm == 1) s += bi * g->gr[iv * ks + m] * bj;
          else s += bi * powi(g->gr[iv * ks + m], mm) * bj;
        }
        rm[irow + ns * jcol] += s;
      }
    }
}

/* ------------------------------------------------------------------ */
/* a5: moments <B_i|r^k|B_j> per interval, packed i<=j, with the        */
/*     scaling law in the geometric region; moments.f90:2-133           */
/* ------------------------------------------------------------------ */
static void moment1(const orc_grid* g, int iv, int k, double* rv) {
  int ks = g->ks;
  double gw[32];
  for (int m = 0; m < ks; m++) {
    gw[m] = g->grw[iv * ks + m];
    if (k > 0) gw[m] = g->grw[iv * ks + m] * powi(g->gr[iv * ks + m], k);
    if (k < 0) gw[m] = g->grw[iv * ks + m] * powi(g->grm[iv * ks + m], -k);
  }
  int ii = 0;
  for (int i = 0; i < ks; i++)
    for (int j = i; j < ks; j++) {
      double s = 0.0;
      for (int m = 0; m < ks; m++)
        s += g->bsp[(iv * ks + m) * ks + i] * (g->bsp[(iv * ks + m) * ks + j] * gw[m]);
      rv[ii++] = s;
    }
}

/* rkm[iv*ik + ii] */
void orc_moments(const orc_grid* g, int k, double* rkm) {
  int ks = g->ks, nv = g->nv, ik = ks * (ks + 1) / 2;
  if (g->me == 0) {
    for (int iv = 0; iv < nv; iv++) moment1(g, iv, k, rkm + (size_t)iv * ik);
    return;
  }
  int ml = g->ml, me = g->me;
  for (int iv = 1; iv <= ml + ks - 1; iv++) moment1(g, iv - 1, k, rkm + (size_t)(iv - 1) * ik);
  double hp = powi(1.0 + g->h, k) * (1.0 + g->h);
  for (int iv = ml + ks; iv <= ml + me - ks + 2; iv++)
    for (int ii = 0; ii < ik; ii++)
      rkm[(size_t)(iv - 1) * ik + ii] = rkm[(size_t)(iv - 2) * ik + ii] * hp;
  for (int iv = ml + me - ks + 3; iv <= nv; iv++) moment1(g, iv - 1, k, rkm + (size_t)(iv - 1) * ik);
}

/* ------------------------------------------------------------------ */
/* a6: diagonal-cell quadrature rk_triang; rk_moments.f90:105-207        */
/* rkd[(iv*ik + jj)*ik + ii]                                            */
/* ------------------------------------------------------------------ */
static void rk_triang(const orc_grid* g, int iv, int k, double* rkd) {
  int ks = g->ks, ik = ks * (ks + 1) / 2;
  const double* t = g->t;
  int left = iv + ks; /* 1-based left = iv(1-based)+ks-1 */
  double xbase = t[left - 1];
  double INT[32 * 32 * 32];
  /* INT[(m*ks + j)*ks + jp] */
  for (int m = 0; m < ks; m++) {
    double gxx[32], gww[32], bt[32][32];
    double r1 = g->gr[iv * ks + m];
    for (int i = 0; i < ks; i++) {
      gxx[i] = (r1 - xbase) * g->gx[i] + xbase;
      bvalues(t, ks, left, gxx[i], bt[i]);
    }
    for (int i = 0; i < ks; i++) gww[i] = (r1 - xbase) * g->gw[i];
    for (int i = 0; i < ks; i++) {
      if (k > 1) gxx[i] = gww[i] * powi(gxx[i], k);
      else if (k == 1) gxx[i] = gww[i] * gxx[i];
      else gxx[i] = gww[i];
    }
    for (int j = 0; j < ks; j++) {
      double gj[32];
      for (int i = 0; i < ks; i++) gj[i] = gxx[i] * bt[i][j];
      for (int jp = j; jp < ks; jp++) {
        double s = 0.0;
        for (int i = 0; i < ks; i++) s += gj[i] * bt[i][jp];
        INT[(m * ks + j) * ks + jp] = s;
      }
    }
  }
  double gx2[32];
  for (int m = 0; m < ks; m++)
    gx2[m] = (k != 0) ? g->grw[iv * ks + m] * powi(g->grm[iv * ks + m], k + 1)
                      : g->grw[iv * ks + m] * g->grm[iv * ks + m];
  double* a = (double*)malloc(sizeof(double) * ik * ik);
  int ii = 0;
  for (int i = 0; i < ks; i++)
    for (int ip = i; ip < ks; ip++) {
      double bi[32];
      for (int m = 0; m < ks; m++)
        bi[m] = g->bsp[(iv * ks + m) * ks + i] * g->bsp[(iv * ks + m) * ks + ip] * gx2[m];
      int jj = 0;
      for (int j = 0; j < ks; j++)
        for (int jp = j; jp < ks; jp++) {
          double s = 0.0;
          for (int m = 0; m < ks; m++) s += bi[m] * INT[(m * ks + j) * ks + jp];
          a[ii + ik * jj] = s;
          jj++;
        }
      ii++;
    }
  for (int jj = 0; jj < ik; jj++)
    for (int i2 = 0; i2 < ik; i2++)
      rkd[(size_t)jj * ik + i2] = a[i2 + ik * jj] + a[jj + ik * i2];
  free(a);
}

/* rk_moments(k): rkd1 = moments(k), rkd2 = moments(-(k+1)), rkd by
 * rk_pdiag with the scaling law rkd(iv)=rkd(iv-1)*(1+h) in the geometric
 * region; src/libs/BS/rk_moments.f90:2-101 (rel=.false.). */
void orc_rk_moments(orc_grid* g, int k, double* rkd1, double* rkd2, double* rkd) {
  int ks = g->ks, nv = g->nv, ik = ks * (ks + 1) / 2;
  orc_moments(g, k, rkd1);
  orc_moments(g, -(k + 1), rkd2);
  size_t cell = (size_t)ik * ik;
  if (g->me == 0) {
    for (int iv = 0; iv < nv; iv++) rk_triang(g, iv, k, rkd + cell * iv);
    return;
  }
  int ml = g->ml, me = g->me;
  for (int iv = 1; iv <= ml + ks - 1; iv++) rk_triang(g, iv - 1, k, rkd + cell * (iv - 1));
  double hp = g->h + 1.0;
  for (int iv = ml + ks; iv <= ml + me - ks + 2; iv++)
    for (size_t q = 0; q < cell; q++) rkd[cell * (iv - 1) + q] = rkd[cell * (iv - 2) + q] * hp;
  for (int iv = ml + me - ks + 3; iv <= nv; iv++) rk_triang(g, iv - 1, k, rkd + cell * (iv - 1));
}

/* ------------------------------------------------------------------ */
/* a7: R^k band by the cell algorithm; mrk_cell.f90:2-82                */
/* ------------------------------------------------------------------ */
static void mrk_cell_internal(orc_grid* g, int k) {
  if (g->krk == k) return;
  int ks = g->ks, ns = g->ns, nv = g->nv, ik = ks * (ks + 1) / 2;
  orc_rk_moments(g, k, g->rkd1, g->rkd2, g->rkd);
  double* rkb = g->rkb;
  memset(rkb, 0, sizeof(double) * ns * ns * ks * ks);
  size_t cell = (size_t)ik * ik;
  for (int jv = 0; jv < nv; jv++) {
    int jj = 0;
    for (int jh = 0; jh < ks; jh++) {
      int j = jv + jh;
      for (int jhp = jh; jhp < ks; jhp++) {
        int jp = jhp - jh;
        for (int iv = 0; iv < nv; iv++) {
          int ii = 0;
          for (int ih = 0; ih < ks; ih++) {
            int i = iv + ih;
            for (int ihp = ih; ihp < ks; ihp++) {
              int ip = ihp - ih;
              double c;
              if (iv < jv) c = g->rkd1[(size_t)iv * ik + ii] * g->rkd2[(size_t)jv * ik + jj];
              else if (iv > jv) c = g->rkd1[(size_t)jv * ik + jj] * g->rkd2[(size_t)iv * ik + ii];
              else c = g->rkd[cell * iv + (size_t)jj * ik + ii];
              rkb[i + (size_t)ns * (j + (size_t)ns * (ip + (size_t)ks * jp))] += c;
              ii++;
            }
          }
        }
        jj++;
      }
    }
  }
  g->krk = k;
}

void orc_mrk_cell(orc_grid* g, int k, double* rkb_out) {
  mrk_cell_internal(g, k);
  if (rkb_out) memcpy(rkb_out, g->rkb, sizeof(double) * g->ns * g->ns * g->ks * g->ks);
}

/* ------------------------------------------------------------------ */
/* a8: density; src/libs/BS/density.f90:2-58                            */
/* ------------------------------------------------------------------ */
void orc_density(int ns, int ks, double* d, const double* p1, const double* p2, char type) {
  if (type == 's') {
    memset(d, 0, sizeof(double) * ns * ks);
    for (int i = 0; i < ns; i++) d[i] = p1[i] * p2[i];
    for (int j = 1; j < ks; j++)
      for (int i = 0; i < ns - j; i++)
        d[i + ns * j] = p1[i] * p2[i + j] + p1[i + j] * p2[i];
  } else if (type == 'n') {
    memset(d, 0, sizeof(double) * ns * (2 * ks - 1));
    for (int j = 1; j <= 2 * ks - 1; j++) {
      int imin = 1 > 1 + ks - j ? 1 : 1 + ks - j;
      int imax = ns < ns + ks - j ? ns : ns + ks - j;
      for (int i = imin; i <= imax; i++)
        d[(i - 1) + ns * (j - 1)] = p1[i - 1] * p2[i + j - ks - 1];
    }
  } else {
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) d[i + ns * j] = p1[i] * p2[j];
  }
}

/* ------------------------------------------------------------------ */
/* a9: convolution of the R^k band with a density, symmetric-symmetric   */
/* cases; src/libs/BS/convol.f90:33-59 (icase 1,2), :163-203 (icase 3,4) */
/* ------------------------------------------------------------------ */
int orc_convol(int ns, int ks, double* a, const double* d, int icase, char sym_i,
               char sym_j, const double* rkb) {
  if (sym_i != 's' || sym_j != 's') return fail(-2, " CONVOL:  unknown case ");
#define RK(i, j, ip, jp) rkb[(i) + (size_t)ns * ((j) + (size_t)ns * ((ip) + (size_t)ks * (jp)))]
  if (icase > 2) memset(a, 0, sizeof(double) * ns * ns);
  if (icase == 1 || icase == 2) {
    for (int ip = 0; ip < ks; ip++)
      for (int i = 0; i < ns - ip; i++) {
        double c = 0.0;
        for (int jp = 0; jp < ks; jp++)
          for (int j = 0; j < ns - jp; j++)
            c += d[j + ns * jp] * (icase == 1 ? RK(i, j, ip, jp) : RK(j, i, jp, ip));
        a[i + ns * ip] = c;
      }
  } else if (icase == 3) {
    for (int jp = 0; jp < ks; jp++)
      for (int j = 0; j < ns - jp; j++) {
        int jj = j + jp;
        for (int ip = 0; ip < ks; ip++)
          for (int i = 0; i < ns - ip; i++) {
            int ii = i + ip;
            double c = RK(i, j, ip, jp);
            a[i + ns * jj] += c * d[ii + ns * j];
            if (ip > 0) a[ii + ns * jj] += c * d[i + ns * j];
            if (jp > 0) a[i + ns * j] += c * d[ii + ns * jj];
            if (ip > 0 && jp > 0) a[ii + ns * j] += c * d[i + ns * jj];
          }
      }
  } else if (icase == 4) {
    for (int jp = 0; jp < ks; jp++)
      for (int j = 0; j < ns - jp; j++) {
        int jj = j + jp;
        for (int ip = 0; ip < ks; ip++)
          for (int i = 0; i < ns - ip; i++) {
            int ii = i + ip;
            double c = RK(i, j, ip, jp);
            a[ii + ns * j] += c * d[i + ns * jj];
            if (ip > 0) a[i + ns * j] += c * d[ii + ns * jj];
            if (jp > 0) a[ii + ns * jj] += c * d[i + ns * j];
            if (ip > 0 && jp > 0) a[i + ns * jj] += c * d[ii + ns * j];
          }
      }
  } else {
    return fail(-2, " CONVOL:  unknown case ");
  }
#undef RK
  return 0;
}

/* ------------------------------------------------------------------ */
/* a10: band algebra                                                     */
/* ------------------------------------------------------------------ */

/* SUM(a*b) over the stored band; sum_amb.f90:2-64 */
double orc_sum_amb(int ns, int ks, const double* a, const double* b, char sym) {
  double x = 0.0;
  if (sym == 'x') {
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < ns; j++) x += a[i + ns * j] * b[i + ns * j];
  } else if (sym == 's') {
    for (int j = 0; j < ks; j++)
      for (int i = 0; i < ns - j; i++) x += a[i + ns * j] * b[i + ns * j];
  } else if (sym == 'l') {
    for (int j = 0; j < ks; j++)
      for (int i = ks - 1 - j; i < ns; i++) x += a[i + ns * j] * b[i + ns * j];
  } else if (sym == 'n') {
    for (int j = 1; j <= 2 * ks - 1; j++) {
      int imin = 1 > 1 + ks - j ? 1 : 1 + ks - j;
      int imax = ns < ns + ks - j ? ns : ns + ks - j;
      for (int i = imin; i <= imax; i++) x += a[(i - 1) + ns * (j - 1)] * b[(i - 1) + ns * (j - 1)];
    }
  }
  return x;
}

/* y = B v for a symmetric upper band ('s') or non-symmetric band;
 * bav.f90:2-79 */
void orc_bav(int ns, int ks, const double* b, const double* v, double* y, char type, char side) {
  if (type == 's') {
    for (int i = 0; i < ns; i++) y[i] = b[i] * v[i];
    for (int ip = 1; ip < ks; ip++)
      for (int i = 0; i < ns - ip; i++) {
        int j = i + ip;
        y[i] += b[i + ns * ip] * v[j];
        y[j] += b[i + ns * ip] * v[i];
      }
  } else {
    for (int i = 0; i < ns; i++) y[i] = 0.0;
    for (int ip = 1; ip <= 2 * ks - 1; ip++) {
      int imin = 1 > 1 + ks - ip ? 1 : 1 + ks - ip;
      int imax = ns < ns + ks - ip ? ns : ns + ks - ip;
      for (int i = imin; i <= imax; i++) {
        int j = i + ip - ks;
        if (side == 'r') y[i - 1] += b[(i - 1) + ns * (ip - 1)] * v[j - 1];
        else y[j - 1] += b[(i - 1) + ns * (ip - 1)] * v[i - 1];
      }
    }
  }
}

/* y = b v for a symmetric lower band; bxv.f90:2-43 */
void orc_bxv(int k, int n, const double* b, const double* v, double* y) {
  for (int i = 0; i < n; i++) y[i] = b[i + n * (k - 1)] * v[i];
  for (int jp = 1; jp <= k - 1; jp++)
    for (int i = k - jp + 1; i <= n; i++) {
      int j = i - k + jp;
      y[i - 1] += b[(i - 1) + n * (jp - 1)] * v[j - 1];
      y[j - 1] += b[(i - 1) + n * (jp - 1)] * v[i - 1];
    }
}

/* <x | array | y> for a symmetric lower band; bvmv.f90:2-67 ('s') */
double orc_bvmv(int n, int k, const double* a, const double* x, const double* y) {
  double s = 0.0;
  for (int jp = 1; jp <= k - 1; jp++)
    for (int i = k + 1 - jp; i <= n; i++) {
      int j = i + jp - k;
      s += a[(i - 1) + n * (jp - 1)] * (x[i - 1] * y[j - 1] + x[j - 1] * y[i - 1]);
    }
  double c = 0.0;
  for (int i = 0; i < n; i++) c += a[i + n * (k - 1)] * x[i] * y[i];
  return s + c;
}

/* <P_i | r^m | P_j>; quadr.f90:2-31 */
double orc_quadr(const orc_grid* g, const double* pi, const double* pj, int m) {
  double* rm = (double*)malloc(sizeof(double) * g->ns * g->ks);
  orc_mrm(g, m, rm);
  double s = orc_bvmv(g->ns, g->ks, rm, pi, pj);
  free(rm);
  return s;
}

/* ------------------------------------------------------------------ */
/* a11-a15: one partial wave (matrix blocks, I_data, O_data, UPDATE_*)  */
/* ------------------------------------------------------------------ */
typedef struct {
  orc_grid* g;
  int ns, ks, nbf, nch, ncp, mk;
  double *pbs, *qbs, *obs; /* (ns,nbf), (ns,nbf), (nbf,nbf) */
  int *iech, *lbs, *ibort; /* nbf, nbf, (nbf,nbf) */
  double eps_ovl, eps_c;
  double *hcc, *hcb, *hbb, *acf, *htarg, *otarg;
  double* qtab; /* quadr cache (nbf,nbf,0:mk) of m=k, computed on demand */
  int* qok;
  /* one-electron operator of Gen_Lval: hl_full (ns,ns,0:mlh), L_int
   * (nbf,nbf), L_vec (ns,nbf); lok once computed */
  int mlh, lok;
  double *hl_full, *lint, *lvec;
} orc_wave;

/* Orbital data as Read_data sets it up (read_data.f90:142-208):
 * qbs = sb*p for bound orbitals (iech==0), OBS(i,j) = <p_i|p_j> for bound
 * orbitals of equal l (|S|<eps_ovl -> 0, diagonal 1).  ibort is the
 * caller's nonzero pattern of allowed channel/bound overlaps. */
orc_wave* orc_wave_new(orc_grid* g, int nbf, const double* pbs, const int* iech, const int* lbs,
                       const int* ibort, int nch, int ncp, int mk, double eps_ovl, double eps_c) {
  orc_wave* w = (orc_wave*)calloc(1, sizeof(orc_wave));
  int ns = g->ns, ks = g->ks;
  w->g = g; w->ns = ns; w->ks = ks; w->nbf = nbf; w->nch = nch; w->ncp = ncp; w->mk = mk;
  w->eps_ovl = eps_ovl; w->eps_c = eps_c;
  w->pbs = (double*)malloc(sizeof(double) * ns * nbf);
  memcpy(w->pbs, pbs, sizeof(double) * ns * nbf);
  w->qbs = (double*)calloc((size_t)ns * nbf, sizeof(double));
  w->obs = (double*)calloc((size_t)nbf * nbf, sizeof(double));
  w->iech = (int*)malloc(sizeof(int) * nbf); memcpy(w->iech, iech, sizeof(int) * nbf);
  w->lbs = (int*)malloc(sizeof(int) * nbf); memcpy(w->lbs, lbs, sizeof(int) * nbf);
  w->ibort = (int*)malloc(sizeof(int) * nbf * nbf); memcpy(w->ibort, ibort, sizeof(int) * nbf * nbf);
  for (int i = 0; i < nbf; i++)
    if (iech[i] == 0) orc_bxv(ks, ns, g->sb, w->pbs + (size_t)ns * i, w->qbs + (size_t)ns * i);
  for (int i = 0; i < nbf; i++) {
    if (iech[i] != 0) continue;
    for (int j = 0; j <= i; j++) {
      if (iech[j] != 0 || lbs[i] != lbs[j]) continue;
      double s = orc_bvmv(ns, ks, g->sb, w->pbs + (size_t)ns * i, w->pbs + (size_t)ns * j);
      if (fabs(s) < eps_ovl) s = 0.0;
      if (i == j) s = 1.0;
      w->obs[i + nbf * j] = s;
      w->obs[j + nbf * i] = s;
    }
  }
  size_t nij = (size_t)nch * (nch + 1) / 2;
  w->hcc = (double*)calloc(nij * ns * ns, sizeof(double));
  w->hcb = (double*)calloc((size_t)ns * nch * (ncp > 0 ? ncp : 1), sizeof(double));
  w->hbb = (double*)calloc((size_t)(ncp > 0 ? ncp * (ncp + 1) / 2 : 1), sizeof(double));
  w->acf = (double*)calloc((size_t)nch * nch * (mk + 1), sizeof(double));
  w->htarg = (double*)calloc(nij, sizeof(double));
  w->otarg = (double*)calloc(nij, sizeof(double));
  w->qtab = (double*)calloc((size_t)nbf * nbf * (mk + 1), sizeof(double));
  w->qok = (int*)calloc((size_t)nbf * nbf * (mk + 1), sizeof(int));
  return w;
}

void orc_wave_free(orc_wave* w) {
  if (!w) return;
  free(w->pbs); free(w->qbs); free(w->obs); free(w->iech); free(w->lbs); free(w->ibort);
  free(w->hcc); free(w->hcb); free(w->hbb); free(w->acf); free(w->htarg); free(w->otarg);
  free(w->qtab); free(w->qok); free(w->hl_full); free(w->lint); free(w->lvec); free(w);
}

void orc_wave_get(const orc_wave* w, double* hcc, double* hcb, double* hbb, double* acf,
                  double* htarg, double* otarg) {
  size_t nij = (size_t)w->nch * (w->nch + 1) / 2;
  if (hcc) memcpy(hcc, w->hcc, sizeof(double) * nij * w->ns * w->ns);
  if (hcb && w->ncp > 0) memcpy(hcb, w->hcb, sizeof(double) * w->ns * w->nch * w->ncp);
  if (hbb && w->ncp > 0) memcpy(hbb, w->hbb, sizeof(double) * w->ncp * (w->ncp + 1) / 2);
  if (acf) memcpy(acf, w->acf, sizeof(double) * w->nch * w->nch * (w->mk + 1));
  if (htarg) memcpy(htarg, w->htarg, sizeof(double) * nij);
  if (otarg) memcpy(otarg, w->otarg, sizeof(double) * nij);
}

void orc_wave_qbs(const orc_wave* w, double* qbs, double* obs) {
  if (qbs) memcpy(qbs, w->qbs, sizeof(double) * w->ns * w->nbf);
  if (obs) memcpy(obs, w->obs, sizeof(double) * w->nbf * w->nbf);
}

static size_t blk(int ich, int jch) {
  int i = ich > jch ? ich : jch, j = ich > jch ? jch : ich;
  return (size_t)i * (i - 1) / 2 + j - 1;
}

/* UPDATE_HX, mod_intmat.f90:69-127 (sym 's' and 'x') */
static int update_hx(orc_wave* w, int ich, int jch, const double* d, char sym) {
  int ns = w->ns, ks = w->ks;
  if (ich < 1 || ich > w->nch || jch < 1 || jch > w->nch)
    return fail(-10, "UPDATE_HX: channel index out of range");
  double* x = (double*)calloc((size_t)ns * ns, sizeof(double));
  if (sym == 's') {
    for (int ith = 0; ith < ks; ith++)
      for (int i = 0; i < ns - ith; i++) {
        int j = i + ith;
        x[i + ns * j] = d[i + ns * ith];
        x[j + ns * i] = d[i + ns * ith];
      }
  } else {
    memcpy(x, d, sizeof(double) * ns * ns);
  }
  double* h = w->hcc + blk(ich, jch) * ns * ns;
  if (ich >= jch)
    for (size_t q = 0; q < (size_t)ns * ns; q++) h[q] += x[q];
  else
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) h[i + ns * j] += x[j + ns * i];
  free(x);
  return 0;
}

/* UPDATE_HL with the Gram band, mod_intmat.f90:132-160 */
int orc_wave_update_hl(orc_wave* w, int ich, int jch, double c) {
  int ns = w->ns, ks = w->ks;
  if (ich < 1 || ich > w->nch || jch < 1 || jch > w->nch)
    return fail(-11, "UPDATE_HL: channel index out of range");
  double* h = w->hcc + blk(ich, jch) * ns * ns;
  const double* sb = w->g->sb;
  for (int ith = 1; ith <= ks; ith++)
    for (int i = ks + 1 - ith; i <= ns; i++) {
      int j = i + ith - ks;
      double s = c * sb[(i - 1) + ns * (ith - 1)];
      h[(i - 1) + ns * (j - 1)] += s;
      if (i != j) h[(j - 1) + ns * (i - 1)] += s;
    }
  return 0;
}

/* UPDATE_HB, mod_intmat.f90:165-180 */
static int update_hb(orc_wave* w, int ic, int jc, double c) {
  if (ic > w->ncp || jc > w->ncp || ic <= 0 || jc <= 0)
    return fail(-12, "UPDATE_HB: indeces out of range");
  int i = ic > jc ? ic : jc, j = ic > jc ? jc : ic;
  w->hbb[(size_t)i * (i - 1) / 2 + j - 1] += c;
  return 0;
}

/* UPDATE_HV, mod_intmat.f90:185-204 */
static int update_hv(orc_wave* w, int ich, int ic, const double* v, double c) {
  if (ich < 1 || ich > w->nch) return fail(-13, "UPDATE_HV: channel index out of range");
  if (ic < 1 || ic > w->ncp) return fail(-14, "UPDATE_HV: bound index out of range");
  double* h = w->hcb + (size_t)(ic - 1) * w->ns * w->nch + (size_t)(ich - 1) * w->ns;
  for (int i = 0; i < w->ns; i++) h[i] += c * v[i];
  return 0;
}

/* UPDATE_HW, mod_intmat.f90:209-240 */
static int update_hw(orc_wave* w, int ich, int jch, const double* v, const double* u) {
  int ns = w->ns;
  if (ich < 1 || ich > w->nch || jch < 1 || jch > w->nch)
    return fail(-15, "UPDATE_HW: channel index out of range");
  double* h = w->hcc + blk(ich, jch) * ns * ns;
  if (ich >= jch) {
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) h[i + ns * j] += v[i] * u[j];
  } else {
    for (int j = 0; j < ns; j++)
      for (int i = 0; i < ns; i++) h[i + ns * j] += v[j] * u[i];
  }
  return 0;
}

/* UPDATE_CF, mod_intmat.f90:245-264 */
static int update_cf(orc_wave* w, int k, int ich, int jch, double c) {
  if (ich < 1 || ich > w->nch || jch < 1 || jch > w->nch)
    return fail(-16, "UPDATE_CF: channel index out of range");
  if (k < 0 || k > w->mk) return fail(-17, "UPDATE_CF: multipole index out of range");
  w->acf[(size_t)k * w->nch * w->nch + (ich - 1) + (size_t)w->nch * (jch - 1)] += c;
  return 0;
}

/* Target_h, mod_intmat.f90:269-293 */
static int target_h(orc_wave* w, int ich, int jch, double c, double cc) {
  if (ich < 1 || ich > w->nch || jch < 1 || jch > w->nch)
    return fail(-16, "UPDATE_CF: channel index out of range");
  size_t ij = blk(ich, jch);
  w->htarg[ij] += c;
  w->otarg[ij] += cc;
  return 0;
}

/* GET_V, get_V.f90:2-27: v = qbs(:,jj) minus overlap projections */
static int get_v(const orc_wave* w, int ii, int jj, double* v) {
  int ns = w->ns, nbf = w->nbf;
  if (ii < 1 || ii > nbf || jj < 1 || jj > nbf) return fail(-18, "get_V: orbital index out of range");
  if (w->ibort[(ii - 1) + nbf * (jj - 1)] == 0) return fail(-19, "get_V: <kl|nl> = 0 ?");
  memcpy(v, w->qbs + (size_t)ns * (jj - 1), sizeof(double) * ns);
  for (int i = 1; i <= nbf; i++) {
    if (i == jj) continue;
    if (w->iech[i - 1] != 0) continue;
    if (w->lbs[i - 1] != w->lbs[ii - 1]) continue;
    if (w->ibort[(i - 1) + nbf * (ii - 1)] != 0) continue;
    double o = w->obs[(i - 1) + nbf * (jj - 1)];
    if (fabs(o) < w->eps_ovl) continue;
    for (int r = 0; r < ns; r++) v[r] -= o * w->qbs[(size_t)ns * (i - 1) + r];
  }
  return 0;
}

static double quadr_cached(orc_wave* w, int i, int j, int m) {
  size_t q = (size_t)(i - 1) + w->nbf * ((size_t)(j - 1) + (size_t)w->nbf * m);
  if (m >= 0 && m <= w->mk) {
    if (!w->qok[q]) {
      w->qtab[q] = orc_quadr(w->g, w->pbs + (size_t)w->ns * (i - 1), w->pbs + (size_t)w->ns * (j - 1), m);
      w->qok[q] = 1;
    }
    return w->qtab[q];
  }
  return orc_quadr(w->g, w->pbs + (size_t)w->ns * (i - 1), w->pbs + (size_t)w->ns * (j - 1), m);
}

static const int IJCASE[9] = {2, 1, 2, 1, 2, 1, 2, 3, 4}; /* mod_cmdata.f90:54 */

/* I_data(jtype,jpol) for R^k (icase 5), src/bsr/BSR_MAT3/I_data.f90:2-224.
 * Records are visited in the order ipt (1-based indices into k1..cdata).
 * coupling='LS', iitar=0 (Target_h only for <i|j> structures). */
int orc_wave_idata(orc_wave* w, int icase, int jtype, int jpol, int nc, const int* ipt,
                   const int* K1, const int* K2, const int* K3, const int* K4, const double* cd) {
  int ns = w->ns, ks = w->ks, rc = 0;
  if (nc == 0) return 0;
  if (icase != 5) return fail(-20, "unknown case in I_data");
  if (jtype < 1 || jtype > 9) return fail(-21, " I_data: unknown itype ");
  mrk_cell_internal(w->g, jpol);
  const double* rkb = w->g->rkb;
  int jcase = IJCASE[jtype - 1];
  char sym_i = 's', sym_j = 's', sym_d, sym_r;
  if (jcase == 1) { sym_d = sym_j; sym_r = sym_i; }
  else if (jcase == 2) { sym_d = sym_i; sym_r = sym_j; }
  else { sym_d = 'x'; sym_r = 'x'; }
  double* dd = (double*)calloc((size_t)ns * ns, sizeof(double));
  double* xx = (double*)calloc((size_t)ns * ns, sizeof(double));
  double* v = (double*)calloc(ns, sizeof(double));
  double* u = (double*)calloc(ns, sizeof(double));
  const double* p = w->pbs;
#define P(n) (p + (size_t)ns * ((n) - 1))
  if (jtype == 1) {
    int iii = 0, jjj = 0;
    double S = 0.0;
    for (int q = 0; q < nc && rc == 0; q++) {
      int i = ipt[q] - 1;
      if (K1[i] != iii) {
        int i1 = K1[i] / IBI, i2 = K1[i] % IBI;
        orc_density(ns, ks, dd, P(i1), P(i2), sym_d);
        orc_convol(ns, ks, xx, dd, jcase, sym_i, sym_j, rkb);
      }
      if (K1[i] != iii || K2[i] != jjj) {
        int j1 = K2[i] / IBI, j2 = K2[i] % IBI;
        orc_density(ns, ks, dd, P(j1), P(j2), sym_r);
        S = orc_sum_amb(ns, ks, xx, dd, sym_r);
      }
      iii = K1[i]; jjj = K2[i];
      double C = S * cd[i];
      int ic = -K3[i], jc = -K4[i], io = -ic, jo = -jc;
      if (ic > 0 && jc > 0) {
        rc = update_hb(w, ic, jc, C);
      } else if (ic < 0 && jc > 0) {
        int i1 = io / IBO, i2 = io % IBO;
        rc = get_v(w, i1, i2, v);
        if (!rc) rc = update_hv(w, w->iech[i1 - 1], jc, v, C);
      } else if (ic < 0 && jc < 0) {
        int i1 = io / IBO, i2 = io % IBO, j1 = jo / IBO, j2 = jo % IBO;
        rc = get_v(w, i1, i2, v);
        if (!rc) rc = get_v(w, j1, j2, u);
        if (!rc) {
          for (int r = 0; r < ns; r++) v[r] *= C;
          rc = update_hw(w, w->iech[i1 - 1], w->iech[j1 - 1], v, u);
        }
      } else if (ic < 0 && jc == 0) {
        int i1 = io / IBO, i2 = io % IBO;
        rc = target_h(w, w->iech[i1 - 1], w->iech[i2 - 1], C, 0.0);
      } else {
        rc = fail(-22, " I_data: unknown structure for itype=1");
      }
    }
  } else if (jtype >= 2 && jtype <= 5) {
    int jjj = 0, iii = 0;
    for (int q = 0; q < nc && rc == 0; q++) {
      int i = ipt[q] - 1;
      int jj = K1[i], ii = K2[i], ich = K3[i], io = K4[i];
      if (jj != jjj || ii != iii) {
        if (jj != jjj) {
          int j1 = jj / IBI, j2 = jj % IBI;
          orc_density(ns, ks, dd, P(j1), P(j2), sym_d);
          orc_convol(ns, ks, xx, dd, jcase, sym_i, sym_j, rkb);
          jjj = jj;
        }
        char sym_v = (icase > 2) ? 'r' : 'l';
        orc_bav(ns, ks, xx, P(ii), v, sym_r, sym_v);
        iii = ii;
      }
      if (io > 0) {
        int j1 = io / IBO, j2 = io % IBO;
        rc = get_v(w, j1, j2, u);
        if (!rc) {
          for (int r = 0; r < ns; r++) u[r] *= cd[i];
          rc = update_hw(w, ich, w->iech[j1 - 1], v, u);
        }
      } else if (io < 0) {
        rc = update_hv(w, ich, -io, v, cd[i]);
      } else {
        rc = fail(-23, " I_data: uknown structure for itype 2-5 ");
      }
    }
  } else {
    memset(xx, 0, sizeof(double) * ns * ns);
    double CA = 0.0;
    for (int q = 0; q < nc && rc == 0; q++) {
      int i = ipt[q] - 1;
      int i1 = K2[i], j1 = K3[i];
      orc_density(ns, ks, dd, P(i1), P(j1), sym_d);
      for (size_t r = 0; r < (size_t)ns * ns; r++) xx[r] += cd[i] * dd[r];
      CA += cd[i] * quadr_cached(w, i1, j1, jpol);
      if (q < nc - 1 && K1[i] == K1[ipt[q + 1] - 1]) continue;
      orc_convol(ns, ks, dd, xx, jcase, sym_i, sym_j, rkb);
      int ich = K1[i] / IBI, jch = K1[i] % IBI;
      rc = update_hx(w, ich, jch, dd, sym_r);
      if (!rc && icase == 5 && jcase <= 2) rc = update_cf(w, jpol, ich, jch, CA);
      memset(xx, 0, sizeof(double) * ns * ns);
      CA = 0.0;
    }
  }
#undef P
  free(dd); free(xx); free(v); free(u);
  return rc;
}

/* O_data, src/bsr/BSR_MAT3/O_data.f90:2-76 (iitar=0) */
int orc_wave_odata(orc_wave* w, int nc, const int* ipt, const int* K3, const int* K4, const double* cd) {
  int ns = w->ns, rc = 0;
  double* v = (double*)calloc(ns, sizeof(double));
  double* u = (double*)calloc(ns, sizeof(double));
  for (int q = 0; q < nc && rc == 0; q++) {
    int i = ipt[q] - 1;
    int ic = -K3[i], jc = -K4[i];
    double C = cd[i];
    if (fabs(C) < w->eps_c) continue;
    if (ic > 0 && jc > 0) {
      rc = update_hb(w, ic, jc, C);
    } else if (ic < 0 && jc > 0) {
      int io = -ic, i1 = io / IBO, i2 = io % IBO;
      rc = get_v(w, i1, i2, v);
      if (!rc) rc = update_hv(w, w->iech[i1 - 1], jc, v, C);
    } else if (ic < 0 && jc < 0) {
      int io = -ic, i1 = io / IBO, i2 = io % IBO, jo = -jc, j1 = jo / IBO, j2 = jo % IBO;
      rc = get_v(w, i1, i2, v);
      if (!rc) rc = get_v(w, j1, j2, u);
      if (!rc) {
        for (int r = 0; r < ns; r++) v[r] *= C;
        rc = update_hw(w, w->iech[i1 - 1], w->iech[j1 - 1], v, u);
      }
    } else if (ic < 0 && jc == 0) {
      int io = -ic, i1 = io / IBO, i2 = io % IBO;
      rc = target_h(w, w->iech[i1 - 1], w->iech[i2 - 1], 0.0, C);
    } else {
      rc = fail(-24, " O_data: unknown structure ");
    }
  }
  free(v); free(u);
  return rc;
}

/* diagonal-block symmetrisation (C + C^T)/2, sub1.f90:80-84 and :187-191 */
void orc_wave_symmetrize_diag(orc_wave* w) {
  int ns = w->ns;
  for (int ich = 1; ich <= w->nch; ich++) {
    double* h = w->hcc + blk(ich, ich) * ns * ns;
    for (int i = 0; i < ns; i++)
      for (int j = 0; j <= i; j++) {
        double c = h[i + ns * j] + h[j + ns * i];
        h[i + ns * j] = c / 2.0;
        h[j + ns * i] = c / 2.0;
      }
  }
}

/* core-energy shift of the overlap blocks, sub1.f90:121-122 */
void orc_wave_scale(orc_wave* w, double ec) {
  size_t n = (size_t)w->nch * (w->nch + 1) / 2 * w->ns * w->ns;
  for (size_t q = 0; q < n; q++) w->hcc[q] *= ec;
  if (w->ncp > 0) {
    for (size_t q = 0; q < (size_t)w->ns * w->nch * w->ncp; q++) w->hcb[q] *= ec;
    for (size_t q = 0; q < (size_t)w->ncp * (w->ncp + 1) / 2; q++) w->hbb[q] *= ec;
  }
}

/* ACF symmetrisation and core screening, sub1.f90:211-219 */
void orc_wave_acf_finish(orc_wave* w, int kcore) {
  int nch = w->nch;
  const double eps_acf = 1e-5;
  for (int k = 0; k <= w->mk; k++)
    for (int i = 0; i < nch; i++)
      for (int j = i; j < nch; j++) {
        double* a = w->acf + (size_t)k * nch * nch;
        double c = a[i + nch * j] + a[j + nch * i];
        if (i != j) c *= 2.0;
        if (fabs(c) < eps_acf) c = 0.0;
        a[i + nch * j] = c;
        a[j + nch * i] = c;
      }
  for (int i = 0; i < nch; i++) w->acf[i + (size_t)nch * i] += kcore;
}

/* Exported accumulators, for tests against the reference's mod_intmat.f90 */
int orc_wave_update_hx(orc_wave* w, int ich, int jch, const double* d, char sym) {
  return update_hx(w, ich, jch, d, sym);
}
int orc_wave_update_hw(orc_wave* w, int ich, int jch, const double* v, const double* u) {
  return update_hw(w, ich, jch, v, u);
}
int orc_wave_update_hv(orc_wave* w, int ich, int ic, const double* v, double c) {
  return update_hv(w, ich, ic, v, c);
}
int orc_wave_update_hb(orc_wave* w, int ic, int jc, double c) { return update_hb(w, ic, jc, c); }
int orc_wave_update_cf(orc_wave* w, int k, int ich, int jch, double c) { return update_cf(w, k, ich, jch, c); }

/* ------------------------------------------------------------------ */
/* One-electron operator with the core and the L-integral lists        */
/* (BSR_MAT3 l_core.f90, L_data.f90), non-relativistic                 */
/* ------------------------------------------------------------------ */

/* Derivative of the order-k B-splines nonzero on interval left from a
 * quantity f of the order-(k-1) ones (f[r] belongs to spline left-k+2+r):
 * D[r] = (k-1) [f_i/(t_{i+k-1}-t_i) - f_{i+1}/(t_{i+k}-t_{i+1})] for the
 * order-k spline i = left-k+1+r (1-based knots), a term with a zero knot
 * span being zero. */
static void bderiv_step(const double* t, int k, int left, const double* f, double* D) {
  for (int r = 0; r < k; r++) {
    const int i = left - k + 1 + r; /* 1-based */
    double v = 0.0;
    if (r >= 1) {
      const double h = t[i + k - 2] - t[i - 1];
      if (h > 0.0) v += f[r - 1] / h;
    }
    if (r <= k - 2) {
      const double h = t[i + k - 1] - t[i];
      if (h > 0.0) v -= f[r] / h;
    }
    D[r] = (k - 1) * v;
  }
}

/* Values, first and second derivatives of the ks B-splines nonzero on knot
 * interval left at x (the quantities vbsplvd returns for nderiv = 3,
 * src/libs/BS/vbsplvd.f90:3-171), from the values of orders ks-2 and ks-1
 * and the derivative formula applied once and twice. */
static void bderivs(const double* t, int ks, int left, double x, double* b0, double* b1, double* b2) {
  double v1[32], v2[32], d1[32];
  bvalues(t, ks, left, x, b0);
  bvalues(t, ks - 1, left, x, v1);
  bvalues(t, ks - 2, left, x, v2);
  bderiv_step(t, ks, left, v1, b1);
  bderiv_step(t, ks - 1, left, v2, d1); /* first derivatives of order ks-1 */
  bderiv_step(t, ks, left, d1, b2);
}

/* First and second derivatives at the Gauss points, [(iv*ks+m)*ks+ith]. */
void orc_grid_bspd(const orc_grid* g, double* d1, double* d2) {
  int ks = g->ks, nv = g->nv;
  for (int iv = 0; iv < nv; iv++)
    for (int m = 0; m < ks; m++) {
      double b0[32], b1[32], b2[32];
      bderivs(g->t, ks, iv + ks, g->gr[iv * ks + m], b0, b1, b2);
      for (int ith = 0; ith < ks; ith++) {
        if (d1) d1[(iv * ks + m) * ks + ith] = b1[ith];
        if (d2) d2[(iv * ks + m) * ks + ith] = b2[ith];
      }
    }
}

/* <B_i|B''_j> in lower band storage (ns,ks), element (i,jp) for
 * j = i+jp-ks, plus in the unused slot (1,1) the element
 * <B_{ns-1}|B''_ns> of the last interval (mdb2,
 * src/libs/BS/define_spline.f90:216-245). */
void orc_db2(const orc_grid* g, double* db2) {
  int ks = g->ks, ns = g->ns, nv = g->nv;
  double* d2 = (double*)malloc(sizeof(double) * nv * ks * ks);
  orc_grid_bspd(g, NULL, d2);
  memset(db2, 0, sizeof(double) * ns * ks);
  for (int ith = 0; ith < ks; ith++)
    for (int jth = 0; jth <= ith; jth++) {
      int jcol = jth - ith + ks - 1;
      for (int iv = 0; iv < nv; iv++) {
        double s = 0.0;
        for (int m = 0; m < ks; m++) {
          size_t q = (size_t)iv * ks + m;
          s += g->grw[q] * g->bsp[q * ks + ith] * d2[q * ks + jth];
        }
        db2[(iv + ith) + (size_t)ns * jcol] += s;
      }
    }
  {
    double s = 0.0;
    for (int m = 0; m < ks; m++) {
      size_t q = (size_t)(nv - 1) * ks + m;
      s += g->grw[q] * g->bsp[q * ks + ks - 2] * d2[q * ks + ks - 1];
    }
    db2[0] += s;
  }
  free(d2);
}

/* ZCB(k1,k2,k3) = (3j symbol (k1 k2 k3; 0 0 0))^2 from the closed form
 * with 2g = k1+k2+k3 even:
 * [(2g-2k1)!(2g-2k2)!(2g-2k3)!/(2g+1)!] [g!/((g-k1)!(g-k2)!(g-k3)!)]^2,
 * zero when 2g is odd or a triangle condition fails. */
double orc_zcb(int k1, int k2, int k3) {
  int J = k1 + k2 + k3;
  if (J % 2) return 0.0;
  int g = J / 2;
  if (g - k1 < 0 || g - k2 < 0 || g - k3 < 0) return 0.0;
  long double f[128];
  f[0] = 1.0L;
  for (int i = 1; i < 128; i++) f[i] = f[i - 1] * i;
  if (J + 1 >= 128) return 0.0;
  long double a = f[J - 2 * k1] * f[J - 2 * k2] * f[J - 2 * k3] / f[J + 1];
  long double b = f[g] / (f[g - k1] * f[g - k2] * f[g - k3]);
  return (double)(a * b * b);
}

/* Gen_Lval with INT_core for every l present (l_core.f90:92-170,
 * :173-257), non-relativistic: hl = db2 - l(l+1) rm2 + 2z rm1 with the
 * Bloch term at the outer boundary, the direct core potential (k = 0) and
 * the core exchange (k up to 7) from the R^k band, the sum expanded into
 * the full hl_full(:,:,l); then L_int and L_vec of the bound orbitals
 * beyond the core.  Orbitals 1..kclosd are the closed core. */
int orc_wave_gen_lval(orc_wave* w, double z, int kclosd) {
  orc_grid* g = w->g;
  int ns = w->ns, ks = w->ks, nbf = w->nbf;
  if (kclosd < 0 || kclosd > nbf) return fail(-60, "Gen_Lval: kclosd outside 0..nbf");
  for (int ip = 0; ip < kclosd; ip++)
    if (w->iech[ip] != 0) return fail(-61, "Gen_Lval: a core orbital is a channel orbital");
  int mlh = 0;
  for (int i = 0; i < nbf; i++) mlh = w->lbs[i] > mlh ? w->lbs[i] : mlh;
  size_t nss = (size_t)ns * ns;
  free(w->hl_full); free(w->lint); free(w->lvec);
  w->mlh = mlh;
  w->hl_full = (double*)calloc(nss * (mlh + 1), sizeof(double));
  w->lint = (double*)calloc((size_t)nbf * nbf, sizeof(double));
  w->lvec = (double*)calloc((size_t)ns * nbf, sizeof(double));
  double *db2 = (double*)malloc(sizeof(double) * ns * ks), *rm1 = (double*)malloc(sizeof(double) * ns * ks),
         *rm2 = (double*)malloc(sizeof(double) * ns * ks), *hl = (double*)malloc(sizeof(double) * ns * ks),
         *hd = (double*)calloc((size_t)ns * ks, sizeof(double)), *d = (double*)malloc(sizeof(double) * nss),
         *x = (double*)malloc(sizeof(double) * nss), *hc = (double*)malloc(sizeof(double) * nss);
  orc_db2(g, db2);
  orc_mrm(g, -1, rm1);
  orc_mrm(g, -2, rm2);
  int have_dir = 0, rc = 0;
  for (int l = 0; l <= mlh && !rc; l++) {
    int present = 0;
    for (int i = 0; i < nbf; i++) present |= w->lbs[i] == l;
    if (!present) continue;
    /* INT_core(l) */
    const double c1 = (double)l * (l + 1), c2 = 2.0 * z;
    for (size_t q = 0; q < (size_t)ns * ks; q++) hl[q] = db2[q] - c1 * rm2[q] + c2 * rm1[q];
    {
      const double cb = db2[0] - db2[(ns - 1) + (size_t)ns * (ks - 2)];
      hl[(ns - 1) + (size_t)ns * (ks - 2)] += cb;
      hl[(ns - 1) + (size_t)ns * (ks - 1)] -= cb;
    }
    memset(hc, 0, sizeof(double) * nss);
    if (kclosd > 0) {
      if (!have_dir) {
        mrk_cell_internal(g, 0);
        for (int ip = 0; ip < kclosd; ip++) {
          const double cc = -2.0 * (4 * w->lbs[ip] + 2);
          const double* p = w->pbs + (size_t)ns * ip;
          orc_density(ns, ks, d, p, p, 's');
          if ((rc = orc_convol(ns, ks, x, d, 2, 's', 's', g->rkb))) break;
          for (size_t q = 0; q < (size_t)ns * ks; q++) hd[q] += cc * x[q];
        }
        have_dir = 1;
      }
      int kmin = 1000, kmax = 0;
      for (int ip = 0; ip < kclosd; ip++) {
        int k = abs(l - w->lbs[ip]);
        if (k < kmin) kmin = k;
        k = l + w->lbs[ip];
        if (k > kmax) kmax = k;
      }
      if (kmax > 7) kmax = 7;
      for (int k = kmin; k <= kmax && !rc; k++) {
        mrk_cell_internal(g, k);
        for (int ip = 0; ip < kclosd; ip++) {
          const double cc = (4 * w->lbs[ip] + 2) * orc_zcb(l, k, w->lbs[ip]);
          if (cc == 0.0) continue;
          const double* p = w->pbs + (size_t)ns * ip;
          orc_density(ns, ks, d, p, p, 'x');
          if ((rc = orc_convol(ns, ks, x, d, 3, 's', 's', g->rkb))) break;
          for (size_t q = 0; q < nss; q++) hc[q] += cc * x[q];
        }
      }
    }
    for (int jp = 1; jp <= ks; jp++)
      for (int i = 1 - jp + ks; i <= ns; i++) {
        int j = i + jp - ks;
        hc[(i - 1) + (size_t)ns * (j - 1)] += hl[(i - 1) + (size_t)ns * (jp - 1)];
        hc[(j - 1) + (size_t)ns * (i - 1)] = hc[(i - 1) + (size_t)ns * (j - 1)];
      }
    if (kclosd > 0)
      for (int jp = 1; jp <= ks; jp++)
        for (int i = 1; i <= ns - jp + 1; i++) {
          int j = i + jp - 1;
          hc[(i - 1) + (size_t)ns * (j - 1)] += hd[(i - 1) + (size_t)ns * (jp - 1)];
          hc[(j - 1) + (size_t)ns * (i - 1)] = hc[(i - 1) + (size_t)ns * (j - 1)];
        }
    /* L_int, L_vec of the bound orbitals with this l beyond the core */
    for (int i = kclosd; i < nbf; i++) {
      if (w->lbs[i] != l || w->iech[i] != 0) continue;
      const double* pi = w->pbs + (size_t)ns * i;
      for (int j = i; j < nbf; j++) {
        if (w->lbs[j] != l || w->iech[j] != 0) continue;
        const double* pj = w->pbs + (size_t)ns * j;
        double c = 0.0;
        for (int m = 0; m < ns; m++)
          for (int k = 0; k < m; k++) c += hc[m + (size_t)ns * k] * (pi[m] * pj[k] + pi[k] * pj[m]);
        for (int m = 0; m < ns; m++) c += hc[m + (size_t)ns * m] * pi[m] * pj[m];
        w->lint[i + (size_t)nbf * j] = c;
        w->lint[j + (size_t)nbf * i] = c;
      }
      for (int m = 0; m < ns; m++) {
        double s = 0.0;
        for (int k = 0; k < ns; k++) s += hc[k + (size_t)ns * m] * pi[k];
        w->lvec[m + (size_t)ns * i] = s;
      }
    }
    memcpy(w->hl_full + nss * l, hc, sizeof(double) * nss);
  }
  free(db2); free(rm1); free(rm2); free(hl); free(hd); free(d); free(x); free(hc);
  w->lok = rc == 0;
  return rc;
}

int orc_wave_lval_mlh(const orc_wave* w) { return w->lok ? w->mlh : -1; }

void orc_wave_lval_get(const orc_wave* w, double* hl_full, double* lint, double* lvec) {
  if (!w->lok) return;
  size_t nss = (size_t)w->ns * w->ns;
  if (hl_full) memcpy(hl_full, w->hl_full, sizeof(double) * nss * (w->mlh + 1));
  if (lint) memcpy(lint, w->lint, sizeof(double) * w->nbf * w->nbf);
  if (lvec) memcpy(lvec, w->lvec, sizeof(double) * w->ns * w->nbf);
}

/* L_data(jtype,l), src/bsr/BSR_MAT3/L_data.f90:2-149, over the records in
 * the order ipt (1-based); iitar as the module bsr_mat holds it
 * (mod_param.f90:33), records with |c| < eps_c skipped. */
int orc_wave_ldata(orc_wave* w, int jtype, int l, int nc, const int* ipt, const int* K1, const int* K2,
                   const int* K3, const int* K4, const double* cd, int iitar, double eps_c) {
  int ns = w->ns, rc = 0;
  if (!w->lok) return fail(-62, "L_data: Gen_Lval has not run");
  if (l < 0 || l > w->mlh) return fail(-63, "L_data: l outside 0..max l");
  if (jtype < 1 || jtype > 9) return fail(-64, " L_data: uknown itype ");
  const double* hc = w->hl_full + (size_t)ns * ns * l;
  double *v = (double*)malloc(sizeof(double) * ns), *u = (double*)malloc(sizeof(double) * ns),
         *xx = (double*)malloc(sizeof(double) * ns * ns);
  for (int q = 0; q < nc && !rc; q++) {
    int i = ipt ? ipt[q] - 1 : q;
    double c = cd[i];
    if (fabs(c) < eps_c) continue;
    if (jtype == 1) {
      int i1 = K1[i] / IBI, i2 = K1[i] % IBI;
      if (i1 < 1 || i1 > w->nbf || i2 < 1 || i2 > w->nbf) { rc = fail(-30, "orbital index out of range"); break; }
      double C = w->lint[(i1 - 1) + (size_t)w->nbf * (i2 - 1)] * c;
      int ic = -K3[i], jc = -K4[i];
      if (ic > 0 && jc > 0) {
        rc = update_hb(w, ic, jc, C);
      } else if (ic < 0 && jc > 0) {
        int io = -ic, j1 = io / IBO, j2 = io % IBO;
        if (!(rc = get_v(w, j1, j2, v))) rc = update_hv(w, w->iech[j1 - 1], jc, v, C);
      } else if (ic < 0 && jc < 0) {
        int io = -ic, a1 = io / IBO, a2 = io % IBO, jo = -jc, b1 = jo / IBO, b2 = jo % IBO;
        if (!(rc = get_v(w, a1, a2, v)) && !(rc = get_v(w, b1, b2, u))) {
          for (int r = 0; r < ns; r++) v[r] *= C;
          rc = update_hw(w, w->iech[a1 - 1], w->iech[b1 - 1], v, u);
        }
      } else if (ic < 0 && jc == 0) {
        int io = -ic, a1 = io / IBO, a2 = io % IBO;
        if (a1 < 1 || a1 > w->nbf || a2 < 1 || a2 > w->nbf) { rc = fail(-30, "orbital index out of range"); break; }
        int ich = w->iech[a1 - 1], jch = w->iech[a2 - 1];
        if (iitar > 0 && ich != jch) rc = orc_wave_update_hl(w, ich, jch, C);
        if (!rc) rc = target_h(w, ich, jch, C, 0.0);
      } else {
        rc = fail(-65, " L_data: unknown structure ");
      }
    } else if (jtype <= 5) {
      int o = K2[i], ich = K3[i], io = K4[i];
      if (o < 1 || o > w->nbf) { rc = fail(-30, "orbital index out of range"); break; }
      const double* lv = w->lvec + (size_t)ns * (o - 1);
      if (io > 0) {
        int j1 = io / IBO, j2 = io % IBO;
        if (!(rc = get_v(w, j1, j2, u))) {
          for (int r = 0; r < ns; r++) u[r] *= c;
          rc = update_hw(w, ich, w->iech[j1 - 1], lv, u);
        }
      } else if (io < 0) {
        rc = update_hv(w, ich, -io, lv, c);
      } else {
        rc = fail(-66, " L_data: uknown structure for itype 2-5 ");
      }
    } else {
      int ich = K1[i] / IBI, jch = K1[i] % IBI;
      if (ich == jch || iitar > 1) {
        for (size_t e = 0; e < (size_t)ns * ns; e++) xx[e] = c * hc[e];
        rc = update_hx(w, ich, jch, xx, 'x');
      } else {
        rc = target_h(w, ich, jch, 0.0, c);
      }
    }
  }
  free(v); free(u); free(xx);
  return rc;
}

/* ------------------------------------------------------------------ */
/* BS_ORTH, src/bsr/BSR_MAT3/bs_orth.f90:2-92: orthogonality of each    */
/* channel orbital to the bound orbitals of its l that IBORT does not  */
/* exempt, by H -> (1 - B zz') H (1 - zz' B) (Bentley).                */
/* ------------------------------------------------------------------ */
static void matmul_nn(int n, const double* a, const double* b, double* c) {
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) {
      double s = 0.0;
      for (int k = 0; k < n; k++) s += a[i + (size_t)n * k] * b[k + (size_t)n * j];
      c[i + (size_t)n * j] = s;
    }
}

int orc_wave_bs_orth(orc_wave* w) {
  int ns = w->ns, ks = w->ks, nbf = w->nbf, nch = w->nch;
  size_t nss = (size_t)ns * ns;
  double *B = (double*)calloc(nss, sizeof(double)), *zz = (double*)malloc(sizeof(double) * nss),
         *x = (double*)malloc(sizeof(double) * nss), *y = (double*)malloc(sizeof(double) * nss),
         *xx = (double*)malloc(sizeof(double) * nss), *z = (double*)malloc(sizeof(double) * ns);
  /* bb: the full Gram matrix from the lower band sb */
  for (int i = 0; i < ns; i++)
    for (int jp = 0; jp < ks; jp++) {
      int j = i + jp - (ks - 1);
      if (j < 0) continue;
      B[i + (size_t)ns * j] = w->g->sb[i + (size_t)ns * jp];
      B[j + (size_t)ns * i] = w->g->sb[i + (size_t)ns * jp];
    }
  int rc = 0;
  for (int ich = 1; ich <= nch && !rc; ich++) {
    int i = 0; /* ipch(ich): the channel orbital */
    for (int o = 1; o <= nbf; o++)
      if (w->iech[o - 1] == ich) { i = o; break; }
    if (!i) { rc = fail(-70, "BS_ORTH: channel without a channel orbital"); break; }
    memset(zz, 0, sizeof(double) * nss);
    int nort = 0;
    for (int j = 1; j <= nbf; j++) {
      if (w->iech[j - 1] > 0) continue;
      if (w->lbs[i - 1] != w->lbs[j - 1]) continue;
      int a = i > j ? i : j, b = i > j ? j : i;
      if (w->ibort[(a - 1) + (size_t)nbf * (b - 1)] != 0) continue;
      nort++;
      const double* p = w->pbs + (size_t)ns * (j - 1);
      for (int ii = 0; ii < ns; ii++)
        for (int jj = 0; jj < ns; jj++) zz[ii + (size_t)ns * jj] += p[ii] * p[jj];
    }
    if (nort == 0) continue;
    matmul_nn(ns, B, zz, x);
    matmul_nn(ns, zz, B, y);
    for (int jch = 1; jch <= nch; jch++) {
      double* h = w->hcc + blk(ich, jch) * nss;
      if (ich >= jch) {
        matmul_nn(ns, x, h, xx);
        for (size_t q = 0; q < nss; q++) h[q] -= xx[q];
      }
      if (ich <= jch) {
        matmul_nn(ns, h, y, xx);
        for (size_t q = 0; q < nss; q++) h[q] -= xx[q];
      }
    }
    for (int ip = 0; ip < w->ncp; ip++) {
      double* v = w->hcb + (size_t)ip * ns * nch + (size_t)(ich - 1) * ns;
      for (int c = 0; c < ns; c++) {
        double s = 0.0;
        for (int r = 0; r < ns; r++) s += v[r] * y[r + (size_t)ns * c];
        z[c] = s;
      }
      for (int c = 0; c < ns; c++) v[c] -= z[c];
    }
  }
  free(B); free(zz); free(x); free(y); free(xx); free(z);
  return rc;
}
