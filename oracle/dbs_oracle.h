/* dbs_oracle.h -- shared declarations of the Dirac B-spline oracle
 * (dbs_oracle.c: bases, moments, bands; dbsr_oracle.c: R_data assembly).
 * TEST INFRASTRUCTURE ONLY. */
#ifndef DBS_ORACLE_H
#define DBS_ORACLE_H
#include <stddef.h>

typedef struct {
  int ks, ns, nv, ksp, ksq, nsp, nsq;
  double *t, *tp, *tq, *gx, *gw, *gr, *grm, *grw, *pbsp, *qbsp;
} OrcDbs;

/* integer power by repeated squaring (Fortran real**integer) */
double orc_dpow(double x, int n);
void orc_dbs_mrk(const OrcDbs* d, int family, int k, double* rkb);

/* BS leaf routines of bsr_oracle.c (the DBS versions compute the same) */
void orc_density(int ns, int ks, double* d, const double* p1, const double* p2, char type);
int orc_convol(int ns, int ks, double* a, const double* d, int icase, char sym_i, char sym_j, const double* rkb);
double orc_sum_amb(int ns, int ks, const double* a, const double* b, char sym);
void orc_bav(int ns, int ks, const double* b, const double* v, double* y, char type, char side);
#endif
