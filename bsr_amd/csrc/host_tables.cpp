// Host-side grid set-up: knots, Gauss-Legendre nodes, B-spline values at
// the Gauss points, and <B_i|r^m|B_j> bands.  Cheap O(nv*ks^3) work done
// once per grid; results are uploaded and stay resident on the device.
#include "bsrg_internal.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace bsrg {

double powi(double a, int n) {
  // Same evaluation order as the integer power of the reference's
  // compiler runtime: square-and-multiply, reciprocal for n < 0.
  bool recip = n < 0;
  unsigned b = recip ? -n : n;
  double r = 1.0;
  for (;;) {
    if (b & 1u) r *= a;
    b >>= 1;
    if (!b) break;
    a *= a;
  }
  return recip ? 1.0 / r : r;
}

// Knot sequences.  kind=1: mkgrid1 (src/libs/BS/define_grid.f90:213-276),
// kind=0: mkgrid (:86-153).  Both build the knots in units of z*r and
// divide by z at the end.
int mkgrid(int kind, int ks, double z, double h, double hmax, double rmax, std::vector<double>& t,
           int& ns, int& ml, int& me, double& hout) {
  if (ks < 2 || ks > kMaxKs || z <= 0 || h <= 0 || rmax <= 0) return fail(-1, "mkgrid: bad arguments");
  if (kind == 1) {
    const double zr = rmax * z;
    double zh = hmax * z;
    const int cap = static_cast<int>((zr / h + 1 + ks) * 2) + 2 * ks + 4;
    std::vector<double> u(cap + 2, 0.0);  // u[n] == t(n), 1-based
    int n = ks;
    ml = static_cast<int>(1.0 / h + 0.5);
    h = 1.0 / ml;
    for (int i = 0; i < ml; ++i) { ++n; u[n] = u[n - 1] + h; }
    if (zh < h) zh = h;
    me = 0;
    for (;;) {
      if (n + 1 >= cap) return fail(-1, "mkgrid: knot overflow");
      u[n + 1] = u[n] * (1.0 + h);
      if (u[n + 1] - u[n] > zh) break;
      ++n;
      if (u[n] > zr) break;
      ++me;
    }
    if (u[n] < zr) {
      for (;;) {
        if (n + 1 >= cap) return fail(-1, "mkgrid: knot overflow");
        u[n + 1] = u[n] + zh;
        ++n;
        if (u[n] >= zr) break;
      }
    }
    if (u[n] > zr) {
      --n; --me;
      for (int i = n + 1; i <= n + ks; ++i) u[i] = zr;
      u[n] = 0.5 * (u[n + 1] + u[n - 1]);
    }
    ns = n;
    t.assign(ns + ks, 0.0);
    for (int i = 1; i <= ns + ks; ++i) t[i - 1] = u[i] / z;
    hout = h;
    return 0;
  }
  // kind 0
  const double h0 = h;
  ml = static_cast<int>(1.0 / h + 0.1);
  h = 1.0 / ml;
  const double hp1 = 1.0 + h, tmax = z * rmax, zh = z * hmax;
  double me1d = std::log(zh / h) / std::log(hp1) + 1.0;
  int me1 = static_cast<int>(me1d > 0 ? me1d : 0.0);
  int me2 = static_cast<int>(std::log(tmax) / std::log(hp1) + 1);
  int m = 0;
  double h2 = 0.0;
  if (me2 <= me1) { me = me2; }
  else {
    me = me1;
    const double tx = powi(hp1, me);
    h2 = h * tx / hp1;
    m = static_cast<int>(std::lround((tmax - tx) / h2));
  }
  const int nlast = ml + me + m + ks - 1;
  ns = nlast;
  std::vector<double> u(ns + ks + 2, 0.0);
  for (int i = ks + 1; i <= ks + ml; ++i) u[i] = u[i - 1] + h0;
  for (int i = ks + ml + 1; i <= ks + me + ml; ++i) u[i] = u[i - 1] * hp1;
  for (int i = ks + me + ml + 1; i <= nlast + 1; ++i) u[i] = u[i - 1] + h2;
  for (int i = nlast + 2; i <= ns + ks; ++i) u[i] = u[nlast + 1];
  t.assign(ns + ks, 0.0);
  for (int i = 1; i <= ns + ks; ++i) t[i - 1] = u[i] / z;
  hout = h;
  return 0;
}

// k-point Gauss-Legendre rule on [0,1], nodes ascending (Newton on P_k).
void gauss_legendre01(int k, double* x, double* w) {
  const double pi = 3.14159265358979323846;
  for (int i = 1; i <= k; ++i) {
    double s = std::cos(pi * (i - 0.25) / (k + 0.5));
    double dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = s;
      for (int n = 2; n <= k; ++n) {
        const double p2 = ((2.0 * n - 1.0) * s * p1 - (n - 1.0) * p0) / n;
        p0 = p1;
        p1 = p2;
      }
      dp = k * (s * p1 - p0) / (s * s - 1.0);
      const double ds = p1 / dp;
      s -= ds;
      if (std::fabs(ds) < 1e-16) break;
    }
    double p0 = 1.0, p1 = s;
    for (int n = 2; n <= k; ++n) {
      const double p2 = ((2.0 * n - 1.0) * s * p1 - (n - 1.0) * p0) / n;
      p0 = p1;
      p1 = p2;
    }
    dp = k * (s * p1 - p0) / (s * s - 1.0);
    x[i - 1] = 0.5 * (1.0 - s);
    w[i - 1] = 1.0 / ((1.0 - s * s) * dp * dp);
  }
}

// The ks B-splines that are nonzero on knot interval [t(left), t(left+1)]
// (1-based left) at x: de Boor's triangular recurrence.
void bspline_values(const double* t, int ks, int left, double x, double* b) {
  double dr[kMaxKs], dl[kMaxKs];
  b[0] = 1.0;
  for (int j = 1; j < ks; ++j) {
    dr[j - 1] = t[left + j - 1] - x;
    dl[j - 1] = x - t[left - j];
    double saved = 0.0;
    for (int r = 0; r < j; ++r) {
      const double term = b[r] / (dr[r] + dl[j - 1 - r]);
      b[r] = saved + dr[r] * term;
      saved = dl[j - 1 - r] * term;
    }
    b[j] = saved;
  }
}

// Gauss-Legendre rule of the Dirac B-spline library: the Numerical Recipes
// method (gauleg, src/libs/ZCOM/gauss_points.f90:2-51) -- Newton from the
// same first guess, stopped at |dz| <= 1e-14, the weight taken from the
// derivative of the last Newton step.  The DBS integrals are defined with
// exactly these nodes and weights.
void gauss_legendre_nr(int n, double x1, double x2, double* x, double* w) {
  const int m = (n + 1) / 2;
  const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
  for (int i = 1; i <= m; ++i) {
    double z = std::cos(3.141592654 * (i - 0.25) / (n + 0.5)), z1, pp;
    do {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      z1 = z;
      z = z1 - p1 / pp;
    } while (std::fabs(z - z1) > 1.0e-14);
    x[i - 1] = xm - xl * z;
    x[n - i] = xm + xl * z;
    w[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
    w[n - i] = w[i - 1];
  }
}

// P (q=0, order ksp) and Q (q=1, order ksq) bases on the knots of order ks
// (Def_pbsp / Def_qbsp, src/libs/DBS/DBS_gauss.f90:153-252): knots
// tb(1:kb) = t(1), tb(kb+1:nsb) = t(ks+1:ns), tb(nsb+1:nsb+kb) = t(ns+1:ns+kb),
// nsb = nv+kb-1; values at the ks Gauss points of every interval.
void DbsTables::build(int ks_, int ns_, int ksp, int ksq, const double* t_) {
  ks = ks_; ns = ns_; nv = ns - ks + 1;
  kb[0] = ksp; kb[1] = ksq;
  t.assign(t_, t_ + ns + ks);
  gx.resize(ks); gw.resize(ks);
  gauss_legendre_nr(ks, 0.0, 1.0, gx.data(), gw.data());
  gr.assign(static_cast<size_t>(nv) * ks, 0.0);
  grm.assign(gr.size(), 0.0);
  grw.assign(gr.size(), 0.0);
  for (int iv = 0; iv < nv; ++iv) {
    const double tl = t[iv + ks - 1], dt = t[iv + ks] - t[iv + ks - 1];
    for (int m = 0; m < ks; ++m) {
      gr[iv * ks + m] = dt * gx[m] + tl;
      grm[iv * ks + m] = 1.0 / gr[iv * ks + m];
      grw[iv * ks + m] = dt * gw[m];
    }
  }
  for (int q = 0; q < 2; ++q) {
    const int k = kb[q];
    nsb[q] = nv + k - 1;
    tb[q].assign(static_cast<size_t>(nsb[q]) + k, 0.0);
    for (int i = 0; i < k; ++i) tb[q][i] = t[0];
    for (int i = k; i < nsb[q]; ++i) tb[q][i] = t[i - k + ks];
    for (int i = 0; i < k; ++i) tb[q][nsb[q] + i] = t[ns + i];
    val[q].assign(static_cast<size_t>(nv) * ks * k, 0.0);
    for (int iv = 0; iv < nv; ++iv)
      for (int m = 0; m < ks; ++m)
        bspline_values(tb[q].data(), k, iv + k, gr[iv * ks + m], &val[q][(static_cast<size_t>(iv) * ks + m) * k]);
  }
}

void GridTables::build(int ks_, int ns_, int ml_, int me_, double h_, const double* t_) {
  ks = ks_; ns = ns_; nv = ns - ks + 1; ml = ml_; me = me_; h = h_;
  t.assign(t_, t_ + ns + ks);
  gx.resize(ks); gw.resize(ks);
  gauss_legendre01(ks, gx.data(), gw.data());
  gr.assign(static_cast<size_t>(nv) * ks, 0.0);
  grm.assign(gr.size(), 0.0);
  grw.assign(gr.size(), 0.0);
  bsp.assign(static_cast<size_t>(nv) * ks * ks, 0.0);
  for (int iv = 0; iv < nv; ++iv) {
    const double tl = t[iv + ks - 1], tr = t[iv + ks];
    for (int m = 0; m < ks; ++m) {
      const double x = (tr - tl) * gx[m] + tl;
      gr[iv * ks + m] = x;
      grm[iv * ks + m] = 1.0 / x;
      grw[iv * ks + m] = (tr - tl) * gw[m];
      bspline_values(t.data(), ks, iv + ks, x, &bsp[(static_cast<size_t>(iv) * ks + m) * ks]);
    }
  }
  sb = mrm(0);
}

// <B_i | r^mm | B_j> in lower band storage (ns,ks): element (i,jp) is
// <B_i|r^mm|B_{i+jp-ks}>.  Follows mrm, src/libs/BS/mrm.f90:2-91.
std::vector<double> GridTables::mrm(int mm) const {
  std::vector<double> rm(static_cast<size_t>(ns) * ks, 0.0);
  for (int ith = 0; ith < ks; ++ith)
    for (int jth = 0; jth <= ith; ++jth) {
      const int jcol = jth - ith + ks - 1;
      for (int iv = 0; iv < nv; ++iv) {
        double s = 0.0;
        for (int m = 0; m < ks; ++m) {
          const size_t q = static_cast<size_t>(iv) * ks + m;
          const double bi = grw[q] * bsp[q * ks + ith], bj = bsp[q * ks + jth];
          double f;
          if (mm == 0) f = bi * bj;
          else if (mm == -1) f = bi * grm[q] * bj;
          else if (mm <= -2) f = bi * powi(grm[q], -mm) * bj;
          else if (mm == 1) f = bi * gr[q] * bj;
          else f = bi * powi(gr[q], mm) * bj;
          s += f;
        }
        rm[(iv + ith) + static_cast<size_t>(ns) * jcol] += s;
      }
    }
  return rm;
}

// Second derivatives of the ks B-splines nonzero on knot interval left
// (1-based) at x.  Differentiating B_{i,k} = w_{i,k} B_{i,k-1} +
// (1 - w_{i+1,k}) B_{i+1,k-1} gives
// B'_{i,k} = (k-1) [B_{i,k-1}/(t_{i+k-1}-t_i) - B_{i+1,k-1}/(t_{i+k}-t_{i+1})];
// applied to the first derivatives of order ks-1 (themselves from the
// values of order ks-2) it gives the second derivatives of order ks.
namespace {
void spline_slope(const double* t, int k, int left, const double* f, double* out) {
  for (int r = 0; r < k; ++r) {
    const int i = left - k + 1 + r;  // 1-based index of the order-k spline
    const double hl = r >= 1 ? t[i + k - 2] - t[i - 1] : 0.0, hr = r <= k - 2 ? t[i + k - 1] - t[i] : 0.0;
    out[r] = (k - 1) * ((hl > 0.0 ? f[r - 1] / hl : 0.0) - (hr > 0.0 ? f[r] / hr : 0.0));
  }
}
}  // namespace

void bspline_second(const double* t, int ks, int left, double x, double* d2) {
  double v[kMaxKs], d1[kMaxKs];
  bspline_values(t, ks - 2, left, x, v);
  spline_slope(t, ks - 1, left, v, d1);
  spline_slope(t, ks, left, d1, d2);
}

// mdb2 (src/libs/BS/define_spline.f90:216-245): <B_i|B''_j> in the lower band
// (i, jp) -> j = i+jp-ks, plus <B_{ns-1}|B''_ns> in the slot (0,0) that the
// band does not use (the Bloch term of INT_core reads it).
std::vector<double> GridTables::db2() const {
  std::vector<double> out(static_cast<size_t>(ns) * ks, 0.0);
  std::vector<double> d2(static_cast<size_t>(nv) * ks * ks);
  for (int iv = 0; iv < nv; ++iv)
    for (int m = 0; m < ks; ++m)
      bspline_second(t.data(), ks, iv + ks, gr[iv * ks + m], &d2[(static_cast<size_t>(iv) * ks + m) * ks]);
  for (int ith = 0; ith < ks; ++ith)
    for (int jth = 0; jth <= ith; ++jth)
      for (int iv = 0; iv < nv; ++iv) {
        double s = 0.0;
        for (int m = 0; m < ks; ++m) {
          const size_t q = static_cast<size_t>(iv) * ks + m;
          s += grw[q] * bsp[q * ks + ith] * d2[q * ks + jth];
        }
        out[(iv + ith) + static_cast<size_t>(ns) * (jth - ith + ks - 1)] += s;
      }
  double s = 0.0;
  for (int m = 0; m < ks; ++m) {
    const size_t q = static_cast<size_t>(nv - 1) * ks + m;
    s += grw[q] * bsp[q * ks + ks - 2] * d2[q * ks + ks - 1];
  }
  out[0] += s;
  return out;
}

// <x|A|y> for a symmetric lower band; bvmv.f90:2-67.
double bvmv_sym(int n, int k, const double* a, const double* x, const double* y) {
  double s = 0.0;
  for (int jp = 1; jp <= k - 1; ++jp)
    for (int i = k + 1 - jp; i <= n; ++i) {
      const int j = i + jp - k;
      s += a[(i - 1) + static_cast<size_t>(n) * (jp - 1)] * (x[i - 1] * y[j - 1] + x[j - 1] * y[i - 1]);
    }
  double c = 0.0;
  for (int i = 0; i < n; ++i) c += a[i + static_cast<size_t>(n) * (k - 1)] * x[i] * y[i];
  return s + c;
}

// y = A v for a symmetric lower band; bxv.f90:2-43.
void bxv_sym(int k, int n, const double* b, const double* v, double* y) {
  for (int i = 0; i < n; ++i) y[i] = b[i + static_cast<size_t>(n) * (k - 1)] * v[i];
  for (int jp = 1; jp <= k - 1; ++jp)
    for (int i = k - jp + 1; i <= n; ++i) {
      const int j = i - k + jp;
      y[i - 1] += b[(i - 1) + static_cast<size_t>(n) * (jp - 1)] * v[j - 1];
      y[j - 1] += b[(i - 1) + static_cast<size_t>(n) * (jp - 1)] * v[i - 1];
    }
}

}  // namespace bsrg
