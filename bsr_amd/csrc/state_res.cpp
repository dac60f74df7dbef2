// The list builder of BSR_MAT3, host side: the int_bnk data bank (angular
// coefficients of the close-coupling expansion) read and expanded over the
// states of one partial wave into the coefficient lists of module cmdata
// that the R^k, L and overlap passes consume (State_res,
// src/bsr/BSR_MAT3/state_res.f90:2-161).
//
//   bank       Read_dets (mod_dets.f90:50-86) after the configuration
//              headers read_symc_LS / Read_symt_LS / Read_oper_LS
//              (ZCONFLS/MOD_symc.f90:175-195, MOD_symt.f90:224-336), then
//              the coefficient records (c, ijt, int, idf) to the end of the
//              file (state_res.f90:37-41); Fortran unformatted sequential
//   walk       for every record of the case: Decode_INT
//              (ZCONFLS/incode_int.f90:19-36), the state pairs of its two
//              symmetries (lower half when they coincide), expansion
//              coefficients WC, channels by Ifind_channel
//              (MOD_channel.f90:246-274), the orbitals of the integral,
//              Jsym_int (jsym_int.f90), the overlap factors by Det_fact
//              (det_fact.f90, check_det.f90, vdet.f90, ZCOM/det.f90), the
//              structure and key by Idef_itype (idef_type.f90)
//   lists      Add_coef (add_coef.f90) into blocks of mb records per
//              (kpol, itype), Add_cdata order within a block, a full set of
//              blocks generating the largest chain as Add_coef does, and
//              Gen_matrix's list step (gen_matrix.f90:17-47: one block as
//              is, several merged by Merge_CDATA) for every chain at the end
//              (state_res.f90:155-159).  Each generated list is a batch.
//
// The arithmetic follows the reference's expressions and order (products
// and sums unfused), so the coefficients agree with the reference bit for
// bit.  The J-dependent relativistic cases (icase 7-10, which read the
// term tables) are not built.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/bsr_gpu.h"

#pragma clang fp contract(off)

namespace bsrg {
int fail(int code, const std::string& msg);
}

namespace {

constexpr int kIbi = 1 << 15;   // packing of orbital and channel pairs (cmdata ibi)
constexpr int kIbc = 1 << 15;   // packing of symmetry pairs (param_LS ibc)
constexpr int kIdet = 1 << 15;  // determinant element packing (dets idet)
constexpr int kIdef = 16;       // overlap factor packing (dets idef)
constexpr int kNtype = 9;       // structures (cmdata ntype)

using Key = std::tuple<int, int, int, int>;

struct Expand {
  int itype, kpol, k1, k2, k3, k4;
  double c;
};

}  // namespace

struct bsrg_lb {
  // bank
  int ndet = 0, kdet = 0, ndef = 0, kdef = 0;
  std::vector<int> kpd, ipd, npd, kpf, ipf, npf;
  std::vector<double> cbuf;
  std::vector<int> ijtb, intb, idfb;
  // orbitals: iech (ief), l, IBORT and OBS (nbf x nbf, column-major)
  int nbf = 0, nelc = 0;
  std::vector<int> iech, lef, ibort;
  std::vector<double> obs;
  // channels and perturbers (module channel)
  int nch = 0, npert = 0;
  std::vector<int> ipconf, ippert;  // ipconf(1..nch), ippert(0..npert)
  // states (module conf_LS)
  std::vector<int> ip_state, no_state, ip_orb, it1, it2, ip_stat;
  std::vector<double> wc;
  // parameters (module bsr_mat)
  double eps_c = 1e-10, eps_det = 1e-10, eps_soo = 1e-3;
  int nb = 2000, mb = 5000, kb = 100;
  // results of the last run
  std::vector<Expand> expansion;
  struct Rec { int batch, jtype, jpol, k1, k2, k3, k4; double c; };
  std::vector<Rec> out;
  int nbatch = 0;
};

namespace {

// ---- the data bank -----------------------------------------------------------

struct Reader {
  FILE* f = nullptr;
  std::vector<char> rec;
  ~Reader() {
    if (f) fclose(f);
  }
  // next record; false at the end of the file
  bool next() {
    int32_t n1 = 0, n2 = 0;
    if (fread(&n1, 4, 1, f) != 1) return false;
    if (n1 < 0) return false;
    rec.resize(n1);
    if (n1 && fread(rec.data(), 1, n1, f) != static_cast<size_t>(n1)) return false;
    if (fread(&n2, 4, 1, f) != 1 || n2 != n1) return false;
    return true;
  }
  int ints(int i) const {
    int32_t v;
    std::memcpy(&v, rec.data() + 4 * i, 4);
    return v;
  }
  size_t nint() const { return rec.size() / 4; }
};

int bad_bank(const std::string& what) { return bsrg::fail(-40, "int_bnk: " + what); }

int read_bank(const char* path, bsrg_lb& L) {
  Reader r;
  r.f = fopen(path, "rb");
  if (!r.f) return bsrg::fail(-41, std::string("int_bnk: cannot open ") + path);
  // read_symc_LS: (nsymc, lsymc) and six arrays
  if (!r.next() || r.nint() < 2) return bad_bank("symc header");
  for (int i = 0; i < 6; ++i)
    if (!r.next()) return bad_bank("symc arrays");
  // Read_symt_LS: (nsymt, lsymt) and three arrays
  if (!r.next() || r.nint() < 2) return bad_bank("symt header");
  const int nsymt = r.ints(0);
  for (int i = 0; i < 3; ++i)
    if (!r.next()) return bad_bank("symt arrays");
  // Read_oper_LS: n and n records
  if (!r.next() || r.nint() < 1) return bad_bank("oper header");
  const int nop = r.ints(0);
  if (static_cast<long>(nsymt) * (nsymt + 1) / 2 < nop) return bad_bank("Read_oper_LS: nsymt too small");
  for (int i = 0; i < nop; ++i)
    if (!r.next()) return bad_bank("oper records");
  // Read_dets: (ndet, ldet, jdet), ndet records (kpd, ipd, jpd, NPD(ipd+1:ipd+kpd)),
  // then the same for the overlap factors unless ndet = 0
  if (!r.next() || r.nint() < 3) return bad_bank("det header");
  L.ndet = r.ints(0);
  L.kdet = r.ints(1);
  L.ndef = L.kdef = 0;
  L.kpd.assign(std::max(L.ndet, 0), 0);
  L.ipd.assign(std::max(L.ndet, 0), 0);
  L.npd.assign(std::max(L.kdet, 0), 0);
  L.kpf.clear(); L.ipf.clear(); L.npf.clear();
  auto read_list = [&](int n, int total, std::vector<int>& kp, std::vector<int>& ipp, std::vector<int>& np,
                       const char* what) {
    for (int i = 0; i < n; ++i) {
      if (!r.next() || r.nint() < 3) return bad_bank(what);
      const int k = r.ints(0), ip = r.ints(1);
      if (k < 0 || ip < 0 || ip + k > total || static_cast<int>(r.nint()) < 3 + k) return bad_bank(what);
      kp[i] = k;
      ipp[i] = ip;
      for (int q = 0; q < k; ++q) np[ip + q] = r.ints(3 + q);
    }
    return 0;
  };
  if (L.ndet > 0) {
    if (int rc = read_list(L.ndet, L.kdet, L.kpd, L.ipd, L.npd, "determinant records")) return rc;
    if (!r.next() || r.nint() < 3) return bad_bank("overlap factor header");
    L.ndef = r.ints(0);
    L.kdef = r.ints(1);
    L.kpf.assign(std::max(L.ndef, 0), 0);
    L.ipf.assign(std::max(L.ndef, 0), 0);
    L.npf.assign(std::max(L.kdef, 0), 0);
    if (L.ndef > 0)
      if (int rc = read_list(L.ndef, L.kdef, L.kpf, L.ipf, L.npf, "overlap factor records")) return rc;
  }
  // coefficient records to the end of the file
  L.cbuf.clear(); L.ijtb.clear(); L.intb.clear(); L.idfb.clear();
  while (r.next()) {
    if (r.rec.size() < 20) return bad_bank("coefficient record");
    double c;
    std::memcpy(&c, r.rec.data(), 8);
    L.cbuf.push_back(c);
    int32_t v[3];
    std::memcpy(v, r.rec.data() + 8, 12);
    L.ijtb.push_back(v[0]);
    L.intb.push_back(v[1]);
    L.idfb.push_back(v[2]);
  }
  return 0;
}

// ---- integral bookkeeping ------------------------------------------------------

// Decode_INT (incode_int.f90:25-36)
inline void decode_int(int v, int& met, int& k, int& i1, int& i2, int& i3, int& i4) {
  i1 = v % 32; v /= 32;
  i2 = v % 32; v /= 32;
  i3 = v % 32; v /= 32;
  i4 = v % 32; v /= 32;
  met = v % 16; k = v / 16;
}

// Jsym_int (jsym_int.f90): the canonical form of the integral
inline void jsym_int(int met, int& i1, int& i2, int& i3, int& i4) {
  int m = 1;
  const int ii = std::min(std::min(i1, i2), std::min(i3, i4));
  switch (met) {
    case 4: case 5:
      if (ii == i4) m = 4;
      if (ii == i3) m = 3;
      if (ii == i2) m = 2;
      if (ii == i1) m = 1;
      break;
    case 3:
      if (ii == i2) m = 2;
      if (ii == i1) m = 1;
      break;
    case 6: case 7: case 8: case 10:
      if (ii == i3) m = 3;
      if (ii == i1) m = 1;
      break;
    default:
      break;
  }
  const int j1 = i1, j2 = i2, j3 = i3, j4 = i4;
  if (m == 2) { i1 = j2; i2 = j1; i3 = j4; i4 = j3; }
  else if (m == 3) { i1 = j3; i2 = j4; i3 = j1; i4 = j2; }
  else if (m == 4) { i1 = j4; i2 = j3; i3 = j2; i4 = j1; }
}

// Idef_itype (idef_type.f90:52-118): 0 when no structure fits
int idef_itype(const std::vector<int>& ief, int i1, int i2, int i3, int i4, int io, int jo, int ic, int jc, int& k1,
               int& k2, int& k3, int& k4) {
  auto f = [&](int i) { return ief[i - 1]; };
  const int ibi = kIbi;
  if (jc > 0) {
    k1 = i1 * ibi + i3; k2 = i2 * ibi + i4; k3 = -ic; k4 = -jc; return 1;
  }
  if (ic > 0) {
    if (io == 0) {
      if (f(i1) > 0) { k1 = i2 * ibi + i4; k2 = i3; k3 = f(i1); k4 = -ic; return 2; }
      if (f(i2) > 0) { k1 = i1 * ibi + i3; k2 = i4; k3 = f(i2); k4 = -ic; return 3; }
      if (f(i3) > 0) { k1 = i2 * ibi + i4; k2 = i1; k3 = f(i3); k4 = -ic; return 4; }
      if (f(i4) > 0) { k1 = i1 * ibi + i3; k2 = i2; k3 = f(i4); k4 = -ic; return 5; }
      return 0;
    }
    if (io > 0) { k1 = i1 * ibi + i3; k2 = i2 * ibi + i4; k3 = io; k4 = -ic; return 1; }
    return 0;
  }
  if (io == 0) {
    if (f(i1) > 0 && f(i3) > 0) { k1 = f(i1) * ibi + f(i3); k2 = i2; k3 = i4; k4 = 0; return 6; }
    if (f(i2) > 0 && f(i4) > 0) { k1 = f(i2) * ibi + f(i4); k2 = i1; k3 = i3; k4 = 0; return 7; }
    if (f(i1) > 0 && f(i4) > 0) { k1 = f(i1) * ibi + f(i4); k2 = i3; k3 = i2; k4 = 0; return 8; }
    if (f(i2) > 0 && f(i3) > 0) { k1 = f(i3) * ibi + f(i2); k2 = i1; k3 = i4; k4 = 0; return 9; }
    return 0;
  }
  if (io > 0 && jo == 0) {
    if (f(i1) > 0) { k1 = i2 * ibi + i4; k2 = i3; k3 = f(i1); k4 = io; return 2; }
    if (f(i2) > 0) { k1 = i1 * ibi + i3; k2 = i4; k3 = f(i2); k4 = io; return 3; }
    if (f(i3) > 0) { k1 = i2 * ibi + i4; k2 = i1; k3 = f(i3); k4 = io; return 4; }
    if (f(i4) > 0) { k1 = i1 * ibi + i3; k2 = i2; k3 = f(i4); k4 = io; return 5; }
    k1 = i1 * ibi + i3; k2 = i2 * ibi + i4; k3 = io; k4 = 0; return 1;
  }
  if (io > 0 && jo > 0) { k1 = i1 * ibi + i3; k2 = i2 * ibi + i4; k3 = io; k4 = jo; return 1; }
  return 0;
}

// x**n for an integer n >= 0 as the Fortran runtime evaluates it: square
// and multiply, lowest bit first
inline double ipow(double x, int n) {
  double y = (n % 2) ? x : 1.0;
  while (n >>= 1) {
    x = x * x;
    if (n % 2) y = y * x;
  }
  return y;
}

inline double sgn(int e) { return (e % 2) ? -1.0 : 1.0; }

struct Dets {
  const bsrg_lb& L;
  // new_dets: one-electron overlap expansions of the determinants of a factor
  std::vector<int> izod, jzod;
  std::vector<double> adet;
  // new_defs: the resulting overlap factors
  std::vector<int> iof, jof;
  std::vector<double> adef;
  std::string err;

  explicit Dets(const bsrg_lb& l) : L(l) {}

  double O(int i, int j) const { return L.obs[(i - 1) + static_cast<size_t>(L.nbf) * (j - 1)]; }
  int IB(int i, int j) const { return L.ibort[(i - 1) + static_cast<size_t>(L.nbf) * (j - 1)]; }

  // DET (ZCOM/det.f90:2-44) of the overlap matrix a(r,c) = OBS(N1(c), N2(r))
  double det(int n, const int* N1, const int* N2) const {
    double A[32 * 32];
    auto a = [&](int r, int c) -> double& { return A[(r - 1) + 32 * (c - 1)]; };
    for (int i = 1; i <= n; ++i)
      for (int j = 1; j <= n; ++j) a(j, i) = O(N1[i - 1], N2[j - 1]);
    double d = 1.0;
    for (int k = 1; k <= n; ++k) {
      double mx = 0.0;
      int jj = k;
      for (int i = k; i <= n; ++i) {
        const double t = a(i, k);
        if (std::fabs(t) > std::fabs(mx)) { mx = t; jj = i; }
      }
      if (mx == 0.0) return 0.0;
      if (jj != k) {
        d = -d;
        for (int i = k; i <= n; ++i) std::swap(a(jj, i), a(k, i));
      }
      for (int i = k + 1; i <= n; ++i) {
        const double t = a(i, k) / mx;
        for (int j = k + 1; j <= n; ++j) a(i, j) = a(i, j) - t * a(k, j);
      }
      d = d * a(k, k);
    }
    return d;
  }

  // VDET (vdet.f90:2-48)
  double vdet(int kd, const int* N1, const int* N2) const {
    if (kd == 0) return 1.0;
    if (kd == 1) return O(N1[0], N2[0]);
    if (kd == 2) return O(N1[0], N2[0]) * O(N1[1], N2[1]) - O(N1[0], N2[1]) * O(N1[1], N2[0]);
    if (kd == 3) {
      const double t1 = O(N1[0], N2[0]) * O(N1[1], N2[1]) * O(N1[2], N2[2]);
      const double t2 = O(N1[0], N2[1]) * O(N1[1], N2[2]) * O(N1[2], N2[0]);
      const double t3 = O(N1[0], N2[2]) * O(N1[1], N2[0]) * O(N1[2], N2[1]);
      const double t4 = O(N1[0], N2[2]) * O(N1[1], N2[1]) * O(N1[2], N2[0]);
      const double t5 = O(N1[0], N2[1]) * O(N1[1], N2[0]) * O(N1[2], N2[2]);
      const double t6 = O(N1[0], N2[0]) * O(N1[1], N2[2]) * O(N1[2], N2[1]);
      return ((((t1 + t2) + t3) - t4) - t5) - t6;
    }
    return det(kd, N1, N2);
  }

  void add_det(int io, int jo, double s) {  // Iadd_ndets (mod_ndet.f90:69-88)
    if (std::fabs(s) < L.eps_det) return;
    izod.push_back(io);
    jzod.push_back(jo);
    adet.push_back(s);
  }
  void add_def(int io, int jo, double s) {  // Iadd_ndefs (mod_ndef.f90:69-98)
    if (std::fabs(s) < L.eps_det) return;
    for (size_t i = 0; i < iof.size(); ++i)
      if (iof[i] == io && jof[i] == jo) {
        adef[i] = adef[i] + s;
        return;
      }
    iof.push_back(io);
    jof.push_back(jo);
    adef.push_back(s);
  }

  static void shift(int n, int m, const int* a, int* b) {  // Shift (check_det.f90:100-112)
    int k = 0;
    for (int i = 1; i <= n; ++i)
      if (i != m) b[k++] = a[i - 1];
  }

  // Check_det (check_det.f90:2-96); false on a stop of the reference
  bool check_det(int kdn, const int* N1, const int* N2, int iext) {
    if (kdn > L.nelc + 1) { err = " Check_det: kdn > nelc"; return false; }
    int j1 = 0, j2 = 0;
    for (int i = 1; i <= kdn; ++i) {
      if (L.iech[N1[i - 1] - 1] != 0) j1 = i;
      if (L.iech[N2[i - 1] - 1] != 0) j2 = i;
    }
    int N3[33], N4[33], N5[33], N6[33];
    if (j1 == 0 && j2 == 0) {
      add_det(0, 0, ipow(vdet(kdn, N1, N2), iext));
    } else if (kdn == 1) {
      if (iext > 1) { err = " Check_det: iext > 1"; return false; }
      const int io = IB(N1[0], N2[0]);
      if (io != 0) add_det(io, 0, 1.0);
    } else if (j1 > 0 && j2 == 0) {
      if (iext > 1) { err = " Check_det: iext > 1"; return false; }
      for (int j = 1; j <= kdn; ++j) {
        const int io = IB(N1[j1 - 1], N2[j - 1]);
        if (io == 0) continue;
        shift(kdn, j1, N1, N3);
        shift(kdn, j, N2, N4);
        add_det(io, 0, vdet(kdn - 1, N3, N4) * sgn(j1 + j));
      }
    } else if (j2 > 0 && j1 == 0) {
      if (iext > 1) { err = " Check_det: iext > 1"; return false; }
      for (int j = 1; j <= kdn; ++j) {
        const int io = IB(N1[j - 1], N2[j2 - 1]);
        if (io == 0) continue;
        shift(kdn, j, N1, N3);
        shift(kdn, j2, N2, N4);
        add_det(io, 0, vdet(kdn - 1, N3, N4) * sgn(j2 + j));
      }
    } else {
      if (iext > 1) { err = " Check_det: iext > 1"; return false; }
      for (int j = 1; j <= kdn; ++j) {
        shift(kdn, j1, N1, N3);
        shift(kdn, j, N2, N4);
        const int io = IB(N1[j1 - 1], N2[j - 1]);
        if (j == j2) {
          if (io == 0) { err = " Orthogonal kl ?"; return false; }
          add_det(io, 0, vdet(kdn - 1, N3, N4) * sgn(j1 + j2));
        } else {
          const int j3 = j < j2 ? j2 - 1 : j2;
          for (int j4 = 1; j4 <= kdn - 1; ++j4) {
            const int jo = IB(N3[j4 - 1], N4[j3 - 1]);
            if (io == 0 || jo == 0) continue;
            shift(kdn - 1, j3, N4, N6);
            shift(kdn - 1, j4, N3, N5);
            add_det(io, jo, vdet(kdn - 2, N5, N6) * sgn(j1 + j + j3 + j4));
          }
        }
      }
    }
    return true;
  }

  // Det_fact (det_fact.f90:2-93); false on a stop of the reference
  bool det_fact(int idf, const int* ipn1, int no1, const int* ipn2, int no2) {
    iof.clear(); jof.clear(); adef.clear();
    if (idf == 0) {
      add_def(0, 0, 1.0);
      return true;
    }
    if (idf < 1 || idf > L.ndef) { err = "Det_fact: overlap factor index out of range"; return false; }
    const int kd = L.kpf[idf - 1], ipf = L.ipf[idf - 1];
    if (kd < 1 || kd > 32) { err = "Det_fact: factor size out of range"; return false; }
    int np[33], idpnt[33];
    for (int j = 0; j < kd; ++j) np[j] = L.npf[ipf + j];
    izod.clear(); jzod.clear(); adet.clear();
    int N1[33], N2[33];
    for (int j = 1; j <= kd; ++j) {
      const int id = np[j - 1] / kIdef, iext = np[j - 1] % kIdef;
      if (id < 1 || id > L.ndet) { err = "Det_fact: determinant index out of range"; return false; }
      const int kn = L.kpd[id - 1], ip = L.ipd[id - 1];
      if (kn > 32) { err = "Det_fact: determinant size out of range"; return false; }
      for (int i = 1; i <= kn; ++i) {
        const int k = L.npd[ip + i - 1], a = k / kIdet, b = k % kIdet;
        if (a < 1 || a > no1 || b < 1 || b > no2) { err = "Det_fact: orbital position out of the states"; return false; }
        N1[i - 1] = ipn1[a - 1];
        N2[i - 1] = ipn2[b - 1];
      }
      if (!check_det(kn, N1, N2, iext)) return false;
      idpnt[j - 1] = static_cast<int>(adet.size());
      if (j > 1 && idpnt[j - 1] == idpnt[j - 2]) return true;  // a zero determinant: no factor
      if (adet.empty()) return true;
    }
    // every choice of one term per determinant
    double df[33];
    int i = 1;
    np[0] = 1;
    for (;;) {
      // label 1
      if (i == 1) {
        df[0] = adet[np[0] - 1];
      } else {
        const int k = idpnt[i - 2] + np[i - 1];
        df[i - 1] = df[i - 2] * adet[k - 1];
      }
      if (i < kd) {
        ++i;
        np[i - 1] = 1;
        continue;
      }
      int io = 0, jo = 0;
      for (int j = 1; j <= kd; ++j) {
        int k = np[j - 1];
        if (j > 1) k = idpnt[j - 2] + np[j - 1];
        const int ii = izod[k - 1];
        if (ii == 0) continue;
        const int jj = jzod[k - 1];
        if (jj == 0) {
          if (io == 0) io = ii;
          else if (io > 0 && jo == 0) jo = ii;
          else { err = "Det_fact: problems with io,jo"; return false; }
        } else {
          if (io == 0) { io = ii; jo = jj; }
          else { err = "Det_fact: problems with io,jo"; return false; }
        }
      }
      add_def(io, jo, df[kd - 1]);
      // label 2: the next choice
      bool more = false;
      while (i > 0) {
        ++np[i - 1];
        int k = np[i - 1];
        if (i > 1) k = idpnt[i - 2] + np[i - 1];
        if (k <= idpnt[i - 1]) { more = true; break; }
        --i;
      }
      if (!more) break;
    }
    return true;
  }
};

// Ifind_channel (MOD_channel.f90:246-274): channel of state ic, nch + ip for
// perturber ip, 0 on a stop of the reference
int find_channel(const bsrg_lb& L, int ic) {
  if (ic <= 0) return 0;
  int r = 1;
  for (int ich = 1; ich <= L.nch; ++ich)
    if (ic > L.ipconf[ich - 1]) r = ich + 1;
  if (r <= L.nch) return r;
  for (int ip = 1; ip <= L.npert; ++ip)
    if (ic > L.ippert[ip]) r = L.nch + ip + 1;
  return r > L.nch + L.npert ? 0 : r;
}

// ---- the blocks of module cmdata (Add_coef, Gen_matrix) ----------------------

struct Blocks {
  bsrg_lb& L;
  int npol;
  int nfree;
  // chain of blocks of each (kpol, itype), acquisition order; a block maps
  // its keys to their sums in arrival order (Add_cdata)
  std::vector<std::vector<std::map<Key, double>>> chain;

  Blocks(bsrg_lb& l, int np) : L(l), npol(np) {
    chain.assign(static_cast<size_t>(npol + 2) * kNtype, {});
    for (auto& c : chain) c.assign(1, {});
    nfree = L.nb - static_cast<int>(chain.size());
  }
  std::vector<std::map<Key, double>>& at(int kpol, int itype) {
    return chain[static_cast<size_t>(kpol + 1) * kNtype + (itype - 1)];
  }

  // Gen_matrix's list step (gen_matrix.f90:17-47): the batch of one chain
  int generate(int itype, int kpol) {
    auto& ch = at(kpol, itype);
    const int n = static_cast<int>(ch.size());
    if (n == 1 && ch[0].empty()) return 0;
    const int batch = L.nbatch;
    if (n == 1) {
      for (const auto& kv : ch[0])
        L.out.push_back({batch, itype, kpol, std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first),
                         std::get<3>(kv.first), kv.second});
    } else {
      std::vector<int> k1, k2, k3, k4, ip, jp;
      std::vector<double> c;
      for (const auto& b : ch) {
        ip.push_back(static_cast<int>(c.size()) + 1);
        for (const auto& kv : b) {
          k1.push_back(std::get<0>(kv.first)); k2.push_back(std::get<1>(kv.first));
          k3.push_back(std::get<2>(kv.first)); k4.push_back(std::get<3>(kv.first));
          c.push_back(kv.second);
        }
        jp.push_back(b.empty() ? 0 : static_cast<int>(c.size()));
      }
      std::vector<int> ipt(std::max<size_t>(c.size(), 1));
      int nc = 0;
      if (int rc = bsrg_cdata_merge(n, ip.data(), jp.data(), static_cast<int>(c.size()), k1.data(), k2.data(),
                                    k3.data(), k4.data(), c.data(), L.eps_c, ipt.data(), &nc))
        return rc;
      for (int q = 0; q < nc; ++q) {
        const int i = ipt[q] - 1;
        L.out.push_back({batch, itype, kpol, k1[i], k2[i], k3[i], k4[i], c[i]});
      }
      nfree += n - 1;
    }
    ch.assign(1, {});
    ++L.nbatch;
    return 0;
  }

  // Add_coef (add_coef.f90:2-59)
  int add(int itype, int kpol, double c, int k1, int k2, int k3, int k4) {
    auto& ch = at(kpol, itype);
    auto& blk = ch.back();
    const Key key(k1, k2, k3, k4);
    auto it = blk.find(key);
    if (it != blk.end()) it->second = it->second + c;
    else blk.emplace(key, c);
    if (static_cast<int>(blk.size()) < L.mb) return 0;
    const int n = static_cast<int>(ch.size());
    if (n < L.kb && nfree > 0) {
      ch.emplace_back();
      --nfree;
      return 0;
    }
    int jtype = itype, jpol = kpol;
    for (int i = 1; i <= kNtype; ++i)
      for (int k = -1; k <= npol; ++k)
        if (static_cast<int>(at(k, i).size()) > n) { jtype = i; jpol = k; }
    if (int rc = generate(jtype, jpol)) return rc;
    if (jtype == itype && jpol == kpol) return 0;
    if (nfree <= 0) return bsrg::fail(-43, " Add_coef: problems with new block ");
    ch.emplace_back();
    --nfree;
    return 0;
  }
};

int run(bsrg_lb& L, int icase, int mk) {
  if (icase < 1 || icase > 11) return bsrg::fail(-1, "bsrg_lb_run: icase out of 1..11");
  if (icase >= 7 && icase <= 10)
    return bsrg::fail(-44, "bsrg_lb_run: the J-dependent relativistic cases (icase 7-10) are not built");
  if (L.nbf < 1) return bsrg::fail(-4, "bsrg_lb_run: orbitals not set");
  if (L.ip_stat.empty() || L.it1.empty()) return bsrg::fail(-4, "bsrg_lb_run: states not set");
  if (L.nch < 1) return bsrg::fail(-4, "bsrg_lb_run: channels not set");
  int maxl = 0;
  for (int l : L.lef) maxl = std::max(maxl, l);
  const int npol = std::max(maxl, mk);
  if (L.nb < (npol + 2) * kNtype) return bsrg::fail(-1, "bsrg_lb_run: nb below (npol+2)*ntype blocks");
  if (L.mb < 1 || L.kb < 1) return bsrg::fail(-1, "bsrg_lb_run: mb and kb must be positive");
  L.expansion.clear();
  L.out.clear();
  L.nbatch = 0;
  Blocks B(L, npol);
  Dets D(L);
  const int nsym = static_cast<int>(L.it1.size()), ncfg = static_cast<int>(L.ip_state.size());
  const int nstat = static_cast<int>(L.ip_stat.size());
  const int nrec = static_cast<int>(L.cbuf.size());
  for (int ibuf = 0; ibuf < nrec; ++ibuf) {
    int jcase, k, i1, i2, i3, i4;
    decode_int(L.intb[ibuf], jcase, k, i1, i2, i3, i4);
    if (jcase != icase) continue;
    int kpol = k;
    if (icase == 4 || icase == 8 || icase == 9 || icase == 10) kpol = k - 1;
    if (icase == 6 || icase == 7 || icase == 11) kpol = 0;
    if (kpol > mk) continue;
    const int ijt = L.ijtb[ibuf], it = ijt / kIbc, jt = ijt % kIbc;
    if (it < 1 || it > nsym || jt < 1 || jt > nsym) return bsrg::fail(-45, "int_bnk: symmetry index out of range");
    const int is1 = L.it1[it - 1], js1 = L.it1[jt - 1];
    if (is1 == 0 || js1 == 0) continue;
    const int is2 = L.it2[it - 1], js2 = L.it2[jt - 1];
    if (is2 > nstat || js2 > nstat) return bsrg::fail(-45, "state range out of IP_stat");
    for (int ik = is1; ik <= is2; ++ik) {
      const int is = L.ip_stat[ik - 1];
      if (is < 1 || is > ncfg) return bsrg::fail(-45, "state index out of range");
      const int* np1 = L.ip_orb.data() + L.ip_state[is - 1];
      const int no1 = L.no_state[is - 1];
      const int ich = find_channel(L, is);
      if (!ich) return bsrg::fail(-46, "Ifind_channel: ic is to big");
      for (int jk = js1; jk <= js2; ++jk) {
        const int js = L.ip_stat[jk - 1];
        if (js < 1 || js > ncfg) return bsrg::fail(-45, "state index out of range");
        const int* np2 = L.ip_orb.data() + L.ip_state[js - 1];
        const int no2 = L.no_state[js - 1];
        const int jch = find_channel(L, js);
        if (!jch) return bsrg::fail(-46, "Ifind_channel: ic is to big");
        if (it == jt && js > is) continue;
        if (icase == 3 || icase == 4) {
          if (std::fabs(L.wc[is - 1] * L.wc[js - 1]) < L.eps_soo) continue;
        }
        double C = L.cbuf[ibuf];
        if (ich == jch && is != js) C = C + C;
        const double CC = C * L.wc[is - 1] * L.wc[js - 1];
        if (std::fabs(CC) < L.eps_c) continue;
        const int ia = ich > L.nch ? ich - L.nch : 0, ja = jch > L.nch ? jch - L.nch : 0;
        const int ic = std::max(ia, ja), jc = std::min(ia, ja);
        int j1 = 1, j2 = 1, j3 = 1, j4 = 1;
        if (icase != 11) {
          if (i1 < 1 || i1 > no1 || i2 < 1 || i2 > no1 || i3 < 1 || i3 > no2 || i4 < 1 || i4 > no2)
            return bsrg::fail(-45, "integral orbital position out of the states");
          j1 = np1[i1 - 1]; j2 = np1[i2 - 1]; j3 = np2[i3 - 1]; j4 = np2[i4 - 1];
        }
        if (icase == 6) kpol = L.lef[j1 - 1];
        jsym_int(icase, j1, j2, j3, j4);
        if (!D.det_fact(L.idfb[ibuf], np1, no1, np2, no2)) return bsrg::fail(-47, D.err);
        for (size_t q = 0; q < D.adef.size(); ++q) {
          const double CCC = CC * D.adef[q];
          int kk1, kk2, kk3, kk4;
          const int itype = idef_itype(L.iech, j1, j2, j3, j4, D.iof[q], D.jof[q], ic, jc, kk1, kk2, kk3, kk4);
          if (!itype) return bsrg::fail(-48, "Idef_itype=0");
          L.expansion.push_back({itype, kpol, kk1, kk2, kk3, kk4, CCC});
          if (int rc = B.add(itype, kpol, CCC, kk1, kk2, kk3, kk4)) return rc;
        }
      }
    }
  }
  for (int jpol = -1; jpol <= npol; ++jpol)
    for (int jtype = 1; jtype <= kNtype; ++jtype)
      if (int rc = B.generate(jtype, jpol)) return rc;
  return 0;
}

}  // namespace

extern "C" {

bsrg_lb* bsrg_lb_create(void) { return new bsrg_lb(); }

void bsrg_lb_destroy(bsrg_lb* lb) { delete lb; }

int bsrg_lb_read_bank(bsrg_lb* lb, const char* path, int* sizes5) {
  if (!lb || !path) return bsrg::fail(-2, "null argument");
  if (int rc = read_bank(path, *lb)) return rc;
  if (sizes5) {
    sizes5[0] = lb->ndet; sizes5[1] = lb->kdet; sizes5[2] = lb->ndef; sizes5[3] = lb->kdef;
    sizes5[4] = static_cast<int>(lb->cbuf.size());
  }
  return 0;
}

int bsrg_lb_bank_get(bsrg_lb* lb, double* c, int* ijt, int* intb, int* idf, int* kpd, int* ipd, int* npd, int* kpf,
                     int* ipf, int* npf) {
  if (!lb) return bsrg::fail(-2, "null argument");
  auto put = [](const auto& v, auto* p) {
    if (p) std::copy(v.begin(), v.end(), p);
  };
  put(lb->cbuf, c); put(lb->ijtb, ijt); put(lb->intb, intb); put(lb->idfb, idf);
  put(lb->kpd, kpd); put(lb->ipd, ipd); put(lb->npd, npd);
  put(lb->kpf, kpf); put(lb->ipf, ipf); put(lb->npf, npf);
  return 0;
}

int bsrg_lb_set_records(bsrg_lb* lb, int nrec, const double* c, const int* ijt, const int* intb, const int* idf) {
  if (!lb || nrec < 0 || (nrec > 0 && (!c || !ijt || !intb || !idf))) return bsrg::fail(-2, "bad arguments");
  lb->cbuf.assign(c, c + nrec); lb->ijtb.assign(ijt, ijt + nrec);
  lb->intb.assign(intb, intb + nrec); lb->idfb.assign(idf, idf + nrec);
  return 0;
}

int bsrg_lb_set_dets(bsrg_lb* lb, int ndet, int kdet, const int* kpd, const int* ipd, const int* npd, int ndef,
                     int kdef, const int* kpf, const int* ipf, const int* npf) {
  if (!lb || ndet < 0 || kdet < 0 || ndef < 0 || kdef < 0) return bsrg::fail(-2, "bad arguments");
  if ((ndet && (!kpd || !ipd)) || (kdet && !npd) || (ndef && (!kpf || !ipf)) || (kdef && !npf))
    return bsrg::fail(-2, "null argument");
  for (int i = 0; i < ndet; ++i)
    if (kpd[i] < 0 || ipd[i] < 0 || ipd[i] + kpd[i] > kdet) return bsrg::fail(-1, "determinant outside NPD");
  for (int i = 0; i < ndef; ++i)
    if (kpf[i] < 0 || ipf[i] < 0 || ipf[i] + kpf[i] > kdef) return bsrg::fail(-1, "overlap factor outside NPF");
  lb->ndet = ndet; lb->kdet = kdet; lb->ndef = ndef; lb->kdef = kdef;
  lb->kpd.assign(kpd, kpd + ndet); lb->ipd.assign(ipd, ipd + ndet); lb->npd.assign(npd, npd + kdet);
  lb->kpf.assign(kpf, kpf + ndef); lb->ipf.assign(ipf, ipf + ndef); lb->npf.assign(npf, npf + kdef);
  return 0;
}

int bsrg_lb_set_orbitals(bsrg_lb* lb, int nbf, const int* iech, const int* lef, const int* ibort, const double* obs,
                         int nelc) {
  if (!lb || nbf < 1 || !iech || !lef || !ibort || !obs || nelc < 0) return bsrg::fail(-2, "bad arguments");
  lb->nbf = nbf;
  lb->nelc = nelc;
  lb->iech.assign(iech, iech + nbf);
  lb->lef.assign(lef, lef + nbf);
  lb->ibort.assign(ibort, ibort + static_cast<size_t>(nbf) * nbf);
  lb->obs.assign(obs, obs + static_cast<size_t>(nbf) * nbf);
  return 0;
}

int bsrg_lb_set_channels(bsrg_lb* lb, int nch, const int* ipconf, int npert, const int* ippert) {
  if (!lb || nch < 1 || !ipconf || npert < 0 || !ippert) return bsrg::fail(-2, "bad arguments");
  lb->nch = nch;
  lb->npert = npert;
  lb->ipconf.assign(ipconf, ipconf + nch);
  lb->ippert.assign(ippert, ippert + npert + 1);
  return 0;
}

int bsrg_lb_set_states(bsrg_lb* lb, int ncfg, const int* ip_state, const int* no_state, const double* wc, int lorb,
                       const int* ip_orb, int nsym, const int* it_state1, const int* it_state2, int nstat,
                       const int* ip_stat) {
  if (!lb || ncfg < 1 || !ip_state || !no_state || !wc || lorb < 1 || !ip_orb || nsym < 1 || !it_state1 ||
      !it_state2 || nstat < 1 || !ip_stat)
    return bsrg::fail(-2, "bad arguments");
  for (int i = 0; i < ncfg; ++i)
    if (ip_state[i] < 0 || no_state[i] < 1 || no_state[i] > 32 || ip_state[i] + no_state[i] > lorb)
      return bsrg::fail(-1, "bsrg_lb_set_states: a state's orbitals lie outside IP_orb");
  for (int i = 0; i < lorb; ++i)
    if (ip_orb[i] < 1 || (lb->nbf && ip_orb[i] > lb->nbf)) return bsrg::fail(-1, "bsrg_lb_set_states: orbital index out of range");
  lb->ip_state.assign(ip_state, ip_state + ncfg);
  lb->no_state.assign(no_state, no_state + ncfg);
  lb->wc.assign(wc, wc + ncfg);
  lb->ip_orb.assign(ip_orb, ip_orb + lorb);
  lb->it1.assign(it_state1, it_state1 + nsym);
  lb->it2.assign(it_state2, it_state2 + nsym);
  lb->ip_stat.assign(ip_stat, ip_stat + nstat);
  return 0;
}

int bsrg_lb_set_params(bsrg_lb* lb, double eps_c, double eps_det, double eps_soo, int nb, int mb, int kb) {
  if (!lb) return bsrg::fail(-2, "null argument");
  lb->eps_c = eps_c; lb->eps_det = eps_det; lb->eps_soo = eps_soo;
  lb->nb = nb; lb->mb = mb; lb->kb = kb;
  return 0;
}

int bsrg_lb_run(bsrg_lb* lb, int icase, int mk, int* sizes3) {
  if (!lb) return bsrg::fail(-2, "null argument");
  if (int rc = run(*lb, icase, mk)) return rc;
  if (sizes3) {
    sizes3[0] = static_cast<int>(lb->expansion.size());
    sizes3[1] = static_cast<int>(lb->out.size());
    sizes3[2] = lb->nbatch;
  }
  return 0;
}

int bsrg_lb_expansion(bsrg_lb* lb, int* itype, int* kpol, int* k1, int* k2, int* k3, int* k4, double* c) {
  if (!lb || !itype || !kpol || !k1 || !k2 || !k3 || !k4 || !c) return bsrg::fail(-2, "null argument");
  for (size_t i = 0; i < lb->expansion.size(); ++i) {
    const auto& e = lb->expansion[i];
    itype[i] = e.itype; kpol[i] = e.kpol; k1[i] = e.k1; k2[i] = e.k2; k3[i] = e.k3; k4[i] = e.k4; c[i] = e.c;
  }
  return 0;
}

int bsrg_lb_lists(bsrg_lb* lb, int* batch, int* jtype, int* jpol, int* k1, int* k2, int* k3, int* k4, double* c) {
  if (!lb || !batch || !jtype || !jpol || !k1 || !k2 || !k3 || !k4 || !c) return bsrg::fail(-2, "null argument");
  for (size_t i = 0; i < lb->out.size(); ++i) {
    const auto& r = lb->out[i];
    batch[i] = r.batch; jtype[i] = r.jtype; jpol[i] = r.jpol;
    k1[i] = r.k1; k2[i] = r.k2; k3[i] = r.k3; k4[i] = r.k4; c[i] = r.c;
  }
  return 0;
}

}  // extern "C"
