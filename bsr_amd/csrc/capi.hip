// C ABI of the MI355X radial core (declarations in include/bsr_gpu.h).
//
// Design: the R^k bands of all multipoles stay resident.  The orbital
// contractions of I_data are linear in the coefficient lists, so the
// contraction of every band with every bound-orbital pair is done once per
// wave ("bases": direct bands D, exchange matrices E, scalars St and
// vectors V).  A list then becomes terms on the channel blocks:
//   exchange (jtype 8,9)  -> coefficient matrix of one FP64 GEMM  E^T * C
//   direct   (jtype 6,7)  -> band terms  c * D   (symmetric band)
//   rank-1   (UPDATE_HW)  -> c * u (x) w
//   vectors  (UPDATE_HV)  -> c * v into hcb
//   scalars  (HB, Target_h, ACF) on the host.
// Each output element is summed by one thread in list order: no atomics,
// results do not depend on what else runs on the device.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <chrono>
#include <string>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/bsr_gpu.h"
#include "bsrg_internal.h"

namespace bsrg {

static thread_local std::string g_err;
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

}  // namespace bsrg

#include "ctx.h"

namespace {

DevGrid devgrid(bsrg_ctx* c) {
  DevGrid g;
  g.ks = c->G.ks; g.ns = c->G.ns; g.nv = c->G.nv; g.ml = c->G.ml; g.me = c->G.me; g.h = c->G.h;
  g.t = c->t.p; g.gx = c->gx.p; g.gw = c->gw.p; g.gr = c->gr.p; g.grm = c->grm.p; g.grw = c->grw.p;
  g.bsp = c->bsp.p;
  return g;
}

int check_ctx(bsrg_ctx* c) {
  if (!c) return fail(-2, "null context");
  return 0;
}

// Bands for k = 0..nk-1 (rebuilt as a whole when more multipoles are needed).
int ensure_bands(bsrg_ctx* c, int nk) {
  if (!c->grid_ok) return fail(-3, "grid not set (bsrg_set_grid)");
  if (nk <= c->band_nk) return 0;
  const int ks = c->G.ks, ns = c->G.ns, nv = c->G.nv, ik = ks * (ks + 1) / 2;
  HIPCHK(c->mom.resize(static_cast<size_t>(nk) * 2 * nv * ik));
  HIPCHK(c->rkd.resize(static_cast<size_t>(nk) * nv * ik * ik));
  HIPCHK(c->rkb.resize(static_cast<size_t>(nk) * ns * ns * ks * ks));
  DevGrid g = devgrid(c);
  launch_moments(g, 0, nk, c->mom.p, c->rkd.p, c->stream);
  HIPCHK(c->msum.resize(static_cast<size_t>(nk) * 2 * ks * ns));
  HIPCHK(c->momT.resize(c->mom.n));
  HIPCHK(c->rdT.resize(c->rkd.n));
  launch_band(g, nk, c->mom.p, c->rkd.p, c->msum.p, c->momT.p, c->rdT.p, c->rkb.p, c->stream);
  HIPCHK(hipGetLastError());
  c->band_nk = nk;
  c->basis_nk = 0;  // bases depend on the bands
  c->exch_nk = 0;
  return 0;
}

int pair_index(bsrg_ctx* c, int o1, int o2, int& p) {
  if (o1 < 1 || o1 > c->nbf || o2 < 1 || o2 > c->nbf) return fail(-30, "orbital index out of range");
  int a = c->bidx[o1 - 1], b = c->bidx[o2 - 1];
  if (a < 0 || b < 0) return fail(-31, "R^k density of a channel orbital (no B-spline expansion)");
  if (a > b) std::swap(a, b);
  p = a * c->nb - a * (a - 1) / 2 + (b - a);
  return 0;
}

int ensure_bases(bsrg_ctx* c, int nk, bool exchange) {
  if (!c->orb_ok) return fail(-4, "orbitals not set (bsrg_set_orbitals)");
  int rc = ensure_bands(c, nk);
  if (rc) return rc;
  const int ks = c->G.ks, ns = c->G.ns, nb = c->nb, np = c->npair;
  if (c->basis_nk < nk) {
    HIPCHK(c->bands.resize((1 + static_cast<size_t>(nk) * np) * ns * ks));
    launch_sb_upper(ns, ks, c->sb.p, c->bands.p, c->stream);
    double* D = c->bands.p + static_cast<size_t>(ns) * ks;
    if (np > 0) {
      HIPCHK(c->dpart.resize(static_cast<size_t>(nk) * np * ks * ns * ks));
      launch_direct_basis(ns, ks, nk, np, c->rkb.p, c->dens.p, c->dpart.p, D, c->stream);
      HIPCHK(c->St.resize(static_cast<size_t>(nk) * np * np));
      launch_stab(ns, ks, nk, np, D, c->dens.p, c->St.p, c->stream);
    }
    const size_t nV = static_cast<size_t>(nk) * np * nb;
    HIPCHK(c->VT.resize((nV + c->wvec.size() / std::max(ns, 1) + 1) * ns));
    if (np > 0) launch_vtab(ns, ks, nk, np, nb, D, c->P.p, c->VT.p, c->stream);
    HIPCHK(hipGetLastError());
    const size_t nst = static_cast<size_t>(nk) * np * np;
    if (nst > c->St_host.cap) {
      HIPCHK(hipStreamSynchronize(c->stream));  // no copy into the old buffer is pending
      HIPCHK(c->St_host.reserve(nst));
    }
    if (np > 0)
      HIPCHK(hipMemcpyAsync(c->St_host.p, c->St.p, nst * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    c->basis_nk = nk;
    c->wvec_uploaded = -1;
  }
  if (exchange && c->exch_nk < nk && nb > 0) {
    HIPCHK(c->E.resize(static_cast<size_t>(nk) * nb * nb * ns * ns));
    launch_exchange_basis(ns, ks, nk, nb, c->rkb.p, c->P.p, c->P.p, c->E.p, c->stream);
    HIPCHK(hipGetLastError());
    c->exch_nk = nk;
    ++c->e_ver;
  }
  return 0;
}

constexpr int kNotRegistered = -1019;  // internal: a GET_V vector is not in the registry yet

// GET_V (src/bsr/BSR_MAT3/get_V.f90:2-27) on the host: qbs(:,jj) minus the
// projections on same-l bound orbitals that the channel orbital ii must be
// orthogonal to.  Returns the vector id in the W part of the vector table.
int get_v(bsrg_ctx* c, int ii, int jj, int& id, bool create = true) {
  const int nbf = c->nbf, ns = c->G.ns;
  if (ii < 1 || ii > nbf || jj < 1 || jj > nbf) return fail(-18, "get_V: orbital index out of range");
  const size_t key = static_cast<size_t>(ii - 1) + static_cast<size_t>(nbf) * (jj - 1);
  if (c->wtab[key] >= 0) { id = c->wtab[key]; return 0; }
  if (c->ibort[(ii - 1) + static_cast<size_t>(nbf) * (jj - 1)] == 0) return fail(-19, "get_V: <kl|nl> = 0 ?");
  if (!create) return fail(kNotRegistered, "get_V: vector not registered");
  std::vector<double> v(c->qbs.begin() + static_cast<size_t>(ns) * (jj - 1),
                        c->qbs.begin() + static_cast<size_t>(ns) * jj);
  for (int i = 1; i <= nbf; ++i) {
    if (i == jj || c->iech[i - 1] != 0 || c->lbs[i - 1] != c->lbs[ii - 1]) continue;
    if (c->ibort[(i - 1) + static_cast<size_t>(nbf) * (ii - 1)] != 0) continue;
    const double o = c->obs[(i - 1) + static_cast<size_t>(nbf) * (jj - 1)];
    if (std::fabs(o) < c->eps_ovl) continue;
    for (int r = 0; r < ns; ++r) v[r] -= o * c->qbs[static_cast<size_t>(ns) * (i - 1) + r];
  }
  id = static_cast<int>(c->wvec.size() / ns);
  c->wvec.insert(c->wvec.end(), v.begin(), v.end());
  c->wtab[key] = id;
  return 0;
}

double quadr(bsrg_ctx* c, int i, int j, int m) {
  if (static_cast<int>(c->Q.size()) <= m) { c->Q.resize(m + 1); c->Qok.resize(m + 1, 0); }
  if (!c->Qok[m]) {
    const int ns = c->G.ns, ks = c->G.ks, nbf = c->nbf;
    std::vector<double> rm = c->G.mrm(m);
    c->Q[m].assign(static_cast<size_t>(nbf) * nbf, 0.0);
    for (int a = 0; a < nbf; ++a) {
      if (c->iech[a] != 0) continue;
      for (int b = 0; b < nbf; ++b) {
        if (c->iech[b] != 0) continue;
        c->Q[m][a + static_cast<size_t>(nbf) * b] =
            bvmv_sym(ns, ks, rm.data(), &c->pbs[static_cast<size_t>(ns) * a], &c->pbs[static_cast<size_t>(ns) * b]);
      }
    }
    c->Qok[m] = 1;
  }
  return c->Q[m][(i - 1) + static_cast<size_t>(c->nbf) * (j - 1)];
}

inline int blk_index(int ich, int jch) {
  const int i = std::max(ich, jch), j = std::min(ich, jch);
  return i * (i - 1) / 2 + j - 1;
}

int chk_ch(const Matrices& M, int ich, int jch, const char* who) {
  if (ich < 1 || ich > M.nch || jch < 1 || jch > M.nch)
    return fail(-10, std::string(who) + ": channel index out of range");
  return 0;
}

// Add one structure-1 record (I_data jtype 1 or O_data) to the terms.
// sidx >= 0 multiplies by the scalar R^k table entry (I_data), -1 for O_data.
int add_struct1(bsrg_ctx* c, const Matrices& M, Terms& T, int K3, int K4, int sidx, double cc, bool odata,
                bool scal_targets) {
  const int ic = -K3, jc = -K4;
  const int nV = c->basis_nk * c->npair * c->nb;
  if (ic > 0 && jc > 0) {
    if (ic > M.ncp || jc > M.ncp) return fail(-12, "UPDATE_HB: indeces out of range");
    const int i = std::max(ic, jc), j = std::min(ic, jc);
    T.scal.push_back({0, i * (i - 1) / 2 + j - 1, sidx, cc});
  } else if (ic < 0 && jc > 0) {
    const int io = -ic, i1 = io / kIbo, i2 = io % kIbo;
    int w, rc = get_v(c, i1, i2, w, !T.deferred);
    if (rc) return rc;
    const int ich = c->iech[i1 - 1];
    if (ich < 1 || ich > M.nch) return fail(-13, "UPDATE_HV: channel index out of range");
    if (jc > M.ncp) return fail(-14, "UPDATE_HV: bound index out of range");
    T.vec.push_back({(jc - 1) * M.nch + (ich - 1), nV + w, sidx, cc});
  } else if (ic < 0 && jc < 0) {
    const int io = -ic, i1 = io / kIbo, i2 = io % kIbo, jo = -jc, j1 = jo / kIbo, j2 = jo % kIbo;
    int v, w, rc = get_v(c, i1, i2, v, !T.deferred);
    if (!rc) rc = get_v(c, j1, j2, w, !T.deferred);
    if (rc) return rc;
    const int ich = c->iech[i1 - 1], jch = c->iech[j1 - 1];
    if ((rc = chk_ch(M, ich, jch, "UPDATE_HW"))) return rc;
    if (ich >= jch) T.r1.push_back({blk_index(ich, jch), nV + v, nV + w, sidx, cc});
    else T.r1.push_back({blk_index(ich, jch), nV + w, nV + v, sidx, cc});
  } else if (ic < 0 && jc == 0) {
    const int io = -ic, i1 = io / kIbo, i2 = io % kIbo;
    if (i1 < 1 || i1 > c->nbf || i2 < 1 || i2 > c->nbf) return fail(-30, "orbital index out of range");
    const int ich = c->iech[i1 - 1], jch = c->iech[i2 - 1];
    int rc = chk_ch(M, ich, jch, "UPDATE_CF");
    if (rc) return rc;
    if (scal_targets) T.scal.push_back({odata ? 2 : 1, blk_index(ich, jch), sidx, cc});
  } else {
    return fail(odata ? -24 : -22, odata ? " O_data: unknown structure " : " I_data: unknown structure for itype=1");
  }
  return 0;
}

const int kIjcase[9] = {2, 1, 2, 1, 2, 1, 2, 3, 4};  // IJCASE, mod_cmdata.f90:54

// I_data for one (jtype, k) list -> terms (I_data.f90:2-224).
int add_idata(bsrg_ctx* c, const Matrices& M, Terms& T, int jtype, int k, int n, const int* ipt, const int* K1,
              const int* K2, const int* K3, const int* K4, const double* cd, double sc) {
  if (k < 0 || k >= c->basis_nk) return fail(-33, "multipole outside the prepared range");
  const int nb = c->nb, np = c->npair, nV = c->basis_nk * np * nb;
  const int jcase = kIjcase[jtype - 1];
  (void)nV;
  if (jtype == 1) {
    for (int q = 0; q < n; ++q) {
      const int i = ipt ? ipt[q] - 1 : q;
      int p1, p2, rc = pair_index(c, K1[i] / kIbi, K1[i] % kIbi, p1);
      if (!rc) rc = pair_index(c, K2[i] / kIbi, K2[i] % kIbi, p2);
      if (!rc) rc = add_struct1(c, M, T, K3[i], K4[i], (k * np + p1) * np + p2, cd[i] * sc, false, true);
      if (rc) return rc;
    }
  } else if (jtype <= 5) {
    for (int q = 0; q < n; ++q) {
      const int i = ipt ? ipt[q] - 1 : q;
      int p, rc = pair_index(c, K1[i] / kIbi, K1[i] % kIbi, p);
      if (rc) return rc;
      const int ii = K2[i];
      if (ii < 1 || ii > c->nbf || c->bidx[ii - 1] < 0) return fail(-31, "BAV with a channel orbital");
      const int vid = (k * np + p) * nb + c->bidx[ii - 1];
      const int ich = K3[i], io = K4[i];
      if (io > 0) {
        const int j1 = io / kIbo, j2 = io % kIbo;
        int w;
        if ((rc = get_v(c, j1, j2, w, !T.deferred))) return rc;
        const int jch = c->iech[j1 - 1];
        if ((rc = chk_ch(M, ich, jch, "UPDATE_HW"))) return rc;
        const int nVv = c->basis_nk * np * nb;
        if (ich >= jch) T.r1.push_back({blk_index(ich, jch), vid, nVv + w, -1, cd[i] * sc});
        else T.r1.push_back({blk_index(ich, jch), nVv + w, vid, -1, cd[i] * sc});
      } else if (io < 0) {
        if (ich < 1 || ich > M.nch) return fail(-13, "UPDATE_HV: channel index out of range");
        if (-io > M.ncp) return fail(-14, "UPDATE_HV: bound index out of range");
        T.vec.push_back({(-io - 1) * M.nch + (ich - 1), vid, -1, cd[i] * sc});
      } else {
        return fail(-23, " I_data: uknown structure for itype 2-5 ");
      }
    }
  } else {
    double CA = 0.0;
    for (int q = 0; q < n; ++q) {
      const int i = ipt ? ipt[q] - 1 : q;
      const int o1 = K2[i], o2 = K3[i];
      if (o1 < 1 || o1 > c->nbf || o2 < 1 || o2 > c->nbf || c->bidx[o1 - 1] < 0 || c->bidx[o2 - 1] < 0)
        return fail(-31, "R^k density of a channel orbital (no B-spline expansion)");
      const int ich = K1[i] / kIbi, jch = K1[i] % kIbi;
      int rc = chk_ch(M, ich, jch, "UPDATE_HX");
      if (rc) return rc;
      const int a = c->bidx[o1 - 1], b = c->bidx[o2 - 1];
      const double cc = cd[i] * sc;
      if (jcase <= 2) {
        int p;
        pair_index(c, o1, o2, p);
        T.band.push_back({blk_index(ich, jch), 1 + k * np + p, cc});
        CA += cd[i] * quadr(c, o1, o2, k);
      } else {
        const size_t nt = static_cast<size_t>(c->basis_nk) * nb * nb;
        const int t = k * nb * nb + (ich >= jch ? a * nb + b : b * nb + a);
        T.gemm_klo = std::min(T.gemm_klo, k);
        T.gemm_khi = std::max(T.gemm_khi, k);
        if (T.deferred) {
          T.gx.push_back({t + nt * blk_index(ich, jch), cc});
        } else {
        if (T.gemm_n == 0) {
          T.gemm_n = nt * M.nblk();
          if (T.gemm.reserve(T.gemm_n) != hipSuccess) return fail(-100, "HIP error: hipHostMalloc (exchange coefficients)");
          std::memset(T.gemm.p, 0, T.gemm_n * sizeof(double));
        }
        T.gemm.p[t + nt * blk_index(ich, jch)] += cc;
        }
      }
      const bool last = (q == n - 1) || (K1[ipt ? ipt[q + 1] - 1 : q + 1] != K1[i]);
      if (last) {
        if (jcase <= 2) {
          if (k > M.mk) return fail(-17, "UPDATE_CF: multipole index out of range");
          const size_t at = static_cast<size_t>(k) * M.nch * M.nch + (ich - 1) + static_cast<size_t>(M.nch) * (jch - 1);
          if (T.deferred) {
            T.ax.push_back({at, CA});
          } else {
            if (T.acf_add.empty()) T.acf_add.assign(M.acf.size(), 0.0);
            T.acf_add[at] += CA;
          }
        }
        CA = 0.0;
      }
    }
  }
  return 0;
}

}  // namespace

namespace bsrg {

// Before a term set or its arena is rewritten: the last upload from them
// has completed.
int wait_staged(Staged& st) {
  if (st.pending) {
    HIPCHK(hipEventSynchronize(st.done));
    st.pending = false;
  }
  return 0;
}

// Host worker threads per context: BSRG_HOST_THREADS (1..64) when set, else
// the hardware threads up to 8 (more made the list decoding of a C2-sized
// wave no faster on a 16-core share: its CPU time grows with the count).
int host_workers() {
  if (const char* e = std::getenv("BSRG_HOST_THREADS")) {
    const int n = std::atoi(e);
    if (n >= 1 && n <= 64) return n;
  }
  return static_cast<int>(std::min(8u, std::max(1u, std::thread::hardware_concurrency())));
}

// Stable counting sort of the terms of all parts (in part order) by key in
// [0, nkeys): position of each term in the sorted order (list order kept
// within a key) and the row pointers ptr[0..nkeys].
template <typename Vec, typename Key>
void counting_sort(const std::vector<Terms*>& parts, Vec vec, int nkeys, Key key, std::vector<int>& cnt, int* pos,
                   int* ptr) {
  cnt.assign(static_cast<size_t>(nkeys) + 1, 0);
  for (const Terms* P : parts)
    for (const auto& x : (P->*vec)) ++cnt[key(x) + 1];
  for (int k = 0; k < nkeys; ++k) cnt[k + 1] += cnt[k];
  std::memcpy(ptr, cnt.data(), (static_cast<size_t>(nkeys) + 1) * sizeof(int));
  size_t q = 0;
  for (const Terms* P : parts)
    for (const auto& x : (P->*vec)) pos[q++] = cnt[key(x)]++;
}

// Byte offsets of the CSR arrays of st (st.nband, nr1, nvec, nblk and nq
// set) in its arena, 256-byte aligned: doubles (coefficients) first, then
// ints, then the live map; sets st.bytes and returns it.
size_t layout_staged(Staged& st, bool live) {
  size_t o = 0;
  auto put = [&o](size_t count, size_t elem) { const size_t at = o; o += ((count * elem + 255) / 256) * 256; return at; };
  st.o_bc = put(st.nband, 8); st.o_rc = put(st.nr1, 8); st.o_vc = put(st.nvec, 8);
  st.o_bptr = put(st.nblk + 1, 4); st.o_bid = put(st.nband, 4);
  st.o_rptr = put(st.nblk + 1, 4); st.o_ru = put(st.nr1, 4); st.o_rw = put(st.nr1, 4); st.o_rs = put(st.nr1, 4);
  st.o_vptr = put(st.nq + 1, 4); st.o_vid = put(st.nvec, 4); st.o_vs = put(st.nvec, 4);
  st.has_live = live;
  st.o_live = put(live ? st.nblk : 0, 1);
  st.bytes = o;
  return o;
}

// Sort the terms of the parts by output and lay the CSR arrays out in the
// pinned arena: doubles (coefficients) first, then ints.  The band, rank-1
// and vector sets are independent and are staged concurrently when a pool
// is given.
int stage_terms(const std::vector<Terms*>& parts, Staged& st, int nblk, int nq, Pool* pool,
                const std::vector<unsigned char>* live) {
  if (int rc = wait_staged(st)) return rc;
  st.nband = st.nr1 = st.nvec = 0;
  for (const Terms* P : parts) {
    st.nband += P->band.size();
    st.nr1 += P->r1.size();
    st.nvec += P->vec.size();
  }
  st.nblk = nblk;
  st.nq = nq;
  size_t o = layout_staged(st, live && live->size() >= static_cast<size_t>(nblk));
  auto put = [&o](size_t count, size_t elem) { const size_t at = o; o += ((count * elem + 255) / 256) * 256; return at; };
  const size_t o_pb = put(st.nband, 4), o_pr = put(st.nr1, 4), o_pv = put(st.nvec, 4);
  if (o > st.h.cap) HIPCHK(st.h.reserve(o));
  if (st.has_live) std::memcpy(st.hp<unsigned char>(st.o_live), live->data(), nblk);
  const std::function<void(int)> task = [&](int which) {
    if (which == 0) {
      int* pos = st.hp<int>(o_pb);
      counting_sort(parts, &Terms::band, st.nblk, [](const Terms::Band& x) { return x.blk; }, st.cnt[0], pos,
                    st.hp<int>(st.o_bptr));
      double* bc = st.hp<double>(st.o_bc);
      int* bid = st.hp<int>(st.o_bid);
      size_t q = 0;
      for (const Terms* P : parts)
        for (const auto& x : P->band) { const int p = pos[q++]; bc[p] = x.c; bid[p] = x.bid; }
    } else if (which == 1) {
      int* pos = st.hp<int>(o_pr);
      counting_sort(parts, &Terms::r1, st.nblk, [](const Terms::R1& x) { return x.blk; }, st.cnt[1], pos,
                    st.hp<int>(st.o_rptr));
      double* rc = st.hp<double>(st.o_rc);
      int *ru = st.hp<int>(st.o_ru), *rw = st.hp<int>(st.o_rw), *rs = st.hp<int>(st.o_rs);
      size_t q = 0;
      for (const Terms* P : parts)
        for (const auto& x : P->r1) { const int p = pos[q++]; rc[p] = x.c; ru[p] = x.u; rw[p] = x.w; rs[p] = x.sidx; }
    } else {
      int* pos = st.hp<int>(o_pv);
      counting_sort(parts, &Terms::vec, st.nq, [](const Terms::Vec& x) { return x.q; }, st.cnt[2], pos,
                    st.hp<int>(st.o_vptr));
      double* vc = st.hp<double>(st.o_vc);
      int *vid = st.hp<int>(st.o_vid), *vs = st.hp<int>(st.o_vs);
      size_t q = 0;
      for (const Terms* P : parts)
        for (const auto& x : P->vec) { const int p = pos[q++]; vc[p] = x.c; vid[p] = x.v; vs[p] = x.sidx; }
    }
  };
  if (pool && pool->size() >= 3) {
    pool->run(3, task);
  } else {
    for (int w = 0; w < 3; ++w) task(w);
  }
  return 0;
}

}  // namespace bsrg

namespace {

// GET_V vectors registered so far -> the W part of the vector table.
int upload_wvec(bsrg_ctx* c) {
  const int ns = c->G.ns;
  hipStream_t s = c->stream;
  const size_t nV = static_cast<size_t>(c->basis_nk) * c->npair * c->nb;
  const size_t nW = c->wvec.size() / std::max(ns, 1);
  if (static_cast<int>(nW) == c->wvec_uploaded) return 0;
  if (c->VT.n < (nV + nW) * ns) {
    DevVec<double> nvt;
    HIPCHK(nvt.resize((nV + nW + 16) * ns));
    if (nV) HIPCHK(hipMemcpyAsync(nvt.p, c->VT.p, nV * ns * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    std::swap(nvt.p, c->VT.p);
    std::swap(nvt.n, c->VT.n);
  }
  if (nW) HIPCHK(hipMemcpyAsync(c->VT.p + nV * ns, c->wvec.data(), nW * ns * sizeof(double), hipMemcpyHostToDevice, s));
  c->wvec_uploaded = static_cast<int>(nW);
  return 0;
}

// nblk+1 zeros: the band-term pointers of a rank-1-only term pass
int ensure_nobands(bsrg_ctx* c, size_t nblk) {
  if (c->xzero.n < nblk + 1) {
    HIPCHK(c->xzero.resize(nblk + 1));
    launch_fill_int(c->xzero.n, 0, c->xzero.p, c->stream);
    HIPCHK(hipGetLastError());
  }
  return 0;
}

// The split-precision workspaces of a staged term set's GEMM (st.gemm_klo
// and st.gemm_khi set).
int gemm_workspaces(bsrg_ctx* c, const Matrices& M, Staged& st) {
  hipStream_t s = c->stream;
  {
    if (c->precision_mode == 1) {
      const int nb = c->nb;
      const size_t nss = static_cast<size_t>(c->G.ns) * c->G.ns;
      const int nr = (st.gemm_khi - st.gemm_klo + 1) * nb * nb;
      HIPCHK(c->ozk.resize(ozaki_workspace_bytes(static_cast<int>(nss), static_cast<int>(M.nblk()), nr,
                                                 c->nslices)));
    } else if (c->precision_mode == 2) {
      const int nss = c->G.ns * c->G.ns, K = c->exch_nk * c->nb * c->nb;
      if (c->oz8a_ver != c->e_ver || c->oz8a_S != c->nslices) {
        HIPCHK(c->oz8a.resize(ozaki8_a_bytes(nss, K, c->nslices)));
        if (int rc = ozaki8_prepare_a(nss, K, c->E.p, nss, c->nslices, c->oz8a.p, s)) return rc;
        c->oz8a_ver = c->e_ver;
        c->oz8a_S = c->nslices;
      }
      HIPCHK(c->oz8b.resize(ozaki8_b_bytes(static_cast<int>(M.nblk()), K, c->nslices)));
    }
  }
  return 0;
}

// The exchange coefficients of a term set (and the split-precision
// workspaces its GEMM needs): all a GEMM launch reads besides the bases.
int upload_gemm(bsrg_ctx* c, const Matrices& M, const Terms& T, Staged& st) {
  st.has_gemm = T.gemm_n > 0 && T.gemm_khi >= 0;
  if (!st.has_gemm) return 0;
  HIPCHK(st.gemm.resize(T.gemm_n));
  HIPCHK(hipMemcpyAsync(st.gemm.p, T.gemm.p, T.gemm_n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  st.gemm_klo = T.gemm_klo;
  st.gemm_khi = T.gemm_khi;
  return gemm_workspaces(c, M, st);
}

// The staged arena of a term set (and the GET_V vectors) for a later
// enqueue, after upload_gemm.
int upload_arena(bsrg_ctx* c, const Matrices& M, Staged& st) {
  hipStream_t s = c->stream;
  if (int rc = upload_wvec(c)) return rc;
  if (st.bytes) {
    HIPCHK(st.d.resize(st.bytes));
    HIPCHK(hipMemcpyAsync(st.d.p, st.h.p, st.bytes, hipMemcpyHostToDevice, s));
  }
  const bool isS = &M == &c->S;
  HIPCHK((isS ? c->bsumS : c->bsum).resize(M.nblk() * c->G.ns * c->G.ks));
  HIPCHK((isS ? c->rfacS : c->rfac).resize(std::max<size_t>(st.nr1, 1)));
  HIPCHK(hipEventRecord(st.done, s));
  st.pending = true;
  return 0;
}

// Upload a staged term set for a later enqueue: one copy of the exchange
// coefficients, one of the arena.
int upload_terms(bsrg_ctx* c, const Matrices& M, const Terms& T, Staged& st) {
  if (int rc = upload_gemm(c, M, T, st)) return rc;
  return upload_arena(c, M, st);
}

// Engine 4 (the exchange contraction fused with the term pass, xasm.hip)
// applies to an overwriting FP64 pass whose used exchange bases fit the LDS
// tile of the kernel.
int xasm_rows(bsrg_ctx* c, const Staged& st) {
  return (st.gemm_khi - st.gemm_klo + 1) * c->nb * c->nb;
}
bool xasm_eligible(bsrg_ctx* c, const Staged& st) {
  return c->precision_mode == 0 && st.has_gemm && xasm_rows(c, st) <= xasm_max_t(c->device);
}
int ensure_xasm(bsrg_ctx* c, const Matrices& M, const Staged& st) {
  const size_t nblk = M.nblk(), n = nblk * static_cast<size_t>(xasm_slot(xasm_rows(c, st)));
  HIPCHK(c->xt.resize(n));
  HIPCHK(c->xc.resize(n));
  HIPCHK(c->xn.resize(std::max<size_t>(nblk, 1)));
  return ensure_nobands(c, nblk);
}

// The exchange-contraction GEMM hcc = E^T C (+hcc): split precision in
// mode 1, else engine 0 = rocBLAS dgemm, 1/2/3 = FP64 MFMA kernel with
// 32/64/128-row workgroup tiles; engine 4 (overwrite only) packs the
// nonzeros of C per block for the fused pass, which does the product.
int run_gemm(bsrg_ctx* c, Matrices& M, const Staged& st, bool overwrite, int engine) {
  const int nb = c->nb;
  const size_t nss = static_cast<size_t>(c->G.ns) * c->G.ns, nblk = M.nblk();
  const size_t nt = static_cast<size_t>(c->basis_nk) * nb * nb;
  const size_t r0 = static_cast<size_t>(st.gemm_klo) * nb * nb;
  const int nr = (st.gemm_khi - st.gemm_klo + 1) * nb * nb;
  const double* A = c->E.p + r0 * nss;
  const double* B = st.gemm.p + r0;
  if (engine == 4 && c->precision_mode == 0) {
    if (overwrite && c->xn.n >= nblk && c->xt.n >= nblk * static_cast<size_t>(xasm_slot(nr))) {
      launch_coef_pack(static_cast<int>(nblk), nr, B, nt, c->xt.p, c->xc.p, c->xn.p, c->stream);
      if (launch_xprod(c->G.ns, static_cast<int>(nblk), nr, A, c->xt.p, c->xc.p, c->xn.p, M.hcc.p, c->stream))
        return fail(-100, "sparse exchange product: bad shape");
      return 0;
    }
    engine = 2;  // an accumulating product: the plain GEMM
  }
  if (c->precision_mode == 1)
    return ozaki_dgemm(static_cast<int>(nss), static_cast<int>(nblk), nr, A, nss, B, nt, M.hcc.p, nss, !overwrite,
                       c->nslices, c->ozk.p, c->stream);
  if (c->precision_mode == 2)  // all multipoles of the bases: their slices are prepared once
    return ozaki8_dgemm(static_cast<int>(nss), static_cast<int>(nblk), c->exch_nk * nb * nb, c->oz8a.p, st.gemm.p,
                        nt, M.hcc.p, nss, !overwrite, c->nslices, c->oz8b.p, c->stream);
  if (engine == 0) {
    const double alpha = 1.0, beta = overwrite ? 0.0 : 1.0;
    rocblas_status rs = rocblas_dgemm(c->blas, rocblas_operation_none, rocblas_operation_none, static_cast<int>(nss),
                                      static_cast<int>(nblk), nr, &alpha, A, static_cast<int>(nss), B,
                                      static_cast<int>(nt), &beta, M.hcc.p, static_cast<int>(nss));
    return rs == rocblas_status_success ? 0 : fail(-101, "rocblas_dgemm failed");
  }
  return mfma_dgemm(static_cast<int>(nss), static_cast<int>(nblk), nr, A, nss, B, nt, M.hcc.p, nss, !overwrite,
                    engine == 1 ? 1 : (engine == 2 ? 2 : 4), c->stream);
}

// Autotuner: time every FP64 engine once per GEMM shape class (overwriting
// the output, which the real launch rewrites) and keep the fastest.  The
// choice is shared by all contexts of the process; the class is (device,
// rows, inner size, column count rounded to a 1.25x geometric bucket,
// capture-safe), so jobs of similar size reuse it.  capture_safe excludes
// the library GEMM.
std::mutex g_tuned_mu;
std::map<std::tuple<int, int, int, int, int>, int> g_tuned;

int autotune_gemm(bsrg_ctx* c, Matrices& M, Staged& st, bool capture_safe) {
  if (!st.has_gemm || c->precision_mode != 0) return 0;
  if (c->forced_engine >= 0) {
    st.engine = c->forced_engine;
    if (st.engine == 4 && !xasm_eligible(c, st)) st.engine = 2;
    if (st.engine == 4) return ensure_xasm(c, M, st);
    return 0;
  }
  if (xasm_eligible(c, st)) {  // the fused pass writes H once: preferred when it fits
    st.engine = 4;
    return ensure_xasm(c, M, st);
  }
  const int nr = (st.gemm_khi - st.gemm_klo + 1) * c->nb * c->nb;
  const int bucket = static_cast<int>(std::floor(std::log(static_cast<double>(M.nblk())) / std::log(1.25)));
  const auto key = std::make_tuple(c->device, c->G.ns, nr, bucket, capture_safe ? 1 : 0);
  {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) { st.engine = it->second; return 0; }
  }
  float best = 1e30f;
  int best_e = 2;
  for (int e = capture_safe ? 1 : 0; e <= 3; ++e) {
    if (int rc = run_gemm(c, M, st, true, e)) return rc;  // warm-up
    (void)hipEventRecord(c->ev[6], c->stream);
    if (int rc = run_gemm(c, M, st, true, e)) return rc;
    (void)hipEventRecord(c->ev[7], c->stream);
    HIPCHK(hipEventSynchronize(c->ev[7]));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, c->ev[6], c->ev[7]);
    if (ms < best) { best = ms; best_e = e; }
  }
  {
    std::lock_guard<std::mutex> g(g_tuned_mu);
    g_tuned[key] = best_e;
  }
  st.engine = best_e;
  return 0;
}

// Device-only work of an uploaded term set (no host synchronisation, no
// allocation): the part of a wave that a hipGraph captures.  The exchange
// GEMM goes first (writing every block on an overwrite), then the term pass
// adds the band and rank-1 terms to what the blocks hold; without a GEMM the
// term pass writes the blocks itself.  gemm_done: the GEMM of this term set
// was already enqueued (enqueue_gemm).
int enqueue_gemm(bsrg_ctx* c, Matrices& M, Staged& st, bool overwrite, hipEvent_t e0, hipEvent_t e1) {
  if (e0) (void)hipEventRecord(e0, c->stream);
  if (int rc = run_gemm(c, M, st, overwrite, st.engine)) return rc;
  c->last_engine = c->precision_mode == 0 ? st.engine : 10 * c->precision_mode + c->nslices;
  if (e1) (void)hipEventRecord(e1, c->stream);
  return 0;
}

int enqueue_terms(bsrg_ctx* c, Matrices& M, Staged& st, bool overwrite, hipEvent_t e0, hipEvent_t e1,
                  bool gemm_done = false) {
  const int ns = c->G.ns, ks = c->G.ks;
  hipStream_t s = c->stream;
  const bool fused = st.has_gemm && !gemm_done && st.engine == 4 && overwrite && c->precision_mode == 0;
  if (fused) {
    // the product phase is the product kernel itself (events around it below)
  } else if (st.has_gemm && !gemm_done) {
    if (int rc = enqueue_gemm(c, M, st, overwrite, e0, e1)) return rc;
  } else if (!st.has_gemm) {
    if (e0) (void)hipEventRecord(e0, s);  // an empty GEMM phase
    if (e1) (void)hipEventRecord(e1, s);
  }
  if (st.has_gemm && !gemm_done && st.engine == 4 && overwrite && c->precision_mode == 0) {
    // engine 4 with nothing enqueued yet (a prepared or captured wave): the
    // exchange product and the band terms in one pass that writes every
    // block once, then the rank-1 terms; a wave whose product already ran
    // (started during the host staging) takes the term pass below
    launch_coef_pack(st.nblk, xasm_rows(c, st), st.gemm.p + static_cast<size_t>(st.gemm_klo) * c->nb * c->nb,
                     static_cast<size_t>(c->basis_nk) * c->nb * c->nb, c->xt.p, c->xc.p, c->xn.p, s);
    c->last_engine = 4;
    const int nb = c->nb;
    const size_t r0 = static_cast<size_t>(st.gemm_klo) * nb * nb, nss = static_cast<size_t>(ns) * ns;
    const int xr = launch_xasm(ns, ks, st.nblk, xasm_rows(c, st), c->E.p + r0 * nss, c->xt.p, c->xc.p, c->xn.p,
                               st.dp<int>(st.o_bptr), st.dp<int>(st.o_bid), st.dp<double>(st.o_bc), c->bands.p,
                               st.dp<int>(st.o_rptr), st.dp<int>(st.o_ru), st.dp<int>(st.o_rw), st.dp<int>(st.o_rs),
                               st.dp<double>(st.o_rc), c->St.p, c->VT.p, M.hcc.p,
                               (&M == &c->S ? c->bsumS : c->bsum).p, (&M == &c->S ? c->rfacS : c->rfac).p,
                               st.has_live && (!c->capturing || c->live_in_capture)
                                   ? st.dp<unsigned char>(st.o_live)
                                   : nullptr,
                               c->xzero.p, s, e0, e1);
    if (xr) return fail(-100, "fused exchange pass: bad shape");
    launch_vec_assemble(ns, st.nq, 0, st.dp<int>(st.o_vptr), st.dp<int>(st.o_vid), st.dp<int>(st.o_vs),
                        st.dp<double>(st.o_vc), c->St.p, c->VT.p, M.hcb.p, s);
    HIPCHK(hipGetLastError());
    return 0;
  }
  // after a GEMM, or when accumulating, the pass adds to the blocks and
  // leaves blocks without terms alone; an overwriting pass without a GEMM
  // writes every block, except empty ones that already hold zeros
  const bool add = st.has_gemm || !overwrite;
  const bool skip = !add && st.has_live && (!c->capturing || c->live_in_capture);
  launch_assemble(ns, ks, st.nblk, add ? 1 : 0, st.dp<int>(st.o_bptr), st.dp<int>(st.o_bid),
                  st.dp<double>(st.o_bc), c->bands.p, st.dp<int>(st.o_rptr), st.dp<int>(st.o_ru), st.dp<int>(st.o_rw),
                  st.dp<int>(st.o_rs), st.dp<double>(st.o_rc), c->St.p, c->VT.p, M.hcc.p,
                  (&M == &c->S ? c->bsumS : c->bsum).p, (&M == &c->S ? c->rfacS : c->rfac).p, s, nullptr, 0, 0,
                  skip ? st.dp<unsigned char>(st.o_live) : nullptr);
  launch_vec_assemble(ns, st.nq, overwrite ? 0 : 1, st.dp<int>(st.o_vptr), st.dp<int>(st.o_vid),
                      st.dp<int>(st.o_vs), st.dp<double>(st.o_vc), c->St.p, c->VT.p, M.hcb.p, s);
  HIPCHK(hipGetLastError());
  return 0;
}

// One-shot waves (bsrg_wave_assemble) start the exchange GEMM of H as soon
// as the exchange records are decoded: it needs only the coefficients, and
// runs while the other records are decoded and the terms sorted and
// uploaded.  A captured launch enqueues the GEMM itself.
int start_early_gemm(bsrg_ctx* c) {
  if (int rc = upload_gemm(c, c->H, c->TH, c->stH)) return rc;
  if (int rc = autotune_gemm(c, c->H, c->stH, c->capture_safe)) return rc;
  if (c->stH.has_gemm) {
    if (int rc = enqueue_gemm(c, c->H, c->stH, true, c->ev[3], c->ev[4])) return rc;
    c->gemm_enqueued = true;
  }
  return 0;
}

// After an overwriting term pass of M (and its GEMM) is enqueued: which
// blocks may now hold nonzeros.  The pass wrote zeros into the empty
// blocks it visited; a GEMM may add to any block; a captured pass may never
// run, so its blocks are unknown.
void update_live(bsrg_ctx* c, Matrices& M, Staged& st) {
  if (M.live.size() < static_cast<size_t>(st.nblk)) M.live.resize(st.nblk, 1);
  const int* bp = st.hp<int>(st.o_bptr);
  const int* rp = st.hp<int>(st.o_rptr);
  const bool unknown = st.has_gemm || (c->capturing && !c->live_in_capture);
  for (int b = 0; b < st.nblk; ++b)
    M.live[b] = unknown || bp[b + 1] > bp[b] || rp[b + 1] > rp[b] ? 1 : 0;
}

// The term pass of S (overwriting) and its diagonal symmetrisation.
int enqueue_s(bsrg_ctx* c) {
  if (int rc = enqueue_terms(c, c->S, c->stS, true, nullptr, nullptr)) return rc;
  update_live(c, c->S, c->stS);
  launch_symmetrize_diag(c->G.ns, c->S.nch, c->S.hcc.p, c->stream);
  return 0;
}

// Host scalars (HB, Target_h, ACF) of the parts in list order; needs
// St_host, i.e. a completed stream.  The ACF additions are in parts[0].
void finish_scalars(bsrg_ctx* c, Matrices& M, const std::vector<Terms*>& parts, bool overwrite) {
  if (overwrite) std::fill(M.hbb.begin(), M.hbb.end(), 0.0);
  for (const Terms* P : parts)
    for (const auto& sc : P->scal) {
      const double f = sc.c * (sc.sidx >= 0 ? c->St_host.p[sc.sidx] : 1.0);
      if (sc.kind == 0) M.hbb[sc.idx] += f;
      else if (sc.kind == 1) M.htarg[sc.idx] += f;
      else M.otarg[sc.idx] += f;
    }
  const Terms& T = *parts[0];
  if (!T.acf_add.empty())
    for (size_t q = 0; q < M.acf.size(); ++q) M.acf[q] += T.acf_add[q];
}

// Per-call path: stage, launch, wait, scalars.
int apply_terms(bsrg_ctx* c, Matrices& M, Terms& T, bool overwrite) {
  const std::vector<Terms*> one{&T};
  if (int rc = stage_terms(one, c->stC, static_cast<int>(M.nblk()), M.nch * M.ncp, nullptr)) return rc;
  if (int rc = upload_terms(c, M, T, c->stC)) return rc;
  c->stC.engine = c->forced_engine >= 0 ? c->forced_engine : 2;
  if (overwrite)
    if (int rc = autotune_gemm(c, M, c->stC, false)) return rc;
  if (int rc = enqueue_terms(c, M, c->stC, overwrite, nullptr, nullptr)) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  finish_scalars(c, M, one, overwrite);
  return 0;
}

// GET_V vectors of every record that needs one, registered in list order
// before the parallel pass (which then only reads the registry).  Failures
// are left for the pass itself to report at the right record.
void register_vectors(bsrg_ctx* c, int nrec, const int* jtype, const int* K3, const int* K4) {
  int id;
  for (int r = 0; r < nrec; ++r) {
    if (jtype[r] == 1) {
      const int ic = -K3[r], jc = -K4[r];
      if (ic < 0) (void)get_v(c, (-ic) / kIbo, (-ic) % kIbo, id);
      if (ic < 0 && jc < 0) (void)get_v(c, (-jc) / kIbo, (-jc) % kIbo, id);
    } else if (jtype[r] <= 5 && K4[r] > 0) {
      (void)get_v(c, K4[r] / kIbo, K4[r] % kIbo, id);
    }
  }
}

// I_data over all lists of a wave on the context's host workers.  Records
// are cut into contiguous ranges at list boundaries or where K1 changes (so
// the ACF group sums of I_data stay whole); each worker decodes its range
// into its own part, and the parts are kept in range order, which is list
// order.  Exchange coefficients and ACF additions are then applied in that
// order.  The workers only read the GET_V registry; vectors missing from it
// (the first wave of new orbitals) are registered in list order and the
// pass is repeated.
// Exchange records (jtype 8, 9) only add exchange coefficients.
inline bool exch_type(int jtype) { return jtype >= 8; }

int idata_parallel(bsrg_ctx* c, int nrec, const int* jtype, const int* kpol, const int* k1, const int* k2,
                   const int* k3, const int* k4, const double* cd, bool any_exch) {
  const auto t1 = std::chrono::steady_clock::now();
  Terms& TH = c->TH;
  const Matrices& M = c->H;
  // lazily built tables the workers read
  for (int r = 0; r < nrec; ++r)
    if (jtype[r] == 6 || jtype[r] == 7) {
      for (int m = 0; m <= M.mk; ++m) (void)quadr(c, 1, 1, m);
      break;
    }
  const auto t2 = std::chrono::steady_clock::now();
  if (!c->pool) c->pool.reset(new Pool(host_workers()));
  // more ranges than workers, taken dynamically: records differ in cost;
  // a range holds exchange records only or none
  const int nth = static_cast<int>(std::min<long>(4L * c->pool->size(), nrec / 4096L + 1));
  std::vector<int> cuts{0};
  for (int t = 1; t < nth; ++t) {
    int b = static_cast<int>(static_cast<long>(nrec) * t / nth);
    if (b <= cuts.back()) continue;
    while (b < nrec && jtype[b] == jtype[b - 1] && kpol[b] == kpol[b - 1] && k1[b] == k1[b - 1]) ++b;
    if (b < nrec) cuts.push_back(b);
  }
  cuts.push_back(nrec);
  {
    std::vector<int> split{0};
    for (size_t q = 0; q + 1 < cuts.size(); ++q) {
      for (int r = cuts[q] + 1; r < cuts[q + 1]; ++r)
        if (exch_type(jtype[r]) != exch_type(jtype[r - 1])) split.push_back(r);
      split.push_back(cuts[q + 1]);
    }
    cuts.swap(split);
  }
  const int np = static_cast<int>(cuts.size()) - 1;
  while (static_cast<int>(c->work.size()) < np) c->work.emplace_back(new Terms());
  c->nwork = np;
  std::vector<int> pexch, pother;  // part indices by kind, in list order
  for (int t = 0; t < np; ++t) (nrec && exch_type(jtype[cuts[t]]) ? pexch : pother).push_back(t);
  const size_t nt = static_cast<size_t>(c->basis_nk) * c->nb * c->nb;
  if (any_exch && c->nb > 0) {
    TH.gemm_n = nt * M.nblk();
    HIPCHK(TH.gemm.reserve(TH.gemm_n));
  }
  std::vector<double> part_ms(np, 0.0);
  auto decode = [&](int t) {
    const auto p0 = std::chrono::steady_clock::now();
    struct Stamp {
      std::chrono::steady_clock::time_point t0;
      double* out;
      ~Stamp() { *out += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
    } stamp{p0, &part_ms[t]};
    Terms& P = *c->work[t];
    P.clear();
    P.deferred = true;
    int r = cuts[t];
    while (r < cuts[t + 1]) {
      int e = r + 1;
      while (e < cuts[t + 1] && jtype[e] == jtype[r] && kpol[e] == kpol[r]) ++e;
      if (int rc = add_idata(c, M, P, jtype[r], kpol[r], e - r, nullptr, k1 + r, k2 + r, k3 + r, k4 + r, cd + r, 1.0)) {
        P.err = rc;
        P.msg = g_err;
        return;
      }
      r = e;
    }
  };
  // phase 1: the exchange parts, while the coefficient matrix is zeroed in
  // slices; their coefficients are summed in list order, and a one-shot
  // wave starts its exchange GEMM now
  const int nx = static_cast<int>(pexch.size()), nz = std::max(nx, 1);
  const std::function<void(int)> phase1 = [&](int i) {
    if (TH.gemm_n) {
      const size_t lo = TH.gemm_n * i / nz, hi = TH.gemm_n * (i + 1) / nz;
      std::memset(TH.gemm.p + lo, 0, (hi - lo) * sizeof(double));
    }
    if (i < nx) decode(pexch[i]);
  };
  c->pool->run(nz, phase1);
  const auto tp1 = std::chrono::steady_clock::now();
  for (int t : pexch) {
    const Terms& P = *c->work[t];
    if (P.err) return fail(P.err, P.msg);
    for (const auto& g : P.gx) TH.gemm.p[g.first] += g.second;
    TH.gemm_klo = std::min(TH.gemm_klo, P.gemm_klo);
    TH.gemm_khi = std::max(TH.gemm_khi, P.gemm_khi);
  }
  if (TH.gemm_khi < 0) TH.gemm_n = 0;
  const auto tp2 = std::chrono::steady_clock::now();
  if (c->early_gemm)
    if (int rc = start_early_gemm(c)) return rc;
  const auto tp3 = std::chrono::steady_clock::now();
  // phase 2: everything else; GET_V vectors not yet in the registry (new
  // orbitals) are registered in list order and the phase is repeated
  const std::function<void(int)> phase2 = [&](int i) { decode(pother[i]); };
  c->pool->run(static_cast<int>(pother.size()), phase2);
  bool missing = false;
  for (int t : pother) missing = missing || c->work[t]->err == kNotRegistered;
  if (missing) {
    register_vectors(c, nrec, jtype, k3, k4);
    c->pool->run(static_cast<int>(pother.size()), phase2);
  }
  const auto t3 = std::chrono::steady_clock::now();
  for (int t : pother) {
    const Terms& P = *c->work[t];
    if (P.err) return fail(P.err, P.msg);
  }
  for (int t = 0; t < np; ++t) {
    const Terms& P = *c->work[t];
    if (!P.ax.empty() && TH.acf_add.empty()) TH.acf_add.assign(M.acf.size(), 0.0);
    for (const auto& a : P.ax) TH.acf_add[a.first] += a.second;
  }
  const auto t4 = std::chrono::steady_clock::now();
  auto ms = [](std::chrono::steady_clock::time_point a, std::chrono::steady_clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
  };
  c->host_ms[2] = ms(t1, t2);
  c->host_ms[3] = ms(t2, t3);
  c->host_ms[4] = ms(t3, t4);
  c->host_ms[10] = ms(t2, tp1);   // exchange records
  c->host_ms[11] = ms(tp1, tp2);  // their coefficients summed
  c->host_ms[12] = ms(tp2, tp3);  // early product enqueued
  c->host_ms[13] = ms(tp3, t3);   // the other records
  c->host_ms[8] = 0.0;
  c->host_ms[9] = 0.0;
  for (double x : part_ms) { c->host_ms[8] += x; c->host_ms[9] = std::max(c->host_ms[9], x); }
  return 0;
}

void band_hl(Terms& T, int ich, int jch, double c) { T.band.push_back({blk_index(ich, jch), 0, c}); }

int add_odata(bsrg_ctx* c, const Matrices& M, Terms& T, int n, const int* ipt, const int* K3, const int* K4,
              const double* cd, double eps_c, double sc, bool targets) {
  for (int q = 0; q < n; ++q) {
    const int i = ipt ? ipt[q] - 1 : q;
    if (std::fabs(cd[i]) < eps_c) continue;
    int rc = add_struct1(c, M, T, K3[i], K4[i], -1, cd[i] * sc, true, targets);
    if (rc) return rc;
  }
  return 0;
}

// ---- R^k lists decoded on the device (lists_dev.hip) ----

// The host decoder's code and text for a record the device decoder rejects.
const char* dl_message(int id, int& code) {
  switch (id) {
    case kDlItype: code = -21; return " I_data: unknown itype ";
    case kDlMultipole: code = -17; return "UPDATE_CF: multipole index out of range";
    case kDlPrepared: code = -33; return "multipole outside the prepared range";
    case kDlOrbRange: code = -30; return "orbital index out of range";
    case kDlChanDensity: code = -31; return "R^k density of a channel orbital (no B-spline expansion)";
    case kDlHb: code = -12; return "UPDATE_HB: indeces out of range";
    case kDlGetvRange: code = -18; return "get_V: orbital index out of range";
    case kDlGetvZero: code = -19; return "get_V: <kl|nl> = 0 ?";
    case kDlHvChannel: code = -13; return "UPDATE_HV: channel index out of range";
    case kDlHvBound: code = -14; return "UPDATE_HV: bound index out of range";
    case kDlHwChannel: code = -10; return "UPDATE_HW: channel index out of range";
    case kDlCfChannel: code = -10; return "UPDATE_CF: channel index out of range";
    case kDlStruct1: code = -22; return " I_data: unknown structure for itype=1";
    case kDlBav: code = -31; return "BAV with a channel orbital";
    case kDlStruct25: code = -23; return " I_data: uknown structure for itype 2-5 ";
    case kDlHxChannel: code = -10; return "UPDATE_HX: channel index out of range";
    default: code = -100; return "device list decoding: unknown rejection";
  }
}

// The smallest b with n < 2^b.
int bits_for(size_t n) {
  int b = 1;
  while (b < 63 && (size_t{1} << b) <= n) ++b;
  return b;
}

// Orbital tables of the device decoder (once per orbital set) and
// <p_a|r^k|p_b> over the bound orbitals for k <= mk.
int dl_tables(bsrg_ctx* c, int mk) {
  auto& R = c->rl;
  const int nbf = c->nbf, ns = c->G.ns, nb = c->nb;
  hipStream_t s = c->stream;
  if (R.tab_ver != c->orb_ver) {
    std::vector<int> it(3 * static_cast<size_t>(nbf) + static_cast<size_t>(nbf) * nbf);
    std::copy(c->bidx.begin(), c->bidx.end(), it.begin());
    std::copy(c->iech.begin(), c->iech.end(), it.begin() + nbf);
    std::copy(c->lbs.begin(), c->lbs.end(), it.begin() + 2 * nbf);
    std::copy(c->ibort.begin(), c->ibort.end(), it.begin() + 3 * nbf);
    std::vector<double> dt(c->obs);
    dt.insert(dt.end(), c->qbs.begin(), c->qbs.end());
    HIPCHK(R.itab.upload(it.data(), it.size(), s));
    HIPCHK(R.dtab.upload(dt.data(), dt.size(), s));
    HIPCHK(hipStreamSynchronize(s));  // the host vectors go out of scope
    R.tab_ver = c->orb_ver;
    R.qb_mk = -1;
  }
  if (R.qb_mk < mk) {
    std::vector<double> q(static_cast<size_t>(mk + 1) * std::max(nb, 1) * std::max(nb, 1), 0.0);
    for (int k = 0; k <= mk; ++k)
      for (int a = 0; a < nb; ++a)
        for (int b = 0; b < nb; ++b)
          q[(static_cast<size_t>(k) * nb + a) * nb + b] = quadr(c, c->bound[a], c->bound[b], k);
    HIPCHK(R.Qb.upload(q.data(), q.size(), s));
    HIPCHK(hipStreamSynchronize(s));
    R.qb_mk = mk;
  }
  (void)ns;
  return 0;
}

// The H terms of the resident R^k lists, decoded on the device, with the
// host terms of H before (TH) and after (THt) them: the term set stH (CSR
// arrays in its device arena, exchange coefficients in its gemm), GET_V
// vectors in the vector table, and the host scalars and ACF sums, read back
// for bsrg_wave_finish.
int dl_decode(bsrg_ctx* c) {
  auto& R = c->rl;
  Matrices& M = c->H;
  Staged& st = c->stH;
  hipStream_t s = c->stream;
  const int nbf = c->nbf, ns = c->G.ns, ks = c->G.ks, nb = c->nb, np = c->npair;
  if (int rc = dl_tables(c, M.mk)) return rc;
  // vector table: the V part, the GET_V vectors of the host terms, then one
  // slot per orbital pair for the records' GET_V vectors
  const size_t nV = static_cast<size_t>(c->basis_nk) * np * nb;
  const size_t nW = c->wvec.size() / std::max(ns, 1);
  const size_t w0 = nV + nW, nvt = w0 + static_cast<size_t>(nbf) * nbf;
  if (w0 + static_cast<size_t>(nbf) * nbf >= (size_t{1} << 31)) return fail(-1, "device list decoding: vector table too large");
  if (c->VT.n < nvt * ns) {
    DevVec<double> nvtab;
    HIPCHK(nvtab.resize((nvt + 16) * ns));
    if (nV) HIPCHK(hipMemcpyAsync(nvtab.p, c->VT.p, nV * ns * sizeof(double), hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    std::swap(nvtab.p, c->VT.p);
    std::swap(nvtab.n, c->VT.n);
    c->wvec_uploaded = -1;
  }
  if (int rc = upload_wvec(c)) return rc;
  // the host terms as items, in the order of their term lists
  const Terms* host[2] = {&c->TH, &c->THt};
  size_t nh[2];
  for (int h = 0; h < 2; ++h) {
    const Terms& T = *host[h];
    nh[h] = T.band.size() + T.r1.size() + T.vec.size() + T.scal.size();
    for (const auto& x : T.scal)
      if (x.kind == 2) return fail(-100, "device list decoding: overlap target term in H");
  }
  const size_t nrec = static_cast<size_t>(R.nrec), nitems = nh[0] + 2 * nrec + nh[1];
  if (nitems >= (size_t{1} << 31)) return fail(-1, "device list decoding: too many records");
  const size_t nblk = M.nblk(), nq = static_cast<size_t>(M.nch) * M.ncp;
  const size_t nt = static_cast<size_t>(c->basis_nk) * nb * nb;
  R.n_hbb = M.hbb.size();
  R.n_htarg = nblk;
  R.n_acf = M.acf.size();
  const size_t kmax = std::max({nblk + 1, nq + 1, R.n_hbb, R.n_acf, R.any_exch ? nt * nblk : size_t{0}});
  const int kb = bits_for(kmax);
  if (kb + 3 > 32) return fail(-1, "device list decoding: too many outputs for the 32-bit sort keys");
  const int nhi = static_cast<int>(nh[0] + nh[1]);
  if (R.hitems_pending) {
    HIPCHK(hipEventSynchronize(R.hitems_done));
    R.hitems_pending = false;
  }
  const size_t hbytes = dl_host_item_bytes(nhi);
  HIPCHK(R.hitems.reserve(std::max<size_t>(hbytes, 8)));
  {
    double* sc = reinterpret_cast<double*>(R.hitems.p);
    unsigned* sk = reinterpret_cast<unsigned*>(sc + nhi);
    int* sa = reinterpret_cast<int*>(sk + nhi);
    int* sb = sa + nhi;
    int* ss = sb + nhi;
    size_t t = 0;
    auto put = [&](int cat, size_t idx, int a, int b, int sidx, double v) {
      sk[t] = (static_cast<unsigned>(cat) << kb) | static_cast<unsigned>(idx);
      sa[t] = a; sb[t] = b; ss[t] = sidx; sc[t] = v;
      ++t;
    };
    for (const Terms* T : host) {
      for (const auto& x : T->band) put(kDlCatBand, x.blk, x.bid, 0, -1, x.c);
      for (const auto& x : T->r1) put(kDlCatR1, x.blk, x.u, x.w, x.sidx, x.c);
      for (const auto& x : T->vec) put(kDlCatVec, x.q, x.v, 0, x.sidx, x.c);
      for (const auto& x : T->scal) put(x.kind == 0 ? kDlCatHbb : kDlCatHtarg, x.idx, 0, 0, x.sidx, x.c);
    }
  }
  // the decoding runs on the second stream, beside the band and basis
  // kernels; only the run sums (which read the scalar R^k table) wait for them
  hipStream_t ds = c->copy_stream;
  if (!R.hitems_done) HIPCHK(hipEventCreateWithFlags(&R.hitems_done, hipEventDisableTiming));
  if (!R.dec_done) HIPCHK(hipEventCreateWithFlags(&R.dec_done, hipEventDisableTiming));
  HIPCHK(R.hstage.resize(std::max<size_t>(hbytes, 8)));
  if (hbytes) HIPCHK(hipMemcpyAsync(R.hstage.p, R.hitems.p, hbytes, hipMemcpyHostToDevice, ds));
  HIPCHK(hipEventRecord(R.hitems_done, ds));
  R.hitems_pending = true;
  // item and sort arrays
  for (DevVec<unsigned>* v : {&R.key, &R.skey, &R.sval, &R.iota}) HIPCHK(v->resize(std::max<size_t>(nitems, 1)));
  for (DevVec<int>* v : {&R.ia, &R.ib, &R.is}) HIPCHK(v->resize(std::max<size_t>(nitems, 1)));
  HIPCHK(R.ic.resize(std::max<size_t>(nitems, 1)));
  HIPCHK(R.need.resize(static_cast<size_t>(nbf) * nbf));
  HIPCHK(R.cb.resize(kDlNumCat + 1));
  HIPCHK(R.err.resize(1));
  if (R.sort_items != static_cast<int>(nitems) || R.sort_bits != kb + 3) {
    R.sort_tmp_bytes = dl_sort_bytes(static_cast<int>(nitems), kb + 3);
    HIPCHK(R.sort_tmp.resize(std::max<size_t>(R.sort_tmp_bytes, 1)));
    R.sort_items = static_cast<int>(nitems);
    R.sort_bits = kb + 3;
  }
  const size_t nout = R.n_hbb + R.n_htarg + R.n_acf;
  HIPCHK(R.outs.resize(std::max<size_t>(nout, 1)));
  launch_fill(std::max<size_t>(nout, 1), 0.0, R.outs.p, ds);
  // the term set: CSR capacities of every item, exchange coefficients
  if (int rc = wait_staged(st)) return rc;
  st.nblk = static_cast<int>(nblk);
  st.nq = static_cast<int>(nq);
  st.nband = st.nr1 = st.nvec = nitems;
  layout_staged(st, false);
  HIPCHK(st.d.resize(st.bytes));
  st.has_gemm = R.any_exch && nb > 0 && R.xkhi >= 0;
  if (st.has_gemm) {
    st.gemm_klo = std::max(0, R.xklo);
    st.gemm_khi = std::min(R.xkhi, c->basis_nk - 1);
    if (st.gemm_khi < st.gemm_klo) st.has_gemm = false;
  }
  if (st.has_gemm) {
    HIPCHK(st.gemm.resize(nt * nblk));
    launch_fill(nt * nblk, 0.0, st.gemm.p, ds);
    if (int rc = gemm_workspaces(c, M, st)) return rc;
  }
  HIPCHK(c->bsum.resize(nblk * ns * ks));
  HIPCHK(c->rfac.resize(std::max<size_t>(nitems, 1)));
  DecodeArgs a;
  a.nrec = R.nrec;
  a.base = static_cast<int>(nh[0]);
  a.jt = R.rec.p;
  a.kp = R.rec.p + nrec;
  a.k1 = R.rec.p + 2 * nrec;
  a.k2 = R.rec.p + 3 * nrec;
  a.k3 = R.rec.p + 4 * nrec;
  a.k4 = R.rec.p + 5 * nrec;
  a.cd = R.cd.p;
  a.nbf = nbf; a.nb = nb; a.np = np; a.nch = M.nch; a.ncp = M.ncp; a.mk = M.mk; a.basis_nk = c->basis_nk; a.ns = ns;
  a.bidx = R.itab.p; a.iech = R.itab.p + nbf; a.lbs = R.itab.p + 2 * nbf; a.ibort = R.itab.p + 3 * nbf;
  a.obs = R.dtab.p; a.qbs = R.dtab.p + static_cast<size_t>(nbf) * nbf;
  a.Qb = R.Qb.p;
  a.St = c->St.p;
  a.eps_ovl = c->eps_ovl;
  a.w0 = static_cast<int>(w0);
  a.nt = static_cast<long>(nt);
  a.kb = kb;
  a.key = R.key.p; a.ia = R.ia.p; a.ib = R.ib.p; a.is = R.is.p; a.ic = R.ic.p;
  a.need = R.need.p;
  a.err = R.err.p;
  launch_dl_host_items(static_cast<int>(nh[0]), static_cast<int>(nh[1]), nh[0] + 2 * nrec, R.hstage.p, a, ds);
  launch_dl_iota(nitems, R.iota.p, ds);
  DecodeSort so;
  so.tmp = R.sort_tmp.p; so.tmp_bytes = R.sort_tmp_bytes;
  so.skey = R.skey.p; so.sval = R.sval.p; so.iota = R.iota.p;
  so.nitems = static_cast<int>(nitems);
  DecodeCsr o;
  o.nblk = static_cast<int>(nblk); o.nq = static_cast<int>(nq);
  char* d = st.d.p;
  o.bc = reinterpret_cast<double*>(d + st.o_bc); o.rc = reinterpret_cast<double*>(d + st.o_rc);
  o.vc = reinterpret_cast<double*>(d + st.o_vc);
  o.bptr = reinterpret_cast<int*>(d + st.o_bptr); o.bid = reinterpret_cast<int*>(d + st.o_bid);
  o.rptr = reinterpret_cast<int*>(d + st.o_rptr); o.ru = reinterpret_cast<int*>(d + st.o_ru);
  o.rw = reinterpret_cast<int*>(d + st.o_rw); o.rs = reinterpret_cast<int*>(d + st.o_rs);
  o.vptr = reinterpret_cast<int*>(d + st.o_vptr); o.vid = reinterpret_cast<int*>(d + st.o_vid);
  o.vs = reinterpret_cast<int*>(d + st.o_vs);
  DecodeOut out;
  out.cb = R.cb.p;
  out.gemm = st.has_gemm ? st.gemm.p : nullptr;
  out.hbb = R.outs.p;
  out.htarg = R.outs.p + R.n_hbb;
  out.acf = R.outs.p + R.n_hbb + R.n_htarg;
  out.VT = c->VT.p;
  if (launch_dl_decode(a, so, o, out, ds)) return fail(-100, "device list decoding: launch failed");
  HIPCHK(hipEventRecord(R.dec_done, ds));
  HIPCHK(hipStreamWaitEvent(s, R.dec_done, 0));
  if (launch_dl_runs(a, so, out, s)) return fail(-100, "device list decoding: launch failed");
  // category starts, rejection word and sums, read back by bsrg_wave_finish
  const size_t back = 64 + nout * sizeof(double);
  if (back > R.back.cap) {
    HIPCHK(hipStreamSynchronize(s));
    HIPCHK(R.back.reserve(back));
  }
  HIPCHK(hipMemcpyAsync(R.back.p, R.cb.p, (kDlNumCat + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(R.back.p + 48, R.err.p, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  if (nout) HIPCHK(hipMemcpyAsync(R.back.p + 64, R.outs.p, nout * sizeof(double), hipMemcpyDeviceToHost, s));
  return 0;
}

// After the stream of a resident wave completed: the rejection word, the
// term counts, and the host scalars of H (HB, target matrix, ACF sums).
int dl_finish(bsrg_ctx* c) {
  auto& R = c->rl;
  Matrices& M = c->H;
  const int* cb = reinterpret_cast<const int*>(R.back.p);
  unsigned long long err;
  std::memcpy(&err, R.back.p + 48, sizeof(err));
  if (err != ~0ull) {
    int code;
    const char* text = dl_message(static_cast<int>(err & 0xFF), code);
    return fail(code, text);
  }
  c->stH.nband = static_cast<size_t>(cb[kDlCatR1] - cb[kDlCatBand]);
  c->stH.nr1 = static_cast<size_t>(cb[kDlCatVec] - cb[kDlCatR1]);
  c->stH.nvec = static_cast<size_t>(cb[kDlCatVec + 1] - cb[kDlCatVec]);
  const double* v = reinterpret_cast<const double*>(R.back.p + 64);
  std::fill(M.hbb.begin(), M.hbb.end(), 0.0);
  for (size_t q = 0; q < R.n_hbb && q < M.hbb.size(); ++q) M.hbb[q] += v[q];
  v += R.n_hbb;
  for (size_t q = 0; q < R.n_htarg && q < M.htarg.size(); ++q) M.htarg[q] += v[q];
  v += R.n_htarg;
  for (size_t q = 0; q < R.n_acf && q < M.acf.size(); ++q) M.acf[q] += v[q];
  return 0;
}

void acf_finish(Matrices& M, int kcore) {
  const int nch = M.nch;
  const double eps_acf = 1e-5;
  for (int k = 0; k <= M.mk; ++k)
    for (int i = 0; i < nch; ++i)
      for (int j = i; j < nch; ++j) {
        double* a = &M.acf[static_cast<size_t>(k) * nch * nch];
        double v = a[i + static_cast<size_t>(nch) * j] + a[j + static_cast<size_t>(nch) * i];
        if (i != j) v *= 2.0;
        if (std::fabs(v) < eps_acf) v = 0.0;
        a[i + static_cast<size_t>(nch) * j] = v;
        a[j + static_cast<size_t>(nch) * i] = v;
      }
  for (int i = 0; i < nch; ++i) M.acf[i + static_cast<size_t>(nch) * i] += kcore;
}

int get_matrices(bsrg_ctx* c, Matrices& M, double* hcc, double* hcb, double* hbb, double* acf, double* htarg,
                 double* otarg) {
  if (hcc && M.hcc_n()) HIPCHK(hipMemcpyAsync(hcc, M.hcc.p, M.hcc_n() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (hcb && M.hcb_n()) HIPCHK(hipMemcpyAsync(hcb, M.hcb.p, M.hcb_n() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (hbb && !M.hbb.empty()) std::memcpy(hbb, M.hbb.data(), M.hbb.size() * sizeof(double));
  if (acf) std::memcpy(acf, M.acf.data(), M.acf.size() * sizeof(double));
  if (htarg) std::memcpy(htarg, M.htarg.data(), M.htarg.size() * sizeof(double));
  if (otarg) std::memcpy(otarg, M.otarg.data(), M.otarg.size() * sizeof(double));
  return 0;
}

}  // namespace

// ---- R^k band of one multipole into caller memory, allocating nothing ----
// The work arrays of the band build (moments, diagonal cells, their pair
// sums and cell-major copies) live in one file-scope device array; each
// device holds its own copy, whose address is looked up once per device
// (bsrg_create) and kept in g_scratch_dev.  Every element of the work
// arrays is written before it is read, so nothing needs clearing.
constexpr size_t kScratchDoubles = size_t(2) << 20;  // 16 MB
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices] = {};

// With the device current: the address of its copy of scratch.
static int resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(-5, "device index beyond the scratch table");
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (g_scratch_dev[device]) return 0;
  void* ptr = nullptr;
  HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fail(-100, "HIP error: no device address for the scratch array");
  g_scratch_dev[device] = static_cast<double*>(ptr);
  return 0;
}

static size_t band_work_doubles(int ks, int ns) {
  const size_t nv = static_cast<size_t>(ns - ks + 1), ik = static_cast<size_t>(ks) * (ks + 1) / 2;
  return 4 * nv * ik + 2 * nv * ik * ik + 2 * static_cast<size_t>(ks) * ns;
}

extern "C" {

const char* bsrg_last_error(void) { return g_err.c_str(); }
int bsrg_abi_version(void) { return 1; }

int bsrg_mkgrid(int kind, int ks, double z, double h, double hmax, double rmax, int cap, double* t, int* ns, int* ml,
                int* me, double* h_out) {
  std::vector<double> tv;
  int n, l, e;
  double hh;
  int rc = mkgrid(kind, ks, z, h, hmax, rmax, tv, n, l, e, hh);
  if (rc) return rc;
  if (static_cast<int>(tv.size()) > cap) return fail(-1, "bsrg_mkgrid: knot buffer too small");
  std::memcpy(t, tv.data(), tv.size() * sizeof(double));
  *ns = n; *ml = l; *me = e; *h_out = hh;
  return 0;
}

int bsrg_grid_tables(int ks, int ns, const double* t, double* gr, double* grw, double* bsp, double* sb) {
  if (ks < 2 || ks > kMaxKs || ns < ks) return fail(-1, "bsrg_grid_tables: bad ks/ns");
  GridTables G;
  G.build(ks, ns, 0, 0, 0.0, t);
  if (gr) std::memcpy(gr, G.gr.data(), G.gr.size() * sizeof(double));
  if (grw) std::memcpy(grw, G.grw.data(), G.grw.size() * sizeof(double));
  if (bsp) std::memcpy(bsp, G.bsp.data(), G.bsp.size() * sizeof(double));
  if (sb) std::memcpy(sb, G.sb.data(), G.sb.size() * sizeof(double));
  return 0;
}

int bsrg_create(bsrg_ctx** out, int device) {
  if (!out) return fail(-2, "null output pointer");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(-5, "no such HIP device");
  HIPCHK(hipSetDevice(device));
  if (int rc = resolve_scratch(device)) return rc;
  std::unique_ptr<bsrg_ctx> c(new bsrg_ctx());
  c->device = device;
  HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  HIPCHK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
  if (rocblas_create_handle(&c->blas) != rocblas_status_success) return fail(-101, "rocblas_create_handle failed");
  rocblas_set_stream(c->blas, c->stream);
  // no atomics in rocBLAS's kernels: the same products on every run
  rocblas_set_atomics_mode(c->blas, rocblas_atomics_not_allowed);
  for (auto& e : c->ev) HIPCHK(hipEventCreate(&e));
  for (Staged* st : {&c->stS, &c->stH, &c->stC, &c->dw.st}) HIPCHK(hipEventCreateWithFlags(&st->done, hipEventDisableTiming));
  *out = c.release();
  return 0;
}

int bsrg_destroy(bsrg_ctx* c) {
  if (!c) return 0;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  if (c->blas) rocblas_destroy_handle(c->blas);
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  for (Staged* st : {&c->stS, &c->stH, &c->stC, &c->dw.st}) if (st->done) (void)hipEventDestroy(st->done);
  if (c->rl.hitems_done) (void)hipEventDestroy(c->rl.hitems_done);
  if (c->rl.dec_done) (void)hipEventDestroy(c->rl.dec_done);
  hipStream_t s = c->stream, cs = c->copy_stream;
  delete c;
  if (s) (void)hipStreamDestroy(s);
  if (cs) (void)hipStreamDestroy(cs);
  return 0;
}

int bsrg_set_grid(bsrg_ctx* c, int ks, int ns, int ml, int me, double h, const double* t) {
  if (int rc = check_ctx(c)) return rc;
  if (ks < 2 || ks > kMaxKs || ns < ks || !t) return fail(-1, "bsrg_set_grid: bad ks/ns");
  for (int i = 1; i < ns + ks; ++i)
    if (t[i] < t[i - 1]) return fail(-1, "bsrg_set_grid: knots must be non-decreasing");
  if (c->grid_ok && c->G.ks == ks && c->G.ns == ns && c->G.ml == ml && c->G.me == me && c->G.h == h &&
      std::equal(c->G.t.begin(), c->G.t.end(), t))
    return 0;  // same grid: tables and R^k bands stay resident
  HIPCHK(hipSetDevice(c->device));
  c->G.build(ks, ns, ml, me, h, t);
  hipStream_t s = c->stream;
  HIPCHK(c->t.upload(c->G.t.data(), c->G.t.size(), s));
  HIPCHK(c->gx.upload(c->G.gx.data(), c->G.gx.size(), s));
  HIPCHK(c->gw.upload(c->G.gw.data(), c->G.gw.size(), s));
  HIPCHK(c->gr.upload(c->G.gr.data(), c->G.gr.size(), s));
  HIPCHK(c->grm.upload(c->G.grm.data(), c->G.grm.size(), s));
  HIPCHK(c->grw.upload(c->G.grw.data(), c->G.grw.size(), s));
  HIPCHK(c->bsp.upload(c->G.bsp.data(), c->G.bsp.size(), s));
  HIPCHK(c->sb.upload(c->G.sb.data(), c->G.sb.size(), s));
  HIPCHK(hipStreamSynchronize(s));
  c->grid_ok = true;
  c->band_nk = 0;
  c->basis_nk = 0;
  c->exch_nk = 0;
  c->orb_ok = false;
  c->lval_ok = false;
  c->wave_ok = false;
  return 0;
}

int bsrg_mrk_cell(bsrg_ctx* c, int k, double* rkb) {
  if (int rc = check_ctx(c)) return rc;
  if (k < 0) return fail(-1, "bsrg_mrk_cell: k < 0");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = ensure_bands(c, k + 1)) return rc;
  const size_t n = static_cast<size_t>(c->G.ns) * c->G.ns * c->G.ks * c->G.ks;
  if (rkb) HIPCHK(hipMemcpyAsync(rkb, c->rkb.p + k * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_mrk_cell_noalloc(bsrg_ctx* c, int k, double* rkb_dev, void* stream) {
  if (int rc = check_ctx(c)) return rc;
  if (k < 0) return fail(-1, "bsrg_mrk_cell_noalloc: k < 0");
  if (!rkb_dev) return fail(-1, "bsrg_mrk_cell_noalloc: null output");
  if (!c->grid_ok) return fail(-3, "grid not set (bsrg_set_grid)");
  const int ks = c->G.ks, ns = c->G.ns, nv = c->G.nv, ik = ks * (ks + 1) / 2;
  if (band_work_doubles(ks, ns) > kScratchDoubles)
    return fail(-6, "bsrg_mrk_cell_noalloc: the grid needs more work space than the 16 MB scratch array");
  HIPCHK(hipSetDevice(c->device));
  double* w = g_scratch_dev[c->device];
  if (!w) return fail(-100, "HIP error: scratch array not resolved for this device");
  double* mom = w;                                         // 2 nv ik
  double* rkd = mom + static_cast<size_t>(2) * nv * ik;    // nv ik^2
  double* msum = rkd + static_cast<size_t>(nv) * ik * ik;  // 2 ks ns
  double* momT = msum + static_cast<size_t>(2) * ks * ns;  // 2 nv ik
  double* rdT = momT + static_cast<size_t>(2) * nv * ik;   // nv ik^2
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c->stream;
  DevGrid g = devgrid(c);
  launch_moments(g, k, 1, mom, rkd, s);
  launch_band(g, 1, mom, rkd, msum, momT, rdT, rkb_dev, s);
  HIPCHK(hipGetLastError());
  if (!stream) HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int bsrg_rk_moments(bsrg_ctx* c, int k, double* rkd1, double* rkd2, double* rkd) {
  if (int rc = check_ctx(c)) return rc;
  if (k < 0) return fail(-1, "bsrg_rk_moments: k < 0");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = ensure_bands(c, k + 1)) return rc;
  const int ks = c->G.ks, nv = c->G.nv, ik = ks * (ks + 1) / 2;
  const double* m1 = c->mom.p + static_cast<size_t>(k * 2) * nv * ik;
  if (rkd1) HIPCHK(hipMemcpyAsync(rkd1, m1, sizeof(double) * nv * ik, hipMemcpyDeviceToHost, c->stream));
  if (rkd2) HIPCHK(hipMemcpyAsync(rkd2, m1 + static_cast<size_t>(nv) * ik, sizeof(double) * nv * ik, hipMemcpyDeviceToHost, c->stream));
  if (rkd) HIPCHK(hipMemcpyAsync(rkd, c->rkd.p + static_cast<size_t>(k) * nv * ik * ik, sizeof(double) * nv * ik * ik,
                                 hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_convol(bsrg_ctx* c, int k, int icase, const double* d, double* a) {
  if (int rc = check_ctx(c)) return rc;
  if (icase < 1 || icase > 4) return fail(-2, " CONVOL:  unknown case ");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = ensure_bands(c, k + 1)) return rc;
  const int ns = c->G.ns, ks = c->G.ks;
  const size_t nin = icase <= 2 ? static_cast<size_t>(ns) * ks : static_cast<size_t>(ns) * ns;
  HIPCHK(c->dtmp.upload(d, nin, c->stream));
  HIPCHK(c->dtmp2.resize(nin));
  const double* R = c->rkb.p + static_cast<size_t>(k) * ns * ns * ks * ks;
  if (icase <= 2) launch_convol_direct(ns, ks, icase, R, c->dtmp.p, c->dtmp2.p, c->stream);
  else launch_convol_exchange(ns, ks, R, c->dtmp.p, c->dtmp2.p, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(a, c->dtmp2.p, nin * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_set_orbitals(bsrg_ctx* c, int nbf, const double* pbs, const int* iech, const int* lbs, const int* ibort,
                      double eps_ovl) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->grid_ok) return fail(-3, "grid not set (bsrg_set_grid)");
  if (nbf < 1 || !pbs || !iech || !lbs || !ibort) return fail(-1, "bsrg_set_orbitals: bad arguments");
  if (c->orb_ok && c->nbf == nbf && c->eps_ovl == eps_ovl &&
      std::equal(c->pbs.begin(), c->pbs.end(), pbs) && std::equal(c->iech.begin(), c->iech.end(), iech) &&
      std::equal(c->lbs.begin(), c->lbs.end(), lbs) && std::equal(c->ibort.begin(), c->ibort.end(), ibort))
    return 0;  // same orbitals: the bases stay resident
  HIPCHK(hipSetDevice(c->device));
  const int ns = c->G.ns, ks = c->G.ks;
  // the bases depend on the grid and the bound orbitals only: they stay
  // resident when a wave brings new channel orbitals for the same target
  std::vector<double> Pnew;
  for (int i = 0; i < nbf; ++i)
    if (iech[i] == 0) Pnew.insert(Pnew.end(), pbs + static_cast<size_t>(ns) * i, pbs + static_cast<size_t>(ns) * (i + 1));
  const bool same_bound = c->orb_ok && Pnew == c->Pbound;
  c->nbf = nbf;
  c->pbs.assign(pbs, pbs + static_cast<size_t>(ns) * nbf);
  c->iech.assign(iech, iech + nbf);
  c->lbs.assign(lbs, lbs + nbf);
  c->ibort.assign(ibort, ibort + static_cast<size_t>(nbf) * nbf);
  c->eps_ovl = eps_ovl;
  c->bound.clear();
  c->bidx.assign(nbf, -1);
  for (int i = 0; i < nbf; ++i)
    if (iech[i] == 0) { c->bidx[i] = static_cast<int>(c->bound.size()); c->bound.push_back(i + 1); }
  c->nb = static_cast<int>(c->bound.size());
  c->npair = c->nb * (c->nb + 1) / 2;
  // qbs = sb * p, OBS for bound orbitals of equal l (read_data.f90:142-208)
  c->qbs.assign(static_cast<size_t>(ns) * nbf, 0.0);
  c->obs.assign(static_cast<size_t>(nbf) * nbf, 0.0);
  for (int i = 0; i < nbf; ++i)
    if (iech[i] == 0) bxv_sym(ks, ns, c->G.sb.data(), &c->pbs[static_cast<size_t>(ns) * i], &c->qbs[static_cast<size_t>(ns) * i]);
  for (int i = 0; i < nbf; ++i) {
    if (iech[i] != 0) continue;
    for (int j = 0; j <= i; ++j) {
      if (iech[j] != 0 || lbs[i] != lbs[j]) continue;
      double s = bvmv_sym(ns, ks, c->G.sb.data(), &c->pbs[static_cast<size_t>(ns) * i], &c->pbs[static_cast<size_t>(ns) * j]);
      if (std::fabs(s) < eps_ovl) s = 0.0;
      if (i == j) s = 1.0;
      c->obs[i + static_cast<size_t>(nbf) * j] = s;
      c->obs[j + static_cast<size_t>(nbf) * i] = s;
    }
  }
  if (!same_bound) {
    std::vector<double> P(static_cast<size_t>(ns) * std::max(c->nb, 1), 0.0);
    if (!Pnew.empty()) std::memcpy(P.data(), Pnew.data(), Pnew.size() * sizeof(double));
    HIPCHK(c->P.upload(P.data(), P.size(), c->stream));
    HIPCHK(c->dens.resize(std::max<size_t>(static_cast<size_t>(c->npair) * ns * ks, 1)));
    if (c->npair) launch_pair_dens(ns, ks, c->nb, c->P.p, c->dens.p, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->Pbound.swap(Pnew);
    c->basis_nk = 0;
    c->exch_nk = 0;
  }
  c->Q.clear();
  c->Qok.clear();
  c->wtab.assign(static_cast<size_t>(nbf) * nbf, -1);
  c->wvec.clear();
  c->wvec_uploaded = -1;
  c->lval_ok = false;
  c->orb_ok = true;
  ++c->orb_ver;
  return 0;
}

int bsrg_alloc_matrix(bsrg_ctx* c, int nch, int ncp, int mk) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->grid_ok) return fail(-3, "grid not set (bsrg_set_grid)");
  if (nch < 1 || ncp < 0 || mk < 0) return fail(-1, "bsrg_alloc_matrix: bad sizes");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(c->cur.alloc(c->G.ns, nch, ncp, mk));
  launch_fill(c->cur.hcc_n(), 0.0, c->cur.hcc.p, c->stream);
  launch_fill(c->cur.hcb_n(), 0.0, c->cur.hcb.p, c->stream);
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

static int cur_ready(bsrg_ctx* c) {
  if (c->cur.nch == 0) return fail(-6, "matrix not allocated (bsrg_alloc_matrix)");
  return 0;
}

int bsrg_update_hl(bsrg_ctx* c, int ich, int jch, double v) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = cur_ready(c)) return rc;
  if (ich < 1 || ich > c->cur.nch || jch < 1 || jch > c->cur.nch) return fail(-11, "UPDATE_HL: channel index out of range");
  HIPCHK(hipSetDevice(c->device));
  if (c->basis_nk == 0) {
    // the band table must at least hold sb
    HIPCHK(c->bands.resize(static_cast<size_t>(c->G.ns) * c->G.ks));
    launch_sb_upper(c->G.ns, c->G.ks, c->sb.p, c->bands.p, c->stream);
  }
  Terms& T = c->TC;
  if (int rc = wait_staged(c->stC)) return rc;
  T.clear();
  band_hl(T, ich, jch, v);
  return apply_terms(c, c->cur, T, false);
}

int bsrg_idata(bsrg_ctx* c, int icase, int jtype, int jpol, int n, const int* ipt, const int* k1, const int* k2,
               const int* k3, const int* k4, const double* cd) {
  if (int rc = check_ctx(c)) return rc;
  if (n == 0) return 0;
  if (icase != 5) return fail(-20, "unknown case in I_data");
  if (jtype < 1 || jtype > 9) return fail(-21, " I_data: unknown itype ");
  if (int rc = cur_ready(c)) return rc;
  if (jpol < 0) return fail(-17, "UPDATE_CF: multipole index out of range");
  HIPCHK(hipSetDevice(c->device));
  const int need = std::max(jpol + 1, c->basis_nk);
  if (int rc = ensure_bases(c, need, jtype >= 8)) return rc;
  Terms& T = c->TC;
  if (int rc = wait_staged(c->stC)) return rc;
  T.clear();
  if (int rc = add_idata(c, c->cur, T, jtype, jpol, n, ipt, k1, k2, k3, k4, cd, 1.0)) return rc;
  return apply_terms(c, c->cur, T, false);
}

int bsrg_odata(bsrg_ctx* c, int n, const int* ipt, const int* k3, const int* k4, const double* cd, double eps_c) {
  if (int rc = check_ctx(c)) return rc;
  if (n == 0) return 0;
  if (int rc = cur_ready(c)) return rc;
  if (!c->orb_ok) return fail(-4, "orbitals not set (bsrg_set_orbitals)");
  HIPCHK(hipSetDevice(c->device));
  if (c->basis_nk == 0)
    if (int rc = ensure_bases(c, 1, false)) return rc;
  Terms& T = c->TC;
  if (int rc = wait_staged(c->stC)) return rc;
  T.clear();
  if (int rc = add_odata(c, c->cur, T, n, ipt, k3, k4, cd, eps_c, 1.0, true)) return rc;
  return apply_terms(c, c->cur, T, false);
}

namespace {

// ZCB(k1,k2,k3) = (3j symbol (k1 k2 k3; 0 0 0))^2: with 2g = k1+k2+k3 even,
// (2g-2k1)!(2g-2k2)!(2g-2k3)!/(2g+1)! * [g!/((g-k1)!(g-k2)!(g-k3)!)]^2;
// zero for odd 2g or a broken triangle.  The factorial ratios are built as
// running products so nothing overflows for the multipoles used here.
double zcb(int k1, int k2, int k3) {
  const int J = k1 + k2 + k3;
  if (J % 2) return 0.0;
  const int g = J / 2;
  if (g < k1 || g < k2 || g < k3) return 0.0;
  auto fact = [](int n) { long double f = 1.0L; for (int i = 2; i <= n; ++i) f *= i; return f; };
  const long double a = fact(J - 2 * k1) * fact(J - 2 * k2) * fact(J - 2 * k3) / fact(J + 1);
  const long double b = fact(g) / (fact(g - k1) * fact(g - k2) * fact(g - k3));
  return static_cast<double>(a * b * b);
}

}  // namespace

int bsrg_gen_lval(bsrg_ctx* c, double z, int kclosd) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->grid_ok) return fail(-3, "grid not set (bsrg_set_grid)");
  if (!c->orb_ok) return fail(-4, "orbitals not set (bsrg_set_orbitals)");
  const int nbf = c->nbf, ns = c->G.ns, ks = c->G.ks, nb = c->nb;
  if (kclosd < 0 || kclosd > nbf) return fail(-60, "Gen_Lval: kclosd outside 0..nbf");
  for (int ip = 0; ip < kclosd; ++ip)
    if (c->iech[ip] != 0) return fail(-61, "Gen_Lval: a core orbital is a channel orbital");
  HIPCHK(hipSetDevice(c->device));
  c->lval_ok = false;
  int mlh = 0, lcore = 0;
  for (int i = 0; i < nbf; ++i) {
    if (c->lbs[i] < 0 || c->lbs[i] > 100) return fail(-62, "Gen_Lval: orbital l outside 0..100");
    mlh = std::max(mlh, c->lbs[i]);
  }
  for (int ip = 0; ip < kclosd; ++ip) lcore = std::max(lcore, c->lbs[ip]);
  // bases: the direct potential of the core needs k = 0, its exchange k up
  // to min(l + l_core, 7)
  if (kclosd > 0) {
    const int nk = std::max(c->basis_nk, std::min(mlh + lcore, 7) + 1);
    if (int rc = ensure_bases(c, nk, true)) return rc;
  }
  const size_t nss = static_cast<size_t>(ns) * ns, nsk = static_cast<size_t>(ns) * ks;
  const std::vector<double> db2 = c->G.db2(), rm1 = c->G.mrm(-1), rm2 = c->G.mrm(-2);
  // per l: the L band (hl, lower storage), the exchange list (E column,
  // factor), k outer and core orbital inner; one direct list for all l
  std::vector<double> hl(static_cast<size_t>(mlh + 1) * nsk, 0.0), ec, dc;
  std::vector<int> eid, eptr{0}, did, lorb, lptr{0}, present(mlh + 1, 0);
  for (int i = 0; i < nbf; ++i) present[c->lbs[i]] = 1;
  for (int ip = 0; ip < kclosd; ++ip) {
    int p;
    if (int rc = pair_index(c, ip + 1, ip + 1, p)) return rc;
    did.push_back(1 + p);  // D(k = 0, pair (ip, ip)) in the band table
    dc.push_back(-2.0 * (4 * c->lbs[ip] + 2));
  }
  for (int l = 0; l <= mlh; ++l) {
    if (present[l]) {
      double* h = &hl[l * nsk];
      const double c1 = static_cast<double>(l) * (l + 1), c2 = 2.0 * z;
      for (size_t q = 0; q < nsk; ++q) h[q] = db2[q] - c1 * rm2[q] + c2 * rm1[q];
      const double cb = db2[0] - db2[(ns - 1) + static_cast<size_t>(ns) * (ks - 2)];  // Bloch term
      h[(ns - 1) + static_cast<size_t>(ns) * (ks - 2)] += cb;
      h[(ns - 1) + static_cast<size_t>(ns) * (ks - 1)] -= cb;
      if (kclosd > 0) {
        int kmin = 1000, kmax = 0;
        for (int ip = 0; ip < kclosd; ++ip) {
          kmin = std::min(kmin, std::abs(l - c->lbs[ip]));
          kmax = std::max(kmax, l + c->lbs[ip]);
        }
        kmax = std::min(kmax, 7);
        for (int k = kmin; k <= kmax; ++k)
          for (int ip = 0; ip < kclosd; ++ip) {
            const double f = (4 * c->lbs[ip] + 2) * zcb(l, k, c->lbs[ip]);
            if (f == 0.0) continue;
            const int a = c->bidx[ip];
            eid.push_back((k * nb + a) * nb + a);
            ec.push_back(f);
          }
      }
      for (int i = kclosd; i < nbf; ++i)
        if (c->lbs[i] == l && c->iech[i] == 0) lorb.push_back(i);
    }
    eptr.push_back(static_cast<int>(eid.size()));
    lptr.push_back(static_cast<int>(lorb.size()));
  }
  const int nlorb = static_cast<int>(lorb.size());
  std::vector<int> pcol(std::max(nlorb, 1), 0);
  for (int o = 0; o < nlorb; ++o) pcol[o] = c->bidx[lorb[o]];
  // one upload of the doubles (bands, factors) and one of the ints
  const size_t o_hl = 0, o_ec = hl.size(), o_dc = o_ec + ec.size(), nd_ = o_dc + dc.size();
  std::vector<double> dbl(nd_ + 1, 0.0);
  std::copy(hl.begin(), hl.end(), dbl.begin() + o_hl);
  std::copy(ec.begin(), ec.end(), dbl.begin() + o_ec);
  std::copy(dc.begin(), dc.end(), dbl.begin() + o_dc);
  const size_t o_eid = 0, o_did = eid.size(), o_pc = o_did + did.size(), ni = o_pc + pcol.size();
  std::vector<int> ints(ni + 1, 0);
  std::copy(eid.begin(), eid.end(), ints.begin() + o_eid);
  std::copy(did.begin(), did.end(), ints.begin() + o_did);
  std::copy(pcol.begin(), pcol.end(), ints.begin() + o_pc);
  hipStream_t s = c->stream;
  HIPCHK(c->lwork.resize(dbl.size() + static_cast<size_t>(std::max(nlorb, 1)) * ns));
  HIPCHK(c->lwork_i.resize(ints.size()));
  HIPCHK(hipMemcpyAsync(c->lwork.p, dbl.data(), dbl.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->lwork_i.p, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HIPCHK(c->hlfull.resize(static_cast<size_t>(mlh + 1) * nss));
  launch_fill(static_cast<size_t>(mlh + 1) * nss, 0.0, c->hlfull.p, s);
  double* lv_dev = c->lwork.p + dbl.size();
  for (int l = 0; l <= mlh; ++l) {
    if (!present[l]) continue;
    double* out = c->hlfull.p + l * nss;
    launch_lcore(ns, ks, eptr[l + 1] - eptr[l], c->lwork_i.p + o_eid + eptr[l], c->lwork.p + o_ec + eptr[l], c->E.p,
                 kclosd, c->lwork_i.p + o_did, c->lwork.p + o_dc, c->bands.p, c->lwork.p + o_hl + l * nsk, out, s);
    launch_lvec(ns, lptr[l + 1] - lptr[l], c->lwork_i.p + o_pc + lptr[l], c->P.p, out, lv_dev + lptr[l] * ns, s);
  }
  HIPCHK(hipGetLastError());
  std::vector<double> hf(static_cast<size_t>(mlh + 1) * nss), lvo(static_cast<size_t>(std::max(nlorb, 1)) * ns);
  HIPCHK(hipMemcpyAsync(hf.data(), c->hlfull.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(lvo.data(), lv_dev, lvo.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  // L_int(i, j) = <p_i|hl_core|p_j> from the lower triangle, as Gen_Lval
  // sums it (l_core.f90:122-130); L_vec(:, i) joins the vector table
  c->lint.assign(static_cast<size_t>(nbf) * nbf, 0.0);
  c->lvec.assign(static_cast<size_t>(ns) * nbf, 0.0);
  c->lvec_id.assign(nbf, -1);
  for (int o = 0; o < nlorb; ++o) {
    const int i = lorb[o], l = c->lbs[i];
    const double* h = &hf[l * nss];
    const double* pi = &c->pbs[static_cast<size_t>(ns) * i];
    for (int q = o; q < lptr[l + 1]; ++q) {
      const int j = lorb[q];
      const double* pj = &c->pbs[static_cast<size_t>(ns) * j];
      double v = 0.0;
      for (int m = 0; m < ns; ++m)
        for (int k = 0; k < m; ++k) v += h[m + static_cast<size_t>(ns) * k] * (pi[m] * pj[k] + pi[k] * pj[m]);
      for (int m = 0; m < ns; ++m) v += h[m + static_cast<size_t>(ns) * m] * pi[m] * pj[m];
      c->lint[i + static_cast<size_t>(nbf) * j] = v;
      c->lint[j + static_cast<size_t>(nbf) * i] = v;
    }
    std::copy(&lvo[static_cast<size_t>(o) * ns], &lvo[static_cast<size_t>(o + 1) * ns],
              &c->lvec[static_cast<size_t>(ns) * i]);
    c->lvec_id[i] = static_cast<int>(c->wvec.size() / ns);
    c->wvec.insert(c->wvec.end(), &lvo[static_cast<size_t>(o) * ns], &lvo[static_cast<size_t>(o + 1) * ns]);
  }
  c->mlh = mlh;
  c->lval_ok = true;
  return 0;
}

int bsrg_lval_get(bsrg_ctx* c, int* mlh, double* hl_full, double* l_int, double* l_vec) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->lval_ok) return fail(-62, "L_data: Gen_Lval has not run (bsrg_gen_lval)");
  HIPCHK(hipSetDevice(c->device));
  if (mlh) *mlh = c->mlh;
  const size_t nss = static_cast<size_t>(c->G.ns) * c->G.ns;
  if (hl_full) {
    HIPCHK(hipMemcpyAsync(hl_full, c->hlfull.p, (c->mlh + 1) * nss * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  if (l_int) std::copy(c->lint.begin(), c->lint.end(), l_int);
  if (l_vec) std::copy(c->lvec.begin(), c->lvec.end(), l_vec);
  return 0;
}

int bsrg_ldata(bsrg_ctx* c, int jtype, int l, int n, const int* ipt, const int* k1, const int* k2, const int* k3,
               const int* k4, const double* cd, int iitar, double eps_c) {
  if (int rc = check_ctx(c)) return rc;
  if (n == 0) return 0;
  if (!c->lval_ok) return fail(-62, "L_data: Gen_Lval has not run (bsrg_gen_lval)");
  if (jtype < 1 || jtype > 9) return fail(-64, " L_data: uknown itype ");
  if (l < 0 || l > c->mlh) return fail(-63, "L_data: l outside 0..max l of the orbitals");
  if (n < 0 || !k1 || !k2 || !k3 || !k4 || !cd) return fail(-1, "bsrg_ldata: missing list arrays");
  if (int rc = cur_ready(c)) return rc;
  HIPCHK(hipSetDevice(c->device));
  if (c->basis_nk == 0)
    if (int rc = ensure_bases(c, 1, false)) return rc;
  Matrices& M = c->cur;
  const int nbf = c->nbf, nV = c->basis_nk * c->npair * c->nb;
  Terms& T = c->TC;
  if (int rc = wait_staged(c->stC)) return rc;
  T.clear();
  struct Dense { int blk, tr; double c; };
  std::vector<Dense> dense;
  for (int q = 0; q < n; ++q) {
    const int i = ipt ? ipt[q] - 1 : q;
    if (i < 0 || i >= n) return fail(-1, "bsrg_ldata: ipt out of range");
    const double cc = cd[i];
    if (std::fabs(cc) < eps_c) continue;
    if (jtype == 1) {
      const int i1 = k1[i] / kIbi, i2 = k1[i] % kIbi;
      if (i1 < 1 || i1 > nbf || i2 < 1 || i2 > nbf) return fail(-30, "orbital index out of range");
      const double C = c->lint[(i1 - 1) + static_cast<size_t>(nbf) * (i2 - 1)] * cc;
      const int ic = -k3[i], jc = -k4[i];
      if (ic < 0 && jc == 0) {  // L(..) <i|j>
        const int io = -ic, a1 = io / kIbo, a2 = io % kIbo;
        if (a1 < 1 || a1 > nbf || a2 < 1 || a2 > nbf) return fail(-30, "orbital index out of range");
        const int ich = c->iech[a1 - 1], jch = c->iech[a2 - 1];
        if (int rc = chk_ch(M, ich, jch, "UPDATE_CF")) return rc;
        if (iitar > 0 && ich != jch) band_hl(T, ich, jch, C);
        T.scal.push_back({1, blk_index(ich, jch), -1, C});
      } else if ((ic > 0 && jc > 0) || (ic < 0 && jc != 0)) {
        if (int rc = add_struct1(c, M, T, k3[i], k4[i], -1, C, false, true)) return rc;
      } else {
        return fail(-65, " L_data: unknown structure ");
      }
    } else if (jtype <= 5) {
      const int o = k2[i], ich = k3[i], io = k4[i];
      if (o < 1 || o > nbf) return fail(-30, "orbital index out of range");
      const int v = c->lvec_id[o - 1];  // -1: L_vec(:, o) is zero
      if (io > 0) {
        const int j1 = io / kIbo, j2 = io % kIbo;
        int w;
        if (int rc = get_v(c, j1, j2, w)) return rc;
        const int jch = c->iech[j1 - 1];
        if (int rc = chk_ch(M, ich, jch, "UPDATE_HW")) return rc;
        if (v >= 0) {
          if (ich >= jch) T.r1.push_back({blk_index(ich, jch), nV + v, nV + w, -1, cc});
          else T.r1.push_back({blk_index(ich, jch), nV + w, nV + v, -1, cc});
        }
      } else if (io < 0) {
        if (ich < 1 || ich > M.nch) return fail(-13, "UPDATE_HV: channel index out of range");
        if (-io > M.ncp) return fail(-14, "UPDATE_HV: bound index out of range");
        if (v >= 0) T.vec.push_back({(-io - 1) * M.nch + (ich - 1), nV + v, -1, cc});
      } else {
        return fail(-66, " L_data: uknown structure for itype 2-5 ");
      }
    } else {
      const int ich = k1[i] / kIbi, jch = k1[i] % kIbi;
      if (int rc = chk_ch(M, ich, jch, "UPDATE_HX")) return rc;
      if (ich == jch || iitar > 1) dense.push_back({blk_index(ich, jch), ich < jch ? 1 : 0, cc});
      else T.scal.push_back({2, blk_index(ich, jch), -1, cc});
    }
  }
  if (int rc = apply_terms(c, M, T, false)) return rc;
  if (dense.empty()) return 0;
  // c * hl_full(l) (or its transpose) per block, in list order within a block
  std::stable_sort(dense.begin(), dense.end(), [](const Dense& a, const Dense& b) { return a.blk < b.blk; });
  std::vector<int> ints, ptr{0}, blk;
  std::vector<double> cv;
  for (size_t t = 0; t < dense.size(); ++t) {
    if (t > 0 && dense[t].blk != dense[t - 1].blk) ptr.push_back(static_cast<int>(t));
    if (t == 0 || dense[t].blk != dense[t - 1].blk) blk.push_back(dense[t].blk);
    cv.push_back(dense[t].c);
  }
  ptr.push_back(static_cast<int>(dense.size()));
  const int nl = static_cast<int>(blk.size()), nt = static_cast<int>(dense.size());
  ints.insert(ints.end(), ptr.begin(), ptr.end());
  ints.insert(ints.end(), blk.begin(), blk.end());
  ints.insert(ints.end(), dense.size(), l);
  for (const Dense& d : dense) ints.push_back(d.tr);
  hipStream_t s = c->stream;
  HIPCHK(c->lwork.resize(cv.size()));
  HIPCHK(c->lwork_i.resize(ints.size()));
  HIPCHK(hipMemcpyAsync(c->lwork.p, cv.data(), cv.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->lwork_i.p, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice, s));
  const int* dp = c->lwork_i.p;
  launch_dense_terms(c->G.ns, nl, dp, dp + nl + 1, dp + 2 * nl + 1, dp + 2 * nl + 1 + nt, c->lwork.p, c->hlfull.p,
                     M.hcc.p, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int bsrg_bs_orth(bsrg_ctx* c, int which) {
  if (int rc = check_ctx(c)) return rc;
  if (which != 0 && which != 1) return fail(-1, "bsrg_bs_orth: which must be 0 (current matrix) or 1 (the wave's H)");
  if (!c->orb_ok) return fail(-4, "orbitals not set (bsrg_set_orbitals)");
  if (which == 0) {
    if (int rc = cur_ready(c)) return rc;
  } else if (!c->wave_ok) {
    return fail(-8, "no assembled wave (bsrg_wave_assemble)");
  }
  Matrices& M = which == 0 ? c->cur : c->H;
  const int ns = c->G.ns, ks = c->G.ks, nbf = c->nbf, nch = M.nch;
  HIPCHK(hipSetDevice(c->device));
  // per channel: the bound orbitals of the channel orbital's l that IBORT
  // does not exempt (bs_orth.f90:40-55)
  std::vector<int> zoff(nch, 0), nz(nch, 0), cols;
  int nzmax = 0;
  for (int ich = 1; ich <= nch; ++ich) {
    int i = 0;
    for (int o = 1; o <= nbf && !i; ++o)
      if (c->iech[o - 1] == ich) i = o;
    if (!i) return fail(-70, "BS_ORTH: channel without a channel orbital");
    zoff[ich - 1] = static_cast<int>(cols.size());
    for (int j = 1; j <= nbf; ++j) {
      if (c->iech[j - 1] > 0 || c->lbs[i - 1] != c->lbs[j - 1]) continue;
      if (c->ibort[(std::max(i, j) - 1) + static_cast<size_t>(nbf) * (std::min(i, j) - 1)] != 0) continue;
      cols.push_back(j);
    }
    nz[ich - 1] = static_cast<int>(cols.size()) - zoff[ich - 1];
    nzmax = std::max(nzmax, nz[ich - 1]);
  }
  if (nzmax == 0) return 0;
  const size_t nc = cols.size();
  std::vector<double> zb(2 * nc * ns);  // Z then B Z, one column per condition
  for (size_t q = 0; q < nc; ++q) {
    const double* p = &c->pbs[static_cast<size_t>(ns) * (cols[q] - 1)];
    std::copy(p, p + ns, &zb[q * ns]);
    bxv_sym(ks, ns, c->G.sb.data(), p, &zb[(nc + q) * ns]);
  }
  std::vector<int> ints(zoff);
  ints.insert(ints.end(), nz.begin(), nz.end());
  hipStream_t s = c->stream;
  HIPCHK(c->lwork.resize(zb.size()));
  HIPCHK(c->lwork_i.resize(ints.size()));
  HIPCHK(hipMemcpyAsync(c->lwork.p, zb.data(), zb.size() * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->lwork_i.p, ints.data(), ints.size() * sizeof(int), hipMemcpyHostToDevice, s));
  const size_t nscr = bs_orth_scratch_doubles(ns, static_cast<int>(M.nblk()), nzmax);
  if (nscr) HIPCHK(c->dtmp2.resize(nscr));
  const int lrc = launch_bs_orth(ns, static_cast<int>(M.nblk()), nch, M.ncp, nzmax, c->lwork_i.p, c->lwork_i.p + nch,
                                 c->lwork.p, c->lwork.p + nc * ns, M.hcc.p, M.hcb.p, nscr ? c->dtmp2.p : nullptr,
                                 nscr, s);
  if (lrc) return fail(-71, "BS_ORTH: no work space for the projectors");
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s));
  return 0;
}

int bsrg_symmetrize_diag(bsrg_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = cur_ready(c)) return rc;
  HIPCHK(hipSetDevice(c->device));
  launch_symmetrize_diag(c->G.ns, c->cur.nch, c->cur.hcc.p, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_scale(bsrg_ctx* c, double ec) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = cur_ready(c)) return rc;
  HIPCHK(hipSetDevice(c->device));
  launch_scale(c->cur.hcc_n(), ec, c->cur.hcc.p, c->stream);
  launch_scale(c->cur.hcb_n(), ec, c->cur.hcb.p, c->stream);
  for (auto& v : c->cur.hbb) v *= ec;
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_acf_finish(bsrg_ctx* c, int kcore) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = cur_ready(c)) return rc;
  acf_finish(c->cur, kcore);
  return 0;
}

int bsrg_get_matrix(bsrg_ctx* c, double* hcc, double* hcb, double* hbb, double* acf, double* htarg, double* otarg) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = cur_ready(c)) return rc;
  HIPCHK(hipSetDevice(c->device));
  return get_matrices(c, c->cur, hcc, hcb, hbb, acf, htarg, otarg);
}

namespace {

// bsrg_wave_prepare; resident: the R^k lists are the context's resident
// ones (bsrg_lists_upload), decoded on the device, and nrec and the list
// arrays are unused.
int wave_prepare(bsrg_ctx* c, bool resident, int nch, int ncp, int mk, int nrec, const int* jtype, const int* kpol,
                 const int* k1, const int* k2, const int* k3, const int* k4, const double* cd, int novl,
                 const int* ok3, const int* ok4, const double* ocd, double eps_c, double ec, int ntarg,
                 const double* etarg, const int* iptar) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->orb_ok) return fail(-4, "orbitals not set (bsrg_set_orbitals)");
  if (resident && c->rl.nrec < 0) return fail(-8, "no R^k lists in device memory (bsrg_lists_upload)");
  if (resident) nrec = 0;
  if (nch < 1 || ncp < 0 || mk < 0 || nrec < 0 || novl < 0) return fail(-1, "bsrg_wave_assemble: bad sizes");
  HIPCHK(hipSetDevice(c->device));
  c->wave_ok = false;
  c->wave_prepared = false;
  c->eager_done = false;
  c->rl.active = false;
  using clk = std::chrono::steady_clock;
  const auto h0 = clk::now();
  // record checks on the host workers (the first bad record in list order
  // is the one reported); the device decoder checks resident records itself
  if (!c->pool) c->pool.reset(new Pool(host_workers()));
  const int nchk = static_cast<int>(std::min<long>(c->pool->size(), nrec / 16384L + 1));
  std::vector<int> bad(nchk, -1), exch(nchk, 0);
  const std::function<void(int)> check = [&](int t) {
    const int r0 = static_cast<int>(static_cast<long>(nrec) * t / nchk);
    const int r1 = static_cast<int>(static_cast<long>(nrec) * (t + 1) / nchk);
    int ex = 0;
    for (int r = r0; r < r1; ++r) {
      if (jtype[r] < 1 || jtype[r] > 9 || kpol[r] < 0 || kpol[r] > mk) { bad[t] = r; return; }
      ex |= jtype[r] >= 8;
    }
    exch[t] = ex;
  };
  if (nrec > 0) c->pool->run(nchk, check);
  bool any_exch = resident && c->rl.any_exch;
  for (int t = 0; t < nchk && nrec > 0; ++t) {
    if (bad[t] >= 0) {
      const int r = bad[t];
      if (jtype[r] < 1 || jtype[r] > 9) return fail(-21, " I_data: unknown itype ");
      return fail(-17, "UPDATE_CF: multipole index out of range");
    }
    any_exch = any_exch || exch[t];
  }
  for (int i = 0; i < nch; ++i)
    if (iptar[i] < 1 || iptar[i] > ntarg) return fail(-7, "iptar out of range");
  (void)hipEventRecord(c->ev[0], c->stream);
  const int nk = mk + 1;
  if (int rc = ensure_bands(c, nk)) return rc;
  (void)hipEventRecord(c->ev[1], c->stream);
  if (int rc = ensure_bases(c, nk, any_exch)) return rc;
  (void)hipEventRecord(c->ev[8], c->stream);
  HIPCHK(c->S.alloc(c->G.ns, nch, ncp, mk));
  HIPCHK(c->H.alloc(c->G.ns, nch, ncp, mk));
  const auto ha = clk::now();
  // list processing on the host overlaps the band and basis kernels
  Terms& TS = c->TS;
  Terms& TH = c->TH;
  Terms& THt = c->THt;
  if (int rc = wait_staged(c->stS)) return rc;
  if (int rc = wait_staged(c->stH)) return rc;
  TS.clear();
  TH.clear();
  THt.clear();
  // S: sb on the channel diagonals + overlaps
  for (int ich = 1; ich <= nch; ++ich) band_hl(TS, ich, ich, 1.0);
  if (int rc = add_odata(c, c->S, TS, novl, nullptr, ok3, ok4, ocd, eps_c, 1.0, true)) return rc;
  // H: ec * S + R^k lists + (E_t - ec) * sb
  for (int ich = 1; ich <= nch; ++ich) band_hl(TH, ich, ich, ec);
  if (int rc = add_odata(c, c->H, TH, novl, nullptr, ok3, ok4, ocd, eps_c, ec, false)) return rc;
  c->partsS.assign(1, &TS);
  c->s_enqueued = false;
  if (c->early_gemm || resident) {
    // a one-shot wave: S is complete, and its term pass runs behind the
    // bases while the R^k lists are decoded
    if (int rc = stage_terms(c->partsS, c->stS, static_cast<int>(c->S.nblk()), c->S.nch * c->S.ncp, nullptr,
                            &c->S.live))
      return rc;
    if (int rc = upload_terms(c, c->S, TS, c->stS)) return rc;
    if (int rc = enqueue_s(c)) return rc;
    c->s_enqueued = true;
  }
  const auto hb = clk::now();
  c->gemm_enqueued = false;
  if (resident) {
    // the lists on the device: H's terms, GET_V vectors and scalars there
    for (int ich = 1; ich <= nch; ++ich) band_hl(THt, ich, ich, etarg[iptar[ich - 1] - 1] - ec);
    if (int rc = dl_decode(c)) return rc;
    if (int rc = autotune_gemm(c, c->H, c->stH, c->capture_safe)) return rc;
    c->partsH.clear();
    c->rl.active = true;
    const auto h3 = clk::now();
    auto ms = [](clk::time_point x, clk::time_point y) { return std::chrono::duration<double, std::milli>(y - x).count(); };
    for (double& x : c->host_ms) x = 0.0;
    c->host_ms[0] = ms(h0, ha);
    c->host_ms[1] = ms(ha, hb);
    c->host_ms[3] = ms(hb, h3);
    c->host_ms[7] = ms(h0, h3);
    c->wave_prepared = true;
    return 0;
  }
  if (int rc = idata_parallel(c, nrec, jtype, kpol, k1, k2, k3, k4, cd, any_exch)) return rc;
  for (int ich = 1; ich <= nch; ++ich) band_hl(THt, ich, ich, etarg[iptar[ich - 1] - 1] - ec);
  c->partsH.assign(1, &TH);
  for (int t = 0; t < c->nwork; ++t) c->partsH.push_back(c->work[t].get());
  c->partsH.push_back(&THt);
  const auto h1 = clk::now();
  if (!c->s_enqueued)
    if (int rc = stage_terms(c->partsS, c->stS, static_cast<int>(c->S.nblk()), c->S.nch * c->S.ncp, nullptr,
                            &c->S.live))
      return rc;
  if (int rc = stage_terms(c->partsH, c->stH, static_cast<int>(c->H.nblk()), c->H.nch * c->H.ncp, c->pool.get(),
                          &c->H.live))
    return rc;
  const auto h2 = clk::now();
  if (!c->s_enqueued)
    if (int rc = upload_terms(c, c->S, TS, c->stS)) return rc;
  if (c->early_gemm) {
    if (int rc = upload_arena(c, c->H, c->stH)) return rc;
  } else {
    if (int rc = upload_terms(c, c->H, TH, c->stH)) return rc;
    if (int rc = autotune_gemm(c, c->H, c->stH, c->capture_safe)) return rc;
  }
  const auto h3 = clk::now();
  auto ms = [](clk::time_point a, clk::time_point b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  c->host_ms[0] = ms(h0, ha);
  c->host_ms[1] = ms(ha, hb);
  c->host_ms[5] = ms(h1, h2);
  c->host_ms[6] = ms(h2, h3);
  c->host_ms[7] = ms(h0, h3);
  if (std::getenv("BSRG_HOST_TRACE"))
    std::fprintf(stderr,
                 "bsrg host ms: checks+launch %.3f overlaps %.3f decode %.3f (exchange %.3f sum %.3f early %.3f "
                 "other %.3f; cpu %.3f longest %.3f) acf %.3f sort %.3f upload %.3f total %.3f\n",
                 c->host_ms[0], c->host_ms[1], c->host_ms[3], c->host_ms[10], c->host_ms[11], c->host_ms[12],
                 c->host_ms[13], c->host_ms[8], c->host_ms[9], c->host_ms[4], c->host_ms[5], c->host_ms[6],
                 c->host_ms[7]);
  c->wave_prepared = true;
  return 0;
}

}  // namespace

int bsrg_wave_prepare(bsrg_ctx* c, int nch, int ncp, int mk, int nrec, const int* jtype, const int* kpol,
                      const int* k1, const int* k2, const int* k3, const int* k4, const double* cd, int novl,
                      const int* ok3, const int* ok4, const double* ocd, double eps_c, double ec, int ntarg,
                      const double* etarg, const int* iptar) {
  return wave_prepare(c, false, nch, ncp, mk, nrec, jtype, kpol, k1, k2, k3, k4, cd, novl, ok3, ok4, ocd, eps_c, ec,
                      ntarg, etarg, iptar);
}

int bsrg_lists_upload(bsrg_ctx* c, int nrec, const int* jtype, const int* kpol, const int* k1, const int* k2,
                      const int* k3, const int* k4, const double* cd) {
  if (int rc = check_ctx(c)) return rc;
  if (nrec < 0) return fail(-1, "bsrg_lists_upload: bad sizes");
  if (nrec > 0 && (!jtype || !kpol || !k1 || !k2 || !k3 || !k4 || !cd)) return fail(-2, "null argument");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));  // no kernel still reads the previous lists
  HIPCHK(hipStreamSynchronize(c->copy_stream));
  auto& R = c->rl;
  R.nrec = -1;
  R.any_exch = false;
  R.xklo = 1 << 30;
  R.xkhi = -1;
  for (int r = 0; r < nrec; ++r)
    if (jtype[r] >= 8 && jtype[r] <= 9) {
      R.any_exch = true;
      R.xklo = std::min(R.xklo, kpol[r]);
      R.xkhi = std::max(R.xkhi, kpol[r]);
    }
  const size_t n = static_cast<size_t>(nrec);
  HIPCHK(R.rec.resize(std::max<size_t>(6 * n, 1)));
  HIPCHK(R.cd.resize(std::max<size_t>(n, 1)));
  const int* src[6] = {jtype, kpol, k1, k2, k3, k4};
  for (int q = 0; q < 6 && n; ++q)
    HIPCHK(hipMemcpyAsync(R.rec.p + q * n, src[q], n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  if (n) HIPCHK(hipMemcpyAsync(R.cd.p, cd, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // the caller's arrays may go away on return
  R.nrec = nrec;
  return 0;
}

int bsrg_wave_assemble_resident(bsrg_ctx* c, int nch, int ncp, int mk, int novl, const int* ok3, const int* ok4,
                                const double* ocd, double eps_c, double ec, int ntarg, const double* etarg,
                                const int* iptar) {
  if (int rc = check_ctx(c)) return rc;
  const int rc = wave_prepare(c, true, nch, ncp, mk, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                              novl, ok3, ok4, ocd, eps_c, ec, ntarg, etarg, iptar);
  if (rc) return rc;
  if (int rc2 = bsrg_wave_launch(c)) return rc2;
  return bsrg_wave_finish(c);
}

int bsrg_wave_launch(bsrg_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->wave_prepared) return fail(-8, "no prepared wave (bsrg_wave_prepare)");
  const bool timed = !c->capturing;  // no event records inside a graph capture
  if (timed) (void)hipEventRecord(c->ev[2], c->stream);
  if (!c->s_enqueued)
    if (int rc = enqueue_s(c)) return rc;
  c->s_enqueued = false;
  if (int rc = enqueue_terms(c, c->H, c->stH, true, timed ? c->ev[3] : nullptr, timed ? c->ev[4] : nullptr,
                             c->gemm_enqueued))
    return rc;
  c->gemm_was_early = c->gemm_enqueued;
  c->gemm_enqueued = false;
  if (c->rl.active) {  // no host copy of the device-made term arrays: any block may hold nonzeros
    if (c->H.live.size() < static_cast<size_t>(c->stH.nblk)) c->H.live.resize(c->stH.nblk, 1);
    std::fill(c->H.live.begin(), c->H.live.end(), 1);
  } else {
    update_live(c, c->H, c->stH);
  }
  launch_symmetrize_diag(c->G.ns, c->H.nch, c->H.hcc.p, c->stream);
  HIPCHK(hipGetLastError());
  if (timed) (void)hipEventRecord(c->ev[5], c->stream);
  if (!c->capturing) c->eager_done = true;
  return 0;
}

int bsrg_wave_finish(bsrg_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->wave_prepared) return fail(-8, "no prepared wave (bsrg_wave_prepare)");
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipStreamSynchronize(c->stream));
  float ms = 0, ms2 = 0;
  if (!c->wave_ok) {
    finish_scalars(c, c->S, c->partsS, true);
    if (c->rl.active) {
      if (int rc = dl_finish(c)) return rc;
    } else {
      finish_scalars(c, c->H, c->partsH, true);
    }
    acf_finish(c->H, 0);
    // the reference keeps one bsr_matrix module: H is recorded after the
    // overlap pass, so it carries the otarg accumulated there
    c->H.otarg = c->S.otarg;
    (void)hipEventElapsedTime(&ms, c->ev[0], c->ev[1]); c->timings[0] = ms;
    (void)hipEventElapsedTime(&ms, c->ev[1], c->ev[8]); c->timings[1] = ms;
  }
  (void)hipEventElapsedTime(&ms, c->ev[3], c->ev[4]); c->timings[2] = ms;
  if (c->gemm_was_early) {  // the GEMM ran before the launch: the launch is all term passes
    (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[5]); c->timings[3] = ms;
  } else {
    (void)hipEventElapsedTime(&ms, c->ev[2], c->ev[3]);
    (void)hipEventElapsedTime(&ms2, c->ev[4], c->ev[5]); c->timings[3] = ms + ms2;
  }
  (void)hipEventElapsedTime(&ms, c->ev[c->wave_ok ? 2 : 0], c->ev[5]); c->timings[4] = ms;
  c->wave_ok = true;
  c->wave_ntarg = c->H.nch > 0 ? c->wave_ntarg : 0;
  return 0;
}

int bsrg_wave_assemble(bsrg_ctx* c, int nch, int ncp, int mk, int nrec, const int* jtype, const int* kpol,
                       const int* k1, const int* k2, const int* k3, const int* k4, const double* cd, int novl,
                       const int* ok3, const int* ok4, const double* ocd, double eps_c, double ec, int ntarg,
                       const double* etarg, const int* iptar) {
  if (int rc = check_ctx(c)) return rc;
  c->early_gemm = true;
  const int rc = bsrg_wave_prepare(c, nch, ncp, mk, nrec, jtype, kpol, k1, k2, k3, k4, cd, novl, ok3, ok4, ocd, eps_c,
                                   ec, ntarg, etarg, iptar);
  c->early_gemm = false;
  if (rc) return rc;
  if (int rc2 = bsrg_wave_launch(c)) return rc2;
  return bsrg_wave_finish(c);
}

int bsrg_wave_engine(bsrg_ctx* c) { return c ? c->last_engine : -1; }

int bsrg_wave_host_timings(bsrg_ctx* c, double* ms10) {
  if (int rc = check_ctx(c)) return rc;
  for (int i = 0; i < 10; ++i) ms10[i] = c->host_ms[i];
  return 0;
}

int bsrg_wave_term_counts(bsrg_ctx* c, int which, long long* n8) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->wave_prepared) return fail(-8, "no prepared wave (bsrg_wave_prepare)");
  if (!n8 || which < 0 || which > 1) return fail(-1, "bsrg_wave_term_counts: bad arguments");
  Staged& st = which == 0 ? c->stS : c->stH;
  const Matrices& M = which == 0 ? c->S : c->H;
  const int* bp = st.hp<int>(st.o_bptr);
  const int* rp = st.hp<int>(st.o_rptr);
  const int* vp = st.hp<int>(st.o_vptr);
  std::vector<int> dptr;
  if (which == 1 && c->rl.active) {  // the term arrays were made on the device
    if (!c->wave_ok) return fail(-8, "bsrg_wave_term_counts: the resident wave has not finished");
    dptr.resize(2 * (static_cast<size_t>(st.nblk) + 1) + st.nq + 1);
    HIPCHK(hipMemcpy(dptr.data(), st.d.p + st.o_bptr, (st.nblk + 1) * sizeof(int), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dptr.data() + st.nblk + 1, st.d.p + st.o_rptr, (st.nblk + 1) * sizeof(int),
                     hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dptr.data() + 2 * (st.nblk + 1), st.d.p + st.o_vptr, (st.nq + 1) * sizeof(int),
                     hipMemcpyDeviceToHost));
    bp = dptr.data();
    rp = bp + st.nblk + 1;
    vp = rp + st.nblk + 1;
  }
  long long bb = 0, rb = 0, vq = 0, live = 0;
  for (int b = 0; b < st.nblk; ++b) {
    bb += bp[b + 1] > bp[b];
    rb += rp[b + 1] > rp[b];
    live += b < static_cast<int>(M.live.size()) && M.live[b];
  }
  for (int q = 0; q < st.nq; ++q) vq += vp[q + 1] > vp[q];
  const long long v[8] = {st.nblk, bb, static_cast<long long>(st.nband), rb, static_cast<long long>(st.nr1),
                          vq, static_cast<long long>(st.nvec), live};
  std::copy(v, v + 8, n8);
  return 0;
}

int bsrg_wave_get(bsrg_ctx* c, int which, double* hcc, double* hcb, double* hbb, double* acf, double* htarg,
                  double* otarg) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->wave_ok) return fail(-8, "no assembled wave (bsrg_wave_assemble)");
  HIPCHK(hipSetDevice(c->device));
  return get_matrices(c, which == 0 ? c->S : c->H, hcc, hcb, hbb, acf, htarg, otarg);
}

}  // extern "C"

namespace bsrg {

bool put_record(std::FILE* f, const void* data, size_t bytes) {
  if (bytes > 0x7fffffffu) return false;
  const int32_t n = static_cast<int32_t>(bytes);
  return std::fwrite(&n, 4, 1, f) == 1 && (bytes == 0 || std::fwrite(data, 1, bytes, f) == bytes) &&
         std::fwrite(&n, 4, 1, f) == 1;
}

}  // namespace bsrg

namespace {

// Record_matrix, src/bsr/BSR_MAT3/record_matrix.f90:2-58: diagonal blocks,
// then non-zero lower blocks with (i,j) headers, channel-perturber vectors,
// perturber scalars, and the (0,idiag) terminator.
bool record_matrix(std::FILE* f, const Matrices& M, const std::vector<double>& hcc, const std::vector<double>& hcb) {
  const int ns = M.ns, kch = M.nch, kcp = M.ncp;
  const size_t nss = static_cast<size_t>(ns) * ns;
  bool ok = true;
  for (int ib = 1; ib <= kch && ok; ++ib) ok = put_record(f, &hcc[(static_cast<size_t>(ib) * (ib + 1) / 2 - 1) * nss], nss * 8);
  int idiag = 1;
  for (int ich = 1; ich <= kch && ok; ++ich)
    for (int jch = 1; jch < ich && ok; ++jch) {
      const double* b = &hcc[(static_cast<size_t>(ich) * (ich - 1) / 2 + jch - 1) * nss];
      double s = 0.0;
      for (size_t q = 0; q < nss; ++q) s += std::fabs(b[q]);
      if (s == 0.0) continue;
      const int32_t hdr[2] = {ich, jch};
      ok = put_record(f, hdr, 8) && put_record(f, b, nss * 8);
      idiag = 0;
    }
  if (kcp > 0) {
    for (int ich = 1; ich <= kch && ok; ++ich)
      for (int ip = 1; ip <= kcp && ok; ++ip) {
        const double* v = &hcb[(static_cast<size_t>(ip - 1) * kch + (ich - 1)) * ns];
        double s = 0.0;
        for (int q = 0; q < ns; ++q) s += std::fabs(v[q]);
        if (s == 0.0) continue;
        const int32_t hdr[2] = {ip + kch, ich};
        ok = put_record(f, hdr, 8) && put_record(f, v, static_cast<size_t>(ns) * 8);
        idiag = 0;
      }
    for (int ip = 1; ip <= kcp && ok; ++ip)
      for (int jp = 1; jp <= ip && ok; ++jp) {
        const double v = M.hbb[static_cast<size_t>(ip) * (ip - 1) / 2 + jp - 1];
        if (v == 0.0) continue;
        const int32_t hdr[2] = {ip + kch, jp + kch};
        ok = put_record(f, hdr, 8) && put_record(f, &v, 8);
        if (ip != jp) idiag = 0;
      }
  }
  const int32_t tail[2] = {0, idiag};
  return ok && put_record(f, tail, 8);
}

}  // namespace

extern "C" {

int bsrg_wave_write(bsrg_ctx* c, const char* path, double tmax, int ntarg, const double* etarg, double ec) {
  if (int rc = check_ctx(c)) return rc;
  if (!c->wave_ok) return fail(-8, "no assembled wave (bsrg_wave_assemble)");
  HIPCHK(hipSetDevice(c->device));
  std::vector<double> shcc(c->S.hcc_n()), shcb(c->S.hcb_n()), hhcc(c->H.hcc_n()), hhcb(c->H.hcb_n());
  if (int rc = get_matrices(c, c->S, shcc.data(), shcb.data(), nullptr, nullptr, nullptr, nullptr)) return rc;
  if (int rc = get_matrices(c, c->H, hhcc.data(), hhcb.data(), nullptr, nullptr, nullptr, nullptr)) return rc;
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return fail(-40, std::string("cannot open ") + path);
  const int ns = c->G.ns;
  const int32_t head[3] = {ns, c->H.nch, c->H.ncp};
  bool ok = put_record(f, head, 12);
  ok = ok && record_matrix(f, c->S, shcc, shcb);
  ok = ok && record_matrix(f, c->H, hhcc, hhcb);
  const int32_t mk = c->H.mk;
  ok = ok && put_record(f, &mk, 4);
  ok = ok && put_record(f, c->H.acf.data(), c->H.acf.size() * 8);
  unsigned char tn[12];
  const int32_t ns32 = ns;
  std::memcpy(tn, &tmax, 8);
  std::memcpy(tn + 8, &ns32, 4);
  ok = ok && put_record(f, tn, 12);
  ok = ok && put_record(f, c->H.htarg.data(), c->H.htarg.size() * 8);
  ok = ok && put_record(f, c->H.otarg.data(), c->H.otarg.size() * 8);
  ok = ok && put_record(f, etarg, static_cast<size_t>(ntarg) * 8);
  ok = ok && put_record(f, &ec, 8);
  if (std::fclose(f) != 0) ok = false;
  if (!ok) return fail(-41, std::string("write failed: ") + path);
  return 0;
}

// ---- interrupted runs (BSR_MAT4) ----------------------------------------------
// The partial matrix of an interrupted run, as BSR_MAT4 leaves it when its
// time limit is reached (state_res.f90:187-196): the header (ns, nch,
// npert), Record_matrix (record_matrix.f90:2-46: every nonzero channel
// block (ich >= jch), channel-perturber vector and perturber scalar behind
// an (i, j) header, perturbers as nch + ip, then (-1, -1)), mk, ACF,
// htarg, otarg, and last (interrupt, intercase, max_nbuf), the record
// read_data.f90:39-41 reads back first.
int bsrg_matrix_record(bsrg_ctx* c, const char* path, int interrupt, int intercase, int max_nbuf) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = cur_ready(c)) return rc;
  if (!path) return fail(-2, "null argument");
  HIPCHK(hipSetDevice(c->device));
  Matrices& M = c->cur;
  std::vector<double> hcc(M.hcc_n()), hcb(M.hcb_n());
  if (int rc = get_matrices(c, M, hcc.data(), hcb.data(), nullptr, nullptr, nullptr, nullptr)) return rc;
  std::FILE* f = std::fopen(path, "wb");
  if (!f) return fail(-40, std::string("cannot open ") + path);
  const int ns = M.ns, kch = M.nch, kcp = M.ncp;
  const size_t nss = static_cast<size_t>(ns) * ns;
  const int32_t head[3] = {ns, kch, kcp};
  bool ok = put_record(f, head, 12);
  for (int ich = 1; ich <= kch && ok; ++ich)
    for (int jch = 1; jch <= ich && ok; ++jch) {
      const double* b = &hcc[(static_cast<size_t>(ich) * (ich - 1) / 2 + jch - 1) * nss];
      double s = 0.0;
      for (size_t q = 0; q < nss; ++q) s += std::fabs(b[q]);
      if (s == 0.0) continue;
      const int32_t hdr[2] = {ich, jch};
      ok = put_record(f, hdr, 8) && put_record(f, b, nss * 8);
    }
  for (int ich = 1; ich <= kch && ok; ++ich)
    for (int ip = 1; ip <= kcp && ok; ++ip) {
      const double* v = &hcb[(static_cast<size_t>(ip - 1) * kch + (ich - 1)) * ns];
      double s = 0.0;
      for (int q = 0; q < ns; ++q) s += std::fabs(v[q]);
      if (s == 0.0) continue;
      const int32_t hdr[2] = {ip + kch, ich};
      ok = put_record(f, hdr, 8) && put_record(f, v, static_cast<size_t>(ns) * 8);
    }
  for (int ip = 1; ip <= kcp && ok; ++ip)
    for (int jp = 1; jp <= ip && ok; ++jp) {
      const double v = M.hbb[static_cast<size_t>(ip) * (ip - 1) / 2 + jp - 1];
      if (v == 0.0) continue;
      const int32_t hdr[2] = {ip + kch, jp + kch};
      ok = put_record(f, hdr, 8) && put_record(f, &v, 8);
    }
  const int32_t end[2] = {-1, -1};
  ok = ok && put_record(f, end, 8);
  const int32_t mk = M.mk;
  ok = ok && put_record(f, &mk, 4);
  ok = ok && put_record(f, M.acf.data(), M.acf.size() * 8);
  ok = ok && put_record(f, M.htarg.data(), M.htarg.size() * 8);
  ok = ok && put_record(f, M.otarg.data(), M.otarg.size() * 8);
  const int32_t tail[3] = {interrupt, intercase, max_nbuf};
  ok = ok && put_record(f, tail, 12);
  if (std::fclose(f) != 0) ok = false;
  if (!ok) return fail(-41, std::string("write failed: ") + path);
  return 0;
}

// Resume from such a file (Read_matrix, read_matrix.f90:2-33, and the
// markers of read_data.f90:39-41): the blocks it holds replace those of the
// current matrix (allocated beforehand with the file's sizes, as
// alloc_dbsr_matrix precedes Read_matrix), the others stay as they are.
int bsrg_matrix_resume(bsrg_ctx* c, const char* path, int* interrupt, int* intercase, int* max_nbuf) {
  if (int rc = check_ctx(c)) return rc;
  if (int rc = cur_ready(c)) return rc;
  if (!path) return fail(-2, "null argument");
  HIPCHK(hipSetDevice(c->device));
  Matrices& M = c->cur;
  std::FILE* f = std::fopen(path, "rb");
  if (!f) return fail(-51, std::string("cannot open ") + path);
  std::vector<char> rec;
  auto next = [&]() {
    int32_t n1 = 0, n2 = 0;
    if (std::fread(&n1, 4, 1, f) != 1 || n1 < 0) return false;
    rec.resize(n1);
    if (n1 && std::fread(rec.data(), 1, n1, f) != static_cast<size_t>(n1)) return false;
    return std::fread(&n2, 4, 1, f) == 1 && n2 == n1;
  };
  auto bad = [&](const char* what) {
    std::fclose(f);
    return fail(-50, std::string("partial matrix file: ") + what);
  };
  auto ints = [&](int i) { int32_t v; std::memcpy(&v, rec.data() + 4 * i, 4); return v; };
  if (!next() || rec.size() < 12) return bad("header");
  if (ints(0) != M.ns || ints(1) != M.nch || ints(2) != M.ncp) return bad("sizes differ from the allocated matrix");
  const int ns = M.ns, kch = M.nch;
  const size_t nss = static_cast<size_t>(ns) * ns;
  for (;;) {
    if (!next() || rec.size() < 8) return bad("block header");
    const int ic = ints(0), jc = ints(1);
    if (ic <= 0) break;
    if (ic > kch && jc > kch) {
      const int ip = ic - kch, jp = jc - kch;
      if (ip > M.ncp || jp > ip || !next() || rec.size() != 8) return bad("perturber scalar");
      std::memcpy(&M.hbb[static_cast<size_t>(ip) * (ip - 1) / 2 + jp - 1], rec.data(), 8);
    } else if (ic > kch) {
      const int ip = ic - kch;
      if (ip > M.ncp || jc < 1 || jc > kch || !next() || rec.size() != static_cast<size_t>(ns) * 8)
        return bad("channel-perturber vector");
      HIPCHK(hipMemcpy(M.hcb.p + (static_cast<size_t>(ip - 1) * kch + (jc - 1)) * ns, rec.data(), rec.size(),
                       hipMemcpyHostToDevice));
    } else {
      if (jc < 1 || jc > ic || !next() || rec.size() != nss * 8) return bad("channel block");
      HIPCHK(hipMemcpy(M.hcc.p + (static_cast<size_t>(ic) * (ic - 1) / 2 + jc - 1) * nss, rec.data(), rec.size(),
                       hipMemcpyHostToDevice));
    }
  }
  if (!next() || rec.size() != 4 || ints(0) != M.mk) return bad("mk differs from the allocated matrix");
  if (!next() || rec.size() != M.acf.size() * 8) return bad("ACF");
  std::memcpy(M.acf.data(), rec.data(), rec.size());
  if (!next() || rec.size() != M.htarg.size() * 8) return bad("htarg");
  std::memcpy(M.htarg.data(), rec.data(), rec.size());
  if (!next() || rec.size() != M.otarg.size() * 8) return bad("otarg");
  std::memcpy(M.otarg.data(), rec.data(), rec.size());
  if (!next() || rec.size() != 12) return bad("interrupt record");
  if (interrupt) *interrupt = ints(0);
  if (intercase) *intercase = ints(1);
  if (max_nbuf) *max_nbuf = ints(2);
  std::fclose(f);
  return 0;
}

int bsrg_wave_timings(bsrg_ctx* c, double* ms5) {
  if (int rc = check_ctx(c)) return rc;
  for (int i = 0; i < 5; ++i) ms5[i] = c->timings[i];
  return 0;
}

int bsrg_set_capture_safe(bsrg_ctx* c, int on) {
  if (int rc = check_ctx(c)) return rc;
  c->capture_safe = on != 0;
  return 0;
}

int bsrg_set_gemm_engine(bsrg_ctx* c, int engine) {
  if (int rc = check_ctx(c)) return rc;
  if (engine < -1 || engine > 4) return fail(-1, "bsrg_set_gemm_engine: engine must be -1..4");
  c->forced_engine = engine;
  return 0;
}

struct bsrg_graph {
  std::vector<bsrg_ctx*> ctx;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t t0 = nullptr, t1 = nullptr, fork = nullptr, fork2 = nullptr;
  std::vector<hipEvent_t> join, join2;
  // batched capture: the waves' product descriptors and the work counter
  DevVec<XWave> xw;
  DevVec<int> counter;
  int batch_bpi = 0;  // blocks per work item chosen by the autotuner (0: the waves' own grids)
  int nbatched = 0;   // waves whose products run in the persistent launch
};

int bsrg_graph_destroy(bsrg_graph* g) {
  if (!g) return 0;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (hipEvent_t e : g->join) if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : g->join2) if (e) (void)hipEventDestroy(e);
  if (g->fork2) (void)hipEventDestroy(g->fork2);
  if (g->t0) (void)hipEventDestroy(g->t0);
  if (g->t1) (void)hipEventDestroy(g->t1);
  if (g->fork) (void)hipEventDestroy(g->fork);
  delete g;
  return 0;
}

int bsrg_graph_capture(int n, bsrg_ctx** ctxs, bsrg_graph** out) {
  if (n < 1 || !ctxs || !out) return fail(-1, "bsrg_graph_capture: bad arguments");
  for (int i = 0; i < n; ++i) {
    if (int rc = check_ctx(ctxs[i])) return rc;
    if (!ctxs[i]->wave_prepared) return fail(-8, "bsrg_graph_capture: wave not prepared");
    if (ctxs[i]->device != ctxs[0]->device) return fail(-5, "bsrg_graph_capture: contexts on different devices");
    if (ctxs[i]->stH.has_gemm && ctxs[i]->precision_mode == 0 && ctxs[i]->stH.engine == 0)
      return fail(-9, "bsrg_graph_capture: prepare with bsrg_set_capture_safe(ctx,1)");
  }
  HIPCHK(hipSetDevice(ctxs[0]->device));
  std::unique_ptr<bsrg_graph, int (*)(bsrg_graph*)> g(new bsrg_graph(), bsrg_graph_destroy);
  g->ctx.assign(ctxs, ctxs + n);
  HIPCHK(hipEventCreate(&g->t0));
  HIPCHK(hipEventCreate(&g->t1));
  HIPCHK(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
  g->join.assign(n, nullptr);
  for (int i = 0; i < n; ++i) HIPCHK(hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming));
  hipStream_t s0 = ctxs[0]->stream;
  for (int i = 0; i < n; ++i) HIPCHK(hipStreamSynchronize(ctxs[i]->stream));
  // a wave that already ran eagerly leaves exact live maps: the blocks they
  // mark hold zeros, and every replay writes the same values into the
  // others, so the captured passes may skip them (their maps go into the
  // staged term sets the captured kernels read)
  for (int i = 0; i < n; ++i) {
    bsrg_ctx* c = ctxs[i];
    if (!c->eager_done) continue;
    for (Staged* st : {&c->stS, &c->stH}) {
      Matrices& M = st == &c->stS ? c->S : c->H;
      if (!st->has_live || M.live.size() < static_cast<size_t>(st->nblk)) continue;
      std::memcpy(st->hp<unsigned char>(st->o_live), M.live.data(), st->nblk);
      HIPCHK(hipMemcpyAsync(st->d.p + st->o_live, st->hp<unsigned char>(st->o_live), st->nblk,
                            hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->live_in_capture = true;
  }
  HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeGlobal));
  int rc = 0;
  for (int i = 0; i < n; ++i) ctxs[i]->capturing = true;
  (void)hipEventRecord(g->fork, s0);
  for (int i = 1; i < n && !rc; ++i)
    if (hipStreamWaitEvent(ctxs[i]->stream, g->fork, 0) != hipSuccess) rc = fail(-100, "graph fork failed");
  for (int i = 0; i < n && !rc; ++i) rc = bsrg_wave_launch(ctxs[i]);
  for (int i = 1; i < n && !rc; ++i) {
    (void)hipEventRecord(g->join[i], ctxs[i]->stream);
    if (hipStreamWaitEvent(s0, g->join[i], 0) != hipSuccess) rc = fail(-100, "graph join failed");
  }
  hipError_t e = hipStreamEndCapture(s0, &g->graph);
  for (int i = 0; i < n; ++i) ctxs[i]->capturing = ctxs[i]->live_in_capture = false;
  if (rc) return rc;
  if (e != hipSuccess) return fail(-100, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  return 0;
}

namespace {

// A prepared FP64 wave whose H pass is the fused sparse product (engine 4,
// nothing enqueued yet): its launch step splits into pre (S pass, H lists
// and summed bands), the product (batched) and post (rank-1 and vector
// terms, symmetrisation), which together enqueue what bsrg_wave_launch does.
bool batchable(bsrg_ctx* c) {
  const Staged& st = c->stH;
  // the persistent kernel's tile rows share the LDS with about 1 KB of its
  // own (the waves' first items)
  return c->precision_mode == 0 && st.has_gemm && st.engine == 4 && !c->gemm_enqueued &&
         xasm_rows(c, st) <= xasm_max_t(c->device) - 4 && c->xzero.n >= static_cast<size_t>(st.nblk) + 1;
}

int h_fused_pre(bsrg_ctx* c) {
  if (!c->s_enqueued)
    if (int rc = enqueue_s(c)) return rc;
  c->s_enqueued = false;
  Staged& st = c->stH;
  const int ns = c->G.ns, ks = c->G.ks, nb = c->nb;
  launch_coef_pack(st.nblk, xasm_rows(c, st), st.gemm.p + static_cast<size_t>(st.gemm_klo) * nb * nb,
                   static_cast<size_t>(c->basis_nk) * nb * nb, c->xt.p, c->xc.p, c->xn.p, c->stream);
  c->last_engine = 4;
  launch_term_prep(ns, ks, st.nblk, st.dp<int>(st.o_bptr), st.dp<int>(st.o_bid), st.dp<double>(st.o_bc), c->bands.p,
                   c->bsum.p, st.dp<int>(st.o_rptr), st.dp<int>(st.o_rs), st.dp<double>(st.o_rc), c->St.p, c->rfac.p,
                   c->stream);
  HIPCHK(hipGetLastError());
  return 0;
}

XWave h_fused_desc(bsrg_ctx* c, int bpi) {
  Staged& st = c->stH;
  const int ns = c->G.ns, ks = c->G.ks, nb = c->nb, nr = xasm_rows(c, st);
  XWave x{};
  x.E = c->E.p + static_cast<size_t>(st.gemm_klo) * nb * nb * static_cast<size_t>(ns) * ns;
  x.xo = c->xt.p;
  x.xc = c->xc.p;
  x.xn = c->xn.p;
  x.te = TermEpi{ns, ks, st.dp<int>(st.o_bptr), c->bsum.p};
  x.rptr = st.dp<int>(st.o_rptr);
  x.live = st.has_live && (!c->capturing || c->live_in_capture) ? st.dp<unsigned char>(st.o_live) : nullptr;
  x.hcc = c->H.hcc.p;
  x.ns = ns;
  x.nblk = st.nblk;
  x.nr = nr;
  x.nrs = xasm_slot(nr);
  xasm_grid(ns, st.nblk, x.ntile, x.ngroup);
  // bpi > 0: about bpi blocks per work item, at least one per wavefront
  if (bpi > 0) x.ngroup = std::max(1, std::min((st.nblk + bpi - 1) / bpi, std::max(1, st.nblk / 16)));
  return x;
}

int h_fused_post(bsrg_ctx* c) {
  Staged& st = c->stH;
  const int ns = c->G.ns, ks = c->G.ks;
  launch_assemble(ns, ks, st.nblk, 1, c->xzero.p, st.dp<int>(st.o_bid), st.dp<double>(st.o_bc), c->bands.p,
                  st.dp<int>(st.o_rptr), st.dp<int>(st.o_ru), st.dp<int>(st.o_rw), st.dp<int>(st.o_rs),
                  st.dp<double>(st.o_rc), c->St.p, c->VT.p, c->H.hcc.p, c->bsum.p, c->rfac.p, c->stream, nullptr, 0,
                  0, nullptr, /*prep=*/false);
  launch_vec_assemble(ns, st.nq, 0, st.dp<int>(st.o_vptr), st.dp<int>(st.o_vid), st.dp<int>(st.o_vs),
                      st.dp<double>(st.o_vc), c->St.p, c->VT.p, c->H.hcb.p, c->stream);
  c->gemm_was_early = false;
  c->gemm_enqueued = false;
  update_live(c, c->H, c->stH);
  launch_symmetrize_diag(ns, c->H.nch, c->H.hcc.p, c->stream);
  HIPCHK(hipGetLastError());
  if (!c->capturing) c->eager_done = true;
  return 0;
}

// The products of the batchable waves `bw` (in descending order of work) as
// descriptors with consecutive item numbers; returns the item count.
int batch_descs(const std::vector<bsrg_ctx*>& bw, int bpi, std::vector<XWave>& xs, int& max_nr) {
  xs.clear();
  max_nr = 0;
  int items = 0;
  for (bsrg_ctx* c : bw) {
    XWave x = h_fused_desc(c, bpi);
    x.item0 = items;
    items += x.ntile * x.ngroup;
    max_nr = std::max(max_nr, x.nr);
    xs.push_back(x);
  }
  return items;
}

// Autotuner of the batched product: blocks per work item, timed once per
// batch shape (the waves' ns, used bases and block counts) and cached.
std::mutex g_batch_mu;
std::map<std::vector<long long>, int> g_batch_bpi;

int tune_batch(bsrg_graph* g, const std::vector<bsrg_ctx*>& bw, int nwg) {
  std::vector<long long> key{bw[0]->device};
  for (bsrg_ctx* c : bw)
    key.push_back((static_cast<long long>(c->G.ns) << 40) | (static_cast<long long>(xasm_rows(c, c->stH)) << 24) |
                  c->stH.nblk);
  {
    std::lock_guard<std::mutex> l(g_batch_mu);
    auto it = g_batch_bpi.find(key);
    if (it != g_batch_bpi.end()) return it->second;
  }
  hipStream_t s0 = bw[0]->stream;
  // every wave's lists and summed bands in place, then each candidate timed
  // on the first stream (the products overwrite H with partial sums; the
  // batch's launches recompute everything)
  for (bsrg_ctx* c : bw) {
    if (int rc = h_fused_pre(c)) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  const int cands[] = {0, 32, 64, 128, 256, 512};
  float best = 1e30f;
  int best_bpi = 0;
  std::vector<XWave> xs;
  for (int bpi : cands) {
    int max_nr = 0;
    const int items = batch_descs(bw, bpi, xs, max_nr);
    HIPCHK(g->xw.resize(xs.size()));
    HIPCHK(hipMemcpy(g->xw.p, xs.data(), xs.size() * sizeof(XWave), hipMemcpyHostToDevice));
    float ms = 0;
    for (int rep = 0; rep < 2; ++rep) {
      (void)hipEventRecord(bw[0]->ev[6], s0);
      if (launch_xasm_batch(g->xw.p, static_cast<int>(xs.size()), items, max_nr, nwg, g->counter.p, s0))
        return fail(-100, "batched product: bad launch");
      (void)hipEventRecord(bw[0]->ev[7], s0);
      HIPCHK(hipEventSynchronize(bw[0]->ev[7]));
      (void)hipEventElapsedTime(&ms, bw[0]->ev[6], bw[0]->ev[7]);
    }
    HIPCHK(hipGetLastError());
    if (ms < best) {
      best = ms;
      best_bpi = bpi;
    }
  }
  std::lock_guard<std::mutex> l(g_batch_mu);
  g_batch_bpi[key] = best_bpi;
  return best_bpi;
}

}  // namespace

int bsrg_graph_capture_batched(int n, bsrg_ctx** ctxs, bsrg_graph** out) {
  if (n < 1 || !ctxs || !out) return fail(-1, "bsrg_graph_capture_batched: bad arguments");
  for (int i = 0; i < n; ++i) {
    if (int rc = check_ctx(ctxs[i])) return rc;
    if (!ctxs[i]->wave_prepared) return fail(-8, "bsrg_graph_capture_batched: wave not prepared");
    if (ctxs[i]->device != ctxs[0]->device) return fail(-5, "bsrg_graph_capture_batched: contexts on different devices");
    if (ctxs[i]->stH.has_gemm && ctxs[i]->precision_mode == 0 && ctxs[i]->stH.engine == 0)
      return fail(-9, "bsrg_graph_capture_batched: prepare with bsrg_set_capture_safe(ctx,1)");
    for (int j = 0; j < i; ++j)
      if (ctxs[j] == ctxs[i]) return fail(-1, "bsrg_graph_capture_batched: a context appears twice");
  }
  HIPCHK(hipSetDevice(ctxs[0]->device));
  std::unique_ptr<bsrg_graph, int (*)(bsrg_graph*)> g(new bsrg_graph(), bsrg_graph_destroy);
  g->ctx.assign(ctxs, ctxs + n);
  HIPCHK(hipEventCreate(&g->t0));
  HIPCHK(hipEventCreate(&g->t1));
  HIPCHK(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&g->fork2, hipEventDisableTiming));
  g->join.assign(n, nullptr);
  g->join2.assign(n, nullptr);
  for (int i = 0; i < n; ++i) {
    HIPCHK(hipEventCreateWithFlags(&g->join[i], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&g->join2[i], hipEventDisableTiming));
  }
  hipStream_t s0 = ctxs[0]->stream;
  for (int i = 0; i < n; ++i) HIPCHK(hipStreamSynchronize(ctxs[i]->stream));
  // the waves whose products join the persistent launch, most work first
  std::vector<bsrg_ctx*> bw;
  for (int i = 0; i < n; ++i)
    if (batchable(ctxs[i]) && static_cast<int>(bw.size()) < 256) bw.push_back(ctxs[i]);
  std::stable_sort(bw.begin(), bw.end(), [](bsrg_ctx* a, bsrg_ctx* b) {
    return static_cast<double>(a->stH.nblk) * a->G.ns * a->G.ns > static_cast<double>(b->stH.nblk) * b->G.ns * b->G.ns;
  });
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctxs[0]->device) != hipSuccess || ncu < 1) {
    (void)hipGetLastError();
    ncu = 256;
  }
  HIPCHK(g->counter.resize(1));
  if (!bw.empty()) {
    const int bpi = tune_batch(g.get(), bw, ncu);
    if (bpi < 0) return bpi;
    g->batch_bpi = bpi;
  }
  // exact live maps of waves that ran eagerly (as in bsrg_graph_capture)
  for (int i = 0; i < n; ++i) {
    bsrg_ctx* c = ctxs[i];
    if (!c->eager_done) continue;
    for (Staged* st : {&c->stS, &c->stH}) {
      Matrices& M = st == &c->stS ? c->S : c->H;
      if (!st->has_live || M.live.size() < static_cast<size_t>(st->nblk)) continue;
      std::memcpy(st->hp<unsigned char>(st->o_live), M.live.data(), st->nblk);
      HIPCHK(hipMemcpyAsync(st->d.p + st->o_live, st->hp<unsigned char>(st->o_live), st->nblk,
                            hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->live_in_capture = true;
  }
  std::vector<XWave> xs;
  int max_nr = 0, items = 0;
  for (bsrg_ctx* c : bw) c->capturing = true;  // descriptors as the capture sees them
  items = batch_descs(bw, g->batch_bpi, xs, max_nr);
  for (bsrg_ctx* c : bw) c->capturing = false;
  if (!xs.empty()) {
    HIPCHK(g->xw.resize(xs.size()));
    HIPCHK(hipMemcpy(g->xw.p, xs.data(), xs.size() * sizeof(XWave), hipMemcpyHostToDevice));
  }
  g->nbatched = static_cast<int>(xs.size());
  auto in_batch = [&](bsrg_ctx* c) { return std::find(bw.begin(), bw.end(), c) != bw.end(); };
  HIPCHK(hipStreamBeginCapture(s0, hipStreamCaptureModeGlobal));
  int rc = 0;
  for (int i = 0; i < n; ++i) ctxs[i]->capturing = true;
  // branches: S passes and the H lists of the batched waves, whole launch
  // steps of the others; then the persistent product; then the rank-1 and
  // vector terms of the batched waves
  (void)hipEventRecord(g->fork, s0);
  for (int i = 1; i < n && !rc; ++i)
    if (hipStreamWaitEvent(ctxs[i]->stream, g->fork, 0) != hipSuccess) rc = fail(-100, "graph fork failed");
  for (int i = 0; i < n && !rc; ++i) rc = in_batch(ctxs[i]) ? h_fused_pre(ctxs[i]) : bsrg_wave_launch(ctxs[i]);
  for (int i = 1; i < n && !rc; ++i) {
    (void)hipEventRecord(g->join[i], ctxs[i]->stream);
    if (hipStreamWaitEvent(s0, g->join[i], 0) != hipSuccess) rc = fail(-100, "graph join failed");
  }
  if (!rc && !xs.empty() &&
      launch_xasm_batch(g->xw.p, static_cast<int>(xs.size()), items, max_nr, ncu, g->counter.p, s0))
    rc = fail(-100, "batched product: bad launch");
  (void)hipEventRecord(g->fork2, s0);
  for (int i = 1; i < n && !rc; ++i)
    if (hipStreamWaitEvent(ctxs[i]->stream, g->fork2, 0) != hipSuccess) rc = fail(-100, "graph fork failed");
  for (int i = 0; i < n && !rc; ++i)
    if (in_batch(ctxs[i])) rc = h_fused_post(ctxs[i]);
  for (int i = 1; i < n && !rc; ++i) {
    (void)hipEventRecord(g->join2[i], ctxs[i]->stream);
    if (hipStreamWaitEvent(s0, g->join2[i], 0) != hipSuccess) rc = fail(-100, "graph join failed");
  }
  hipError_t e = hipStreamEndCapture(s0, &g->graph);
  for (int i = 0; i < n; ++i) ctxs[i]->capturing = ctxs[i]->live_in_capture = false;
  if (rc) return rc;
  if (e != hipSuccess) return fail(-100, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  HIPCHK(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  *out = g.release();
  return 0;
}

int bsrg_graph_batch_info(bsrg_graph* g, int* nbatched, int* blocks_per_item) {
  if (!g) return fail(-1, "bsrg_graph_batch_info: no graph");
  if (nbatched) *nbatched = g->nbatched;
  if (blocks_per_item) *blocks_per_item = g->batch_bpi;
  return 0;
}

int bsrg_graph_launch(bsrg_graph* g) {
  if (!g || !g->exec) return fail(-1, "bsrg_graph_launch: no graph");
  HIPCHK(hipSetDevice(g->ctx[0]->device));
  HIPCHK(hipEventRecord(g->t0, g->ctx[0]->stream));
  HIPCHK(hipGraphLaunch(g->exec, g->ctx[0]->stream));
  HIPCHK(hipEventRecord(g->t1, g->ctx[0]->stream));
  return 0;
}

int bsrg_graph_sync(bsrg_graph* g, double* wave_ms, double* total_ms) {
  if (!g || !g->exec) return fail(-1, "bsrg_graph_sync: no graph");
  HIPCHK(hipStreamSynchronize(g->ctx[0]->stream));
  for (size_t i = 0; i < g->ctx.size(); ++i) {
    bsrg_ctx* c = g->ctx[i];
    if (!c->wave_ok)
      if (int rc = bsrg_wave_finish(c)) return rc;
    // device time of the wave's launch step when it ran alone (eager)
    if (wave_ms) wave_ms[i] = c->timings[2] + c->timings[3];
  }
  if (total_ms) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, g->t0, g->t1));
    *total_ms = ms;
  }
  return 0;
}

int bsrg_wave_diff(bsrg_ctx* a, bsrg_ctx* ref, int which, double floor_frac, double* rel) {
  if (int rc = check_ctx(a)) return rc;
  if (int rc = check_ctx(ref)) return rc;
  if (!a->wave_ok || !ref->wave_ok) return fail(-8, "no assembled wave (bsrg_wave_assemble)");
  if (a->device != ref->device) return fail(-5, "bsrg_wave_diff: contexts on different devices");
  const Matrices& A = which == 0 ? a->S : a->H;
  const Matrices& B = which == 0 ? ref->S : ref->H;
  if (A.ns != B.ns || A.nch != B.nch || A.ncp != B.ncp) return fail(-1, "bsrg_wave_diff: different shapes");
  HIPCHK(hipSetDevice(ref->device));
  HIPCHK(hipStreamSynchronize(a->stream));
  hipStream_t s = ref->stream;
  HIPCHK(ref->dtmp.resize(2 * kReduceBlocks));
  std::vector<double> part(2 * kReduceBlocks);
  // scale of the reference: largest element of its channel blocks, vectors and hbb
  launch_absmax(B.hcc_n(), B.hcc.p, ref->dtmp.p, s);
  launch_absmax(B.hcb_n(), B.hcb.p, ref->dtmp.p + kReduceBlocks, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(part.data(), ref->dtmp.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double mx = 0.0;
  for (double v : part) mx = std::max(mx, v);
  for (double v : B.hbb) mx = std::max(mx, std::fabs(v));
  const double floor = std::max(floor_frac * mx, 1e-300);
  launch_reldiff(B.hcc_n(), A.hcc.p, B.hcc.p, floor, ref->dtmp.p, s);
  launch_reldiff(B.hcb_n(), A.hcb.p, B.hcb.p, floor, ref->dtmp.p + kReduceBlocks, s);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(part.data(), ref->dtmp.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double r = 0.0;
  for (double v : part) r = std::max(r, v);
  for (size_t q = 0; q < B.hbb.size(); ++q) r = std::max(r, std::fabs(A.hbb[q] - B.hbb[q]) / std::max(std::fabs(B.hbb[q]), floor));
  *rel = r;
  return 0;
}

}  // extern "C"

namespace bsrg {

DevDbs devdbs(bsrg_ctx* c) {
  auto& D = c->dbs;
  DevDbs g;
  g.ks = D.T.ks; g.ns = D.T.ns; g.nv = D.T.nv;
  for (int q = 0; q < 2; ++q) { g.kb[q] = D.T.kb[q]; g.nsb[q] = D.T.nsb[q]; g.tb[q] = D.tb[q].p; g.val[q] = D.val[q].p; }
  g.t = D.t.p; g.gx = D.gx.p; g.gw = D.gw.p; g.gr = D.gr.p; g.grm = D.grm.p; g.grw = D.grw.p;
  return g;
}

// moments of both bases for k < nk, then the bands of family f for k < nk
int ensure_dbs_band(bsrg_ctx* c, int f, int nk) {
  auto& D = c->dbs;
  if (!D.ok) return fail(-3, "Dirac grid not set (bsrg_dbs_set_grid)");
  if (f < 0 || f > 3) return fail(-1, "family must be 0 (pppp), 1 (qqqq), 2 (pqpq) or 3 (qpqp)");
  if (nk <= D.band_nk[f]) return 0;
  const int ks = D.T.ks, ns = D.T.ns, nv = D.T.nv;
  const DevDbs g = devdbs(c);
  if (D.mom_nk < nk) {
    for (int q = 0; q < 2; ++q) {
      const size_t ikb = static_cast<size_t>(D.T.kb[q]) * (D.T.kb[q] + 1) / 2;
      HIPCHK(D.mom[q].resize(static_cast<size_t>(nk) * 2 * nv * ikb));
      HIPCHK(D.momT[q].resize(D.mom[q].n));
      launch_dbs_moments(g, q, nk, D.mom[q].p, D.momT[q].p, c->stream);
    }
    D.mom_nk = nk;
  }
  const int q1 = (f == 1 || f == 3) ? 1 : 0, q2 = (f == 1 || f == 2) ? 1 : 0;
  const size_t ik1 = static_cast<size_t>(D.T.kb[q1]) * (D.T.kb[q1] + 1) / 2;
  const size_t ik2 = static_cast<size_t>(D.T.kb[q2]) * (D.T.kb[q2] + 1) / 2;
  HIPCHK(D.rkd.resize(static_cast<size_t>(nk) * nv * ik1 * ik2));
  HIPCHK(D.rdT.resize(D.rkd.n));
  for (int q = 0; q < 2; ++q)
    for (int side = 0; side < 2; ++side) HIPCHK(D.msum[q][side].resize(static_cast<size_t>(nk) * 2 * ks * ns));
  HIPCHK(D.rkb[f].resize(static_cast<size_t>(nk) * ns * ns * ks * ks));
  launch_dbs_band(g, f, nk, D.mom[q1].p, D.mom[q2].p, D.momT[q1].p, D.momT[q2].p, D.msum[q1][0].p, D.msum[q2][1].p,
                  D.rkd.p, D.rdT.p, D.rkb[f].p, c->stream);
  HIPCHK(hipGetLastError());
  D.band_nk[f] = nk;
  return 0;
}

}  // namespace bsrg

extern "C" {

int bsrg_dbs_set_grid(bsrg_ctx* c, int ks, int ns, int ksp, int ksq, const double* t) {
  if (int rc = check_ctx(c)) return rc;
  if (ks < 2 || ks > 12 || ns < ks || !t) return fail(-1, "bsrg_dbs_set_grid: bad ks/ns (2 <= ks <= 12)");
  if (ksp < 2 || ksp > ks || ksq < 2 || ksq > ks) return fail(-1, "bsrg_dbs_set_grid: ksp, ksq must be in 2..ks");
  for (int i = 1; i < ns + ks; ++i)
    if (t[i] < t[i - 1]) return fail(-1, "bsrg_dbs_set_grid: knots must be non-decreasing");
  auto& D = c->dbs;
  if (D.ok && D.T.ks == ks && D.T.ns == ns && D.T.kb[0] == ksp && D.T.kb[1] == ksq &&
      std::equal(D.T.t.begin(), D.T.t.end(), t))
    return 0;  // same grid: bands stay resident
  HIPCHK(hipSetDevice(c->device));
  D.T.build(ks, ns, ksp, ksq, t);
  hipStream_t s = c->stream;
  HIPCHK(D.t.upload(D.T.t.data(), D.T.t.size(), s));
  HIPCHK(D.gx.upload(D.T.gx.data(), D.T.gx.size(), s));
  HIPCHK(D.gw.upload(D.T.gw.data(), D.T.gw.size(), s));
  HIPCHK(D.gr.upload(D.T.gr.data(), D.T.gr.size(), s));
  HIPCHK(D.grm.upload(D.T.grm.data(), D.T.grm.size(), s));
  HIPCHK(D.grw.upload(D.T.grw.data(), D.T.grw.size(), s));
  for (int q = 0; q < 2; ++q) {
    HIPCHK(D.tb[q].upload(D.T.tb[q].data(), D.T.tb[q].size(), s));
    HIPCHK(D.val[q].upload(D.T.val[q].data(), D.T.val[q].size(), s));
  }
  HIPCHK(hipStreamSynchronize(s));
  D.ok = true;
  D.mom_nk = 0;
  for (int& n : D.band_nk) n = 0;
  // the Dirac wave's orbitals and bases belong to the old grid
  c->dw.orb_ok = false;
  c->dw.basis_nk = c->dw.exch_nk = 0;
  c->dw.ok = false;
  c->dw.s_ok = c->dw.sol_ok = c->dw.ovl_ok = false;
  return 0;
}

int bsrg_dbs_mrk(bsrg_ctx* c, int family, int k, double* rkb) {
  if (int rc = check_ctx(c)) return rc;
  if (k < 0) return fail(-1, "bsrg_dbs_mrk: k < 0");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = ensure_dbs_band(c, family, k + 1)) return rc;
  const auto& D = c->dbs;
  const size_t n = static_cast<size_t>(D.T.ns) * D.T.ns * D.T.ks * D.T.ks;
  if (rkb) HIPCHK(hipMemcpyAsync(rkb, D.rkb[family].p + k * n, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_dbs_moments(bsrg_ctx* c, int q, int k, double* rk, double* rkm) {
  if (int rc = check_ctx(c)) return rc;
  if (q < 0 || q > 1 || k < 0) return fail(-1, "bsrg_dbs_moments: q must be 0 (P) or 1 (Q), k >= 0");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = ensure_dbs_band(c, 0, k + 1)) return rc;  // builds the moments of both bases
  const auto& D = c->dbs;
  const size_t ikb = static_cast<size_t>(D.T.kb[q]) * (D.T.kb[q] + 1) / 2, nv = D.T.nv;
  const double* m = D.mom[q].p + static_cast<size_t>(k) * 2 * nv * ikb;
  if (rk) HIPCHK(hipMemcpyAsync(rk, m, nv * ikb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (rkm) HIPCHK(hipMemcpyAsync(rkm, m + nv * ikb, nv * ikb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_dbs_convol(bsrg_ctx* c, int family, int k, int icase, const double* d, double* a) {
  if (int rc = check_ctx(c)) return rc;
  if (icase < 1 || icase > 4) return fail(-2, " CONVOL:  unknown case ");
  if (k < 0) return fail(-1, "bsrg_dbs_convol: k < 0");
  HIPCHK(hipSetDevice(c->device));
  if (int rc = ensure_dbs_band(c, family, k + 1)) return rc;
  const auto& D = c->dbs;
  const int ns = D.T.ns, ks = D.T.ks;
  const size_t nin = icase <= 2 ? static_cast<size_t>(ns) * ks : static_cast<size_t>(ns) * ns;
  HIPCHK(c->dtmp.upload(d, nin, c->stream));
  HIPCHK(c->dtmp2.resize(nin));
  const double* R = D.rkb[family].p + static_cast<size_t>(k) * ns * ns * ks * ks;
  if (icase <= 2) launch_convol_direct(ns, ks, icase, R, c->dtmp.p, c->dtmp2.p, c->stream);
  else launch_convol_exchange(ns, ks, R, c->dtmp.p, c->dtmp2.p, c->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(a, c->dtmp2.p, nin * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int bsrg_drop_caches(bsrg_ctx* c) {
  if (int rc = check_ctx(c)) return rc;
  c->band_nk = 0;
  c->basis_nk = 0;
  c->exch_nk = 0;
  // Dirac: moments and bands of the four families, bases, quadr_pq tables
  c->dbs.mom_nk = 0;
  for (int& n : c->dbs.band_nk) n = 0;
  c->dw.basis_nk = c->dw.exch_nk = 0;
  c->dw.Q.clear();
  return 0;
}

int bsrg_set_precision(bsrg_ctx* c, int mode, int nslices) {
  if (int rc = check_ctx(c)) return rc;
  if (mode < 0 || mode > 2) return fail(-9, "precision mode must be 0 (FP64), 1 (bf16 split) or 2 (int8 split)");
  if (mode != 0 && nslices == 0) nslices = 6;
  if (mode != 0 && (nslices < 2 || nslices > 8)) return fail(-9, "split precision: slices must be 2..8");
  c->precision_mode = mode;
  c->nslices = nslices;
  return 0;
}

}  // extern "C"
