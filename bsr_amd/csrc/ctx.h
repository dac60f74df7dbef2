// Internal state of the C ABI: device/pinned buffers, term sets, staging
// arenas, the host worker pool and the context struct.  Shared by capi.hip
// (BS partial waves) and dirac.hip (Dirac partial waves).
#pragma once
#include <rocblas/rocblas.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

#include "bsrg_internal.h"

namespace bsrg {

int fail(int code, const std::string& msg);
#define HIPCHK(x)                                                                        \
  do {                                                                                   \
    hipError_t e_ = (x);                                                                 \
    if (e_ != hipSuccess) return fail(-100, std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

template <typename T>
struct DevVec {
  T* p = nullptr;
  size_t n = 0;
  DevVec() = default;
  DevVec(const DevVec&) = delete;
  DevVec& operator=(const DevVec&) = delete;
  ~DevVec() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  // Grows only; a regrowth allocates 25% headroom so that a run of jobs of
  // slowly increasing size does not free and allocate (both synchronise)
  // every time.
  hipError_t resize(size_t m) {
    if (m <= n && p) return hipSuccess;
    const size_t want = n ? std::max(m, n + n / 4) : m;
    release();
    if (m == 0) return hipSuccess;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), want * sizeof(T));
    if (e != hipSuccess) {
      (void)hipGetLastError();  // the failed attempt must not surface later
      p = nullptr;
      e = hipMalloc(reinterpret_cast<void**>(&p), m * sizeof(T));
      if (e == hipSuccess) n = m;
      return e;
    }
    n = want;
    return e;
  }
  hipError_t upload(const T* h, size_t m, hipStream_t s) {
    hipError_t e = resize(m);
    if (e != hipSuccess || m == 0) return e;
    return hipMemcpyAsync(p, h, m * sizeof(T), hipMemcpyHostToDevice, s);
  }
};

struct Matrices {
  int nch = 0, ncp = 0, mk = 0, ns = 0;
  DevVec<double> hcc, hcb;
  std::vector<double> hbb, acf, htarg, otarg;
  // live[blk] = 0: channel block blk of hcc is known to hold zeros (blocks
  // keep their numbering when nch changes, so the map outlives a wave)
  std::vector<unsigned char> live;
  size_t nblk() const { return static_cast<size_t>(nch) * (nch + 1) / 2; }
  size_t hcc_n() const { return nblk() * ns * ns; }
  size_t hcb_n() const { return static_cast<size_t>(ns) * nch * ncp; }
  hipError_t alloc(int ns_, int nch_, int ncp_, int mk_) {
    const double* before = hcc.p;
    const int ns_before = ns;
    ns = ns_; nch = nch_; ncp = ncp_; mk = mk_;
    hipError_t e = hcc.resize(std::max<size_t>(hcc_n(), 1));
    if (e != hipSuccess) return e;
    if (hcc.p != before || ns != ns_before) live.assign(nblk(), 1);
    if (live.size() < nblk()) live.resize(nblk(), 1);
    e = hcb.resize(std::max<size_t>(hcb_n(), 1));
    hbb.assign(ncp > 0 ? static_cast<size_t>(ncp) * (ncp + 1) / 2 : 0, 0.0);
    acf.assign(static_cast<size_t>(nch) * nch * (mk + 1), 0.0);
    htarg.assign(nblk(), 0.0);
    otarg.assign(nblk(), 0.0);
    return e;
  }
};

// Page-locked host buffer: uploads from it are asynchronous.  The owner
// makes sure no copy from the old buffer is pending before it grows.
template <typename T>
struct PinnedVec {
  T* p = nullptr;
  size_t cap = 0;
  PinnedVec() = default;
  PinnedVec(const PinnedVec&) = delete;
  PinnedVec& operator=(const PinnedVec&) = delete;
  ~PinnedVec() { if (p) (void)hipHostFree(p); }
  hipError_t reserve(size_t m) {
    if (m <= cap) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    const size_t want = std::max(m, cap + cap / 2);
    hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&p), want * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) cap = want;
    return e;
  }
};

// Term lists for one matrix set.
struct Terms {
  struct Band { int blk, bid; double c; };
  struct R1 { int blk, u, w, sidx; double c; };
  struct Vec { int q, v, sidx; double c; };
  struct Scal { int kind, idx, sidx; double c; };  // kind 0 hbb, 1 htarg, 2 otarg
  PinnedVec<double> gemm;    // nt x nblk, t = kk*nb*nb + a*nb + b (pinned: uploaded as is)
  size_t gemm_n = 0;         // 0 while no exchange term was added
  int gemm_klo = 1 << 30, gemm_khi = -1;
  std::vector<Band> band;
  std::vector<R1> r1;
  std::vector<Vec> vec;
  std::vector<Scal> scal;
  std::vector<double> acf_add;  // same shape as acf
  // A worker part of a parallel pass lists its exchange coefficients and
  // ACF additions (index, value) instead of applying them; they are applied
  // afterwards in part order, i.e. in list order.
  bool deferred = false;
  std::vector<std::pair<size_t, double>> gx, ax;
  int err = 0;
  std::string msg;
  void clear() {
    gemm_n = 0;
    gemm_klo = 1 << 30;
    gemm_khi = -1;
    band.clear(); r1.clear(); vec.clear(); scal.clear(); acf_add.clear();
    gx.clear(); ax.clear();
    err = 0;
    msg.clear();
  }
};

// CSR arrays of one term set, sorted by output (list order kept within an
// output), row pointers over every output, in one pinned host arena that
// goes to the device in one copy.
struct Staged {
  PinnedVec<char> h;
  DevVec<char> d;
  DevVec<double> gemm;
  hipEvent_t done = nullptr;  // the last upload from h / the term set's gemm has completed
  bool pending = false;
  std::vector<int> cnt[3];    // counting-sort scratch (band, rank-1, vector sets)
  int nblk = 0, nq = 0;
  size_t nband = 0, nr1 = 0, nvec = 0, bytes = 0;
  // byte offsets of the arrays in the arena
  size_t o_bc = 0, o_rc = 0, o_vc = 0, o_bptr = 0, o_bid = 0, o_rptr = 0, o_ru = 0, o_rw = 0, o_rs = 0, o_vptr = 0,
         o_vid = 0, o_vs = 0, o_live = 0;
  bool has_gemm = false, has_live = false;
  int gemm_klo = 0, gemm_khi = -1, engine = 2;
  template <typename T>
  T* hp(size_t o) { return reinterpret_cast<T*>(h.p + o); }
  template <typename T>
  const T* dp(size_t o) const { return reinterpret_cast<const T*>(d.p + o); }
};

// Fixed set of host worker threads owned by a context: run(n, f) calls
// f(0..n-1), each index once, on the workers and the calling thread (indices
// are taken from a shared counter), and returns when all are done.
class Pool {
 public:
  explicit Pool(int nthreads) {
    for (int t = 1; t < nthreads; ++t) th_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(m_);
      stop_ = true;
    }
    cv_.notify_all();
    for (auto& x : th_) x.join();
  }
  int size() const { return static_cast<int>(th_.size()) + 1; }
  void run(int n, const std::function<void(int)>& f) {
    {
      std::lock_guard<std::mutex> g(m_);
      f_ = &f;
      n_ = n;
      next_.store(0);
      busy_ = static_cast<int>(th_.size());
      ++gen_;
    }
    cv_.notify_all();
    for (int i; (i = next_.fetch_add(1)) < n;) f(i);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return busy_ == 0; });
    f_ = nullptr;
  }

 private:
  void loop() {
    unsigned seen = 0;
    for (;;) {
      const std::function<void(int)>* f;
      int n;
      {
        std::unique_lock<std::mutex> lk(m_);
        cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        f = f_;
        n = n_;
      }
      for (int i; (i = next_.fetch_add(1)) < n;) (*f)(i);
      std::lock_guard<std::mutex> g(m_);
      if (--busy_ == 0) done_.notify_one();
    }
  }
  std::vector<std::thread> th_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  const std::function<void(int)>* f_ = nullptr;
  std::atomic<int> next_{0};
  int n_ = 0, busy_ = 0;
  unsigned gen_ = 0;
  bool stop_ = false;
};

// A Dirac partial wave (R^k part of DBSR_MAT3).  Channel blocks are ms x ms
// (ms = 2 ns: large then small component); on the device they are kept
// quadrant-major, hq[q][blk] with q = 2a+b (a, b = 0 P, 1 Q) an ns x ns
// block, so that the exchange GEMM of each family writes plain columns; the
// output hch[blk] is the reference's ms x ms layout.  Blocks are numbered as
// alloc_dbsr_matrix numbers them on one process: diagonal blocks first.
struct DiracWave {
  // orbitals
  bool orb_ok = false;
  int nbf = 0, nb = 0, npair = 0, nvch = 0;
  std::vector<double> pq;      // (ns,2,nbf)
  std::vector<int> ipbs, bidx; // bidx: orbital -> bound index, -1 for channel orbitals
  std::vector<int> kcol;       // (nbf*nbf): 1-based V_ch column of (i,j), 0 if none
  std::vector<double> vch;     // (ms, nvch)
  DevVec<double> P;            // [c][b](x): component c of bound orbital b
  DevVec<double> dens;         // [c][p](ns*ks): pair densities of component c
  std::vector<std::vector<double>> Q;  // Q[k][a*nb+b] = quadr_pq(a,b,k), bound orbitals
  // bases of the four families for k < basis_nk (exchange for k < exch_nk)
  int basis_nk = 0, exch_nk = 0;
  long e_ver = 0;
  DevVec<double> D, St, Vf, E, dpart;  // per family
  // families summed where a record uses them with one coefficient: direct
  // bands per quadrant Dc[h][k][p] (h = r1 component), scalars Sc[k][pB][pA]
  // over the four families, vectors VT = [Vc[h][k][p][b] | V_ch halves]
  DevVec<double> Dc, Sc, VT;
  PinnedVec<double> St_host;
  // matrices
  int nch = 0, npert = 0, mk = 0, ntarg = 0;
  size_t nblk = 0;
  std::vector<int> iptar;      // target of each channel
  DevVec<double> hq, hv, hch, hcp, coef;
  std::vector<double> hp, acf, htarg;
  // lists of the last wave, decoded in parts (contiguous record ranges, in
  // list order) on the host workers
  struct XC { int blk, t; double c; };
  struct Part {
    Terms T;
    std::vector<XC> xc;                         // exchange coefficients (block, basis row, value)
    std::vector<std::pair<size_t, double>> ax;  // ACF additions (index, value)
    int klo[4], khi[4];
    int err = 0;
    std::string msg;
  };
  std::vector<std::unique_ptr<Part>> parts;
  int nparts = 0;
  Staged st;
  PinnedVec<char> xh;          // CSR of the exchange coefficients over blocks, uploaded in one copy
  DevVec<char> xd;
  size_t xbytes = 0, o_xptr = 0, o_xt = 0, o_xc = 0;
  std::vector<int> xcnt;
  int klo[4] = {0, 0, 0, 0}, khi[4] = {-1, -1, -1, -1};  // multipole range of each family's exchange terms
  // split precision: int8 slices of each family's exchange matrices, kept
  // while the matrices are unchanged (e_ver), the coefficient-side and bf16
  // workspaces
  DevVec<char> oz8a[4], oz8b, ozk;
  long oz8a_ver[4] = {-1, -1, -1, -1};
  int oz8a_S[4] = {0, 0, 0, 0};
  // sparse FP64 product: per-block nonzero lists of the coefficients
  DevVec<int> qxo, qxn;
  DevVec<double> qxc;
  int engine = -1;
  bool ok = false;
  bool fused = false;  // the last wave's products wrote the output layout directly (sparse FP64, int8 split)
  double ms[5] = {0, 0, 0, 0, 0};
  double host_ms[3] = {0, 0, 0};  // list decoding, staging and uploads, whole call
  // overlap pass: S channel blocks, channel-perturber vectors (device),
  // perturber and target overlaps (host); the B-spline overlaps fppqq and
  // V_ch on the device; term lists of the block passes
  bool s_ok = false;
  int s_nch = 0, s_npert = 0, s_ntarg = 0;
  std::vector<int> s_iptar;
  DevVec<double> fpq, vchd, shch, shcp, dmat;
  std::vector<double> shp, sotarg;
  DevVec<int> ti, tb;
  DevVec<int4> tk;
  DevVec<double> tcd;
  DevVec<const double*> mp;
  // channel solutions diag(ms,ms,nch), their counts, the kept B-splines
  bool sol_ok = false;
  DevVec<double> sol, dwork, bmax;
  DevVec<int> keepd, posd, ksold, isold, iwork;
  std::vector<int> ipsol, iprm;
  // largest transformed overlaps (nch+npert)^2 of the last S transform
  bool ovl_ok = false;
  int ovl_n = 0;
  std::vector<double> ovl;
  double dms[4] = {0, 0, 0, 0};  // overlap pass, channel diagonalisation, S and H transforms
  bool tr[2] = {false, false};    // S / H transformed with the current solutions
  DevVec<double> rec;             // Record_matrix staging: block corners, then their maxima
  DevVec<int> reci;
  DevVec<size_t> reco;
};

}  // namespace bsrg

using namespace bsrg;

struct bsrg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // uploads that may run beside the compute stream (the Dirac term arena
  // while the exchange GEMMs compute); default priority like the other
  hipStream_t copy_stream = nullptr;
  rocblas_handle blas = nullptr;
  // eigenvalues: 0 automatic (block Lanczos from nhm 2048 on, else the dense
  // solver), 1 dense (sygvdx), 2 block Lanczos; what the last call used, its
  // Krylov dimension and largest residual bound
  int eigen_method = 0, eigen_used = 0, eigen_dim = 0;
  double eigen_resid = 0.0;
  hipEvent_t ev[10] = {};  // 0-5 phases of a wave, 6-7 autotuning, 8 end of the bases
  double timings[5] = {0, 0, 0, 0, 0};
  // host ms of the last prepare: [0] checks and band/basis launches, [1] S and
  // the overlap terms of H, [2] registry of GET_V vectors, [3] parallel list
  // decoding, [4] merge of the parts, [5] sort and staging, [6] uploads and
  // GEMM autotune, [7] total, [8] decoding time summed over the parts (cpu
  // ms), [9] the longest part
  double host_ms[16] = {};  // 0-9 as bsrg_wave_host_timings reports them; 10-13 the decode phases
  int precision_mode = 0, nslices = 0;

  GridTables G;
  bool grid_ok = false;
  DevVec<double> t, gx, gw, gr, grm, grw, bsp, sb;
  int band_nk = 0;  // bands built for k = 0..band_nk-1
  DevVec<double> rkb, mom, rkd, msum, momT, rdT;

  // orbitals
  bool orb_ok = false;
  int nbf = 0, nb = 0, npair = 0;
  std::vector<double> pbs, qbs, obs;
  std::vector<double> Pbound;  // coefficients of the bound orbitals the bases were built for
  std::vector<int> iech, lbs, ibort, bound, bidx;
  double eps_ovl = 1e-8;
  DevVec<double> P, dens;
  // bases for k = 0..basis_nk-1
  int basis_nk = 0, exch_nk = 0;
  DevVec<double> bands;  // [sbu | D(kk,p)] each ns*ks
  DevVec<double> E, St, VT, dpart;
  PinnedVec<double> St_host;           // scalar table, read back for the host scalars
  std::vector<std::vector<double>> Q;  // Q[k][(i-1)+nbf*(j-1)] = <p_i|r^k|p_j>, computed lazily
  std::vector<char> Qok;
  std::vector<int> wtab;                // GET_V registry (i1,i2) -> id, -1 if not yet made
  std::vector<double> wvec;             // nW x ns
  int wvec_uploaded = -1;
  // one-electron operator with the core (Gen_Lval): hl_full(:,:,l) for
  // l = 0..mlh on the device, L_int on the host, and L_vec(:,i) as vectors
  // of the W part of the vector table (lvec_id[i], -1 where L_vec is zero)
  bool lval_ok = false;
  int mlh = -1;
  DevVec<double> hlfull, lwork;
  DevVec<int> lwork_i;
  std::vector<double> lint, lvec;
  std::vector<int> lvec_id;

  Matrices cur, S, H;
  bool wave_ok = false, wave_prepared = false, capture_safe = false, capturing = false;
  // the prepared wave ran eagerly (not captured) since bsrg_wave_prepare, so
  // the live maps say exactly which blocks hold zeros; a capture may then
  // skip them (live_in_capture, set for the duration of a capture)
  bool eager_done = false, live_in_capture = false;
  // one-shot waves (bsrg_wave_assemble) enqueue the H exchange GEMM during
  // the prepare, before the terms are staged
  bool early_gemm = false, gemm_enqueued = false, gemm_was_early = false;
  bool s_enqueued = false;  // the S pass of a one-shot wave was enqueued during the prepare
  int wave_ntarg = 0;

  // staged term uploads: S and H of the fused wave, and the per-call path
  Staged stS, stH, stC;
  DevVec<char> ozk;  // split-precision workspace (bf16 slices)
  // int8 split precision: slices of the exchange bases, kept while the
  // bases are unchanged (e_ver counts their rebuilds), and of the coefficients
  DevVec<char> oz8a, oz8b;
  long e_ver = 0, oz8a_ver = -1;
  int oz8a_S = 0;
  Terms TS, TH;      // term sets of the prepared wave (host scalars at finish)
  Terms TC;          // term set of the per-call path
  Terms THt;         // H terms after the lists: (E_t - ec) * sb on the diagonal
  std::vector<std::unique_ptr<Terms>> work;  // worker parts of the parallel list pass
  int nwork = 0;                             // worker parts used by the last prepare
  std::unique_ptr<Pool> pool;                // host workers of the list pass
  std::vector<Terms*> partsS, partsH;        // term parts of S and H in list order
  int last_engine = -1, forced_engine = -1;
  DevVec<double> dtmp, dtmp2;
  DevVec<double> bsum, rfac;  // term-pass scratch: summed band of each block, factor of each rank-1 term
  // engine 4 (exchange fused with the term pass): nonzeros of each block's
  // exchange coefficients, slots of nr per block
  DevVec<int> xt, xn, xzero;  // xzero: nblk+1 zeros (band pointers of a rank-1-only pass)
  DevVec<double> xc;
  DevVec<double> bsumS, rfacS;  // the same for S, whose pass may be queued while H's scratch grows

  // orbital set counter: changes whenever bsrg_set_orbitals installs new orbitals
  long orb_ver = 0;
  // R^k lists kept in device memory (bsrg_lists_upload) and decoded there
  // by bsrg_wave_assemble_resident (lists_dev.hip)
  struct ResidentLists {
    int nrec = -1;  // -1: nothing uploaded
    bool any_exch = false;
    int xklo = 1 << 30, xkhi = -1;  // multipoles of the exchange records
    DevVec<int> rec;                // jtype, kpol, k1, k2, k3, k4: nrec each
    DevVec<double> cd;
    long tab_ver = -1;              // orbital set the tables below hold
    DevVec<int> itab;               // bidx, iech, lbs (nbf each), ibort (nbf^2)
    DevVec<double> dtab;            // obs (nbf^2), qbs (ns x nbf)
    DevVec<double> Qb;              // (mk+1) x nb x nb
    int qb_mk = -1;
    DevVec<unsigned char> need;
    DevVec<unsigned> key, skey, sval, iota;
    DevVec<int> ia, ib, is;
    DevVec<double> ic;
    DevVec<char> sort_tmp;
    size_t sort_tmp_bytes = 0;
    int sort_items = -1, sort_bits = -1;
    DevVec<double> outs;            // sums of HB (perturber pairs), target-matrix and ACF items
    DevVec<int> cb;
    DevVec<unsigned long long> err;
    PinnedVec<char> hitems;         // the host terms of H as items
    DevVec<char> hstage;            // and their device copy
    hipEvent_t hitems_done = nullptr;
    hipEvent_t dec_done = nullptr;  // end of the decoding on the second stream
    bool hitems_pending = false;
    PinnedVec<char> back;           // category starts, rejection word and the sums, read back
    size_t n_hbb = 0, n_htarg = 0, n_acf = 0;
    bool active = false;            // the prepared wave decodes its lists on the device
  } rl;

  // Dirac B-spline grid and the resident bands of the four families
  struct Dbs {
    bool ok = false;
    DbsTables T;
    DevVec<double> t, gx, gw, gr, grm, grw, tb[2], val[2];
    int mom_nk = 0;                       // moments of both bases for k < mom_nk
    DevVec<double> mom[2], momT[2], msum[2][2];
    int band_nk[4] = {0, 0, 0, 0};       // bands of family f for k < band_nk[f]
    DevVec<double> rkd, rdT, rkb[4];
  } dbs;
  DiracWave dw;
};

namespace bsrg {
// Before a term set or its arena is rewritten: the last upload from them
// has completed.
int wait_staged(Staged& st);
// Host worker threads of a context's pool.
int host_workers();
// Sort the terms of all parts by output block (nblk blocks) and vector
// output (nq vectors), list order kept within an output, into the pinned
// arena of st (CSR arrays; see Staged).
int stage_terms(const std::vector<Terms*>& parts, Staged& st, int nblk, int nq, Pool* pool,
                const std::vector<unsigned char>* live = nullptr);
}  // namespace bsrg

namespace bsrg {
// Dirac grid on the device and the bands of family f for k < nk (built on
// the context's stream, kept resident).
DevDbs devdbs(bsrg_ctx* c);
int ensure_dbs_band(bsrg_ctx* c, int f, int nk);
}  // namespace bsrg

