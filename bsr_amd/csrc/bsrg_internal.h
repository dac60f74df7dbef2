// Internal declarations shared by the host set-up, the kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdio>
#include <cstdint>
#include <string>
#include <vector>

namespace bsrg {

constexpr int kMaxKs = 16;    // largest spline order supported
constexpr int kIbi = 32768;   // packing base of orbital pairs (cmdata ibi)
constexpr int kIbo = 32768;   // packing base of overlaps (bsr_mat ibo)

int fail(int code, const std::string& msg);
double powi(double a, int n);
int mkgrid(int kind, int ks, double z, double h, double hmax, double rmax, std::vector<double>& t,
           int& ns, int& ml, int& me, double& hout);
void gauss_legendre01(int k, double* x, double* w);
void bspline_values(const double* t, int ks, int left, double x, double* b);
double bvmv_sym(int n, int k, const double* a, const double* x, const double* y);
void bxv_sym(int k, int n, const double* b, const double* v, double* y);

// Host copy of the grid tables.  gr/grm/grw: [iv*ks+m];
// bsp: [(iv*ks+m)*ks+ith] = B_{iv+ith}(gr(iv,m)) (0-based spline index);
// sb: (ns,ks) lower band.
struct GridTables {
  int ks = 0, ns = 0, nv = 0, ml = 0, me = 0;
  double h = 0;
  std::vector<double> t, gx, gw, gr, grm, grw, bsp, sb;
  void build(int ks, int ns, int ml, int me, double h, const double* t);
  std::vector<double> mrm(int mm) const;
  // <B_i|B''_j> in lower band storage, slot (0,0) holding <B_{ns-1}|B''_ns>
  std::vector<double> db2() const;
};

// Dirac B-spline tables: Gauss points of the intervals (gauleg rule), the
// P (q=0, order kb[0]=ksp) and Q (q=1, order kb[1]=ksq) bases: knots tb[q],
// values val[q][(iv*ks+m)*kb[q]+ith] of spline iv+ith at point m of interval iv.
void gauss_legendre_nr(int n, double x1, double x2, double* x, double* w);
struct DbsTables {
  int ks = 0, ns = 0, nv = 0, kb[2] = {0, 0}, nsb[2] = {0, 0};
  std::vector<double> t, gx, gw, gr, grm, grw, tb[2], val[2];
  void build(int ks, int ns, int ksp, int ksq, const double* t);
};

// ---- kernel launchers (kernels.hip) -------------------------------------
struct DevGrid {
  int ks, ns, nv, ml, me;
  double h;
  const double *t, *gx, *gw, *gr, *grm, *grw, *bsp;
};

// moments for multipoles k0..k0+nk-1: mom[((kk*2+s)*nv+iv)*ik+ii],
// s=0: r^k, s=1: r^-(k+1); diagonal cells rkd[((kk*nv+iv)*ik+jj)*ik+ii].
void launch_moments(const DevGrid& g, int k0, int nk, double* mom, double* rkd, hipStream_t s);
// R^k bands for the same multipoles: band kk at rkb + kk*ns*ns*ks*ks.
// Scratch: msum nk*2*ks*ns doubles (per-pair sums of the moments), momT and
// rdT cell-major copies of mom and rkd (same sizes).
void launch_band(const DevGrid& g, int nk, const double* mom, const double* rkd, double* msum, double* momT,
                 double* rdT, double* rkb, hipStream_t s);
// upper-band densities dens[p](j,jp) of all unordered bound pairs p.
void launch_pair_dens(int ns, int ks, int nb, const double* P, double* dens, hipStream_t s);
// direct basis D[kk][p](i,ip) = sum d_p(j,jp) rkb_k(i,j,ip,jp);
// part: scratch of nk*npair*ks*ns*ks doubles
void launch_direct_basis(int ns, int ks, int nk, int npair, const double* rkb, const double* dens,
                         double* part, double* D, hipStream_t s);
// exchange basis E[kk][a][b](x,y) = sum K(x,z,y,w) p_a(z) q_b(w), p = P (r1
// pair), q = P2 (r2 pair)
void launch_exchange_basis(int ns, int ks, int nk, int nb, const double* rkb, const double* P,
                           const double* P2, double* E, hipStream_t s);
// scalar table St[kk][p1][p2] = sum_band D[kk][p1]*dens[p2]
void launch_stab(int ns, int ks, int nk, int npair, const double* D, const double* dens, double* St,
                 hipStream_t s);
// vector table V[(kk*npair+p)*nb+b](x) = (D[kk][p] as symmetric band) * p_b
void launch_vtab(int ns, int ks, int nk, int npair, int nb, const double* D, const double* P,
                 double* V, hipStream_t s);
// sb (lower band) -> upper-band form sbu(i,ith) = <B_i|B_{i+ith}>
void launch_sb_upper(int ns, int ks, const double* sb, double* sbu, hipStream_t s);
// Channel blocks in one pass: hcc[blk] = base + sum_t c_t * band_t (symmetric band)
// + sum_r c_r * sfac_r * V[u_r] (x) V[w_r]; CSR term lists over all nblk blocks;
// base = current hcc if has_base, else 0 (then every block is written).
// qnblk > 0 (Dirac): the nblk = 4*qnblk blocks are quadrants q = 2a+b of
// channel blocks of order 2ns; base[blk] (ns x ns, quadrant-major) is read
// where bit q of qmask is set, and the sum is written to rows a*ns..,
// columns b*ns.. of hcc block blk % qnblk.
// Scratch: bsum of nblk*ns*ks doubles (the summed band of each block), rf
// of one double per rank-1 term (its factor).  live (optional, one byte per
// block, qnblk = 0 and has_base = 0 only): 0 marks a block that already
// holds zeros, which is then not rewritten when it has no terms.
void launch_assemble(int ns, int ks, int nblk, int has_base, const int* bptr, const int* bid, const double* bc,
                     const double* bands, const int* rptr, const int* ru, const int* rw, const int* rs,
                     const double* rc, const double* stab, const double* V, double* hcc, double* bsum,
                     double* rf, hipStream_t s, const double* base = nullptr, int qnblk = 0, int qmask = 0,
                     const unsigned char* live = nullptr, bool prep = true);
// The per-block preparation of a term pass alone (k_term_prep): summed
// band bsum of every block with band terms, factor rf of every rank-1 term.
void launch_term_prep(int ns, int ks, int nblk, const int* bptr, const int* bid, const double* bc,
                      const double* bands, double* bsum, const int* rptr, const int* rs, const double* rc,
                      const double* stab, double* rf, hipStream_t s);
// Exchange contraction fused with the term pass (xasm.hip), H overwritten:
// launch_coef_pack turns the dense coefficient matrix C (rows 0..nr-1 of
// columns blk, leading dimension ldc) into per-block nonzero lists (slots of
// xasm_slot(nr) per block: xo = LDS byte offset of the basis, xc = value;
// xn = count); launch_xasm then writes every block once with the exchange
// product from the bases E (columns of ns^2, the nr used ones) and its band
// terms (TermEpi), and adds the rank-1 terms with the term pass (nobands:
// nblk+1 zeros).  The term arena is a term set's, as for launch_assemble.
// xasm_max_t(device): the largest nr whose LDS tile fits (64 nr doubles).
int xasm_max_t(int device);
int xasm_slot(int nr);
void launch_coef_pack(int nblk, int nr, const double* C, size_t ldc, int* xo, double* xc, int* xn, hipStream_t s);
// launch_xprod: the exchange product alone, H = E C over the packed lists,
// every block written (the GEMM's place in the wave; the term pass follows).
// The Dirac form: launch_xquad writes quadrant (qa, qb) of the channel
// blocks (order 2ns) of hch with the product of the family's bases E
// (columns of ns^2) and the lists launch_coef_pack_quad packed from rows
// [0, nr) of C (slots of xasm_slot(nr) per block) for tiles of wt positions,
// wt = xquad_width(nr, device) (64, 48 or 32; 0 when nr > xquad_max_t).
int xquad_width(int nr, int device);
int xquad_max_t(int device);
void launch_coef_pack_quad(int nblk, int nr, int wt, const double* C, size_t ldc, int* xo, double* xc, int* xn,
                           hipStream_t s);
int launch_xquad(int ns, int nblk, int nr, int wt, const double* E, const int* xo, const double* xc, const int* xn,
                 int qa, int qb, double* hch, hipStream_t s);
int launch_xprod(int ns, int nblk, int nr, const double* E, const int* xo, const double* xc, const int* xn,
                 double* hcc, hipStream_t s);
int launch_xasm(int ns, int ks, int nblk, int nr, const double* E, const int* xo, const double* xc, const int* xn,
                const int* bptr, const int* bid, const double* bc, const double* bands, const int* rptr,
                const int* ru, const int* rw, const int* rs, const double* rc, const double* stab, const double* V,
                double* hcc, double* bsum, double* rf, const unsigned char* live, const int* nobands,
                hipStream_t s, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
// The band terms of an overwriting exchange product, added in the product
// kernel's epilogue (term_epi.h): per element, the product, then its band
// term -- the first step of k_assemble after a stored product, which then
// adds the rank-1 terms; a pass with the bands off (bptr all zeros) adds the
// rank-1 terms after such a product, so both routes give the same bits.
// bsum comes from k_term_prep (launch_term_prep).
struct TermEpi {
  int ns, ks;
  const int* bptr;      // band-term CSR over the blocks (only emptiness is read)
  const double* bsum;   // summed band of each block with band terms, ns*ks per block
};
// One wave's fused exchange product in a persistent batched launch
// (launch_xasm_batch): the arguments of its own k_xasm launch and the first
// of its work items (ntile * ngroup items: tile = i % ntile, group = i /
// ntile).
struct XWave {
  const double* E;
  const int* xo;
  const double* xc;
  const int* xn;
  TermEpi te;
  const int* rptr;
  const unsigned char* live;
  double* hcc;
  int ns, nblk, nr, nrs, ntile, ngroup, item0;
};
// The 64-position tiles of an ns^2 block and the default block groups of a
// wave's own launch.
void xasm_grid(int ns, int nblk, int& ntile, int& ngroup);
// The fused products of nw waves (device array `waves`, items numbered
// consecutively in array order, `total` in all) by nwg persistent
// workgroups pulling items from `counter` (one int of device memory,
// cleared here by a kernel); max_nr: the largest nr (LDS tile rows, at
// most xasm_max_t); nw at most 256.
int launch_xasm_batch(const XWave* waves, int nw, int total, int max_nr, int nwg, int* counter, hipStream_t s);
// hcb[q] = base + sum_r c_r * sfac_r * V[v_r]; CSR over all nq positions
void launch_vec_assemble(int ns, int nq, int has_base, const int* ptr, const int* vid, const int* sidx,
                         const double* c, const double* stab, const double* V, double* hcb, hipStream_t s);
void launch_symmetrize_diag(int ns, int nch, double* hcc, hipStream_t s);
// BS_ORTH on all channel blocks (nblk = nch(nch+1)/2) and channel-perturber
// vectors: channel c is projected with the nz[c] columns of Z (and BZ = B Z)
// starting at column zoff[c]; nzmax = max nz (any number).  The work arrays
// of a block go to LDS when they fit, else to scratch (nscratch doubles, at
// least bs_orth_scratch_doubles(ns, nblk, nzmax)).  Returns -1 when nzmax < 1,
// -2 when the scratch is too small.
size_t bs_orth_scratch_doubles(int ns, int nblk, int nzmax);
int launch_bs_orth(int ns, int nblk, int nch, int ncp, int nzmax, const int* zoff, const int* nz, const double* Z,
                   const double* BZ, double* hcc, double* hcb, double* scratch, size_t nscratch, hipStream_t s);
// One-electron operator of one l (Gen_Lval/INT_core): out (ns x ns) =
// sum_t ec[t] E(eid[t]) + the band hl (lower storage) + sum_u dc[u]
// bands(did[u]) (upper storage), the band region taken from the lower
// element so that it is exactly symmetric.
void launch_lcore(int ns, int ks, int nt, const int* eid, const double* ec, const double* E, int nd,
                  const int* did, const double* dc, const double* bands, const double* hl, double* out,
                  hipStream_t s);
// out(:, o) = hl^T P(:, pcol[o]) for o < n.
void launch_lvec(int ns, int n, const int* pcol, const double* P, const double* hl, double* out, hipStream_t s);
// hcc[blk[b]] += sum over t in [ptr[b], ptr[b+1]) of c[t] * L(lid[t]) (its
// transpose when tr[t]), L(l) = hlfull + l*ns*ns; nlisted blocks.
void launch_dense_terms(int ns, int nlisted, const int* ptr, const int* blk, const int* lid, const int* tr,
                        const double* c, const double* hlfull, double* hcc, hipStream_t s);
// max-reductions over n doubles into kReduceBlocks partial results
constexpr int kReduceBlocks = 1024;
void launch_absmax(size_t n, const double* x, double* partial, hipStream_t s);
void launch_reldiff(size_t n, const double* a, const double* b, double floor, double* partial, hipStream_t s);
void launch_scale(size_t n, double a, double* x, hipStream_t s);
void launch_fill(size_t n, double v, double* x, hipStream_t s);
void launch_fill_int(size_t n, int v, int* x, hipStream_t s);
// leaf convolutions of one band with a general density
void launch_convol_direct(int ns, int ks, int jcase, const double* rkb, const double* d, double* a,
                          hipStream_t s);
void launch_convol_exchange(int ns, int ks, const double* rkb, const double* d, double* a,
                            hipStream_t s);

// cell-major copy: in[(b*nv+iv)*w+e] -> out[(b*w+e)*nv+iv] for b < nb
void launch_cells_transpose(int nv, int w, int nb, const double* in, double* out, hipStream_t s);

// ---- Dirac B-spline bands (dbs.hip) --------------------------------------
struct DevDbs {
  int ks, ns, nv, kb[2], nsb[2];
  const double *t, *gx, *gw, *gr, *grm, *grw, *tb[2], *val[2];
};
// moments of basis q (0 P, 1 Q) for k < nk: mom[((kk*2+s)*nv+iv)*ikb+ii]
// (s=0 r^k, s=1 r^-(k+1)) and their cell-major copy momT
void launch_dbs_moments(const DevDbs& g, int q, int nk, double* mom, double* momT, hipStream_t s);
// bands of family f (0 pppp, 1 qqqq, 2 pqpq, 3 qpqp) for k < nk: band kk at
// rkb + kk*ns*ns*ks*ks.  mom1/momT1 of the r1 basis, mom2/momT2 of the r2
// basis; scratch msum1/2 (nk*2*kb*ns), rkd and rdT (nk*nv*ik1*ik2).
void launch_dbs_band(const DevDbs& g, int f, int nk, const double* mom1, const double* mom2, const double* momT1,
                     const double* momT2, double* msum1, double* msum2, double* rkd, double* rdT, double* rkb,
                     hipStream_t s);

// FP64 MFMA GEMM (ozaki.hip): C = A*B (+C), column-major; tm = 1, 2, 4
int mfma_dgemm(int m, int n, int K, const double* A, size_t lda, const double* B, size_t ldb, double* C,
               size_t ldc, bool accumulate, int tm, hipStream_t s);
// split-precision FP64 GEMM (ozaki.hip): C = A*B (+C), column-major
// int8 split GEMM: A-side slices prepared once (kept while A is unchanged),
// B side sliced per call into bbuf; C (+)= A*B.
size_t ozaki8_a_bytes(int m, int K, int S);
size_t ozaki8_b_bytes(int n, int K, int S);
int ozaki8_prepare_a(int m, int K, const double* A, size_t lda, int S, void* abuf, hipStream_t s);
// qns > 0: C points at the origin of a qns x qns submatrix of block 0 of a
// sequence of blocks of order ldo (ldo^2 apart), m = qns^2, and row
// x + qns y of column col lands at (x, y) of that submatrix in block col.
int ozaki8_dgemm(int m, int n, int K, const void* abuf, const double* B, size_t ldb, double* C, size_t ldc,
                 bool accumulate, int S, void* bbuf, hipStream_t s, int qns = 0, int ldo = 0);
size_t ozaki_workspace_bytes(int m, int n, int K, int S);
int ozaki_dgemm(int m, int n, int K, const double* A, size_t lda, const double* B, size_t ldb, double* C,
                size_t ldc, bool accumulate, int S, void* work, hipStream_t s);

// One Fortran unformatted sequential record (4-byte length, data, length);
// false on a write error or a record of 2 GiB or more.
bool put_record(std::FILE* f, const void* data, size_t bytes);

// ---- R^k lists decoded on the device (lists_dev.hip) ----
// Item categories: the term kinds of the term pass (band and rank-1 terms
// by block, vector terms by channel-perturber position), exchange
// coefficients (by slot of the coefficient matrix), host scalars (HB by
// perturber pair, target matrix by block) and ACF group sums (by multipole
// and channel pair).  An item's sort key is (category << kb) | index.
enum { kDlCatBand = 0, kDlCatR1, kDlCatVec, kDlCatGemm, kDlCatHbb, kDlCatHtarg, kDlCatAcf, kDlNumCat };
// Messages of a record the device decoder rejects (the host decoder's
// codes and texts for the same record, capi.hip dl_message).
enum {
  kDlItype = 1, kDlMultipole, kDlPrepared, kDlOrbRange, kDlChanDensity, kDlHb, kDlGetvRange, kDlGetvZero,
  kDlHvChannel, kDlHvBound, kDlHwChannel, kDlCfChannel, kDlStruct1, kDlBav, kDlStruct25, kDlHxChannel
};
struct DecodeArgs {
  int nrec = 0, base = 0;  // records; record r's items are base + 2r and base + 2r + 1
  const int *jt = nullptr, *kp = nullptr, *k1 = nullptr, *k2 = nullptr, *k3 = nullptr, *k4 = nullptr;
  const double* cd = nullptr;
  int nbf = 0, nb = 0, np = 0, nch = 0, ncp = 0, mk = 0, basis_nk = 0, ns = 0;
  const int *bidx = nullptr, *iech = nullptr, *lbs = nullptr, *ibort = nullptr;  // nbf, nbf, nbf, nbf^2
  const double *obs = nullptr, *qbs = nullptr;  // nbf^2, ns x nbf
  const double* Qb = nullptr;                   // <p_a|r^k|p_b> over bound orbitals, (k*nb + a)*nb + b
  const double* St = nullptr;                   // scalar R^k table
  double eps_ovl = 0.0;
  int w0 = 0;       // vector-table id of GET_V pair (ii, jj): w0 + (ii-1) + nbf*(jj-1)
  long nt = 0;      // exchange coefficient rows per block
  int kb = 0;       // index bits of a sort key
  unsigned* key = nullptr;  // items
  int *ia = nullptr, *ib = nullptr, *is = nullptr;
  double* ic = nullptr;
  unsigned char* need = nullptr;  // nbf^2: GET_V pairs the records name
  unsigned long long* err = nullptr;  // min over rejected records of (record << 8 | message)
};
struct DecodeSort {
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  unsigned *skey = nullptr, *sval = nullptr;
  const unsigned* iota = nullptr;
  int nitems = 0;
};
struct DecodeCsr {  // the term arrays of the term pass (Staged layout)
  int nblk = 0, nq = 0;
  double *bc = nullptr, *rc = nullptr, *vc = nullptr;
  int *bptr = nullptr, *bid = nullptr, *rptr = nullptr, *ru = nullptr, *rw = nullptr, *rs = nullptr,
      *vptr = nullptr, *vid = nullptr, *vs = nullptr;
};
struct DecodeOut {
  int* cb = nullptr;  // kDlNumCat + 1 category starts in the sorted items
  double *gemm = nullptr, *hbb = nullptr, *htarg = nullptr, *acf = nullptr;  // dense, zeroed by the caller
  double* VT = nullptr;                                                     // vector table (GET_V rows)
};
size_t dl_sort_bytes(int nitems, int end_bit);
// The caller's items (nh0 before the records' items, nh1 from item tail on),
// staged in device memory as ic (doubles), key, ia, ib, is.
size_t dl_host_item_bytes(int nh);
void launch_dl_host_items(int nh0, int nh1, size_t tail, const void* staged, const DecodeArgs& a, hipStream_t s);
void launch_dl_iota(size_t n, unsigned* v, hipStream_t s);
// Items of the records, GET_V vectors, sort and CSR arrays (the items
// before base and after base + 2 nrec are the caller's: host terms); then,
// once the scalar R^k table is made, the run sums.
int launch_dl_decode(const DecodeArgs& a, const DecodeSort& so, const DecodeCsr& o, const DecodeOut& out,
                     hipStream_t s);
int launch_dl_runs(const DecodeArgs& a, const DecodeSort& so, const DecodeOut& out, hipStream_t s);

}  // namespace bsrg
