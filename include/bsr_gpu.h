/*
 * bsr_gpu.h -- C ABI of the MI355X radial core of bsr_mat.
 *
 * Plain pointers and sizes only.  Arrays use the reference's Fortran
 * column-major layouts so a Fortran host can pass its module arrays
 * directly (ISO_C_BINDING, see bsr_amd/fortran/bsr_gpu.f90):
 *   t(ns+ks), pbs(ns,nbf), rkb(ns,ns,ks,ks), d(ns,ks)|d(ns,ns),
 *   hcc(ns,ns,nch*(nch+1)/2), hcb(ns,nch,ncp), hbb(ncp*(ncp+1)/2),
 *   acf(nch,nch,0:mk), htarg/otarg(nch*(nch+1)/2).
 * Orbital, channel and perturber indices are 1-based as in the reference;
 * K1..K4 use the reference packing (ibi = ibo = 2**15).
 *
 * Every call returns 0 on success or a negative status; bsrg_last_error()
 * gives the message (the reference's Stop text where one exists).
 * A context is used from one host thread at a time.  All arrays passed in
 * are host memory unless the name says _dev.
 *
 * Reference paths below are relative to the jagot/bsr checkout.
 */
#ifndef BSR_GPU_H
#define BSR_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct bsrg_ctx bsrg_ctx;

/* Message for the last failing call on this thread. */
const char* bsrg_last_error(void);
int bsrg_abi_version(void);

/* ---- host-only helpers (no device needed) ---------------------------- */

/* Knots of the semi-exponential grid.
 * Replaces mkgrid1, src/libs/BS/define_grid.f90:213-276 (kind=1) and
 * mkgrid, :86-153 (kind=0).  t must hold cap doubles. */
int bsrg_mkgrid(int kind, int ks, double z, double h, double hmax, double rmax, int cap,
                double* t, int* ns, int* ml, int* me, double* h_out);

/* Gauss points/weights per interval, B-spline values there and the Gram
 * band.  Replaces initvb + mrm(0,sb), src/libs/BS/define_spline.f90:67-129,
 * :162.  gr,grw: (ks,nv) [m fastest]; bsp: (ks,ks,nv) [ith fastest, then m];
 * sb: (ns,ks) lower band. */
int bsrg_grid_tables(int ks, int ns, const double* t, double* gr, double* grw, double* bsp,
                     double* sb);

/* ---- context ---------------------------------------------------------- */

int bsrg_create(bsrg_ctx** ctx, int device);
int bsrg_destroy(bsrg_ctx* ctx);

/* Knot grid and spline tables for the following calls; keeps them resident.
 * Replaces define_grid + define_spline module state (spline_param,
 * spline_grid, spline_galerkin), src/libs/BS/define_spline.f90:15-63.
 * ml, me, h describe the geometric region used by the scaling law. */
int bsrg_set_grid(bsrg_ctx* ctx, int ks, int ns, int ml, int me, double h, const double* t);

/* ---- R^k band (a5-a7) --------------------------------------------------- */

/* Build (or reuse) the resident R^k band for multipole k and optionally
 * copy it to rkb(ns,ns,ks,ks).  Replaces MRK_cell(k),
 * src/libs/BS/mrk_cell.f90:2-82 (with rk_moments, rk_moments.f90:2-101). */
int bsrg_mrk_cell(bsrg_ctx* ctx, int k, double* rkb);

/* The R^k band of multipole k, as bsrg_mrk_cell computes it, written to
 * device memory rkb_dev (ns*ns*ks*ks doubles, the same layout), allocating
 * nothing: no hipMalloc, hipFree or allocating library call.  No
 * counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops).  The grid must be set
 * (bsrg_set_grid) beforehand.  Work arrays (moments, diagonal cells) live
 * in a 16 MB device array of the library, one per GPU, shared by every call
 * on that GPU: calls that overlap on one GPU race.  Grids whose work arrays
 * exceed it (4 nv ik + 2 nv ik^2 + 2 ks ns doubles, ik = ks(ks+1)/2) are
 * rejected (status -6).  stream: a hipStream_t to enqueue on (the call
 * then returns without synchronising), or NULL for the context's stream,
 * synchronised before return. */
int bsrg_mrk_cell_noalloc(bsrg_ctx* ctx, int k, double* rkb_dev, void* stream);

/* Moments and diagonal-cell integrals of multipole k, packed as the
 * reference's first ik=ks(ks+1)/2 rows: rkd1,rkd2 (ik,nv), rkd (ik,ik,nv).
 * Replaces rk_moments(k), src/libs/BS/rk_moments.f90:2-50. */
int bsrg_rk_moments(bsrg_ctx* ctx, int k, double* rkd1, double* rkd2, double* rkd);

/* Convolution of the R^k band with a density, symmetric-symmetric cases.
 * icase 1,2: d(ns,ks) -> a(ns,ks); icase 3,4: d(ns,ns) -> a(ns,ns).
 * Replaces convol(ns,ks,a,d,icase,'s','s'), src/libs/BS/convol.f90:2-330. */
int bsrg_convol(bsrg_ctx* ctx, int k, int icase, const double* d, double* a);

/* ---- Dirac B-spline bands (dbs_mat) ---------------------------------------
 * Knots t(ns+ks) of order ks (ks <= 12); P functions are B-splines of order
 * ksp, Q functions of order ksq, on the same intervals.  Families:
 * 0 PPPP, 1 QQQQ, 2 PQPQ (P pair on r1, Q pair on r2), 3 QPQP.  Bands stay
 * resident while the grid is unchanged.
 * Replaces the grid/spline module state of DBS_grid and DBS_gauss,
 * src/libs/DBS/DBS_gauss.f90:72-252 (alloc_DBS_gauss, Def_pbsp, Def_qbsp). */
int bsrg_dbs_set_grid(bsrg_ctx* ctx, int ks, int ns, int ksp, int ksq, const double* t);

/* rkb(ns,ns,ks,ks) of family f for multipole k (entries beyond the orders
 * are zero).  Replaces mrk_pppp / mrk_qqqq / mrk_pqpq / mrk_qpqp,
 * src/libs/DBS/mrk_pppp.f90:2-77, mrk_qqqq.f90, mrk_pqpq.f90, mrk_qpqp.f90. */
int bsrg_dbs_mrk(bsrg_ctx* ctx, int family, int k, double* rkb);

/* Moments of basis q (0 P, 1 Q) over every cell, pairs i<=j packed:
 * rk(ikb,nv) for r^k and rkm(ikb,nv) for r^-(k+1), ikb = kb(kb+1)/2.
 * Replaces moments_pp / moments_qq, src/libs/DBS/DBS_moments.f90:146-244. */
int bsrg_dbs_moments(bsrg_ctx* ctx, int q, int k, double* rk, double* rkm);

/* convol(ns,ks,a,d,icase,'s','s') with the band of family f, multipole k.
 * Replaces src/libs/DBS/convol.f90:2-268 for the symmetric cases. */
int bsrg_dbs_convol(bsrg_ctx* ctx, int family, int k, int icase, const double* d, double* a);

/* ---- one partial wave (a11-a15) ---------------------------------------- */

/* ---- Dirac partial waves (DBSR_MAT3 R^k part) ------------------------------
 * Grid from bsrg_dbs_set_grid.  Orbitals: pq(ns,2,nbf) large (1) and small
 * (2) components, ipbs(i) = channel of orbital i (0 for bound orbitals), the
 * <.|p> vectors V_ch(2ns,nv_ch) with their keys (i_ch, j_ch) = (channel
 * orbital, bound orbital), strictly increasing as KBORT requires.  Replaces
 * the module state R_data reads: DBS_orbitals_pq (pq, ipbs, V_ch, i_ch,
 * j_ch; src/libs/DBS/DBS_orbitals_pq.f90:2-38, Get_v_ch :383-408). */
int bsrg_dirac_set_orbitals(bsrg_ctx* ctx, int nbf, const double* pq, const int* ipbs, int nv_ch,
                            const int* i_ch, const int* j_ch, const double* v_ch);

/* The R^k part of one Dirac partial wave: every (itype, kpol) list through
 * R_data for the four families PPPP, QQQQ, PQPQ, QPQP, as Gen_matrix does
 * for icase 2 (src/dbsr/DBSR_MAT3_MPI/gen_matrix.f90:11-14, R_data.f90:2-290),
 * accumulated from zero into hch(ms,ms,nblk), hcp(ms,nch*npert), hp,
 * acf(nblk,0:mk), htarg with the UPDATE_* rules of updates.f90:79-289 and
 * Target_h (target_subs.f90:2-31); ms = 2 ns, nblk = nch(nch+1)/2 numbered as
 * alloc_dbsr_matrix does on one process (diagonal blocks first).  Records as
 * in module c_data: record r has structure itype[r] (1..4) and multipole
 * kpol[r]; records of one list are contiguous and in IPT order.  iptar(nch)
 * maps channels to targets 1..ntarg.  iitar = 0 (target-structure records
 * only update htarg).  Exchange records (itype 4) with ich < jch are
 * rejected: the reference's list builder never writes them
 * (idef_type.f90:268-296) and its UPDATE_HS would place their transpose in
 * the untransposed quadrant. */
int bsrg_dirac_assemble(bsrg_ctx* ctx, int nch, int npert, int mk, int ntarg, const int* iptar, int nrec,
                        const int* itype, const int* kpol, const int* k1, const int* k2, const int* k3,
                        const int* k4, const double* cdata);

/* After bsrg_dirac_assemble: copy the accumulators to host arrays in the
 * reference layout (any pointer may be NULL): hch(ms,ms,nblk),
 * hcp(ms,nch*npert) with column (ich-1)*npert+ic, hp(npert(npert+1)/2),
 * acf(nblk,0:mk), htarg(ntarg(ntarg+1)/2). */
int bsrg_dirac_get(bsrg_ctx* ctx, double* hch, double* hcp, double* hp, double* acf, double* htarg);

/* Times (ms) of the last bsrg_dirac_assemble.  Device, from HIP events on
 * the context's stream: [0] bands and bases, [1] exchange GEMMs, [2] band
 * and rank-1 terms with the output layout, [3] symmetrisation of the
 * diagonal blocks and the channel-perturber vectors, [4] first to last event
 * (host list work in between included).  Host: [5] list decoding on the
 * host workers, [6] staging and uploads, [7] the whole call. */
int bsrg_dirac_timings(bsrg_ctx* ctx, double* ms8);

/* Engine of the last Dirac exchange contraction, as bsrg_wave_engine (the
 * precision set by bsrg_set_precision applies to Dirac waves too). */
int bsrg_dirac_engine(bsrg_ctx* ctx);

/* Error of one assembled Dirac wave against a reference result of the same
 * wave (e.g. split precision against FP64): rel = max over the elements of
 * |a - r| / max(|r|, floor_frac * scale), scale = max|r| over hch, hcp and
 * hp together, over acf, over htarg.  Both contexts on one device.  No
 * counterpart in the reference. */
int bsrg_dirac_diff(bsrg_ctx* a, bsrg_ctx* ref, double floor_frac, double* rel);

/* ---- the channel-block steps of DBSR_MAT3's SUB1 ----------------------------
 * S (the overlap matrix) lives beside H in the context, in the same layout.
 *
 * Overlap pass: S from zero, fppqq (the P and Q B-spline overlaps,
 * src/libs/DBS/DBS_gauss.f90:256-280) on every channel diagonal block
 * (src/dbsr/DBSR_MAT3_MPI/sub1.f90:77-78), then O_data (O_data.f90:2-90)
 * over the overlap records in list order: itype 1 <ic|jc> (K1=ic, K2=jc),
 * 2 <i|.> jc (K1=io, K2=jc), 3 <i|.><.|j> (K1=io, K2=jo), 4 <i|j> (K1=io)
 * with io = i1*2^15 + i2 as in module c_data; records with |c| < eps_c are
 * skipped; iitar > 1 adds c*fppqq for itype 4 off the diagonal.  Errors as
 * the reference's stops (-10 range, -12 KBORT=0, -21 unknown itype).
 * Results: shch(ms,ms,nblk), shcp(ms,nch*npert), shp, otarg. */
int bsrg_dirac_overlap(bsrg_ctx* ctx, int nch, int npert, int ntarg, const int* iptar, int nrec, const int* itype,
                       const int* k1, const int* k2, const double* cdata, double eps_c, int iitar);
/* Copy S to host arrays in the layout of bsrg_dirac_get (NULL skips one);
 * otarg(ntarg(ntarg+1)/2). */
int bsrg_dirac_overlap_get(bsrg_ctx* ctx, double* hch, double* hcp, double* hp, double* otarg);

/* UPDATE_HX (updates.f90:79-106) for nrec records in order: block (ich,jch)
 * of S (which = 0) or H (which = 1) += c * D_m (D_m^T when ich < jch),
 * D = d(ms,ms,nmat), m = imat[r] (1-based).  The update L_data makes for
 * its L(ij) records with the one-electron matrices hd(:,:,kappa)
 * (L_data.f90:112-128). */
int bsrg_dirac_update_hx(bsrg_ctx* ctx, int which, int nmat, const double* d, int nrec, const int* ich,
                         const int* jch, const int* imat, const double* c);

/* H_ii += Ecore S_ii + (Etarg(iptar(i)) - Ecore) fppqq for every channel:
 * the core and target energies of the diagonal blocks (sub1.f90:96,
 * :128-130).  etarg(ntarg). */
int bsrg_dirac_energy_shift(bsrg_ctx* ctx, double ecore, const double* etarg);

/* Diag_channels (diag.f90:2-81) after Zero_cond (zero_cond.f90:2-56): per
 * channel, H_ii c = E S_ii c on the B-splines kept (lch, kch: l and kappa of
 * each channel; ilzero, jlzero < 0 take the defaults from l and kappa;
 * ibzero, jbzero deleted at the boundary); solutions with E in [edmin,
 * edmax] and |E| >= egap, largest component made positive.  The reference's
 * defaults: ilzero = jlzero = -1, ibzero = jbzero = 1, edmin = -2 c^2,
 * edmax = 1e10, egap = 1e-3.  ipsol(nch) (may be NULL), *nsol = sum. */
int bsrg_dirac_diag_channels(bsrg_ctx* ctx, const int* lch, const int* kch, int ilzero, int jlzero, int ibzero,
                             int jbzero, double edmin, double edmax, double egap, int* ipsol, int* nsol);
/* The solutions: diag(ms,ms,nch) as the reference holds it (column k <
 * ipsol(ich) solution k, energies in column ms, the rest of S_ii), eval(nsol)
 * in channel order (sub1.f90:148-152), iprm(nch*ms) of Zero_cond; any may
 * be NULL. */
int bsrg_dirac_solutions(bsrg_ctx* ctx, double* diag, double* eval, int* iprm);

/* transform_matrix (transform_matrix.f90:2-66) of S (which = 0) or H
 * (which = 1) with the last solutions: off-diagonal blocks with an element
 * of at least 1e-20 become C_i^T H_ij C_j in their top-left isol x jsol
 * corner and zeros elsewhere; vectors likewise C_i^T v.  For S also the
 * overlaps matrix (nch+npert)^2 the reference's Check_mat reads. */
int bsrg_dirac_transform(bsrg_ctx* ctx, int which);
/* overlaps(nch+npert, nch+npert) of the last S transform. */
int bsrg_dirac_overlaps(bsrg_ctx* ctx, double* ov);
/* Device ms of the last overlap pass, channel diagonalisation, S and H
 * transforms. */
int bsrg_dirac_diag_timings(bsrg_ctx* ctx, double* ms4);
/* Record_matrix of DBSR (src/dbsr/DBSR_MAT3_MPI/record_matrix.f90:2-51)
 * for the last transform of S (which = 0) or H (which = 1), as sub1.f90
 * writes dbsr_mat.nnn (unformatted sequential records):
 *   which = 0 starts the file (sub1.f90:154-158: ns, nch, npert, nsp, nsq;
 *   nsol; jpsol(0:nch); eval(nsol); the solution vectors) and adds the
 *   transformed S (sub1.f90:182);
 *   which = 1 appends the transformed H (sub1.f90:271), then mk and the
 *   asymptotic coefficients acf(icc(ich,jch),k) for all ich, jch, k
 *   (sub1.f90:283-284) after the core term 2*nelc_core on the diagonal k=0
 *   coefficients (:249-253) and the |acf| < 1e-5 cut (:277-279).
 * Record_matrix writes (ich,jch) and the isol(ich) x isol(jch) corner of
 * every off-diagonal channel block with an element of at least 1e-20, then
 * (ip+nch,ich) and the corner of each channel-perturber vector likewise,
 * (ip+nch,jp+nch) and each non-zero perturber element, and (0,0).
 * -6 when that matrix has not been transformed with the current solutions
 * (an update, a new assembly or new solutions since); -40/-41 file errors. */
int bsrg_dirac_record(bsrg_ctx* ctx, const char* path, int which, int nelc_core);

/* Orbitals of the wave: pbs(ns,nbf); iech(i)=channel of orbital i (0 for
 * bound orbitals); lbs(i)=l; ibort(nbf,nbf) nonzero where the overlap
 * <kl|nl> of channel orbital kl and bound orbital nl is kept.
 * Sets up qbs=sb*pbs and the bound overlaps OBS as Read_data does,
 * src/bsr/BSR_MAT3/read_data.f90:142-208, and keeps pbs resident. */
int bsrg_set_orbitals(bsrg_ctx* ctx, int nbf, const double* pbs, const int* iech,
                      const int* lbs, const int* ibort, double eps_ovl);

/* Allocate and clear hcc,hcb,hbb,acf,htarg,otarg on the device.
 * Replaces allocate_matrix, src/bsr/BSR_MAT3/mod_intmat.f90:33-59. */
int bsrg_alloc_matrix(bsrg_ctx* ctx, int nch, int ncp, int mk);

/* hcc(:,:,ij(ich,jch)) += c * sb (symmetric band).
 * Replaces UPDATE_HL(ich,jch,ns,ks,sb,c), mod_intmat.f90:132-160. */
int bsrg_update_hl(bsrg_ctx* ctx, int ich, int jch, double c);

/* Process one sorted coefficient list of structure jtype (1..9) and
 * multipole jpol for R^k (icase 5), accumulating into the device matrices.
 * Records are visited in the order ipt (1-based); K1..K4 and cdata as in
 * module cmdata.  Replaces I_data(jtype,jpol),
 * src/bsr/BSR_MAT3/I_data.f90:2-224 (coupling LS, iitar=0). */
int bsrg_idata(bsrg_ctx* ctx, int icase, int jtype, int jpol, int ncdata, const int* ipt,
               const int* k1, const int* k2, const int* k3, const int* k4, const double* cdata);

/* The coefficient lists of module cmdata (mod_cmdata.f90), host side.
 * bsrg_cdata_build: n successive Add_cdata calls into an empty block
 * (src/bsr/BSR_MAT3/add_cdata.f90:2-56): the records sorted on
 * (K1,K2,K3,K4), the coefficients of equal keys summed in arrival order;
 * outputs hold up to n entries, *nout of them used.
 * bsrg_cdata_merge: Merge_CDATA (merge_cdata.f90:2-91) over nn sorted
 * blocks [ip(m), jp(m)] (1-based, within the n records) of k1..k4/cdata
 * (ipt holds up to n entries): equal keys of different
 * blocks summed into the first block's entry (the others set to 0) as the
 * reference orders those sums, ipt(1..*nc) the surviving entries with
 * |cdata| > eps_c in key order; ip is consumed (set to 0), cdata updated. */
int bsrg_cdata_build(int n, const int* j1, const int* j2, const int* j3, const int* j4, const double* c,
                     int* k1, int* k2, int* k3, int* k4, double* cout, int* nout);
int bsrg_cdata_merge(int nn, int* ip, const int* jp, int n, const int* k1, const int* k2, const int* k3,
                     const int* k4, double* cdata, double eps_c, int* ipt, int* nc);

/* ---- list builder: int_bnk records -> cmdata lists (host) -------------------
 * The record loop of State_res (src/bsr/BSR_MAT3/state_res.f90:2-161) for one
 * icase, on a builder object holding the module state that loop reads.  The
 * relativistic J-dependent cases 7-10 (which read the term tables) are
 * rejected (-44); the reference's stops come back as errors with its
 * message (-45..-48). */
typedef struct bsrg_lb bsrg_lb;
bsrg_lb* bsrg_lb_create(void);
void bsrg_lb_destroy(bsrg_lb* lb);
/* Read an int_bnk file: the configuration headers (read_symc_LS,
 * Read_symt_LS, Read_oper_LS, src/libs/ZCONFLS/MOD_symc.f90:175-195,
 * MOD_symt.f90:224-336, skipped), the determinants and overlap factors
 * (Read_dets, src/bsr/BSR_MAT3/mod_dets.f90:50-86), then the coefficient
 * records (c, ijt, int, idf) to the end (state_res.f90:37-41).  Fortran
 * unformatted sequential, 4-byte record markers.  sizes5 = ndet, kdet,
 * ndef, kdef, nrec.  -40 malformed, -41 cannot open. */
int bsrg_lb_read_bank(bsrg_lb* lb, const char* path, int* sizes5);
/* The bank as read (any pointer may be NULL). */
int bsrg_lb_bank_get(bsrg_lb* lb, double* c, int* ijt, int* intb, int* idf, int* kpd, int* ipd, int* npd, int* kpf,
                     int* ipf, int* npf);
/* The same data given directly (module dets: KPD, IPD, NPD, KPF, IPF, NPF,
 * mod_dets.f90:2-34; 0-based pointers ipd/ipf as stored there). */
int bsrg_lb_set_records(bsrg_lb* lb, int nrec, const double* c, const int* ijt, const int* intb, const int* idf);
int bsrg_lb_set_dets(bsrg_lb* lb, int ndet, int kdet, const int* kpd, const int* ipd, const int* npd, int ndef,
                     int kdef, const int* kpf, const int* ipf, const int* npf);
/* Orbitals: iech (= ief, channel of a continuum orbital, 0 bound), l, the
 * packed IBORT and OBS (nbf x nbf, as Read_data leaves them,
 * read_data.f90:150-208) and the electron count nelc (module target). */
int bsrg_lb_set_orbitals(bsrg_lb* lb, int nbf, const int* iech, const int* lef, const int* ibort, const double* obs,
                         int nelc);
/* Last configuration of each channel ipconf(nch) and of each perturber
 * ippert(0:npert) (module channel, MOD_channel.f90:2-46; Ifind_channel). */
int bsrg_lb_set_channels(bsrg_lb* lb, int nch, const int* ipconf, int npert, const int* ippert);
/* States (module conf_LS): state is has the no_state(is) orbitals
 * ip_orb(ip_state(is)+1 ..) and expansion coefficient wc(is); symmetry it
 * covers the states ip_stat(it_state1(it) .. it_state2(it)) (it_state1 = 0:
 * none). */
int bsrg_lb_set_states(bsrg_lb* lb, int ncfg, const int* ip_state, const int* no_state, const double* wc, int lorb,
                       const int* ip_orb, int nsym, const int* it_state1, const int* it_state2, int nstat,
                       const int* ip_stat);
/* eps_c, eps_det, eps_soo and the cmdata blocks nb, mb, kb (module bsr_mat,
 * mod_param.f90:55-66; defaults 1e-10, 1e-10, 1e-3, 2000, 5000, 100). */
int bsrg_lb_set_params(bsrg_lb* lb, double eps_c, double eps_det, double eps_soo, int nb, int mb, int kb);
/* Run the loop for icase with multipoles up to mk.  sizes3 = coefficients
 * passed to Add_coef, list records produced, lists (batches). */
int bsrg_lb_run(bsrg_lb* lb, int icase, int mk, int* sizes3);
/* The coefficients in the order State_res passes them to Add_coef
 * (structure, multipole, K1..K4, c). */
int bsrg_lb_expansion(bsrg_lb* lb, int* itype, int* kpol, int* k1, int* k2, int* k3, int* k4, double* c);
/* The lists in generation order: each batch is what one Gen_matrix call
 * hands to I_data/L_data/O_data (gen_matrix.f90:17-66), records in IPT
 * order. */
int bsrg_lb_lists(bsrg_lb* lb, int* batch, int* jtype, int* jpol, int* k1, int* k2, int* k3, int* k4, double* c);

/* Overlap records (structure 1 only).  Replaces O_data,
 * src/bsr/BSR_MAT3/O_data.f90:2-76; records with |c| < eps_c skipped. */
int bsrg_odata(bsrg_ctx* ctx, int ncdata, const int* ipt, const int* k3, const int* k4,
               const double* cdata, double eps_c);

/* One-electron operator with the closed core, for every l of the orbitals:
 * hl_full(:,:,l) = d2/dr2 - l(l+1)/r^2 + 2z/r (with the Bloch term at the
 * outer boundary) + the direct and exchange potential of the core orbitals
 * 1..kclosd (R^k, k = 0 and |l-l_c| .. min(l+l_c, 7)); then
 * L_int(i,j) = <p_i|hl_full|p_j> and L_vec(:,i) = hl_full p_i for the
 * bound orbitals beyond the core.  Replaces Gen_Lval and INT_core,
 * src/bsr/BSR_MAT3/l_core.f90:92-170, :173-257 (non-relativistic: the
 * mass-velocity branch taken when rel is set is not reproduced).  Needs
 * bsrg_set_orbitals; invalidated by a new grid or new orbitals. */
int bsrg_gen_lval(bsrg_ctx* ctx, double z, int kclosd);

/* Results of bsrg_gen_lval: mlh = largest l; hl_full (ns,ns,0:mlh),
 * L_int (nbf,nbf), L_vec (ns,nbf), column-major (module L_core,
 * l_core.f90:2-48); any output pointer may be NULL. */
int bsrg_lval_get(bsrg_ctx* ctx, int* mlh, double* hl_full, double* l_int, double* l_vec);

/* L-integral records of one (jtype, l) list into the current matrix.
 * Replaces L_data(jtype,l), src/bsr/BSR_MAT3/L_data.f90:2-149: jtype 1
 * scales L_int into HB/HV/HW/<i|j> (UPDATE_HL with sb when iitar > 0 and
 * the channels differ, then Target_h), jtype 2-5 use L_vec (HW/HV), jtype
 * 6-9 add c*hl_full(:,:,l) to block (ich,jch) when ich == jch or
 * iitar > 1, else Target_h(ich,jch,0,c).  iitar as bsr_mat holds it
 * (mod_param.f90:33); records with |c| < eps_c skipped. */
int bsrg_ldata(bsrg_ctx* ctx, int jtype, int l, int ncdata, const int* ipt, const int* k1,
               const int* k2, const int* k3, const int* k4, const double* cdata, int iitar,
               double eps_c);

/* Orthogonality of every channel orbital to the bound orbitals of its l
 * that IBORT does not exempt: H(i,j) -> (1 - B zz_i) H(i,j) (1 - zz_j B)
 * on the channel blocks and hcb(:,i,ip) -> (1 - B zz_i) hcb(:,i,ip), with
 * zz_i = sum of p p^T over those orbitals and B the B-spline Gram matrix.
 * Replaces BS_ORTH, src/bsr/BSR_MAT3/bs_orth.f90:2-92 (called by SUB1 after
 * the diagonal symmetrisation, sub1.f90:195).  which = 0: the matrix of
 * bsrg_alloc_matrix; 1: H of the last bsrg_wave_assemble.  At most 16
 * conditions per channel; a channel without a channel orbital (iech) is an
 * error. */
int bsrg_bs_orth(bsrg_ctx* ctx, int which);

/* Diagonal channel blocks -> (C+C^T)/2.  Replaces sub1.f90:80-84, :187-191. */
int bsrg_symmetrize_diag(bsrg_ctx* ctx);

/* hcc,hcb,hbb *= ec.  Replaces the core-energy shift, sub1.f90:121-122. */
int bsrg_scale(bsrg_ctx* ctx, double ec);

/* ACF symmetrisation, eps_acf=1e-5 cut and k=0 core term.
 * Replaces sub1.f90:211-219. */
int bsrg_acf_finish(bsrg_ctx* ctx, int kcore);

/* Copy the matrices to host arrays (any pointer may be NULL). */
int bsrg_get_matrix(bsrg_ctx* ctx, double* hcc, double* hcb, double* hbb, double* acf,
                    double* htarg, double* otarg);

/* ---- fused partial wave --------------------------------------------------
 * The whole of SUB1's S and H assembly for R^k (sub1.f90:70-219) in one
 * pass: S = sb on channel diagonals + O_data, symmetrised; H = ec*S + all
 * R^k lists (jpol outer, jtype inner, state_res.f90:155-159) +
 * (E_t - ec)*sb on diagonals, symmetrised; ACF finished (kcore=0).
 * Lists are concatenated: record r has structure jtype[r] and multipole
 * kpol[r]; records of one (jtype,kpol) are contiguous and in cmdata order.
 * Orbitals and grid must be set; matrices are (re)allocated here. */
int bsrg_wave_assemble(bsrg_ctx* ctx, int nch, int ncp, int mk, int nrec, const int* jtype,
                       const int* kpol, const int* k1, const int* k2, const int* k3,
                       const int* k4, const double* cdata, int novl, const int* ok3,
                       const int* ok4, const double* ocdata, double eps_c, double ec,
                       int ntarg, const double* etarg, const int* iptar);

/* The R^k lists of a partial wave (arguments as bsrg_wave_assemble's
 * nrec .. cdata) copied into device memory owned by the context, replacing
 * the ones uploaded before; they stay there until the next upload or
 * bsrg_destroy, for any number of bsrg_wave_assemble_resident calls.  No
 * counterpart in the reference: its bsr_mat reads the lists of int_bnk on
 * the host (State_res, src/bsr/BSR_MAT3/state_res.f90:2-160). */
int bsrg_lists_upload(bsrg_ctx* ctx, int nrec, const int* jtype, const int* kpol, const int* k1,
                      const int* k2, const int* k3, const int* k4, const double* cdata);

/* bsrg_wave_assemble with the R^k lists of the last bsrg_lists_upload,
 * decoded on the device: I_data (src/bsr/BSR_MAT3/I_data.f90:2-224) and
 * GET_V (get_V.f90:2-27) per record in one kernel, a stable sort of the
 * terms by output, and sums in list order -- S, H and the host scalars
 * equal bsrg_wave_assemble's bit for bit.  A record the host decoder would
 * reject returns that decoder's status and message (for the first such
 * record in list order), found when the wave finishes. */
int bsrg_wave_assemble_resident(bsrg_ctx* ctx, int nch, int ncp, int mk, int novl, const int* ok3,
                                const int* ok4, const double* ocdata, double eps_c, double ec,
                                int ntarg, const double* etarg, const int* iptar);

/* After bsrg_wave_assemble: copy S (which=0) or H (which=1) to host. */
int bsrg_wave_get(bsrg_ctx* ctx, int which, double* hcc, double* hcb, double* hbb,
                  double* acf, double* htarg, double* otarg);

/* After bsrg_wave_assemble: write bsr_mat.nnn in the reference format
 * (Fortran unformatted sequential, 4-byte markers, little endian).
 * Replaces Record_matrix + the trailing writes of SUB1,
 * src/bsr/BSR_MAT3/record_matrix.f90:2-58, sub1.f90:107-110,202,258,266-270. */
int bsrg_wave_write(bsrg_ctx* ctx, const char* path, double tmax, int ntarg,
                    const double* etarg, double ec);

/* Device time (ms) of the phases of the last bsrg_wave_assemble, measured
 * with HIP events on the context's stream: [0] bands, [1] orbital bases,
 * [2] exchange GEMM, [3] band/rank-1/vector terms, [4] total. */
int bsrg_wave_timings(bsrg_ctx* ctx, double* ms5);

/* bsrg_wave_assemble in three steps, for batching:
 *   prepare  host list processing, uploads, bands and bases, GEMM engine
 *            choice (autotuned per shape); leaves the inputs resident;
 *   launch   device-only work of S and H on the context's stream (no
 *            host synchronisation or allocation: capturable in a hipGraph);
 *   finish   wait, host scalars (hbb, htarg, otarg, ACF), timings.
 * Arguments of prepare as bsrg_wave_assemble.  No counterpart in the
 * reference (its SUB1 does the whole wave in one pass). */
int bsrg_wave_prepare(bsrg_ctx* ctx, int nch, int ncp, int mk, int nrec, const int* jtype,
                      const int* kpol, const int* k1, const int* k2, const int* k3,
                      const int* k4, const double* cdata, int novl, const int* ok3,
                      const int* ok4, const double* ocdata, double eps_c, double ec,
                      int ntarg, const double* etarg, const int* iptar);
int bsrg_wave_launch(bsrg_ctx* ctx);
int bsrg_wave_finish(bsrg_ctx* ctx);

/* Engine of the last exchange contraction: 0 library FP64 GEMM, 1-3 FP64
 * MFMA kernel (32/64/128-row tiles), 4 FP64 product over each block's
 * nonzero coefficients fused with the term pass (H written once), 10+S
 * bf16 split precision and 20+S int8 split precision with S slices. */
int bsrg_wave_engine(bsrg_ctx* ctx);

/* Force the FP64 exchange-contraction engine (0..4 as above) or -1 to let
 * the library choose (default: engine 4 when the wave's exchange bases fit
 * its LDS tile, else the autotuner's fastest of 0..3; 1..4 when capture-safe). */
int bsrg_set_gemm_engine(bsrg_ctx* ctx, int engine);

/* Host time (ms) of the last prepare: [0] checks and band/basis launches,
 * [1] S and the overlap terms of H, [2] registry of GET_V vectors,
 * [3] parallel list decoding (overlapping the band and basis kernels),
 * [4] merge of the decoded parts, [5] sorting and staging, [6] uploads and
 * GEMM autotune, [7] total, [8] decoding time summed over the parts,
 * [9] the longest part. */
int bsrg_wave_host_timings(bsrg_ctx* ctx, double* ms10);

/* After bsrg_wave_prepare: sizes of the staged term sets of S (which=0) or
 * H (which=1): n8 = {channel blocks, blocks with band terms, band terms,
 * blocks with rank-1 terms, rank-1 terms, channel-perturber vectors with
 * terms, vector terms, blocks that may hold nonzeros after the last
 * launch}.  No counterpart in the reference. */
int bsrg_wave_term_counts(bsrg_ctx* ctx, int which, long long* n8);

/* 1: the autotuner of later prepares only considers engines that can be
 * captured in a hipGraph (the hand-written kernels). */
int bsrg_set_capture_safe(bsrg_ctx* ctx, int on);

/* One hipGraph holding the launch step of n prepared waves (one branch per
 * context/stream, joined on the first context's stream).  launch enqueues
 * it; sync waits, finishes every wave and returns per-wave device time and
 * the whole graph's time (ms).  No counterpart in the reference. */
typedef struct bsrg_graph bsrg_graph;
int bsrg_graph_capture(int n, bsrg_ctx** ctxs, bsrg_graph** graph);
int bsrg_graph_launch(bsrg_graph* graph);
int bsrg_graph_sync(bsrg_graph* graph, double* wave_ms, double* total_ms);
int bsrg_graph_destroy(bsrg_graph* graph);

/* As bsrg_graph_capture, but the exchange products of the FP64 waves whose
 * H pass is the fused sparse product (engine 4) run in ONE persistent
 * launch: persistent workgroups (one per CU) take work items (a 64-position
 * tile of a group of blocks of one wave) from a counter, the waves with the
 * most work first.  Branches before it run the S passes and the waves'
 * lists and summed bands, branches after it the rank-1 and vector terms;
 * other waves keep their whole launch step in their branch.  The
 * granularity (blocks per work item) is chosen by timing candidates on the
 * first capture of a batch shape and cached; the timing leaves partial
 * products in H until the graph runs.  Each element is summed as in the
 * wave's own launch: the results equal bsrg_graph_capture's bit for bit.
 * batch_info: the number of waves in the persistent launch and the chosen
 * blocks per item (0: the waves' own grids).  No counterpart in the
 * reference (its waves are separate program runs, bsr_mat.f90:82-91). */
int bsrg_graph_capture_batched(int n, bsrg_ctx** ctxs, bsrg_graph** graph);
int bsrg_graph_batch_info(bsrg_graph* graph, int* nbatched, int* blocks_per_item);

/* Error of one assembled result against a reference result of the same
 * wave (e.g. split precision against FP64), S (which=0) or H (which=1):
 * rel = max over hcc, hcb, hbb of |a - r| / max(|r|, floor_frac * max|r|).
 * Both contexts must be on one device.  No counterpart in the reference. */
int bsrg_wave_diff(bsrg_ctx* a, bsrg_ctx* ref, int which, double floor_frac, double* rel);

/* After bsrg_wave_assemble: the nev lowest eigenvalues e[0..nev) of the
 * generalized problem H x = E S x of the wave, FP64 on the device: the final
 * DPOTRF / DSYGST / DSYEV steps of BSR_HD3's diag_hd
 * (src/bsr/BSR_HD3/diag_hd.f90:271-312) on the full basis (nhm = nch*ns +
 * ncp, no per-channel truncation as diag_hd.f90:78-190 makes first).  Uses
 * 16 nhm^2 bytes of device memory during the call.  -40: S not positive
 * definite; -41: no convergence. */
int bsrg_wave_eigen(bsrg_ctx* ctx, int nev, double* e);

/* The same eigenvalues for the S and H blocks of a bsr_mat.nnn file, read
 * as BSR_HD3 reads it (src/bsr/BSR_HD3/diag_hd.f90:37-63: kch diagonal
 * blocks, then (ic, jc) headers with ns x ns, ns or scalar records, up to
 * ic <= 0, for S then H).  -50: malformed file; -51: cannot open. */
int bsrg_matfile_eigen(bsrg_ctx* ctx, const char* path, int nev, double* e);

/* How bsrg_wave_eigen / bsrg_matfile_eigen solve: 1 dense (DPOTRF, DSYGST
 * and the tridiagonal reduction of the whole standard-form matrix,
 * rocSOLVER sygvdx); 2 block Lanczos on the standard-form matrix (DPOTRF
 * and DSYGST as above, then blocks of max(16, nev+6) vectors with full
 * reorthogonalisation until each of the nev lowest Ritz values has a
 * residual bound below 1e-11 max(1, |E|) -- an eigenvalue lies within that
 * bound -- falling back to the dense solver when its Krylov space is
 * exhausted first); 0 (default) block Lanczos from nhm = 2048 on, dense
 * below.  eigen_info: the method the last call used, its Krylov dimension
 * (nhm for the dense solver) and largest residual bound.  No counterpart in
 * the reference (diag_hd.f90:312 always calls DSYEV). */
int bsrg_set_eigen_method(bsrg_ctx* ctx, int method);
int bsrg_eigen_info(bsrg_ctx* ctx, int* method, int* dim, double* resid);

/* Forget the resident R^k bands and orbital bases (of the BS and of the
 * Dirac partial waves) so the next call rebuilds them (bsrg_set_grid,
 * bsrg_set_orbitals, bsrg_dbs_set_grid and bsrg_dirac_set_orbitals keep them
 * when called again with identical inputs).  No counterpart in the
 * reference. */
int bsrg_drop_caches(bsrg_ctx* ctx);

/* Interrupted runs (BSR_MAT4's time limit, src/bsr/BSR_MAT4/state_res.f90:
 * 187-196): write the matrix of bsrg_alloc_matrix as it stands -- header
 * (ns, nch, npert), Record_matrix (record_matrix.f90:2-46; nonzero channel
 * blocks, channel-perturber vectors and perturber scalars behind (i, j)
 * headers, perturbers as nch + ip, ended by (-1, -1)), mk, ACF, htarg,
 * otarg, and last (interrupt, intercase, max_nbuf).  -40 cannot open, -41
 * write failed. */
int bsrg_matrix_record(bsrg_ctx* ctx, const char* path, int interrupt, int intercase, int max_nbuf);
/* Resume: read such a file into the matrix of bsrg_alloc_matrix (allocated
 * with the file's ns, nch, npert and mk), as Read_matrix does
 * (read_matrix.f90:2-33), and return the markers (read_data.f90:39-41);
 * the caller skips what was done and continues.  -50 malformed or other
 * sizes, -51 cannot open. */
int bsrg_matrix_resume(bsrg_ctx* ctx, const char* path, int* interrupt, int* intercase, int* max_nbuf);

/* Precision of the exchange contraction for later bsrg_wave_assemble
 * calls: 0 = plain FP64 (default); 1 = split precision, Ozaki slices on
 * bf16 MFMA with FP32 accumulation and FP64 recombination; 2 = split
 * precision, the same slices on int8 MFMA with exact int32 accumulation
 * and FP64 recombination.  nslices = 0 chooses 6 (2..8 allowed). */
int bsrg_set_precision(bsrg_ctx* ctx, int mode, int nslices);

#ifdef __cplusplus
}
#endif
#endif
