"""Host mirror of the reference's bsr_mat interface over the C ABI.

Method names follow the reference routines they replace (MRK_cell,
rk_moments, convol, I_data, O_data, UPDATE_HL, ...); arguments keep the
reference's meaning and 1-based indices, and errors raise BsrGpuError with
the reference's Stop text where one exists.
"""
from __future__ import annotations

import ctypes as C

import os

import numpy as np

from . import lib as L
from .lib import f64, i32


class Context:
    """One device context: resident grid tables, R^k bands, orbital bases
    and the matrices of the current partial wave."""

    def __init__(self, device: int = 0):
        self._p = C.c_void_p()
        L.check(L.lib().bsrg_create(C.byref(self._p), device))
        self.ks = self.ns = 0
        self.nch = self.ncp = self.mk = 0
        self._wave = None

    def close(self):
        if self._p:
            L.lib().bsrg_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- grid and integrals ------------------------------------------------
    def set_grid(self, ks, ns, ml, me, h, t):
        L.check(L.lib().bsrg_set_grid(self._p, ks, ns, ml, me, h, f64(t)))
        self.ks, self.ns = ks, ns

    def mrk_cell(self, k, copy=True):
        """R^k band rkb(ns,ns,ks,ks) (mrk_cell.f90:2-82)."""
        out = np.zeros((self.ns, self.ns, self.ks, self.ks), order="F") if copy else None
        L.check(L.lib().bsrg_mrk_cell(self._p, k, out))
        return out

    def rk_moments(self, k):
        ks, ns = self.ks, self.ns
        nv, ik = ns - ks + 1, ks * (ks + 1) // 2
        r1 = np.zeros(nv * ik, order="F")
        r2 = np.zeros(nv * ik, order="F")
        rd = np.zeros(nv * ik * ik, order="F")
        L.check(L.lib().bsrg_rk_moments(self._p, k, r1, r2, rd))
        return r1.reshape(nv, ik), r2.reshape(nv, ik), rd.reshape(nv, ik, ik)

    def convol(self, k, d, icase):
        """convol(ns,ks,a,d,icase,'s','s') for the band of multipole k."""
        ns, ks = self.ns, self.ks
        shape = (ns, ks) if icase <= 2 else (ns, ns)
        d = f64(d)
        if d.shape != shape:
            raise ValueError(f"density must have shape {shape}")
        a = np.zeros(shape, order="F")
        L.check(L.lib().bsrg_convol(self._p, k, icase, d, a))
        return a

    # ---- Dirac B-spline bands (four families) ---------------------------------
    def dbs_set_grid(self, ks, ns, ksp, ksq, t):
        """Knots t(ns+ks) of order ks; P functions of order ksp, Q of order
        ksq (DBS_grid / DBS_gauss module state)."""
        L.check(L.lib().bsrg_dbs_set_grid(self._p, ks, ns, ksp, ksq, f64(t)))
        self.dbs = (ks, ns, ksp, ksq)

    def dbs_mrk(self, family, k):
        """rkb(ns,ns,ks,ks) of family 0 pppp, 1 qqqq, 2 pqpq, 3 qpqp
        (mrk_pppp.f90 and siblings)."""
        ks, ns = self.dbs[0], self.dbs[1]
        out = np.zeros((ns, ns, ks, ks), order="F")
        L.check(L.lib().bsrg_dbs_mrk(self._p, family, k, out))
        return out

    def dbs_moments(self, q, k):
        """Moments of basis q (0 P, 1 Q), packed pairs x cells: (r^k, r^-(k+1))."""
        ks, ns, ksp, ksq = self.dbs
        kb = ksq if q else ksp
        nv, ikb = ns - ks + 1, kb * (kb + 1) // 2
        a, b = np.zeros((ikb, nv), order="F"), np.zeros((ikb, nv), order="F")
        L.check(L.lib().bsrg_dbs_moments(self._p, q, k, a, b))
        return a, b

    def dirac_set_orbitals(self, pq, ipbs, i_ch, j_ch, v_ch):
        """Two-component orbitals pq(ns,2,nbf), channel pointers ipbs, and the
        <.|p> vectors V_ch(2ns,nv_ch) keyed by sorted (i_ch, j_ch)."""
        ns = self.dbs[1]
        pq = f64(pq)
        nv = len(i_ch)
        v = f64(np.asarray(v_ch, np.float64).reshape(2 * ns, -1)) if nv else np.zeros((2 * ns, 1), order="F")
        L.check(L.lib().bsrg_dirac_set_orbitals(self._p, pq.shape[2], pq, i32(ipbs), nv, i32(i_ch if nv else [0]),
                                                i32(j_ch if nv else [0]), v))

    def dirac_assemble(self, nch, npert, mk, ntarg, iptar, itype, kpol, k1, k2, k3, k4, c):
        """R^k part of a Dirac partial wave (R_data over the four families)."""
        n = len(c)
        z = np.zeros(1, np.int32)
        L.check(L.lib().bsrg_dirac_assemble(
            self._p, nch, npert, mk, ntarg, i32(iptar), n, *(i32(a) if n else z for a in (itype, kpol, k1, k2, k3, k4)),
            f64(c) if n else np.zeros(1)))
        self.dirac_dims = (nch, npert, mk, ntarg)

    def dirac_wave(self, w):
        """Grid, orbitals and lists of a synthetic Dirac wave (synth_dirac.DiracWave)."""
        self.dbs_set_grid(w.ks, w.ns, w.ksp, w.ksq, w.t)
        self.dirac_set_orbitals(w.pq, w.ipbs, w.i_ch, w.j_ch, w.v_ch)
        self.dirac_assemble(w.nch, w.npert, w.mk, w.ntarg, w.iptar, *w.concat())

    def dirac_get(self):
        """hch(ms,ms,nblk), hcp(ms,nch*npert), hp, acf(nblk,0:mk), htarg."""
        ns = self.dbs[1]
        nch, npert, mk, ntarg = self.dirac_dims
        ms, nblk = 2 * ns, nch * (nch + 1) // 2
        hch = np.zeros((ms, ms, nblk), order="F")
        hcp = np.zeros((ms, max(nch * npert, 1)), order="F")
        hp = np.zeros(max(npert * (npert + 1) // 2, 1))
        acf = np.zeros((nblk, mk + 1), order="F")
        htarg = np.zeros(ntarg * (ntarg + 1) // 2)
        L.check(L.lib().bsrg_dirac_get(self._p, hch, hcp, hp, acf, htarg))
        return dict(hch=hch, hcp=hcp[:, : nch * npert], hp=hp[: npert * (npert + 1) // 2], acf=acf, htarg=htarg)

    def dirac_engine(self):
        return int(L.lib().bsrg_dirac_engine(self._p))

    def dirac_diff(self, ref, floor_frac=1e-3):
        """Largest relative element error of this context's Dirac wave
        against ref's (same wave), floor floor_frac * max|ref|."""
        r = C.c_double()
        L.check(L.lib().bsrg_dirac_diff(self._p, ref._p, floor_frac, C.byref(r)))
        return r.value

    def dirac_timings(self):
        """ms: device bands+bases, exchange GEMMs, terms with the output
        layout, symmetrisation+vectors, first to last event; host decoding,
        staging+uploads, whole call."""
        out = np.zeros(8)
        L.check(L.lib().bsrg_dirac_timings(self._p, out))
        return out

    # ---- channel-block steps of the Dirac wave (overlap, diag, transform) -----
    def dirac_overlap(self, nch, npert, ntarg, iptar, itype, k1, k2, c, eps_c=1e-10, iitar=0):
        """Overlap pass: fppqq on the channel diagonal blocks, then O_data
        over the overlap records (itype per record)."""
        n = len(c)
        z = np.zeros(1, np.int32)
        L.check(L.lib().bsrg_dirac_overlap(self._p, nch, npert, ntarg, i32(iptar), n,
                                           *(i32(a) if n else z for a in (itype, k1, k2)),
                                           f64(c) if n else np.zeros(1), float(eps_c), int(iitar)))
        self.dirac_sdims = (nch, npert, ntarg)

    def dirac_overlap_wave(self, w, eps_c=1e-10, iitar=0):
        """Overlap pass of a synthetic Dirac wave (its olists)."""
        ls = [L_ for L_ in w.olists if len(L_)]
        it = np.concatenate([np.full(len(L_), L_.itype, np.int32) for L_ in ls]) if ls else np.zeros(0, np.int32)
        cat = [np.concatenate([getattr(L_, a) for L_ in ls]) if ls else np.zeros(0) for a in ("k1", "k2", "c")]
        self.dirac_overlap(w.nch, w.npert, w.ntarg, w.iptar, it, *cat, eps_c=eps_c, iitar=iitar)

    def dirac_overlap_get(self):
        """S: hch(ms,ms,nblk), hcp(ms,nch*npert), hp, otarg."""
        ns = self.dbs[1]
        nch, npert, ntarg = self.dirac_sdims
        ms, nblk = 2 * ns, nch * (nch + 1) // 2
        hch = np.zeros((ms, ms, nblk), order="F")
        hcp = np.zeros((ms, max(nch * npert, 1)), order="F")
        hp = np.zeros(max(npert * (npert + 1) // 2, 1))
        otarg = np.zeros(ntarg * (ntarg + 1) // 2)
        L.check(L.lib().bsrg_dirac_overlap_get(self._p, hch, hcp, hp, otarg))
        return dict(hch=hch, hcp=hcp[:, : nch * npert], hp=hp[: npert * (npert + 1) // 2], otarg=otarg)

    def dirac_update_hx(self, which, d, ich, jch, imat, c):
        """UPDATE_HX records: block (ich,jch) of S (0) or H (1) += c * d[:,:,imat-1]."""
        d = f64(np.asarray(d, np.float64).reshape(d.shape[0], d.shape[1], -1))
        n = len(c)
        z = np.zeros(1, np.int32)
        L.check(L.lib().bsrg_dirac_update_hx(self._p, which, d.shape[2], d, n,
                                             *(i32(a) if n else z for a in (ich, jch, imat)),
                                             f64(c) if n else np.zeros(1)))

    def dirac_energy_shift(self, ecore, etarg):
        L.check(L.lib().bsrg_dirac_energy_shift(self._p, float(ecore), f64(etarg)))

    def dirac_diag_channels(self, lch, kch, ilzero=-1, jlzero=-1, ibzero=1, jbzero=1,
                            edmin=-2 * 137.035999084 ** 2, edmax=1e10, egap=1e-3):
        """Diag_channels after Zero_cond; returns ipsol (nch,)."""
        nch = len(lch)
        ipsol = np.zeros(nch, np.int32)
        ns = C.c_int()
        L.check(L.lib().bsrg_dirac_diag_channels(self._p, i32(lch), i32(kch), ilzero, jlzero, ibzero, jbzero,
                                                 float(edmin), float(edmax), float(egap), ipsol, C.byref(ns)))
        self._ipsol = ipsol
        return ipsol

    def dirac_solutions(self):
        """diag(ms,ms,nch), eval(nsol), iprm(nch*ms)."""
        ms = 2 * self.dbs[1]
        nch = len(self._ipsol)
        diag = np.zeros((ms, ms, nch), order="F")
        ev = np.zeros(max(int(self._ipsol.sum()), 1))
        iprm = np.zeros(nch * ms, np.int32)
        L.check(L.lib().bsrg_dirac_solutions(self._p, diag, ev, iprm))
        return diag, ev[: int(self._ipsol.sum())], iprm

    def dirac_transform(self, which):
        L.check(L.lib().bsrg_dirac_transform(self._p, int(which)))

    def dirac_overlaps(self):
        nch, npert, _ = self.dirac_sdims
        out = np.zeros((nch + npert, nch + npert), order="F")
        L.check(L.lib().bsrg_dirac_overlaps(self._p, out))
        return out

    def dirac_diag_timings(self):
        """Device ms: overlap pass, channel diagonalisation, S and H transforms."""
        out = np.zeros(4)
        L.check(L.lib().bsrg_dirac_diag_timings(self._p, out))
        return out

    def dirac_record(self, path, which, nelc_core=0):
        """Write (which=0: start the file with the solutions, then S) or
        append (which=1: H, mk and the asymptotic coefficients) the
        transformed matrix as DBSR's Record_matrix
        (src/dbsr/DBSR_MAT3_MPI/record_matrix.f90, sub1.f90:154-182,271-284)."""
        L.check(L.lib().bsrg_dirac_record(self._p, os.fsencode(str(path)), int(which), int(nelc_core)))

    def dbs_convol(self, family, k, d, icase):
        ks, ns = self.dbs[0], self.dbs[1]
        shape = (ns, ks) if icase <= 2 else (ns, ns)
        d = f64(d)
        if d.shape != shape:
            raise ValueError(f"density must have shape {shape}")
        a = np.zeros(shape, order="F")
        L.check(L.lib().bsrg_dbs_convol(self._p, family, k, icase, d, a))
        return a

    # ---- one partial wave, call by call --------------------------------------
    def set_orbitals(self, pbs, iech, lbs, ibort, eps_ovl=1e-8):
        pbs = f64(pbs)
        L.check(L.lib().bsrg_set_orbitals(self._p, pbs.shape[1], pbs, i32(iech), i32(lbs), i32(ibort), eps_ovl))
        self._nbf = pbs.shape[1]

    def allocate_matrix(self, nch, ncp, mk):
        L.check(L.lib().bsrg_alloc_matrix(self._p, nch, ncp, mk))
        self.nch, self.ncp, self.mk = nch, ncp, mk

    def update_hl(self, ich, jch, c):
        L.check(L.lib().bsrg_update_hl(self._p, ich, jch, float(c)))

    def i_data(self, jtype, jpol, lst, icase=5, ipt=None):
        n = len(lst.c)
        ipt = np.arange(1, n + 1) if ipt is None else ipt
        L.check(L.lib().bsrg_idata(self._p, icase, jtype, jpol, n, i32(ipt), i32(lst.k1), i32(lst.k2),
                                   i32(lst.k3), i32(lst.k4), f64(lst.c)))

    def o_data(self, lst, eps_c=1e-10, ipt=None):
        n = len(lst.c)
        ipt = np.arange(1, n + 1) if ipt is None else ipt
        L.check(L.lib().bsrg_odata(self._p, n, i32(ipt), i32(lst.k3), i32(lst.k4), f64(lst.c), eps_c))

    def gen_lval(self, z, kclosd):
        """Gen_Lval with INT_core (l_core.f90:92-257): the one-electron
        operator with the closed core (orbitals 1..kclosd) for every l."""
        L.check(L.lib().bsrg_gen_lval(self._p, float(z), int(kclosd)))

    def lval(self):
        """(mlh, hl_full (ns,ns,mlh+1), L_int (nbf,nbf), L_vec (ns,nbf))."""
        mlh = C.c_int()
        L.check(L.lib().bsrg_lval_get(self._p, mlh, None, None, None))
        ns, nbf = self.ns, self._nbf
        hl = np.zeros((ns, ns, mlh.value + 1), order="F")
        li = np.zeros((nbf, nbf), order="F")
        lv = np.zeros((ns, nbf), order="F")
        L.check(L.lib().bsrg_lval_get(self._p, mlh, hl, li, lv))
        return mlh.value, hl, li, lv

    def l_data(self, jtype, l, lst, iitar=0, eps_c=0.0, ipt=None):
        """L_data(jtype, l) (L_data.f90:2-149) on the current matrix."""
        n = len(lst.c)
        ipt = np.arange(1, n + 1) if ipt is None else ipt
        L.check(L.lib().bsrg_ldata(self._p, jtype, l, n, i32(ipt), i32(lst.k1), i32(lst.k2), i32(lst.k3),
                                   i32(lst.k4), f64(lst.c), int(iitar), float(eps_c)))

    def bs_orth(self, which=0):
        """BS_ORTH (bs_orth.f90:2-92): which = 0 the matrix of
        allocate_matrix, 1 the H of the last wave_assemble."""
        L.check(L.lib().bsrg_bs_orth(self._p, int(which)))

    def symmetrize_diag(self):
        L.check(L.lib().bsrg_symmetrize_diag(self._p))

    def scale(self, ec):
        L.check(L.lib().bsrg_scale(self._p, float(ec)))

    def acf_finish(self, kcore=0):
        L.check(L.lib().bsrg_acf_finish(self._p, kcore))

    def _alloc_out(self, nch, ncp, mk):
        from .matrices import WaveMatrices
        return WaveMatrices(self.ns, nch, ncp, mk)

    def matrix_record(self, path, interrupt, intercase, max_nbuf=0):
        """The current matrix as an interrupted run leaves it (BSR_MAT4)."""
        L.check(L.lib().bsrg_matrix_record(self._p, str(path).encode(), int(interrupt), int(intercase),
                                           int(max_nbuf)))

    def matrix_resume(self, path):
        """Read an interrupted run's matrix into the current one: returns
        (interrupt, intercase, max_nbuf)."""
        a, b, m = C.c_int(), C.c_int(), C.c_int()
        L.check(L.lib().bsrg_matrix_resume(self._p, str(path).encode(), C.byref(a), C.byref(b), C.byref(m)))
        return a.value, b.value, m.value

    def get_matrix(self):
        m = self._alloc_out(self.nch, self.ncp, self.mk)
        L.check(L.lib().bsrg_get_matrix(self._p, m.hcc, m.hcb if self.ncp else None, m.hbb if self.ncp else None,
                                        m.acf, m.htarg, m.otarg))
        return m

    # ---- one partial wave, fused ---------------------------------------------
    def wave_assemble(self, w, set_inputs=True):
        """S and H of a PartialWave in one pass (SUB1, sub1.f90:70-219)."""
        if set_inputs:
            self.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
            self.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
        jt, kp, k1, k2, k3, k4, c = w.concat()
        o = w.olist
        L.check(L.lib().bsrg_wave_assemble(
            self._p, w.nch, w.ncp, w.mk, len(c), i32(jt), i32(kp), i32(k1), i32(k2), i32(k3), i32(k4), f64(c),
            len(o.c), i32(o.k3), i32(o.k4), f64(o.c), w.eps_c, w.ec, w.ntarg, f64(w.etarg), i32(w.iptar)))
        self._wave = (w.nch, w.ncp, w.mk)

    def lists_upload(self, w):
        """The R^k lists of a PartialWave into device memory (bsrg_lists_upload)."""
        jt, kp, k1, k2, k3, k4, c = w.concat()
        L.check(L.lib().bsrg_lists_upload(self._p, len(c), i32(jt), i32(kp), i32(k1), i32(k2), i32(k3), i32(k4),
                                          f64(c)))

    def wave_assemble_resident(self, w, set_inputs=True):
        """wave_assemble with the lists of the last lists_upload, decoded on
        the device (the same S and H, bit for bit)."""
        if set_inputs:
            self.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
            self.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
        o = w.olist
        L.check(L.lib().bsrg_wave_assemble_resident(
            self._p, w.nch, w.ncp, w.mk, len(o.c), i32(o.k3), i32(o.k4), f64(o.c), w.eps_c, w.ec, w.ntarg,
            f64(w.etarg), i32(w.iptar)))
        self._wave = (w.nch, w.ncp, w.mk)

    def wave_prepare(self, w, set_inputs=True):
        """Host lists, uploads, bands and bases; inputs stay resident."""
        if set_inputs:
            self.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
            self.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
        jt, kp, k1, k2, k3, k4, c = w.concat()
        o = w.olist
        L.check(L.lib().bsrg_wave_prepare(
            self._p, w.nch, w.ncp, w.mk, len(c), i32(jt), i32(kp), i32(k1), i32(k2), i32(k3), i32(k4), f64(c),
            len(o.c), i32(o.k3), i32(o.k4), f64(o.c), w.eps_c, w.ec, w.ntarg, f64(w.etarg), i32(w.iptar)))
        self._wave = (w.nch, w.ncp, w.mk)

    def wave_launch(self):
        L.check(L.lib().bsrg_wave_launch(self._p))

    def wave_finish(self):
        L.check(L.lib().bsrg_wave_finish(self._p))

    def wave_engine(self):
        return L.lib().bsrg_wave_engine(self._p)

    def set_gemm_engine(self, engine=-1):
        """0 library FP64 GEMM, 1-3 FP64 MFMA kernel tiles, 4 sparse product
        over the nonzero coefficients, -1 automatic (BS waves: autotuned,
        4 preferred where its tile fits; Dirac waves: 4)."""
        L.check(L.lib().bsrg_set_gemm_engine(self._p, engine))

    def set_capture_safe(self, on=True):
        L.check(L.lib().bsrg_set_capture_safe(self._p, 1 if on else 0))

    def wave_get(self, which):
        nch, ncp, mk = self._wave
        m = self._alloc_out(nch, ncp, mk)
        L.check(L.lib().bsrg_wave_get(self._p, which, m.hcc, m.hcb if ncp else None, m.hbb if ncp else None,
                                      m.acf, m.htarg, m.otarg))
        return m

    def wave_write(self, path, w):
        L.check(L.lib().bsrg_wave_write(self._p, str(path).encode(), float(w.t[w.ns]), w.ntarg, f64(w.etarg),
                                        float(w.ec)))

    def wave_timings(self):
        t = np.zeros(5, order="F")
        L.check(L.lib().bsrg_wave_timings(self._p, t))
        return dict(bands=t[0], bases=t[1], gemm=t[2], terms=t[3], total=t[4])

    def wave_host_timings(self):
        t = np.zeros(10, order="F")
        L.check(L.lib().bsrg_wave_host_timings(self._p, t))
        return dict(launch=t[0], overlaps=t[1], registry=t[2], decode=t[3], merge=t[4], sort=t[5],
                    upload=t[6], total=t[7], decode_cpu=t[8], decode_longest=t[9])

    def wave_term_counts(self, which: int = 1) -> dict:
        """Sizes of the staged term sets of S (which=0) or H (which=1)."""
        n = np.zeros(8, np.int64)
        L.check(L.lib().bsrg_wave_term_counts(self._p, which, n))
        keys = ("blocks", "band_blocks", "band_terms", "rank1_blocks", "rank1_terms", "vec_positions", "vec_terms",
                "live_blocks")
        return {k: int(v) for k, v in zip(keys, n)}

    def wave_diff(self, ref: "Context", which: int = 1, floor_frac: float = 1e-3) -> float:
        """Largest relative element error of this context's last wave against
        the same wave assembled by `ref` (S for which=0, H for which=1), with
        the floor floor_frac * max|ref|."""
        out = C.c_double()
        L.check(L.lib().bsrg_wave_diff(self._p, ref._p, which, floor_frac, C.byref(out)))
        return out.value

    def wave_eigen(self, nev: int = 10) -> np.ndarray:
        """The nev lowest eigenvalues of H x = E S x of the last wave (FP64,
        on the device; the closing steps of diag_hd,
        src/bsr/BSR_HD3/diag_hd.f90:271-312, on the full basis)."""
        e = np.zeros(nev)
        L.check(L.lib().bsrg_wave_eigen(self._p, nev, e))
        return e

    def set_eigen_method(self, method: int = 0):
        """0 automatic (block Lanczos from nhm 2048 on), 1 dense (sygvdx), 2
        block Lanczos (falls back to dense when it does not converge)."""
        L.check(L.lib().bsrg_set_eigen_method(self._p, method))

    def eigen_info(self) -> dict:
        """Method the last eigenvalue call used (1 dense, 2 block Lanczos),
        its Krylov dimension and its largest residual bound."""
        m, d, r = C.c_int(), C.c_int(), C.c_double()
        L.check(L.lib().bsrg_eigen_info(self._p, C.byref(m), C.byref(d), C.byref(r)))
        return dict(method=m.value, dim=d.value, resid=r.value)

    def matfile_eigen(self, path, nev: int = 10) -> np.ndarray:
        """The nev lowest eigenvalues of H x = E S x of a bsr_mat.nnn file
        (read as BSR_HD3 reads it, solved on the device)."""
        e = np.zeros(nev)
        L.check(L.lib().bsrg_matfile_eigen(self._p, str(path).encode(), nev, e))
        return e

    def drop_caches(self):
        """Forget resident bands and bases (the next wave rebuilds them)."""
        L.check(L.lib().bsrg_drop_caches(self._p))

    def set_precision(self, mode, nslices=0):
        L.check(L.lib().bsrg_set_precision(self._p, mode, nslices))


def cdata_build(j1, j2, j3, j4, c):
    """Add_cdata (add_cdata.f90:2-56) of every record into one empty block:
    ((k1, k2, k3, k4), c) sorted on the key, equal keys summed."""
    n = len(c)
    k = [np.zeros(n, np.int32) for _ in range(4)]
    co = np.zeros(n)
    nout = C.c_int()
    L.check(L.lib().bsrg_cdata_build(n, i32(j1), i32(j2), i32(j3), i32(j4), f64(c), *k, co, nout))
    m = nout.value
    return [x[:m] for x in k], co[:m]


def cdata_merge(ip, jp, k1, k2, k3, k4, c, eps_c):
    """Merge_CDATA (merge_cdata.f90:2-91) over the sorted blocks
    [ip(m), jp(m)]: (ipt, merged coefficients)."""
    n = len(c)
    cc = f64(c).copy()
    ipt = np.zeros(max(n, 1), np.int32)
    nc = C.c_int()
    L.check(L.lib().bsrg_cdata_merge(len(ip), i32(ip).copy(), i32(jp), n, i32(k1), i32(k2), i32(k3), i32(k4), cc,
                                     float(eps_c), ipt, nc))
    return ipt[: nc.value].copy(), cc

