"""CPU oracle for the bsr_mat radial core -- TEST INFRASTRUCTURE ONLY.

Thin ctypes layer over ``liborc.so`` (the C restatement in ``bsr_oracle.c``)
plus the partial-wave driver that strings the reference steps together in the
order of ``src/bsr/BSR_MAT3/sub1.f90`` and the ``bsr_mat.nnn`` writer/reader
of ``src/bsr/BSR_MAT3/record_matrix.f90``.  Only ``tests/``, the smoke check
and the CPU-baseline leg of ``bench.py`` use this module, and only as the
checker; the product package never imports it.

Parity status of each function is recorded in the test files that pin it
(``tests/test_oracle_*.py``).
"""
from __future__ import annotations

import ctypes as C
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
IBI = 2 ** 15
IBO = 2 ** 15

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


def build() -> str:
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile")], check=True)
    return os.path.join(HERE, "liborc.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        sig = {
            "orc_last_error": (C.c_char_p, []),
            "orc_mkgrid1": (i, [i, d, d, d, d, i, _dp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
            "orc_mkgrid0": (i, [i, d, d, d, d, i, _dp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
            "orc_gauss": (None, [i, _dp, _dp]),
            "orc_grid_new": (vp, [i, i, i, i, d, _dp]),
            "orc_grid_free": (None, [vp]),
            "orc_grid_nv": (i, [vp]),
            "orc_grid_tables": (None, [vp, _dp, _dp, _dp, _dp]),
            "orc_bspline_at": (None, [vp, d, _dp]),
            "orc_mrm": (None, [vp, i, _dp]),
            "orc_moments": (None, [vp, i, _dp]),
            "orc_rk_moments": (None, [vp, i, _dp, _dp, _dp]),
            "orc_mrk_cell": (None, [vp, i, _dp]),
            "orc_density": (None, [i, i, _dp, _dp, _dp, C.c_char]),
            "orc_convol": (i, [i, i, _dp, _dp, i, C.c_char, C.c_char, _dp]),
            "orc_sum_amb": (d, [i, i, _dp, _dp, C.c_char]),
            "orc_bav": (None, [i, i, _dp, _dp, _dp, C.c_char, C.c_char]),
            "orc_bxv": (None, [i, i, _dp, _dp, _dp]),
            "orc_bvmv": (d, [i, i, _dp, _dp, _dp]),
            "orc_quadr": (d, [vp, _dp, _dp, i]),
            "orc_wave_new": (vp, [vp, i, _dp, _ip, _ip, _ip, i, i, i, d, d]),
            "orc_wave_free": (None, [vp]),
            "orc_wave_get": (None, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_wave_qbs": (None, [vp, _dp, _dp]),
            "orc_wave_update_hl": (i, [vp, i, i, d]),
            "orc_wave_idata": (i, [vp, i, i, i, i, _ip, _ip, _ip, _ip, _ip, _dp]),
            "orc_wave_odata": (i, [vp, i, _ip, _ip, _ip, _dp]),
            "orc_wave_symmetrize_diag": (None, [vp]),
            "orc_wave_scale": (None, [vp, d]),
            "orc_wave_acf_finish": (None, [vp, i]),
            "orc_wave_update_hx": (i, [vp, i, i, _dp, C.c_char]),
            "orc_wave_update_hw": (i, [vp, i, i, _dp, _dp]),
            "orc_wave_update_hv": (i, [vp, i, i, _dp, d]),
            "orc_wave_update_hb": (i, [vp, i, i, d]),
            "orc_wave_update_cf": (i, [vp, i, i, i, d]),
            "orc_grid_bspd": (None, [vp, _dp, _dp]),
            "orc_db2": (None, [vp, _dp]),
            "orc_zcb": (d, [i, i, i]),
            "orc_wave_gen_lval": (i, [vp, d, i]),
            "orc_wave_lval_mlh": (i, [vp]),
            "orc_wave_lval_get": (None, [vp, _dp, _dp, _dp]),
            "orc_wave_ldata": (i, [vp, i, i, i, _ip, _ip, _ip, _ip, _ip, _dp, i, d]),
            "orc_wave_bs_orth": (i, [vp]),
            "orc_gauleg": (None, [d, d, _dp, _dp, i]),
            "orc_dbs_new": (vp, [i, i, i, i, _dp]),
            "orc_dbs_free": (None, [vp]),
            "orc_dbs_tables": (None, [vp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_dbs_moments": (None, [vp, i, i, _dp]),
            "orc_dbs_triang": (None, [vp, i, i, i, _dp]),
            "orc_dbs_mrk": (None, [vp, i, i, _dp]),
            "orc_dbsr_new": (vp, [vp, i, _dp, _ip, i, _ip, _ip, _dp, i, i, i, i, _ip]),
            "orc_dbsr_free": (None, [vp]),
            "orc_dbsr_sizes": (None, [vp, _ip]),
            "orc_dbsr_get": (None, [vp, _dp, _dp, _dp, _dp, _dp]),
            "orc_dbsr_rdata": (i, [vp, i, i, i, _ip, _ip, _ip, _ip, _dp]),
            "orc_dbsr_quadr_pq": (d, [vp, i, i, i]),
            "orc_dbsr_last_error": (C.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _f(a, dtype=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


class OracleError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise OracleError(lib().orc_last_error().decode())


# ---------------------------------------------------------------- grid

def mkgrid1(ks, z, h, hmax, rmax):
    """Knots of mkgrid1 (define_grid.f90:213-276): (t, ns, ml, me, h)."""
    cap = int((rmax * z / h + 1 + ks) * 2) + 2 * ks + 8
    t = np.zeros(cap, order="F")
    ns, ml, me, hh = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    _check(lib().orc_mkgrid1(ks, z, h, hmax, rmax, cap, t, ns, ml, me, hh))
    return t[: ns.value + ks].copy(), ns.value, ml.value, me.value, hh.value


def mkgrid0(ks, z, h, hmax, rmax):
    """Knots of mkgrid (define_grid.f90:86-153): (t, ns, ml, me, h)."""
    cap = 100000
    t = np.zeros(cap, order="F")
    ns, ml, me, hh = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    _check(lib().orc_mkgrid0(ks, z, h, hmax, rmax, cap, t, ns, ml, me, hh))
    return t[: ns.value + ks].copy(), ns.value, ml.value, me.value, hh.value


def zcb(k1, k2, k3):
    """(3j symbol (k1 k2 k3; 0 0 0))^2, the reference's ZCB."""
    return lib().orc_zcb(k1, k2, k3)


def gauss(k):
    x = np.zeros(k, order="F")
    w = np.zeros(k, order="F")
    lib().orc_gauss(k, x, w)
    return x, w


FAMILIES = ("pppp", "qqqq", "pqpq", "qpqp")


def gauleg(n, x1=0.0, x2=1.0):
    x, w = np.zeros(n), np.zeros(n)
    lib().orc_gauleg(x1, x2, x, w, n)
    return x, w


class DbsGrid:
    """Dirac B-spline grid: knots t of order ks, P functions of order ksp
    and Q functions of order ksq on the same intervals (DBS_gauss.f90)."""

    def __init__(self, ks, ns, ksp, ksq, t):
        self.ks, self.ns, self.ksp, self.ksq = ks, ns, ksp, ksq
        self.nv = ns - ks + 1
        self.nsp, self.nsq = self.nv + ksp - 1, self.nv + ksq - 1
        self.t = np.asfortranarray(t, dtype=np.float64)
        self._p = lib().orc_dbs_new(ks, ns, ksp, ksq, self.t)

    def __del__(self):
        try:
            lib().orc_dbs_free(self._p)
        except Exception:
            pass

    def tables(self):
        nv, ks = self.nv, self.ks
        gr, grw = np.zeros((nv, ks), order="F"), np.zeros((nv, ks), order="F")
        pb = np.zeros((nv + 1, ks, self.ksp), order="F")
        qb = np.zeros((nv + 1, ks, self.ksq), order="F")
        tp, tq = np.zeros(self.nsp + self.ksp), np.zeros(self.nsq + self.ksq)
        lib().orc_dbs_tables(self._p, gr, grw, pb, qb, tp, tq)
        return dict(gr=gr, grw=grw, pbsp=pb, qbsp=qb, tp=tp, tq=tq)

    def moments(self, q, k):
        """rkm(ks*ks, nv) of basis q (0 P, 1 Q) for r^k (moment_pp/qq)."""
        out = np.zeros((self.ks * self.ks, self.nv), order="F")
        lib().orc_dbs_moments(self._p, q, k, out)
        return out

    def triang(self, family, k, iv):
        kk = self.ks * self.ks
        out = np.zeros((kk, kk), order="F")
        lib().orc_dbs_triang(self._p, family, k, iv, out)
        return out

    def mrk(self, family, k):
        """rkb(ns,ns,ks,ks) of family 0 pppp, 1 qqqq, 2 pqpq, 3 qpqp."""
        ns, ks = self.ns, self.ks
        out = np.zeros((ns, ns, ks, ks), order="F")
        lib().orc_dbs_mrk(self._p, family, k, out)
        return out


def _i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32).reshape(-1))


class DbsrWave:
    """Dirac R^k assembly (R_data over the four families) on one grid:
    orbitals pq(ns,2,nbf), channel pointers ipbs, <.|p> vectors V_ch(ms,nv_ch)
    keyed by sorted (i_ch, j_ch), target pointers iptar.  Accumulates
    hch(ms,ms,iicc), hcp(ms,iicb), hp, acf(iicc,0:mk), htarg in the
    reference's numbering (dbsr_oracle.c)."""

    def __init__(self, grid: DbsGrid, pq, ipbs, i_ch, j_ch, v_ch, nch, npert, mk, ntarg, iptar):
        self.grid, self.mk, self.ntarg = grid, mk, ntarg
        pq = np.asfortranarray(pq, dtype=np.float64)
        v = np.asfortranarray(np.asarray(v_ch, dtype=np.float64).reshape(2 * grid.ns, -1))
        if v.size == 0:
            v = np.zeros((2 * grid.ns, 1), order="F")
        self._p = lib().orc_dbsr_new(grid._p, pq.shape[2], pq, _i32(ipbs), len(i_ch), _i32(list(i_ch) or [0]),
                                     _i32(list(j_ch) or [0]), v, nch, npert, mk, ntarg, _i32(iptar))
        sz = np.zeros(4, dtype=np.int32)
        lib().orc_dbsr_sizes(self._p, sz)
        self.iicc, self.iicb, self.iibb, self.ms = (int(x) for x in sz)

    def __del__(self):
        try:
            lib().orc_dbsr_free(self._p)
        except Exception:
            pass

    def rdata(self, itype, kpol, k1, k2, k3, k4, c):
        rc = lib().orc_dbsr_rdata(self._p, itype, kpol, len(c), _i32(k1), _i32(k2), _i32(k3), _i32(k4),
                                  np.asfortranarray(c, dtype=np.float64))
        if rc:
            raise RuntimeError(f"orc_dbsr_rdata {rc}: {lib().orc_dbsr_last_error().decode()}")

    def quadr_pq(self, io, jo, m):
        return lib().orc_dbsr_quadr_pq(self._p, io, jo, m)

    def get(self):
        ms = self.ms
        hch = np.zeros((ms, ms, max(self.iicc, 1)), order="F")
        hcp = np.zeros((ms, max(self.iicb, 1)), order="F")
        hp = np.zeros(max(self.iibb, 1), order="F")
        acf = np.zeros((max(self.iicc, 1), self.mk + 1), order="F")
        htarg = np.zeros(max(self.ntarg * (self.ntarg + 1) // 2, 1), order="F")
        lib().orc_dbsr_get(self._p, hch.reshape(-1, order="F"), hcp.reshape(-1, order="F"), hp,
                           acf.reshape(-1, order="F"), htarg)
        return dict(hch=hch[:, :, : self.iicc], hcp=hcp[:, : self.iicb], hp=hp[: self.iibb],
                    acf=acf[: self.iicc], htarg=htarg[: self.ntarg * (self.ntarg + 1) // 2])


def dirac_assemble(w):
    """Oracle run of a synthetic Dirac wave (bsr_amd.synth.DiracWave): every
    (itype, kpol) list through R_data for the four families, in list order."""
    g = DbsGrid(w.ks, w.ns, w.ksp, w.ksq, w.t)
    o = DbsrWave(g, w.pq, w.ipbs, w.i_ch, w.j_ch, w.v_ch, w.nch, w.npert, w.mk, w.ntarg, w.iptar)
    for L in w.lists:
        o.rdata(L.itype, L.kpol, L.k1, L.k2, L.k3, L.k4, L.c)
    out = o.get()
    out["_grid"] = g
    return out


class Grid:
    """Gauss tables, B-spline values and the Gram band on one knot grid."""

    def __init__(self, ks, ns, ml, me, h, t):
        self.ks, self.ns, self.ml, self.me, self.h = ks, ns, ml, me, h
        self.t = _f(t)
        self.nv = ns - ks + 1
        self.ptr = lib().orc_grid_new(ks, ns, ml, me, h, self.t)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_grid_free(self.ptr)
            self.ptr = None

    @property
    def ik(self):
        return self.ks * (self.ks + 1) // 2

    def tables(self):
        ks, nv, ns = self.ks, self.nv, self.ns
        gr = np.zeros(nv * ks, order="F")
        grw = np.zeros(nv * ks, order="F")
        bsp = np.zeros(nv * ks * ks, order="F")
        sb = np.zeros((ns, ks), order="F")
        lib().orc_grid_tables(self.ptr, gr, grw, bsp, sb)
        return gr.reshape(nv, ks), grw.reshape(nv, ks), bsp.reshape(nv, ks, ks), sb

    def bsplines_at(self, r):
        out = np.zeros(self.ns, order="F")
        lib().orc_bspline_at(self.ptr, float(r), out)
        return out

    def mrm(self, m):
        rm = np.zeros((self.ns, self.ks), order="F")
        lib().orc_mrm(self.ptr, m, rm)
        return rm

    def bspd(self):
        """First and second derivatives of the B-splines at the Gauss
        points, [iv, m, ith] like bsp."""
        ks, nv = self.ks, self.nv
        d1 = np.zeros(nv * ks * ks)
        d2 = np.zeros(nv * ks * ks)
        lib().orc_grid_bspd(self.ptr, d1, d2)
        return d1.reshape(nv, ks, ks), d2.reshape(nv, ks, ks)

    def db2(self):
        """<B_i|B''_j> in lower band storage (mdb2), slot (1,1) holding
        <B_{ns-1}|B''_ns>."""
        out = np.zeros((self.ns, self.ks), order="F")
        lib().orc_db2(self.ptr, out)
        return out

    def moments(self, k):
        rkm = np.zeros(self.ik * self.nv, order="F")
        lib().orc_moments(self.ptr, k, rkm)
        return rkm.reshape(self.nv, self.ik)

    def rk_moments(self, k):
        ik, nv = self.ik, self.nv
        r1 = np.zeros(ik * nv, order="F")
        r2 = np.zeros(ik * nv, order="F")
        rd = np.zeros(ik * ik * nv, order="F")
        lib().orc_rk_moments(self.ptr, k, r1, r2, rd)
        return r1.reshape(nv, ik), r2.reshape(nv, ik), rd.reshape(nv, ik, ik)

    def mrk_cell(self, k):
        """rkb(ns,ns,ks,ks), Fortran order."""
        rkb = np.zeros((self.ns, self.ns, self.ks, self.ks), order="F")
        lib().orc_mrk_cell(self.ptr, k, rkb)
        return rkb

    def quadr(self, p1, p2, m):
        return lib().orc_quadr(self.ptr, _f(p1), _f(p2), m)


# ---------------------------------------------------------------- leaves

def density(ns, ks, p1, p2, typ):
    cols = {"s": ks, "n": 2 * ks - 1, "x": ns}[typ]
    d = np.zeros((ns, cols), order="F")
    lib().orc_density(ns, ks, d, _f(p1), _f(p2), typ.encode())
    return d


def convol(ns, ks, d, icase, rkb):
    a = np.zeros((ns, ns), order="F")
    dd = np.zeros((ns, ns), order="F")
    dd[:, : d.shape[1]] = d
    _check(lib().orc_convol(ns, ks, a, dd, icase, b"s", b"s", _f(rkb)))
    return a if icase > 2 else a[:, :ks].copy(order="F")


def sum_amb(ns, ks, a, b, sym):
    return lib().orc_sum_amb(ns, ks, _f(a), _f(b), sym.encode())


def bav(ns, ks, b, v, typ="s", side="r"):
    y = np.zeros(ns, order="F")
    lib().orc_bav(ns, ks, _f(b), _f(v), y, typ.encode(), side.encode())
    return y


def bxv(ks, ns, b, v):
    y = np.zeros(ns, order="F")
    lib().orc_bxv(ks, ns, _f(b), _f(v), y)
    return y


def bvmv(ns, ks, a, x, y):
    return lib().orc_bvmv(ns, ks, _f(a), _f(x), _f(y))


# ---------------------------------------------------------------- a wave

class WaveMatrices:
    """hcc(ns,ns,nch(nch+1)/2), hcb(ns,nch,ncp), hbb, acf(nch,nch,0:mk),
    htarg, otarg -- the contents of module bsr_matrix."""

    def __init__(self, ns, nch, ncp, mk):
        nij = nch * (nch + 1) // 2
        self.hcc = np.zeros((ns, ns, nij), order="F")
        self.hcb = np.zeros((ns, nch, max(ncp, 1)), order="F")
        self.hbb = np.zeros(max(ncp * (ncp + 1) // 2, 1), order="F")
        self.acf = np.zeros((nch, nch, mk + 1), order="F")
        self.htarg = np.zeros(nij, order="F")
        self.otarg = np.zeros(nij, order="F")
        self.ns, self.nch, self.ncp, self.mk = ns, nch, ncp, mk


class Wave:
    """State of one partial wave in the oracle (modules bsr_matrix,
    spline_orbitals and cmdata of the reference)."""

    def __init__(self, grid: Grid, job):
        self.grid = grid
        self.job = job
        self.ptr = lib().orc_wave_new(
            grid.ptr, job.nbf, _f(job.pbs), _f(job.iech, np.int32), _f(job.lbs, np.int32),
            _f(job.ibort, np.int32), job.nch, job.ncp, job.mk, job.eps_ovl, job.eps_c)

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_wave_free(self.ptr)
            self.ptr = None

    def get(self) -> WaveMatrices:
        j = self.job
        m = WaveMatrices(j.ns, j.nch, j.ncp, j.mk)
        lib().orc_wave_get(self.ptr, m.hcc, m.hcb, m.hbb, m.acf, m.htarg, m.otarg)
        return m

    def update_hl(self, ich, jch, c):
        _check(lib().orc_wave_update_hl(self.ptr, ich, jch, c))

    def update_hx(self, ich, jch, d, sym):
        dd = np.zeros((self.job.ns, self.job.ns), order="F")
        dd[:, : d.shape[1]] = d
        _check(lib().orc_wave_update_hx(self.ptr, ich, jch, dd, sym.encode()))

    def update_hw(self, ich, jch, v, u):
        _check(lib().orc_wave_update_hw(self.ptr, ich, jch, _f(v), _f(u)))

    def update_hv(self, ich, ic, v, c):
        _check(lib().orc_wave_update_hv(self.ptr, ich, ic, _f(v), c))

    def update_hb(self, ic, jc, c):
        _check(lib().orc_wave_update_hb(self.ptr, ic, jc, c))

    def update_cf(self, k, ich, jch, c):
        _check(lib().orc_wave_update_cf(self.ptr, k, ich, jch, c))

    def idata(self, jtype, k, lst, icase=5):
        n = len(lst.c)
        ipt = np.arange(1, n + 1, dtype=np.int32)
        _check(lib().orc_wave_idata(self.ptr, icase, jtype, k, n, ipt, _f(lst.k1, np.int32),
                                    _f(lst.k2, np.int32), _f(lst.k3, np.int32),
                                    _f(lst.k4, np.int32), _f(lst.c)))

    def odata(self, lst):
        n = len(lst.c)
        ipt = np.arange(1, n + 1, dtype=np.int32)
        _check(lib().orc_wave_odata(self.ptr, n, ipt, _f(lst.k3, np.int32), _f(lst.k4, np.int32),
                                    _f(lst.c)))

    def gen_lval(self, z, kclosd):
        """Gen_Lval / INT_core (l_core.f90:92-257), non-relativistic."""
        _check(lib().orc_wave_gen_lval(self.ptr, float(z), int(kclosd)))

    def lval(self):
        """(mlh, hl_full (ns,ns,mlh+1), L_int (nbf,nbf), L_vec (ns,nbf))."""
        j = self.job
        mlh = lib().orc_wave_lval_mlh(self.ptr)
        if mlh < 0:
            raise OracleError("Gen_Lval has not run")
        hl = np.zeros((j.ns, j.ns, mlh + 1), order="F")
        li = np.zeros((j.nbf, j.nbf), order="F")
        lv = np.zeros((j.ns, j.nbf), order="F")
        lib().orc_wave_lval_get(self.ptr, hl, li, lv)
        return mlh, hl, li, lv

    def ldata(self, jtype, l, lst, iitar=0, eps_c=0.0):
        """L_data(jtype, l) (L_data.f90:2-149)."""
        n = len(lst.c)
        ipt = np.arange(1, n + 1, dtype=np.int32)
        _check(lib().orc_wave_ldata(self.ptr, jtype, l, n, ipt, _f(lst.k1, np.int32), _f(lst.k2, np.int32),
                                    _f(lst.k3, np.int32), _f(lst.k4, np.int32), _f(lst.c), int(iitar),
                                    float(eps_c)))

    def bs_orth(self):
        """BS_ORTH (bs_orth.f90:2-92) on the accumulated matrix."""
        _check(lib().orc_wave_bs_orth(self.ptr))

    def symmetrize_diag(self):
        lib().orc_wave_symmetrize_diag(self.ptr)

    def scale(self, ec):
        lib().orc_wave_scale(self.ptr, ec)

    def acf_finish(self, kcore=0):
        lib().orc_wave_acf_finish(self.ptr, kcore)


def core_acf(lbs, kclosd: int) -> int:
    """Core-screening term added to ACF(i,i,0): sum over the closed shells
    1..kclosd of 2(4l+2) (sub1.f90:218-219)."""
    return int(sum(2 * (4 * int(lbs[i]) + 2) for i in range(kclosd)))


def run_wave(job, grid: Grid | None = None, orth: bool = False, l_lists: dict | None = None, z: float = 0.0,
             kclosd: int = 0, iitar: int = 0):
    """One partial wave in the order of SUB1 (src/bsr/BSR_MAT3/sub1.f90:70-219):
    S = sb on the channel diagonals + O_data, symmetrised; H = EC*S + R^k
    contributions (jpol outer, jtype inner, state_res.f90:155-159) +
    (E_t - EC)*sb, symmetrised (then BS_ORTH when orth, sub1.f90:195); ACF
    finished with the core term of the closed shells 1..kclosd.  With l_lists, Gen_Lval (nuclear charge z, closed core
    1..kclosd) and the L-integral lists (keyed (jtype, l)) come between the
    EC scaling and the R^k lists (sub1.f90:121-129).  Returns (S, H) as
    WaveMatrices (S.htarg/acf unused)."""
    if grid is None:
        grid = Grid(job.ks, job.ns, job.ml, job.me, job.h, job.t)
    w = Wave(grid, job)
    for ich in range(1, job.nch + 1):
        w.update_hl(ich, ich, 1.0)
    if job.olist is not None and len(job.olist.c):
        w.odata(job.olist)
    w.symmetrize_diag()
    S = w.get()
    w.scale(job.ec)
    if l_lists:
        w.gen_lval(z, kclosd)
        for (jt, l), lst in sorted(l_lists.items(), key=lambda kv: (kv[0][1], kv[0][0])):  # l outer, jtype inner
            if len(lst):
                w.ldata(jt, l, lst, iitar=iitar, eps_c=job.eps_c)
    for k in range(0, job.mk + 1):
        for jtype in range(1, 10):
            lst = job.rk_lists.get((jtype, k))
            if lst is not None and len(lst.c):
                w.idata(jtype, k, lst)
    for ich in range(1, job.nch + 1):
        w.update_hl(ich, ich, float(job.etarg[job.iptar[ich - 1] - 1] - job.ec))
    w.symmetrize_diag()
    if orth:
        w.bs_orth()
    w.acf_finish(core_acf(job.lbs, kclosd))
    H = w.get()
    return S, H


# ---------------------------------------------------------------- file format

def _rec(f, payload: bytes):
    n = struct.pack("<i", len(payload))
    f.write(n + payload + n)


def _record_matrix(f, m: WaveMatrices, hcc, hcb, hbb):
    """Record_matrix, src/bsr/BSR_MAT3/record_matrix.f90:2-58."""
    kch, kcp = m.nch, m.ncp
    for ib in range(1, kch + 1):
        _rec(f, np.asfortranarray(hcc[:, :, ib * (ib + 1) // 2 - 1]).tobytes(order="F"))
    idiag = 1
    for ich in range(1, kch + 1):
        for jch in range(1, ich):
            ib = ich * (ich - 1) // 2 + jch - 1
            blk = hcc[:, :, ib]
            if np.sum(np.abs(blk)) == 0.0:
                continue
            _rec(f, struct.pack("<2i", ich, jch))
            _rec(f, np.asfortranarray(blk).tobytes(order="F"))
            idiag = 0
    if kcp > 0:
        for ich in range(1, kch + 1):
            for ip in range(1, kcp + 1):
                v = hcb[:, ich - 1, ip - 1]
                if np.sum(np.abs(v)) == 0.0:
                    continue
                _rec(f, struct.pack("<2i", ip + kch, ich))
                _rec(f, v.tobytes())
                idiag = 0
        for ip in range(1, kcp + 1):
            for jp in range(1, ip + 1):
                ib = ip * (ip - 1) // 2 + jp - 1
                if hbb[ib] == 0.0:
                    continue
                _rec(f, struct.pack("<2i", ip + kch, jp + kch))
                _rec(f, struct.pack("<d", hbb[ib]))
                if ip != jp:
                    idiag = 0
    _rec(f, struct.pack("<2i", 0, idiag))


def write_bsr_mat(path, job, S: WaveMatrices, H: WaveMatrices):
    """bsr_mat.nnn as SUB1 writes it (sub1.f90:107-110, 202, 258, 266-270)."""
    with open(path, "wb") as f:
        _rec(f, struct.pack("<3i", job.ns, job.nch, job.ncp))
        _record_matrix(f, S, S.hcc, S.hcb, S.hbb)
        _record_matrix(f, H, H.hcc, H.hcb, H.hbb)
        _rec(f, struct.pack("<i", job.mk))
        _rec(f, np.asfortranarray(H.acf).tobytes(order="F"))
        _rec(f, struct.pack("<di", float(job.t[job.ns]), job.ns))
        _rec(f, H.htarg.tobytes())
        _rec(f, H.otarg.tobytes())
        _rec(f, np.asarray(job.etarg, np.float64).tobytes())
        _rec(f, struct.pack("<d", job.ec))


def read_records(path):
    """All records of a Fortran unformatted sequential file (4-byte markers)."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        out.append(data[pos + 4: pos + 4 + n])
        (m,) = struct.unpack_from("<i", data, pos + 4 + n)
        if m != n:
            raise ValueError("corrupt record markers")
        pos += 8 + n
    return out


def read_bsr_mat(path):
    """Parse bsr_mat.nnn the way BSR_HD reads it (src/bsr/BSR_HD3/diag_hd.f90:37-64):
    returns dict with ns, nch, ncp, S and H (each: dict of diagonal blocks,
    off-diagonal blocks keyed (i,j), idiag) and the trailing records."""
    recs = read_records(path)
    ns, nch, ncp = struct.unpack("<3i", recs[0])
    pos = 1

    def one():
        nonlocal pos
        m = {"diag": [], "blocks": {}, "vec": {}, "scal": {}}
        for _ in range(nch):
            m["diag"].append(np.frombuffer(recs[pos], np.float64).reshape(ns, ns, order="F"))
            pos += 1
        while True:
            i, j = struct.unpack("<2i", recs[pos])
            pos += 1
            if i == 0:
                m["idiag"] = j
                return m
            body = recs[pos]
            pos += 1
            if i <= nch:
                m["blocks"][(i, j)] = np.frombuffer(body, np.float64).reshape(ns, ns, order="F")
            elif j <= nch:
                m["vec"][(i, j)] = np.frombuffer(body, np.float64).copy()
            else:
                m["scal"][(i, j)] = struct.unpack("<d", body)[0]

    S = one()
    H = one()
    (mk,) = struct.unpack("<i", recs[pos]); pos += 1
    acf = np.frombuffer(recs[pos], np.float64).reshape(nch, nch, mk + 1, order="F"); pos += 1
    tmax, ns2 = struct.unpack("<di", recs[pos]); pos += 1
    htarg = np.frombuffer(recs[pos], np.float64); pos += 1
    otarg = np.frombuffer(recs[pos], np.float64); pos += 1
    etarg = np.frombuffer(recs[pos], np.float64); pos += 1
    (ec,) = struct.unpack("<d", recs[pos]); pos += 1
    return dict(ns=ns, nch=nch, ncp=ncp, S=S, H=H, mk=mk, acf=acf, tmax=tmax,
                ns2=ns2, htarg=htarg, otarg=otarg, etarg=etarg, ec=ec)


def dense_from_wave(m: WaveMatrices):
    """Full symmetric nhm x nhm matrix (nhm = nch*ns + ncp) from the block
    arrays: channel blocks, channel-perturber vectors, perturber scalars."""
    ns, nch, ncp = m.ns, m.nch, m.ncp
    n = nch * ns + ncp
    A = np.zeros((n, n))
    for i in range(1, nch + 1):
        for j in range(1, i + 1):
            b = m.hcc[:, :, i * (i - 1) // 2 + j - 1]
            A[(i - 1) * ns: i * ns, (j - 1) * ns: j * ns] = b
            if i != j:
                A[(j - 1) * ns: j * ns, (i - 1) * ns: i * ns] = b.T
    for ip in range(ncp):
        for ich in range(nch):
            v = m.hcb[:, ich, ip]
            A[nch * ns + ip, ich * ns:(ich + 1) * ns] = v
            A[ich * ns:(ich + 1) * ns, nch * ns + ip] = v
        for jp in range(ip + 1):
            c = m.hbb[(ip + 1) * ip // 2 + jp]
            A[nch * ns + ip, nch * ns + jp] = c
            A[nch * ns + jp, nch * ns + ip] = c
    return A
