"""ctypes access to oracle/_ref/libbsr_ref.so -- TEST INFRASTRUCTURE ONLY.

The library is the reference's own Fortran (compiled by oracle/ref.mk from
the reference checkout) driven through oracle/ref_driver.f90.  It exists
only where it was built; ``available()`` says whether it can be loaded.
Module state lives inside the library, so one grid is active at a time.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PATH = os.path.join(HERE, "_ref", "libbsr_ref.so")
_L = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")


def build() -> bool:
    """Build from the reference checkout when it is present."""
    if not os.path.isdir("/root/reference/src"):
        return os.path.exists(PATH)
    subprocess.run(["make", "-s", "-f", os.path.join(HERE, "ref.mk")], check=True)
    return True


def available() -> bool:
    return os.path.exists(PATH)


def lib():
    global _L
    if _L is None:
        L = C.CDLL(PATH)
        i, d, ch = C.c_int, C.c_double, C.c_char
        sig = {
            "ref_mkgrid": (None, [i, i, d, d, d, d, i, _dp, C.POINTER(i), C.POINTER(i), C.POINTER(i), C.POINTER(d)]),
            "ref_gauss": (None, [i, _dp, _dp]),
            "ref_init_grid": (None, [i, i, i, i, d, _dp]),
            "ref_tables": (None, [_dp, _dp, _dp, _dp]),
            "ref_moments": (None, [i, _dp]),
            "ref_rk_moments": (None, [i, _dp, _dp, _dp]),
            "ref_mrk_cell": (None, [i, _dp]),
            "ref_density": (None, [i, i, _dp, _dp, _dp, ch]),
            "ref_convol": (None, [i, i, i, _dp, _dp, i]),
            "ref_sum_amb": (d, [i, i, _dp, _dp, ch]),
            "ref_bav": (None, [i, i, _dp, _dp, _dp, ch, ch]),
            "ref_bxv": (None, [i, i, _dp, _dp, _dp]),
            "ref_bvmv": (d, [i, i, _dp, _dp, _dp]),
            "ref_set_orbitals": (None, [i, _dp]),
            "ref_quadr": (d, [i, i, i]),
            "ref_alloc_matrix": (None, [i, i, i]),
            "ref_update_hx": (None, [i, i, _dp, ch]),
            "ref_update_hl": (None, [i, i, d]),
            "ref_update_hw": (None, [i, i, _dp, _dp]),
            "ref_update_hv": (None, [i, i, _dp, d]),
            "ref_update_hb": (None, [i, i, d]),
            "ref_update_cf": (None, [i, i, i, d]),
            "ref_get_matrix": (None, [_dp, _dp, _dp, _dp]),
            "ref_bspd": (None, [_dp, _dp]),
            "ref_add_cdata": (None, [i, _ip, _ip, _ip, _ip, _dp, _ip, _ip, _ip, _ip, _dp, C.POINTER(i)]),
            "ref_merge_cdata": (None, [i, _ip, _ip, i, _ip, _ip, _ip, _ip, _dp, d, _ip, C.POINTER(i)]),
            "ref_galerkin": (None, [_dp, _dp]),
            "ref_set_lcore": (None, [d, i, i, _ip, _ip, _dp]),
            "ref_gen_lval": (None, [C.POINTER(i), _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _L = L
    return _L


def _f(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def mkgrid(kind, ks, z, h, hmax, rmax):
    cap = 100000
    t = np.zeros(cap, order="F")
    ns, ml, me, hh = C.c_int(), C.c_int(), C.c_int(), C.c_double()
    lib().ref_mkgrid(kind, ks, z, h, hmax, rmax, cap, t, ns, ml, me, hh)
    return t[: ns.value + ks].copy(), ns.value, ml.value, me.value, hh.value


def gauss(k):
    x = np.zeros(k, order="F")
    w = np.zeros(k, order="F")
    lib().ref_gauss(k, x, w)
    return x, w


class RefGrid:
    def __init__(self, ks, ns, ml, me, h, t):
        self.ks, self.ns, self.nv = ks, ns, ns - ks + 1
        lib().ref_init_grid(ks, ns, ml, me, h, _f(t))

    @property
    def ik(self):
        return self.ks * (self.ks + 1) // 2

    def tables(self):
        ks, nv, ns = self.ks, self.nv, self.ns
        gr = np.zeros((nv, ks), order="F")
        grw = np.zeros((nv, ks), order="F")
        bsp = np.zeros((nv, ks, ks), order="F")
        sb = np.zeros((ns, ks), order="F")
        lib().ref_tables(gr, grw, bsp, sb)
        return gr, grw, bsp, sb

    def moments(self, k):
        r = np.zeros((self.ks * self.ks, self.nv), order="F")
        lib().ref_moments(k, r)
        return r[: self.ik, :].T.copy()

    def rk_moments(self, k):
        ik, nv = self.ik, self.nv
        r1 = np.zeros((ik, nv), order="F")
        r2 = np.zeros((ik, nv), order="F")
        rd = np.zeros((ik, ik, nv), order="F")
        lib().ref_rk_moments(k, r1, r2, rd)
        # to the oracle's [iv][jj][ii] layout
        return r1.T.copy(), r2.T.copy(), np.transpose(rd, (2, 1, 0)).copy()

    def mrk_cell(self, k):
        rkb = np.zeros((self.ns, self.ns, self.ks, self.ks), order="F")
        lib().ref_mrk_cell(k, rkb)
        return rkb

    def convol(self, k, d, icase):
        ns, ks = self.ns, self.ks
        a = np.zeros((ns, ns), order="F")
        dd = np.zeros((ns, ns), order="F")
        dd[:, : d.shape[1]] = d
        lib().ref_convol(k, ns, ks, a, dd, icase)
        return a if icase > 2 else a[:, :ks].copy(order="F")

    def set_orbitals(self, pbs):
        lib().ref_set_orbitals(pbs.shape[1], _f(pbs))

    def quadr(self, i, j, m):
        return lib().ref_quadr(i, j, m)

    def bspd(self):
        ks, nv = self.ks, self.nv
        d1 = np.zeros((nv, ks, ks), order="F")
        d2 = np.zeros((nv, ks, ks), order="F")
        lib().ref_bspd(d1, d2)
        return d1, d2

    def rm(self):
        """rm1 = <B_i|1/r|B_j>, rm2 = <B_i|1/r^2|B_j> (lower band)."""
        rm1 = np.zeros((self.ns, self.ks), order="F")
        rm2 = np.zeros((self.ns, self.ks), order="F")
        lib().ref_galerkin(rm1, rm2)
        return rm1, rm2

    def gen_lval(self, pbs, lbs, iech, z, kclosd, db2):
        """The reference's Gen_Lval (l_core.f90:92-170) on orbitals pbs
        with the given db2: (mlh, hl_full, L_int, L_vec)."""
        nbf = pbs.shape[1]
        self.set_orbitals(pbs)
        lib().ref_set_lcore(float(z), int(kclosd), nbf, np.ascontiguousarray(lbs, dtype=np.int32),
                            np.ascontiguousarray(iech, dtype=np.int32), _f(db2))
        ns = self.ns
        mlh = C.c_int()
        hl = np.zeros((ns, ns, int(max(lbs)) + 1), order="F")
        li = np.zeros((nbf, nbf), order="F")
        lv = np.zeros((ns, nbf), order="F")
        lib().ref_gen_lval(mlh, hl, li, lv)
        return mlh.value, hl, li, lv


def add_cdata(j1, j2, j3, j4, c):
    """The reference's Add_cdata, record by record, into one block."""
    n = len(c)
    i32 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.int32))
    k = [np.zeros(n, np.int32) for _ in range(4)]
    co = np.zeros(n, order="F")
    nout = C.c_int()
    lib().ref_add_cdata(n, i32(j1), i32(j2), i32(j3), i32(j4), _f(c), *k, co, nout)
    m = nout.value
    return [x[:m] for x in k], co[:m]


def merge_cdata(ip, jp, k1, k2, k3, k4, c, eps):
    """The reference's Merge_CDATA: (ipt, merged cdata)."""
    i32 = lambda a: np.asfortranarray(np.asarray(a, dtype=np.int32))
    n = len(c)
    cc = _f(c).copy()
    ipt = np.zeros(n, np.int32)
    nc = C.c_int()
    lib().ref_merge_cdata(len(ip), i32(ip), i32(jp), n, i32(k1), i32(k2), i32(k3), i32(k4), cc, float(eps), ipt, nc)
    return ipt[: nc.value].copy(), cc


def density(ns, ks, p1, p2, typ):
    cols = {"s": ks, "n": 2 * ks - 1, "x": ns}[typ]
    d = np.zeros((ns, cols), order="F")
    lib().ref_density(ns, ks, d, _f(p1), _f(p2), typ.encode())
    return d


def sum_amb(ns, ks, a, b, sym):
    return lib().ref_sum_amb(ns, ks, _f(a), _f(b), sym.encode())


def bav(ns, ks, b, v, typ="s", side="r"):
    y = np.zeros(ns, order="F")
    lib().ref_bav(ns, ks, _f(b), _f(v), y, typ.encode(), side.encode())
    return y


def bxv(ks, ns, b, v):
    y = np.zeros(ns, order="F")
    lib().ref_bxv(ks, ns, _f(b), _f(v), y)
    return y


def bvmv(ns, ks, a, x, y):
    return lib().ref_bvmv(ns, ks, _f(a), _f(x), _f(y))


class RefMatrix:
    """module bsr_matrix with the UPDATE_* accumulators."""

    def __init__(self, ns, nch, ncp, mk):
        self.ns, self.nch, self.ncp, self.mk = ns, nch, ncp, mk
        lib().ref_alloc_matrix(nch, ncp, mk)

    def update_hx(self, ich, jch, d, sym):
        dd = np.zeros((self.ns, self.ns), order="F")
        dd[:, : d.shape[1]] = d
        lib().ref_update_hx(ich, jch, dd, sym.encode())

    def update_hl(self, ich, jch, c):
        lib().ref_update_hl(ich, jch, c)

    def update_hw(self, ich, jch, v, w):
        lib().ref_update_hw(ich, jch, _f(v), _f(w))

    def update_hv(self, ich, ic, v, c):
        lib().ref_update_hv(ich, ic, _f(v), c)

    def update_hb(self, ic, jc, c):
        lib().ref_update_hb(ic, jc, c)

    def update_cf(self, k, ich, jch, c):
        lib().ref_update_cf(k, ich, jch, c)

    def get(self):
        ns, nch, ncp, mk = self.ns, self.nch, self.ncp, self.mk
        nij = nch * (nch + 1) // 2
        hcc = np.zeros((ns, ns, nij), order="F")
        hcb = np.zeros((ns, nch, max(ncp, 1)), order="F")
        hbb = np.zeros(max(ncp * (ncp + 1) // 2, 1), order="F")
        acf = np.zeros((nch, nch, mk + 1), order="F")
        lib().ref_get_matrix(hcc.reshape(-1, order="F"), hcb.reshape(-1, order="F"),
                             hbb, acf.reshape(-1, order="F"))
        return hcc, hcb, hbb, acf


# ---- Dirac B-spline routines: oracle/_ref/libdbs_ref.so ---------------------
DBS_PATH = os.path.join(HERE, "_ref", "libdbs_ref.so")
_D = None


def dbs_available() -> bool:
    return os.path.exists(DBS_PATH)


def dbs_lib():
    global _D
    if _D is None:
        L = C.CDLL(DBS_PATH)
        i = C.c_int
        sig = {
            "ref_dbs_init": (None, [i, i, i, i, _dp]),
            "ref_dbs_tables": (None, [_dp, _dp, _dp, _dp]),
            "ref_dbs_moments": (None, [i, i, _dp]),
            "ref_dbs_mrk": (None, [i, i, _dp]),
            "ref_dbs_convol": (None, [i, i, i, _dp, _dp]),
            "ref_dbsr_setup": (None, [i, _dp, _ip, i, _ip, _ip, _dp, i, i, i, i, _ip, _ip]),
            "ref_dbsr_rdata": (None, [i, i, i, _ip, _ip, _ip, _ip, _dp]),
            "ref_dbsr_get": (None, [_dp, _dp, _dp, _dp, _dp]),
            "ref_dbsr_odata": (None, [i, i, _ip, _ip, _dp, C.c_double, i]),
            "ref_dbsr_diag_fppqq": (None, []),
            "ref_dbsr_get_otarg": (None, [_dp]),
            "ref_dbsr_zero_cond": (None, [_ip, _ip, i, i, i, i, _ip, _ip]),
            "ref_dbsr_transform": (None, [i, _dp, _ip]),
            "ref_dbsr_get_overlaps": (None, [i, _dp]),
            "ref_dbsr_record": (None, [C.c_char_p, i]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _D = L
    return _D


class RefDbs:
    """The reference's DBS module state for one grid (one active at a time)."""

    def __init__(self, ks, ns, ksp, ksq, t):
        self.ks, self.ns, self.ksp, self.ksq = ks, ns, ksp, ksq
        self.nv = ns - ks + 1
        dbs_lib().ref_dbs_init(ks, ns, ksp, ksq, np.asfortranarray(t, dtype=np.float64))

    def tables(self):
        nv, ks = self.nv, self.ks
        gr, grw = np.zeros((nv, ks), order="F"), np.zeros((nv, ks), order="F")
        pb = np.zeros((nv + 1, ks, self.ksp), order="F")
        qb = np.zeros((nv + 1, ks, self.ksq), order="F")
        dbs_lib().ref_dbs_tables(gr, grw, pb, qb)
        return dict(gr=gr, grw=grw, pbsp=pb, qbsp=qb)

    def moments(self, q, k):
        out = np.zeros((self.ks * self.ks, self.nv), order="F")
        dbs_lib().ref_dbs_moments(q, k, out)
        return out

    def mrk(self, family, k):
        out = np.zeros((self.ns, self.ns, self.ks, self.ks), order="F")
        dbs_lib().ref_dbs_mrk(family, k, out)
        return out

    def convol(self, family, k, icase, d):
        ns, ks = self.ns, self.ks
        a = np.zeros((ns, ks if icase <= 2 else ns), order="F")
        dbs_lib().ref_dbs_convol(family, k, icase, np.asfortranarray(d), a)
        return a


def _i32(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int32).reshape(-1))


class RefDbsr:
    """The reference's R_data (DBSR_MAT3) on the active DBS grid: module
    state filled by ref_dbsr_setup, lists run through R_data(itype,kpol,0..3)
    as Gen_matrix does, accumulators copied out by ref_dbsr_get."""

    def __init__(self, pq, ipbs, i_ch, j_ch, v_ch, nch, npert, mk, ntarg, iptar):
        pq = np.asfortranarray(pq, dtype=np.float64)
        self.ns = pq.shape[0]
        self.mk, self.ntarg, self.npert = mk, ntarg, npert
        v = np.asfortranarray(np.asarray(v_ch, dtype=np.float64).reshape(2 * self.ns, -1))
        if v.size == 0:
            v = np.zeros((2 * self.ns, 1), order="F")
        sz = np.zeros(4, dtype=np.int32)
        dbs_lib().ref_dbsr_setup(pq.shape[2], pq, _i32(ipbs), len(i_ch), _i32(list(i_ch) or [0]),
                                 _i32(list(j_ch) or [0]), v, nch, npert, mk, ntarg, _i32(iptar), sz)
        self.iicc, self.iicb, self.iibb, self.ms = (int(x) for x in sz)

    def rdata(self, itype, kpol, k1, k2, k3, k4, c):
        dbs_lib().ref_dbsr_rdata(itype, kpol, len(c), _i32(k1), _i32(k2), _i32(k3), _i32(k4),
                                 np.asfortranarray(c, dtype=np.float64))

    def diag_fppqq(self):
        """UPDATE_HX(ich,ich,fppqq,1) for every channel (sub1.f90:78)."""
        dbs_lib().ref_dbsr_diag_fppqq()

    def odata(self, itype, k1, k2, c, eps_c=1e-10, iitar=0):
        """O_data(itype) on one overlap list (O_data.f90:2-90)."""
        n = len(c)
        if n == 0:
            return
        dbs_lib().ref_dbsr_odata(itype, n, _i32(k1), _i32(k2), np.asfortranarray(c, dtype=np.float64),
                                 float(eps_c), int(iitar))

    def otarg(self):
        out = np.zeros(max(self.ntarg * (self.ntarg + 1) // 2, 1))
        dbs_lib().ref_dbsr_get_otarg(out)
        return out[: self.ntarg * (self.ntarg + 1) // 2]

    def zero_cond(self, lch, kch, ilzero=-1, jlzero=-1, ibzero=1, jbzero=1):
        """Zero_cond (zero_cond.f90:2-56): iprm(nch*ms) and nhm."""
        nch = len(lch)
        iprm = np.zeros(nch * self.ms, np.int32)
        nhm = np.zeros(1, np.int32)
        dbs_lib().ref_dbsr_zero_cond(_i32(lch), _i32(kch), ilzero, jlzero, ibzero, jbzero, iprm, nhm)
        return iprm, int(nhm[0])

    def transform(self, icase, diag, ipsol):
        """transform_matrix (transform_matrix.f90:2-66) on the module matrices."""
        dbs_lib().ref_dbsr_transform(int(icase), np.asfortranarray(diag, dtype=np.float64), _i32(ipsol))

    def record(self, path):
        """Record_matrix (record_matrix.f90:2-51) of the module matrices into a new file."""
        b = os.fsencode(str(path))
        dbs_lib().ref_dbsr_record(b, len(b))

    def overlaps(self, nch):
        n = nch + self.npert
        out = np.zeros((n, n), order="F")
        dbs_lib().ref_dbsr_get_overlaps(n, out)
        return out

    def get(self):
        ms = self.ms
        hch = np.zeros((ms, ms, max(self.iicc, 1)), order="F")
        hcp = np.zeros((ms, max(self.iicb, 1)), order="F")
        hp = np.zeros(max(self.iibb, 1), order="F")
        acf = np.zeros((max(self.iicc, 1), self.mk + 1), order="F")
        htarg = np.zeros(max(self.ntarg * (self.ntarg + 1) // 2, 1), order="F")
        dbs_lib().ref_dbsr_get(hch.reshape(-1, order="F"), hcp.reshape(-1, order="F"), hp,
                               acf.reshape(-1, order="F"), htarg)
        return dict(hch=hch[:, :, : self.iicc], hcp=hcp[:, : self.iicb], hp=hp[: self.iibb],
                    acf=acf[: self.iicc], htarg=htarg[: self.ntarg * (self.ntarg + 1) // 2])


def dirac_assemble(w):
    """Reference run of a synthetic Dirac wave (bsr_amd.synth.DiracWave)."""
    RefDbs(w.ks, w.ns, w.ksp, w.ksq, w.t)
    r = RefDbsr(w.pq, w.ipbs, w.i_ch, w.j_ch, w.v_ch, w.nch, w.npert, w.mk, w.ntarg, w.iptar)
    for L in w.lists:
        r.rdata(L.itype, L.kpol, L.k1, L.k2, L.k3, L.k4, L.c)
    return r.get()


def dirac_overlap(w, eps_c=1e-10, iitar=0, keep=False):
    """Reference overlap pass of a synthetic Dirac wave: fppqq on the channel
    diagonal blocks, then O_data on every overlap list (sub1.f90:78-79).
    keep=True also returns the RefDbsr that holds the result."""
    RefDbs(w.ks, w.ns, w.ksp, w.ksq, w.t)
    r = RefDbsr(w.pq, w.ipbs, w.i_ch, w.j_ch, w.v_ch, w.nch, w.npert, w.mk, w.ntarg, w.iptar)
    r.diag_fppqq()
    for L in w.olists:
        r.odata(L.itype, L.k1, L.k2, L.c, eps_c, iitar)
    out = r.get()
    out.pop("acf"), out.pop("htarg")
    out["otarg"] = r.otarg()
    return (out, r) if keep else out


# ---- list builder routines: oracle/_ref/liblb_ref.so --------------------------
LB_PATH = os.path.join(HERE, "_ref", "liblb_ref.so")
_LB = None


def lb_available() -> bool:
    return os.path.exists(LB_PATH)


def lb_lib():
    global _LB
    if _LB is None:
        L = C.CDLL(LB_PATH)
        i = C.c_int
        sig = {
            "ref_lb_read": (None, [i, _dp, _ip, _ip, _ip, C.POINTER(C.c_int), _ip]),
            "ref_lb_get_dets": (None, [_ip, _ip, _ip, _ip, _ip, _ip]),
            "ref_lb_set_dets": (None, [i, i, _ip, _ip, _ip, i, i, _ip, _ip, _ip]),
            "ref_lb_setup": (None, [i, _ip, _ip, _ip, _dp, i, i, _ip, i, _ip]),
            "ref_lb_decode": (None, [i, _ip, _ip]),
            "ref_lb_jsym": (None, [i, _ip, _ip]),
            "ref_lb_expand": (None, [i, i, C.c_double, C.c_double, C.c_double, i, _dp, _ip, _ip, _ip, i, _ip, _ip,
                                     _dp, i, _ip, i, _ip, _ip, i, _ip, i, _ip, _ip, _ip, _dp,
                                     C.POINTER(C.c_int)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LB = L
    return _LB


def lb_read(dirpath, cap=1 << 20):
    """Read_dets on dirpath/int_bnk.tst and the coefficient records after it:
    (c, ijt, int, idf), (kpd, ipd, npd, kpf, ipf, npf)."""
    cwd = os.getcwd()
    c = np.zeros(cap)
    a = [np.zeros(cap, np.int32) for _ in range(3)]
    n = C.c_int()
    sz = np.zeros(4, np.int32)
    try:
        os.chdir(dirpath)
        lb_lib().ref_lb_read(cap, c, *a, C.byref(n), sz)
    finally:
        os.chdir(cwd)
    m = n.value
    ndet, kdet, ndef, kdef = (int(x) for x in sz)
    d = [np.zeros(max(x, 1), np.int32) for x in (ndet, ndet, kdet, ndef, ndef, kdef)]
    lb_lib().ref_lb_get_dets(*d)
    d = [x[:k] for x, k in zip(d, (ndet, ndet, kdet, ndef, ndef, kdef))]
    return (c[:m], a[0][:m], a[1][:m], a[2][:m]), tuple(d)


def lb_decode(ints):
    ints = _i32(ints)
    out = np.zeros((6, len(ints)), np.int32, order="F")
    lb_lib().ref_lb_decode(len(ints), ints, out)
    return out


def lb_jsym(met, quads):
    q = np.asfortranarray(np.asarray(quads, np.int32).reshape(4, -1))
    lb_lib().ref_lb_jsym(q.shape[1], _i32(met), q)
    return q


def lb_expand(b, icase, mk, eps_c=1e-10, eps_det=1e-10, eps_soo=1e-3, cap=1 << 21):
    """The reference's routines walked in State_res order over a synthetic
    bank (bsr_amd.synth_bank.BankWave): (itype, kpol, k1, k2, k3, k4, c)."""
    w = b.wave
    L = lb_lib()
    L.ref_lb_setup(w.nbf, _i32(w.iech), _i32(w.lbs), _i32(np.asarray(b.ibort).reshape(-1, order="F")),
                   np.asfortranarray(np.asarray(b.obs, np.float64).reshape(-1, order="F")), b.nelc,
                   len(b.ipconf), _i32(b.ipconf), len(b.ippert) - 1, _i32(b.ippert))
    z = np.zeros(1, np.int32)
    L.ref_lb_set_dets(len(b.kpd), len(b.npd), _i32(b.kpd) if len(b.kpd) else z, _i32(b.ipd) if len(b.ipd) else z,
                      _i32(b.npd) if len(b.npd) else z, len(b.kpf), len(b.npf), _i32(b.kpf) if len(b.kpf) else z,
                      _i32(b.ipf) if len(b.ipf) else z, _i32(b.npf) if len(b.npf) else z)
    t = np.zeros(cap, np.int32)
    kp = np.zeros(cap, np.int32)
    kk = np.zeros((4, cap), np.int32, order="F")
    cc = np.zeros(cap)
    n = C.c_int()
    L.ref_lb_expand(icase, mk, eps_c, eps_det, eps_soo, len(b.c), _f(b.c), _i32(b.ijt), _i32(b.intb), _i32(b.idf),
                    len(b.ip_state), _i32(b.ip_state), _i32(b.no_state), _f(b.wc), len(b.ip_orb), _i32(b.ip_orb),
                    len(b.it1), _i32(b.it1), _i32(b.it2), len(b.ip_stat), _i32(b.ip_stat), cap, t, kp,
                    kk.reshape(-1, order="F"), cc, C.byref(n))
    m = n.value
    if m > cap:
        raise RuntimeError("lb_expand: cap too small")
    return t[:m], kp[:m], kk[0, :m], kk[1, :m], kk[2, :m], kk[3, :m], cc[:m]
