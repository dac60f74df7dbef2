"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the DBSR_MAT3 steps around
the Dirac R^k assembly: the B-spline overlap matrix fppqq
(src/libs/DBS/DBS_gauss.f90:256-280 with ZINTYm, zinty.f90:2-27), the
overlap pass O_data (src/dbsr/DBSR_MAT3_MPI/O_data.f90:2-90, with
UPDATE_HX/HW/HV/HB and Target_h of updates.f90 and target_subs.f90), the
zero conditions (zero_cond.f90:2-56), the channel diagonalisation
(diag.f90:2-81; LAPACK DSYGV through scipy) and the basis transform
(transform_matrix.f90:2-66).

Matrices use the layout of the GPU library: channel blocks ms x ms
(ms = 2 ns, P rows and columns first), numbered diagonal blocks first
(alloc_dbsr_matrix on one process), channel-perturber vectors hcp(ms,
(ich-1)*npert + ip), perturber scalars hp(i(i-1)/2 + j), target matrices
(it, jt) packed the same way.  Products are written as the reference
evaluates them (x = v w^T, then C x), so the overlap pass agrees with the
reference bit for bit on the same fppqq.
"""
from __future__ import annotations

import numpy as np

IBO = 2 ** 15


def dblk(nch, ich, jch):
    """Block of channel pair (ich, jch), 1-based channels."""
    if ich == jch:
        return ich - 1
    i, j = max(ich, jch), min(ich, jch)
    return nch + (i - 1) * (i - 2) // 2 + (j - 1)


def fppqq(tables, ks, ns, ksp, ksq):
    """The overlap matrix of the P and Q B-spline bases, ms x ms:
    fpbs in the PP quadrant, fqbs in the QQ quadrant (ZINTYm with the
    Gauss weights)."""
    grw, pb, qb = tables["grw"], tables["pbsp"], tables["qbsp"]
    nv = ns - ks + 1
    out = np.zeros((2 * ns, 2 * ns))
    for off, b, kb in ((0, pb, ksp), (ns, qb, ksq)):
        for iv in range(nv):
            for ith in range(kb):
                for jth in range(kb):
                    s = 0.0
                    for m in range(ks):
                        s += grw[iv, m] * b[iv, m, ith] * b[iv, m, jth]
                    out[off + iv + ith, off + iv + jth] += s
    return out


class Overlap:
    """The overlap matrices of a Dirac wave: S channel blocks hch, hcp, hp,
    otarg; fppqq on the channel diagonal blocks (sub1.f90:78), then O_data
    lists."""

    def __init__(self, w, fpq):
        self.w, self.fpq = w, fpq
        ms, nch, npert = w.ms, w.nch, w.npert
        self.nblk = nch * (nch + 1) // 2
        self.hch = np.zeros((ms, ms, self.nblk), order="F")
        self.hcp = np.zeros((ms, max(nch * npert, 1)), order="F")
        self.hp = np.zeros(max(npert * (npert + 1) // 2, 1))
        self.otarg = np.zeros(w.ntarg * (w.ntarg + 1) // 2)
        for ich in range(1, nch + 1):
            self.update_hx(ich, ich, fpq, 1.0)
        self.kbort = {(int(i), int(j)): q for q, (i, j) in enumerate(zip(w.i_ch, w.j_ch))}

    def update_hx(self, ich, jch, d, c):
        """UPDATE_HX, updates.f90:79-106"""
        xx = c * d if ich >= jch else c * d.T
        self.hch[:, :, dblk(self.w.nch, ich, jch)] += xx

    def update_hw(self, ich, jch, v, w, c):
        """UPDATE_HW, updates.f90:176-204"""
        x = np.outer(v, w)
        if ich > jch:
            xx = c * x
        elif jch > ich:
            xx = c * x.T
        else:
            xx = (c / 2.0) * (x + x.T)
        self.hch[:, :, dblk(self.w.nch, ich, jch)] += xx

    def vch(self, io):
        i1, i2 = divmod(int(io), IBO)
        q = self.kbort.get((i1, i2))
        if q is None:
            raise ValueError("O: KBORT=0")
        return i1, self.w.v_ch[:, q]

    def odata(self, itype, k1, k2, c, eps_c=1e-10, iitar=0):
        """O_data(itype), O_data.f90:2-90"""
        w = self.w
        for K1, K2, cc in zip(k1, k2, c):
            if abs(cc) < eps_c:
                continue
            if itype == 1:
                ic, jc = int(K1), int(K2)
                i, j = max(ic, jc), min(ic, jc)
                self.hp[i * (i - 1) // 2 + j - 1] += cc
            elif itype == 2:
                i1, v = self.vch(K1)
                ich, jc = int(w.ipbs[i1 - 1]), int(K2)
                self.hcp[:, (ich - 1) * w.npert + jc - 1] += cc * v
            elif itype == 3:
                i1, v = self.vch(K1)
                j1, u = self.vch(K2)
                self.update_hw(int(w.ipbs[i1 - 1]), int(w.ipbs[j1 - 1]), v, u, cc)
            elif itype == 4:
                i1, i2 = divmod(int(K1), IBO)
                ich, jch = int(w.ipbs[i1 - 1]), int(w.ipbs[i2 - 1])
                it, jt = int(w.iptar[ich - 1]), int(w.iptar[jch - 1])
                i, j = max(it, jt), min(it, jt)
                self.otarg[(i - 1) * i // 2 + j - 1] += cc
                if iitar > 1 and ich != jch:
                    self.update_hx(ich, jch, self.fpq, cc)
            else:
                raise ValueError("O_data: unknown structure (itype)")


def zero_cond(ns, nsp, nsq, ksp, ksq, lch, kch, ilzero=-1, jlzero=-1, ibzero=1, jbzero=1):
    """iprm (nch*ms, 1 = kept, 0 = deleted) and nhm - npert, zero_cond.f90."""
    nch = len(lch)
    iprm = np.ones(nch * 2 * ns, dtype=np.int32)
    ip, jp = 0, 0
    for ich in range(nch):
        mi = 0
        if ilzero >= 0:
            il = ilzero
        else:
            il = lch[ich] + 1
            if il > ksp - 1:
                il, mi = 1, 1
        mj = 0
        if jlzero >= 0:
            jl = jlzero
        else:
            jl = 0
            if kch[ich] < 0:
                jl = lch[ich] + 2
            if kch[ich] >= 1:
                jl = lch[ich]
            if jl > ksq - 1:
                jl, mj = 1, 1
        if mi == 1 or mj == 1:
            il, jl = 1, 1
        for lim, mm in ((il, nsp + 1 - ibzero), (jl, nsq + 1 - jbzero)):
            for i in range(1, ns + 1):
                if i <= lim or i >= mm:
                    iprm[ip] = 0
                jp += iprm[ip]
                ip += 1
    return iprm, jp


def diag_channels(hch, sdiag, iprm, nch, edmin, edmax, egap):
    """Diag_channels (diag.f90:2-81): per channel the generalized problem of
    the kept rows and columns of the H and S diagonal blocks; solutions in
    the window, sign fixed so that the largest component is positive,
    stored as columns 1..ksol of diag(:,:,ich) with the energies in its
    last column.  diag starts as the S diagonal blocks (sub1.f90:84)."""
    import scipy.linalg as sl
    ms = hch.shape[0]
    diag = np.array(sdiag, order="F", copy=True)
    ipsol = np.zeros(nch, dtype=np.int32)
    for ich in range(nch):
        keep = np.nonzero(iprm[ich * ms:(ich + 1) * ms])[0]
        a = hch[np.ix_(keep, keep, [ich])][:, :, 0]
        b = sdiag[np.ix_(keep, keep, [ich])][:, :, 0]
        wv, vv = sl.eigh(a, b, lower=True)
        ksol = 0
        for i in range(len(keep)):
            if wv[i] < edmin or wv[i] > edmax or abs(wv[i]) < egap:
                continue
            v = np.zeros(ms)
            v[keep] = vv[:, i]
            if v[np.argmax(np.abs(v))] < 0.0:
                v = -v
            diag[:, ksol, ich] = v
            diag[ksol, ms - 1, ich] = wv[i]
            ksol += 1
        ipsol[ich] = ksol
    return diag, ipsol


def _maxabs(a):
    # Fortran maxval of an empty section is -huge(0d0)
    return float(np.max(np.abs(a))) if a.size else -np.finfo(np.float64).max


def transform(hch, hcp, diag, ipsol, nch, npert, hp=None):
    """transform_matrix (transform_matrix.f90:2-66) on copies of hch, hcp;
    also the overlaps matrix (nch+npert)^2 it records for icase 0: the
    largest |element| of each transformed block and vector, |hp| of the
    nonzero perturber pairs."""
    ms = hch.shape[0]
    h, v = np.array(hch, copy=True), np.array(hcp, copy=True)
    ov = np.zeros((nch + npert, nch + npert), order="F")
    for ich in range(1, nch + 1):
        isol = ipsol[ich - 1]
        for jch in range(1, ich):
            jsol = ipsol[jch - 1]
            k = dblk(nch, ich, jch)
            if np.max(np.abs(h[:, :, k])) < 1e-20:
                continue
            c = np.zeros((ms, ms))
            c[:isol, :jsol] = (diag[:, :isol, ich - 1].T @ h[:, :, k]) @ diag[:, :jsol, jch - 1]
            h[:, :, k] = c
            ov[ich - 1, jch - 1] = _maxabs(c[:isol, :jsol])
    for ich in range(1, nch + 1):
        isol = ipsol[ich - 1]
        for ip in range(1, npert + 1):
            q = (ich - 1) * npert + ip - 1
            if np.max(np.abs(v[:, q])) < 1e-20:
                continue
            wv = np.zeros(ms)
            wv[:isol] = diag[:, :isol, ich - 1].T @ v[:, q]
            v[:, q] = wv
            ov[nch + ip - 1, ich - 1] = _maxabs(wv[:isol])
    if hp is not None:
        for i in range(1, npert + 1):
            for j in range(1, i + 1):
                x = hp[i * (i - 1) // 2 + j - 1]
                if x != 0.0:
                    ov[nch + i - 1, nch + j - 1] = abs(x)
    return h, v, ov


def energy_shift(hch, shch, fpq, iptar, ecore, etarg):
    """The diagonal channel blocks of H with the core and target energies:
    H_ii + Ecore S_ii + (Etarg(iptar(i)) - Ecore) fppqq (sub1.f90:96 and
    :128-130; there the Ecore S term comes first and the integrals are
    added to it)."""
    h = np.array(hch, copy=True)
    for ich in range(len(iptar)):
        c = etarg[int(iptar[ich]) - 1] - ecore
        h[:, :, ich] = h[:, :, ich] + ecore * shch[:, :, ich] + c * fpq
    return h


def record_matrix(hch, hcp, hp, ipsol, nch, npert):
    """Record_matrix (src/dbsr/DBSR_MAT3_MPI/record_matrix.f90:2-51) as the
    list of records it writes: for every off-diagonal channel block whose
    isol(ich) x isol(jch) corner has an element of at least 1e-20, (ich,
    jch) and the corner (column-major); then (ip+nch, ich) and the corner of
    each channel-perturber vector likewise; (ip+nch, jp+nch) and every
    non-zero perturber element; (0, 0)."""
    recs = []
    for ich in range(2, nch + 1):
        ni = ipsol[ich - 1]
        for jch in range(1, ich):
            nj = ipsol[jch - 1]
            c = hch[:ni, :nj, dblk(nch, ich, jch)]
            if _maxabs(c) < 1e-20:
                continue
            recs += [np.array([ich, jch], np.int32), np.asfortranarray(c).ravel(order="F")]
    for ich in range(1, nch + 1):
        ni = ipsol[ich - 1]
        for ip in range(1, npert + 1):
            v = hcp[:ni, (ich - 1) * npert + ip - 1]
            if _maxabs(v) < 1e-20:
                continue
            recs += [np.array([ip + nch, ich], np.int32), np.array(v)]
    for ip in range(1, npert + 1):
        for jp in range(1, ip + 1):
            x = hp[ip * (ip - 1) // 2 + jp - 1]
            if x == 0.0:
                continue
            recs += [np.array([ip + nch, jp + nch], np.int32), np.array([x])]
    recs.append(np.array([0, 0], np.int32))
    return recs


def solution_records(ns, nch, npert, nsp, nsq, diag, ipsol):
    """The records sub1.f90:154-158 writes before the transformed overlap
    matrix: (ns, nch, npert, nsp, nsq); nsol; jpsol(0:nch); eval(nsol) =
    diag(is, ms, ich) in channel order; the solution vectors diag(:, js, ich)."""
    ms = diag.shape[0]
    jpsol = np.concatenate([[0], np.cumsum(ipsol)]).astype(np.int32)
    ev = np.concatenate([diag[:k, ms - 1, i] for i, k in enumerate(ipsol)] + [np.zeros(0)])
    vec = np.concatenate([diag[:, :k, i].ravel(order="F") for i, k in enumerate(ipsol)] + [np.zeros(0)])
    return [np.array([ns, nch, npert, nsp, nsq], np.int32), np.array([jpsol[-1]], np.int32), jpsol, ev, vec]


def acf_records(acf, nch, mk, nelc_core):
    """sub1.f90:249-253 and :277-284: acf(i,0) += 2 nelc_core on the diagonal
    blocks, |acf| < 1e-5 set to zero, then mk and acf(icc(ich,jch),k) for
    ich fastest, then jch, then k; acf[blk + nblk*k] in the block numbering
    of dblk."""
    nblk = nch * (nch + 1) // 2
    a = np.array(acf, dtype=np.float64, copy=True)
    for ich in range(1, nch + 1):
        a[dblk(nch, ich, ich)] += 2 * nelc_core
    a[np.abs(a) < 0.00001] = 0.0
    full = np.array([a[dblk(nch, ich, jch) + nblk * k] for k in range(mk + 1) for jch in range(1, nch + 1)
                     for ich in range(1, nch + 1)])
    return [np.array([mk], np.int32), full]


def read_records(path):
    """The records of a Fortran unformatted sequential file (4-byte
    markers), as raw bytes."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    p = 0
    while p < len(data):
        n = int(np.frombuffer(data[p:p + 4], np.int32)[0])
        out.append(data[p + 4:p + 4 + n])
        assert int(np.frombuffer(data[p + 4 + n:p + 8 + n], np.int32)[0]) == n
        p += n + 8
    return out
