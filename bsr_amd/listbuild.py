"""Host mirror of the BSR_MAT3 list builder (State_res,
src/bsr/BSR_MAT3/state_res.f90:2-161): an int_bnk data bank and the state
tables of one partial wave in, the coefficient lists of module cmdata out,
in the batches Gen_matrix hands to I_data/O_data.  The work is done by the
native library (bsr_amd/csrc/state_res.cpp, bsrg_lb_* entry points); this
module packs the arrays and unpacks the results.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib as L

IBO = 2 ** 15


def pack_ibort(iech, lbs, kept, bound_iort=None):
    """IBORT as Read_data leaves it (src/bsr/BSR_MAT3/read_data.f90:172-193):
    bound pairs keep their orthogonality flag (bound_iort, default 1 on equal
    l), a continuum orbital i and a bound orbital j of equal l whose overlap
    is kept get the code i*ibo+j (continuum first), two continuum orbitals of
    equal l get i*ibo+j; everything else 0.  kept(nbf,nbf): nonzero where a
    continuum/bound overlap is kept."""
    iech = np.asarray(iech)
    lbs = np.asarray(lbs)
    nbf = len(iech)
    ib = np.zeros((nbf, nbf), np.int32, order="F")
    for i in range(nbf):
        for j in range(nbf):
            if lbs[i] != lbs[j]:
                continue
            ci, cj = iech[i] != 0, iech[j] != 0
            if not ci and not cj:
                ib[i, j] = 1 if bound_iort is None else bound_iort[i, j]
            elif ci and cj:
                ib[i, j] = (i + 1) * IBO + (j + 1)
            elif cj and kept[i, j]:
                ib[i, j] = (j + 1) * IBO + (i + 1)
            elif ci and kept[i, j]:
                ib[i, j] = (i + 1) * IBO + (j + 1)
    return ib


class ListBuilder:
    """bsrg_lb: module state of State_res plus its results."""

    def __init__(self):
        self._p = L.lib().bsrg_lb_create()
        if not self._p:
            raise MemoryError("bsrg_lb_create")

    def close(self):
        if self._p:
            L.lib().bsrg_lb_destroy(self._p)
            self._p = None

    def __del__(self):
        self.close()

    def read_bank(self, path):
        """int_bnk file -> (ndet, kdet, ndef, kdef, nrec)."""
        sz = np.zeros(5, np.int32)
        L.check(L.lib().bsrg_lb_read_bank(self._p, str(path).encode(), sz))
        self.sizes = tuple(int(x) for x in sz)
        return self.sizes

    def bank(self):
        ndet, kdet, ndef, kdef, nrec = self.sizes
        out = dict(c=np.zeros(max(nrec, 1)), ijt=np.zeros(max(nrec, 1), np.int32),
                   intb=np.zeros(max(nrec, 1), np.int32), idf=np.zeros(max(nrec, 1), np.int32),
                   kpd=np.zeros(max(ndet, 1), np.int32), ipd=np.zeros(max(ndet, 1), np.int32),
                   npd=np.zeros(max(kdet, 1), np.int32), kpf=np.zeros(max(ndef, 1), np.int32),
                   ipf=np.zeros(max(ndef, 1), np.int32), npf=np.zeros(max(kdef, 1), np.int32))
        L.check(L.lib().bsrg_lb_bank_get(self._p, *(out[k] for k in ("c", "ijt", "intb", "idf", "kpd", "ipd", "npd",
                                                                      "kpf", "ipf", "npf"))))
        n = dict(c=nrec, ijt=nrec, intb=nrec, idf=nrec, kpd=ndet, ipd=ndet, npd=kdet, kpf=ndef, ipf=ndef, npf=kdef)
        return {k: v[: n[k]] for k, v in out.items()}

    def set_records(self, c, ijt, intb, idf):
        n = len(c)
        z = np.zeros(1, np.int32)
        L.check(L.lib().bsrg_lb_set_records(self._p, n, L.f64(c) if n else np.zeros(1),
                                             *(L.i32(a) if n else z for a in (ijt, intb, idf))))

    def set_dets(self, kpd, ipd, npd, kpf, ipf, npf):
        z = np.zeros(1, np.int32)
        a = [L.i32(x) if len(x) else z for x in (kpd, ipd, npd, kpf, ipf, npf)]
        L.check(L.lib().bsrg_lb_set_dets(self._p, len(kpd), len(npd), a[0], a[1], a[2], len(kpf), len(npf), a[3],
                                         a[4], a[5]))

    def set_orbitals(self, iech, lef, ibort, obs, nelc):
        nbf = len(iech)
        L.check(L.lib().bsrg_lb_set_orbitals(self._p, nbf, L.i32(iech), L.i32(lef),
                                             L.i32(np.asarray(ibort).reshape(-1, order="F")),
                                             L.f64(np.asarray(obs).reshape(-1, order="F")), int(nelc)))

    def set_channels(self, ipconf, ippert):
        L.check(L.lib().bsrg_lb_set_channels(self._p, len(ipconf), L.i32(ipconf), len(ippert) - 1, L.i32(ippert)))

    def set_states(self, ip_state, no_state, wc, ip_orb, it1, it2, ip_stat):
        L.check(L.lib().bsrg_lb_set_states(self._p, len(ip_state), L.i32(ip_state), L.i32(no_state), L.f64(wc),
                                           len(ip_orb), L.i32(ip_orb), len(it1), L.i32(it1), L.i32(it2),
                                           len(ip_stat), L.i32(ip_stat)))

    def set_params(self, eps_c=1e-10, eps_det=1e-10, eps_soo=1e-3, nb=2000, mb=5000, kb=100):
        L.check(L.lib().bsrg_lb_set_params(self._p, eps_c, eps_det, eps_soo, nb, mb, kb))

    def run(self, icase, mk):
        """State_res for icase: (coefficients passed on, list records, batches)."""
        sz = np.zeros(3, np.int32)
        L.check(L.lib().bsrg_lb_run(self._p, int(icase), int(mk), sz))
        self.run_sizes = tuple(int(x) for x in sz)
        return self.run_sizes

    def expansion(self):
        n = self.run_sizes[0]
        out = [np.zeros(max(n, 1), np.int32) for _ in range(6)] + [np.zeros(max(n, 1))]
        L.check(L.lib().bsrg_lb_expansion(self._p, *out))
        return tuple(a[:n] for a in out)   # itype, kpol, k1, k2, k3, k4, c

    def lists(self):
        """(batch, jtype, jpol, k1, k2, k3, k4, c) in generation order."""
        n = self.run_sizes[1]
        out = [np.zeros(max(n, 1), np.int32) for _ in range(7)] + [np.zeros(max(n, 1))]
        L.check(L.lib().bsrg_lb_lists(self._p, *out))
        return tuple(a[:n] for a in out)


def lists_by_type(lists):
    """{(jtype, jpol): (k1, k2, k3, k4, c)} with the batches of one
    structure and multipole joined in generation order."""
    batch, jt, jp, k1, k2, k3, k4, c = lists
    out = {}
    for key in sorted(set(zip(jt.tolist(), jp.tolist()))):
        m = (jt == key[0]) & (jp == key[1])
        out[key] = (k1[m], k2[m], k3[m], k4[m], c[m])
    return out
