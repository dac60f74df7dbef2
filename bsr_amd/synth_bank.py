"""Seeded synthetic int_bnk data banks with the state tables of one partial
wave: the input of the list builder (State_res,
src/bsr/BSR_MAT3/state_res.f90:2-161) in the reference's formats.

The orbitals, grid and target data are those of a synthetic partial wave
(synth.make_wave).  On top of them, from numpy.random.default_rng([seed, 2]):

  orbitals     bound orbitals in two sets (odd / even index); OBS = p^T S p
               for equal l and different sets, 0 within a set, 1 on the
               diagonal (Read_data, read_data.f90:150-208); IBORT packed as
               Read_data packs it from the kept overlaps of the wave
  states       every state has nelc = 4 orbitals.  Channel ich has one or
               two configurations (w.p. 1/2 each): three distinct bound
               orbitals and the channel orbital last; each perturber one
               configuration of four bound orbitals.  WC ~ +-U[0.5, 1].
               Channel states and perturber states fall into symmetries of
               their own kind, about two states each; IP_stat lists the
               states by symmetry
  records      for every symmetry pair it >= jt, w.p. 0.6, one to three
               records of the kinds that pair allows (below), c ~ U[-1,1],
               multipole k ~ U{0..mk} for R^k (icase 5); overlap records
               (icase 11) w.p. 1/2 and one-electron records (icase 6,
               L(i1,i3) with i2 = i1, i4 = i3) w.p. 0.3 per record
  dets         the overlap factors the records name, built from
               determinants of orbital positions (row = position in the
               first state, column = position in the second):
               bound-only 1x1..3x3 (iext 1 or 2 for 1x1, 2x2), 4x4 for
               perturber pairs, the continuum overlap <kl|nl> (1x1), the
               continuum row of a 2x2, and <kl|k'l'> (1x1)

Kinds (cc = two channel states, pc = perturber then channel state, pp):
  cc  R^k  direct/exchange integral over both channel orbitals, factor 1
           or bound-only; one channel orbital in the integral and the other
           in an overlap; both in overlaps and a bound integral
  cc  O    both channel orbitals in overlaps; <kl|k'l'>
  cc  L    L(kl, k'l);   pc L: L(nl, kl);   pp L: L(nl, n'l)
  pc  R^k  the channel orbital in the integral; or in an overlap
  pc  O    the channel orbital in an overlap
  pp  R^k  any positions, factor 1 or bound-only;   pp O: likewise
Every record is valid for every state pair of its two symmetries: the
structures it produces are those Idef_itype accepts.

write() produces the file the reference reads: the configuration headers
read_symc_LS / Read_symt_LS / Read_oper_LS expect (one symmetry per
configuration symmetry, no operator records), Read_dets' determinants and
factors, then the coefficient records, as Fortran unformatted sequential
records with 4-byte markers.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass

import numpy as np

from . import grid as G
from . import synth
from .listbuild import pack_ibort

IDET = 2 ** 15
IDEF = 16
IBC = 2 ** 15
NELC = 4


def incode_int(met, k, i1, i2, i3, i4):
    """Incode_INT (src/libs/ZCONFLS/incode_int.f90:2-16)."""
    return ((((k * 16 + met) * 32 + i4) * 32 + i3) * 32 + i2) * 32 + i1


@dataclass
class BankWave:
    wave: synth.PartialWave
    nelc: int
    ibort: np.ndarray
    obs: np.ndarray
    ipconf: np.ndarray
    ippert: np.ndarray
    ip_state: np.ndarray
    no_state: np.ndarray
    wc: np.ndarray
    ip_orb: np.ndarray
    it1: np.ndarray
    it2: np.ndarray
    ip_stat: np.ndarray
    kpd: np.ndarray
    ipd: np.ndarray
    npd: np.ndarray
    kpf: np.ndarray
    ipf: np.ndarray
    npf: np.ndarray
    c: np.ndarray
    ijt: np.ndarray
    intb: np.ndarray
    idf: np.ndarray

    def write(self, path):
        def rec(f, payload):
            f.write(struct.pack("<i", len(payload)))
            f.write(payload)
            f.write(struct.pack("<i", len(payload)))

        def ints(*a):
            v = np.concatenate([np.atleast_1d(np.asarray(x, np.int32)) for x in a]) if a else np.zeros(0, np.int32)
            return v.astype("<i4").tobytes()

        nsym = len(self.it1)
        with open(path, "wb") as f:
            rec(f, ints(1, 1))                        # read_symc_LS
            for _ in range(6):
                rec(f, ints(0))
            rec(f, ints(nsym, nsym))                  # Read_symt_LS
            rec(f, ints(np.ones(nsym)))
            rec(f, ints(np.arange(nsym)))
            rec(f, ints(np.zeros(5 * nsym)))
            rec(f, ints(0))                           # Read_oper_LS
            nd, nf = len(self.kpd), len(self.kpf)
            rec(f, ints(nd, len(self.npd), nd))       # Read_dets
            for i in range(nd):
                rec(f, ints(self.kpd[i], self.ipd[i], i + 1, self.npd[self.ipd[i]:self.ipd[i] + self.kpd[i]]))
            if nd:
                rec(f, ints(nf, len(self.npf), nf))
                for i in range(nf):
                    rec(f, ints(self.kpf[i], self.ipf[i], i + 1, self.npf[self.ipf[i]:self.ipf[i] + self.kpf[i]]))
            for c, ijt, ib, idf in zip(self.c, self.ijt, self.intb, self.idf):
                rec(f, struct.pack("<d", float(c)) + ints(ijt, ib, idf))


def make_bank(seed: int, ks: int = 8, ns: int = 40, nch: int = 6, ncp: int = 3, mk: int = 3, nb: int = 6,
              ntarg: int = 3, z: float = 4.0, rmax: float = 20.0, p_pair: float = 0.6) -> BankWave:
    w = synth.make_wave(seed, ks, ns, z, rmax, nch, ncp, mk, nb=nb, ntarg=ntarg, fill=0.0, name="bank")
    rng = np.random.default_rng([seed, 2])
    nbf = w.nbf
    iech, lbs = w.iech, w.lbs
    # orbitals: IBORT packed from the kept overlaps; OBS between the sets
    bset = np.arange(nbf) % 2 + 1
    iort = np.zeros((nbf, nbf), np.int32)
    for i in range(nbf):
        for j in range(nbf):
            if iech[i] == 0 and iech[j] == 0 and lbs[i] == lbs[j] and (i == j or bset[i] != bset[j]):
                iort[i, j] = 1
    ibort = pack_ibort(iech, lbs, w.ibort, iort)
    S = G.band_to_dense_lower(G.tables(w.grid)[3])
    obs = np.zeros((nbf, nbf), order="F")
    for i in range(nbf):
        for j in range(nbf):
            if iech[i] or iech[j] or lbs[i] != lbs[j]:
                continue
            if i == j:
                obs[i, j] = 1.0
            elif iort[i, j]:
                s = float(w.pbs[:, i] @ S @ w.pbs[:, j])
                obs[i, j] = s if abs(s) >= w.eps_ovl else 0.0
    # states
    cfgs, ipconf = [], []
    for ich in range(1, nch + 1):
        for _ in range(1 + int(rng.random() < 0.5)):
            b = rng.choice(nb, 3, replace=False) + 1
            cfgs.append([*b.tolist(), nb + ich])
        ipconf.append(len(cfgs))
    nchs = len(cfgs)
    for _ in range(ncp):
        cfgs.append((rng.choice(nb, 4, replace=False) + 1).tolist())
    ippert = [ipconf[-1]] + [ipconf[-1] + ip for ip in range(1, ncp + 1)]
    ncfg = len(cfgs)
    ip_orb = np.asarray([o for cfg in cfgs for o in cfg], np.int32)
    no_state = np.full(ncfg, NELC, np.int32)
    ip_state = (np.arange(ncfg) * NELC).astype(np.int32)
    wc = rng.uniform(0.5, 1.0, ncfg) * rng.choice([-1.0, 1.0], ncfg)
    # symmetries of one kind each
    nsc = max(1, nchs // 2)
    nsp = max(1, ncp // 2) if ncp else 0
    sym = np.concatenate([rng.integers(1, nsc + 1, nchs), nsc + rng.integers(1, nsp + 1, ncp) if ncp else
                          np.zeros(0, np.int64)])
    nsym = nsc + nsp
    order = np.argsort(sym, kind="stable")
    ip_stat = (order + 1).astype(np.int32)
    it1 = np.zeros(nsym, np.int32)
    it2 = np.zeros(nsym, np.int32)
    for pos, s in enumerate(sym[order]):
        if it1[s - 1] == 0:
            it1[s - 1] = pos + 1
        it2[s - 1] = pos + 1
    kind = ["c"] * nsc + ["p"] * nsp
    # determinants and factors
    dets, npd, defs, npf = {}, [], {}, []
    kpd, ipd, kpf, ipf = [], [], [], []

    def det(pos):
        key = tuple(pos)
        if key not in dets:
            ipd.append(len(npd))
            kpd.append(len(pos))
            npd.extend(a * IDET + b for a, b in pos)
            dets[key] = len(kpd)
        return dets[key]

    def factor(items):
        key = tuple(items)
        if key not in defs:
            ipf.append(len(npf))
            kpf.append(len(items))
            npf.extend(d * IDEF + e for d, e in items)
            defs[key] = len(kpf)
        return defs[key]

    def bound_det(rows, cols, n):
        a = rng.choice(rows, n, replace=False)
        b = rng.choice(cols, n, replace=False)
        return [(int(x), int(y)) for x, y in zip(a, b)]

    def bound_items(rows, cols, big=False):
        items = []
        for _ in range(int(rng.integers(1, 3))):
            n = int(rng.integers(1, 5 if big else 4))
            e = int(rng.integers(1, 3)) if n <= 2 else 1
            items.append((det(bound_det(rows, cols, n)), e))
        return items

    b3, b4 = [1, 2, 3], [1, 2, 3, 4]
    recs = []   # (c, ijt, int, idf)

    def add(it, jt, icase, k, i, items):
        idf = factor(items) if items else 0
        recs.append((float(rng.uniform(-1, 1)), it * IBC + jt, incode_int(icase, k, *i), idf))

    for it in range(1, nsym + 1):
        for jt in range(1, it + 1):
            if rng.random() >= p_pair:
                continue
            pair = kind[it - 1] + kind[jt - 1]
            for _ in range(int(rng.integers(1, 4))):
                k = int(rng.integers(0, mk + 1))
                v = int(rng.integers(0, 3))
                if pair == "cc":
                    if v == 0:      # both channel orbitals in the integral
                        a, b = (int(x) for x in rng.choice(b3, 2))
                        i = [(4, a, 4, b), (a, 4, b, 4), (4, a, b, 4), (a, 4, 4, b)][int(rng.integers(0, 4))]
                        add(it, jt, 5, k, i, bound_items(b3, b3) if rng.random() < 0.5 else [])
                    elif v == 1:    # the first in an overlap, the second in the integral
                        a, b, c = (int(x) for x in rng.choice(b3, 3))
                        i = (a, b, 4, c) if rng.random() < 0.5 else (a, b, c, 4)
                        if rng.random() < 0.5:
                            items = [(det([(4, int(rng.choice(b3)))]), 1)]
                        else:
                            items = [(det([(4, 1), (2, 3)]), 1)]
                        add(it, jt, 5, k, i, items + (bound_items(b3, b3)[:1] if rng.random() < 0.3 else []))
                    else:           # both in overlaps
                        i = tuple(int(x) for x in rng.choice(b3, 4))
                        items = [(det([(4, int(rng.choice(b3)))]), 1), (det([(int(rng.choice(b3)), 4)]), 1)]
                        add(it, jt, 5, k, i, items)
                    if rng.random() < 0.3:   # L(kl, k'l) between the channel orbitals
                        add(it, jt, 6, 0, (4, 4, 4, 4), bound_items(b3, b3) if rng.random() < 0.5 else [])
                    if rng.random() < 0.5:
                        if rng.random() < 0.7:
                            items = [(det([(4, int(rng.choice(b3)))]), 1), (det([(int(rng.choice(b3)), 4)]), 1)]
                        else:
                            items = [(det([(4, 4)]), 1)]
                        add(it, jt, 11, 0, (1, 1, 1, 1), items)
                elif pair == "pc":
                    if v < 2:       # the channel orbital in the integral
                        a, b, c = (int(x) for x in rng.choice(b4, 3))
                        i = (a, b, 4, int(rng.choice(b3))) if rng.random() < 0.5 else (a, b, int(rng.choice(b3)), 4)
                        add(it, jt, 5, k, i, bound_items(b4, b3) if rng.random() < 0.5 else [])
                    else:           # in an overlap
                        i = (int(rng.choice(b4)), int(rng.choice(b4)), int(rng.choice(b3)), int(rng.choice(b3)))
                        add(it, jt, 5, k, i, [(det([(int(rng.choice(b4)), 4)]), 1)])
                    if rng.random() < 0.3:   # L(nl, kl)
                        a = int(rng.choice(b4))
                        add(it, jt, 6, 0, (a, a, 4, 4), bound_items(b4, b3) if rng.random() < 0.5 else [])
                    if rng.random() < 0.5:
                        add(it, jt, 11, 0, (1, 1, 1, 1), [(det([(int(rng.choice(b4)), 4)]), 1)])
                else:
                    i = tuple(int(x) for x in rng.choice(b4, 4))
                    add(it, jt, 5, k, i, bound_items(b4, b4, big=True) if rng.random() < 0.5 else [])
                    if rng.random() < 0.3:   # L(nl, n'l)
                        a, c = (int(x) for x in rng.choice(b4, 2))
                        add(it, jt, 6, 0, (a, a, c, c), bound_items(b4, b4) if rng.random() < 0.5 else [])
                    if rng.random() < 0.5:
                        add(it, jt, 11, 0, (1, 1, 1, 1), bound_items(b4, b4) if rng.random() < 0.5 else [])
    c, ijt, intb, idf = (np.asarray(x) for x in zip(*recs))
    i32 = lambda a: np.asarray(a, np.int32)
    return BankWave(w, NELC, ibort, obs, i32(ipconf), i32(ippert), ip_state, no_state, wc, ip_orb, it1, it2,
                    ip_stat, i32(kpd), i32(ipd), i32(npd), i32(kpf), i32(ipf), i32(npf), c.astype(np.float64),
                    i32(ijt), i32(intb), i32(idf))


def builder_for(b: BankWave, path=None, **params):
    """A ListBuilder holding the bank (read from path when given, else set
    directly) and the state tables of b."""
    from .listbuild import ListBuilder
    lb = ListBuilder()
    if path is not None:
        lb.read_bank(path)
    else:
        lb.set_records(b.c, b.ijt, b.intb, b.idf)
        lb.set_dets(b.kpd, b.ipd, b.npd, b.kpf, b.ipf, b.npf)
    w = b.wave
    lb.set_orbitals(w.iech, w.lbs, b.ibort, b.obs, b.nelc)
    lb.set_channels(b.ipconf, b.ippert)
    lb.set_states(b.ip_state, b.no_state, b.wc, b.ip_orb, b.it1, b.it2, b.ip_stat)
    lb.set_params(**params)
    return lb


def wave_from_lists(b: BankWave, rk, ov, name="bank_wave"):
    """The partial wave of b with the builder's lists: rk from icase 5
    (structures 1-9 by multipole), ov from icase 11 (structure 1)."""
    from dataclasses import replace
    from .listbuild import lists_by_type
    rk_lists = {}
    for (jt, k), (k1, k2, k3, k4, c) in lists_by_type(rk).items():
        rk_lists[(jt, k)] = synth.RecordList(k1, k2, k3, k4, c)
    parts = lists_by_type(ov)
    if parts:
        ks = sorted(parts)
        cat = [np.concatenate([parts[key][q] for key in ks]) for q in range(5)]
        olist = synth.RecordList(*cat)
    else:
        z = np.zeros(0, np.int32)
        olist = synth.RecordList(z, z, z, z, np.zeros(0))
    return replace(b.wave, name=name, rk_lists=rk_lists, olist=olist)
