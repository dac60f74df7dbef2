"""Seeded synthetic partial waves in the reference's input formats.

A partial wave is what BSR_MAT3's SUB1 holds for one klsp: the knot grid,
the orbital list (bound orbitals with B-spline coefficients, channel
orbitals with iech>0), the target data, and the sorted coefficient lists of
module cmdata for the overlap pass (O_data) and for every R^k structure
jtype 1..9 and multipole k (I_data).  The packing of K1..K4 follows
src/bsr/BSR_MAT3/idef_type.f90:5-51 with ibi = ibo = 2**15.

Workload definition (one draw of numpy.random.default_rng(seed)):
  bound orbitals   nb, l ~ U{0..3}, coefficients N(0,1) with p(1..l+1)=0
                   and p(ns)=0, normalised to p^T S p = 1
  channels         nch, l ~ U{0..3}; target index ~ U{1..ntarg}
  overlaps kept    channel/bound orbital pairs of equal l, each w.p. 0.7
  jtype 6-9        every channel pair (i>=j) x bound pair (a,b) x k,
                   included w.p. fill, c ~ U[-1,1]; the order of the pair in
                   K1 is flipped w.p. 1/2 for i != j
  jtype 1          perturber pairs (HB), kept-overlap x perturber (HV),
                   kept-overlap pairs (HW), channel-orbital pairs (<i|j>),
                   each included w.p. fill with a random bound quadruple
  jtype 2-5        channel x perturber (HV) and channel x kept overlap (HW),
                   w.p. fill, random bound pair and bound orbital
  overlaps (S)     perturber diagonals 1.0; HB/HV/HW/<i|j> records w.p. fill
                   with c ~ 0.01*U[-1,1] (keeps S positive definite)
Duplicates (equal K1..K4) are merged by summing, as Merge_cdata does.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import grid as G

IBI = 2 ** 15
IBO = 2 ** 15


@dataclass
class RecordList:
    k1: np.ndarray
    k2: np.ndarray
    k3: np.ndarray
    k4: np.ndarray
    c: np.ndarray

    def __len__(self):
        return len(self.c)

    @staticmethod
    def build(k1, k2, k3, k4, c) -> "RecordList":
        """Sort lexicographically on (K1,K2,K3,K4) and merge duplicates."""
        k = np.stack([np.asarray(x, np.int64) for x in (k1, k2, k3, k4)], axis=1)
        c = np.asarray(c, np.float64)
        if len(c) == 0:
            z = np.zeros(0, np.int32)
            return RecordList(z, z.copy(), z.copy(), z.copy(), np.zeros(0))
        keys, inv = np.unique(k, axis=0, return_inverse=True)
        cs = np.zeros(len(keys))
        np.add.at(cs, inv.ravel(), c)
        return RecordList(*(keys[:, q].astype(np.int32) for q in range(4)), cs)


@dataclass
class PartialWave:
    name: str
    seed: int
    grid: G.Grid
    nb: int
    nch: int
    ncp: int
    mk: int
    pbs: np.ndarray            # (ns, nbf)
    iech: np.ndarray           # (nbf,)
    lbs: np.ndarray            # (nbf,)
    ibort: np.ndarray          # (nbf, nbf)
    ntarg: int
    etarg: np.ndarray
    iptar: np.ndarray          # (nch,) 1-based
    ec: float
    olist: RecordList
    rk_lists: dict = field(default_factory=dict)   # (jtype, k) -> RecordList
    eps_c: float = 1e-10
    eps_ovl: float = 1e-8

    # reference-style attribute names
    @property
    def ks(self): return self.grid.ks
    @property
    def ns(self): return self.grid.ns
    @property
    def ml(self): return self.grid.ml
    @property
    def me(self): return self.grid.me
    @property
    def h(self): return self.grid.h
    @property
    def t(self): return self.grid.t
    @property
    def nbf(self): return self.pbs.shape[1]
    @property
    def nhm(self): return self.nch * self.ns + self.ncp
    @property
    def n_elements(self): return self.nhm * (self.nhm + 1) // 2

    def n_records(self):
        return sum(len(v) for v in self.rk_lists.values())

    def concat(self):
        """All R^k records in call order (k outer, jtype inner,
        state_res.f90:155-159): (jtype, kpol, k1, k2, k3, k4, c), packed
        once per wave as contiguous int32/float64 arrays."""
        packed = self.__dict__.get("_packed")
        if packed is None:
            packed = tuple(np.ascontiguousarray(a, dtype=np.float64 if q == 6 else np.int32)
                           for q, a in enumerate(self._concat()))
            self.__dict__["_packed"] = packed
        return packed

    def _concat(self):
        parts = [[], [], [], [], [], [], []]
        for k in range(self.mk + 1):
            for jt in range(1, 10):
                lst = self.rk_lists.get((jt, k))
                if lst is None or len(lst) == 0:
                    continue
                n = len(lst)
                parts[0].append(np.full(n, jt, np.int32))
                parts[1].append(np.full(n, k, np.int32))
                for q, a in enumerate((lst.k1, lst.k2, lst.k3, lst.k4)):
                    parts[2 + q].append(a)
                parts[6].append(lst.c)
        if not parts[0]:
            z = np.zeros(0, np.int32)
            return z, z, z, z, z, z, np.zeros(0)
        return tuple(np.concatenate(p) for p in parts)

    def footprint_bytes(self):
        """Device bytes the wave needs (matrices S and H, bands, bases, the
        summed band of each block)."""
        ns, ks = self.ns, self.ks
        nblk = self.nch * (self.nch + 1) // 2
        mats = 2 * 8 * (nblk * ns * ns + ns * self.nch * self.ncp)
        bands = 8 * (self.mk + 1) * ns * ns * ks * ks
        basis = 8 * (self.mk + 1) * self.nb * self.nb * ns * ns
        return mats + bands + basis + 8 * nblk * ns * ks


def _bound_orbitals(rng, g: G.Grid, nb: int):
    _, _, _, sb = G.tables(g)
    S = G.band_to_dense_lower(sb)
    lb = rng.integers(0, 4, nb)
    P = rng.standard_normal((g.ns, nb))
    for a in range(nb):
        P[: lb[a] + 1, a] = 0.0
        P[-1, a] = 0.0
        P[:, a] /= np.sqrt(P[:, a] @ S @ P[:, a])
    return P, lb


def make_wave(seed: int, ks: int, ns: int, z: float, rmax: float, nch: int, ncp: int, mk: int,
              nb: int = 5, ntarg: int = 5, fill: float = 0.05, h: float = 0.05,
              name: str = "wave", target_seed: int | None = None, grid: G.Grid | None = None) -> PartialWave:
    """target_seed, when given, draws the bound (target) orbitals from their
    own stream, so that waves of one target share them.  grid, when given,
    replaces the fixed-ns grid (ks, ns, z, rmax and h are then taken from it)."""
    rng = np.random.default_rng(seed)
    g = G.fixed_ns_grid(ks, ns, z, rmax, h) if grid is None else grid
    P, lb = _bound_orbitals(rng if target_seed is None else np.random.default_rng(target_seed), g, nb)
    lch = rng.integers(0, 4, nch)
    nbf = nb + nch
    pbs = np.zeros((g.ns, nbf), order="F")
    pbs[:, :nb] = P
    iech = np.concatenate([np.zeros(nb, np.int32), np.arange(1, nch + 1, dtype=np.int32)])
    lbs = np.concatenate([lb, lch]).astype(np.int32)
    ibort = np.zeros((nbf, nbf), np.int32, order="F")
    kept = []  # (channel orbital, bound orbital), 1-based
    for ich in range(nch):
        kl = nb + ich
        for a in range(nb):
            if lbs[kl] == lbs[a] and rng.random() < 0.7:
                ibort[kl, a] = ibort[a, kl] = 1
                kept.append((kl + 1, a + 1))
    kept_io = np.array([kl * IBO + nl for kl, nl in kept], np.int64)
    iptar = rng.integers(1, ntarg + 1, nch).astype(np.int32)
    etarg = np.sort(rng.uniform(-2.0, -0.5, ntarg))
    ec = float(rng.uniform(-3.0, -1.0))

    ch_i, ch_j = np.tril_indices(nch)  # 0-based, ch_i >= ch_j
    rand_pair = lambda n: (rng.integers(1, nb + 1, n) * IBI + rng.integers(1, nb + 1, n))
    rk = {}
    for k in range(mk + 1):
        # jtype 6-9: channel pair x bound pair
        for jt in (6, 7, 8, 9):
            m = rng.random((len(ch_i), nb, nb)) < fill
            q, a, b = np.nonzero(m)
            flip = rng.random(len(ch_i)) < 0.5
            ci, cj = ch_i[q] + 1, ch_j[q] + 1
            f = flip[q] & (ci != cj)
            k1 = np.where(f, cj * IBI + ci, ci * IBI + cj)
            rk[(jt, k)] = RecordList.build(k1, a + 1, b + 1, np.zeros(len(q)), rng.uniform(-1, 1, len(q)))
        # jtype 1
        parts = []
        if ncp > 0:
            pi, pj = np.tril_indices(ncp)
            sel = rng.random(len(pi)) < fill
            n = int(sel.sum())
            parts.append((rand_pair(n), rand_pair(n), -(pi[sel] + 1), -(pj[sel] + 1)))
            if len(kept_io):
                m = rng.random((len(kept_io), ncp)) < fill
                io, jc = np.nonzero(m)
                n = len(io)
                parts.append((rand_pair(n), rand_pair(n), kept_io[io], -(jc + 1)))
        if len(kept_io):
            m = rng.random((len(kept_io), len(kept_io))) < fill
            io, jo = np.nonzero(m)
            n = len(io)
            parts.append((rand_pair(n), rand_pair(n), kept_io[io], kept_io[jo]))
        m = rng.random((nch, nch)) < fill
        i1, i2 = np.nonzero(m)
        n = len(i1)
        parts.append((rand_pair(n), rand_pair(n), (nb + i1 + 1) * IBO + (nb + i2 + 1), np.zeros(n)))
        cols = [np.concatenate([p[q] for p in parts]) for q in range(4)]
        rk[(1, k)] = RecordList.build(*cols, rng.uniform(-1, 1, len(cols[0])))
        # jtype 2-5
        for jt in (2, 3, 4, 5):
            parts = []
            if ncp > 0:
                m = rng.random((nch, ncp)) < fill
                ich, ic = np.nonzero(m)
                n = len(ich)
                parts.append((rand_pair(n), rng.integers(1, nb + 1, n), ich + 1, -(ic + 1)))
            if len(kept_io):
                m = rng.random((nch, len(kept_io))) < fill
                ich, io = np.nonzero(m)
                n = len(ich)
                parts.append((rand_pair(n), rng.integers(1, nb + 1, n), ich + 1, kept_io[io]))
            if parts:
                cols = [np.concatenate([p[q] for p in parts]) for q in range(4)]
                rk[(jt, k)] = RecordList.build(*cols, rng.uniform(-1, 1, len(cols[0])))
    # overlaps (S)
    parts, cs = [], []
    if ncp > 0:
        d = np.arange(1, ncp + 1)
        parts.append((-d, -d)); cs.append(np.ones(ncp))
        pi, pj = np.tril_indices(ncp, -1)
        sel = rng.random(len(pi)) < fill * 0.2
        parts.append((-(pi[sel] + 1), -(pj[sel] + 1))); cs.append(0.01 * rng.uniform(-1, 1, int(sel.sum())))
        if len(kept_io):
            m = rng.random((len(kept_io), ncp)) < fill
            io, jc = np.nonzero(m)
            parts.append((kept_io[io], -(jc + 1))); cs.append(0.01 * rng.uniform(-1, 1, len(io)))
    if len(kept_io):
        m = rng.random((len(kept_io), len(kept_io))) < fill
        io, jo = np.nonzero(m)
        parts.append((kept_io[io], kept_io[jo])); cs.append(0.01 * rng.uniform(-1, 1, len(io)))
    m = rng.random((nch, nch)) < fill
    i1, i2 = np.nonzero(m)
    parts.append(((nb + i1 + 1) * IBO + (nb + i2 + 1), np.zeros(len(i1)))); cs.append(rng.uniform(-1, 1, len(i1)))
    k3 = np.concatenate([p[0] for p in parts])
    k4 = np.concatenate([p[1] for p in parts])
    olist = RecordList.build(np.zeros(len(k3)), np.zeros(len(k3)), k3, k4, np.concatenate(cs))
    return PartialWave(name, seed, g, nb, nch, ncp, mk, pbs, iech, lbs, ibort, ntarg, etarg, iptar, ec,
                       olist, rk)


# ---- the workloads ---------------------------------------------------------

def l_lists(w: PartialWave, seed: int, fill: float = 0.2) -> dict:
    """L-integral records (icase 6, L_data.f90:2-149) for a wave, keyed
    (jtype, l) for l = 0..max l of the orbitals, with the structures of
    L_data: jtype 1 a bound pair with HB/HV/HW/<i|j> parts, jtype 2-5 a
    bound orbital with HV/HW parts, jtype 6-9 a channel pair (both orders).
    Values U[-1,1], sorted and merged like the R^k lists."""
    rng = np.random.default_rng(seed)
    nb, nch, ncp, nbf = w.nb, w.nch, w.ncp, w.nbf
    kept = [(i + 1, a + 1) for i in range(nb, nbf) for a in range(nb) if w.ibort[i, a]]
    kept_io = np.array([kl * IBO + nl for kl, nl in kept], np.int64)
    rand_pair = lambda n: (rng.integers(1, nb + 1, n) * IBI + rng.integers(1, nb + 1, n))
    ch_i, ch_j = np.tril_indices(nch)
    out = {}
    for l in range(int(w.lbs.max()) + 1):
        parts = []
        if ncp > 0:
            pi, pj = np.tril_indices(ncp)
            sel = rng.random(len(pi)) < fill
            parts.append((rand_pair(int(sel.sum())), -(pi[sel] + 1), -(pj[sel] + 1)))
            if len(kept_io):
                io, jc = np.nonzero(rng.random((len(kept_io), ncp)) < fill)
                parts.append((rand_pair(len(io)), kept_io[io], -(jc + 1)))
        if len(kept_io):
            io, jo = np.nonzero(rng.random((len(kept_io), len(kept_io))) < fill)
            parts.append((rand_pair(len(io)), kept_io[io], kept_io[jo]))
        i1, i2 = np.nonzero(rng.random((nch, nch)) < fill)
        parts.append((rand_pair(len(i1)), (nb + i1 + 1) * IBO + (nb + i2 + 1), np.zeros(len(i1))))
        k1 = np.concatenate([p[0] for p in parts])
        k3 = np.concatenate([p[1] for p in parts])
        k4 = np.concatenate([p[2] for p in parts])
        out[(1, l)] = RecordList.build(k1, np.zeros(len(k1)), k3, k4, rng.uniform(-1, 1, len(k1)))
        for jt in (2, 3, 4, 5):
            parts = []
            if ncp > 0:
                ich, ic = np.nonzero(rng.random((nch, ncp)) < fill)
                parts.append((rng.integers(1, nb + 1, len(ich)), ich + 1, -(ic + 1)))
            if len(kept_io):
                ich, io = np.nonzero(rng.random((nch, len(kept_io))) < fill)
                parts.append((rng.integers(1, nb + 1, len(ich)), ich + 1, kept_io[io]))
            if parts:
                k2 = np.concatenate([p[0] for p in parts])
                k3 = np.concatenate([p[1] for p in parts])
                k4 = np.concatenate([p[2] for p in parts])
                out[(jt, l)] = RecordList.build(np.zeros(len(k2)), k2, k3, k4, rng.uniform(-1, 1, len(k2)))
        for jt in (6, 7, 8, 9):
            sel = rng.random(len(ch_i)) < fill
            ci, cj = ch_i[sel] + 1, ch_j[sel] + 1
            flip = rng.random(len(ci)) < 0.5
            k1 = np.where(flip & (ci != cj), cj * IBI + ci, ci * IBI + cj)
            z = np.zeros(len(k1))
            out[(jt, l)] = RecordList.build(k1, z, z, z, rng.uniform(-1, 1, len(k1)))
    return out


def c1(seed: int = 1) -> PartialWave:
    """Plain FP64 small: ks=8, ns=60, Z=6, Rmax=20, k=0..4, 20 channels,
    5 target states, 10% fill, dim 1200."""
    return make_wave(seed, 8, 60, 6.0, 20.0, nch=20, ncp=0, mk=4, nb=5, ntarg=5, fill=0.10, name="c1")


def c2(seed: int = 2, nch: int = 60, ncp: int = 400) -> PartialWave:
    """Plain FP64 production: ks=8, ns=200, Z=18, Rmax=40, k=0..10,
    60 channels + 400 perturbers, 5% fill, dim 12400."""
    return make_wave(seed, 8, 200, 18.0, 40.0, nch=nch, ncp=ncp, mk=10, nb=5, ntarg=5, fill=0.05, name="c2")


def c3_specs(seed: int = 3, nwaves: int = 48):
    """Uneven batch: per wave nch ~ U[10,120], ns ~ U[100,250], ks=8,
    k=0..8, per-wave seeds."""
    rng = np.random.default_rng(seed)
    return [dict(seed=int(seed * 1000 + w), nch=int(rng.integers(10, 121)), ns=int(rng.integers(100, 251)))
            for w in range(nwaves)]


def c3_wave(spec, ncp: int = 0) -> PartialWave:
    w = make_wave(spec["seed"], 8, spec["ns"], 18.0, 40.0, nch=spec["nch"], ncp=ncp, mk=8, nb=5, ntarg=5,
                  fill=0.05, name=f"c3-{spec['seed']}")
    w.concat()  # a job arrives with its lists packed
    return w


def c4_stream(seed: int = 4, njobs: int = 60, rate: float = 20.0, ndistinct: int | None = None):
    """Served stream: Poisson arrivals at `rate` jobs/s of partial waves of
    one target drawn around the production wave -- same grid (ks=8, ns=200,
    Z=18, Rmax=40) and bound orbitals, k=0..10, 5% fill, nch ~ U[40,80],
    ncp ~ U[200,400].  The arrivals are a Poisson process conditioned on
    njobs arrivals in the window [0, njobs/rate): njobs sorted uniform
    times, so the realised rate is the specified one.  At most `ndistinct`
    waves are generated and the stream cycles through them (every job is
    assembled anew).  Returns [(arrival_s, PartialWave)]."""
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0.0, njobs / rate, njobs))
    nd = njobs if ndistinct is None else max(1, min(njobs, ndistinct))
    waves = []
    for j in range(nd):
        w = make_wave(seed * 100000 + j, 8, 200, 18.0, 40.0, nch=int(rng.integers(40, 81)),
                      ncp=int(rng.integers(200, 401)), mk=10, nb=5, ntarg=5, fill=0.05,
                      name=f"c4-{j}", target_seed=seed)
        w.concat()  # a request arrives with its lists packed
        waves.append(w)
    return [(float(t[j]), waves[j % nd]) for j in range(njobs)]
