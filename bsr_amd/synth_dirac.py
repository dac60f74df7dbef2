"""Seeded synthetic Dirac partial waves in the reference's input formats.

A Dirac partial wave is what DBSR_MAT3's SUB1 holds for the R^k pass: the
knot grid with the orders of the large (P, order ksp) and small (Q, order
ksq) component B-splines, the two-component orbitals pq(ns,2,nbf) with their
channel pointers ipbs (0 for bound orbitals), the <.|p> vectors V_ch(ms,nv_ch)
keyed by sorted (i_ch, j_ch) = (channel orbital, bound orbital), the target
pointers iptar, and the sorted R^k lists of module c_data for itype 1..4 and
every multipole k.  K1..K4 are packed as src/dbsr/DBSR_MAT3_MPI/
idef_type.f90:168-186 describes (ibo = 2**15):

  itype 1  K1=(i1,i3) K2=(i2,i4) bound pairs; K3,K4 = -ic,-jc (perturber
           pair), io,-ic (overlap, perturber), io,jo (two overlaps) or io,0
           (channel-orbital pair, target structure)
  itype 2  K1=(j1,j2) bound pair, K2 bound orbital, K3 channel, K4 = io
           (overlap) or -ic (perturber)
  itype 3  K1=(ich,jch) ich>=jch, K2<=K3 bound orbitals    (direct)
  itype 4  K1=(ich,jch) ich>=jch, K2,K3 bound orbitals      (exchange)

Workload definition (one draw of numpy.random.default_rng(seed)):
  grid        ks-fold knots at 0 and rmax, nv-1 interior knots geometric
              from r0 = 0.05/z to rmax
  bound       nb orbitals; P coefficients N(0,1) times a bump, Q = 0.1 x
              N(0,1) times the same bump; coefficient 1 and the coefficients
              past each component's function count are 0
  channels    one channel orbital per channel (ipbs = ich), random pq;
              target of a channel ~ U{1..ntarg}
  V_ch        each (channel orbital, bound orbital) pair w.p. 0.7, a
              vector 0.1 x N(0,1) of length ms = 2 ns
  itype 3/4   every channel pair (i>=j) x bound pair (a<=b for itype 3,
              ordered for itype 4) x k, w.p. fill, c ~ U[-1,1]
  itype 1/2   records per structure w.p. fill with random bound orbitals
              (perturber structures only when npert > 0)
Duplicates (equal K1..K4) are merged by summing, as Merge_cdata does.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .synth import IBO, RecordList


@dataclass
class DiracList:
    itype: int
    kpol: int
    k1: np.ndarray
    k2: np.ndarray
    k3: np.ndarray
    k4: np.ndarray
    c: np.ndarray

    def __len__(self):
        return len(self.c)


@dataclass
class DiracWave:
    name: str
    seed: int
    ks: int
    ns: int
    ksp: int
    ksq: int
    t: np.ndarray              # (ns+ks,)
    nb: int
    nch: int
    npert: int
    mk: int
    ntarg: int
    pq: np.ndarray             # (ns, 2, nbf) Fortran order
    ipbs: np.ndarray           # (nbf,)
    i_ch: np.ndarray           # (nv_ch,) channel orbital (1-based)
    j_ch: np.ndarray           # (nv_ch,) bound orbital (1-based)
    v_ch: np.ndarray           # (ms, nv_ch) Fortran order
    iptar: np.ndarray          # (nch,)
    lists: list = field(default_factory=list)
    _cat: tuple | None = None
    # the overlap pass (O_data lists, kpol 0, K3 = K4 = 0), the channel
    # quantum numbers l and kappa, core and target energies
    olists: list = field(default_factory=list)
    lch: np.ndarray | None = None
    kch: np.ndarray | None = None
    ecore: float = 0.0
    etarg: np.ndarray | None = None
    # one-electron matrices hd(ms,ms,nkap) of the channel kappas and the
    # L(ij) records (ich, jch, index into hd, c) that add them to H
    hd: np.ndarray | None = None
    lij: tuple | None = None

    @property
    def ms(self): return 2 * self.ns

    @property
    def nbf(self): return self.pq.shape[2]

    @property
    def nhm(self): return self.nch * self.ms + self.npert

    @property
    def n_elements(self):
        """Unique elements of the symmetric H of the R^k part."""
        return self.nhm * (self.nhm + 1) // 2

    @property
    def n_records(self):
        return sum(len(L) for L in self.lists)

    def concat(self):
        """All lists as one record set: itype, kpol, K1..K4, c (int32/float64)."""
        if self._cat is None:
            ls = [L for L in self.lists if len(L)]
            if not ls:
                z = np.zeros(0, np.int32)
                self._cat = (z, z, z, z, z, z, np.zeros(0))
            else:
                self._cat = (
                    np.concatenate([np.full(len(L), L.itype, np.int32) for L in ls]),
                    np.concatenate([np.full(len(L), L.kpol, np.int32) for L in ls]),
                    *(np.ascontiguousarray(np.concatenate([getattr(L, a) for L in ls]), dtype=np.int32)
                      for a in ("k1", "k2", "k3", "k4")),
                    np.ascontiguousarray(np.concatenate([L.c for L in ls]), dtype=np.float64),
                )
        return self._cat

    def footprint_bytes(self):
        """Device memory of one assembly in FP64: bands of four families,
        exchange bases, channel blocks (quadrant-major and in the output
        layout), channel-perturber vectors, the exchange coefficients and the
        summed bands of the quadrant blocks.  The int8 split form needs no
        quadrant-major copy (half of `blocks`) but its slices of the exchange
        bases instead."""
        ns, ks, nk, nb = self.ns, self.ks, self.mk + 1, self.nb
        nblk = self.nch * (self.nch + 1) // 2
        bands = 4 * nk * ns * ns * ks * ks
        exch = 4 * nk * nb * nb * ns * ns
        blocks = 2 * nblk * self.ms * self.ms
        vecs = 2 * self.ms * self.nch * self.npert
        coef = 4 * nk * nb * nb * nblk
        bsum = 4 * nblk * ns * ks
        return 8 * (bands + exch + blocks + vecs + coef + bsum)


def dirac_grid(ks: int, ns: int, z: float, rmax: float) -> np.ndarray:
    nv = ns - ks + 1
    inner = np.geomspace(0.05 / z, rmax, nv)[:-1]
    return np.concatenate([np.zeros(ks), inner, np.full(ks, rmax)])


def make_dirac_wave(seed: int, ks: int, ns: int, ksp: int, ksq: int, z: float, rmax: float, nb: int, nch: int,
                    npert: int, mk: int, fill: float, ntarg: int = 3, name: str = "dirac") -> DiracWave:
    rng = np.random.default_rng(seed)
    nv = ns - ks + 1
    nsp, nsq = nv + ksp - 1, nv + ksq - 1
    t = dirac_grid(ks, ns, z, rmax)
    nbf = nb + nch
    pq = np.zeros((ns, 2, nbf), order="F")
    x = np.linspace(0.0, 1.0, ns)
    for o in range(nbf):
        bump = np.exp(-((x - rng.uniform(0.05, 0.6)) / rng.uniform(0.05, 0.3)) ** 2)
        pq[:, 0, o] = rng.standard_normal(ns) * bump
        pq[:, 1, o] = 0.1 * rng.standard_normal(ns) * bump
        pq[0, :, o] = 0.0
        pq[nsp:, 0, o] = 0.0
        pq[nsq:, 1, o] = 0.0
    ipbs = np.zeros(nbf, np.int32)
    ipbs[nb:] = np.arange(1, nch + 1)
    iptar = rng.integers(1, ntarg + 1, nch).astype(np.int32)
    ich, jch = [], []
    for c in range(nch):
        for b in range(nb):
            if rng.random() < 0.7:
                ich.append(nb + c + 1)
                jch.append(b + 1)
    nvch = len(ich)
    v_ch = np.asfortranarray(0.1 * rng.standard_normal((2 * ns, nvch)))
    ovl = list(zip(ich, jch))
    w = DiracWave(name, seed, ks, ns, ksp, ksq, t, nb, nch, npert, mk, ntarg, pq, ipbs,
                  np.asarray(ich, np.int32), np.asarray(jch, np.int32), v_ch, iptar)

    def pairs(n):
        """n random bound pairs a <= b, packed a*ibo + b"""
        x = rng.integers(1, nb + 1, (2, n))
        return np.minimum(x[0], x[1]) * IBO + np.maximum(x[0], x[1])

    ov_code = np.asarray([int(i) * IBO + int(j) for i, j in ovl], np.int64)
    nov = len(ovl)
    for k in range(mk + 1):
        for itype in (1, 2, 3, 4):
            parts = []  # (k1, k2, k3, k4) arrays of the records kept

            def keep(n):
                return np.nonzero(rng.random(n) < fill)[0]

            if itype in (3, 4):
                # every channel pair x bound pair, w.p. fill
                ii, jj = np.tril_indices(nch)
                if itype == 3:
                    aa, bb = np.triu_indices(nb)
                else:
                    aa, bb = np.meshgrid(np.arange(nb), np.arange(nb), indexing="ij")
                    aa, bb = aa.ravel(), bb.ravel()
                r, s_ = np.nonzero(rng.random((len(ii), len(aa))) < fill)
                parts.append(((ii[r] + 1) * IBO + (jj[r] + 1), aa[s_] + 1, bb[s_] + 1, np.zeros(len(r), np.int64)))
            elif itype == 1:
                if npert:
                    ic, jc = np.tril_indices(npert)
                    q = keep(len(ic))
                    parts.append((pairs(len(q)), pairs(len(q)), -(ic[q] + 1), -(jc[q] + 1)))
                    o, ic = np.meshgrid(np.arange(nov), np.arange(npert), indexing="ij")
                    o, ic = o.ravel(), ic.ravel()
                    q = keep(len(o))
                    parts.append((pairs(len(q)), pairs(len(q)), ov_code[o[q]], -(ic[q] + 1)))
                p1, p2 = np.tril_indices(nov)
                q = keep(len(p1))
                parts.append((pairs(len(q)), pairs(len(q)), ov_code[p1[q]], ov_code[p2[q]]))
                c1, c2 = np.tril_indices(nch)
                q = keep(len(c1))
                parts.append((pairs(len(q)), pairs(len(q)), (nb + c1[q] + 1) * IBO + (nb + c2[q] + 1),
                              np.zeros(len(q), np.int64)))
            else:
                if npert:
                    ch, ic = np.meshgrid(np.arange(nch), np.arange(npert), indexing="ij")
                    ch, ic = ch.ravel(), ic.ravel()
                    q = keep(len(ch))
                    parts.append((pairs(len(q)), rng.integers(1, nb + 1, len(q)), ch[q] + 1, -(ic[q] + 1)))
                ch, o = np.meshgrid(np.arange(nch), np.arange(nov), indexing="ij")
                ch, o = ch.ravel(), o.ravel()
                q = keep(len(ch))
                parts.append((pairs(len(q)), rng.integers(1, nb + 1, len(q)), ch[q] + 1, ov_code[o[q]]))
            cols = [np.concatenate([p_[i] for p_ in parts]) if parts else np.zeros(0, np.int64) for i in range(4)]
            rl = RecordList.build(*cols, rng.uniform(-1.0, 1.0, len(cols[0])))
            w.lists.append(DiracList(itype, k, rl.k1, rl.k2, rl.k3, rl.k4, rl.c))
    add_overlap_pass(w, fill)
    return w


def add_overlap_pass(w: DiracWave, fill: float) -> None:
    """The overlap lists of the wave (O_data structures, O_data.f90:7-12),
    the channel kappa and l, and the energies, from their own stream
    (default_rng([seed, 1])):

      kappa       ~ U{-4..-1, 1..4}, l = kappa for kappa > 0, -kappa-1 else
      itype 1     every perturber pair ic >= jc w.p. fill, c ~ U[-0.05,0.05],
                  1 on the diagonal
      itype 2     every (overlap, perturber) pair w.p. fill, c ~ U[-0.05,0.05]
      itype 3     every overlap io with itself w.p. 1/2, c ~ U(0, 0.05] (a
                  positive rank-1 term keeps the channel's S block positive
                  definite), and every ordered pair of overlaps of different
                  channels w.p. fill, c ~ U[-0.05,0.05]
      itype 4     every channel-orbital pair (ich >= jch): c = 1 on the
                  diagonal, w.p. fill c ~ U[-0.05,0.05] off it
    Ecore ~ U[-120,-80], Etarg(it) = Ecore + U[0,2]; per distinct kappa a
    one-electron matrix hd: symmetric, N(0,1) within 8 of the diagonal (a
    stand-in for the Dirac operator that spreads the channel spectra), and
    an L(ij) record (ich, ich, hd of its kappa, 1) per channel."""
    rng = np.random.default_rng([w.seed, 1])
    nch, npert, nb = w.nch, w.npert, w.nb
    kap = rng.integers(1, 5, nch) * rng.choice([-1, 1], nch)
    w.kch = kap.astype(np.int32)
    w.lch = np.where(kap > 0, kap, -kap - 1).astype(np.int32)
    w.ecore = float(rng.uniform(-120.0, -80.0))
    w.etarg = w.ecore + rng.uniform(0.0, 2.0, w.ntarg)
    ov = [int(i) * IBO + int(j) for i, j in zip(w.i_ch, w.j_ch)]
    och = [int(w.ipbs[int(i) - 1]) for i in w.i_ch]

    def small(n):
        return rng.uniform(-0.05, 0.05, n)

    lists = []
    # itype 1: perturber pairs
    k1, k2, c = [], [], []
    for ic in range(1, npert + 1):
        for jc in range(1, ic + 1):
            if ic == jc:
                k1.append(ic), k2.append(jc), c.append(1.0)
            elif rng.random() < fill:
                k1.append(ic), k2.append(jc), c.append(float(small(1)[0]))
    lists.append((1, np.asarray(k1, np.int64), np.asarray(k2, np.int64), np.asarray(c)))
    # itype 2: overlap x perturber
    k1, k2 = [], []
    for o in ov:
        for ic in range(1, npert + 1):
            if rng.random() < fill:
                k1.append(o), k2.append(ic)
    lists.append((2, np.asarray(k1, np.int64), np.asarray(k2, np.int64), small(len(k1))))
    # itype 3: overlap x overlap
    k1, k2, c = [], [], []
    for a, (o, ca) in enumerate(zip(ov, och)):
        if rng.random() < 0.5:
            k1.append(o), k2.append(o), c.append(float(rng.uniform(1e-3, 0.05)))
        for o2, cb in zip(ov, och):
            if cb != ca and rng.random() < fill:
                k1.append(o), k2.append(o2), c.append(float(small(1)[0]))
    lists.append((3, np.asarray(k1, np.int64), np.asarray(k2, np.int64), np.asarray(c)))
    # itype 4: channel-orbital pairs (target structure)
    k1, c = [], []
    for ich in range(1, nch + 1):
        for jch in range(1, ich + 1):
            if ich == jch:
                k1.append((nb + ich) * IBO + nb + jch), c.append(1.0)
            elif rng.random() < fill:
                k1.append((nb + ich) * IBO + nb + jch), c.append(float(small(1)[0]))
    lists.append((4, np.asarray(k1, np.int64), np.zeros(len(k1), np.int64), np.asarray(c)))
    kaps = sorted(set(int(k) for k in w.kch))
    ms = w.ms
    hd = np.zeros((ms, ms, len(kaps)), order="F")
    band = np.abs(np.subtract.outer(np.arange(ms), np.arange(ms))) <= 8
    for m in range(len(kaps)):
        x = rng.standard_normal((ms, ms)) * band
        hd[:, :, m] = (x + x.T) / 2
    w.hd = hd
    ich = np.arange(1, nch + 1, dtype=np.int32)
    w.lij = (ich, ich.copy(), np.asarray([kaps.index(int(k)) + 1 for k in w.kch], np.int32), np.ones(nch))
    for itype, k1, k2, c in lists:
        rl = RecordList.build(k1, k2, np.zeros(len(c)), np.zeros(len(c)), c)
        w.olists.append(DiracList(itype, 0, rl.k1, rl.k2, rl.k3, rl.k4, rl.c))


def small(seed: int = 11, npert: int = 3) -> DiracWave:
    """Parity case: ks=6, ns=24 (P order 5, Q order 6), 3 bound orbitals,
    5 channels, k=0..2, fill 0.3."""
    return make_dirac_wave(seed, 6, 24, 5, 6, 10.0, 20.0, 3, 5, npert, 2, 0.3, name="dirac_small")


def c5(seed: int = 5, nch: int = 150, nb: int = 5, fill: float = 0.03) -> DiracWave:
    """Dirac wave of the C5 shape: ks=9, ns=150 per component (P order 8,
    Q order 9), Z=54, rmax=30, k=0..12, nch=150 (ms=300, nhm=45 000), no
    perturbers, 5 bound orbitals, fill 3%."""
    return make_dirac_wave(seed, 9, 150, 8, 9, 54.0, 30.0, nb, nch, 0, 12, fill, name="dirac_c5")
