"""Stream file describing one partial wave, read by the Fortran driver
bsr_amd/fortran/bsr_mat_gpu.f90 (little endian, no record markers):

  int32  ks ns ml me nbf nch ncp mk ntarg nlists novl
  f64    h ec eps_c eps_ovl
  f64    t(ns+ks) pbs(ns,nbf); int32 iech(nbf) lbs(nbf) ibort(nbf,nbf) iptar(nch); f64 etarg(ntarg)
  int32  ok3(novl) ok4(novl); f64 oc(novl)
  int32  nllists kclosd iitar orth; f64 z
  nllists x { int32 jtype l n; int32 k1(n) k2(n) k3(n) k4(n); f64 c(n) }
  nlists x { int32 jtype kpol n; int32 k1(n) k2(n) k3(n) k4(n); f64 c(n) }

The L-integral lists (SUB1's icase 6, after Gen_Lval with the closed core
1..kclosd and nuclear charge z) come before the R^k lists; both appear in
call order (l or k outer, jtype inner, state_res.f90:155-159); orth = 1 applies BS_ORTH after the final
symmetrisation.
"""
from __future__ import annotations

import numpy as np


def write(path, w, l_lists: dict | None = None, z: float = 0.0, kclosd: int = 0, iitar: int = 0,
          orth: bool = False) -> None:
    # State_res generates jpol (here l) outer, jtype inner (state_res.f90:155-159)
    llists = sorted(((key, lst) for key, lst in (l_lists or {}).items() if len(lst)), key=lambda kv: (kv[0][1], kv[0][0]))
    lists = [((jt, k), w.rk_lists[(jt, k)]) for k in range(w.mk + 1) for jt in range(1, 10)
             if (jt, k) in w.rk_lists and len(w.rk_lists[(jt, k)])]
    i4 = lambda a: np.asarray(a, "<i4").tobytes(order="F")
    f8 = lambda a: np.asarray(a, "<f8").tobytes(order="F")
    with open(path, "wb") as f:
        f.write(i4([w.ks, w.ns, w.ml, w.me, w.nbf, w.nch, w.ncp, w.mk, w.ntarg, len(lists), len(w.olist)]))
        f.write(f8([w.h, w.ec, w.eps_c, w.eps_ovl]))
        f.write(f8(w.t))
        f.write(f8(np.asfortranarray(w.pbs)))
        f.write(i4(w.iech) + i4(w.lbs) + i4(np.asfortranarray(w.ibort)) + i4(w.iptar))
        f.write(f8(w.etarg))
        f.write(i4(w.olist.k3) + i4(w.olist.k4) + f8(w.olist.c))
        f.write(i4([len(llists), kclosd, iitar, 1 if orth else 0]) + f8([z]))
        for (jt, l), lst in llists:
            f.write(i4([jt, l, len(lst)]))
            f.write(i4(lst.k1) + i4(lst.k2) + i4(lst.k3) + i4(lst.k4) + f8(lst.c))
        for (jt, k), lst in lists:
            f.write(i4([jt, k, len(lst)]))
            f.write(i4(lst.k1) + i4(lst.k2) + i4(lst.k3) + i4(lst.k4) + f8(lst.c))
