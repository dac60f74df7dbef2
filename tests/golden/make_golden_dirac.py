"""Write dirac_vectors.npz: inputs and outputs of the reference's own Dirac
R^k assembly -- R_data (src/dbsr/DBSR_MAT3_MPI/R_data.f90) over the four
families for every list, as Gen_matrix runs it, with the UPDATE_* routines
of updates.f90 and Target_h -- compiled by oracle/ref.mk into
oracle/_ref/libdbs_ref.so and driven through oracle/ref_dbsr_driver.f90.
Run in the container that has the reference checkout; the fixture is
committed, the reference is not.

Waves (bsr_amd.synth_dirac): "p" = small(11, npert=3) and "n" = small(12,
npert=0): ks=6, ns=24, P order 5, Q order 6, 3 bound orbitals, 5 channels,
k = 0..2, all four itypes (HB/HV/HW/Target_h, overlaps, direct, exchange).
Stored per wave: every input the assembly reads (grid, orbitals, V_ch and
keys, sizes, iptar, the concatenated lists) and the outputs hch (every 5th
element), hcp, hp, acf, htarg.

Usage: python tests/golden/make_golden_dirac.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    from bsr_amd import synth_dirac as SD
    from oracle import refbind as R
    R.build()
    out = {}
    for tag, (seed, npert) in {"p": (11, 3), "n": (12, 0)}.items():
        w = SD.small(seed, npert)
        out[f"{tag}_ints"] = np.array([w.ks, w.ns, w.ksp, w.ksq, w.nch, w.npert, w.mk, w.ntarg])
        out[f"{tag}_t"] = w.t
        out[f"{tag}_pq"] = w.pq
        out[f"{tag}_ipbs"] = w.ipbs
        out[f"{tag}_ich"] = w.i_ch
        out[f"{tag}_jch"] = w.j_ch
        out[f"{tag}_vch"] = w.v_ch
        out[f"{tag}_iptar"] = w.iptar
        it, kp, k1, k2, k3, k4, c = w.concat()
        out[f"{tag}_lists"] = np.stack([it, kp, k1, k2, k3, k4])
        out[f"{tag}_c"] = c
        r = R.dirac_assemble(w)
        out[f"{tag}_hch5"] = r["hch"].reshape(-1, order="F")[::5]
        for key in ("hcp", "hp", "acf", "htarg"):
            out[f"{tag}_{key}"] = r[key]
    np.savez_compressed(os.path.join(HERE, "dirac_vectors.npz"), **out)
    print("wrote", os.path.join(HERE, "dirac_vectors.npz"))


if __name__ == "__main__":
    main()
