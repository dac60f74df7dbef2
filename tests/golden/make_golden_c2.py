"""Write tests/golden/c2_full_sample.npz: the oracle's S and H of the full
production wave (synth.c2(): ks=8, ns=200, k=0..10, 60 channels, 400
perturbers, 5% fill, nhm = 12 400), summarised so that the file stays
small:

  * the Frobenius norm of every channel block of S and H (1830 each), of
    every channel-perturber vector (60 x 400), of the perturber block and
    of the ACF (by multipole);
  * htarg and otarg in full;
  * a seeded sample of elements (positions and values) of the channel
    blocks, the channel-perturber vectors, the perturber block and ACF;
  * the maximum |element| of each array (the floor of the relative bound).

The oracle runs the reference's loop structure on one core (about ten
minutes).  Usage: python tests/golden/make_golden_c2.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

OUT = os.path.join(HERE, "c2_full_sample.npz")


def summarize(M, rng, nsamp):
    ns, nch, ncp = M.ns, M.nch, M.ncp
    out = {}
    out["hcc_norm"] = np.sqrt(np.einsum("xyb,xyb->b", M.hcc, M.hcc))
    out["hcc_max"] = np.abs(M.hcc).max()
    idx = rng.integers(0, M.hcc.size, nsamp)
    out["hcc_idx"] = idx.astype(np.int64)
    out["hcc_val"] = M.hcc.ravel(order="F")[idx]
    if ncp > 0:
        out["hcb_norm"] = np.sqrt(np.einsum("xcp,xcp->cp", M.hcb, M.hcb))
        out["hcb_max"] = np.abs(M.hcb).max()
        idx = rng.integers(0, M.hcb.size, nsamp // 4)
        out["hcb_idx"] = idx.astype(np.int64)
        out["hcb_val"] = M.hcb.ravel(order="F")[idx]
        out["hbb_norm"] = np.linalg.norm(M.hbb)
        out["hbb_max"] = np.abs(M.hbb).max()
        idx = rng.integers(0, M.hbb.size, nsamp // 4)
        out["hbb_idx"] = idx.astype(np.int64)
        out["hbb_val"] = M.hbb.ravel(order="F")[idx]
    out["acf_norm"] = np.sqrt(np.einsum("ijk,ijk->k", M.acf, M.acf))
    out["acf_max"] = np.abs(M.acf).max()
    idx = rng.integers(0, M.acf.size, nsamp // 4)
    out["acf_idx"] = idx.astype(np.int64)
    out["acf_val"] = M.acf.ravel(order="F")[idx]
    out["htarg"] = M.htarg.copy()
    out["otarg"] = M.otarg.copy()
    return out


def main():
    from bsr_amd import synth
    from oracle import oracle as O
    w = synth.c2()
    t0 = time.perf_counter()
    S, H = O.run_wave(w)
    dt = time.perf_counter() - t0
    rng = np.random.default_rng(20)
    data = {"seconds": dt, "nhm": w.nhm, "records": w.n_records()}
    for name, M in (("S", S), ("H", H)):
        for key, v in summarize(M, rng, 20000).items():
            data[f"{name}_{key}"] = v
    np.savez_compressed(OUT, **data)
    print(f"{OUT}: oracle {dt:.0f} s, {os.path.getsize(OUT) / 1e6:.2f} MB", flush=True)


if __name__ == "__main__":
    main()
