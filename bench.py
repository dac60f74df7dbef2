"""Benchmark: partial-wave assembly of H and S on one MI355X.

Workload (default): one production-size partial wave -- ks=8, ns=200
B-splines on a semi-exponential grid to Rmax=40 a0 (Z=18), multipoles
k=0..10, 60 channels + 400 perturbers, 5 bound orbitals, angular lists at 5%
fill with coefficients U[-1,1]: H and S of dimension 12400 (synthetic, seeded).

A step is the whole job from resident inputs: R^k bands for every k,
orbital contractions, list processing, and the S and H blocks complete on
the device.  value = unique H elements nhm(nhm+1)/2 per second.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_MFMA = 78.6   # TFLOP/s, MI355X dense FP64 matrix
PEAK_HBM = 8000.0       # GB/s


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(seconds_hint=20.0):
    """The oracle (1 thread, the reference's loop structure) on a bounded
    sample of the same workload: same grid, k range, fill and orbital set,
    10 channels and 50 perturbers instead of 60 and 400."""
    from bsr_amd import synth
    from oracle import oracle as O
    w = synth.c2(nch=10, ncp=50)
    t0 = time.perf_counter()
    O.run_wave(w)
    dt = time.perf_counter() - t0
    return dict(value=w.n_elements / dt, unit="elements/s", cores=1, kind="port",
                sample=f"nch=10 ncp=50 ns=200 k<=10 (nhm={w.nhm}); {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mode", choices=["fp64", "split"], default="fp64")
    ap.add_argument("--nch", type=int, default=60)
    ap.add_argument("--ncp", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); partial waves shard as whole jobs "
              "per GPU, so the per-GPU figure is the N=1 line.", file=sys.stderr)
        sys.exit(2)
    from bsr_amd import Context, synth
    log("generating workload")
    w = synth.c2(nch=a.nch, ncp=a.ncp)
    ctx = Context(0)
    if a.mode == "split":
        ctx.set_precision(1, 0)

    def step():
        # a new partial wave: grid and orbitals arrive, bands and bases are
        # rebuilt, lists processed, S and H assembled on the device
        ctx.drop_caches()
        ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
        ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
        ctx.wave_assemble(w, set_inputs=False)

    for i in range(a.warmup):
        step()
        log(f"warmup {i + 1}/{a.warmup}")
    dev, host = [], []
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
        dev.append(ctx.wave_timings())
        host.append(ctx.wave_host_timings())
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / a.steps
    log(f"timed {a.steps} steps: {ms:.3f} ms/step")
    log("device ms: " + json.dumps({k: round(sum(d[k] for d in dev) / len(dev), 3) for k in dev[0]}))
    log("host ms:   " + json.dumps({k: round(sum(d[k] for d in host) / len(host), 3) for k in host[0]}))
    gemm_ms = sum(d["gemm"] for d in dev) / len(dev)
    total_dev = sum(d["total"] for d in dev) / len(dev)
    nss = w.ns * w.ns
    nblk = w.nch * (w.nch + 1) // 2
    nt = (w.mk + 1) * w.nb * w.nb
    flops = 2.0 * nss * nblk * nt
    ach = flops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0
    hbm_bytes = 8.0 * (nss * nt + nt * nblk + nss * nblk)
    out = {
        "metric": "matrix elems/s FP64 vs split-prec @<=1e-10 rel err; p50/p99 partial-wave latency",
        "value": w.n_elements / (ms * 1e-3),
        "unit": "elements/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.mode == "fp64" else "f64(split bf16)",
        "data": "synthetic",
        "config": {"workload": f"partial wave ks=8 ns=200 k<=10 nch={w.nch} ncp={w.ncp} nhm={w.nhm}",
                   "mode": a.mode, "device_ms_per_step": total_dev},
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP64_MFMA, "traffic": None,
                     "kernel": "exchange contraction GEMM", "kernel_ms": gemm_ms,
                     "algorithmic_hbm_GBps": hbm_bytes / (gemm_ms * 1e-3) / 1e9 if gemm_ms > 0 else 0.0},
    }
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
