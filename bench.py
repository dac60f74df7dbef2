"""Benchmark: partial-wave assembly of H and S on one MI355X.

Headline workload: one production-size partial wave -- ks=8, ns=200
B-splines on a semi-exponential grid to Rmax=40 a0 (Z=18), multipoles
k=0..10, 60 channels + 400 perturbers, 5 bound orbitals, angular lists at 5%
fill with coefficients U[-1,1]: H and S of dimension 12400 (synthetic,
seeded).  A step is the whole job from inputs resident in HBM (grid,
orbitals, and the R^k lists uploaded once): R^k bands for every k, orbital
contractions, the lists decoded into terms on the device, and the S and H
blocks complete on the device, in plain FP64.  value = unique H elements
nhm(nhm+1)/2 per second.

Beside the headline line's required fields, the JSON line carries:
  roofline the exchange contraction (the step's largest kernel; by default
           the product over each block's nonzero coefficients with the band
           terms in its epilogue) against the HBM bound: algorithmic bytes
           (bases read once, nonzero list, summed bands read once, H written
           once) over its event-timed duration, with the counter traffic of
           the same kernel;
  host_lists  the same step from host lists decoded by the host workers
           (the lists cross the C ABI as host arrays in every step);
  step_hbm the step's algorithmic bytes (bands read once, S and H written
           once, lists, orbitals) over the step time, against 8 TB/s;
  latency  the device-resident step, and the extra time to copy S and H to
           the host and to write bsr_mat.nnn;
  fp64_engines  the same step with the FP64 product forced to the library
           GEMM (0), the MFMA kernel (2) and the default (4);
  eigen    the lowest 10 eigenvalues of H x = E S x of the wave (FP64 on
           the device) and the time of that solve;
  split    the same wave with the exchange contraction in split precision
           (int8 Ozaki slices; split_bf16: bf16 slices), its largest
           relative element error against the FP64 result and (int8, the
           served form) its largest lowest-10 eigenvalue error;
  serving  a Poisson stream (20 jobs/s) of waves of the same target
           (nch ~ U[40,80], ncp ~ U[200,400]) through the serving loop (6
           slots), in
           FP64 and in split precision: p50/p99 latency at that rate, the
           saturated capacity (all jobs queued at once), and the error of
           every split-precision job against its FP64 result;
  l_data   the one-electron operator with a closed core (Gen_Lval) and the
           L-integral lists of the wave applied call by call (L_data);
  bs_orth  BS_ORTH (the channel orbitals' orthogonality conditions) on the
           wave's H: time per call and the algorithmic HBM rate of the
           blocks it rewrites;
  batch    48 uneven waves (nch ~ U[10,120], ns ~ U[100,250], k=0..8) as one
           hipGraph: total elements/s of a replay, per-wave kernel time;
  dirac    the Dirac (DBSR) R^k assembly of one C5-shaped wave: ks=9, ns=150
           per component (P order 8, Q order 9), Z=54, k=0..12, 150 channels
           (ms=300, nhm=45 000), 5 bound orbitals, lists at 3% fill: the four
           R^k band families, their orbital contractions, list processing and
           hch/acf/htarg complete on the device, FP64; elements/s of H, the
           device phases, the exchange GEMM against the FP64 MFMA peak, and
           the oracle on an 8-channel sample of the same shape (time and the
           largest element error of the GPU result on that sample); and a
           stream of 6 C5-shaped waves served on 2 slots, split precision
           (7 int8 slices) and FP64, against the plain FP64 wave;
  cpu_baseline  the oracle (the reference's loop structure, FP64) on a
           10-channel, 50-perturber sample of the headline wave, on all host
           cores (processes owning blocks round robin, as BSR_MAT3_MPI
           ranks do) and on one core, and on the full C1 wave.
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
PEAK_FP64_VECTOR = 78.6  # TFLOP/s, MI355X FP64 vector
PEAK_HBM = 8000.0       # GB/s
PEAK_INT8_MFMA = 5000.0  # TOP/s, MI355X dense int8 matrix
PEAK_BF16_MFMA = 2500.0  # TFLOP/s, MI355X dense bf16 matrix
# HBM bytes of one launch of the sparse exchange product with the band terms
# in its epilogue (k_xasm<true>) at the headline shape: two rocprofv3 counter
# runs (FETCH_SIZE, WRITE_SIZE) scaled by a 1 GiB device copy in the same
# process (read 515.7 MB, write 585.6 MB)
XPROD_TRAFFIC_BYTES = 1101.3e6


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def _owner_split(w, nparts):
    """The wave's records dealt to nparts owners by output location, round
    robin over channel-channel blocks as BSR_MAT3_MPI deals them
    (src/bsr/BSR_MAT3_MPI/mod_intmat_mpi.f90:62-67), channel-perturber
    vectors and perturber pairs likewise; one sub-wave per owner."""
    import copy
    import numpy as np
    from bsr_amd import synth
    IBI, IBO = synth.IBI, synth.IBO
    iech = np.asarray(w.iech)
    nblk = w.nch * (w.nch + 1) // 2

    def blk(i, j):
        i, j = np.maximum(i, j), np.minimum(i, j)
        return i * (i - 1) // 2 + j - 1

    def owner_rk(jt, L):
        if jt >= 6:
            return blk(L.k1 // IBI, L.k1 % IBI)
        if jt >= 2:
            ch = L.k3
            jch = np.where(L.k4 > 0, iech[np.maximum(L.k4 // IBO, 1) - 1], 0)
            return np.where(L.k4 > 0, blk(ch, np.maximum(jch, 1)), nblk + (ch - 1) * max(w.ncp, 1) - L.k4 - 1)
        return _owner_struct1(L.k3, L.k4)

    def _owner_struct1(k3, k4):
        ic, jc = -k3, -k4
        out = np.zeros(len(k3), np.int64)
        pp = (ic > 0) & (jc > 0)
        out[pp] = nblk + w.nch * max(w.ncp, 1) + np.maximum(ic[pp], jc[pp])
        cv = (ic < 0) & (jc > 0)
        out[cv] = nblk + (iech[(-ic[cv]) // IBO - 1] - 1) * max(w.ncp, 1) + jc[cv] - 1
        cc = (ic < 0) & (jc < 0)
        out[cc] = blk(iech[(-ic[cc]) // IBO - 1], iech[(-jc[cc]) // IBO - 1])
        tg = (ic < 0) & (jc == 0)
        out[tg] = blk(iech[(-ic[tg]) // IBO - 1], iech[(-ic[tg]) % IBO - 1])
        return out

    parts = []
    for r in range(nparts):
        sub = copy.copy(w)
        sub.__dict__.pop("_packed", None)
        sub.rk_lists = {}
        for key, L in w.rk_lists.items():
            sel = owner_rk(key[0], L) % nparts == r
            sub.rk_lists[key] = synth.RecordList(L.k1[sel], L.k2[sel], L.k3[sel], L.k4[sel], L.c[sel])
        o = w.olist
        sel = _owner_struct1(o.k3, o.k4) % nparts == r
        sub.olist = synth.RecordList(o.k1[sel], o.k2[sel], o.k3[sel], o.k4[sel], o.c[sel])
        parts.append(sub)
    return parts


def _oracle_run(job):
    from oracle import oracle as O
    t0 = time.perf_counter()
    O.run_wave(job)
    return time.perf_counter() - t0


def cpu_time(w, cores):
    """Seconds of the oracle (the reference's loop structure, FP64) on wave w:
    one process, or `cores` processes each owning a share of the output
    blocks (every process builds its own R^k bands, as MPI ranks do); the
    wall time of the slowest."""
    if cores <= 1:
        return _oracle_run(w)
    import multiprocessing as mp
    parts = _owner_split(w, cores)
    with mp.get_context("fork").Pool(cores) as pool:
        t0 = time.perf_counter()
        pool.map(_oracle_run, parts)
        return time.perf_counter() - t0


def host_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_baseline():
    """The oracle on a bounded sample of the headline workload (same grid, k
    range, fill and orbitals; 10 channels and 50 perturbers instead of 60
    and 400) on all host cores, block ownership as BSR_MAT3_MPI; beside it
    the same sample on one core, and the full C1 wave (ks=8, ns=60, 20
    channels) on one core and on all cores."""
    from bsr_amd import synth
    cores = host_cores()
    w = synth.c2(nch=10, ncp=50)
    dt_all = cpu_time(w, cores)
    dt_one = cpu_time(w, 1)
    c1 = synth.c1()
    c1_one = cpu_time(c1, 1)
    c1_all = cpu_time(c1, cores)
    return dict(value=w.n_elements / dt_all, unit="elements/s", cores=cores, kind="port",
                sample=f"nch=10 ncp=50 ns=200 k<=10 (nhm={w.nhm}), {cores} processes owning blocks round robin: "
                       f"{dt_all:.2f} s; one core {dt_one:.2f} s",
                one_core_elements_per_s=w.n_elements / dt_one,
                c1_full={"nhm": c1.nhm, "one_core_s": c1_one, "one_core_elements_per_s": c1.n_elements / c1_one,
                         "all_cores_s": c1_all, "all_cores_elements_per_s": c1.n_elements / c1_all,
                         "cores": cores})


def copy_calibration(nbytes):
    """One device-to-device copy of nbytes by the HIP runtime's copy kernel:
    the reference reading for the HBM byte counters (FETCH_SIZE and
    WRITE_SIZE of a kernel that reads and writes nbytes)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(C.byref(p), C.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def mean_dict(ds):
    return {k: round(sum(d[k] for d in ds) / len(ds), 4) for k in ds[0]}


def _orth_conditions(w):
    """Number of BS_ORTH conditions of each channel (bs_orth.f90:40-55):
    bound orbitals of the channel orbital's l with IBORT = 0."""
    import numpy as np
    out = []
    for ich in range(1, w.nch + 1):
        i = int(np.nonzero(np.asarray(w.iech) == ich)[0][0]) + 1
        n = 0
        for j in range(1, w.nbf + 1):
            if w.iech[j - 1] > 0 or w.lbs[j - 1] != w.lbs[i - 1]:
                continue
            if w.ibort[max(i, j) - 1, min(i, j) - 1] == 0:
                n += 1
        out.append(n)
    return out if any(out) else []


def exchange_nonzeros(w):
    """Nonzeros of the exchange coefficient matrix C(t, blk), t = (k, a, b):
    the distinct (block, multipole, ordered bound pair) of the jtype 8/9
    records (I_data.f90:192-216 contracts each group of one channel pair)."""
    import numpy as np
    from bsr_amd import synth
    nb, keys = w.nb, []
    for (jt, k), L in w.rk_lists.items():
        if jt < 8 or len(L) == 0:
            continue
        ich, jch = L.k1 // synth.IBI, L.k1 % synth.IBI
        i, j = np.maximum(ich, jch), np.minimum(ich, jch)
        a, b = L.k2.astype(np.int64) - 1, L.k3.astype(np.int64) - 1
        t = k * nb * nb + np.where(ich >= jch, a * nb + b, b * nb + a)
        keys.append((i * (i - 1) // 2 + j - 1).astype(np.int64) * 100000 + t)
    return int(len(np.unique(np.concatenate(keys)))) if keys else 0


def step_bytes(w):
    """Algorithmic HBM bytes of one wave: every R^k band read once, S and H
    written once, the lists and the orbitals read."""
    ns, ks, nch, ncp = w.ns, w.ks, w.nch, w.ncp
    bands = 8 * ns * ns * ks * ks * (w.mk + 1)
    mats = 2 * 8 * (nch * (nch + 1) // 2 * ns * ns + nch * ncp * ns + ncp * (ncp + 1) // 2)
    return bands + mats + 24 * w.n_records() + 8 * ns * (w.nb + ncp)


def run_steps(ctx, w, warmup, steps, resident=False):
    """resident: the R^k lists are in device memory (ctx.lists_upload(w)
    before) and decoded there; else they are host arrays decoded by the host
    workers within each step."""
    def step():
        # a new partial wave: grid and orbitals arrive, bands and bases are
        # rebuilt, lists processed, S and H assembled on the device
        ctx.drop_caches()
        ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
        ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
        if resident:
            ctx.wave_assemble_resident(w, set_inputs=False)
        else:
            ctx.wave_assemble(w, set_inputs=False)

    for _ in range(warmup):
        step()
    dev, host = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        dev.append(ctx.wave_timings())
        host.append(ctx.wave_host_timings())
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return ms, mean_dict(dev), mean_dict(host)


def serving(stream, precision, window_s, nslots=6):
    """One pass of the stream at its arrival times (window_s: the arrival
    window), then the same jobs all queued at once (capacity)."""
    from bsr_amd.serve import Server, summarize
    srv = Server(nslots=nslots, precision=precision)
    try:
        # warm every slot: grid tables, R^k bands and bases become resident
        warm = [srv.submit(w) for _, w in stream[:nslots]]
        for j in warm:
            j.event.wait()
        d0 = srv.device_list_jobs
        paced = summarize(srv.run_stream(stream))
        d1 = srv.device_list_jobs
        burst = summarize(srv.run_stream([(0.0, w) for _, w in stream]))
        d2 = srv.device_list_jobs
    finally:
        srv.close()
    # the slots decode a job's lists on the device when no job waits, else on the host
    return dict(rate_jobs_per_s=round(len(stream) / window_s, 2),
                device_list_jobs={"paced": d1 - d0, "capacity": d2 - d1, "of": len(stream)},
                p50_ms=paced["p50_ms"],
                p99_ms=paced["p99_ms"], elements_per_s_at_rate=paced["elements_per_s"],
                capacity_elements_per_s=burst["elements_per_s"], capacity_p50_ms=burst["p50_ms"],
                capacity_p99_ms=burst["p99_ms"], service_ms_mean=paced["service_ms_mean"],
                device_ms_mean=paced["device_ms_mean"], failed=paced["failed"] + burst["failed"])


def split_errors(stream):
    """Each distinct wave of the stream in FP64 and in split precision: the
    largest relative element error of S and H (floor 1e-3 of the largest
    element)."""
    from bsr_amd import Context
    ref, spl = Context(0), Context(0)
    spl.set_precision(2, 0)
    errs = []
    seen = set()
    try:
        for _, w in stream:
            if id(w) in seen:
                continue
            seen.add(id(w))
            ref.wave_assemble(w)
            spl.wave_assemble(w)
            errs.append(max(spl.wave_diff(ref, 0), spl.wave_diff(ref, 1)))
    finally:
        ref.close()
        spl.close()
    return errs


def batch_c3(nwaves):
    from bsr_amd import synth
    from bsr_amd.batch import Batch
    waves = [synth.c3_wave(s) for s in synth.c3_specs(nwaves=nwaves)]
    b = Batch(waves)
    try:
        b.prepare()
        out = {}
        # the exchange products of all waves in one persistent launch (work
        # items from one counter, granularity from the autotuner), and one
        # graph branch per wave
        for key, persistent in (("persistent", True), ("per_wave", False)):
            t0 = time.perf_counter()
            b.capture(persistent=persistent)
            cap_s = time.perf_counter() - t0
            b.run()
            reps = [b.run() for _ in range(3)]
            tot = min(r[1] for r in reps)
            out[key] = dict(graph_ms=tot, elements_per_s=b.n_elements() / (tot * 1e-3), capture_s=cap_s)
            if persistent:
                nb, bpi = b.batch_info()
                out[key].update(waves_in_persistent_launch=nb, blocks_per_item=bpi)
        best = min(out.values(), key=lambda d: d["graph_ms"])
        per = sorted(b.eager_ms)
        return dict(waves=len(waves), graph_ms=best["graph_ms"], elements_per_s=best["elements_per_s"],
                    graphs=out, wave_kernel_ms_p50=per[len(per) // 2], wave_kernel_ms_max=per[-1],
                    eager_kernel_ms_sum=sum(per), engines=sorted(set(b.engines())))
    finally:
        b.close()


def rel_err(a, b, floor_frac=1e-3):
    import numpy as np
    a, b = np.asarray(a, float), np.asarray(b, float)
    if b.size == 0:
        return 0.0
    scale = np.abs(b).max()
    if scale == 0.0:
        return float(np.abs(a).max())
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor_frac * scale)))


def dirac_serving(plain_elements_per_s, njobs=6, nslots=2):
    from bsr_amd import synth_dirac as SD
    from bsr_amd.serve import Server, summarize
    waves = [SD.c5(seed=100 + j) for j in range(njobs)]
    out = {"jobs": njobs, "slots": nslots}
    for name, prec, nsl in (("split", 2, 7), ("fp64", 0, 0)):
        srv = Server(nslots=nslots, precision=prec, nslices=nsl)
        try:
            warm = [srv.submit(w) for w in waves[:nslots]]
            for j in warm:
                j.event.wait()
            sm = summarize(srv.run_stream([(0.0, w) for w in waves]))
        finally:
            srv.close()
        out[name] = {"capacity_elements_per_s": sm["elements_per_s"], "ms_per_job": sm["span_s"] * 1e3 / njobs,
                     "failed": sm["failed"]}
    out["split_capacity_vs_fp64_plain"] = out["split"]["capacity_elements_per_s"] / plain_elements_per_s
    return out


def dirac_c5(steps, warmup, with_cpu):
    from bsr_amd import Context, synth_dirac as SD
    w = SD.c5()
    ctx = Context(0)
    try:
        def step():
            # a new Dirac wave: bands of the four families, bases, lists
            ctx.drop_caches()
            ctx.dirac_wave(w)

        for _ in range(warmup):
            step()
        dev = []
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
            dev.append(dict(zip(("bases", "gemm", "terms_layout", "symm_vec", "total", "host_decode", "host_stage", "call"),
                                map(float, ctx.dirac_timings()))))
        ms = (time.perf_counter() - t0) * 1e3 / steps
        dv = mean_dict(dev)
        nblk = w.nch * (w.nch + 1) // 2
        nss = w.ns * w.ns
        nt = (w.mk + 1) * w.nb * w.nb
        flops = 4 * 2.0 * nss * nblk * nt
        eng = ctx.dirac_engine()
        out = {"workload": f"Dirac wave ks=9 ns=150 (ksp=8 ksq=9) k<=12 nch={w.nch} ms={w.ms} nhm={w.nhm} "
                           f"nb={w.nb} fill=3%",
               "records": w.n_records, "elements": w.n_elements, "ms_per_wave": ms,
               "elements_per_s": w.n_elements / (ms * 1e-3), "device_ms": dv, "engine": eng,
               "device_bytes": w.footprint_bytes()}
        if eng == 4:
            # the sparse product writes every quadrant of every channel block
            # once and reads each family's bases once: its algorithmic bytes
            nbytes = 8.0 * (4 * nblk * nss + 4 * nt * nss)
            out["product"] = {"kind": "sparse exchange product over the nonzero coefficients (4 families)",
                              "ms": dv["gemm"], "algorithmic_bytes": nbytes,
                              "GB_s": nbytes / (dv["gemm"] * 1e-3) / 1e9,
                              "frac_hbm_peak": nbytes / (dv["gemm"] * 1e-3) / 1e9 / PEAK_HBM,
                              "dense_equivalent_flops": flops}
        else:
            tf = flops / (dv["gemm"] * 1e-3) / 1e12 if dv["gemm"] > 0 else 0.0
            out["product"] = {"kind": "dense FP64 GEMM (4 families)", "ms": dv["gemm"], "tflops": tf,
                              "frac_fp64_peak": tf / PEAK_FP64_MFMA, "flops": flops}
        # the same wave with the library FP64 GEMM (the dense product) for comparison
        lctx = Context(0)
        try:
            lctx.set_gemm_engine(0)

            def lstep():
                lctx.drop_caches()
                lctx.dirac_wave(w)

            for _ in range(warmup):
                lstep()
            ldev = []
            t0 = time.perf_counter()
            for _ in range(steps):
                lstep()
                ldev.append(dict(zip(("bases", "gemm", "terms_layout", "symm_vec", "total", "host_decode",
                                      "host_stage", "call"), map(float, lctx.dirac_timings()))))
            lms = (time.perf_counter() - t0) * 1e3 / steps
            ldv = mean_dict(ldev)
            ltf = flops / (ldv["gemm"] * 1e-3) / 1e12 if ldv["gemm"] > 0 else 0.0
            out["library_gemm"] = {"engine": lctx.dirac_engine(), "ms_per_wave": lms, "device_ms": ldv,
                                   "gemm_tflops": ltf, "gemm_frac_fp64_peak": ltf / PEAK_FP64_MFMA,
                                   "max_rel_err_vs_sparse": lctx.dirac_diff(ctx),
                                   "sparse_speedup": lms / ms}
        finally:
            lctx.close()
        # the same wave with the exchange GEMMs in split precision (int8
        # slices; 7 of them keep Dirac waves within 1e-10), and its largest
        # element error against the FP64 result
        sctx = Context(0)
        try:
            sctx.set_precision(2, 7)

            def sstep():
                sctx.drop_caches()
                sctx.dirac_wave(w)

            for _ in range(warmup):
                sstep()
            sdev = []
            t0 = time.perf_counter()
            for _ in range(steps):
                sstep()
                sdev.append(dict(zip(("bases", "gemm", "terms_layout", "symm_vec", "total", "host_decode",
                                      "host_stage", "call"), map(float, sctx.dirac_timings()))))
            sms = (time.perf_counter() - t0) * 1e3 / steps
            out["split"] = {"engine": sctx.dirac_engine(), "ms_per_wave": sms,
                            "elements_per_s": w.n_elements / (sms * 1e-3), "device_ms": mean_dict(sdev),
                            "max_rel_err_vs_fp64": sctx.dirac_diff(ctx), "vs_fp64_plain": ms / sms}
        finally:
            sctx.close()
        # served: a stream of C5-shaped waves (own seeds and orbitals, the
        # same grid) queued at once on two slots, in split precision and in
        # FP64; capacity against the plain FP64 wave above
        out["serving"] = dirac_serving(out["elements_per_s"])
        if with_cpu:
            from oracle import oracle as O
            ws = SD.c5(nch=8)
            ctx.dirac_wave(ws)
            got = ctx.dirac_get()
            t0 = time.perf_counter()
            ref = O.dirac_assemble(ws)
            dt = time.perf_counter() - t0
            out["sample_max_rel_err_vs_oracle"] = max(rel_err(got[k], ref[k]) for k in
                                                      ("hch", "hcp", "hp", "acf", "htarg"))
            out["cpu_baseline"] = dict(value=ws.n_elements / dt, unit="elements/s", cores=1, kind="port",
                                       sample=f"nch=8 of the same shape (nhm={ws.nhm}, {ws.n_records} records); "
                                              f"{dt:.1f} s")
        return out
    finally:
        ctx.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nch", type=int, default=60)
    ap.add_argument("--ncp", type=int, default=400)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-eigen", action="store_true", help="skip the eigenvalue check of the split forms")
    ap.add_argument("--no-eigen-dense", action="store_true", help="skip the dense eigen-solve beside block Lanczos")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--batch-waves", type=int, default=48)
    ap.add_argument("--no-dirac", action="store_true")
    ap.add_argument("--dirac-only", action="store_true", help="only the Dirac leg (printed as its own JSON line)")
    ap.add_argument("--dirac-steps", type=int, default=3)
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="finish with one device-to-device copy of this many MiB (calibrates HBM byte counters)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {a.gpus}); partial waves shard as whole jobs "
              "per GPU, so the per-GPU figure is the N=1 line.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    from bsr_amd import Context, synth
    if a.dirac_only:
        d = dirac_c5(a.dirac_steps, 1, not a.no_cpu)
        log("dirac: " + json.dumps(d))
        print(json.dumps({"dirac": d}), flush=True)
        return
    log("generating workload")
    w = synth.c2(nch=a.nch, ncp=a.ncp)
    ctx = Context(0)
    # the inputs in HBM before the timed steps: grid, orbitals and the R^k
    # lists (decoded on the device in each step)
    ctx.set_grid(w.ks, w.ns, w.ml, w.me, w.h, w.t)
    ctx.set_orbitals(w.pbs, w.iech, w.lbs, w.ibort, w.eps_ovl)
    ctx.lists_upload(w)
    ms, dev, host = run_steps(ctx, w, a.warmup, a.steps, resident=True)
    log(f"fp64: {a.steps} steps, {ms:.3f} ms/step (lists decoded on the device)")
    # the same step from host lists, decoded by the host workers
    hms, hdev, hhost = run_steps(ctx, w, a.warmup, a.steps)
    log(f"fp64 from host lists: {hms:.3f} ms/step; device ms {json.dumps(hdev)}; host ms {json.dumps(hhost)}")
    terms = {"S": ctx.wave_term_counts(0), "H": ctx.wave_term_counts(1)}
    log("  terms: " + json.dumps(terms))
    log("  device ms: " + json.dumps(dev))
    log("  host ms:   " + json.dumps(host))
    gemm_ms = dev["gemm"]
    engine = ctx.wave_engine()
    nss = w.ns * w.ns
    nblk = w.nch * (w.nch + 1) // 2
    nt = (w.mk + 1) * w.nb * w.nb
    nnz = exchange_nonzeros(w)
    # the exchange contraction with the band terms in its epilogue (the
    # largest kernel of the step): algorithmic bytes = the used exchange bases
    # read once + the nonzero list + the summed band of every block read once
    # + every H block written once
    xbytes = 8.0 * nss * nt + 12.0 * nnz + 8.0 * w.ns * w.ks * nblk + 8.0 * nss * nblk
    ach = xbytes / (gemm_ms * 1e-3) / 1e9 if gemm_ms > 0 else 0.0
    xflops = 2.0 * nss * nnz
    sbytes = step_bytes(w)
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
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"partial wave ks=8 ns=200 k<=10 nch={w.nch} ncp={w.ncp} nhm={w.nhm}",
                   "device_ms_per_step": dev["total"], "device_ms": dev, "host_ms": host, "terms": terms},
        "roofline": {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM, "unit": "GB/s", "frac": ach / PEAK_HBM,
                     "traffic": XPROD_TRAFFIC_BYTES if engine == 4 and (w.nch, w.ncp) == (60, 400) else None,
                     "traffic_source": "rocprofv3 FETCH_SIZE/WRITE_SIZE of k_xasm<true> at this shape, "
                                       "copy-calibrated: profiles/wave_c2_hbm_traffic.md",
                     "kernel": "exchange contraction over the nonzero coefficients with the band terms in its "
                               "epilogue, k_xasm<true> (engine 4)",
                     "engine": engine, "kernel_ms": gemm_ms, "algorithmic_bytes": xbytes,
                     "coefficient_nonzeros": nnz, "coefficient_density": nnz / (nt * nblk),
                     "fp64_flops": xflops,
                     "fp64_tflops": xflops / (gemm_ms * 1e-3) / 1e12 if gemm_ms > 0 else 0.0,
                     "dense_gemm_flops": 2.0 * nss * nblk * nt},
        "step_hbm": {"algorithmic_bytes": sbytes, "GB_s": sbytes / (ms * 1e-3) / 1e9,
                     "frac": sbytes / (ms * 1e-3) / 1e9 / PEAK_HBM,
                     "device_GB_s": sbytes / (dev["total"] * 1e-3) / 1e9},
    }
    # the S and H of the wave to host memory and to a bsr_mat file: the
    # latency a caller sees beyond the device-resident result
    t0 = time.perf_counter()
    ctx.wave_get(0)
    ctx.wave_get(1)
    d2h = (time.perf_counter() - t0) * 1e3
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        t0 = time.perf_counter()
        ctx.wave_write(os.path.join(td, "bsr_mat.001"), w)
        fms = (time.perf_counter() - t0) * 1e3
    out["host_lists"] = {"ms_per_step": hms, "elements_per_s": w.n_elements / (hms * 1e-3), "device_ms": hdev,
                         "host_ms": hhost}
    out["latency"] = {"device_resident_ms": ms, "d2h_ms": d2h, "file_ms": fms}
    log(f"latency: step {ms:.2f} ms, +D2H {d2h:.1f} ms, +file {fms:.1f} ms")
    # the same step with each FP64 product engine forced: library GEMM (0),
    # MFMA kernel with 64-row tiles (2), and the default sparse product (4)
    eng = {}
    for e in (0, 2):
        ctx.set_gemm_engine(e)
        ems, edev, _ = run_steps(ctx, w, 2, max(3, a.steps // 2))
        eng[str(e)] = {"ms_per_step": ems, "product_ms": edev["gemm"], "terms_ms": edev["terms"]}
    ctx.set_gemm_engine(-1)
    eng[str(engine)] = {"ms_per_step": hms, "product_ms": hdev["gemm"], "terms_ms": hdev["terms"]}
    out["fp64_engines"] = eng
    log("engines: " + json.dumps(eng))
    e_ref = None
    if not a.no_eigen:
        # lowest 10 eigenvalues of H x = E S x (FP64 on the device), the
        # reference point of the split forms' eigenvalue error: block
        # Lanczos (the default at this size), then the dense solver on the
        # same matrices as an independent check
        t0 = time.perf_counter()
        e_ref = ctx.wave_eigen(10)
        eig_s = time.perf_counter() - t0
        # again, without the first call's loading of the solver libraries' kernels
        t0 = time.perf_counter()
        e_rep = ctx.wave_eigen(10)
        eig_rep_s = time.perf_counter() - t0
        if not np.array_equal(e_rep, e_ref):
            raise SystemExit("eigen: a repeated solve gave other eigenvalues")
        info = ctx.eigen_info()
        log(f"eigen: lowest {e_ref[:3]} ... in {eig_s:.2f} s, again {eig_rep_s:.2f} s (nhm={w.nhm}, {info})")
        out["eigen"] = {"lowest10": [float(x) for x in e_ref], "solve_s": eig_s, "repeat_solve_s": eig_rep_s,
                        "method": "block Lanczos" if info["method"] == 2 else "dense (sygvdx)",
                        "krylov_dim": info["dim"], "residual_bound": info["resid"]}
        if not a.no_eigen_dense:
            ctx.set_eigen_method(1)
            t0 = time.perf_counter()
            e_dense = ctx.wave_eigen(10)
            dense_s = time.perf_counter() - t0
            ctx.set_eigen_method(0)
            de = float(np.max(np.abs(e_dense - e_ref)))
            out["eigen"]["dense"] = {"solve_s": dense_s, "max_abs_diff": de}
            log(f"eigen dense: {dense_s:.2f} s, max |dE| against Lanczos {de:.2e}")
    if not a.no_split:
        # int8 slices (the split form used below) and bf16 slices
        for key, mode in (("split", 2), ("split_bf16", 1)):
            sctx = Context(0)
            sctx.set_precision(mode, 0)
            sms, sdev, _ = run_steps(sctx, w, a.warmup, a.steps)
            err = max(sctx.wave_diff(ctx, 0), sctx.wave_diff(ctx, 1))
            log(f"{key}: {sms:.3f} ms/step, GEMM {sdev['gemm']:.3f} ms, max rel element error vs FP64 {err:.2e}")
            out[key] = {"value": w.n_elements / (sms * 1e-3), "ms_per_step": sms, "gemm_ms": sdev["gemm"],
                        "engine": sctx.wave_engine(), "max_rel_err_vs_fp64": err, "vs_fp64_plain": ms / sms}
            if sdev["gemm"] > 0:
                # slice products issued by the split GEMM: S(S+1)/2 per level
                # pair over the padded inner size (64 for int8, 32 for bf16)
                S = sctx.wave_engine() - (20 if mode == 2 else 10)
                pad = 64 if mode == 2 else 32
                kp = (nt + pad - 1) // pad * pad
                ops = S * (S + 1) // 2 * 2.0 * nss * ((nblk + 63) // 64 * 64) * kp
                rate = ops / (sdev["gemm"] * 1e-3) / 1e12
                peak = PEAK_INT8_MFMA if mode == 2 else PEAK_BF16_MFMA
                out[key]["mfma"] = {("int8_tops" if mode == 2 else "bf16_tflops"): rate, "peak": peak,
                                    "frac": rate / peak, "counters": "profiles/mfma_utilisation_c2.md"}
            if e_ref is not None and key == "split":
                de = float(np.max(np.abs(sctx.wave_eigen(10) - e_ref)))
                out[key]["max_eig_err_vs_fp64"] = de
                log(f"{key}: lowest-10 eigenvalue error vs FP64 {de:.2e}")
            sctx.close()
    # BS_ORTH on the wave's H (after the legs that compare against it): the
    # channel orbitals' orthogonality conditions; the kernel reads every
    # touched block twice and writes it once
    conds = _orth_conditions(w)
    if conds:
        reps = 5
        ctx.bs_orth(1)  # warm-up
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.bs_orth(1)
        oms = (time.perf_counter() - t0) * 1e3 / reps
        touched = sum(1 for i in range(1, w.nch + 1) for j in range(1, i + 1) if conds[i - 1] or conds[j - 1])
        gbs = 3 * touched * w.ns * w.ns * 8 / (oms * 1e-3) / 1e9
        out["bs_orth"] = {"ms": oms, "channels_with_conditions": sum(1 for q in conds if q),
                          "blocks": touched, "algorithmic_GB_s": gbs}
        log(f"bs_orth: {oms:.3f} ms per call, {touched} blocks, {gbs:.0f} GB/s (2 reads + 1 write per block)")
    # one-electron operator with a closed core of the first two bound
    # orbitals (Gen_Lval) and the L-integral lists of the wave (L_data, call
    # by call into a matrix of the wave's size)
    llists = synth.l_lists(w, 6, fill=0.05)
    nrec_l = sum(len(x) for x in llists.values())
    ctx.allocate_matrix(w.nch, w.ncp, w.mk)
    t0 = time.perf_counter()
    ctx.gen_lval(18.0, 2)
    gl_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    for (jt, l), lst in sorted(llists.items()):
        if len(lst):
            ctx.l_data(jt, l, lst, iitar=1)
    ld_ms = (time.perf_counter() - t0) * 1e3
    out["l_data"] = {"gen_lval_ms": gl_ms, "l_data_ms": ld_ms, "records": nrec_l, "lists": len(llists)}
    log(f"gen_lval: {gl_ms:.2f} ms; l_data: {ld_ms:.2f} ms for {nrec_l} records in {len(llists)} lists")
    ctx.close()
    if not a.no_serve:
        log("generating served stream")
        rate = 20.0
        stream = synth.c4_stream(njobs=a.serve_jobs, rate=rate, ndistinct=60)
        sv = {"jobs": len(stream), "distinct_waves": len({id(w) for _, w in stream}),
              "arrivals": f"Poisson at {rate:g} jobs/s conditioned on {len(stream)} arrivals in {len(stream) / rate:g} s"}
        for name, prec in (("fp64", 0), ("split", 2)):
            sv[name] = serving(stream, prec, len(stream) / rate)
            log(f"serving {name}: " + json.dumps(sv[name]))
        errs = split_errors(stream)
        sv["split"]["max_rel_err_vs_fp64"] = max(errs)
        sv["split"]["rel_err_per_job"] = [float(f"{e:.2e}") for e in errs]
        sv["split_capacity_vs_fp64_plain"] = sv["split"]["capacity_elements_per_s"] / out["value"]
        out["serving"] = sv
    if not a.no_batch:
        log("batch of uneven waves")
        out["batch"] = batch_c3(a.batch_waves)
        log("batch: " + json.dumps(out["batch"]))
    if not a.no_dirac:
        log("dirac wave")
        out["dirac"] = dirac_c5(a.dirac_steps, 1, not a.no_cpu)
        log("dirac: " + json.dumps(out["dirac"]))
    if not a.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline()
    if a.copy_mb:
        copy_calibration(a.copy_mb << 20)
        log(f"copied {a.copy_mb} MiB device to device")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
