"""Throughput benchmark of the batched Smith-Waterman engine on one MI355X.

One step = one whole job over the staged batch (unpack, forward pass, 32-bit
reruns, reverse pass, finalize) on inputs already resident in HBM.  Prints one
JSON line.  Headline workload: the length-skewed 4M-pair batch (windows
log-normal with median 400 bp clipped to [150, 1024], reads 150/250 bp split
70/30, 25% of reads with a 20-60 bp random chimeric tail) with mhm2's
1/3/5/2/2 scoring, autotuned, compared with the plain one-launch form on the
same GPU and with the reference CPU aligner (SSW) on the host cores, every
pair checked bit-exact.  The other batch workloads (the 10k plain pairs, the
uneven 1M-pair batch and the 500k overflow pairs) run as secondary legs with
the same checks; local assembly and the serving loop follow.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "batched SW GCUPS + pairs/s, bit-exact vs SSW; p99 latency at fixed QPS"
# the synthetic generator's workload numbers (mhmsw_synth_generate) and what each one is
WORKLOADS = {
    1: ("plain_10k", "plain SW: 10k pairs, 150 bp reads vs 350 bp windows, 1% subs, 0.1% indels"),
    2: ("uneven_1M", "uneven batched launch: 1M pairs, reads U{100..250} vs windows U{150..1024}, 0.5% N, "
                     "2% subs, 0.2% indels, 10% unrelated"),
    3: ("skewed_4M", "length-skewed batch: 4M pairs, windows lognormal(median 400) in [150,1024], reads 150/250 "
                     "(70/30), 25% chimeric tails"),
    4: ("overflow_500k", "overflow-exact: 500k pairs, 250 bp reads vs 1024 bp tandem-repeat windows, "
                         "5% 4000 bp reads"),
}
HEADLINE = 3
VALU_OPS_PER_CELL = 10  # t=H-gO, E, F (2 each), d=Hdiag+s, H=max of 4, best=max(best,H)
NOMINAL_PEAK = 78.6     # 256 CUs x 64 lanes x 2 int16 per lane x 2.4 GHz, T lane-ops/s
SCORING_TEXT = "match 1, mismatch 3, gap open 5, gap extend 2, N 2"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def host_cores():
    """Host cores this job may use: the affinity mask, bounded by a cgroup CPU quota when one is set and
    by the host share of one GPU on the GPU machines (16 cores)."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
        if q != "max":
            quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        quota = None
    cores = min(aff, quota) if quota else aff
    share = 16
    info = {"os_cpu_count": os.cpu_count(), "affinity": aff, "cgroup_quota": quota, "gpu_host_share": share}
    return min(cores, share), info


def spread(n, k):
    """k pair indices spread evenly over [0, n)."""
    if k >= n:
        return np.arange(n, dtype=np.int64)
    return np.unique(np.linspace(0, n - 1, k).astype(np.int64))


def ref_run(batch, scoring, idx, nthreads):
    """Reference SSW (oracle/_ref) or, when absent, the oracle port on the pairs idx; (results, kind, seconds)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib
    kind = "reference" if oracle_lib.have_ref() else "port"
    fn = oracle_lib.ref_batch if kind == "reference" else oracle_lib.oracle_batch
    sel = slice(None) if idx is None else idx
    t = time.perf_counter()
    res = fn(batch.reads, batch.read_off[sel], batch.read_len[sel], batch.ctgs, batch.ctg_off[sel],
             batch.ctg_len[sel], scoring, nthreads)
    return res, kind, time.perf_counter() - t


def cells_of(batch, idx=None):
    sel = slice(None) if idx is None else idx
    return int(np.dot(batch.read_len[sel].astype(np.int64), batch.ctg_len[sel].astype(np.int64)))


def compare(res, ref, idx=None):
    got = np.stack([res[c] for c in ("score", "read_begin", "read_end", "ctg_begin", "ctg_end")], 1)
    if idx is not None:
        got = got[idx]
    want = np.stack([ref[c] for c in ("score", "ref_begin", "ref_end", "query_begin", "query_end")], 1)
    return int((got != want).any(1).sum())


def cpu_legs(batch, scoring, res, cores, one_core_seconds=8.0):
    """cpu_baseline (every pair on all usable host cores, which is also the bit-exact check) and a one-core
    figure on pairs spread over the batch, sized for about one_core_seconds."""
    n = batch.npairs
    ref, kind, dt = ref_run(batch, scoring, None, cores)
    cells = cells_of(batch)
    out = {"cpu_baseline": {"value": round(cells / dt / 1e9, 3), "unit": "GCUPS", "cores": cores, "kind": kind,
                            "pairs_per_s": round(n / dt, 1),
                            "sample": f"all {n} pairs of the same workload ({cells} cells), {dt:.1f} s, one pair "
                                      f"per task over {cores} threads"},
           "bit_exact_check": {"pairs": n, "mismatches": compare(res, ref), "against": kind,
                               "fields": "score, read_begin, read_end, ctg_begin, ctg_end"}}
    # one core: calibrate on a small spread sample, then time one sized for one_core_seconds
    k = min(n, 500)
    idx = spread(n, k)
    _, _, t1 = ref_run(batch, scoring, idx, 1)
    k = int(min(n, max(k, k * one_core_seconds / max(t1, 1e-6))))
    idx = spread(n, k)
    r1, _, t1 = ref_run(batch, scoring, idx, 1)
    c1 = cells_of(batch, idx)
    out["cpu_one_core"] = {"value": round(c1 / t1 / 1e9, 3), "unit": "GCUPS", "cores": 1, "kind": kind,
                           "pairs_per_s": round(len(idx) / t1, 1),
                           "sample": f"{len(idx)} pairs spread over the batch ({c1} cells), {t1:.1f} s",
                           "mismatches": compare(res, r1, idx)}
    return out


def kernel_of(st, k):
    """The kernel that ran forward class k: sw_fast_kernel<lanes per pair, packed registers per lane>."""
    lanes = st["class_lanes"][k]
    if k >= 16 or not lanes:
        return "sw_plain_kernel"
    return f"sw_fast_kernel<{lanes},{st['class_regs'][k]}>"


def class_roofline(st, class_ms, peak_measured, traffic):
    """Roofline of the forward class with the most device time (HIP events around its launch on the context
    stream): achieved = 10 int16 lane-ops per cell of the class / its launch's duration."""
    from mhm2_staging_amd import _native
    cm = np.mean(np.array(class_ms), 0)
    k = int(np.argmax(cm))
    kcells = st["class_cells"][k]
    achieved = VALU_OPS_PER_CELL * kcells / (cm[k] * 1e-3) / 1e12
    name = _native.class_name(k)
    rl = {"bound": "valu", "kernel": kernel_of(st, k), "class": name, "achieved": round(achieved, 2),
          "peak": NOMINAL_PEAK,
          "unit": "TOP/s (int16 lane-ops)", "frac": round(achieved / NOMINAL_PEAK, 4), "traffic": None,
          "kernel_ms": round(float(cm[k]), 3), "kernel_cells": int(kcells), "kernel_pairs": int(st["class_pairs"][k]),
          "ops_per_cell": VALU_OPS_PER_CELL, "h_scale_bits": int(st["class_tag_bits"][k])}
    if peak_measured:
        rl["peak_measured"] = round(peak_measured["lane_ops_per_s"] / 1e12, 2)
        rl["frac_of_measured"] = round(achieved * 1e12 / peak_measured["lane_ops_per_s"], 4)
    if traffic and name in traffic.get("per_launch_bytes", {}):
        rl["traffic"] = traffic["per_launch_bytes"][name]
        rl["traffic_source"] = traffic["source"]
    classes = {_native.class_name(i): {"pairs": st["class_pairs"][i], "cells": st["class_cells"][i],
                                       "ms": round(float(cm[i]), 3), "kernel": kernel_of(st, i),
                                       "h_scale_bits": int(st["class_tag_bits"][i])}
               for i in range(_native.NUM_CLASSES) if st["class_pairs"][i]}
    return rl, classes


def load_traffic(key):
    """HBM bytes per launch of the forward classes of a workload, from rocprofv3 PMC runs under profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        tr = json.load(f)
    return tr.get("workloads", {}).get(key)


def timed_jobs(al, steps):
    """Time `steps` jobs on the staged batch (eager launches, so every kernel's HIP events are live)."""
    job_ms, class_ms = [], []
    t = time.perf_counter()
    for _ in range(steps):
        al.align_staged()
        al.wait()
        st = al.stats()
        job_ms.append(st["job_ms"])
        class_ms.append(st["class_ms"])
    return time.perf_counter() - t, st, job_ms, class_ms


def plain_leg(al, batch, res, gcups):
    """The plain form (one 32-bit launch per batch, one wavefront per pair) on the same staged batch."""
    import mhm2_staging_amd as m
    al.set_mode(m.MODE_PLAIN)
    al.align_staged()
    al.wait()
    t = time.perf_counter()
    al.align_staged()
    al.wait()
    pe = time.perf_counter() - t
    pres = al.results()
    al.set_mode(m.MODE_FAST)
    plain = batch.cells() / pe / 1e9
    return {"value": round(plain, 2), "unit": "GCUPS", "ms_per_step": round(pe * 1e3, 3),
            "fast_over_plain": round(gcups / plain, 2), "identical_results": bool((pres == res).all())}


def batch_leg(cfg, args, cores, peak, headline=False):
    """One batch workload: autotuned fast form (timed), plain form, CPU legs and the bit-exact check."""
    import mhm2_staging_amd as m
    key, text = WORKLOADS[cfg]
    scoring = m.MHM2_SCORING
    steps = args.steps if headline else max(3, min(args.steps, 10))
    t0 = time.perf_counter()
    batch = m.generate(cfg, npairs=args.npairs if headline else 0)
    log(f"[bench] {key}: generated {batch.npairs} pairs, {batch.cells():.3e} cells in {time.perf_counter() - t0:.1f}s")
    al = m.Aligner(device=0, scoring=scoring, read_is_ssw_ref=True, mode=m.MODE_FAST)
    # eager launches in the timed region so each kernel's HIP events are live
    # (a replayed job graph reports only its total time); the graph-replay
    # rate of the same job is reported separately below
    al.set_graphs(False)
    t0 = time.perf_counter()
    al.stage(batch)
    stage_s = time.perf_counter() - t0
    out = {"workload": text, "npairs": batch.npairs, "cells": batch.cells(), "stage_s": round(stage_s, 3)}
    default_tuning = al.get_tuning()
    if headline and not args.tuning and not args.no_autotune:
        # the default scheduling parameters, timed before autotuning changes them
        for _ in range(args.warmup):
            al.align_staged()
            al.wait()
        el, st, _, _ = timed_jobs(al, steps)
        out["default_tuning"] = {"tuning": default_tuning, "value": round(batch.cells() * steps / el / 1e9, 2),
                                 "unit": "GCUPS", "ms_per_step": round(el / steps * 1e3, 3)}
    if headline and args.tuning:
        b_, g_, s_ = (int(x) for x in args.tuning.split(","))
        al.set_tuning(blocks_per_cu=b_, g64_min_rows=g_, bin_shift=s_)
    elif not args.no_autotune:
        t0 = time.perf_counter()
        al.autotune(repeats=2)
        log(f"[bench] {key}: autotuned {al.get_tuning()} in {time.perf_counter() - t0:.1f}s")
    out["tuning"] = al.get_tuning()
    for _ in range(args.warmup):
        al.align_staged()
        al.wait()
    el, st, job_ms, class_ms = timed_jobs(al, steps)
    res = al.results()
    cells = batch.cells()
    gcups = cells * steps / el / 1e9
    log(f"[bench] {key}: fast form {gcups:.1f} GCUPS, {el / steps * 1e3:.2f} ms/step, job {st['job_ms']:.2f} ms")
    if "default_tuning" in out:
        out["default_tuning"]["autotuned_over_default"] = round(gcups / out["default_tuning"]["value"], 3)
    rl, classes = class_roofline(st, class_ms, peak, load_traffic(key))
    out.update({"value": round(gcups, 2), "unit": "GCUPS", "steps": steps,
                "ms_per_step": round(el / steps * 1e3, 3), "pairs_per_s": round(batch.npairs * steps / el, 1),
                "device_job_ms": round(float(np.mean(job_ms)), 3),
                "stats": {kk: v for kk, v in st.items() if not kk.startswith("class_")},
                "classes": classes, "roofline": rl})
    if headline:
        al.set_graphs(True)
        for _ in range(2):  # a geometry's first job runs eagerly, the second is captured
            al.align_staged()
            al.wait()
        t = time.perf_counter()
        for _ in range(steps):
            al.align_staged()
            al.wait()
        ge = time.perf_counter() - t
        out["graph_replay"] = {"value": round(cells * steps / ge / 1e9, 2), "unit": "GCUPS",
                               "ms_per_step": round(ge / steps * 1e3, 3),
                               "identical_results": bool((al.results() == res).all())}
        al.set_graphs(False)
        # host-inclusive rate: host strings in, host results out (2-bit packing on
        # the host, PCIe both ways, the job); never the headline value
        t = time.perf_counter()
        e2e_steps = 2
        for _ in range(e2e_steps):
            al.stage(batch)
            al.align_staged()
            al.wait()
            eres = al.results()
        ee = time.perf_counter() - t
        out["end_to_end"] = {"value": round(cells * e2e_steps / ee / 1e9, 2), "unit": "GCUPS",
                             "ms_per_step": round(ee / e2e_steps * 1e3, 3),
                             "includes": "host 2-bit packing, H2D copy, job, D2H of results",
                             "identical_results": bool((eres == res).all())}
        log(f"[bench] {key}: end to end {out['end_to_end']}")
    if not args.no_plain:
        out["plain_form"] = plain_leg(al, batch, res, gcups)
        log(f"[bench] {key}: plain form {out['plain_form']}")
    if headline and not args.no_cigar:
        out["cigar_pass"] = cigar_leg(al, batch, res, scoring, steps, not args.no_cpu, cores)
        log(f"[bench] {key}: cigar pass {out['cigar_pass']}")
    al.close()
    if not args.no_cpu:
        out.update(cpu_legs(batch, scoring, res, cores, 8.0 if headline else 3.0))
        log(f"[bench] {key}: cpu {out['cpu_baseline']} one core {out['cpu_one_core']['value']} GCUPS, "
            f"bit-exact {out['bit_exact_check']}")
    return out


def cigar_leg(al, batch, res, scoring, steps, check, cores):
    """Jobs with the CIGAR pass on (SSW's CIGAR + mismatches, klign with report_cigar): its device time,
    the job's rate, and -- when check -- the reference SSW with report_cigar on host cores on a bounded
    sample spread over the batch, every GPU CIGAR of that sample compared with the reference's."""
    al.set_cigar(True)
    al.align_staged()
    al.wait()
    cms, jms = [], []
    t = time.perf_counter()
    for _ in range(steps):
        al.align_staged()
        al.wait()
        st = al.stats()
        cms.append(st["cigar_ms"])
        jms.append(st["job_ms"])
    el = time.perf_counter() - t
    cres = al.results()
    mism, off, ops = al.cigars()
    al.set_cigar(False)
    cells = batch.cells()
    out = {"cigar_ms": round(float(np.mean(cms)), 3), "job_ms": round(float(np.mean(jms)), 3),
           "value": round(cells * steps / el / 1e9, 2), "unit": "GCUPS (jobs with the CIGAR pass)",
           "pairs_per_s": round(batch.npairs * steps / el, 1),
           "cigar_pairs_per_s_device": round(batch.npairs / (float(np.mean(cms)) * 1e-3), 1),
           "ops": int(len(ops)), "reruns": int(st["cigar_reruns"]),
           "identical_results": bool((cres == res).all())}
    if check:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        if oracle_lib.have_ref():
            n = batch.npairs

            def run(idx, maxops):
                return oracle_lib.ref_cigar_batch(batch.reads, batch.read_off[idx], batch.read_len[idx], batch.ctgs,
                                                  batch.ctg_off[idx], batch.ctg_len[idx], scoring, maxops=maxops,
                                                  nthreads=cores)
            idx = spread(n, min(n, 2000))
            t = time.perf_counter()
            run(idx, 64)
            dt = max(time.perf_counter() - t, 1e-6)
            idx = spread(n, int(min(n, max(len(idx), len(idx) * 10.0 / dt))))
            nops_gpu = np.diff(off)[idx]
            maxops = int(nops_gpu.max(initial=0)) + 8
            t = time.perf_counter()
            rres, rmism, rnops, rops = run(idx, maxops)
            dt = time.perf_counter() - t
            kc = cells_of(batch, idx)
            bad = (nops_gpu != rnops) | (mism[idx] != rmism)
            for q in np.nonzero(~bad)[0]:
                p = idx[q]
                if not np.array_equal(ops[off[p]:off[p + 1]], rops[q, :rnops[q]]):
                    bad[q] = True
            out["cpu_reference"] = {"value": round(kc / dt / 1e9, 3), "unit": "GCUPS", "cores": cores,
                                    "pairs_per_s": round(len(idx) / dt, 1),
                                    "sample": f"{len(idx)} of {n} pairs spread over the batch, reference SSW "
                                              f"with report_cigar, {dt:.1f} s"}
            out["bit_exact_check"] = {"pairs": int(len(idx)), "mismatches": int(bad.sum()), "against": "reference",
                                      "fields": "CIGAR ops, mismatches"}
    return out


def klign_call_leg(calls=40, npairs=20000):
    """klign-sized calls: new batches of 20k pairs of 150 bp reads vs 350 bp windows per call, std::string in and
    short results out through the drop-in adept_sw::GPUDriver (mhm2_staging_amd/klign_call_bench, the call
    sequence of klign.cpp:356-388); and the same batches through the Python API (stage = pack + copy, the job,
    the result copy)."""
    import subprocess
    import mhm2_staging_amd as m
    prog = os.path.join(ROOT, "mhm2_staging_amd", "klign_call_bench")
    r = subprocess.run([prog, "--calls", str(calls), "--npairs", str(npairs)], capture_output=True, text=True,
                       timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"klign_call_bench failed ({r.returncode}): {r.stderr[-2000:]}")
    out = json.loads(r.stdout.strip().splitlines()[-1])
    out.pop("checksum", None)
    r = subprocess.run([prog, "--calls", str(calls), "--npairs", str(npairs), "--same-geometry", "1"],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        raise RuntimeError(f"klign_call_bench failed ({r.returncode}): {r.stderr[-2000:]}")
    same = json.loads(r.stdout.strip().splitlines()[-1])
    same.pop("checksum", None)
    out["equal_geometry"] = same
    batches = [m.generate(1, npairs=npairs, seed=101 + k) for k in range(4)]
    al = m.Aligner(device=0, scoring=m.MHM2_SCORING, read_is_ssw_ref=True)
    for b in batches:  # every buffer at its size, the worker pool started
        al.align(b)
    stage, job, take, wall = [], [], [], []
    for i in range(calls):
        b = batches[i % len(batches)]
        t0 = time.perf_counter()
        al.stage(b)
        t1 = time.perf_counter()
        al.align_staged()
        al.wait()
        t2 = time.perf_counter()
        al.results()
        t3 = time.perf_counter()
        stage.append(t1 - t0)
        job.append(t2 - t1)
        take.append(t3 - t2)
        wall.append(t3 - t0)
    al.close()
    cells = float(np.mean([b.cells() for b in batches]))
    w = float(np.median(wall))
    out["python_api"] = {"workload": "the same batches as flat arrays through the Python API (Aligner.stage, "
                                     "align_staged, wait, results)",
                         "ms_per_call": round(w * 1e3, 3), "value": round(cells / w / 1e9, 1), "unit": "GCUPS",
                         "stage_ms": round(float(np.median(stage)) * 1e3, 3),
                         "job_ms": round(float(np.median(job)) * 1e3, 3),
                         "results_ms": round(float(np.median(take)) * 1e3, 3)}
    return out


def locassm_leg(nctgs, steps, warmup, check, cores):
    """Local assembly (contig end extension) on a generated batch resident in
    HBM: one step = both sides of every contig walked at every mer length
    the schedule visits (what local_assem_driver does for a slice)."""
    from mhm2_staging_amd import locassm
    from mhm2_staging_amd import locassm_data as ld

    t = time.perf_counter()
    b = ld.generate(nctgs, seed=7)
    log(f"[bench] local assembly batch: {b.nctgs} contigs, {b.nreads} reads ({time.perf_counter() - t:.1f} s)")
    kw = dict(kmer_len=33, walk_len_limit=360)
    la = locassm.LocalAssembler(0)
    t = time.perf_counter()
    la.stage(b)
    stage_s = time.perf_counter() - t
    p = locassm.params(**kw)
    for _ in range(warmup):
        la.run(p)
        la.wait()
    ms = []
    t = time.perf_counter()
    for _ in range(steps):
        la.run(p)
        la.wait()
        ms.append(la.stats()["job_ms"])
    wall = (time.perf_counter() - t) / steps
    st = la.stats()
    got = la.results()
    la.close()
    dev_ms = float(np.mean(ms))
    out = {"workload": f"{b.nctgs} contigs U[200,2000] bp, both ends; per side 0 (10%), 1-9 (55%) or 10-80 (35%) "
                       "150 bp reads, 1% substitutions, 0.1% N; 10% of sides with a two-haplotype fork, 5% with a "
                       "repeat; kmer_len 33, walk_len_limit 360, depth U[2,50)",
           "contigs": b.nctgs, "reads": b.nreads, "sides_on_device": st["sides"], "launches": st["launches"],
           "mers_per_step": st["mers"], "walks_per_step": st["walks"], "walk_steps_per_step": st["walk_steps"],
           "device_ms": round(dev_ms, 3), "ms_per_step": round(wall * 1e3, 3),
           "stage_ms": round(stage_s * 1e3, 1),
           "value": round(b.nctgs / (dev_ms / 1e3), 1), "unit": "contigs/s (both ends)",
           "mers_per_s": round(st["mers"] / (dev_ms / 1e3), 1),
           "scratch_bytes": st["scratch_bytes"], "table_resizes": st["table_resizes"],
           "phase_share": dict(zip(("clear", "count", "decide", "walk"),
                                   (round(x / max(1, sum(st["phase_cycles"])), 3) for x in st["phase_cycles"]))),
           "extended_sides": int((got.left_len > 0).sum() + (got.right_len > 0).sum()),
           "bases_added": int(got.left_len.sum() + got.right_len.sum())}
    # roofline: random transactions into per-side tables far larger than the
    # caches.  Algorithmic bytes = two 64-byte sectors per mer occurrence (its
    # slot probe and its vote atomic, the count phase's two accesses into the
    # table); traffic = the whole launch's HBM bytes from the committed PMC
    # profile of this workload
    alg = 2 * 64 * st["mers"]
    rl = {"bound": "hbm", "kernel": "walk_kernel", "unit": "GB/s", "peak": 8000.0,
          "algorithmic_bytes": int(alg), "bytes_per_mer_occurrence": 128,
          "achieved": round(alg / (dev_ms * 1e-3) / 1e9, 1)}
    rl["frac"] = round(rl["achieved"] / rl["peak"], 4)
    prof = load_traffic("locassm_50k")
    if prof and nctgs == 50000:
        rl["traffic"] = prof["per_launch_bytes"]["walk_kernel"]
        rl["traffic_source"] = prof["source"]
        rl["traffic_gbs"] = round(rl["traffic"] / (dev_ms * 1e-3) / 1e9, 1)
    out["roofline"] = rl
    if check:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_lib
        t = time.perf_counter()
        ll, le, rl, re, tm = oracle_lib.la_extend_batch(oracle_lib.la_params(**kw), b, nthreads=cores, with_terms=True)
        dt = time.perf_counter() - t
        bad = int(((ll != got.left_len) | (rl != got.right_len) | (tm != got.terms).any(axis=(1, 2))).sum())
        for c in np.nonzero((ll == got.left_len) & (rl == got.right_len))[0]:
            if (not np.array_equal(le[c, :ll[c]], got.left[c, :ll[c]])
                    or not np.array_equal(re[c, :rl[c]], got.right[c, :rl[c]])):
                bad += 1
        out["cpu_baseline"] = {"value": round(b.nctgs / dt, 1), "unit": "contigs/s (both ends)", "cores": cores,
                               "kind": "port", "sample": f"all {b.nctgs} contigs of the same batch, {dt:.1f} s "
                               "(oracle/locassm_oracle.c, the reference kernel's semantics, one contig per task)"}
        out["bit_exact_check"] = {"contigs": b.nctgs, "mismatches": bad, "against": "oracle",
                                  "fields": "both extensions (bytes and lengths), walk terminations"}
    return out


def serving_bench(qps_list, duration, slots=4, seed=5):
    """Serving loop: 64k resident contigs, Poisson requests of log-uniform 1e3..1e5 pairs, at each request rate
    of qps_list; plus the HBM rate of the window extraction from the resident contigs."""
    import mhm2_staging_amd as m
    rng = np.random.default_rng(seed)
    cs = m.contigs(nctgs=64000, seed=seed)
    res = m.ResidentContigs(cs, device=0)
    # a pool of distinct requests, resubmitted cyclically (generation is not timed)
    pool = []
    for k in range(48):
        n = int(round(np.exp(rng.uniform(np.log(1e3), np.log(1e5)))))
        pool.append(m.window_pairs(cs, n, seed=1000 + k))
    # window extraction (resident 2-bit codes + N bitmap -> nibble codes of every window) on one large
    # windowed batch, eager launches so its HIP events are live
    big = m.window_pairs(cs, 1 << 20, seed=999)
    al = m.Aligner(device=0, scoring=m.MHM2_SCORING)
    al.set_graphs(False)
    al.stage_windows(res, big)
    wms = []
    for _ in range(4):
        al.align_staged()
        al.wait()
        wms.append(al.stats()["window_ms"])
    st = al.stats()
    al.close()
    wbases = st["window_bases"]
    wbytes = wbases * (0.25 + 0.125 + 0.5)  # read 2-bit codes and N bits, write 4-bit codes
    wmed = float(np.median(wms[1:]))
    extraction = {"pairs": big.npairs, "window_bases": int(wbases), "ms": round(wmed, 4),
                  "algorithmic_bytes": int(wbytes), "achieved_gbs": round(wbytes / (wmed * 1e-3) / 1e9, 1),
                  "peak_gbs": 8000.0, "frac": round(wbytes / (wmed * 1e-3) / 8e12, 4),
                  "bytes_per_base": 0.875}
    log(f"[bench] serving: window extraction {extraction}")
    srv = m.Server(res, scoring=m.MHM2_SCORING, slots=slots)
    rid = srv.submit(pool[0])  # warm up every code path once
    srv.results(rid)
    points = []
    for qps in qps_list:
        gaps = rng.exponential(1.0 / qps, size=int(qps * duration) + 1)
        arrivals = np.cumsum(gaps)
        arrivals = arrivals[arrivals < duration]
        t0 = time.perf_counter()
        c0 = srv.clock()  # server clock at t0: latencies run from the scheduled arrival
        ids = []
        for i, a in enumerate(arrivals):
            d = t0 + a - time.perf_counter()
            if d > 0:
                time.sleep(d)
            b = pool[i % len(pool)]
            ids.append((srv.submit(b), b.cells(), b.npairs, c0 + a))
        lat, cells, first, last, pairs, dev = [], 0, None, None, 0, 0.0
        for rid, c, n, arrival in ids:
            _, t = srv.results(rid)
            lat.append(t["done_s"] - arrival)
            cells += c
            pairs += n
            first = t["submit_s"] if first is None else min(first, t["submit_s"])
            last = t["done_s"] if last is None else max(last, t["done_s"])
        lat = np.array(lat) * 1e3
        span = last - first
        offered = sum(c for _, c, _, _ in ids) / duration
        points.append({"qps": qps, "requests": len(ids), "p50_ms": round(float(np.percentile(lat, 50)), 2),
                       "p99_ms": round(float(np.percentile(lat, 99)), 2),
                       "offered_gcups": round(offered / 1e9, 1),
                       "sustained_gcups": round(cells / span / 1e9, 1), "pairs_per_s": round(pairs / span, 1)})
        log(f"[bench] serving {points[-1]}")
    srv.close()
    res.close()
    return {"workload": "64k resident contigs (log-uniform 500..20000 bp), Poisson requests of log-uniform "
                        "1e3..1e5 pairs (150 bp reads, 1% error) vs U[300,1024] windows",
            "slots": slots, "duration_s_per_point": duration, "points": points, "window_extraction": extraction}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=HEADLINE, choices=sorted(WORKLOADS),
                    help="headline batch workload (generator number: 1 plain 10k, 2 uneven 1M, 3 skewed 4M, "
                         "4 overflow 500k)")
    ap.add_argument("--npairs", type=int, default=0, help="headline batch size (0: the workload's own)")
    ap.add_argument("--secondary", type=str, default="", metavar="N,N",
                    help="other batch workloads to run after the headline (default: every other one)")
    ap.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison runs")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines and bit-exact checks")
    ap.add_argument("--no-serving", action="store_true", help="skip the serving-loop measurement")
    ap.add_argument("--no-cigar", action="store_true", help="skip the CIGAR-pass measurement")
    ap.add_argument("--no-locassm", action="store_true", help="skip the local assembly measurement")
    ap.add_argument("--no-klign", action="store_true", help="skip the klign-sized call measurement")
    ap.add_argument("--only-headline", action="store_true", help="the headline batch workload only")
    ap.add_argument("--locassm-only", action="store_true", help="only the local assembly measurement")
    ap.add_argument("--klign-only", action="store_true", help="only the klign-sized call measurement")
    ap.add_argument("--locassm-contigs", type=int, default=50000)
    ap.add_argument("--no-autotune", action="store_true", help="time the default scheduling parameters")
    ap.add_argument("--tuning", type=str, default="", metavar="B,G,S",
                    help="fixed scheduling parameters (blocks_per_cu, g64_min_rows, bin_shift) instead of autotuning")
    ap.add_argument("--serving-qps", type=str, default="50,400,1000,1600,2000,2400", help="comma-separated request rates")
    ap.add_argument("--serving-seconds", type=float, default=4.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB (3 times) to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 3}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (asked for {args.gpus}); the path shards by pairs "
              "with no exchange, so multi-GPU throughput is N times the one-GPU value by construction.",
              file=sys.stderr)
        sys.exit(2)
    cores, core_info = host_cores()
    log(f"[bench] host cores used: {cores} ({core_info})")

    if args.klign_only:
        kc = klign_call_leg()
        log(f"[bench] klign-sized calls: {kc}")
        print(json.dumps({"klign_call": kc}), flush=True)
        return

    if args.locassm_only:
        la = locassm_leg(args.locassm_contigs, args.steps, args.warmup, not args.no_cpu, cores)
        log(f"[bench] local assembly: {la}")
        print(json.dumps({"local_assembly": la}), flush=True)
        return

    from mhm2_staging_amd import _native
    peak = _native.measure_valu_peak(0)
    log(f"[bench] packed int16 VALU peak measured: {peak['lane_ops_per_s'] / 1e12:.2f} T lane-ops/s at "
        f"{peak['clock_mhz']:.0f} MHz (nominal {NOMINAL_PEAK} T at 2400 MHz)")

    h = batch_leg(args.config, args, cores, peak, headline=True)
    key, text = WORKLOADS[args.config]
    out = {"metric": METRIC, "value": h["value"], "unit": "GCUPS", "n_gpus": 1, "steps": h["steps"],
           "warmup": args.warmup, "ms_per_step": h["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "i16 (packed, exact i32 rerun)",
           "data": "synthetic (seeded generator)",
           "config": {"workload": text, "name": key, "npairs": h["npairs"], "cells": h["cells"],
                      "scoring": SCORING_TEXT, "form": "fast", "tuning": h["tuning"]},
           "pairs_per_s": h["pairs_per_s"], "device_job_ms": h["device_job_ms"], "stats": h["stats"],
           "classes": h["classes"], "roofline": h["roofline"],
           "valu_peak_measured": {"lane_ops_per_s": peak["lane_ops_per_s"], "clock_mhz": round(peak["clock_mhz"], 1),
                                  "max_clock_mhz": peak["max_clock_mhz"], "cus": peak["cus"]},
           "host_cores": dict(core_info, used=cores)}
    for k in ("default_tuning", "graph_replay", "end_to_end", "plain_form", "cigar_pass", "cpu_baseline",
              "cpu_one_core", "bit_exact_check"):
        if k in h:
            out[k] = h[k]
    if not args.only_headline:
        sec = [int(x) for x in args.secondary.split(",") if x] if args.secondary else \
            [c for c in sorted(WORKLOADS) if c != args.config]
        out["workloads"] = {}
        for c in sec:
            leg = batch_leg(c, args, cores, peak)
            out["workloads"][WORKLOADS[c][0]] = leg
        if not args.no_klign:
            out["klign_call"] = klign_call_leg()
            log(f"[bench] klign-sized calls: {out['klign_call']}")
        if not args.no_locassm:
            out["local_assembly"] = locassm_leg(args.locassm_contigs, min(args.steps, 10), min(args.warmup, 2),
                                                not args.no_cpu, cores)
            log(f"[bench] local assembly: {out['local_assembly']}")
        if not args.no_serving:
            out["serving"] = serving_bench([float(x) for x in args.serving_qps.split(",")], args.serving_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
