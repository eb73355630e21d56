"""The batch workloads at their full sizes, every pair against the reference SSW on the host cores.

The reference SSW is compiled from its own sources into oracle/_ref by oracle/Makefile and travels with the
tree as a built file; each test aligns the whole batch on the GPU (packed form, autotuned scheduling) and
compares every (score, begin, end) with the reference's.
"""
import os
import time

import numpy as np
import pytest

import mhm2_staging_amd as m
import oracle_lib

pytestmark = pytest.mark.gpu
HOST_THREADS = max(1, min(len(os.sched_getaffinity(0)), 16))


@pytest.mark.parametrize("config,npairs", [(3, 4_000_000), (4, 500_000), (2, 1_000_000)])
def test_full_batch_bit_exact_vs_reference(config, npairs):
    if not oracle_lib.have_ref():
        pytest.skip("oracle/_ref/libssw_ref.so is not built (make -C oracle where the reference sources are)")
    t = time.perf_counter()
    b = m.generate(config, npairs=npairs)
    assert b.npairs == npairs
    print(f"[gpu] workload {config}: {b.npairs} pairs generated in {time.perf_counter() - t:.1f}s", flush=True)
    with m.Aligner(device=0, scoring=m.MHM2_SCORING) as al:
        al.stage(b)
        al.autotune(repeats=1)
        al.align_staged()
        al.wait()
        got = al.results()
        st = al.stats()
    print(f"[gpu] job {st['job_ms']:.1f} ms; reference on {HOST_THREADS} threads", flush=True)
    t = time.perf_counter()
    ref = oracle_lib.ref_batch(b.reads, b.read_off, b.read_len, b.ctgs, b.ctg_off, b.ctg_len, m.MHM2_SCORING,
                               HOST_THREADS)
    print(f"[gpu] reference took {time.perf_counter() - t:.1f}s", flush=True)
    g = np.stack([got[c] for c in ("score", "read_begin", "read_end", "ctg_begin", "ctg_end")], 1)
    w = np.stack([ref[c] for c in ("score", "ref_begin", "ref_end", "query_begin", "query_end")], 1)
    bad = np.nonzero((g != w).any(1))[0]
    assert bad.size == 0, f"{bad.size} pairs differ, first {bad[:3]}: {g[bad[:3]]} vs {w[bad[:3]]}"
    if config == 4:
        assert st["fast_pairs"] > 0 and st["plain_pairs"] + st["fast_pairs"] == b.npairs
