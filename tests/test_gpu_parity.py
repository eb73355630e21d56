"""HIP path (libmhmsw.so) against the oracle and the reference's golden vectors.

Bit-exact on (score, read begin/end, contig begin/end) for every pair, in both
execution forms (packed 16-bit fast form, plain 32-bit form).
"""
import json
import os

import numpy as np
import pytest

import mhm2_staging_amd as m
from golden.kat_cases import MHM2_DEFAULT_SCORING, KAT_CASES
from oracle_lib import oracle_batch

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
MODES = [m.MODE_FAST, m.MODE_PLAIN]


def gpu_table(res, read_is_ref=True):
    """GPU results in SSW order (score, ref_begin, ref_end, query_begin, query_end)."""
    if read_is_ref:
        cols = ("score", "read_begin", "read_end", "ctg_begin", "ctg_end")
    else:
        cols = ("score", "ctg_begin", "ctg_end", "read_begin", "read_end")
    return np.stack([res[c] for c in cols], 1).astype(np.int32)


def oracle_table(batch, scoring, read_is_ref=True):
    if read_is_ref:
        o = oracle_batch(batch.reads, batch.read_off, batch.read_len, batch.ctgs, batch.ctg_off, batch.ctg_len, scoring)
    else:
        o = oracle_batch(batch.ctgs, batch.ctg_off, batch.ctg_len, batch.reads, batch.read_off, batch.read_len, scoring)
    return np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin", "query_end")], 1).astype(np.int32)


def run(al, batch, scoring, mode, read_is_ref=True):
    print(f"[gpu] npairs={batch.npairs} scoring={scoring} mode={mode} read_is_ref={read_is_ref}", flush=True)
    al.set_scoring(scoring)
    al.set_orientation(read_is_ref)
    al.set_mode(mode)
    res = al.align(batch)
    return gpu_table(res, read_is_ref)


def assert_same(got, want, batch=None):
    bad = np.nonzero((got != want).any(1))[0]
    if len(bad):
        p = int(bad[0])
        detail = ""
        if batch is not None:
            detail = f" read={batch.read(p)[:80]} ctg={batch.ctg(p)[:80]}"
        raise AssertionError(f"{len(bad)} of {len(got)} pairs differ; first {p}: got {got[p].tolist()} "
                             f"want {want[p].tolist()}{detail}")


@pytest.mark.parametrize("mode", MODES)
def test_kats_vs_reference_golden(gpu_aligner, mode):
    with open(os.path.join(GOLD, "golden_kat.json")) as f:
        kat = json.load(f)
    refs = [c[1] for c in KAT_CASES]
    qrys = [c[0] for c in KAT_CASES]
    for name, sc in kat["scorings"].items():
        # reads play SSW-ref: reads = KAT refs, contigs = KAT queries
        b = m.from_strings(refs, qrys)
        assert_same(run(gpu_aligner, b, tuple(sc), mode), np.array(kat["results"][name]))
        # same inputs, opposite roles through the orientation switch
        b2 = m.from_strings(qrys, refs)
        assert_same(run(gpu_aligner, b2, tuple(sc), mode, read_is_ref=False), np.array(kat["results"][name]))
        assert_same(run(gpu_aligner, b, tuple(sc), mode, read_is_ref=False),
                    np.array(kat["results"][name + "_swapped"]))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", ["ssw_default", "mhm2_tree", "cigar", "baseline"])
def test_small_corpus_vs_reference_golden(gpu_aligner, mode, name):
    z = np.load(os.path.join(GOLD, "golden_small.npz"))
    b = m.PairBatch(z["ref_buf"], z["ref_off"], z["ref_len"], z["qry_buf"], z["qry_off"], z["qry_len"])
    sc = tuple(int(x) for x in z[f"scoring_{name}"])
    assert_same(run(gpu_aligner, b, sc, mode), z[f"res_{name}"], b)


@pytest.mark.parametrize("mode", MODES)
def test_config1_vs_reference_golden(gpu_aligner, mode):
    z = np.load(os.path.join(GOLD, "golden_config1.npz"))
    b = m.generate(1)
    assert_same(run(gpu_aligner, b, MHM2_DEFAULT_SCORING, mode), z["results"], b)


@pytest.mark.parametrize("config,npairs", [(2, 6000), (3, 6000), (4, 1500), (5, 4000)])
@pytest.mark.parametrize("mode", MODES)
def test_configs_vs_oracle(gpu_aligner, config, npairs, mode):
    b = m.generate(config, npairs=npairs)
    assert_same(run(gpu_aligner, b, MHM2_DEFAULT_SCORING, mode), oracle_table(b, MHM2_DEFAULT_SCORING), b)


@pytest.mark.parametrize("scoring", [(1, 1, 1, 1, 1), (2, 2, 3, 1, 2), (2, 4, 4, 2, 1)])
def test_other_scorings_vs_oracle(gpu_aligner, scoring):
    b = m.generate(2, npairs=3000, seed=11)
    want = oracle_table(b, scoring)
    for mode in MODES:
        assert_same(run(gpu_aligner, b, scoring, mode), want, b)


def test_orientation_ctg_as_ssw_ref_vs_oracle(gpu_aligner):
    b = m.generate(3, npairs=3000, seed=5)
    want = oracle_table(b, MHM2_DEFAULT_SCORING, read_is_ref=False)
    for mode in MODES:
        assert_same(run(gpu_aligner, b, MHM2_DEFAULT_SCORING, mode, read_is_ref=False), want, b)


def test_overflow_rerun_exact(gpu_aligner):
    """match=5 on near-identical 1024 bp repeats: 16-bit lanes saturate and those pairs rerun in 32-bit."""
    sc = (5, 3, 5, 2, 2)
    b = m.generate(4, npairs=600)
    want = oracle_table(b, sc)
    got = run(gpu_aligner, b, sc, m.MODE_FAST)
    st = gpu_aligner.stats()
    assert st["rerun_pairs"] > 0
    assert want[:, 0].max() > 2047
    assert_same(got, want, b)


def test_edge_cases_vs_oracle(gpu_aligner):
    rng = np.random.default_rng(3)

    def rs(n, alph="ACGT"):
        return "".join(rng.choice(list(alph), n)) if n else ""

    reads, ctgs = [], []
    # empty, single bases, all-N, lowercase/U, length boundaries of the packed classes
    for r, c in [("", "ACGT"), ("ACGT", ""), ("", ""), ("A", "A"), ("A", "C"), ("N", "N"), ("NNNN", "NNNN"),
                 ("acgu", "ACGT"), ("ACGTX", "ACGT-")]:
        reads.append(r)
        ctgs.append(c)
    for L in [31, 32, 33, 63, 64, 65, 255, 256, 257, 383, 384, 385, 1023, 1024, 1025]:
        w = rs(L + 200)
        reads.append(w[100:100 + L])
        ctgs.append(w)
        reads.append(w)          # read longer than the window
        ctgs.append(w[50:50 + L])
    w = rs(5000)
    reads.append(w[:4000])
    ctgs.append(w[1000:2024])
    reads.append(w[:1500])        # both sides above 1024: 32-bit form
    ctgs.append(w[200:2200])
    reads.append(rs(300, "AC"))   # many equal-score ends
    ctgs.append(rs(900, "AC"))
    b = m.from_strings(reads, ctgs)
    want = oracle_table(b, MHM2_DEFAULT_SCORING)
    for mode in MODES:
        assert_same(run(gpu_aligner, b, MHM2_DEFAULT_SCORING, mode), want, b)


def test_repeat_staged_runs_identical(gpu_aligner):
    b = m.generate(3, npairs=4000, seed=9)
    gpu_aligner.set_scoring(MHM2_DEFAULT_SCORING)
    gpu_aligner.set_orientation(True)
    gpu_aligner.set_mode(m.MODE_FAST)
    gpu_aligner.stage(b)
    outs = []
    for _ in range(3):
        gpu_aligner.align_staged()
        gpu_aligner.wait()
        outs.append(gpu_aligner.results())
    assert (outs[0] == outs[1]).all() and (outs[1] == outs[2]).all()


def test_full_size_config2_properties(gpu_aligner):
    """1M-pair uneven batch: fast == plain on every pair, sample == oracle, coordinate invariants."""
    b = m.generate(2)
    fast = run(gpu_aligner, b, MHM2_DEFAULT_SCORING, m.MODE_FAST)
    st = gpu_aligner.stats()
    assert st["npairs"] == b.npairs and st["cells"] == b.cells()
    plain = run(gpu_aligner, b, MHM2_DEFAULT_SCORING, m.MODE_PLAIN)
    assert_same(fast, plain, b)
    pos = fast[:, 0] > 0
    assert (fast[pos, 1] <= fast[pos, 2]).all() and (fast[pos, 3] <= fast[pos, 4]).all()
    assert (fast[pos, 2] < b.read_len[pos]).all() and (fast[pos, 4] < b.ctg_len[pos]).all()
    assert (fast[:, 0] <= np.minimum(b.read_len, b.ctg_len)).all()
    idx = np.random.default_rng(0).choice(b.npairs, 3000, replace=False)
    sub = b.subset(idx)
    assert_same(fast[idx], oracle_table(sub, MHM2_DEFAULT_SCORING), sub)


def test_graph_replays_match_eager_launches(gpu_aligner):
    """The captured job graph, replayed with eager work in between, matches eager launches and the oracle."""
    a = m.generate(3, npairs=3000, seed=31)
    bb = m.generate(2, npairs=2000, seed=32)
    al = gpu_aligner
    al.set_scoring(MHM2_DEFAULT_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    try:
        al.set_graphs(True)
        outs = []
        al.stage(a)
        for _ in range(3):  # capture, then replays
            al.align_staged()
            al.wait()
            outs.append(gpu_table(al.results()))
        al.set_graphs(False)
        al.align_staged()
        al.wait()
        outs.append(gpu_table(al.results()))
        al.stage(bb)  # eager work on another batch
        al.align_staged()
        al.wait()
        other = gpu_table(al.results())
        al.set_graphs(True)
        al.stage(a)
        for _ in range(2):
            al.align_staged()
            al.wait()
            outs.append(gpu_table(al.results()))
        want = oracle_table(a, MHM2_DEFAULT_SCORING)
        for o in outs:
            assert_same(o, want, a)
        assert_same(other, oracle_table(bb, MHM2_DEFAULT_SCORING), bb)
    finally:
        al.set_graphs(True)


def test_tunings_give_identical_results_and_autotune(gpu_aligner):
    """Scheduling parameters change speed only: every tuning matches the oracle."""
    b = m.generate(2, npairs=4000, seed=41)
    want = oracle_table(b, MHM2_DEFAULT_SCORING)
    al = gpu_aligner
    al.set_scoring(MHM2_DEFAULT_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    al.stage(b)
    try:
        for t in [(2, 257, 4), (8, 129, 6), (12, 193, 5), (1, 33, 3)]:
            print(f"[gpu] tuning {t}", flush=True)
            al.set_tuning(*t)
            al.align_staged()
            al.wait()
            assert_same(gpu_table(al.results()), want, b)
        best = al.autotune(repeats=1)
        assert 1 <= best["blocks_per_cu"] <= 16
        al.align_staged()
        al.wait()
        assert_same(gpu_table(al.results()), want, b)
    finally:
        al.set_tuning()


def test_phase_timing_off_same_results_job_time_only(gpu_aligner):
    """mhmsw_set_phase_timing(ctx, 0): an eager job records only its own two events (the drop-in
    driver's setting); results are unchanged and the stats report job_ms without per-phase times."""
    print("[gpu] phase timing off, 10k pairs", flush=True)
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    al.set_score2(False)
    b = m.generate(1)
    al.set_graphs(False)
    try:
        want = al.align(b)
        assert al.stats()["fwd_fast_ms"] > 0
        al._check(al._L.mhmsw_set_phase_timing(al._ctx, 0), "set_phase_timing")
        got = al.align(b)
        st = al.stats()
    finally:
        al._check(al._L.mhmsw_set_phase_timing(al._ctx, 1), "set_phase_timing")
        al.set_graphs(True)
    for k in ("score", "read_begin", "read_end", "ctg_begin", "ctg_end"):
        assert np.array_equal(got[k], want[k])
    assert st["job_ms"] > 0 and st["fwd_fast_ms"] == 0 and st["rev_ms"] == 0
