"""SSW's second-best score on the GPU (mhmsw_set_score2), against the reference
SSW's own values (golden_score2.npz) and the oracle, in every form.

score2 is the largest per-SSW-ref-position maximum outside the window
[end - maskLen, end + maskLen] (16-bit pass: one position narrower at the
top), maxima taken over SSW's zero-scoring query padding too
(ssw_core.cpp:512-536, :822-842; klign.cpp:257)."""
import os

import numpy as np
import pytest

import mhm2_staging_amd as m
from golden.kat_cases import MHM2_DEFAULT_SCORING
from oracle_lib import oracle_batch2
from test_gpu_parity import assert_same, gpu_table

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
SCORINGS = [MHM2_DEFAULT_SCORING, (1, 1, 1, 1, 1), (2, 2, 3, 1, 2), (2, 4, 4, 2, 1)]


def strings_batch(refs_buf, refs_off, refs_len, qry_buf, qry_off, qry_len):
    """A batch whose reads are the SSW-refs and whose windows are the SSW-queries."""
    reads = [bytes(refs_buf[o:o + n]) for o, n in zip(refs_off, refs_len)]
    ctgs = [bytes(qry_buf[o:o + n]) for o, n in zip(qry_off, qry_len)]
    return m.from_strings(reads, ctgs)


def run_score2(al, batch, scoring, mode=m.MODE_FAST, read_is_ref=True):
    print(f"[gpu] score2 npairs={batch.npairs} scoring={scoring} mode={mode} read_is_ref={read_is_ref}", flush=True)
    al.set_scoring(scoring)
    al.set_mode(mode)
    al.set_orientation(read_is_ref)
    al.set_score2(True)
    al.stage(batch)
    al.align_staged()
    al.wait()
    res, s2 = al.results(), al.score2()
    al.set_score2(False)
    return res, s2


def oracle_s2(batch, scoring, read_is_ref=True):
    if read_is_ref:
        o, s2 = oracle_batch2(batch.reads, batch.read_off, batch.read_len, batch.ctgs, batch.ctg_off, batch.ctg_len,
                              scoring)
    else:
        o, s2 = oracle_batch2(batch.ctgs, batch.ctg_off, batch.ctg_len, batch.reads, batch.read_off, batch.read_len,
                              scoring)
    return np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin", "query_end")], 1), s2


def check(al, batch, scoring, mode=m.MODE_FAST, read_is_ref=True):
    res, s2 = run_score2(al, batch, scoring, mode, read_is_ref)
    want, want_s2 = oracle_s2(batch, scoring, read_is_ref)
    assert_same(gpu_table(res, read_is_ref), want.astype(np.int32), batch)
    bad = np.nonzero(s2 != want_s2)[0]
    assert len(bad) == 0, f"{len(bad)} score2 mismatches, first: pair {bad[0]} got {s2[bad[0]]} want {want_s2[bad[0]]}"


@pytest.mark.parametrize("mode", [m.MODE_FAST, m.MODE_PLAIN])
def test_score2_reference_fixtures(gpu_aligner, mode):
    z = np.load(os.path.join(GOLD, "golden_score2.npz"))
    small = np.load(os.path.join(GOLD, "golden_small.npz"))
    b = strings_batch(small["ref_buf"], small["ref_off"], small["ref_len"], small["qry_buf"], small["qry_off"],
                      small["qry_len"])
    for name in ("ssw_default", "mhm2_tree", "cigar", "baseline"):
        sc = tuple(int(v) for v in small["scoring_" + name])
        res, s2 = run_score2(gpu_aligner, b, sc, mode)
        np.testing.assert_array_equal(gpu_table(res), small["res_" + name])
        np.testing.assert_array_equal(s2, z["small_score2_" + name])
    for name in ("baseline", "ssw_default", "cigar"):
        p = f"edge_{name}_"
        b = strings_batch(z[p + "ref_buf"], z[p + "ref_off"], z[p + "ref_len"], z[p + "qry_buf"], z[p + "qry_off"],
                          z[p + "qry_len"])
        res, s2 = run_score2(gpu_aligner, b, tuple(int(v) for v in z[p + "scoring"]), mode)
        np.testing.assert_array_equal(gpu_table(res), z[p + "res"])
        np.testing.assert_array_equal(s2, z[p + "score2"])
    b = strings_batch(z["short_ref_buf"], z["short_ref_off"], z["short_ref_len"], z["short_qry_buf"],
                      z["short_qry_off"], z["short_qry_len"])
    for name in ("baseline", "mhm2_tree"):
        res, s2 = run_score2(gpu_aligner, b, tuple(int(v) for v in z[f"short_{name}_scoring"]), mode)
        np.testing.assert_array_equal(gpu_table(res), z[f"short_{name}_res"])
        np.testing.assert_array_equal(s2, z[f"short_{name}_score2"])


@pytest.mark.parametrize("cfg", [1, 2, 3, 4])
def test_score2_configs_against_oracle(gpu_aligner, cfg):
    batch = m.generate(cfg, npairs=4000, seed=90 + cfg)
    for sc in SCORINGS:
        check(gpu_aligner, batch, sc)


def test_score2_plain_form_and_other_orientation(gpu_aligner):
    batch = m.generate(2, npairs=3000, seed=97)
    check(gpu_aligner, batch, MHM2_DEFAULT_SCORING, mode=m.MODE_PLAIN)
    check(gpu_aligner, batch, (2, 2, 3, 1, 2), read_is_ref=False)


def test_score2_overflow_reruns(gpu_aligner):
    """Scores past the packed range rerun in 32-bit, score2 with them."""
    batch = m.generate(4, npairs=3000, seed=98)
    check(gpu_aligner, batch, (5, 3, 5, 2, 2))


def test_score2_leaves_other_results_unchanged(gpu_aligner):
    batch = m.generate(3, npairs=5000, seed=99)
    al = gpu_aligner
    al.set_scoring(MHM2_DEFAULT_SCORING)
    al.set_mode(m.MODE_FAST)
    al.set_orientation(True)
    al.set_score2(False)
    al.stage(batch)
    al.align_staged()
    al.wait()
    plain_res = al.results()
    res, _ = run_score2(al, batch, MHM2_DEFAULT_SCORING)
    assert (res == plain_res).all()


def test_score2_graph_replays(gpu_aligner):
    batch = m.generate(2, npairs=3000, seed=100)
    al = gpu_aligner
    al.set_graphs(True)
    try:
        res0, s20 = run_score2(al, batch, MHM2_DEFAULT_SCORING)
        al.set_score2(True)
        for _ in range(3):
            al.align_staged()
            al.wait()
            assert (al.results() == res0).all()
            assert (al.score2() == s20).all()
        al.set_score2(False)
    finally:
        al.set_graphs(False)
    _, want = oracle_s2(batch, MHM2_DEFAULT_SCORING)
    assert (s20 == want).all()
