"""SSW's CIGAR and mismatch count on the GPU (mhmsw_set_cigar), against the
reference SSW's own CIGARs (golden_cigar.npz) and the oracle, bit-exact in
every op and every count, in both forms and both orientations.

The CIGAR is what klign's call with report_cigar set returns
(ssw_core.cpp:1180-1192 banded_sw, ssw.cpp:57-93 and :129-254)."""
import os

import numpy as np
import pytest

import mhm2_staging_amd as m
from golden.kat_cases import MHM2_DEFAULT_SCORING
from oracle_lib import cigar_string, oracle_cigar_batch
from test_gpu_parity import assert_same, gpu_table

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ("score", "ref_begin", "ref_end", "query_begin", "query_end")


def strings_batch(refs_buf, refs_off, refs_len, qry_buf, qry_off, qry_len):
    """A batch whose reads are the SSW-refs and whose windows are the SSW-queries."""
    reads = [bytes(refs_buf[o:o + n]) for o, n in zip(refs_off, refs_len)]
    ctgs = [bytes(qry_buf[o:o + n]) for o, n in zip(qry_off, qry_len)]
    return m.from_strings(reads, ctgs)


def run_cigar(al, batch, scoring, mode=m.MODE_FAST, read_is_ref=True):
    print(f"[gpu] cigar npairs={batch.npairs} scoring={scoring} mode={mode} read_is_ref={read_is_ref}", flush=True)
    al.set_scoring(scoring)
    al.set_mode(mode)
    al.set_orientation(read_is_ref)
    al.set_cigar(True)
    try:
        al.stage(batch)
        al.align_staged()
        al.wait()
        res = al.results()
        mism, off, ops = al.cigars()
        stats = al.stats()
    finally:
        al.set_cigar(False)
    return res, mism, off, ops, stats


def assert_cigars(batch, mism, off, ops, want_mism, want_nops, want_ops_of):
    """want_ops_of(p) -> the expected ops of pair p."""
    nops = np.diff(off)
    bad = np.nonzero((nops != want_nops) | (mism != want_mism))[0]
    if len(bad) == 0:
        for p in range(len(nops)):
            if not np.array_equal(ops[off[p]:off[p + 1]], want_ops_of(p)):
                bad = np.array([p])
                break
    if len(bad):
        p = int(bad[0])
        raise AssertionError(
            f"pair {p}: got {cigar_string(ops[off[p]:off[p + 1]])} mism {mism[p]}, want "
            f"{cigar_string(want_ops_of(p))} mism {want_mism[p]} (read={batch.read(p)[:60]} ctg={batch.ctg(p)[:60]})")


def oracle_cigars(batch, scoring, read_is_ref=True):
    if read_is_ref:
        args = (batch.reads, batch.read_off, batch.read_len, batch.ctgs, batch.ctg_off, batch.ctg_len)
    else:
        args = (batch.ctgs, batch.ctg_off, batch.ctg_len, batch.reads, batch.read_off, batch.read_len)
    res, mism, nops, ops = oracle_cigar_batch(*args, scoring)
    assert (nops >= 0).all()
    return np.stack([res[k] for k in FIELDS], 1).astype(np.int32), mism, nops, ops


def check_vs_oracle(al, batch, scoring, mode=m.MODE_FAST, read_is_ref=True):
    res, mism, off, ops, stats = run_cigar(al, batch, scoring, mode, read_is_ref)
    want, wmism, wnops, wops = oracle_cigars(batch, scoring, read_is_ref)
    assert_same(gpu_table(res, read_is_ref), want, batch)
    assert_cigars(batch, mism, off, ops, wmism, wnops, lambda p: wops[p, :wnops[p]])
    return stats


@pytest.mark.parametrize("mode", [m.MODE_FAST, m.MODE_PLAIN])
def test_cigar_reference_fixtures(gpu_aligner, mode):
    g = np.load(os.path.join(GOLD, "golden_cigar.npz"))
    s = np.load(os.path.join(GOLD, "golden_small.npz"))
    sets = [("small_" + n, strings_batch(s["ref_buf"], s["ref_off"], s["ref_len"], s["qry_buf"], s["qry_off"],
                                         s["qry_len"])) for n in ("ssw_default", "mhm2_tree", "cigar", "baseline")]
    tiny = strings_batch(g["tiny_ref_buf"], g["tiny_ref_off"], g["tiny_ref_len"], g["tiny_qry_buf"],
                         g["tiny_qry_off"], g["tiny_qry_len"])
    sets += [("tiny_" + n, tiny) for n in ("baseline", "mhm2_tree", "ssw_default", "cigar", "bigmatch", "match5")]
    for name, b in sets:
        sc = tuple(int(v) for v in g[name + "_scoring"])
        res, mism, off, ops, _ = run_cigar(gpu_aligner, b, sc, mode)
        np.testing.assert_array_equal(gpu_table(res), g[name + "_res"])
        wn = g[name + "_nops"]
        woff = np.concatenate([[0], np.cumsum(wn)])
        wops = g[name + "_ops"]
        assert_cigars(b, mism, off, ops, g[name + "_mism"], wn, lambda p: wops[woff[p]:woff[p + 1]])


@pytest.mark.parametrize("config,npairs", [(1, 6000), (2, 6000), (3, 6000), (4, 1500), (5, 4000)])
def test_cigar_configs_vs_oracle(gpu_aligner, config, npairs):
    b = m.generate(config, npairs=npairs, seed=21)
    for sc in (MHM2_DEFAULT_SCORING, (1, 1, 1, 1, 1), (10, 2, 3, 1, 2)):
        check_vs_oracle(gpu_aligner, b, sc)


def test_cigar_plain_form_and_other_orientation(gpu_aligner):
    b = m.generate(2, npairs=3000, seed=22)
    check_vs_oracle(gpu_aligner, b, MHM2_DEFAULT_SCORING, mode=m.MODE_PLAIN)
    check_vs_oracle(gpu_aligner, b, MHM2_DEFAULT_SCORING, read_is_ref=False)
    check_vs_oracle(gpu_aligner, b, (2, 2, 3, 1, 2), read_is_ref=False)


def test_cigar_edge_cases(gpu_aligner):
    rng = np.random.default_rng(5)

    def rs(n, alph="ACGT"):
        return "".join(rng.choice(list(alph), n)) if n else ""

    reads, ctgs = [], []
    for r, c in [("", "ACGT"), ("", "CAGT"), ("A", "A"), ("A", "C"), ("N", "N"), ("NNNN", "NNNN"), ("acgu", "ACGT"),
                 ("ACGTX", "ACGT-"), ("ACGT", "TTTT")]:
        reads.append(r)
        ctgs.append(c)
    w = rs(5000)
    reads.append(w[:4000])          # 4000-base read against a 1024-base window
    ctgs.append(w[1000:2024])
    reads.append(w[:1500])          # both sides above 1024: 32-bit form
    ctgs.append(w[200:2200])
    reads.append(rs(300, "AC"))     # many equal-score paths
    ctgs.append(rs(900, "AC"))
    # deletions and insertions of 40..600 bases: wide bands of 1, 2, 16 chunks of
    # 64 diagonals and, for 600, wider than the lanes (serial path)
    for gap in (40, 100, 300):
        reads.append(w[:300] + w[300 + gap:600 + gap])
        ctgs.append(w[:600 + gap])
        reads.append(w[:300] + rs(gap) + w[300:600])
        ctgs.append(w[:600])
    reads.append(w[:1500] + w[2100:3600])
    ctgs.append(w[:3600])
    b = m.from_strings(reads, ctgs)
    for sc in (MHM2_DEFAULT_SCORING, (1, 1, 1, 1, 1), (10, 2, 3, 1, 2)):
        for mode in (m.MODE_FAST, m.MODE_PLAIN):
            check_vs_oracle(gpu_aligner, b, sc, mode=mode)


def test_cigar_pools_grow_and_rerun(gpu_aligner):
    """Mismatch cost 0 makes CIGARs of alternating = and X runs (far more ops
    than the pool's first size of 12 per pair); the pass reruns with a pool
    of the size it counted, and every CIGAR still matches the oracle."""
    rng = np.random.default_rng(6)
    reads, ctgs = [], []
    for _ in range(2000):
        r = "".join(rng.choice(list("ACGT"), 200))
        c = "".join(ch if k % 2 == 0 else "ACGT"[("ACGT".index(ch) + 1) % 4] for k, ch in enumerate(r))
        reads.append(r)
        ctgs.append("GG" + c + "TT")
    b = m.from_strings(reads, ctgs)
    stats = check_vs_oracle(gpu_aligner, b, (1, 0, 5, 2, 0))
    assert stats["cigar_reruns"] >= 1


def test_cigar_results_unchanged_and_graph_replays(gpu_aligner):
    """The CIGAR pass leaves the alignment results as they are, and replays of
    the captured job (CIGAR pass inside) match an eager run every time."""
    b = m.generate(3, npairs=5000, seed=23)
    al = gpu_aligner
    al.set_scoring(MHM2_DEFAULT_SCORING)
    al.set_mode(m.MODE_FAST)
    al.set_orientation(True)
    base = al.align(b)
    al.set_graphs(False)
    eager = run_cigar(al, b, MHM2_DEFAULT_SCORING)
    al.set_graphs(True)
    assert np.array_equal(gpu_table(eager[0]), gpu_table(base))
    al.set_cigar(True)
    try:
        al.stage(b)
        for k in range(3):
            al.align_staged()
            al.wait()
            got = al.cigars()
            for a, e in zip(got, eager[1:4]):
                assert np.array_equal(a, e), f"replay {k}"
            np.testing.assert_array_equal(gpu_table(al.results()), gpu_table(base))
            _ = m.generate(1, npairs=64, seed=k)  # host work between replays
    finally:
        al.set_cigar(False)
