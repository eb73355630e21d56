"""Every kernel shape of the packed forward pass against the 16-lane one and the oracle.

A class of many pairs runs narrow groups (2, 4 or 8 lanes per pair, 8..19
packed registers per lane), and the 97..128-row class a trimmed shape (4 x 13)
when its rows all fit 104; small classes run 16-lane groups.  The 129..152-row
class runs 4 x 19 and rows 153..160 belong to the 153..192-row class.  Each
batch below holds one read length, enough pairs for its class to take the
narrow (or trimmed) shape; the same pairs aligned in small slices take the
16-lane shape, and the results must agree on every pair, and with the oracle
on a sample.  Windows carry N bases in a tenth of the pairs, so both
substitution forms of the kernel run."""
import numpy as np
import pytest

import mhm2_staging_amd as m
from oracle_lib import oracle_batch, oracle_batch2

pytestmark = pytest.mark.gpu

FIELDS = ("score", "read_begin", "read_end", "ctg_begin", "ctg_end")
NPAIRS = 70_000    # above the narrow-shape bound of every class on 256 CUs (65,536 pairs for 2-lane groups)
SLICE = 4_000      # below the bound of every class: 16-lane groups
WINDOW = 300

# read length -> (lanes, packed registers) of the forward kernel its class runs
SHAPES = {30: (2, 8), 60: (2, 16), 90: (4, 12), 100: (4, 13), 104: (4, 13), 105: (4, 16), 120: (4, 16),
          129: (4, 19), 150: (4, 19), 152: (4, 19), 153: (8, 12), 160: (8, 12), 192: (8, 12), 193: (8, 14),
          250: (8, 16), 256: (8, 16)}


def class_of(rows):
    """The packed class of a pair with `rows` rows (fast16_class_of in sw_common.h; the 64-lane
    classes, 8..15, above 256 rows)."""
    if rows > 256:
        return 8 + (rows + 127) // 128 - 1
    if rows <= 128:
        return (rows + 31) // 32 - 1
    return 4 if rows <= 152 else 5 if rows <= 192 else (rows + 31) // 32 - 1


def batch_of(read_len, seed, n=NPAIRS, W=WINDOW):
    rng = np.random.default_rng(seed)
    L = read_len
    win = rng.integers(0, 4, (n, W), dtype=np.uint8)
    start = rng.integers(0, W - L + 1, n)
    rd = np.take_along_axis(win, start[:, None] + np.arange(L)[None, :], 1)
    sub = rng.random((n, L)) < 0.02
    rd[sub] = (rd[sub] + rng.integers(1, 4, int(sub.sum()), dtype=np.uint8)) % 4
    unrelated = rng.random(n) < 0.1
    rd[unrelated] = rng.integers(0, 4, (int(unrelated.sum()), L), dtype=np.uint8)
    letters = np.frombuffer(b"ACGTN", np.uint8)
    w = letters[win]
    with_n = np.nonzero(rng.random(n) < 0.1)[0]
    w[with_n, rng.integers(0, W, len(with_n))] = ord("N")
    reads = np.concatenate([letters[rd].ravel(), [0]]).astype(np.uint8)
    ctgs = np.concatenate([w.ravel(), [0]]).astype(np.uint8)
    return m.PairBatch(reads, np.arange(n, dtype=np.int64) * L, np.full(n, L, np.int32),
                       ctgs, np.arange(n, dtype=np.int64) * W, np.full(n, W, np.int32))


def table(res):
    return np.stack([res[k] for k in FIELDS], 1)


def sliced(b, lo, hi):
    return m.PairBatch(b.reads, b.read_off[lo:hi], b.read_len[lo:hi], b.ctgs, b.ctg_off[lo:hi], b.ctg_len[lo:hi])


@pytest.mark.parametrize("read_len", sorted(SHAPES))
def test_forward_shape_matches_16_lane_groups_and_oracle(gpu_aligner, read_len):
    print(f"[gpu] forward shape, reads of {read_len} bases, {NPAIRS} pairs", flush=True)
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    al.set_score2(False)
    b = batch_of(read_len, 1000 + read_len)
    got = table(al.align(b))
    st = al.stats()
    cls = class_of(read_len)
    assert st["class_pairs"][cls] == NPAIRS
    assert (st["class_lanes"][cls], st["class_regs"][cls]) == SHAPES[read_len]
    assert st["class_tag_bits"][cls] == 5  # 32 H: match + gap open = 6, rows * match < 960
    want = np.concatenate([table(al.align(sliced(b, lo, min(lo + SLICE, NPAIRS))))
                           for lo in range(0, NPAIRS, SLICE)])
    st = al.stats()
    assert (st["class_lanes"][cls], st["class_regs"][cls]) == (16, cls + 1)
    assert st["class_tag_bits"][cls] == 4
    assert np.array_equal(got, want)
    assert (got[:, 0] > 0).mean() > 0.9
    idx = np.linspace(0, NPAIRS - 1, 200).astype(np.int64)
    o = oracle_batch(b.reads, b.read_off[idx], b.read_len[idx], b.ctgs, b.ctg_off[idx], b.ctg_len[idx],
                     m.MHM2_SCORING)
    assert np.array_equal(got[idx], np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin",
                                                              "query_end")], 1))


# (scoring, read length, H scale of the narrow shape: 32 H needs match + gap open <= 7 and
# rows * match < 960)
TAG_BITS_CASES = [((2, 3, 6, 2, 2), 150, 4), ((2, 3, 6, 2, 2), 250, 4), ((4, 3, 3, 2, 2), 150, 5),
                  ((4, 3, 3, 2, 2), 250, 4), ((1, 2, 6, 1, 1), 60, 5), ((3, 1, 4, 2, 1), 193, 5)]


@pytest.mark.parametrize("scoring,read_len,tag_bits", TAG_BITS_CASES)
def test_forward_shape_h_scale_by_scoring(gpu_aligner, scoring, read_len, tag_bits):
    """The narrow shapes keep 16 H or 32 H as the scoring and the class's rows allow: both scales
    against 16-lane groups (16 H) and the oracle under the same scoring."""
    print(f"[gpu] forward shape, scoring {scoring}, reads of {read_len} bases", flush=True)
    al = gpu_aligner
    al.set_scoring(scoring)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    al.set_score2(False)
    try:
        b = batch_of(read_len, 5000 + read_len + 7 * scoring[0])
        got = table(al.align(b))
        st = al.stats()
        cls = class_of(read_len)
        assert (st["class_lanes"][cls], st["class_regs"][cls]) == SHAPES[read_len]
        assert st["class_tag_bits"][cls] == tag_bits
        want = np.concatenate([table(al.align(sliced(b, lo, min(lo + SLICE, NPAIRS))))
                               for lo in range(0, NPAIRS, SLICE)])
        assert np.array_equal(got, want)
        assert (got[:, 0] > 0).mean() > 0.9
        idx = np.linspace(0, NPAIRS - 1, 200).astype(np.int64)
        o = oracle_batch(b.reads, b.read_off[idx], b.read_len[idx], b.ctgs, b.ctg_off[idx], b.ctg_len[idx],
                         scoring)
        assert np.array_equal(got[idx], np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin",
                                                                  "query_end")], 1))
    finally:
        al.set_scoring(m.MHM2_SCORING)


@pytest.mark.parametrize("read_len", [30, 60, 100, 150, 193, 250])
def test_forward_shape_score2_matches_16_lane_groups_and_oracle(gpu_aligner, read_len):
    """SSW's second-best score (score2: the padded edge chunks and the per-row maxima over a
    group's lanes) on the narrow shapes, against 16-lane groups and the oracle."""
    print(f"[gpu] forward shape with score2, reads of {read_len} bases, {NPAIRS} pairs", flush=True)
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    al.set_score2(True)
    try:
        b = batch_of(read_len, 2000 + read_len)

        def run(batch):
            al.stage(batch)
            al.align_staged()
            al.wait()
            return table(al.results()), al.score2()

        got, got2 = run(b)
        st = al.stats()
        cls = class_of(read_len)
        assert (st["class_lanes"][cls], st["class_regs"][cls]) == SHAPES[read_len]
        parts = [run(sliced(b, lo, min(lo + SLICE, NPAIRS))) for lo in range(0, NPAIRS, SLICE)]
        assert np.array_equal(got, np.concatenate([p[0] for p in parts]))
        assert np.array_equal(got2, np.concatenate([p[1] for p in parts]))
        idx = np.linspace(0, NPAIRS - 1, 200).astype(np.int64)
        o, o2 = oracle_batch2(b.reads, b.read_off[idx], b.read_len[idx], b.ctgs, b.ctg_off[idx], b.ctg_len[idx],
                              m.MHM2_SCORING)
        assert np.array_equal(got[idx], np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin",
                                                                  "query_end")], 1))
        assert np.array_equal(got2[idx], o2)
    finally:
        al.set_score2(False)


@pytest.mark.parametrize("read_len", [60, 100, 150, 250])
def test_forward_shape_contig_as_ssw_ref(gpu_aligner, read_len):
    """The other orientation (the window is SSW's reference): rows are still the shorter read, but
    the end point's tie-break runs over columns first."""
    print(f"[gpu] forward shape, window as SSW-ref, reads of {read_len} bases", flush=True)
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_mode(m.MODE_FAST)
    al.set_score2(False)
    al.set_orientation(False)
    try:
        b = batch_of(read_len, 3000 + read_len)
        got = table(al.align(b))
        st = al.stats()
        cls = class_of(read_len)
        assert (st["class_lanes"][cls], st["class_regs"][cls]) == SHAPES[read_len]
        want = np.concatenate([table(al.align(sliced(b, lo, min(lo + SLICE, NPAIRS))))
                               for lo in range(0, NPAIRS, SLICE)])
        assert np.array_equal(got, want)
        idx = np.linspace(0, NPAIRS - 1, 200).astype(np.int64)
        o = oracle_batch(b.ctgs, b.ctg_off[idx], b.ctg_len[idx], b.reads, b.read_off[idx], b.read_len[idx],
                         m.MHM2_SCORING)
        assert np.array_equal(got[idx], np.stack([o[k] for k in ("score", "query_begin", "query_end", "ref_begin",
                                                                  "ref_end")], 1))
    finally:
        al.set_orientation(True)


# rows above 256 (the 64-lane classes): read length -> forward shape
SHAPES64 = {300: (16, 12), 450: (16, 16), 600: (16, 20), 700: (32, 12), 850: (32, 14), 1000: (32, 16)}


@pytest.mark.parametrize("read_len", sorted(SHAPES64))
def test_forward_shape_64_lane_classes(gpu_aligner, read_len):
    """Reads of 257..1024 bases: 16 or 32 lanes per pair forward, against 64-lane groups (slices
    of 4,000 pairs) and the oracle."""
    n, W = 20_000, 1024
    print(f"[gpu] forward shape, reads of {read_len} bases against {W}-base windows, {n} pairs", flush=True)
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    al.set_score2(False)
    b = batch_of(read_len, 4000 + read_len, n, W)
    got = table(al.align(b))
    st = al.stats()
    cls = class_of(read_len)
    assert st["class_pairs"][cls] == n
    assert (st["class_lanes"][cls], st["class_regs"][cls]) == SHAPES64[read_len]
    want = np.concatenate([table(al.align(sliced(b, lo, min(lo + SLICE, n)))) for lo in range(0, n, SLICE)])
    st = al.stats()
    assert (st["class_lanes"][cls], st["class_regs"][cls]) == (64, cls - 7)
    assert np.array_equal(got, want)
    assert (got[:, 0] > 0).mean() > 0.85
    idx = np.linspace(0, n - 1, 100).astype(np.int64)
    o = oracle_batch(b.reads, b.read_off[idx], b.read_len[idx], b.ctgs, b.ctg_off[idx], b.ctg_len[idx],
                     m.MHM2_SCORING)
    assert np.array_equal(got[idx], np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin",
                                                              "query_end")], 1))


def mixed_batch(n, seed):
    """n pairs of mixed lengths: reads of 1..300 bases, windows of 1..700, reads copied from their
    window (2% substitutions) where it is long enough, else random; N in a few windows."""
    rng = np.random.default_rng(seed)
    rl = rng.integers(1, 301, n)
    wl = rng.integers(1, 701, n)
    letters = np.frombuffer(b"ACGT", np.uint8)
    win = letters[rng.integers(0, 4, int(wl.sum()))]
    wo = np.concatenate([[0], np.cumsum(wl)[:-1]])
    reads = letters[rng.integers(0, 4, int(rl.sum()))]
    ro = np.concatenate([[0], np.cumsum(rl)[:-1]])
    fits = np.nonzero((wl >= rl) & (rng.random(n) < 0.9))[0]
    for p in fits:
        s = wo[p] + rng.integers(0, wl[p] - rl[p] + 1)
        reads[ro[p]:ro[p] + rl[p]] = win[s:s + rl[p]]
    sub = np.nonzero(rng.random(len(reads)) < 0.02)[0]
    reads[sub] = letters[rng.integers(0, 4, len(sub))]
    nw = rng.integers(0, len(win), len(win) // 2000)
    win[nw] = ord("N")
    return m.PairBatch(np.concatenate([reads, [0]]).astype(np.uint8), ro.astype(np.int64), rl.astype(np.int32),
                       np.concatenate([win, [0]]).astype(np.uint8), wo.astype(np.int64), wl.astype(np.int32))


@pytest.mark.parametrize("score2", [False, True])
def test_mixed_lengths_every_class_in_one_job(gpu_aligner, score2):
    """600k pairs of mixed lengths: every 16-lane class and a 64-lane one in one job, each large
    enough for its narrow shape, the task ranges of the classes side by side; against the same
    pairs in slices of 4,000 (16- and 64-lane groups) and the oracle."""
    n = 600_000
    print(f"[gpu] mixed lengths, {n} pairs, score2={score2}", flush=True)
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    al.set_score2(score2)
    try:
        b = mixed_batch(n, 77 + score2)

        def run(batch):
            al.stage(batch)
            al.align_staged()
            al.wait()
            return table(al.results()), (al.score2() if score2 else None)

        got, got2 = run(b)
        st = al.stats()
        narrow = [k for k in range(8) if st["class_pairs"][k] and st["class_lanes"][k] < 16]
        assert len(narrow) >= 4, (st["class_pairs"][:8], st["class_lanes"][:8])
        parts = [run(sliced(b, lo, min(lo + SLICE, n))) for lo in range(0, n, SLICE)]
        assert np.array_equal(got, np.concatenate([p[0] for p in parts]))
        if score2:
            assert np.array_equal(got2, np.concatenate([p[1] for p in parts]))
        idx = np.linspace(0, n - 1, 300).astype(np.int64)
        if score2:
            o, o2 = oracle_batch2(b.reads, b.read_off[idx], b.read_len[idx], b.ctgs, b.ctg_off[idx],
                                  b.ctg_len[idx], m.MHM2_SCORING)
            assert np.array_equal(got2[idx], o2)
        else:
            o = oracle_batch(b.reads, b.read_off[idx], b.read_len[idx], b.ctgs, b.ctg_off[idx], b.ctg_len[idx],
                             m.MHM2_SCORING)
        assert np.array_equal(got[idx], np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin",
                                                                  "query_end")], 1))
    finally:
        al.set_score2(False)
