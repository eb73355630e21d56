"""The local assembly oracle (oracle/locassm_oracle.c) against answers derived
by hand from the reference's rules (kernel.hpp:71-161, kernel.cpp:243-545,
driver.cpp:141-319) and against closed forms computed from read geometry
alone.  The reference's local assembly cannot be built here and ships no
tests or fixtures, so these cases are what pins the restatement."""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
from mhm2_staging_amd import locassm_data as ld  # noqa: E402
from oracle_lib import la_ext, la_extend_batch, la_params, la_rating, la_sort_keys, la_walk_side  # noqa: E402

COMP = str.maketrans("ACGTN", "TGCAN")


def rc(s):
    return s.translate(COMP)[::-1]


def rand_seq(rng, n):
    return "".join(rng.choice(list("ACGT"), n))


# ---- ratings: get_base_rating, kernel.hpp:71-82 ---------------------------
# depth 10: min_viable = max(2.0, 2.0) = 2, min_expected = max(3.0, 2.0) = 3
# depth 20: min_viable = 4, min_expected = 6
# depth 12: min_viable = 2.4, min_expected = 3.6
RATING_CASES = [
    (10, 0, 0, 0),   # no votes
    (10, 1, 0, 1),   # one vote
    (10, 1, 1, 1),
    (20, 3, 3, 2),   # 3 < min_viable 4
    (20, 4, 3, 3),   # 4 <= n < 6, hi 3 < 4
    (20, 5, 4, 4),   # 4 <= n < 6, hi 4 >= 4
    (20, 6, 3, 5),   # n >= 6, hi < 4
    (20, 6, 5, 6),   # n >= 6, 4 < hi < 6
    (20, 6, 4, 7),   # n >= 6, hi == min_viable: neither 5 (hi < 4) nor 6 (4 < hi): 7
    (20, 6, 6, 7),   # hi >= min_expected
    (12, 3, 3, 4),   # 2.4 <= 3 < 3.6, hi 3 >= 2.4
    (12, 4, 2, 5),   # n >= 3.6, hi 2 < 2.4
    (12, 4, 3, 6),   # 2.4 < 3 < 3.6
    (12, 4, 4, 7),
    (19, 6, 4, 6),   # min_viable 3.8, min_expected 5.7: 3.8 < 4 < 5.7
    (0, 2, 2, 7),    # depth 0: both thresholds 2
    (0, 2, 1, 5),
    (10, 2, 2, 4),   # 2 <= n < 3, hi 2 >= 2
    (10, 2, 1, 3),
]


@pytest.mark.parametrize("depth,n,hi,want", RATING_CASES)
def test_rating(depth, n, hi, want):
    assert la_rating(la_params(), n, hi, depth) == want


# ---- extension choice: set_ext, kernel.hpp:112-161 ------------------------
# (depth, low votes ACGT, high votes ACGT, expected ext, expected count)
EXT_CASES = [
    (10, [0, 0, 0, 0], [0, 0, 0, 0], "X", 0),      # top rating 0 is not above the threshold 0
    (10, [0, 1, 0, 0], [0, 0, 0, 0], "C", 1),      # top 1 <= 3 and runner-up 0
    (10, [1, 1, 0, 0], [1, 1, 0, 0], "X", 0),      # top 1, runner-up 1: contested
    (20, [3, 0, 0, 0], [3, 0, 0, 0], "A", 3),      # rating 2, uncontested
    (20, [3, 1, 0, 0], [3, 0, 0, 0], "X", 0),      # rating 2 vs 1
    (20, [0, 0, 6, 3], [0, 0, 3, 0], "G", 6),      # top 5 (G), runner-up 2 (T: n 3 < 4)
    (20, [0, 0, 6, 4], [0, 0, 3, 3], "X", 0),      # top 5, runner-up 3 (T: 4 <= n < 6, hi < 4)
    (20, [6, 4, 0, 0], [5, 3, 0, 0], "A", 6),      # top 6, runner-up 3
    (20, [6, 5, 0, 0], [5, 4, 0, 0], "X", 0),      # top 6, runner-up 4
    (10, [5, 1, 0, 0], [5, 0, 0, 0], "A", 5),      # top 7, runner-up 1
    (10, [0, 6, 0, 4], [0, 6, 0, 4], "C", 6),      # two 7s, third 0, C has more votes
    (10, [4, 6, 0, 0], [4, 3, 0, 0], "C", 6),      # two 7s: A sorts first (hi 4 > 3) but C has more votes
    (10, [5, 5, 0, 0], [5, 4, 0, 0], "F", 0),      # two 7s with equal votes: fork
    (10, [3, 3, 9, 0], [3, 3, 9, 0], "F", 0),      # three 7s: fork
    (10, [3, 0, 0, 9], [3, 0, 0, 3], "T", 9),      # two 7s tied on hi: T sorts first on votes and wins
]


@pytest.mark.parametrize("depth,low,hi,want,count", EXT_CASES)
def test_ext(depth, low, hi, want, count):
    assert la_ext(la_params(), low, hi, depth) == (want, count)


def test_exchange_sort_orders_keys():
    """kernel.hpp:100-110 visits every ordered pair and swaps when the first
    ranks before (or ties with) the second: on every key combination of a
    small range the result is in descending (rating, hi, low) order, so the
    extension depends on the keys alone and not on the sort's tie handling."""
    vals = [(r, h, n) for r in (0, 1, 7) for h in (0, 1) for n in (0, 2)]
    for combo in itertools.product(vals, repeat=4):
        out = [tuple(int(x) for x in row) for row in la_sort_keys(np.array(combo).reshape(-1))]
        assert out == sorted(combo, reverse=True), combo
    rng = np.random.default_rng(3)
    for _ in range(3000):
        combo = rng.integers(0, 4, (4, 3))
        out = [tuple(int(x) for x in row) for row in la_sort_keys(combo.reshape(-1))]
        assert out == sorted([tuple(int(x) for x in r) for r in combo], reverse=True)


# ---- closed-form walks -------------------------------------------------------

def coverage_walk(c, k, intervals, limit):
    """Bases appended from contig end c at mer length k when error-free reads
    [s, e) of a repeat-free sequence are all the reads: window x..x+k (k+1
    bases, a mer and the base after it) must lie inside one read."""
    x, w = c - k, 0
    while w < limit and any(s <= x + w and x + w + k + 1 <= e for s, e in intervals):
        w += 1
    return w


def repeat_free(g, k=21):
    mers = [g[i:i + k] for i in range(len(g) - k + 1)]
    return len(set(mers)) == len(mers)


@pytest.mark.parametrize("seed", range(6))
def test_walk_matches_read_coverage(seed):
    rng = np.random.default_rng(seed)
    g = rand_seq(rng, 3000)
    assert repeat_free(g)
    c = 1200
    reads = []
    for _ in range(int(rng.integers(1, 12))):
        s = int(rng.integers(c - 150, c + 400))
        L = int(rng.integers(60, 151))
        reads.append((s, s + L))
    p = la_params(kmer_len=int(rng.choice([21, 33, 45, 55])), walk_len_limit=int(rng.choice([50, 300])))
    ks = list(range(p.kmer_len, 20, -8))   # every walk ends in X: the schedule only downshifts
    want = max(coverage_walk(c, k, reads, p.walk_len_limit) for k in ks)
    walk, term, iters, _ = la_walk_side(p, g[:c], 10.0, [(g[s:e], "I" * (e - s)) for s, e in reads])
    assert iters == len(ks)
    assert term == "X"
    assert walk == g[c:c + want]


def test_left_walk_is_reverse_complemented():
    rng = np.random.default_rng(11)
    g = rand_seq(rng, 2000)
    ctg = g[500:1500]
    # left reads are given in the orientation of the reverse-complemented contig
    left = [(rc(g[440:600]), "I" * 160), (rc(g[400:560]), "I" * 160)]
    b = ld.from_contigs([(ctg, 10.0, left, [])])
    p = la_params(kmer_len=33, walk_len_limit=300)
    ll, le, rl, _ = la_extend_batch(p, b)
    want = coverage_walk(1000, 33, [(1500 - 600, 1500 - 440), (1500 - 560, 1500 - 400)], 300)
    assert want == 100 and rl[0] == 0
    assert ll[0] == want and bytes(le[0, :want]).decode() == g[500 - want:500]


def test_fork_upshifts_to_the_largest_mer():
    """Two haplotypes, 3 reads each, differ 50 bases past the contig end.
    At depth 10 each branch has 3 votes, all high quality: both rate 7 with
    equal votes, a fork, at every mer length.  The schedule upshifts from 33
    by 8 while walks end in forks, until 129 > 121 leaves the loop: 12 walks,
    each stopping at the fork after 50 bases."""
    rng = np.random.default_rng(12)
    g = rand_seq(rng, 3000)
    c = 1000
    alt = g[:c + 50] + ("A" if g[c + 50] != "A" else "C") + g[c + 51:]
    reads = []
    for s in (c - 140, c - 130, c - 125):
        reads.append((g[s:s + 300], "I" * 300))
        reads.append((alt[s:s + 300], "I" * 300))
    walk, term, iters, _ = la_walk_side(la_params(kmer_len=33, walk_len_limit=300), g[:c], 10.0, reads)
    assert (walk, term, iters) == (g[c:c + 50], "F", 12)


def test_repeat_stops_after_one_period():
    """A periodic continuation (period 25) makes the walk meet its start mer
    again after 25 bases ('R'); repeats upshift like forks."""
    rng = np.random.default_rng(13)
    unit = rand_seq(rng, 25)
    per = unit * 40
    ctg = rand_seq(rng, 200) + per[:150]
    reads = [(per[:300], "I" * 300)]
    walk, term, iters, _ = la_walk_side(la_params(kmer_len=33, walk_len_limit=300), ctg, 10.0, reads)
    assert (walk, term, iters) == (per[150:175], "R", 12)


def test_contested_branch_downshifts_then_stops():
    """A fork at k=33 followed by X at k=41 (upshifting): the loop stops."""
    rng = np.random.default_rng(14)
    g = rand_seq(rng, 3000)
    c = 1000
    alt = g[:c + 50] + ("A" if g[c + 50] != "A" else "C") + g[c + 51:]
    # reads start 100 bases inside: windows of 42 bases reach to read end - 42
    reads = []
    for s in (c - 36, c - 35, c - 34):
        reads.append((g[s:s + 120], "I" * 120))
        reads.append((alt[s:s + 120], "I" * 120))
    walk, term, iters, _ = la_walk_side(la_params(kmer_len=33, walk_len_limit=300), g[:c], 10.0, reads)
    # k=33: start mer at c-33 lies in the reads starting at c-36..c-34 -> fork at +50
    # k=41: start mer at c-41 precedes every read: X at once; shift was +8: stop
    assert (walk, term, iters) == (g[c:c + 50], "X", 2)


def test_quality_and_n_rules():
    rng = np.random.default_rng(15)
    g = rand_seq(rng, 1000)
    c = 400
    s = c - 60
    read = g[s:s + 150]
    # the base 20 past the contig end has quality 9 (below 10): no vote, dead end
    q = ["I"] * 150
    q[c + 20 - s] = chr(33 + 9)
    p = la_params(kmer_len=33, walk_len_limit=300)
    walk = la_walk_side(p, g[:c], 10.0, [(read, "".join(q))])[0]
    assert walk == g[c:c + 20]
    # quality 10 is a low-quality vote: one vote still extends (rating 1, uncontested)
    q[c + 20 - s] = chr(33 + 10)
    walk = la_walk_side(p, g[:c], 10.0, [(read, "".join(q))])[0]
    assert walk == g[c:c + 90] and coverage_walk(c, 33, [(s, s + 150)], 300) == 90
    # an N as the next base gives no vote; an N inside a mer is part of the key
    # (the GPU kernel counts mers that hold N, kernel.cpp:343-360)
    ctg = g[:c - 10] + "N" + g[c - 9:c]
    read_n = ctg[s:] + g[c:s + 150]
    walk = la_walk_side(p, ctg, 10.0, [(read_n, "I" * len(read_n))])[0]
    assert walk == g[c:c + 90]


def test_depth_is_truncated_to_int():
    """set_ext takes an int depth (kernel.hpp:112, called with the double
    contig depth): 20.9 behaves as 20 and 19.9 as 19.  With 6 votes of which
    4 high quality: depth 20 rates 7 (hi == min_viable 4), depth 19 rates 6
    (3.8 < 4 < 5.7); against a second base rated 4 (5 votes, 4 high), 7 wins
    and 6 does not."""
    rng = np.random.default_rng(16)
    g = rand_seq(rng, 1000)
    c = 400
    alt = g[:c + 30] + ("A" if g[c + 30] != "A" else "C") + g[c + 31:]
    reads = []
    for i in range(6):
        qv = "I" if i < 4 else chr(33 + 15)
        r = g[c - 50:c + 100]
        q = ["I"] * 150
        q[80] = qv
        reads.append((r, "".join(q)))
    for i in range(5):
        r = alt[c - 50:c + 100]
        q = ["I"] * 150
        q[80] = "I" if i < 4 else chr(33 + 15)
        reads.append((r, "".join(q)))
    p = la_params(kmer_len=33, walk_len_limit=300)
    w20 = la_walk_side(p, g[:c], 20.9, reads)[0]
    w19 = la_walk_side(p, g[:c], 19.9, reads)[0]
    assert len(w19) == 30 and len(w20) > 30 and w20[:31] == g[c:c + 31]


def test_max_reads_limit():
    rng = np.random.default_rng(17)
    g = rand_seq(rng, 1000)
    c = 400
    short = [(g[c - 40:c + 10], "I" * 50)] * 3000
    far = (g[c - 40:c + 110], "I" * 150)
    p = la_params(kmer_len=33, walk_len_limit=300)
    assert la_walk_side(p, g[:c], 10.0, short + [far])[0] == g[c:c + 10]  # the 3001st read is not counted
    assert la_walk_side(p, g[:c], 10.0, [far] + short)[0] == g[c:c + coverage_walk(c, 33, [(c - 40, c + 110)], 300)]


def test_degenerate_sides():
    p = la_params(kmer_len=33, walk_len_limit=300)
    g = "ACGT" * 20
    assert la_walk_side(p, g[:20], 10.0, [(g, "I" * len(g))])[2] == 0      # contig shorter than 21
    assert la_walk_side(la_params(kmer_len=41), g[:40], 10.0, [(g, "I" * 80)])[2] == 0  # kmer_len > contig
    assert la_walk_side(p, g, 10.0, [])[:2] == ("", "X")
    assert la_walk_side(la_params(kmer_len=33, walk_len_limit=0), g, 10.0, [(g, "I" * 80)])[0] == ""
    b = ld.from_contigs([("ACGTacgt" * 10, 10.0, [], [])])
    with pytest.raises(ValueError):
        la_extend_batch(p, b)


def test_read_order_does_not_matter():
    b = ld.generate(300, seed=5)
    p = la_params(kmer_len=33, walk_len_limit=360)
    want = la_extend_batch(p, b)
    rng = np.random.default_rng(1)
    idx = b.read_idx.copy()
    for first in (b.left_first, b.right_first):
        for c in range(b.nctgs):
            seg = idx[first[c]:first[c + 1]]
            if len(seg) <= p.max_reads:
                rng.shuffle(seg)
    b2 = ld.LocAssmBatch(**{**b.__dict__, "read_idx": idx})
    got = la_extend_batch(p, b2)
    for a, w in zip(got, want):
        assert np.array_equal(a, w)


def test_generated_batch_reaches_every_outcome():
    b = ld.generate(400, seed=2)
    p = la_params(kmer_len=33, walk_len_limit=360)
    terms = np.zeros(3, np.int64)
    for c in range(b.nctgs):
        for side in ("left", "right"):
            seq = b.ctg(c) if side == "right" else rc(b.ctg(c))
            terms += la_walk_side(p, seq, float(b.depth[c]), b.reads_of(c, side))[3]
    assert (terms > 0).all(), terms
    ll, _, rl, _ = la_extend_batch(p, b)
    assert (ll > 0).mean() > 0.3 and (rl > 0).mean() > 0.3 and (rl == 360).any()


def test_kernel_extension_rule_matches_oracle():
    """mhmla::choose_ext (the kernel's rule, built for the host) against the
    oracle's set_ext on every vote combination up to 4 votes per base at 12
    depths and 2.5M random ones (tests/cpp/locassm_core_check.cpp)."""
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "locassm_core_check")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/locassm_core_check not built (run __graft_entry__.build())")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0 and out.stdout.startswith("ok"), out.stdout + out.stderr
