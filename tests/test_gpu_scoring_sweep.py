"""Alignment results, second-best scores and CIGARs on the GPU under many
scorings (random draws inside the accepted range plus extremes) against the
oracle, and rejection of the scorings SSW's lazy-F rule makes inexact
(include/mhmsw.h)."""
import numpy as np
import pytest

import mhm2_staging_amd as m
from test_gpu_cigar import check_vs_oracle
from test_gpu_parity import assert_same, gpu_table, oracle_table

pytestmark = pytest.mark.gpu


def accepted(sc):
    match, mismatch, gap_open, gap_ext, amb = sc
    worst = max(mismatch, amb)
    return (1 <= match <= 127 and 0 <= mismatch <= 127 and 1 <= gap_open <= 127 and 1 <= gap_ext <= 127
            and 0 <= amb <= 127 and gap_ext <= gap_open and gap_open + gap_ext >= worst and 2 * gap_ext >= worst)


def sweep_scorings():
    rng = np.random.default_rng(77)
    out = []
    while len(out) < 20:
        sc = (int(rng.integers(1, 9)), int(rng.integers(0, 10)), int(rng.integers(1, 14)), int(rng.integers(1, 8)),
              int(rng.integers(0, 8)))
        if accepted(sc) and sc not in out:
            out.append(sc)
    # extremes: the largest costs, free mismatches and N, match far above the gap costs
    out += [(127, 127, 127, 127, 127), (1, 0, 1, 1, 0), (3, 6, 6, 3, 6), (50, 10, 20, 10, 10), (9, 1, 2, 1, 1)]
    return out


def mixed_batch(seed):
    """Config-2 pairs plus short, N-rich, two-letter and unrelated pairs."""
    rng = np.random.default_rng(seed)
    b = m.generate(2, npairs=1200, seed=seed)
    reads = [b.read(p) for p in range(b.npairs)]
    ctgs = [b.ctg(p) for p in range(b.npairs)]
    for _ in range(600):
        alph = list(rng.choice(["ACGT", "ACGTN", "AC", "ACGTNNN"]))
        w = "".join(rng.choice(alph, int(rng.integers(1, 80))))
        r = list(w[int(rng.integers(0, max(1, len(w) // 3))):])
        for _ in range(int(rng.integers(0, 6))):
            if r:
                r[int(rng.integers(0, len(r)))] = str(rng.choice(alph))
        reads.append("".join(r) if r else "A")
        ctgs.append(w)
    return m.from_strings(reads, ctgs)


@pytest.mark.parametrize("sc", sweep_scorings())
def test_scoring_sweep_vs_oracle(gpu_aligner, sc):
    b = mixed_batch(sum(sc))
    al = gpu_aligner
    al.set_orientation(True)
    print(f"[gpu] scoring sweep {sc}", flush=True)
    want = oracle_table(b, sc)
    for mode in (m.MODE_FAST, m.MODE_PLAIN):
        al.set_scoring(sc)
        al.set_mode(mode)
        res = al.align(b)
        assert_same(gpu_table(res), want, b)
    al.set_mode(m.MODE_FAST)
    # CIGARs and mismatch counts of the same pairs
    check_vs_oracle(al, b, sc)


def test_rejected_scorings(gpu_aligner):
    for sc in [(1, 3, 1, 1, 2),    # gap_open + gap_ext < mismatch
               (1, 5, 4, 2, 2),    # 2 * gap_ext < mismatch
               (1, 1, 4, 1, 3),    # 2 * gap_ext < ambiguity
               (1, 1, 2, 3, 1),    # gap_ext > gap_open
               (0, 1, 1, 1, 1),    # match 0
               (1, 128, 128, 128, 1)]:
        assert not accepted(sc)
        with pytest.raises(m.MhmswError):
            gpu_aligner.set_scoring(sc)
    gpu_aligner.set_scoring(m.MHM2_SCORING)
