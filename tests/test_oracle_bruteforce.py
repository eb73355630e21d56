"""First-principles check of the oracle on tiny inputs.

Enumerates every local alignment path (diagonal steps and gap runs, affine
cost gap_open + (L-1)*gap_ext, first and last step diagonal) and derives:
  score      = best path score (0 if none is positive),
  end        = least (SSW-ref index, SSW-query index) among ends of best paths,
  begin      = among best paths ending there, the greatest SSW-ref start,
               then the greatest SSW-query start (what SSW's reverse pass
               returns, ssw_core.cpp:1162-1178).
A mistake in a recurrence term, a tie-break direction or a swapped role fails
here independently of the DP the oracle uses.
"""
import numpy as np
import pytest

from oracle_lib import oracle_pairs

CODE = {"A": 0, "C": 1, "G": 2, "T": 3}


def subst(x, y, sc):
    m, mm, go, ge, amb = sc
    cx, cy = CODE.get(x, 4), CODE.get(y, 4)
    if cx == 4 or cy == 4:
        return -amb
    return m if cx == cy else -mm


def all_paths(r, q, sc):
    """Yield (score, start_i, start_j, end_i, end_j) for every path."""
    m, n = len(r), len(q)
    go, ge = sc[2], sc[3]
    out = []

    def extend(i, j, i0, j0, score, last_gap):
        # (i, j) is the last diagonal cell; record it as an end
        out.append((score, i0, j0, i, j))
        # next diagonal step
        if i + 1 < m and j + 1 < n:
            extend(i + 1, j + 1, i0, j0, score + subst(r[i + 1], q[j + 1], sc), None)
        # gap run along ref (skip L ref bases), then a diagonal step; optional opposite gap in between
        for L in range(1, m):
            for L2 in range(0, n):
                ni, nj = i + L + 1, j + L2 + 1
                if ni >= m or nj >= n:
                    break
                cost = go + (L - 1) * ge + (go + (L2 - 1) * ge if L2 else 0)
                extend(ni, nj, i0, j0, score - cost + subst(r[ni], q[nj], sc), None)
        # gap run along query, then an optional opposite gap, then a diagonal step
        for L in range(1, n):
            for L2 in range(0, m):
                ni, nj = i + L2 + 1, j + L + 1
                if ni >= m or nj >= n:
                    break
                cost = go + (L - 1) * ge + (go + (L2 - 1) * ge if L2 else 0)
                extend(ni, nj, i0, j0, score - cost + subst(r[ni], q[nj], sc), None)

    for i0 in range(m):
        for j0 in range(n):
            extend(i0, j0, i0, j0, subst(r[i0], q[j0], sc), None)
    return out


def brute(r, q, sc):
    paths = all_paths(r, q, sc)
    best = max([p[0] for p in paths] + [0])
    if best <= 0:
        return (0, -1, -1, 0, 0)
    ends = sorted((p[3], p[4]) for p in paths if p[0] == best)
    ei, ej = ends[0]
    starts = sorted(((p[1], p[2]) for p in paths if p[0] == best and (p[3], p[4]) == (ei, ej)), reverse=True)
    bi, bj = starts[0]
    return (best, bi, ei, bj, ej)


@pytest.mark.parametrize("sc", [(1, 3, 5, 2, 2), (1, 1, 1, 1, 1), (2, 2, 3, 1, 2), (2, 4, 4, 2, 1)])
def test_oracle_equals_path_enumeration(sc):
    rng = np.random.default_rng(hash(sc) % 2**32)
    refs, qrys = [], []
    for k in range(120):
        alph = ["ACGT", "AC", "ACN", "AAC"][k % 4]
        refs.append("".join(rng.choice(list(alph), int(rng.integers(1, 6)))))
        qrys.append("".join(rng.choice(list(alph), int(rng.integers(1, 6)))))
    res = oracle_pairs(refs, qrys, sc)
    for r, q, o in zip(refs, qrys, res):
        got = (int(o["score"]), int(o["ref_begin"]), int(o["ref_end"]), int(o["query_begin"]), int(o["query_end"]))
        assert got == brute(r, q, sc), (r, q, sc)


def test_bruteforce_catches_a_planted_tie_break_flip():
    """The check has teeth: flipping the end tie-break to 'greatest index' disagrees somewhere."""
    sc = (1, 1, 1, 1, 1)
    rng = np.random.default_rng(7)
    differs = 0
    for _ in range(60):
        r = "".join(rng.choice(list("AC"), 5))
        q = "".join(rng.choice(list("AC"), 5))
        paths = all_paths(r, q, sc)
        best = max([p[0] for p in paths] + [0])
        if best <= 0:
            continue
        ends = sorted((p[3], p[4]) for p in paths if p[0] == best)
        if ends[0] != ends[-1]:
            differs += 1
    assert differs > 0
