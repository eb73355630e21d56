"""The oracle's CIGAR and mismatch count (klign's call with report_cigar set).

* fixtures recorded from the reference SSW (tests/golden/make_golden_cigar.py);
* when oracle/_ref is built (this container), fresh seeded corpora of every
  workload generator against the reference directly;
* invariants that hold for any correct answer, independent of how the
  restatement computes it: the CIGAR walks the whole SSW-query (soft clips
  before the begin and after the end) and exactly the aligned SSW-ref span,
  rescoring its path gives the pair's score, '=' / 'X' agree with the bases,
  and mismatches = X + I + D positions.
"""
import os

import numpy as np
import pytest

from oracle_lib import (have_ref, oracle_batch, oracle_cigar_batch, ref_cigar_batch)

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FIELDS = ("score", "ref_begin", "ref_end", "query_begin", "query_end")
OP_M, OP_I, OP_D, OP_S, OP_EQ, OP_X = 0, 1, 2, 4, 7, 8


def table(res):
    return np.stack([res[k] for k in FIELDS], 1).astype(np.int32)


def load(name):
    z = np.load(os.path.join(GOLD, name))
    return {k: z[k] for k in z.files}


def split_ops(nops, flat):
    off = np.concatenate([[0], np.cumsum(nops)])
    return [flat[off[p]:off[p + 1]] for p in range(len(nops))]


def check_against(o, mism, nops, ops_list):
    res, omism, onops, oops = o
    np.testing.assert_array_equal(omism, mism)
    np.testing.assert_array_equal(onops, nops)
    for p, want in enumerate(ops_list):
        np.testing.assert_array_equal(oops[p, :len(want)], want, err_msg=f"pair {p}")


@pytest.mark.parametrize("name", ["ssw_default", "mhm2_tree", "cigar", "baseline"])
def test_oracle_cigar_matches_reference_golden_small(name):
    g = load("golden_cigar.npz")
    s = load("golden_small.npz")
    sc = tuple(int(v) for v in g[f"small_{name}_scoring"])
    o = oracle_cigar_batch(s["ref_buf"], s["ref_off"], s["ref_len"], s["qry_buf"], s["qry_off"], s["qry_len"], sc)
    np.testing.assert_array_equal(table(o[0]), g[f"small_{name}_res"])
    check_against(o, g[f"small_{name}_mism"], g[f"small_{name}_nops"],
                  split_ops(g[f"small_{name}_nops"], g[f"small_{name}_ops"]))


@pytest.mark.parametrize("name", ["baseline", "mhm2_tree", "ssw_default", "cigar", "bigmatch", "match5"])
def test_oracle_cigar_matches_reference_golden_tiny(name):
    g = load("golden_cigar.npz")
    sc = tuple(int(v) for v in g[f"tiny_{name}_scoring"])
    o = oracle_cigar_batch(g["tiny_ref_buf"], g["tiny_ref_off"], g["tiny_ref_len"], g["tiny_qry_buf"],
                           g["tiny_qry_off"], g["tiny_qry_len"], sc)
    np.testing.assert_array_equal(table(o[0]), g[f"tiny_{name}_res"])
    check_against(o, g[f"tiny_{name}_mism"], g[f"tiny_{name}_nops"],
                  split_ops(g[f"tiny_{name}_nops"], g[f"tiny_{name}_ops"]))


def test_golden_tiny_covers_score0_and_truncated_traces():
    """The fixtures hold what the rules single out: score-0 pairs (the CIGAR
    compares SSW-query[0] with the byte before SSW-ref) and, for match >
    gap_open, CIGARs whose SSW-ref span is shorter than the aligned one."""
    g = load("golden_cigar.npz")
    res = g["tiny_bigmatch_res"]
    assert (res[:, 0] == 0).sum() > 20
    short = 0
    for p, ops in enumerate(split_ops(g["tiny_bigmatch_nops"], g["tiny_bigmatch_ops"])):
        if res[p, 0] > 0:
            span = sum(int(x) >> 4 for x in ops if int(x) & 15 in (OP_EQ, OP_X, OP_D))
            short += span != res[p, 2] - res[p, 1] + 1
    assert short > 20


def codes(buf, off, n):
    t = np.full(256, 4, np.int8)
    for ch, v in zip(b"ACGTUacgtu", (0, 1, 2, 3, 0, 0, 1, 2, 3, 0)):
        t[ch] = v
    return t[buf[off:off + n]]


def rescore(ops, R, Q, rb, qb, sc):
    match, mismatch, go, ge, amb = sc
    i, j, s = qb, rb, 0
    for o in ops:
        ln, op = int(o) >> 4, int(o) & 15
        if op == OP_S:
            continue
        if op in (OP_EQ, OP_X):
            for _ in range(ln):
                a, b = R[j], Q[i]
                assert (a == b) == (op == OP_EQ)
                s += -amb if (a == 4 or b == 4) else (match if a == b else -mismatch)
                i += 1
                j += 1
        elif op == OP_I:
            s -= go + (ln - 1) * ge
            i += ln
        elif op == OP_D:
            s -= go + (ln - 1) * ge
            j += ln
        else:
            raise AssertionError(f"unexpected op {op}")
    return s, i, j


@pytest.mark.parametrize("config", [1, 2, 3, 5])
@pytest.mark.parametrize("sc", [(1, 3, 5, 2, 2), (1, 1, 1, 1, 1), (2, 4, 4, 2, 1)])
def test_oracle_cigar_invariants(config, sc):
    """For match <= gap_open every optimal path begins at the begin point, so
    SSW's traceback ends at the rectangle's corner: the CIGAR is a path from
    begin to end that scores exactly the score."""
    import mhm2_staging_amd as m
    b = m.generate(config, npairs=800, seed=11)
    res, mism, nops, ops = oracle_cigar_batch(b.reads, b.read_off, b.read_len, b.ctgs, b.ctg_off, b.ctg_len, sc)
    np.testing.assert_array_equal(table(res), table(oracle_batch(b.reads, b.read_off, b.read_len, b.ctgs,
                                                                 b.ctg_off, b.ctg_len, sc)))
    for p in range(b.npairs):
        R = codes(b.reads, b.read_off[p], b.read_len[p])
        Q = codes(b.ctgs, b.ctg_off[p], b.ctg_len[p])
        r = res[p]
        o = ops[p, :nops[p]]
        qlen = sum(int(x) >> 4 for x in o if int(x) & 15 in (OP_S, OP_EQ, OP_X, OP_I))
        assert qlen == len(Q)
        if r["score"] == 0:
            continue
        head = int(o[0]) >> 4 if int(o[0]) & 15 == OP_S else 0
        tail = int(o[-1]) >> 4 if int(o[-1]) & 15 == OP_S else 0
        assert head == r["query_begin"] and tail == len(Q) - 1 - r["query_end"]
        s, i, j = rescore(o, R, Q, r["ref_begin"], r["query_begin"], sc)
        assert (s, i, j) == (r["score"], r["query_end"] + 1, r["ref_end"] + 1), p
        assert mism[p] == sum(int(x) >> 4 for x in o if int(x) & 15 in (OP_X, OP_I, OP_D))
        kinds = [int(x) & 15 for x in o]
        assert all(a != c for a, c in zip(kinds, kinds[1:])), "adjacent ops of one kind"


@pytest.mark.skipif(not have_ref(), reason="oracle/_ref (reference SSW) not built here")
@pytest.mark.parametrize("config", [1, 2, 3, 4, 5])
def test_oracle_cigar_matches_reference_fresh(config):
    import mhm2_staging_amd as m
    b = m.generate(config, npairs=1500 if config != 4 else 300, seed=99)
    for sc in ((1, 3, 5, 2, 2), (1, 1, 1, 1, 1), (10, 2, 3, 1, 2)):
        args = (b.reads, b.read_off, b.read_len, b.ctgs, b.ctg_off, b.ctg_len, sc)
        o = oracle_cigar_batch(*args)
        r = ref_cigar_batch(*args)
        np.testing.assert_array_equal(table(o[0]), table(r[0]))
        np.testing.assert_array_equal(o[1], r[1])
        np.testing.assert_array_equal(o[2], r[2])
        np.testing.assert_array_equal(o[3], r[3])
