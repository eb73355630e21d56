"""Write SSW's second-best scores (sw_score_next_best) from the reference SSW.

Run where /root/reference exists (``make -C oracle`` builds
oracle/_ref/libssw_ref.so from the reference's own sources).  Output:
tests/golden/golden_score2.npz with, per scoring of golden_small.npz, the
score2 of its 1560 pairs, plus two constructed sets whose expected values
also come from the reference:

* ``edge``: SSW-refs whose best alignment scores at least 255 - bias (SSW's
  16-bit pass) and whose second-best row sits exactly at, or one past, the
  upper edge of the masked window around the end (the 16-bit pass includes
  position end + maskLen, the 8-bit pass does not, ssw_core.cpp:531 / :841);
* ``short``: SSW-refs of 10..29 bases with a repeated motif, where maskLen is
  15 rather than half the length (klign.cpp:257).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import have_ref, pack_strings, ref_batch2  # noqa: E402

EDGE_SCORINGS = {"baseline": (1, 3, 5, 2, 2), "ssw_default": (2, 2, 3, 1, 2), "cigar": (2, 4, 4, 2, 1)}


def rand_seq(rng, n, alph="ACGT"):
    return "".join(rng.choice(list(alph), n))


def edge_cases(rng, sc, n):
    """Reads (SSW-ref) = X + filler + Y + tail; window (SSW-query) = X + W1,
    Y a copy of a piece of W1 whose last base is placed at end + maskLen + d."""
    match = sc[0]
    bias = max(0, sc[1], sc[4])
    refs, qrys = [], []
    for i in range(n):
        nx = (255 - bias) // match + 1 + int(rng.integers(0, 12))  # forces the 16-bit pass
        x = rand_seq(rng, nx)
        w1 = rand_seq(rng, 40)
        ylen = int(rng.integers(8, 25))
        y = w1[5:5 + ylen]
        tail = int(rng.integers(1, 8))
        d = int(rng.integers(-1, 2))
        # solve m = p + 1 + tail with p = (nx - 1) + m // 2 + d
        m = None
        for cand in range(nx + ylen + tail, 4 * nx + 200):
            if cand - 1 - tail == (nx - 1) + cand // 2 + d:
                m = cand
                break
        if m is None:
            continue
        p = m - 1 - tail
        nfill = p - (ylen - 1) - nx
        if nfill < 1:
            continue
        # filler's first base differs from W1's so the main alignment ends with X
        f0 = rng.choice([b for b in "ACGT" if b != w1[0]])
        read = x + f0 + rand_seq(rng, nfill - 1) + y + rand_seq(rng, tail)
        refs.append(read)
        qrys.append(x + w1)
    return refs, qrys


def short_cases(rng, n):
    refs, qrys = [], []
    for i in range(n):
        m = int(rng.integers(10, 30))
        motif = rand_seq(rng, int(rng.integers(3, 7)))
        read = list(rand_seq(rng, m))
        for _ in range(int(rng.integers(2, 4))):
            a = int(rng.integers(0, max(1, m - len(motif))))
            read[a:a + len(motif)] = list(motif)
        read = "".join(read)[:m]
        win = rand_seq(rng, int(rng.integers(2, 10))) + motif + rand_seq(rng, int(rng.integers(2, 10)))
        refs.append(read)
        qrys.append(win)
    return refs, qrys


def main():
    if not have_ref():
        sys.exit("oracle/_ref/libssw_ref.so is missing: build it with make -C oracle")
    out = {}
    small = np.load(os.path.join(HERE, "golden_small.npz"))
    for name in ("ssw_default", "mhm2_tree", "cigar", "baseline"):
        sc = tuple(int(v) for v in small["scoring_" + name])
        res, s2 = ref_batch2(small["ref_buf"], small["ref_off"], small["ref_len"], small["qry_buf"],
                             small["qry_off"], small["qry_len"], sc, 1)
        assert (np.stack([res[k] for k in ("score", "ref_begin", "ref_end", "query_begin", "query_end")], 1)
                == small["res_" + name]).all()
        out["small_score2_" + name] = s2
    rng = np.random.default_rng(2024)
    for name, sc in EDGE_SCORINGS.items():
        refs, qrys = edge_cases(rng, sc, 120)
        a, b = pack_strings(refs), pack_strings(qrys)
        res, s2 = ref_batch2(*a, *b, sc, 1)
        out.update({f"edge_{name}_ref_buf": a[0], f"edge_{name}_ref_off": a[1], f"edge_{name}_ref_len": a[2],
                    f"edge_{name}_qry_buf": b[0], f"edge_{name}_qry_off": b[1], f"edge_{name}_qry_len": b[2],
                    f"edge_{name}_scoring": np.array(sc, np.int32),
                    f"edge_{name}_res": np.stack([res[k] for k in ("score", "ref_begin", "ref_end", "query_begin",
                                                                   "query_end")], 1).astype(np.int32),
                    f"edge_{name}_score2": s2})
    refs, qrys = short_cases(rng, 400)
    a, b = pack_strings(refs), pack_strings(qrys)
    for name, sc in (("baseline", (1, 3, 5, 2, 2)), ("mhm2_tree", (1, 1, 1, 1, 1))):
        res, s2 = ref_batch2(*a, *b, sc, 1)
        out[f"short_{name}_scoring"] = np.array(sc, np.int32)
        out[f"short_{name}_res"] = np.stack([res[k] for k in ("score", "ref_begin", "ref_end", "query_begin",
                                                              "query_end")], 1).astype(np.int32)
        out[f"short_{name}_score2"] = s2
    out.update({"short_ref_buf": a[0], "short_ref_off": a[1], "short_ref_len": a[2],
                "short_qry_buf": b[0], "short_qry_off": b[1], "short_qry_len": b[2]})
    np.savez_compressed(os.path.join(HERE, "golden_score2.npz"), **out)
    print("wrote golden_score2.npz:", len(out), "arrays")


if __name__ == "__main__":
    main()
