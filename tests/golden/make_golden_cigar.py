"""Write SSW's CIGARs and mismatch counts from the reference SSW.

Run where /root/reference exists (``make -C oracle`` builds
oracle/_ref/libssw_ref.so from the reference's own sources; its
ssw_ref_cigar_batch makes klign's call with report_cigar = true).  Output:
tests/golden/golden_cigar.npz holding, for every set and scoring, the
reference's Alignment::mismatches, the op count and the BAM-encoded ops of
Alignment::cigar, concatenated:

* ``small``: the 1560 pairs of golden_small.npz under its four scorings;
* ``tiny``: 3000 seeded pairs of 1..29 bases (related and unrelated, ACGT,
  ACGTN and two-letter alphabets, so score-0 pairs and ties are common) under
  six scorings, two of them with match > gap_open, where SSW's traceback stops
  on row 0 away from column 0 and its CIGAR leaves out the rest.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle_lib import have_ref, pack_strings, ref_cigar_batch  # noqa: E402

TINY_SCORINGS = {"baseline": (1, 3, 5, 2, 2), "mhm2_tree": (1, 1, 1, 1, 1), "ssw_default": (2, 2, 3, 1, 2),
                 "cigar": (2, 4, 4, 2, 1), "bigmatch": (10, 2, 3, 1, 2), "match5": (5, 1, 2, 1, 1)}
FIELDS = ("score", "ref_begin", "ref_end", "query_begin", "query_end")


def tiny_pairs(rng, n):
    refs, qrys = [], []
    for _ in range(n):
        mode = int(rng.integers(0, 4))
        alph = ["ACGT", "ACGT", "ACGTN", "AC"][mode]
        r = "".join(rng.choice(list(alph), int(rng.integers(1, 30))))
        if rng.random() < 0.5:
            q = list(r)
            for _ in range(int(rng.integers(0, 5))):
                pos = int(rng.integers(0, len(q) + 1))
                op = int(rng.integers(0, 3))
                if op == 0 and len(q) > 1:
                    del q[min(pos, len(q) - 1)]
                elif op == 1:
                    q.insert(pos, str(rng.choice(list(alph))))
                elif q:
                    q[min(pos, len(q) - 1)] = str(rng.choice(list(alph)))
            q = "".join(q) or "A"
        else:
            q = "".join(rng.choice(list(alph), int(rng.integers(1, 30))))
        refs.append(r)
        qrys.append(q)
    return refs, qrys


def record(out, prefix, a, b, sc):
    res, mism, nops, ops = ref_cigar_batch(*a, *b, sc, nthreads=1)
    assert (nops >= 0).all(), "the reference refused a pair"
    out[prefix + "_res"] = np.stack([res[k] for k in FIELDS], 1).astype(np.int32)
    out[prefix + "_mism"] = mism
    out[prefix + "_nops"] = nops
    out[prefix + "_ops"] = np.concatenate([ops[p, :nops[p]] for p in range(len(nops))]).astype(np.uint32)
    out[prefix + "_scoring"] = np.array(sc, np.int32)


def main():
    if not have_ref():
        sys.exit("oracle/_ref/libssw_ref.so is missing: build it with make -C oracle")
    out = {}
    small = np.load(os.path.join(HERE, "golden_small.npz"))
    a = (small["ref_buf"], small["ref_off"], small["ref_len"])
    b = (small["qry_buf"], small["qry_off"], small["qry_len"])
    for name in ("ssw_default", "mhm2_tree", "cigar", "baseline"):
        record(out, "small_" + name, a, b, tuple(int(v) for v in small["scoring_" + name]))
    rng = np.random.default_rng(4242)
    refs, qrys = tiny_pairs(rng, 3000)
    a, b = pack_strings(refs), pack_strings(qrys)
    out.update({"tiny_ref_buf": a[0], "tiny_ref_off": a[1], "tiny_ref_len": a[2],
                "tiny_qry_buf": b[0], "tiny_qry_off": b[1], "tiny_qry_len": b[2]})
    for name, sc in TINY_SCORINGS.items():
        record(out, "tiny_" + name, a, b, sc)
    np.savez_compressed(os.path.join(HERE, "golden_cigar.npz"), **out)
    print("wrote golden_cigar.npz:", len(out), "arrays")


if __name__ == "__main__":
    main()
