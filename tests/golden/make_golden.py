"""Write golden vectors from the reference SSW (oracle/_ref/libssw_ref.so).

Run in a container that has /root/reference (``make -C oracle`` builds the
reference from its own sources first) and libmhmsw.so (for the config-1
generator).  Output: tests/golden/golden_small.npz (seeded small corpora at
four scorings, both roles), tests/golden/golden_kat.json and
tests/golden/golden_config1.npz (10,000 pairs of the plain 150 x 350 bp
workload under mhm2's default scoring in klign's orientation, with a sha256
of the result table).
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle_lib import have_ref, pack_strings, ref_batch  # noqa: E402
from golden.kat_cases import MHM2_DEFAULT_SCORING, KAT_CASES, KAT_SCORINGS  # noqa: E402

SCORINGS = dict(KAT_SCORINGS)
SCORINGS["baseline"] = MHM2_DEFAULT_SCORING


def result_hash(res):
    arr = np.stack([res[k].astype("<i4") for k in ("score", "ref_begin", "ref_end", "query_begin", "query_end")], 1)
    return hashlib.sha256(np.ascontiguousarray(arr).tobytes()).hexdigest()


def small_corpus(rng, n):
    """Short pairs rich in ties, Ns, score 0 and near-repeats."""
    refs, qrys = [], []
    alphabets = ["ACGT", "ACGTN", "AC", "AAAAC", "ACGTNNNN", "acgtu"]
    for i in range(n):
        alph = alphabets[i % len(alphabets)]
        m = int(rng.integers(1, 70))
        k = int(rng.integers(1, 70))
        r = "".join(rng.choice(list(alph), m))
        kind = i % 4
        if kind == 0:
            q = "".join(rng.choice(list(alph), k))
        elif kind == 1:  # query contains a mutated piece of ref
            a = int(rng.integers(0, m))
            piece = list(r[a:a + int(rng.integers(1, m + 1))])
            for t in range(len(piece)):
                if rng.random() < 0.1:
                    piece[t] = rng.choice(list("ACGTN"))
            q = "".join(rng.choice(list(alph), int(rng.integers(0, 10)))) + "".join(piece) + \
                "".join(rng.choice(list(alph), int(rng.integers(0, 10))))
        elif kind == 2:  # tandem repeats: many equal-score ends
            unit = "".join(rng.choice(list("ACGT"), int(rng.integers(1, 5))))
            r = (unit * 40)[:m]
            q = (unit * 40)[int(rng.integers(0, 4)):][:k]
        else:  # indels
            q = list(r)
            for _ in range(int(rng.integers(1, 4))):
                if q and rng.random() < 0.5:
                    del q[int(rng.integers(0, len(q)))]
                else:
                    q.insert(int(rng.integers(0, len(q) + 1)), rng.choice(list("ACGT")))
            q = "".join(q) or "A"
        refs.append(r)
        qrys.append(q)
    return refs, qrys


def long_corpus(rng, n):
    """Longer pairs crossing SSW's 8-bit -> 16-bit switch (score >= 252)."""
    refs, qrys = [], []
    for i in range(n):
        L = int(rng.integers(200, 700))
        q = "".join(rng.choice(list("ACGT"), L + 300))
        a = int(rng.integers(0, 300))
        r = list(q[a:a + L])
        for t in range(len(r)):
            if rng.random() < 0.01:
                r[t] = rng.choice(list("ACGTN"))
        refs.append("".join(r))
        qrys.append(q)
    return refs, qrys


def main():
    if not have_ref():
        raise SystemExit("oracle/_ref/libssw_ref.so is not built (make -C oracle with the reference present)")
    out = {}
    rng = np.random.default_rng(20240601)
    refs, qrys = small_corpus(rng, 1500)
    lrefs, lqrys = long_corpus(rng, 60)
    refs += lrefs
    qrys += lqrys
    a = pack_strings(refs)
    b = pack_strings(qrys)
    out["ref_buf"], out["ref_off"], out["ref_len"] = a
    out["qry_buf"], out["qry_off"], out["qry_len"] = b
    for name, sc in SCORINGS.items():
        res = ref_batch(*a, *b, sc, 8)
        out[f"res_{name}"] = np.stack([res[k] for k in res.dtype.names], 1).astype(np.int32)
        out[f"scoring_{name}"] = np.array(sc, dtype=np.int32)
    np.savez_compressed(os.path.join(HERE, "golden_small.npz"), **out)

    kat = {"cases": [list(c) for c in KAT_CASES], "scorings": {k: list(v) for k, v in SCORINGS.items()}, "results": {}}
    for name, sc in SCORINGS.items():
        res = ref_batch(*pack_strings([c[1] for c in KAT_CASES]), *pack_strings([c[0] for c in KAT_CASES]), sc, 1)
        kat["results"][name] = [[int(x) for x in r] for r in res.tolist()]
        # swapped roles (SSW-ref = the test's query)
        res2 = ref_batch(*pack_strings([c[0] for c in KAT_CASES]), *pack_strings([c[1] for c in KAT_CASES]), sc, 1)
        kat["results"][name + "_swapped"] = [[int(x) for x in r] for r in res2.tolist()]
    with open(os.path.join(HERE, "golden_kat.json"), "w") as f:
        json.dump(kat, f, indent=1)

    import mhm2_staging_amd as m
    batch = m.generate(1)
    # klign orientation: SSW-ref = read, SSW-query = contig window (klign.cpp:257)
    res = ref_batch(batch.reads, batch.read_off, batch.read_len, batch.ctgs, batch.ctg_off, batch.ctg_len,
                    MHM2_DEFAULT_SCORING, 8)
    table = np.stack([res[k] for k in res.dtype.names], 1).astype(np.int32)
    np.savez_compressed(os.path.join(HERE, "golden_config1.npz"), results=table,
                        sha256=np.array(result_hash(res)), scoring=np.array(MHM2_DEFAULT_SCORING, dtype=np.int32))
    print("wrote golden_small.npz, golden_kat.json, golden_config1.npz; config1 sha256", result_hash(res))


if __name__ == "__main__":
    main()
