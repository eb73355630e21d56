"""mhmsw_submit_seqs (sequences by pointer, the drop-in driver's path) against
mhmsw_stage + align_staged on flat buffers and against the oracle."""
import ctypes

import numpy as np
import pytest

import mhm2_staging_amd as m
from mhm2_staging_amd import _native
from oracle_lib import oracle_batch

pytestmark = pytest.mark.gpu

FIELDS = ("score", "read_begin", "read_end", "ctg_begin", "ctg_end")


def as_lists(b):
    return [b.read(p).encode() for p in range(b.npairs)], [b.ctg(p).encode() for p in range(b.npairs)]


def table(res):
    return np.stack([res[k] for k in FIELDS], 1)


def oracle_of(reads, ctgs):
    buf_r = b"".join(reads) + b"\0"
    buf_c = b"".join(ctgs) + b"\0"
    rl = np.array([len(x) for x in reads], np.int32)
    cl = np.array([len(x) for x in ctgs], np.int32)
    ro = np.concatenate([[0], np.cumsum(rl)[:-1]]).astype(np.int64)
    co = np.concatenate([[0], np.cumsum(cl)[:-1]]).astype(np.int64)
    o = oracle_batch(np.frombuffer(buf_r, np.uint8), ro, rl, np.frombuffer(buf_c, np.uint8), co, cl,
                     m.MHM2_SCORING)
    return np.stack([o[k] for k in ("score", "ref_begin", "ref_end", "query_begin", "query_end")], 1)


@pytest.mark.parametrize("config,npairs", [(2, 6000), (3, 20000), (4, 1500)])
def test_submit_seqs_matches_flat_batches(gpu_aligner, config, npairs):
    print(f"[gpu] submit_seqs config={config} npairs={npairs}", flush=True)
    b = m.generate(config, npairs=npairs)
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    want = table(al.align(b))
    reads, ctgs = as_lists(b)
    got = table(al.align_seqs(reads, ctgs))
    assert np.array_equal(got, want)
    # and again on the flat form after it (the job's pinned result copy is per call)
    assert np.array_equal(table(al.align(b)), want)
    idx = np.linspace(0, npairs - 1, 400).astype(int)
    o = oracle_of([reads[i] for i in idx], [ctgs[i] for i in idx])
    assert np.array_equal(got[idx], o)


def test_submit_seqs_edge_cases_vs_oracle(gpu_aligner):
    print("[gpu] submit_seqs edge cases", flush=True)
    rng = np.random.default_rng(7)

    def rs(n, alpha=b"ACGT"):
        return bytes(rng.choice(list(alpha), n))

    reads, ctgs = [], []
    for _ in range(300):
        s = rs(int(rng.integers(1, 260)))
        w = rs(int(rng.integers(1, 700)), b"ACGTNacgtnU")
        reads.append(s)
        ctgs.append(w)
    reads += [b"A", b"ACGT", b"N" * 40, b"acgtacgtacgt", rs(150)]
    ctgs += [b"A", b"TTTT", b"N" * 40, b"ACGTACGTACGT", rs(150)]
    al = gpu_aligner
    al.set_scoring(m.MHM2_SCORING)
    al.set_orientation(True)
    al.set_mode(m.MODE_FAST)
    got = table(al.align_seqs(reads, ctgs))
    assert np.array_equal(got, oracle_of(reads, ctgs))
    # empty sequences, given as NULL or as b"": as the same pairs in flat buffers
    er, ec = [None, b"", b"ACGT", b"ACGTT"], [b"ACGT", b"ACGT", None, b"ACGTT"]
    got = table(al.align_seqs(er, ec))
    want = table(al.align(m.from_strings([x or b"" for x in er], [x or b"" for x in ec])))
    assert np.array_equal(got, want)
    assert (got[:3, 0] == 0).all()
    assert got[3].tolist() == oracle_of([b"ACGTT"], [b"ACGTT"])[0].tolist()


def test_submit_seqs_rejects_null_pointer_with_length(gpu_aligner):
    print("[gpu] submit_seqs null pointer", flush=True)
    L = _native.lib()
    rp = (ctypes.c_char_p * 2)(b"ACGT", None)
    cp = (ctypes.c_char_p * 2)(b"ACGT", b"ACGT")
    rl = np.array([4, 4], np.int32)
    cl = np.array([4, 4], np.int32)
    rc = L.mhmsw_submit_seqs(gpu_aligner._ctx, 2, ctypes.cast(rp, ctypes.c_void_p), rl.ctypes.data,
                             ctypes.cast(cp, ctypes.c_void_p), cl.ctypes.data)
    assert rc == _native.MHMSW_E_ARG
    assert b"null sequence pointer" in L.mhmsw_last_error(gpu_aligner._ctx)
