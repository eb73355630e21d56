"""mhmsw_submit_noalloc (the entry point that allocates nothing) against mhmsw_submit, through the C ABI.

Inputs and outputs are device buffers the test allocates; the entry point's own working memory is its fixed
scratch array, shared by consecutive calls on the GPU, so each test makes several calls in a row (the second
smaller than the first) and checks every one bit-exact against the allocating entry point.
"""
import ctypes

import numpy as np
import pytest

import mhm2_staging_amd as m
from mhm2_staging_amd import _native
from hip_mem import DeviceArray, synchronize


def noalloc(al, b):
    """One call on the context's stream; (results, status[2])."""
    L = _native.lib()
    n = b.npairs
    bufs = [DeviceArray(x) for x in (b.reads, b.read_off, b.read_len, b.ctgs, b.ctg_off, b.ctg_len)]
    out = DeviceArray(nbytes=max(n, 1) * _native.RESULT_DTYPE.itemsize)
    status = DeviceArray(np.full(2, 7, np.int32))
    try:
        print(f"noalloc call: {n} pairs", flush=True)
        rc = L.mhmsw_submit_noalloc(al._ctx, n, *[d.ptr for d in bufs], out.ptr, status.ptr, None)
        if rc != 0:
            raise _native.MhmswError(rc, L.mhmsw_last_error(al._ctx).decode())
        synchronize()
        return out.to_host(_native.RESULT_DTYPE, n), status.to_host(np.int32, 2)
    finally:
        for d in bufs + [out, status]:
            d.free()


@pytest.mark.gpu
@pytest.mark.parametrize("scoring", [m.MHM2_SCORING, m.MHM2_TREE_SCORING])
def test_noalloc_matches_submit_consecutive_calls(scoring):
    big = m.generate(2, npairs=20000)           # uneven lengths, N bases, unrelated pairs
    small = m.generate(1, npairs=3000, seed=7)  # smaller batch of other lengths
    tiny = m.from_strings(["ACGTNACGT", "", "A", "GATTACA" * 40, "acgu" * 9],
                          ["ACGTACGT", "ACGT", "", "GATTACA" * 50, "ACGT" * 9])
    with m.Aligner(device=0, scoring=scoring) as al:
        for b in (big, small, tiny, big):
            want = al.align(b)
            got, st = noalloc(al, b)
            assert list(st) == [0, -1]
            assert got.shape == want.shape
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, (b.npairs, bad[:5], got[bad[:5]], want[bad[:5]])


@pytest.mark.gpu
def test_noalloc_orientation_and_plain_mode():
    b = m.generate(3, npairs=5000)
    with m.Aligner(device=0, read_is_ssw_ref=False) as al:
        want = al.align(b)
        got, st = noalloc(al, b)
        assert list(st) == [0, -1] and (got == want).all()
        al.set_mode(m.MODE_PLAIN)
        got2, st2 = noalloc(al, m.generate(1, npairs=500))
        want2 = al.align(m.generate(1, npairs=500))
        assert list(st2) == [0, -1] and (got2 == want2).all()


@pytest.mark.gpu
def test_noalloc_rejects_oversized_and_invalid_batches():
    with m.Aligner(device=0) as al:
        L = _native.lib()
        # more pairs than the scratch holds: refused on the host, nothing enqueued
        rc = L.mhmsw_submit_noalloc(al._ctx, _native.NOALLOC_MAX_PAIRS + 1, *([ctypes.c_void_p(8)] * 9), None)
        assert rc == _native.MHMSW_E_ARG
        assert b"at most" in L.mhmsw_last_error(al._ctx)
        # a pair whose two sequences both exceed 8192 bases: status names it, every result is the sentinel
        long = "ACGT" * 2100
        b = m.from_strings(["ACGT" * 30, long, "ACGT"], ["ACGT" * 40, long, "ACG"])
        got, st = noalloc(al, b)
        assert list(st) == [_native.MHMSW_E_ARG, 1]
        assert (got["score"] == 0).all() and (got["read_end"] == -1).all()
        # more padded bases than the scratch holds: status without a pair index
        n = 40
        reads = np.frombuffer(b"A" * 1_000_000 + b"\0", np.uint8).copy()
        ctgs = np.frombuffer(b"C" * 100 + b"\0", np.uint8).copy()
        over = m.PairBatch(reads, np.zeros(n, np.int64), np.full(n, 1_000_000, np.int32), ctgs,
                           np.zeros(n, np.int64), np.full(n, 100, np.int32))
        got, st = noalloc(al, over)
        assert list(st) == [_native.MHMSW_E_ARG, -1] and (got["score"] == 0).all()
        # the scratch is usable again by the next call
        small = m.generate(1, npairs=2000, seed=11)
        want = al.align(small)
        got, st = noalloc(al, small)
        assert list(st) == [0, -1] and (got == want).all()
