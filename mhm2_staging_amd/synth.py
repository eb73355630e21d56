"""Seeded synthetic read / contig-window batches (wraps mhmsw_synth_generate)."""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native


@dataclass
class PairBatch:
    """Pairs as concatenated ASCII buffers, the layout every batch entry point takes."""
    reads: np.ndarray      # uint8, NUL-terminated
    read_off: np.ndarray   # int64
    read_len: np.ndarray   # int32
    ctgs: np.ndarray
    ctg_off: np.ndarray
    ctg_len: np.ndarray

    @property
    def npairs(self):
        return int(self.read_len.shape[0])

    def cells(self):
        return int(np.dot(self.read_len.astype(np.int64), self.ctg_len.astype(np.int64)))

    def read(self, p):
        return bytes(self.reads[self.read_off[p]:self.read_off[p] + self.read_len[p]]).decode()

    def ctg(self, p):
        return bytes(self.ctgs[self.ctg_off[p]:self.ctg_off[p] + self.ctg_len[p]]).decode()

    def subset(self, idx):
        idx = np.asarray(idx)
        return from_strings([self.read(int(p)) for p in idx], [self.ctg(int(p)) for p in idx])


def _pack(seqs):
    seqs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.int32)
    offs = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    buf = np.frombuffer(b"".join(seqs) + b"\0", dtype=np.uint8).copy()
    return buf, offs, lens


def from_strings(reads, ctgs):
    if len(reads) != len(ctgs):
        raise ValueError("reads and contig windows differ in count")
    r = _pack(reads)
    c = _pack(ctgs)
    return PairBatch(r[0], r[1], r[2], c[0], c[1], c[2])


def generate(config, npairs=0, seed=0, threads=0):
    """Workload of BASELINE config 1..5 (plain, uneven, skewed, repeats, serving pairs)."""
    L = _native.lib()
    b = _native.SynthBatch()
    rc = L.mhmsw_synth_generate(int(config), int(npairs), int(seed), int(threads), ctypes.byref(b))
    if rc != 0:
        raise _native.MhmswError(rc, f"synthetic generation failed for config {config}")
    try:
        n = b.npairs

        def arr(ptr, ctype, count):
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (count,)).copy()

        return PairBatch(arr(b.reads, ctypes.c_uint8, b.reads_bytes + 1), arr(b.read_off, ctypes.c_int64, n),
                         arr(b.read_len, ctypes.c_int32, n), arr(b.ctgs, ctypes.c_uint8, b.ctgs_bytes + 1),
                         arr(b.ctg_off, ctypes.c_int64, n), arr(b.ctg_len, ctypes.c_int32, n))
    finally:
        L.mhmsw_synth_free(ctypes.byref(b))
