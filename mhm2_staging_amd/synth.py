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
    """Seeded workload 1..5 (plain, uneven, skewed, repeats, serving pairs)."""
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


@dataclass
class ContigSet:
    """Contigs as one concatenated ASCII buffer."""
    seqs: np.ndarray     # uint8, NUL-terminated
    off: np.ndarray      # int64
    len: np.ndarray      # int32

    @property
    def nctgs(self):
        return int(self.len.shape[0])

    def contig(self, c):
        return bytes(self.seqs[self.off[c]:self.off[c] + self.len[c]]).decode()


@dataclass
class WindowBatch:
    """Pairs whose contig windows are (contig id, offset, length) ranges of a ContigSet."""
    reads: np.ndarray
    read_off: np.ndarray
    read_len: np.ndarray
    ctg_id: np.ndarray   # int32
    win_off: np.ndarray  # int32
    win_len: np.ndarray  # int32

    @property
    def npairs(self):
        return int(self.read_len.shape[0])

    def cells(self):
        return int(np.dot(self.read_len.astype(np.int64), self.win_len.astype(np.int64)))

    def as_strings(self, contigs):
        """The same pairs with the windows copied out as strings."""
        reads = [bytes(self.reads[self.read_off[p]:self.read_off[p] + self.read_len[p]]).decode()
                 for p in range(self.npairs)]
        wins = []
        for p in range(self.npairs):
            c = int(self.ctg_id[p])
            o = int(contigs.off[c]) + int(self.win_off[p])
            wins.append(bytes(contigs.seqs[o:o + int(self.win_len[p])]).decode())
        return from_strings(reads, wins)

    def slice(self, a, b):
        """Pairs a..b-1 (reads re-based)."""
        ro = self.read_off[a:b]
        rl = self.read_len[a:b]
        buf = np.concatenate([self.reads[o:o + n] for o, n in zip(ro, rl)] + [np.zeros(1, np.uint8)])
        off = np.zeros(b - a, np.int64)
        if b - a > 1:
            off[1:] = np.cumsum(rl[:-1], dtype=np.int64)
        return WindowBatch(buf, off, rl.copy(), self.ctg_id[a:b].copy(), self.win_off[a:b].copy(),
                           self.win_len[a:b].copy())


def contigs(nctgs=64000, seed=5, min_len=500, max_len=20000, threads=0):
    """Serving contig set: i.i.d. ACGT, lengths log-uniform in [min_len, max_len]."""
    L = _native.lib()
    c = _native.SynthContigs()
    rc = L.mhmsw_synth_contigs(int(seed), int(nctgs), int(min_len), int(max_len), int(threads), ctypes.byref(c))
    if rc != 0:
        raise _native.MhmswError(rc, "contig generation failed")
    try:
        def arr(ptr, ctype, count):
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (count,)).copy()
        return ContigSet(arr(c.seqs, ctypes.c_uint8, c.bytes + 1), arr(c.off, ctypes.c_int64, c.nctgs),
                         arr(c.len, ctypes.c_int32, c.nctgs))
    finally:
        L.mhmsw_synth_contigs_free(ctypes.byref(c))


def window_pairs(cs: ContigSet, npairs, seed, threads=0):
    """Serving pairs: 150 bp reads at 1% error against U[300,1024] windows of cs."""
    L = _native.lib()
    c = _native.SynthContigs(cs.nctgs, cs.seqs.ctypes.data, cs.off.ctypes.data, cs.len.ctypes.data,
                             int(cs.off[-1] + cs.len[-1]) if cs.nctgs else 0)
    w = _native.SynthWindows()
    rc = L.mhmsw_synth_window_pairs(ctypes.byref(c), int(seed), int(npairs), int(threads), ctypes.byref(w))
    if rc != 0:
        raise _native.MhmswError(rc, "window pair generation failed")
    try:
        n = w.npairs

        def arr(ptr, ctype, count):
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctype)), (count,)).copy()
        return WindowBatch(arr(w.reads, ctypes.c_uint8, w.reads_bytes + 1), arr(w.read_off, ctypes.c_int64, n),
                           arr(w.read_len, ctypes.c_int32, n), arr(w.ctg_id, ctypes.c_int32, n),
                           arr(w.win_off, ctypes.c_int32, n), arr(w.win_len, ctypes.c_int32, n))
    finally:
        L.mhmsw_synth_window_pairs_free(ctypes.byref(w))
