"""Local assembly (contig end extension) on the GPU: ctypes binding of the
mhmla C ABI in libmhmsw.so (include/mhmla.h).

``LocalAssembler.extend(batch)`` is what ``locassm_driver::local_assem_driver``
(src/gpu_loc_assem/driver.hpp:60) computes for a slice of contigs: per
contig, the left extension (to prepend) and the right extension (to append).
There is no CPU fallback: every call raises when the library or the GPU is
missing.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native
from .locassm_data import LocAssmBatch, from_contigs, generate  # noqa: F401

MHMLA_EXPORTS = [
    "mhmla_default_params", "mhmla_device_count", "mhmla_create", "mhmla_destroy", "mhmla_last_error",
    "mhmla_stage", "mhmla_run", "mhmla_wait", "mhmla_results", "mhmla_extend", "mhmla_get_stats",
]


class Params(ctypes.Structure):
    _fields_ = [("kmer_len", ctypes.c_int32), ("min_kmer_len", ctypes.c_int32), ("max_kmer_len", ctypes.c_int32),
                ("shift", ctypes.c_int32), ("walk_len_limit", ctypes.c_int32), ("qual_offset", ctypes.c_int32),
                ("min_qual", ctypes.c_int32), ("min_hi_qual", ctypes.c_int32), ("rating_thres", ctypes.c_int32),
                ("max_reads", ctypes.c_int32), ("min_viable_depth", ctypes.c_double),
                ("min_expected_depth", ctypes.c_double)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Stats(ctypes.Structure):
    _fields_ = [("job_ms", ctypes.c_double), ("contigs", ctypes.c_int64), ("sides", ctypes.c_int64),
                ("launches", ctypes.c_int64), ("mers", ctypes.c_int64), ("walk_steps", ctypes.c_int64),
                ("walks", ctypes.c_int64), ("scratch_bytes", ctypes.c_int64), ("phase_cycles", ctypes.c_int64 * 4),
                ("table_resizes", ctypes.c_int64)]

    def as_dict(self):
        d = {n: getattr(self, n) for n, _ in self._fields_}
        d["phase_cycles"] = list(self.phase_cycles)
        return d


class LocAssmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mhmla error {code}: {msg}")
        self.code = code


_bound = False


def lib():
    """libmhmsw.so with the mhmla signatures declared."""
    global _bound
    L = _native.lib()
    if not _bound:
        vp, i64 = ctypes.c_void_p, ctypes.c_int64
        P = ctypes.POINTER(Params)
        sigs = {
            "mhmla_device_count": [],
            "mhmla_create": [ctypes.c_int, ctypes.POINTER(vp)],
            "mhmla_stage": [vp, i64, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp],
            "mhmla_run": [vp, P],
            "mhmla_wait": [vp],
            "mhmla_results": [vp, vp, vp, vp, vp, vp],
            "mhmla_extend": [vp, P, i64, vp, vp, vp, vp, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp],
            "mhmla_get_stats": [vp, ctypes.POINTER(Stats)],
        }
        for name, args in sigs.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.mhmla_default_params.argtypes = [P]
        L.mhmla_default_params.restype = None
        L.mhmla_destroy.argtypes = [vp]
        L.mhmla_destroy.restype = None
        L.mhmla_last_error.argtypes = [vp]
        L.mhmla_last_error.restype = ctypes.c_char_p
        _bound = True
    return L


def params(**kw):
    """Parameters with the reference's defaults (include/mhmla.h), overridden by kw."""
    p = Params()
    lib().mhmla_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown local assembly parameter {k}")
        setattr(p, k, v)
    return p


@dataclass
class Extensions:
    left_len: np.ndarray   # int32 [nctgs]
    left: np.ndarray       # uint8 [nctgs, walk_len_limit], reverse-complemented: prepend
    right_len: np.ndarray  # int32 [nctgs]
    right: np.ndarray      # uint8 [nctgs, walk_len_limit]: append
    terms: np.ndarray      # int32 [nctgs, 2, 3]: walks ending in X, F, R per side (left, right)

    def left_str(self, c):
        return bytes(self.left[c, :self.left_len[c]]).decode()

    def right_str(self, c):
        return bytes(self.right[c, :self.right_len[c]]).decode()

    def extended(self, batch, c):
        """Contig c with both extensions attached, as driver.cpp:290-306 stitches it."""
        return self.left_str(c) + batch.ctg(c) + self.right_str(c)


class LocalAssembler:
    """One device context for local assembly walks."""

    def __init__(self, device=0):
        L = lib()
        ctx = ctypes.c_void_p()
        rc = L.mhmla_create(int(device), ctypes.byref(ctx))
        if rc != 0:
            raise LocAssmError(rc, "no usable GPU" if rc == -4 else "context creation failed")
        self._ctx = ctx
        self._L = L
        self._batch = None
        self._params = None

    def close(self):
        if getattr(self, "_ctx", None):
            self._L.mhmla_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != 0:
            raise LocAssmError(rc, self._L.mhmla_last_error(self._ctx).decode())

    def stage(self, b: LocAssmBatch):
        self._check(self._L.mhmla_stage(self._ctx, b.nctgs, b.ctg_seq.ctypes.data, b.ctg_off.ctypes.data,
                                        b.ctg_len.ctypes.data, b.depth.ctypes.data, b.left_first.ctypes.data,
                                        b.right_first.ctypes.data, b.read_idx.ctypes.data, b.nreads,
                                        b.read_seq.ctypes.data, b.read_qual.ctypes.data, b.read_off.ctypes.data,
                                        b.read_len.ctypes.data))
        self._batch = b

    def run(self, p=None):
        p = p if p is not None else params()
        self._check(self._L.mhmla_run(self._ctx, ctypes.byref(p)))
        self._params = p

    def wait(self):
        self._check(self._L.mhmla_wait(self._ctx))

    def results(self):
        n = self._batch.nctgs
        W = max(int(self._params.walk_len_limit), 1)
        ll = np.zeros(n, np.int32)
        rl = np.zeros(n, np.int32)
        le = np.zeros((n, W), np.uint8)
        re = np.zeros((n, W), np.uint8)
        tm = np.zeros((n, 2, 3), np.int32)
        self._check(self._L.mhmla_results(self._ctx, ll.ctypes.data, le.ctypes.data, rl.ctypes.data, re.ctypes.data,
                                          tm.ctypes.data))
        return Extensions(ll, le, rl, re, tm)

    def extend(self, b: LocAssmBatch, p=None):
        self.stage(b)
        self.run(p)
        self.wait()
        return self.results()

    def stats(self):
        s = Stats()
        self._check(self._L.mhmla_get_stats(self._ctx, ctypes.byref(s)))
        return s.as_dict()
