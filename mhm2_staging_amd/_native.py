"""ctypes binding of libmhmsw.so (include/mhmsw.h, include/mhmsw_synth.h).

The library is built in-tree by ``make -C mhm2_staging_amd``.  There is no
fallback: every entry point raises if the library is missing.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmhmsw.so")

MHMSW_OK = 0
MHMSW_E_ARG = -1
NOALLOC_MAX_PAIRS = 131072       # MHMSW_NOALLOC_MAX_PAIRS
NOALLOC_MAX_BASES = 33554432     # MHMSW_NOALLOC_MAX_BASES
MODE_FAST = 0
MODE_PLAIN = 1
NUM_CLASSES = 17


def class_name(k):
    """Work class k of the forward pass (include/mhmsw.h)."""
    if k < 8:
        return f"fast16_P{k + 1}"
    if k < 16:
        return f"fast64_P{k - 7}"
    return "plain32"

RESULT_DTYPE = np.dtype([("score", "<i4"), ("read_begin", "<i4"), ("read_end", "<i4"),
                         ("ctg_begin", "<i4"), ("ctg_end", "<i4")])


class Scoring(ctypes.Structure):
    _fields_ = [("match", ctypes.c_int32), ("mismatch", ctypes.c_int32), ("gap_open", ctypes.c_int32),
                ("gap_ext", ctypes.c_int32), ("ambiguity", ctypes.c_int32)]


class Stats(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_int64), ("cells", ctypes.c_int64), ("reverse_cells", ctypes.c_int64),
                ("fast_pairs", ctypes.c_int64), ("plain_pairs", ctypes.c_int64), ("rerun_pairs", ctypes.c_int64),
                ("job_ms", ctypes.c_double), ("fwd_fast_ms", ctypes.c_double), ("fwd_plain_ms", ctypes.c_double),
                ("rev_ms", ctypes.c_double), ("prep_ms", ctypes.c_double),
                ("class_pairs", ctypes.c_int64 * 17), ("class_cells", ctypes.c_int64 * 17),
                ("class_ms", ctypes.c_double * 17), ("cigar_ms", ctypes.c_double), ("cigar_reruns", ctypes.c_int64),
                ("window_ms", ctypes.c_double), ("window_bases", ctypes.c_int64), ("packed_scoring", ctypes.c_int32),
                ("device_bytes", ctypes.c_int64), ("host_pack_ms", ctypes.c_double),
                ("host_upload_ms", ctypes.c_double), ("host_launch_ms", ctypes.c_double),
                ("class_lanes", ctypes.c_int32 * 17), ("class_regs", ctypes.c_int32 * 17),
                ("class_tag_bits", ctypes.c_int32 * 17)]

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if name.startswith("class_") else v
        return d


class Tuning(ctypes.Structure):
    _fields_ = [("blocks_per_cu", ctypes.c_int32), ("g64_min_rows", ctypes.c_int32), ("bin_shift", ctypes.c_int32)]

    def as_dict(self):
        return {"blocks_per_cu": self.blocks_per_cu, "g64_min_rows": self.g64_min_rows, "bin_shift": self.bin_shift}


class ReqTimes(ctypes.Structure):
    _fields_ = [("submit_s", ctypes.c_double), ("admit_s", ctypes.c_double), ("done_s", ctypes.c_double),
                ("device_ms", ctypes.c_double), ("launched_s", ctypes.c_double), ("finished_s", ctypes.c_double),
                ("job_est_bytes", ctypes.c_double)]


class ValuPeak(ctypes.Structure):
    _fields_ = [("lane_ops_per_s", ctypes.c_double), ("clock_mhz", ctypes.c_double),
                ("max_clock_mhz", ctypes.c_double), ("kernel_ms", ctypes.c_double), ("cus", ctypes.c_int32)]


class SynthContigs(ctypes.Structure):
    _fields_ = [("nctgs", ctypes.c_int64), ("seqs", ctypes.c_void_p), ("off", ctypes.c_void_p),
                ("len", ctypes.c_void_p), ("bytes", ctypes.c_int64)]


class SynthWindows(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_int64), ("reads", ctypes.c_void_p), ("read_off", ctypes.c_void_p),
                ("read_len", ctypes.c_void_p), ("ctg_id", ctypes.c_void_p), ("win_off", ctypes.c_void_p),
                ("win_len", ctypes.c_void_p), ("reads_bytes", ctypes.c_int64)]


class SynthBatch(ctypes.Structure):
    _fields_ = [("npairs", ctypes.c_int64), ("reads", ctypes.c_void_p), ("read_off", ctypes.c_void_p),
                ("read_len", ctypes.c_void_p), ("ctgs", ctypes.c_void_p), ("ctg_off", ctypes.c_void_p),
                ("ctg_len", ctypes.c_void_p), ("reads_bytes", ctypes.c_int64), ("ctgs_bytes", ctypes.c_int64)]


# every symbol include/mhmsw.h and include/mhmsw_synth.h declare
EXPORTS = [
    "mhmsw_device_count", "mhmsw_device_total_mem", "mhmsw_device_name", "mhmsw_create", "mhmsw_destroy",
    "mhmsw_last_error", "mhmsw_set_scoring", "mhmsw_set_orientation", "mhmsw_set_mode", "mhmsw_set_score2", "mhmsw_set_cigar", "mhmsw_set_graphs", "mhmsw_set_phase_timing",
    "mhmsw_set_tuning", "mhmsw_get_tuning", "mhmsw_autotune", "mhmsw_stage",
    "mhmsw_align_staged", "mhmsw_submit", "mhmsw_stage_seqs", "mhmsw_submit_seqs", "mhmsw_query", "mhmsw_wait", "mhmsw_results", "mhmsw_results_score2", "mhmsw_results_cigar", "mhmsw_get_stats",
    "mhmsw_synth_generate", "mhmsw_synth_free",
    "mhmsw_contigs_load", "mhmsw_contigs_free", "mhmsw_contigs_bytes", "mhmsw_stage_windows",
    "mhmsw_server_create", "mhmsw_server_submit", "mhmsw_server_status", "mhmsw_server_wait",
    "mhmsw_server_results", "mhmsw_server_clock", "mhmsw_server_last_error", "mhmsw_server_destroy",
    "mhmsw_server_set_outputs", "mhmsw_server_results_extra",
    "mhmsw_synth_contigs", "mhmsw_synth_contigs_free", "mhmsw_synth_window_pairs", "mhmsw_synth_window_pairs_free",
    "mhmsw_measure_valu_peak", "mhmsw_submit_noalloc",
]

_lib = None


def lib():
    """Load libmhmsw.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C mhm2_staging_amd` "
                           "(or __graft_entry__.build())")
    L = ctypes.CDLL(LIB_PATH)
    vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    sigs = {
        "mhmsw_device_count": [ctypes.POINTER(ctypes.c_int)],
        "mhmsw_device_total_mem": [ctypes.c_int, ctypes.POINTER(ctypes.c_size_t)],
        "mhmsw_device_name": [ctypes.c_int, ctypes.c_char_p, ctypes.c_size_t],
        "mhmsw_create": [ctypes.c_int, ctypes.POINTER(vp)],
        "mhmsw_destroy": [vp],
        "mhmsw_set_scoring": [vp, ctypes.POINTER(Scoring)],
        "mhmsw_set_orientation": [vp, ctypes.c_int],
        "mhmsw_set_mode": [vp, ctypes.c_int],
        "mhmsw_set_score2": [vp, ctypes.c_int],
        "mhmsw_set_cigar": [vp, ctypes.c_int],
        "mhmsw_set_graphs": [vp, ctypes.c_int],
        "mhmsw_set_phase_timing": [vp, ctypes.c_int],
        "mhmsw_set_tuning": [vp, ctypes.POINTER(Tuning)],
        "mhmsw_get_tuning": [vp, ctypes.POINTER(Tuning)],
        "mhmsw_autotune": [vp, ctypes.c_int, ctypes.POINTER(Tuning)],
        "mhmsw_stage": [vp, i64, vp, vp, vp, vp, vp, vp],
        "mhmsw_align_staged": [vp],
        "mhmsw_submit": [vp, i64, vp, vp, vp, vp, vp, vp],
        "mhmsw_stage_seqs": [vp, i64, vp, vp, vp, vp],
        "mhmsw_submit_seqs": [vp, i64, vp, vp, vp, vp],
        "mhmsw_query": [vp],
        "mhmsw_wait": [vp],
        "mhmsw_results": [vp, vp, i64],
        "mhmsw_results_score2": [vp, vp, i64],
        "mhmsw_results_cigar": [vp, i64, vp, vp, vp, i64, ctypes.POINTER(i64)],
        "mhmsw_get_stats": [vp, ctypes.POINTER(Stats)],
        "mhmsw_synth_generate": [ctypes.c_int, i64, ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(SynthBatch)],
        "mhmsw_contigs_load": [ctypes.c_int, i64, vp, vp, vp, ctypes.POINTER(vp)],
        "mhmsw_contigs_free": [vp],
        "mhmsw_stage_windows": [vp, vp, i64, vp, vp, vp, vp, vp, vp],
        "mhmsw_server_create": [ctypes.c_int, ctypes.POINTER(Scoring), vp, ctypes.c_int, i64, ctypes.POINTER(vp)],
        "mhmsw_server_submit": [vp, i64, vp, vp, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_uint64)],
        "mhmsw_server_status": [vp, ctypes.c_uint64],
        "mhmsw_server_wait": [vp, ctypes.c_uint64],
        "mhmsw_server_results": [vp, ctypes.c_uint64, vp, i64, ctypes.POINTER(ReqTimes)],
        "mhmsw_server_destroy": [vp],
        "mhmsw_server_set_outputs": [vp, ctypes.c_int, ctypes.c_int],
        "mhmsw_server_results_extra": [vp, ctypes.c_uint64, i64, vp, vp, vp, vp, i64, ctypes.POINTER(i64)],
        "mhmsw_synth_contigs": [ctypes.c_uint64, i64, i32, i32, ctypes.c_int, ctypes.POINTER(SynthContigs)],
        "mhmsw_synth_window_pairs": [ctypes.POINTER(SynthContigs), ctypes.c_uint64, i64, ctypes.c_int,
                                     ctypes.POINTER(SynthWindows)],
        "mhmsw_measure_valu_peak": [ctypes.c_int, ctypes.POINTER(ValuPeak)],
        "mhmsw_submit_noalloc": [vp, i64, vp, vp, vp, vp, vp, vp, vp, vp, vp],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.mhmsw_last_error.argtypes = [vp]
    L.mhmsw_last_error.restype = ctypes.c_char_p
    L.mhmsw_synth_free.argtypes = [ctypes.POINTER(SynthBatch)]
    L.mhmsw_synth_free.restype = None
    L.mhmsw_synth_contigs_free.argtypes = [ctypes.POINTER(SynthContigs)]
    L.mhmsw_synth_contigs_free.restype = None
    L.mhmsw_synth_window_pairs_free.argtypes = [ctypes.POINTER(SynthWindows)]
    L.mhmsw_synth_window_pairs_free.restype = None
    L.mhmsw_contigs_bytes.argtypes = [vp]
    L.mhmsw_contigs_bytes.restype = ctypes.c_int64
    L.mhmsw_server_clock.argtypes = []
    L.mhmsw_server_clock.restype = ctypes.c_double
    L.mhmsw_server_last_error.argtypes = [vp]
    L.mhmsw_server_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


class MhmswError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mhmsw error {code}: {msg}")
        self.code = code


def measure_valu_peak(device=0):
    """Sustained packed-int16 VALU rate of the device (mhmsw_measure_valu_peak) as a dict."""
    p = ValuPeak()
    rc = lib().mhmsw_measure_valu_peak(int(device), ctypes.byref(p))
    if rc != 0:
        raise MhmswError(rc, "mhmsw_measure_valu_peak failed")
    return {"lane_ops_per_s": p.lane_ops_per_s, "clock_mhz": p.clock_mhz, "max_clock_mhz": p.max_clock_mhz,
            "kernel_ms": p.kernel_ms, "cus": p.cus}
