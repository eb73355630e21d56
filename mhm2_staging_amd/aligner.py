"""Host-side mirror of mhm2's GPU Smith-Waterman interface.

``Aligner`` is a thin owner of one mhmsw context (include/mhmsw.h).
``GPUDriver`` mirrors adept_sw::GPUDriver (src/adept-sw/driver.hpp:39-55):
same method names, argument meaning (penalties passed negative to ``init``)
and call protocol; results come back as int16 arrays named as in
AlignmentResults, where "ref" is the contig window and "query" the read.
"""
import ctypes
import time

import numpy as np

from . import _native
from .synth import ContigSet, PairBatch, WindowBatch, from_strings

# mhm2's alignment scoring, its "previous defaults" (CMakeDefinitions.txt:118-122)
MHM2_SCORING = (1, 3, 5, 2, 2)
# the values active in the reference tree (CMakeDefinitions.txt:113-117)
MHM2_TREE_SCORING = (1, 1, 1, 1, 1)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Aligner:
    """One device context: scoring, orientation, mode, and one batch at a time."""

    def __init__(self, device=0, scoring=MHM2_SCORING, read_is_ssw_ref=True, mode=_native.MODE_FAST):
        self._L = _native.lib()
        self._ctx = ctypes.c_void_p()
        rc = self._L.mhmsw_create(int(device), ctypes.byref(self._ctx))
        if rc != 0:
            raise _native.MhmswError(rc, f"mhmsw_create(device={device}) failed")
        self._batch = None
        self.set_scoring(scoring)
        self.set_orientation(read_is_ssw_ref)
        self.set_mode(mode)

    def _check(self, rc, what):
        if rc < 0:
            msg = self._L.mhmsw_last_error(self._ctx)
            raise _native.MhmswError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if self._ctx:
            self._L.mhmsw_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def set_scoring(self, scoring):
        sc = _native.Scoring(*[int(x) for x in scoring])
        self._check(self._L.mhmsw_set_scoring(self._ctx, ctypes.byref(sc)), "set_scoring")
        self.scoring = tuple(int(x) for x in scoring)

    def set_orientation(self, read_is_ssw_ref):
        self._check(self._L.mhmsw_set_orientation(self._ctx, 1 if read_is_ssw_ref else 0), "set_orientation")

    def set_mode(self, mode):
        self._check(self._L.mhmsw_set_mode(self._ctx, int(mode)), "set_mode")

    def set_graphs(self, enable):
        self._check(self._L.mhmsw_set_graphs(self._ctx, 1 if enable else 0), "set_graphs")

    def set_score2(self, enable):
        """Also compute SSW's second-best score (sw_score_next_best) per pair."""
        self._check(self._L.mhmsw_set_score2(self._ctx, 1 if enable else 0), "set_score2")

    def set_cigar(self, enable):
        """Also compute SSW's CIGAR and mismatch count per pair (klign with report_cigar)."""
        self._check(self._L.mhmsw_set_cigar(self._ctx, 1 if enable else 0), "set_cigar")

    def set_tuning(self, blocks_per_cu=8, g64_min_rows=257, bin_shift=5):
        t = _native.Tuning(int(blocks_per_cu), int(g64_min_rows), int(bin_shift))
        self._check(self._L.mhmsw_set_tuning(self._ctx, ctypes.byref(t)), "set_tuning")

    def get_tuning(self):
        t = _native.Tuning()
        self._check(self._L.mhmsw_get_tuning(self._ctx, ctypes.byref(t)), "get_tuning")
        return t.as_dict()

    def autotune(self, repeats=2):
        """Pick the fastest scheduling parameters for the staged batch (kept on the context)."""
        t = _native.Tuning()
        self._check(self._L.mhmsw_autotune(self._ctx, int(repeats), ctypes.byref(t)), "autotune")
        return t.as_dict()

    def stage(self, batch: PairBatch):
        self._batch = batch  # keep the host buffers alive while referenced
        self._check(self._L.mhmsw_stage(self._ctx, batch.npairs, _ptr(batch.reads), _ptr(batch.read_off),
                                        _ptr(batch.read_len), _ptr(batch.ctgs), _ptr(batch.ctg_off),
                                        _ptr(batch.ctg_len)), "stage")

    def stage_windows(self, contigs, batch: WindowBatch):
        """Stage pairs whose windows are ranges of device-resident contigs."""
        self._batch = batch
        self._check(self._L.mhmsw_stage_windows(self._ctx, contigs.handle, batch.npairs, _ptr(batch.reads),
                                                _ptr(batch.read_off), _ptr(batch.read_len), _ptr(batch.ctg_id),
                                                _ptr(batch.win_off), _ptr(batch.win_len)), "stage_windows")

    def align_staged(self):
        self._check(self._L.mhmsw_align_staged(self._ctx), "align_staged")

    def submit_seqs(self, reads, ctgs):
        """Submit pairs given as sequences of bytes objects (read p = reads[p], window p = ctgs[p]; None for an
        empty one) through mhmsw_submit_seqs, which packs every sequence where it lies; the job also copies its
        results to pinned host memory."""
        if len(reads) != len(ctgs):
            raise ValueError("reads and ctgs differ in count")
        n = len(reads)
        rp = (ctypes.c_char_p * max(n, 1))(*reads)
        cp = (ctypes.c_char_p * max(n, 1))(*ctgs)
        rl = np.array([len(x) if x is not None else 0 for x in reads], dtype=np.int32)
        cl = np.array([len(x) if x is not None else 0 for x in ctgs], dtype=np.int32)
        self._batch = _SeqBatch(n, (rp, cp, rl, cl, reads, ctgs))
        self._check(self._L.mhmsw_submit_seqs(self._ctx, n, ctypes.cast(rp, ctypes.c_void_p), _ptr(rl),
                                              ctypes.cast(cp, ctypes.c_void_p), _ptr(cl)), "submit_seqs")

    def align_seqs(self, reads, ctgs):
        self.submit_seqs(reads, ctgs)
        self.wait()
        return self.results()

    def submit(self, batch: PairBatch):
        self.stage(batch)
        self.align_staged()

    def query(self):
        return self._check(self._L.mhmsw_query(self._ctx), "query") == 1

    def wait(self):
        self._check(self._L.mhmsw_wait(self._ctx), "wait")

    def results(self):
        n = self._batch.npairs if self._batch is not None else 0
        out = np.zeros(n, dtype=_native.RESULT_DTYPE)
        self._check(self._L.mhmsw_results(self._ctx, _ptr(out), n), "results")
        return out

    def score2(self):
        """Second-best scores of the last batch (int32 per pair); needs set_score2(True)."""
        n = self._batch.npairs if self._batch is not None else 0
        out = np.zeros(n, dtype=np.int32)
        self._check(self._L.mhmsw_results_score2(self._ctx, _ptr(out), n), "results_score2")
        return out

    def cigars(self):
        """(mismatches int32[n], op_off int64[n+1], ops uint32[total]) of the last batch, ops BAM-encoded
        (length << 4 | op, op 0 M, 1 I, 2 D, 4 S, 7 =, 8 X) in pair order; needs set_cigar(True)."""
        n = self._batch.npairs if self._batch is not None else 0
        mism = np.zeros(n, dtype=np.int32)
        off = np.zeros(n + 1, dtype=np.int64)
        total = ctypes.c_int64(0)
        self._check(self._L.mhmsw_results_cigar(self._ctx, n, _ptr(mism), _ptr(off), None, 0, ctypes.byref(total)),
                    "results_cigar")
        ops = np.zeros(max(int(total.value), 1), dtype=np.uint32)
        self._check(self._L.mhmsw_results_cigar(self._ctx, n, None, None, _ptr(ops), int(total.value),
                                                ctypes.byref(total)), "results_cigar")
        return mism, off, ops[:int(total.value)]

    def stats(self):
        st = _native.Stats()
        self._check(self._L.mhmsw_get_stats(self._ctx, ctypes.byref(st)), "get_stats")
        return st.as_dict()

    def align(self, batch: PairBatch):
        self.submit(batch)
        self.wait()
        return self.results()


class _SeqBatch:
    """The batch of a submit_seqs call: its size, and the host objects it reads."""

    def __init__(self, npairs, keep):
        self.npairs = npairs
        self.keep = keep


CIGAR_OPS = "MIDNSHP=X"


def cigar_string(ops):
    """BAM-encoded ops -> CIGAR text as SSW's Alignment::cigar_string spells it (src/ssw/ssw.cpp:57-93)."""
    return "".join(f"{int(o) >> 4}{CIGAR_OPS[int(o) & 0xF]}" for o in ops)


class AlignmentResults:
    """int16 arrays as in adept_sw::AlignmentResults (driver.hpp:17-23)."""

    def __init__(self, res):
        # the reference's short fields (its 1024-base limit kept them in range): a value past the int16 range
        # fails as the reference's errors do
        for f in ("score", "ctg_end", "read_end"):
            if res.shape[0] and int(res[f].max()) > np.iinfo(np.int16).max:
                p = int(np.argmax(res[f]))
                raise RuntimeError(f"GPU assert pair {p}: {f} {int(res[f][p])} does not fit the short fields of "
                                   "AlignmentResults")
        self.ref_begin = res["ctg_begin"].astype(np.int16)
        self.query_begin = res["read_begin"].astype(np.int16)
        self.ref_end = res["ctg_end"].astype(np.int16)
        self.query_end = res["read_end"].astype(np.int16)
        self.top_scores = res["score"].astype(np.int16)


class GPUDriver:
    """Python mirror of adept_sw::GPUDriver (src/adept-sw/driver.hpp:39-55)."""

    def __init__(self):
        self._al = None
        self._results = None
        self._submitted = False

    def init(self, upcxx_rank_me, upcxx_rank_n, match_score, mismatch_score, gap_opening_score,
             gap_extending_score, rlen_limit, ambiguity=1):
        t = time.perf_counter()
        ndev = get_num_node_gpus()
        if ndev <= 0:
            raise RuntimeError("GPU assert no GPU device available")
        scoring = (match_score, -mismatch_score, -gap_opening_score, -gap_extending_score, ambiguity)
        self._al = Aligner(device=upcxx_rank_me % ndev, scoring=scoring, read_is_ssw_ref=True)
        return time.perf_counter() - t

    def run_kernel_forwards(self, reads, contigs, max_read_size=0, max_contig_size=0):
        if self._al is None:
            raise RuntimeError("GPU assert run_kernel_forwards before init")
        self._al.submit(from_strings(reads, contigs))
        self._submitted = True
        self._results = None
        self._n = len(reads)

    def run_kernel_backwards(self, reads, contigs, max_read_size=0, max_contig_size=0):
        if not self._submitted:
            raise RuntimeError("GPU assert run_kernel_backwards before run_kernel_forwards")
        if len(reads) != self._n or len(contigs) != self._n:
            raise RuntimeError("GPU assert run_kernel_backwards with a different batch")

    def kernel_is_done(self):
        if not self._submitted:
            return True
        if not self._al.query():
            return False
        self._fetch()
        return True

    def kernel_block(self):
        if not self._submitted:
            return
        self._al.wait()
        self._fetch()

    def _fetch(self):
        if self._results is None:
            self._results = AlignmentResults(self._al.results())

    def get_aln_results(self):
        return self._results


def get_num_node_gpus():
    L = _native.lib()
    n = ctypes.c_int(0)
    if L.mhmsw_device_count(ctypes.byref(n)) != 0:
        return 0
    return n.value


def get_tot_gpu_mem():
    L = _native.lib()
    b = ctypes.c_size_t(0)
    if L.mhmsw_device_total_mem(0, ctypes.byref(b)) != 0:
        return 0
    return b.value


def get_avail_gpu_mem_per_rank(tot_ranks, num_devices=0):
    if num_devices == 0:
        num_devices = get_num_node_gpus()
    if not num_devices:
        return 0
    return int(get_tot_gpu_mem() * 0.8) // max(1, tot_ranks // num_devices)


class ResidentContigs:
    """Contigs packed to 2-bit codes + N bitmap and kept in device memory."""

    def __init__(self, cs: ContigSet, device=0):
        self._L = _native.lib()
        self.handle = ctypes.c_void_p()
        self.nctgs = cs.nctgs
        rc = self._L.mhmsw_contigs_load(int(device), cs.nctgs, _ptr(cs.seqs), _ptr(cs.off), _ptr(cs.len),
                                        ctypes.byref(self.handle))
        if rc != 0:
            raise _native.MhmswError(rc, "mhmsw_contigs_load failed")
        self.device = device
        self.bytes = self._L.mhmsw_contigs_bytes(self.handle)

    def close(self):
        if self.handle:
            self._L.mhmsw_contigs_free(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching serving loop over device-resident contigs (mhmsw_server_*)."""

    def __init__(self, contigs: ResidentContigs, scoring=MHM2_SCORING, slots=4, hbm_budget=0):
        self._L = _native.lib()
        self._h = ctypes.c_void_p()
        self._contigs = contigs
        sc = _native.Scoring(*[int(x) for x in scoring])
        rc = self._L.mhmsw_server_create(int(contigs.device), ctypes.byref(sc), contigs.handle, int(slots),
                                         int(hbm_budget), ctypes.byref(self._h))
        if rc != 0:
            raise _native.MhmswError(rc, "mhmsw_server_create failed")
        self._pending = {}
        self._extra = {}
        self._outputs = (False, False)

    def set_outputs(self, score2=False, cigar=False):
        """Second-best scores and/or CIGARs for requests admitted from now on (mhmsw_server_set_outputs)."""
        rc = self._L.mhmsw_server_set_outputs(self._h, int(bool(score2)), int(bool(cigar)))
        if rc != 0:
            raise _native.MhmswError(rc, "mhmsw_server_set_outputs failed")
        self._outputs = (bool(score2), bool(cigar))

    def submit(self, batch: WindowBatch):
        rid = ctypes.c_uint64(0)
        rc = self._L.mhmsw_server_submit(self._h, batch.npairs, _ptr(batch.reads), _ptr(batch.read_off),
                                         _ptr(batch.read_len), _ptr(batch.ctg_id), _ptr(batch.win_off),
                                         _ptr(batch.win_len), ctypes.byref(rid))
        if rc != 0:
            msg = self._L.mhmsw_server_last_error(self._h)
            raise _native.MhmswError(rc, f"submit: {msg.decode() if msg else ''}")
        self._pending[rid.value] = batch.npairs
        self._extra[rid.value] = self._outputs
        return rid.value

    def status(self, rid):
        return self._L.mhmsw_server_status(self._h, ctypes.c_uint64(rid))

    def results(self, rid):
        """Wait for request rid; return (results, times dict)."""
        n = self._pending.pop(rid)
        out = np.zeros(n, dtype=_native.RESULT_DTYPE)
        t = _native.ReqTimes()
        rc = self._L.mhmsw_server_results(self._h, ctypes.c_uint64(rid), _ptr(out), n, ctypes.byref(t))
        if rc != 0:
            msg = self._L.mhmsw_server_last_error(self._h)
            raise _native.MhmswError(rc, f"request {rid}: {msg.decode() if msg else ''}")
        return out, {"submit_s": t.submit_s, "admit_s": t.admit_s, "done_s": t.done_s, "device_ms": t.device_ms,
                     "launched_s": t.launched_s, "finished_s": t.finished_s, "job_est_bytes": t.job_est_bytes}

    def extras(self, rid, npairs):
        """After results(rid): (score2 or None, mismatches, op_off, ops) or CIGAR fields None when off."""
        s2_on, cig_on = self._extra.pop(rid, (False, False))
        s2 = np.zeros(npairs, np.int32) if s2_on else None
        mism = off = ops = None
        if cig_on:
            mism = np.zeros(npairs, np.int32)
            off = np.zeros(npairs + 1, np.int64)
            tot = ctypes.c_int64(0)
            rc = self._L.mhmsw_server_results_extra(self._h, ctypes.c_uint64(rid), npairs, _ptr(s2), _ptr(mism),
                                                    _ptr(off), None, 0, ctypes.byref(tot))
            if rc == 0:
                ops = np.zeros(max(tot.value, 1), np.uint32)
                rc = self._L.mhmsw_server_results_extra(self._h, ctypes.c_uint64(rid), npairs, None, None, None,
                                                        _ptr(ops), tot.value, ctypes.byref(tot))
                ops = ops[:tot.value]
        elif s2_on:
            rc = self._L.mhmsw_server_results_extra(self._h, ctypes.c_uint64(rid), npairs, _ptr(s2), None, None,
                                                    None, 0, None)
        else:
            return None, None, None, None
        if rc != 0:
            msg = self._L.mhmsw_server_last_error(self._h)
            raise _native.MhmswError(rc, f"request {rid}: {msg.decode() if msg else ''}")
        return s2, mism, off, ops

    def clock(self):
        return self._L.mhmsw_server_clock()

    def close(self):
        if self._h:
            self._L.mhmsw_server_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
