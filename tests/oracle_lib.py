"""ctypes access to the oracle (test infrastructure only).

``liboracle.so`` is the scalar C restatement in ``oracle/sw_oracle.c``;
``_ref/libssw_ref.so`` is the reference SSW compiled from its own sources by
``oracle/Makefile``.  Both take the same batch arguments: pair p aligns
SSW-ref ``a[a_off[p]:a_off[p]+a_len[p]]`` against SSW-query ``b[...]``.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libssw_ref.so")

RESULT_DTYPE = np.dtype([("score", "<i4"), ("ref_begin", "<i4"), ("ref_end", "<i4"),
                         ("query_begin", "<i4"), ("query_end", "<i4")])


class Scoring(ctypes.Structure):
    _fields_ = [("match", ctypes.c_int32), ("mismatch", ctypes.c_int32), ("gap_open", ctypes.c_int32),
                ("gap_ext", ctypes.c_int32), ("ambiguity", ctypes.c_int32)]


_libs = {}


def _load(path, fn, with_score2=False):
    key = (path, fn)
    if key not in _libs:
        if not os.path.exists(path):
            return None
        lib = ctypes.CDLL(path)
        f = getattr(lib, fn)
        f.restype = ctypes.c_int
        args = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(Scoring), ctypes.c_void_p]
        if with_score2:
            args.append(ctypes.c_void_p)
        f.argtypes = args + [ctypes.c_int]
        _libs[key] = f
    return _libs[key]


def have_ref():
    return os.path.exists(REF_SO)


def pack_strings(seqs):
    """Concatenate byte strings; return (buffer, offsets int64, lengths int32)."""
    seqs = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    lens = np.array([len(s) for s in seqs], dtype=np.int32)
    offs = np.zeros(len(seqs), dtype=np.int64)
    if len(seqs) > 1:
        offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
    buf = np.frombuffer(b"".join(seqs) + b"\0", dtype=np.uint8).copy()
    return buf, offs, lens


def _run(fn, a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads, score2=None):
    n = len(a_len)
    out = np.zeros(n, dtype=RESULT_DTYPE)
    sc = Scoring(*scoring)
    a_buf = np.ascontiguousarray(a_buf, dtype=np.uint8)
    b_buf = np.ascontiguousarray(b_buf, dtype=np.uint8)
    a_off = np.ascontiguousarray(a_off, dtype=np.int64)
    b_off = np.ascontiguousarray(b_off, dtype=np.int64)
    a_len = np.ascontiguousarray(a_len, dtype=np.int32)
    b_len = np.ascontiguousarray(b_len, dtype=np.int32)
    extra = () if score2 is None else (score2.ctypes.data,)
    fn(n, a_buf.ctypes.data, a_off.ctypes.data, a_len.ctypes.data, b_buf.ctypes.data, b_off.ctypes.data,
       b_len.ctypes.data, ctypes.byref(sc), out.ctypes.data, *extra, nthreads)
    return out


def oracle_batch2(a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads=0):
    """(results, score2): the oracle with SSW's second-best score."""
    fn = _load(ORACLE_SO, "oracle_align_batch2", True)
    if fn is None:
        raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
    s2 = np.zeros(len(a_len), dtype=np.int32)
    return _run(fn, a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads, s2), s2


def ref_batch2(a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads=0):
    """(results, score2) from the reference SSW (sw_score_next_best)."""
    fn = _load(REF_SO, "ssw_ref_align_batch2", True)
    if fn is None:
        raise RuntimeError("oracle/_ref/libssw_ref.so is not built")
    s2 = np.zeros(len(a_len), dtype=np.int32)
    return _run(fn, a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads, s2), s2


def oracle_batch(a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads=0):
    fn = _load(ORACLE_SO, "oracle_align_batch")
    if fn is None:
        raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
    return _run(fn, a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads)


def ref_batch(a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads=0):
    fn = _load(REF_SO, "ssw_ref_align_batch")
    if fn is None:
        raise RuntimeError("oracle/_ref/libssw_ref.so is not built")
    return _run(fn, a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, nthreads)


def oracle_pairs(refs, queries, scoring, nthreads=0):
    a = pack_strings(refs)
    b = pack_strings(queries)
    return oracle_batch(*a, *b, scoring, nthreads)


def ref_pairs(refs, queries, scoring, nthreads=0):
    a = pack_strings(refs)
    b = pack_strings(queries)
    return ref_batch(*a, *b, scoring, nthreads)


CIGAR_OPS = "MIDNSHP=X"


def cigar_string(ops):
    """BAM-encoded ops -> CIGAR text, as ssw.cpp's ConvertAlignment prints them."""
    return "".join(f"{int(o) >> 4}{CIGAR_OPS[int(o) & 0xF]}" for o in ops)


def _cigar_run(path, name, a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, maxops, nthreads):
    key = (path, name)
    if key not in _libs:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built")
        f = getattr(ctypes.CDLL(path), name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_int64] + [ctypes.c_void_p] * 6 + [ctypes.POINTER(Scoring)] + [ctypes.c_void_p] * 4 + \
            [ctypes.c_int, ctypes.c_int]
        _libs[key] = f
    fn = _libs[key]
    n = len(a_len)
    if maxops is None:
        maxops = int(2 * (max(np.max(a_len, initial=0), np.max(b_len, initial=0))) + 8)
    out = np.zeros(n, dtype=RESULT_DTYPE)
    mism = np.zeros(n, dtype=np.int32)
    nops = np.zeros(n, dtype=np.int32)
    ops = np.zeros((n, maxops), dtype=np.uint32)
    sc = Scoring(*scoring)
    arrs = [np.ascontiguousarray(x, dtype=d) for x, d in ((a_buf, np.uint8), (a_off, np.int64), (a_len, np.int32),
                                                          (b_buf, np.uint8), (b_off, np.int64), (b_len, np.int32))]
    fn(n, *[x.ctypes.data for x in arrs], ctypes.byref(sc), out.ctypes.data, mism.ctypes.data, nops.ctypes.data,
       ops.ctypes.data, maxops, nthreads)
    return out, mism, nops, ops


def oracle_cigar_batch(a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, maxops=None, nthreads=0):
    """(results, mismatches, nops, ops[npairs, maxops]) from the oracle's CIGAR restatement."""
    return _cigar_run(ORACLE_SO, "oracle_cigar_batch", a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, maxops,
                      nthreads)


def ref_cigar_batch(a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, maxops=None, nthreads=0):
    """The same from the reference SSW called with report_cigar = true."""
    return _cigar_run(REF_SO, "ssw_ref_cigar_batch", a_buf, a_off, a_len, b_buf, b_off, b_len, scoring, maxops,
                      nthreads)


# ---- local assembly (oracle/locassm_oracle.c) ------------------------------

class LaParams(ctypes.Structure):
    _fields_ = [("kmer_len", ctypes.c_int32), ("min_kmer_len", ctypes.c_int32), ("max_kmer_len", ctypes.c_int32),
                ("shift", ctypes.c_int32), ("walk_len_limit", ctypes.c_int32), ("qual_offset", ctypes.c_int32),
                ("min_qual", ctypes.c_int32), ("min_hi_qual", ctypes.c_int32), ("rating_thres", ctypes.c_int32),
                ("max_reads", ctypes.c_int32), ("min_viable_depth", ctypes.c_double),
                ("min_expected_depth", ctypes.c_double)]


def la_params(**kw):
    p = LaParams()
    _la().oracle_la_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


_la_lib = None


def _la():
    global _la_lib
    if _la_lib is None:
        L = ctypes.CDLL(ORACLE_SO)
        P = ctypes.POINTER(LaParams)
        L.oracle_la_default_params.argtypes = [P]
        L.oracle_la_rating.argtypes = [P, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int]
        L.oracle_la_rating.restype = ctypes.c_int
        L.oracle_la_ext.argtypes = [P, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.oracle_la_ext.restype = ctypes.c_int
        L.oracle_la_sort_keys.argtypes = [ctypes.c_void_p]
        L.oracle_la_walk_side.argtypes = [P, ctypes.c_char_p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        L.oracle_la_walk_side.restype = ctypes.c_int
        L.oracle_la_extend_batch.argtypes = [P, ctypes.c_int64] + [ctypes.c_void_p] * 16 + [ctypes.c_int]
        L.oracle_la_extend_batch.restype = ctypes.c_int
        _la_lib = L
    return _la_lib


def la_rating(p, nvotes, hi, depth):
    return _la().oracle_la_rating(ctypes.byref(p), nvotes, hi, depth)


def la_ext(p, low, hi, depth):
    lo = np.ascontiguousarray(low, dtype=np.uint32)
    h = np.ascontiguousarray(hi, dtype=np.uint32)
    cnt = np.zeros(1, dtype=np.uint32)
    e = _la().oracle_la_ext(ctypes.byref(p), lo.ctypes.data, h.ctypes.data, depth, cnt.ctypes.data)
    return chr(e), int(cnt[0])


def la_sort_keys(keys):
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint32).reshape(12)).copy()
    _la().oracle_la_sort_keys(k.ctypes.data)
    return k.reshape(4, 3)


def la_walk_side(p, seq, depth, reads):
    """reads: [(bases, quals)]; returns (walk, termination of the last walk,
    number of walks, (walks ending in X, F, R))."""
    seq = seq.encode() if isinstance(seq, str) else bytes(seq)
    rs = [r.encode() if isinstance(r, str) else bytes(r) for r, _ in reads]
    qs = [q.encode() if isinstance(q, str) else bytes(q) for _, q in reads]
    rb, ro, rl = pack_strings(rs)
    qb, _, _ = pack_strings(qs)
    idx = np.arange(len(rs), dtype=np.int64)
    walk = ctypes.create_string_buffer(p.walk_len_limit + 1)
    term = ctypes.create_string_buffer(1)
    tc = np.zeros(3, np.int32)
    n = _la().oracle_la_walk_side(ctypes.byref(p), seq, len(seq), float(depth), len(rs), rb.ctypes.data,
                                  qb.ctypes.data, ro.ctypes.data, rl.ctypes.data, idx.ctypes.data,
                                  ctypes.cast(walk, ctypes.c_void_p), tc.ctypes.data, ctypes.cast(term, ctypes.c_void_p))
    return walk.raw[:n].decode(), term.raw[:1].decode(), int(tc.sum()), tuple(int(x) for x in tc)


def la_extend_batch(p, b, nthreads=0, with_terms=False):
    """b: mhm2_staging_amd.locassm_data.LocAssmBatch -> (left_len, left_ext, right_len, right_ext[, terms]);
    *_ext are uint8 [nctgs, walk_len_limit]; terms int32 [nctgs, 2, 3] (X F R, left then right)."""
    n = b.nctgs
    W = p.walk_len_limit
    ll = np.zeros(n, np.int32)
    rl = np.zeros(n, np.int32)
    le = np.zeros((n, max(W, 1)), np.uint8)
    re = np.zeros((n, max(W, 1)), np.uint8)
    tm = np.zeros((n, 2, 3), np.int32)
    rc = _la().oracle_la_extend_batch(ctypes.byref(p), n, b.ctg_seq.ctypes.data, b.ctg_off.ctypes.data,
                                      b.ctg_len.ctypes.data, b.depth.ctypes.data, b.left_first.ctypes.data,
                                      b.right_first.ctypes.data, b.read_idx.ctypes.data, b.read_seq.ctypes.data,
                                      b.read_qual.ctypes.data, b.read_off.ctypes.data, b.read_len.ctypes.data,
                                      ll.ctypes.data, le.ctypes.data, rl.ctypes.data, re.ctypes.data, tm.ctypes.data,
                                      int(nthreads))
    if rc != 0:
        raise ValueError("a contig holds a byte outside ACGTN and the IUPAC codes")
    return (ll, le, rl, re, tm) if with_terms else (ll, le, rl, re)
