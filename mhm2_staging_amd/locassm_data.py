"""Local assembly inputs: contigs with the reads that hang over each end.

The layout is the one the local assembly entry points take
(``mhmla_stage``): all contigs in one ASCII buffer, all reads (bases and
quality bytes at the same offsets) in two more, and per contig the ranges of
``read_idx`` that list its left and its right reads.  It carries what
``CtgWithReads`` carries in the reference (``gpu_loc_assem/locassem_struct.hpp``:
id, sequence, depth, reads_left, reads_right).
"""
from dataclasses import dataclass

import numpy as np

_COMP = np.full(256, 0, dtype=np.uint8)
for _a, _b in zip(b"ACGTN", b"TGCAN"):
    _COMP[_a] = _b


@dataclass
class LocAssmBatch:
    cid: np.ndarray          # int64 contig ids
    ctg_seq: np.ndarray      # uint8, NUL-terminated
    ctg_off: np.ndarray      # int64
    ctg_len: np.ndarray      # int32
    depth: np.ndarray        # float64
    left_first: np.ndarray   # int64 [nctgs + 1] into read_idx
    right_first: np.ndarray  # int64 [nctgs + 1] into read_idx
    read_idx: np.ndarray     # int64 read numbers
    read_seq: np.ndarray     # uint8
    read_qual: np.ndarray    # uint8, same offsets as read_seq
    read_off: np.ndarray     # int64
    read_len: np.ndarray     # int32

    @property
    def nctgs(self):
        return int(self.ctg_len.shape[0])

    @property
    def nreads(self):
        return int(self.read_len.shape[0])

    def ctg(self, c):
        return bytes(self.ctg_seq[self.ctg_off[c]:self.ctg_off[c] + self.ctg_len[c]]).decode()

    def reads_of(self, c, side):
        first = self.left_first if side == "left" else self.right_first
        out = []
        for r in self.read_idx[first[c]:first[c + 1]]:
            o, n = int(self.read_off[r]), int(self.read_len[r])
            out.append((bytes(self.read_seq[o:o + n]).decode(), bytes(self.read_qual[o:o + n]).decode()))
        return out

    def mer_occurrences(self, k, max_reads=3000):
        """Mers counted at mer length k over both sides (first max_reads reads of each)."""
        total = 0
        for first in (self.left_first, self.right_first):
            counts = np.minimum(np.diff(first), max_reads)
            for c in np.nonzero(counts)[0]:
                lens = self.read_len[self.read_idx[first[c]:first[c] + counts[c]]].astype(np.int64)
                total += int(np.maximum(lens - k, 0).sum())
        return total

    def subset(self, idx):
        return from_contigs([(self.ctg(int(c)), float(self.depth[c]), self.reads_of(int(c), "left"),
                              self.reads_of(int(c), "right")) for c in idx],
                            cids=[int(self.cid[c]) for c in idx])


def from_contigs(contigs, cids=None):
    """contigs: [(seq, depth, [(read, qual), ...] left, [(read, qual), ...] right)]."""
    ctg = [c[0].encode() if isinstance(c[0], str) else bytes(c[0]) for c in contigs]
    reads, quals, left_first, right_first, idx = [], [], [0], [0], []
    lefts, rights = [], []
    for c in contigs:
        for side, lst in ((2, lefts), (3, rights)):
            ids = []
            for r, q in c[side]:
                r = r.encode() if isinstance(r, str) else bytes(r)
                q = q.encode() if isinstance(q, str) else bytes(q)
                if len(r) != len(q):
                    raise ValueError("a read and its quality string differ in length")
                ids.append(len(reads))
                reads.append(r)
                quals.append(q)
            lst.append(ids)
    for ids in lefts:
        idx.extend(ids)
        left_first.append(len(idx))
    right_first = [len(idx)]
    for ids in rights:
        idx.extend(ids)
        right_first.append(len(idx))

    def pack(seqs):
        lens = np.array([len(s) for s in seqs], dtype=np.int32)
        offs = np.zeros(len(seqs), dtype=np.int64)
        if len(seqs) > 1:
            offs[1:] = np.cumsum(lens[:-1], dtype=np.int64)
        return np.frombuffer(b"".join(seqs) + b"\0", dtype=np.uint8).copy(), offs, lens

    cs, co, cl = pack(ctg)
    rs, ro, rl = pack(reads)
    qs, _, _ = pack(quals)
    return LocAssmBatch(np.asarray(cids if cids is not None else range(len(ctg)), dtype=np.int64), cs, co, cl,
                        np.array([float(c[1]) for c in contigs], dtype=np.float64),
                        np.array(left_first, dtype=np.int64), np.array(right_first, dtype=np.int64),
                        np.array(idx, dtype=np.int64), rs, qs, ro, rl)


def generate(nctgs, seed=0, ctg_len=(200, 2000), read_len=150, flank=520, sub_rate=0.01, n_rate=0.001,
             fork_rate=0.1, repeat_rate=0.05, qual_offset=33):
    """Seeded local assembly batch shaped like MHM2's localassm input.

    Each contig is the middle of a random source sequence with ``flank``
    bases on each side.  Each side gets 0 reads (10% of sides), 1..9 reads
    (55%) or 10..80 reads (35%), of ``read_len`` bases starting uniformly
    between 140 bases inside the contig end and 350 bases past it, with
    substitutions at ``sub_rate`` (quality 2..15), N at ``n_rate`` (quality
    2), other bases at quality 25..40 except 3% at 2..19.  Left reads
    are drawn from the reverse complement of the source, the orientation in
    which the left side is walked.  ``fork_rate`` of the sides draw every
    second read from a second haplotype with one substitution in the flank,
    at the previous read's position (equal support: a fork); ``repeat_rate`` of the sides copy the contig's last 60 bases
    into the flank 200 bases out (a repeat for mer lengths under 60).
    Depth is uniform in [2, 50).
    """
    flank = max(flank, 350 + read_len)   # reads start at most 350 bases past the contig end
    rng = np.random.default_rng(seed)
    alpha = np.frombuffer(b"ACGT", dtype=np.uint8)
    clen = rng.integers(ctg_len[0], ctg_len[1] + 1, nctgs).astype(np.int64)
    slen = clen + 2 * flank
    soff = np.zeros(nctgs + 1, dtype=np.int64)
    soff[1:] = np.cumsum(slen)
    src = alpha[rng.integers(0, 4, int(soff[-1]))]
    depth = rng.uniform(2.0, 50.0, nctgs)

    # per side (2c = left, 2c + 1 = right): read counts
    kind = rng.random(2 * nctgs)
    nreads = np.where(kind < 0.10, 0, np.where(kind < 0.65, rng.integers(1, 10, 2 * nctgs),
                                                rng.integers(10, 81, 2 * nctgs)))
    fork = rng.random(2 * nctgs) < fork_rate
    rep = rng.random(2 * nctgs) < repeat_rate

    # repeats: copy the last 60 contig bases (in the side's orientation) 200 bases past the end
    for s in np.nonzero(rep)[0]:
        c, right = s // 2, s % 2 == 1
        o = soff[c]
        if right:
            end = o + flank + clen[c]
            src[end + 200:end + 260] = src[end - 60:end]
        else:  # in the reverse complement the contig's start is the end
            beg = o + flank
            src[beg - 260:beg - 200] = src[beg:beg + 60]

    side_of_read = np.repeat(np.arange(2 * nctgs), nreads)
    nr = side_of_read.shape[0]
    ctg_of_read = side_of_read // 2
    is_right = (side_of_read % 2) == 1
    # start relative to the contig end, in the side's orientation
    rel = rng.integers(-140, 351, nr)
    # read position in source coordinates (right) or reverse-complement coordinates (left)
    end_in_side = flank + clen[ctg_of_read]          # contig end, same in both orientations
    start = end_in_side + rel                        # start of the read in side coordinates
    # fork sides: every second read is the previous one's start on the other haplotype
    side_first = np.zeros(2 * nctgs + 1, dtype=np.int64)
    side_first[1:] = np.cumsum(nreads)
    within = np.arange(nr) - side_first[side_of_read]
    alt = fork[side_of_read] & (within % 2 == 1)
    start = np.where(alt, np.roll(start, 1), start)
    snp_at = end_in_side + 60 + (side_of_read % 97)  # one fixed position per side, 60..156 past the end

    read_seq = np.empty(nr * read_len, dtype=np.uint8)
    read_qual = np.empty(nr * read_len, dtype=np.uint8)
    code = np.zeros(256, dtype=np.int64)
    code[alpha] = np.arange(4)
    chunk = 50000
    ar = np.arange(read_len)
    for a in range(0, nr, chunk):
        b = min(nr, a + chunk)
        pos = start[a:b, None] + ar[None, :]                      # side coordinates
        c = ctg_of_read[a:b, None]
        right = is_right[a:b, None]
        src_pos = np.where(right, soff[c] + pos, soff[c] + slen[c] - 1 - pos)
        bases = src[src_pos]
        bases = np.where(right, bases, _COMP[bases])
        hit = alt[a:b, None] & (pos == snp_at[a:b, None])
        bases[hit] = alpha[(code[bases[hit]] + 1) % 4]
        q = rng.integers(25, 41, (b - a, read_len), dtype=np.uint8)
        u = rng.random((b - a, read_len), dtype=np.float32)
        sub = u < sub_rate
        nn = (u >= sub_rate) & (u < sub_rate + n_rate)
        low = (u >= sub_rate + n_rate) & (u < sub_rate + n_rate + 0.03)
        ns = int(sub.sum())
        bases[sub] = alpha[(code[bases[sub]] + rng.integers(1, 4, ns)) % 4]
        q[sub] = rng.integers(2, 16, ns, dtype=np.uint8)
        q[low] = rng.integers(2, 20, int(low.sum()), dtype=np.uint8)
        bases[nn] = ord("N")
        q[nn] = 2
        read_seq[a * read_len:b * read_len] = bases.reshape(-1)
        read_qual[a * read_len:b * read_len] = (q + qual_offset).reshape(-1)

    ctg_seq = np.empty(int(clen.sum()) + 1, dtype=np.uint8)
    ctg_off = np.zeros(nctgs, dtype=np.int64)
    ctg_off[1:] = np.cumsum(clen[:-1])
    for c in range(nctgs):
        ctg_seq[ctg_off[c]:ctg_off[c] + clen[c]] = src[soff[c] + flank:soff[c] + flank + clen[c]]
    ctg_seq[-1] = 0

    # reads are numbered in side order; left reads of contig c are the reads of side 2c
    read_off = np.arange(nr, dtype=np.int64) * read_len
    rl = np.full(nr, read_len, dtype=np.int32)
    order_left = np.nonzero(~is_right)[0]
    order_right = np.nonzero(is_right)[0]
    read_idx = np.concatenate([order_left, order_right]).astype(np.int64)
    nl = nreads[0::2].astype(np.int64)
    nrr = nreads[1::2].astype(np.int64)
    left_first = np.zeros(nctgs + 1, dtype=np.int64)
    left_first[1:] = np.cumsum(nl)
    right_first = np.zeros(nctgs + 1, dtype=np.int64)
    right_first[1:] = left_first[-1] + np.cumsum(nrr)
    right_first[0] = left_first[-1]
    seq = np.concatenate([read_seq, np.zeros(1, np.uint8)])
    qual = np.concatenate([read_qual, np.zeros(1, np.uint8)])
    return LocAssmBatch(np.arange(nctgs, dtype=np.int64), ctg_seq, ctg_off, clen.astype(np.int32), depth,
                        left_first, right_first, read_idx, seq, qual, read_off, rl)
