// Per-pair CIGAR work shared by the HIP kernels (cigar_kernels.hip) and a
// CPU check of the same code (tests/cpp/cigar_core_check.cpp).
//
// What it computes is SSW's CIGAR for klign's call with report_cigar set
// (src/ssw/ssw_core.cpp:1180-1192 and banded_sw :854-1053, then
// ConvertAlignment src/ssw/ssw.cpp:57-93 and CalculateNumberMismatch
// :129-254):
//   * rectangle: rows i over the SSW-query segment [qb, qe], columns j over
//     the SSW-ref segment [rb, re]; band |j - i| <= w, w starting at
//     |columns - rows| + 1 and doubling until the band holds the score;
//   * band DP: every cell outside the band, above row 0 or left of column 0
//     counts as H = E = F = 0; E(i,j) = max(H(i-1,j) - gO, E(i-1,j) - gE),
//     F(i,j) = max(H(i,j-1) - gO, F(i,j-1) - gE), H = max(E, F, 0,
//     H(i-1,j-1) + s).  E and F are kept clamped at 0: SSW keeps them
//     unclamped but only uses max(E, 0) and max(F, 0) in H, and its
//     open/extend choice for a cell differs from the clamped one only where
//     the gap value is <= 0, which no traceback reads (a traceback enters a
//     gap state only where that gap value is positive);
//   * per cell 4 direction bits: bits 0-1 what H took -- 0 the diagonal
//     (max(E, F) <= H(i-1,j-1) + s) with equal codes, 1 the diagonal with
//     different codes, 2 E (E > F), 3 F -- bit 2 "E was opened" (H(i-1,j) -
//     gO > E(i-1,j) - gE), bit 3 "F was opened";
//   * the band holds the score exactly when the rectangle's last cell does:
//     every cell of the rectangle is at most the forward score and only the
//     end point reaches it (the end is the least such cell);
//   * traceback from the last cell in state H while the row is above 0
//     (diagonal = M, E = I, F = D); the cell where it stops is one more M.
//     The CIGAR then walks from (rb, qb): M positions compare the codes
//     ('=' when equal, N against N included, else 'X'), so when the
//     traceback stops at column jf > 0 every compared SSW-ref position is
//     shifted by jf, as in SSW.  A traceback that stops at column 0 (always,
//     when match <= gap_open: a path that starts with a gap never beats the
//     one that skips it) takes '=' / 'X' from the cells' equality bits; one
//     that stops elsewhere is walked again comparing shifted codes;
//   * soft clips: qb bases before, n - 1 - qe after; mismatches = X + I + D.
//
// Band cells are addressed by diagonal d = j - i + BW in [0, 2*BW]; a row
// keeps H and E of its 2*BW+1 diagonals, so cell (i, d) reads (i-1, d) for
// the diagonal, (i-1, d+1) for E and (i, d-1) for F.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define MHMSW_HD __host__ __device__
#else
#define MHMSW_HD
#endif

namespace mhmsw {

// BAM CIGAR op codes (MAPSTR "MIDNSHP=X", src/ssw/ssw_core.hpp:23)
constexpr uint32_t kOpM = 0, kOpI = 1, kOpD = 2, kOpS = 4, kOpEq = 7, kOpX = 8;
// SSW's distance filter for klign's Filter (src/ssw/ssw.hpp:59-64)
constexpr int32_t kCigarDistanceFilter = 32767;
// band classes: half-width 1..7 one lane per pair with the band row in
// registers (class bw - 1); wider bands, and rectangles taller than
// kCigNarrowMaxRows, one wavefront per pair (the wide class)
constexpr int kCigExactMaxBw = 7;
constexpr int kCigClassWide = 7;
constexpr int kCigClasses = 8;
constexpr int32_t kCigNarrowMaxRows = 2048;

MHMSW_HD inline int cigar_class(int32_t bw, int32_t rows) {
  if (rows > kCigNarrowMaxRows || bw > kCigExactMaxBw) return kCigClassWide;
  return bw - 1;
}

MHMSW_HD inline int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }
MHMSW_HD inline int32_t imax3(int32_t a, int32_t b, int32_t c) { return imax(imax(a, b), c); }

// One pair's rectangle in SSW naming.
struct CigarGeo {
  int64_t rbase, qbase;  // nibble index of SSW-ref / SSW-query position 0
  int32_t rb, re, qb, qe;
  int32_t n;             // SSW-query length
  int32_t score;
  MHMSW_HD int32_t rl() const { return re - rb + 1; }
  MHMSW_HD int32_t ql() const { return qe - qb + 1; }
};

MHMSW_HD inline uint32_t nib_at(const uint32_t* codes, int64_t idx) {
  return (codes[idx >> 3] >> ((uint32_t)(idx & 7) * 4u)) & 0xFu;
}

// Sequential nibble reader that loads each dword once.
struct NibCursor {
  int64_t dw = -1;
  uint32_t val = 0;
  MHMSW_HD uint32_t get(const uint32_t* codes, int64_t idx) {
    if ((idx >> 3) != dw) {
      dw = idx >> 3;
      val = codes[dw];
    }
    return (val >> ((uint32_t)(idx & 7) * 4u)) & 0xFu;
  }
};

// Substitution scores of one row against SSW-ref codes 0..7, biased by +128,
// as the 8 bytes (lo, hi) of a byte-select table: codes 0..3 score match or
// -mismatch, 4 (N) -ambiguity, 5 (left of column 0) a value <= 0.  A row
// whose code is N scores -ambiguity everywhere.
MHMSW_HD inline void subst_table(uint32_t qc, int32_t match, int32_t mismatch, int32_t amb, uint32_t& lo,
                                 uint32_t& hi) {
  const uint32_t bx = (uint32_t)(128 - (qc < 4 ? mismatch : amb));
  const uint32_t ba = (uint32_t)(128 - amb);
  lo = bx * 0x01010101u;
  if (qc < 4) lo = (lo & ~(0xFFu << (8 * qc))) | ((uint32_t)(128 + match) << (8 * qc));
  hi = ba | (bx << 8);
}

MHMSW_HD inline int32_t subst_lookup(uint32_t lo, uint32_t hi, uint32_t rc) {
#ifdef __HIP_DEVICE_COMPILE__
  return (int32_t)(__builtin_amdgcn_perm(hi, lo, rc) & 0xFFu) - 128;
#else
  const uint64_t t = ((uint64_t)hi << 32) | lo;
  return (int32_t)((t >> (8 * rc)) & 0xFFu) - 128;
#endif
}

constexpr uint32_t kCigKill = 5;  // SSW-ref code left of column 0

// The 4 direction bits of a cell (t1 = max(E, F), t2 = H(i-1,j-1) + s).
MHMSW_HD inline uint32_t cell_bits(int32_t t1, int32_t t2, int32_t e, int32_t f, bool eq, bool eopen, bool fopen) {
  const uint32_t h = t1 > t2 ? (e > f ? 2u : 3u) : (eq ? 0u : 1u);
  return h | (eopen ? 4u : 0u) | (fopen ? 8u : 0u);
}

// One traceback step at a cell with bits nib from state (0 E, 1 F, 2 H):
// the op (kOpEq / kOpX for the diagonal as the equality bit says, kOpI,
// kOpD); state becomes the state at the next cell.
MHMSW_HD inline uint32_t step_op(uint32_t nib, int& state) {
  if (state == 2) {
    const uint32_t h = nib & 3u;
    if (h < 2u) return h == 0u ? kOpEq : kOpX;
    if (h == 2u) {
      state = (nib & 4u) ? 2 : 0;
      return kOpI;
    }
    state = (nib & 8u) ? 2 : 1;
    return kOpD;
  }
  if (state == 0) {
    state = (nib & 4u) ? 2 : 0;
    return kOpI;
  }
  state = (nib & 8u) ? 2 : 1;
  return kOpD;
}

// 8 nibble codes starting at nibble index a (any alignment): two dwords and
// a funnel shift.
MHMSW_HD inline uint32_t load8(const uint32_t* codes, int64_t a) {
  const uint32_t lo = codes[a >> 3], hi = codes[(a >> 3) + 1];
  const uint32_t sh = (uint32_t)(a & 7) * 4u;
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
  return (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
#endif
}

// Direction layouts: 4 bits per cell, diagonal d at bit 4*(d % 8) of word
// d / 8 of its row; a row takes DS words (at least the (4W+31)/32 its W
// cells need).
//  * blocked (narrow bands, one lane per pair): rows 8b..8b+7 of a lane are
//    8*DS consecutive words, the blocks of the 64 lanes of a wavefront follow
//    each other, so a lane reads or writes 8 rows with 2*DS 16-byte
//    accesses; dirs points at the lane's first block;
//  * row-major (the serial path of the widest bands): row i at i * DS.
MHMSW_HD inline int64_t dir_at(bool blocked, int DS, int32_t i, int32_t k) {
  return blocked ? (int64_t)(i >> 3) * 512 * DS + (i & 7) * DS + k : (int64_t)i * DS + k;
}

struct Words4 {
  uint32_t w[4];
};

// Band DP for half-width BW, directions in the blocked layout with DS words
// per row.
// Returns H of the rectangle's last cell.  Rows go in blocks of 8 whose codes
// (one SSW-query code per row, the SSW-ref code entering the band) are loaded
// one block ahead, so no row waits for memory.
template <int BW, int DS>
MHMSW_HD int32_t cigar_band_dp(const uint32_t* codes, const CigarGeo& g, int32_t match, int32_t mismatch, int32_t amb,
                               int32_t gO, int32_t gE, uint32_t* dirs) {
  constexpr int W = 2 * BW + 1;
  static_assert((4 * W + 31) / 32 <= DS, "a row's words");
  const int32_t rl = g.rl(), ql = g.ql();
  int32_t H[W], E[W];
  uint32_t rc[W];
#pragma unroll
  for (int d = 0; d < W; d++) {
    H[d] = 0;
    E[d] = 0;
    const int32_t j = d - BW;
    rc[d] = j < 0 ? kCigKill : nib_at(codes, g.rbase + g.rb + j);
  }
  // row i's SSW-query code at qa + i; the SSW-ref code entering the band
  // after row i (column i + 1 + BW) at ra + i
  const int64_t qa = g.qbase + g.qb, ra = g.rbase + g.rb + BW + 1;
  uint32_t qn = load8(codes, qa), rn = load8(codes, ra);
  for (int32_t i0 = 0; i0 < ql; i0 += 8) {
    const uint32_t qw = qn, rw = rn;
    qn = load8(codes, qa + i0 + 8);
    rn = load8(codes, ra + i0 + 8);
    uint32_t* out = dirs + dir_at(true, DS, i0, 0);
#pragma unroll
    for (int r = 0; r < 8; r++) {
      if (i0 + r >= ql) break;
      const uint32_t qc = (qw >> (4 * r)) & 0xFu;
      uint32_t tlo, thi;
      subst_table(qc, match, mismatch, amb, tlo, thi);
      int32_t hl = 0, fl = 0;
      uint32_t dw[DS];
#pragma unroll
      for (int k = 0; k < DS; k++) dw[k] = 0;
#pragma unroll
      for (int d = 0; d < W; d++) {
        const int32_t hu = d + 1 < W ? H[d + 1] : 0;
        const int32_t eu = d + 1 < W ? E[d + 1] : 0;
        const int32_t eo = hu - gO, ee = eu - gE;
        const int32_t e = imax3(eo, ee, 0);
        const int32_t fo = hl - gO, fe = fl - gE;
        const int32_t f = imax3(fo, fe, 0);
        const int32_t t2 = H[d] + subst_lookup(tlo, thi, rc[d]);
        const int32_t t1 = imax(e, f);
        const int32_t h = imax(t1, t2);
        dw[(4 * d) / 32] |= cell_bits(t1, t2, e, f, rc[d] == qc, eo > ee, fo > fe) << ((4 * d) % 32);
        H[d] = h;
        E[d] = e;
        hl = h;
        fl = f;
      }
#pragma unroll
      for (int k = 0; k < DS; k++) out[r * DS + k] = dw[k];
#pragma unroll
      for (int d = 0; d < W - 1; d++) rc[d] = rc[d + 1];
      rc[W - 1] = (rw >> (4 * r)) & 0xFu;
    }
  }
  // last cell: row ql - 1, column rl - 1
  const int32_t dc = (rl - 1) - (ql - 1) + BW;
  int32_t corner = -1;
#pragma unroll
  for (int d = 0; d < W; d++) corner = d == dc ? H[d] : corner;
  return corner;
}

// Run-length emitter of the CIGAR, fed in reverse order (last op first).
struct CigarEmit {
  uint32_t* out;   // null: count only
  int32_t nops;    // total ops (for placing them from the end), write mode
  int32_t count = 0;
  int32_t mism = 0;
  uint32_t op = 0xFFu;
  uint32_t len = 0;
  MHMSW_HD void flush() {
    if (len == 0) return;
    if (out && count < nops) out[nops - 1 - count] = (len << 4) | op;
    count++;
    if (op == kOpX || op == kOpI || op == kOpD) mism += (int32_t)len;
    len = 0;
  }
  MHMSW_HD void push(uint32_t o, uint32_t l) {
    if (l == 0) return;
    if (len && o != op) flush();
    op = o;
    len += l;
  }
};

// Run-length emitter for narrow bands that keeps the CIGAR in registers:
// up to 16 ops between the soft clips as 16-bit halves (length << 4 | op;
// lengths below 4096, which rectangles of at most kCigNarrowMaxRows rows
// guarantee) in q[0..7], newest first, and the two clips apart.  The walk
// emits backwards, so the newest op is the first in CIGAR order: op m of
// the part between the clips is half m of q.  fits() tells whether the
// whole CIGAR was kept.
struct CigarEmitQ {
  int32_t count = 0;
  int32_t mism = 0;
  int32_t body = 0;
  uint32_t op = 0xFFu;
  uint32_t len = 0;
  uint32_t clip_tail = 0, clip_head = 0;
  bool long_run = false;
  uint32_t q[8];
  MHMSW_HD CigarEmitQ() {
#pragma unroll
    for (int k = 0; k < 8; k++) q[k] = 0;
  }
  MHMSW_HD void flush() {
    if (len == 0) return;
    count++;
    if (op == kOpS) {
      if (body == 0 && clip_tail == 0 && count == 1) {
        clip_tail = len;
      } else {
        clip_head = len;
      }
      len = 0;
      return;
    }
    if (op == kOpX || op == kOpI || op == kOpD) mism += (int32_t)len;
    long_run = long_run || len >= 4096u;
    const uint32_t v = ((len << 4) | op) & 0xFFFFu;
#pragma unroll
    for (int k = 7; k > 0; k--) q[k] = (q[k] << 16) | (q[k - 1] >> 16);
    q[0] = (q[0] << 16) | v;
    body++;
    len = 0;
  }
  MHMSW_HD void push(uint32_t o, uint32_t l) {
    if (l == 0) return;
    if (len && o != op) flush();
    op = o;
    len += l;
  }
  MHMSW_HD bool fits() const { return body <= 16 && !long_run; }
  // the CIGAR in order into out[0 .. count)
  MHMSW_HD void write(uint32_t* out) const {
    int32_t at = 0;
    if (clip_head) out[at++] = (clip_head << 4) | kOpS;
#pragma unroll
    for (int m = 0; m < 16; m++)
      if (m < body) out[at + m] = (q[m >> 1] >> (16 * (m & 1))) & 0xFFFFu;
    at += body;
    if (clip_tail) out[at] = (clip_tail << 4) | kOpS;
  }
};

// Traceback over the direction bits of a band laid out with half-width bw
// (diagonal index d = j - i + bw; blocked or row-major layout with DW words
// per row), emitting the final CIGAR backwards.
// shift = the column where the traceback stops: 0 takes '=' / 'X' from the
// equality bits (eq00 = SSW-query[qb] == SSW-ref[rb] for the last M), any
// other value compares the codes at rb + j - shift.  Returns the column where
// the traceback stopped, or -1 if it would leave the band or the rectangle
// (no correct input does; reads stay inside the band's storage either way).
MHMSW_HD inline int32_t cigar_walk(const uint32_t* codes, const CigarGeo& g, int32_t bw, const uint32_t* dirs, int DW,
                                   bool blocked, int32_t shift, bool eq00, CigarEmit& em) {
  NibCursor rcur, qcur;
  em.push(kOpS, (uint32_t)(g.n - 1 - g.qe));
  int32_t i = g.ql() - 1, j = g.rl() - 1;
  int32_t d = j - i + bw;
  int state = 2;
  int64_t cached_at = -1;
  uint32_t cached = 0;
  while (i > 0) {
    if (d < 0 || d > 2 * bw || j < 0) return -1;
    const int64_t at = dir_at(blocked, DW, i, d >> 3);
    if (at != cached_at) {
      cached_at = at;
      cached = dirs[at];
    }
    uint32_t op = step_op((cached >> ((uint32_t)(d & 7) * 4u)) & 0xFu, state);
    if (op == kOpEq || op == kOpX) {
      if (shift != 0) {
        const uint32_t a = qcur.get(codes, g.qbase + g.qb + i);
        const uint32_t b = rcur.get(codes, g.rbase + g.rb + j - shift);
        op = a == b ? kOpEq : kOpX;
      }
      i--;
      j--;
    } else if (op == kOpI) {
      i--;
      d++;
    } else {
      j--;
      d--;
    }
    em.push(op, 1);
  }
  bool last_eq = eq00;
  if (shift != 0) last_eq = nib_at(codes, g.qbase + g.qb) == nib_at(codes, g.rbase + g.rb + j - shift);
  em.push(last_eq ? kOpEq : kOpX, 1);
  em.push(kOpS, (uint32_t)g.qb);
  em.flush();
  return j;
}

// The same traceback over a narrow band (blocked layout, DS words per row)
// with shift 0: each block of 8 rows comes in with 2*DS 16-byte loads, so a
// traceback waits for memory once per 8 rows.
template <int DS, class Emit>
MHMSW_HD int32_t cigar_walk_rows(const CigarGeo& g, int32_t bw, const uint32_t* dirs, bool eq00, Emit& em) {
  constexpr int DW = DS;
  em.push(kOpS, (uint32_t)(g.n - 1 - g.qe));
  int32_t i = g.ql() - 1, j = g.rl() - 1;
  int32_t d = j - i + bw;
  int state = 2;
  uint32_t rows[8 * DW];
  uint32_t cur[DW];
#pragma unroll
  for (int w = 0; w < DW; w++) cur[w] = 0;
  int32_t blk_lo = i + 1, cur_row = -1;
  while (i > 0) {
    if (d < 0 || d > 2 * bw || j < 0) return -1;
    if (i != cur_row) {
      if (i < blk_lo) {
        blk_lo = i & ~7;
        const Words4* src = reinterpret_cast<const Words4*>(dirs + dir_at(true, DW, blk_lo, 0));
#pragma unroll
        for (int k = 0; k < 2 * DW; k++) {
          const Words4 x = src[k];
          rows[4 * k] = x.w[0];
          rows[4 * k + 1] = x.w[1];
          rows[4 * k + 2] = x.w[2];
          rows[4 * k + 3] = x.w[3];
        }
      }
#pragma unroll
      for (int r = 0; r < 8; r++) {
#pragma unroll
        for (int w = 0; w < DW; w++) cur[w] = (i & 7) == r ? rows[r * DW + w] : cur[w];
      }
      cur_row = i;
    }
    uint32_t word = cur[0];
#pragma unroll
    for (int w = 1; w < DW; w++) word = (d >> 3) == w ? cur[w] : word;
    const uint32_t op = step_op((word >> ((uint32_t)(d & 7) * 4u)) & 0xFu, state);
    if (op == kOpEq || op == kOpX) {
      i--;
      j--;
    } else if (op == kOpI) {
      i--;
      d++;
    } else {
      j--;
      d--;
    }
    em.push(op, 1);
  }
  em.push(eq00 ? kOpEq : kOpX, 1);
  em.push(kOpS, (uint32_t)g.qb);
  em.flush();
  return j;
}

// Pairs whose CIGAR needs no band: an empty SSW-query (SSW refuses it: no
// CIGAR), spans beyond the distance filter (no CIGAR, 0 mismatches) and
// score 0.  A score-0 pair has rb = -1, and SSW's CIGAR is one M comparing
// SSW-query[0] with the byte before its SSW-ref buffer -- the top byte of
// glibc's chunk size word, 0 (code A) -- then the clip of the rest.
// Returns true when the pair is one of these; em then holds its CIGAR.
MHMSW_HD inline bool cigar_trivial(const uint32_t* codes, const CigarGeo& g, CigarEmit& em) {
  if (g.n <= 0) return true;
  if (g.re - g.rb > kCigarDistanceFilter || g.qe - g.qb > kCigarDistanceFilter) return true;
  if (g.score > 0) return false;
  em.push(kOpS, (uint32_t)(g.n - 1 - g.qe));
  em.push(nib_at(codes, g.qbase + g.qb) == 0 ? kOpEq : kOpX, 1);
  em.push(kOpS, (uint32_t)g.qb);
  em.flush();
  return true;
}

// Band DP for any half-width, one cell at a time (the serial path for bands
// wider than the wavefront kernel takes): rows of H and E in hrow/erow
// (2*bw+1 entries each), direction dwords row-major with DW =
// ceil((2*bw+1)/8) dwords per row.  Returns H of the last cell.
MHMSW_HD inline int32_t cigar_band_dp_any(const uint32_t* codes, const CigarGeo& g, int32_t bw, int32_t match,
                                          int32_t mismatch, int32_t amb, int32_t gO, int32_t gE, int32_t* hrow,
                                          int32_t* erow, uint32_t* dirs) {
  const int32_t W = 2 * bw + 1, DW = (4 * W + 31) / 32;
  const int32_t rl = g.rl(), ql = g.ql();
  for (int32_t d = 0; d < W; d++) {
    hrow[d] = 0;
    erow[d] = 0;
  }
  for (int32_t i = 0; i < ql; i++) {
    uint32_t tlo, thi;
    const uint32_t qc = nib_at(codes, g.qbase + g.qb + i);
    subst_table(qc, match, mismatch, amb, tlo, thi);
    int32_t hl = 0, fl = 0;
    uint32_t cur = 0;
    for (int32_t d = 0; d < W; d++) {
      // columns right of the rectangle never feed it: not read
      const int32_t j = i - bw + d;
      const uint32_t r = (j < 0 || j >= rl) ? kCigKill : nib_at(codes, g.rbase + g.rb + j);
      const int32_t hu = d + 1 < W ? hrow[d + 1] : 0;
      const int32_t eu = d + 1 < W ? erow[d + 1] : 0;
      const int32_t eo = hu - gO, ee = eu - gE;
      const int32_t e = imax3(eo, ee, 0);
      const int32_t fo = hl - gO, fe = fl - gE;
      const int32_t f = imax3(fo, fe, 0);
      const int32_t t2 = hrow[d] + subst_lookup(tlo, thi, r);
      const int32_t t1 = imax(e, f);
      const int32_t h = imax(t1, t2);
      const uint32_t bits = cell_bits(t1, t2, e, f, r == qc, eo > ee, fo > fe);
      cur |= bits << ((4 * (d & 7)));
      if ((d & 7) == 7 || d == W - 1) {
        dirs[(int64_t)i * DW + (d >> 3)] = cur;
        cur = 0;
      }
      hrow[d] = h;
      erow[d] = e;
      hl = h;
      fl = f;
    }
  }
  return hrow[(rl - 1) - (ql - 1) + bw];
}

}  // namespace mhmsw
