// HIP kernels for batched Smith-Waterman with SSW result semantics, gfx950.
//
// Replaces gpu_bsw::sequence_dna_kernel / sequence_dna_reverse
// (src/adept-sw/kernel.cpp:160-509) and the host-side packing and reverse
// sizing of adept_sw::GPUDriver (src/adept-sw/driver.cpp:281-417).
//
// Recurrence (local Gotoh, SSW's, ssw_core.cpp:321-379):
//   E(i,j) = max(E(i-1,j) - gE, H(i-1,j) - gO)
//   F(i,j) = max(F(i,j-1) - gE, H(i,j-1) - gO)
//   H(i,j) = max(0, H(i-1,j-1) + s(i,j), E(i,j), F(i,j))
// End point: the lexicographic optimum (max H, least SSW-ref index i, least
// SSW-query index j), which is what ssw_core.cpp:464-506 selects.  Begin point:
// the same optimum on the reversed prefixes ending at the end point
// (ssw_core.cpp:1162-1178).
//
// Two forms:
//  * fast (packed 16-bit): a wavefront is split into 64/G groups of G lanes,
//    each group aligns one pair.  Lane l of a group holds 2P consecutive rows
//    as P packed registers (low half: rows 2Pl .. 2Pl+P-1, high half: rows
//    2Pl+P .. 2Pl+2P-1).  The halves are two stages of a systolic array: at
//    step t the low half of lane l works on column t-2l and the high half on
//    column t-2l-1, so every dependency crosses one stage per step and is
//    carried by one DPP row/wave shift (groups narrower than a 16-lane DPP
//    row: a row shift and a v_perm that drops what a group's lane 0 got).
//    Scores are kept multiplied by 16; the low 4 bits of a running per-row
//    maximum hold 15 - (t mod 16), so the maximum also names the first step
//    that reached it.  Every 16 steps the per-row maxima are folded into
//    32-bit (score, first step) records.  Gap values are clamped at 0 with
//    unsigned saturating subtracts (as SSW's are), so H = max(d, E, F) needs
//    no extra max with 0: 12 VALU ops per packed register per step, i.e. 6
//    per cell (10 in the forward kernel, whose H = max(d, E, F) is one
//    v_pk_maximum3_f16 over biased values).  The forward pass runs 2 to 32
//    lanes per pair (fast_shape), the reverse pass 16 or 64.
//  * plain (32-bit): one wavefront per pair, one row per lane, rows in tiles
//    of 64, the last row of a tile passed to the next through LDS.
//
// The packed kernel's instantiations are compiled in parts (sw_fast_*.hip:
// this file with MHMSW_FAST_PART defined, one group size and a list of P
// each), so that they build in parallel; this file alone holds everything
// else and dispatches to them.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>

#include "sw_common.h"
#include "sw_launch.h"

namespace mhmsw {

// ----------------------------------------------------------------------------
// small helpers
// ----------------------------------------------------------------------------
typedef short v2i16 __attribute__((ext_vector_type(2)));
typedef unsigned short v2u16 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t u32(v2i16 x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ v2i16 vi(uint32_t x) { return __builtin_bit_cast(v2i16, x); }
__device__ __forceinline__ v2u16 vu(uint32_t x) { return __builtin_bit_cast(v2u16, x); }
__device__ __forceinline__ uint32_t u32u(v2u16 x) { return __builtin_bit_cast(uint32_t, x); }

__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) { return u32(__builtin_elementwise_max(vi(a), vi(b))); }
__device__ __forceinline__ uint32_t pk_maxu(uint32_t a, uint32_t b) {
  return u32u(__builtin_elementwise_max(vu(a), vu(b)));
}
__device__ __forceinline__ uint32_t pk_adds(uint32_t a, uint32_t b) {
  return u32(__builtin_elementwise_add_sat(vi(a), vi(b)));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) { return u32(vi(a) - vi(b)); }
__device__ __forceinline__ uint32_t pk_addu_sat(uint32_t a, uint32_t b) {
  return u32u(__builtin_elementwise_add_sat(vu(a), vu(b)));
}
// max of three packed u16 values, each below 0x7C00: as f16 bit patterns
// these are finite and non-negative (denormals included, which gfx950 kernels
// keep), and f16 maximum returns one of its inputs
typedef _Float16 v2f16 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_max3_u16f(uint32_t a, uint32_t b, uint32_t c) {
  const v2f16 x = __builtin_bit_cast(v2f16, a), y = __builtin_bit_cast(v2f16, b), z = __builtin_bit_cast(v2f16, c);
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_maximum(__builtin_elementwise_maximum(x, y), z));
}
// unsigned saturating subtract: gap values are kept >= 0 (as SSW's are)
__device__ __forceinline__ uint32_t pk_subu_sat(uint32_t a, uint32_t b) {
  return u32u(__builtin_elementwise_sub_sat(vu(a), vu(b)));
}

// Substitution selector: for 16-bit halves of codes a (rows) and b (column),
// bits 0-1 become a^b and every other bit a|b.  bitop3 truth table with
// src0=0xF0, src1=0xCC, src2=0xAA: src2 ? (src0^src1) : (src0|src1) = 0x7C.
__device__ __forceinline__ uint32_t subst_sel(uint32_t a, uint32_t b) {
  return __builtin_amdgcn_bitop3_b32(a, b, 0x00030003u, 0x7C);
}

// A column outside the pair, or a padding row, compares as N
// (s' = 16*(gO - ambiguity), a finite non-negative value).  Its cells then
// hold values no larger than the real cells they follow from, so they can
// never beat a real cell, and in front of column 0 they stay at the boundary.
constexpr uint32_t kKillColN = (uint32_t)kCodeN;       // selector 4..7: ambiguity
constexpr uint32_t kKillRowN = (uint32_t)kCodeN | 0x0C00u;
// The packed kernels keep H, E and F as unsigned 16-bit multiples of
// 16 below 0x7C00, where the bit patterns of finite non-negative f16 values
// order as the integers do, so H = max(d, E, F) is one v_pk_maximum3_f16.  A
// pair whose score reaches kFastOverflowScore (with the largest substitution
// still below 0x7C00) is rerun exactly in 32-bit.
constexpr uint32_t kRowHi = 0x0C00u;     // high selector byte of a row code: 0x0C -> byte 0x00

__device__ __forceinline__ uint32_t code_at(const uint32_t* __restrict__ codes, int64_t idx) {
  return (codes[idx >> 3] >> ((uint32_t)(idx & 7) * 4u)) & 0xFu;
}

// the 8 nibbles of a dword in reverse order (byte reverse, then swap the
// nibbles of each byte): reverse tasks stream the forward codes backwards
__device__ __forceinline__ uint32_t nib_reverse(uint32_t v) {
  const uint32_t y = __builtin_amdgcn_perm(v, v, 0x00010203u);
  return ((y >> 4) & 0x0F0F0F0Fu) | ((y << 4) & 0xF0F0F0F0u);
}

template <int G>
__device__ __forceinline__ uint32_t shift_in(uint32_t old, uint32_t src) {
  // lane l receives lane l-1 of its group; lane 0 of a group keeps `old`
  if constexpr (G == 16) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, 0x111, 0xF, 0xF, false);  // row_shr:1
  } else {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, 0x138, 0xF, 0xF, false);  // wave_shr:1
  }
}

__device__ __forceinline__ int wave_max_i32(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    int w = __shfl_xor(v, o, 64);
    v = v > w ? v : w;
  }
  return v;
}
__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    int w = __shfl_xor(v, o, 64);
    v = v < w ? v : w;
  }
  return v;
}

// Wave-uniform copies (SGPR) of values every lane holds or that lane 0 holds.
// Loop bounds and work indices must be uniform so that the wavefront never
// splits around the persistent work loop.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// A persistent wavefront's next task number from a class's work counter:
// lane 0 alone adds 1, and the others read its old value.  (Every lane
// calling atomicAdd with lane 0's 1 and the rest's 0 compiles to a scan over
// the 64 lanes' values, hundreds of scalar instructions per fetch.)
__device__ __forceinline__ int32_t next_wave_task(int32_t* counter, int lane) {
  int32_t old = 0;
  if (lane == 0) old = atomicAdd(counter, 1);
  return __builtin_amdgcn_readlane(old, 0);
}
__device__ __forceinline__ int32_t from_lane0(int32_t v) { return __builtin_amdgcn_readlane(v, 0); }
__device__ __forceinline__ int64_t uni64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)v);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int32_t)(uint32_t)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint64_t end_key(int32_t score, int32_t i, int32_t j) {
  return ((uint64_t)(uint32_t)score << 42) | ((uint64_t)(0x1FFFFFu - (uint32_t)i) << 21) |
         (uint64_t)(0x1FFFFFu - (uint32_t)j);
}
__device__ __forceinline__ void key_decode(uint64_t k, int32_t& score, int32_t& i, int32_t& j) {
  score = (int32_t)(k >> 42);
  i = (int32_t)(0x1FFFFFu - (uint32_t)((k >> 21) & 0x1FFFFFu));
  j = (int32_t)(0x1FFFFFu - (uint32_t)(k & 0x1FFFFFu));
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m, int w) {
  uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
  lo = (uint32_t)__shfl_xor((int)lo, m, w);
  hi = (uint32_t)__shfl_xor((int)hi, m, w);
  return ((uint64_t)hi << 32) | lo;
}

#ifndef MHMSW_FAST_PART  // compiled once; the packed kernels' parts skip it

// ----------------------------------------------------------------------------
// unpacking: 2-bit + N mask -> nibble codes
// ----------------------------------------------------------------------------
// Eight 2-bit codes (16 bits) and their N flags (8 bits) -> eight nibble
// codes.  The packer stores N bases as code 0 with their flag set
// (pack2bit.h), so a flag moved to bit 2 of its nibble makes the code 4.
__device__ __forceinline__ uint32_t unpack8(uint32_t two16, uint32_t n8) {
  uint32_t x = two16;  // code k in bits 2k..2k+1 -> bits 4k..4k+1
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  uint32_t n = n8;     // flag k in bit k -> bit 4k
  n = (n | (n << 12)) & 0x000F000Fu;
  n = (n | (n << 6)) & 0x03030303u;
  n = (n | (n << 3)) & 0x11111111u;
  return x | (n << 2);
}

// 32 bases per thread: two dwords of 2-bit codes and one of N flags in, four
// code dwords out (one 16-byte store)
__global__ void unpack_kernel(BatchView b) {
  const int64_t ng = b.total >> 5;  // total is a multiple of 32
  const uint2* __restrict__ two = reinterpret_cast<const uint2*>(b.packed2);
  uint4* __restrict__ out = reinterpret_cast<uint4*>(b.codes);
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < ng; g += (int64_t)gridDim.x * blockDim.x) {
    const uint2 t = two[g];
    const uint32_t nm = b.nmask[g];
    uint4 o;
    o.x = unpack8(t.x & 0xFFFFu, nm & 0xFFu);
    o.y = unpack8(t.x >> 16, (nm >> 8) & 0xFFu);
    o.z = unpack8(t.y & 0xFFFFu, (nm >> 16) & 0xFFu);
    o.w = unpack8(t.y >> 16, nm >> 24);
    out[g] = o;
  }
}

// Contig windows of a windowed batch: output dword w of pair p's window holds
// bases win_off+8w .. +7 of resident contig win_ctg[p], taken from the
// resident 2-bit codes and N bitmap with one funnel shift each.
__global__ void extract_windows_kernel(BatchView b) {
  const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x / 64);
  const int lane = threadIdx.x & 63;
  for (int64_t p = (int64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64; p < b.npairs; p += nwaves) {
    const int64_t s = 2 * p + 1;
    const int64_t base = b.seq_base[s];
    const int32_t len = b.seq_len[s];
    const int64_t src0 = b.res_base[b.win_ctg[p]] + b.win_off[p];
    const int32_t ndw = (len + 7) >> 3;
    for (int32_t w = lane; w < ndw; w += 64) {
      const int64_t sb = src0 + 8 * (int64_t)w;
      const uint32_t two = __builtin_amdgcn_alignbit(b.res2[(sb >> 4) + 1], b.res2[sb >> 4], (uint32_t)(sb & 15) * 2u);
      const uint32_t nm = __builtin_amdgcn_alignbit(b.resN[(sb >> 5) + 1], b.resN[sb >> 5], (uint32_t)(sb & 31));
      uint32_t out = 0;
#pragma unroll
      for (int k = 0; k < 8; k++) {
        if (8 * w + k < len) {
          const uint32_t c = ((nm >> k) & 1u) ? (uint32_t)kCodeN : (two >> (2 * k)) & 3u;
          out |= c << (4 * k);
        }
      }
      b.codes[(base >> 3) + w] = out;
    }
  }
}

// ----------------------------------------------------------------------------
// scheduling: classify -> schedule -> scatter
// ----------------------------------------------------------------------------
// Geometry of pair p in pass `pass`; false when the pair has no task.
// `direct`, when given, receives the reverse pass's answer for a pair that
// needs no task (make_task then returns false); `probe_out`, when given, is set
// for a probe task of pass kPassRev.
__device__ bool make_task(const BatchView& b, int pass, int64_t p, SwTask& t, int& cls, int& bin,
                          PairEnd* direct = nullptr, bool* probe_out = nullptr) {
  const int64_t sref = b.sc.read_is_ref ? 2 * p : 2 * p + 1;  // SSW-ref sequence
  const int64_t sqry = b.sc.read_is_ref ? 2 * p + 1 : 2 * p;  // SSW-query sequence
  const int32_t lref = b.seq_len[sref], lqry = b.seq_len[sqry];
  const PairEnd fe = b.fwd[p];
  int32_t ref_len, qry_len;
  int64_t ref_start, qry_start;
  bool wide = b.mode == 1 || !b.sc.fast_ok;
  if (pass == kPassFwd) {
    if (lref <= 0 || lqry <= 0) return false;
    ref_len = lref;
    qry_len = lqry;
    ref_start = b.seq_base[sref];
    qry_start = b.seq_base[sqry];
  } else if (pass == kPassRerun) {
    if (!(fe.flags & (kEndOverflow | kEndScore2))) return false;
    wide = true;
    ref_len = lref;
    qry_len = lqry;
    ref_start = b.seq_base[sref];
    qry_start = b.seq_base[sqry];
  } else {
    if (fe.score <= 0) return false;
    if (fe.i < 0 || fe.i >= lref || fe.j < 0 || fe.j >= lqry) {
      if (pass == kPassRev) atomicAdd(&b.errors[2], 1);  // an end point outside the pair: never reverse from it
      return false;
    }
    if (fe.flags & kEndWide) wide = true;
    // reversed prefixes [0..i] and [0..j]: R'[k] = R[i-k]
    ref_len = fe.i + 1;
    qry_len = fe.j + 1;
    // Only cells with H' = score matter, and every alignment reaching the
    // score in the reversed rectangle starts at its corner (the forward end
    // is the only cell holding the score there).  An alignment over k+1 rows
    // and l+1 columns scores at most match*min(k+1,l+1) minus the cost of
    // |k-l| gap positions, so |k-l| <= D; rows and columns beyond the other
    // side plus D cannot reach the score and are dropped (exact: DP values
    // inside the kept rectangle do not depend on the dropped cells).  The
    // plain form keeps the reference's full prefix rectangle (ssw_core.cpp:1167-1173).
    // When D is small the band |k-l| <= D itself is all that is computed
    // (sw_band_kernel): cells outside it never carry the score, and values
    // inside it computed without them are lower bounds that are exact on
    // every path reaching the score.
    int64_t D = -1;
    if (b.mode != 1) {
      const int32_t L = ref_len < qry_len ? ref_len : qry_len;
      const int64_t slack = (int64_t)b.sc.match * L - fe.score;
      D = 0;
      if (slack >= b.sc.gap_open) D = 1 + (slack - b.sc.gap_open) / b.sc.gap_ext;
      const int64_t rcap = (int64_t)qry_len + D, qcap = (int64_t)ref_len + D;
      if (ref_len > rcap) ref_len = (int32_t)rcap;
      if (qry_len > qcap) qry_len = (int32_t)qcap;
    }
    // reverse tasks read the forward codes backwards: R'[k] = codes[ref_start - k]
    ref_start = b.seq_base[sref] + fe.i;
    qry_start = b.seq_base[sqry] + fe.j;
    bool band = !wide && D >= 0 && D <= b.band_max_d && ref_len + D < 0xFFF0;
    // A large job probes a wide-band pair whose score is at most
    // match * kProbeRows on the first kProbeRows rows of its rectangle.  The
    // values there are the full rectangle's (a cell depends only on cells of
    // lower rows and columns).  A cell (k, l) holding the score with
    // k < kProbeRows ends an alignment over at most kProbeRows rows, which
    // scores at most match * kProbeRows minus the cost of |l - k| gap
    // positions: so |l - k| <= Dp (D's bound over kProbeRows rows), and the
    // probe is the band |l - k| <= Dp of those rows.  When it reaches the
    // score, its least (k, l) is the rectangle's; pass kPassRev2 runs the full
    // rectangle of the others.  (An unrelated pair's short local alignment
    // ends within the probe; a gapped real one scores too high to be probed.)
    const bool probe = b.rev_probe && !wide && !band && D >= 0 && ref_len > kProbeRows &&
                       (int64_t)b.sc.match * kProbeRows >= fe.score;
    if (pass == kPassRev && probe && probe_out) *probe_out = true;  // a probe kPassRev2 checks
    if (pass == kPassRev2) {
      if (!probe || b.rev[p].score == fe.score) return false;
      // the few probes that miss: every one in the largest reverse class
    } else if (probe) {
      const int64_t slack = (int64_t)b.sc.match * kProbeRows - fe.score;
      D = slack >= b.sc.gap_open ? 1 + (slack - b.sc.gap_open) / b.sc.gap_ext : 0;
      ref_len = kProbeRows;
      if ((int64_t)qry_len > (int64_t)kProbeRows + D) qry_len = (int32_t)(kProbeRows + D);
      band = D <= b.band_max_d;
    }
    // A single-diagonal pair (D = 0: no gap reaches the score) whose score
    // leaves no room for a mismatch or an N either is score/match matches in
    // a row: each non-match costs at least c = match + min(mismatch,
    // ambiguity) against a match, and an alignment of b non-matches within
    // the rectangle's L = min(rows, columns) positions scores at most
    // match*L - b*c.  Along its diagonal H' = (k+1)*match, so the least cell
    // holding the score is k = l = score/match - 1, with no codes to read.
    if (pass == kPassRev && band && D == 0 && direct && b.sc.match > 0) {
      const int32_t L = ref_len < qry_len ? ref_len : qry_len;
      const int64_t c = (int64_t)b.sc.match + (b.sc.mismatch < b.sc.ambiguity ? b.sc.mismatch : b.sc.ambiguity);
      const int32_t a = fe.score / b.sc.match;
      if ((int64_t)fe.score + c > (int64_t)b.sc.match * L && fe.score % b.sc.match == 0 && a <= L) {
        *direct = PairEnd{fe.score, a - 1, a - 1, 0};
        return false;
      }
    }
    if (pass == kPassRev && band) {
      // rows = reversed SSW-ref, columns = reversed SSW-query, one lane per pair
      t.pair = (int32_t)p;
      // the forward score rides in the flags (below 2^18: packed pairs stay
      // under kFastOverflowScore), so the band kernels need no load of it
      t.flags = kTaskReverse | kTaskRowsAreRef | ((int32_t)D << kTaskBandShift) | (fe.score << kTaskScoreShift);
      t.row_start = ref_start;
      t.row_len = ref_len;
      t.col_start = qry_start;
      t.col_len = qry_len;
      // 4P diagonals cover |l - k| <= 2P - 1; D = 0 is the single diagonal
      cls = D == 0 ? kClassDiag : kClassBand + (D <= 1 ? 0 : D <= 3 ? 1 : D <= 7 ? 2 : D <= 15 ? 3 : 4);
      const int bn = (ref_len + (int32_t)D) >> b.bin_shift;  // the lane's loop runs K + D steps
      bin = bn < kNumBins ? bn : kNumBins - 1;
      return true;
    }
  }
  t.pair = (int32_t)p;
  t.flags = pass >= kPassRev ? kTaskReverse : 0;
  if (b.score2 && pass < kPassRev) {
    // score2 reads per-SSW-ref-position maxima: rows are SSW-ref, columns
    // SSW-query plus up to kMaxQueryPad zero-scoring columns
    if (!wide && ref_len <= kFast16MaxRows && ref_len < b.g64_min_rows && qry_len + kMaxQueryPad <= kFastMaxCols) {
      cls = kClassFast16 + fast16_class_of(ref_len);
    } else if (!wide && ref_len <= kFast64MaxRows && qry_len + kMaxQueryPad <= kFastMaxCols) {
      cls = kClassFast64 + (ref_len + 127) / 128 - 1;
    } else {
      cls = kClassPlain;
    }
    t.flags |= kTaskRowsAreRef;
    t.row_start = ref_start;
    t.row_len = ref_len;
    t.col_start = qry_start;
    t.col_len = qry_len;
    const int bn = t.col_len >> b.bin_shift;
    bin = bn < kNumBins ? bn : kNumBins - 1;
    return true;
  }
  const int32_t shorter = ref_len <= qry_len ? ref_len : qry_len;
  const int32_t longer = ref_len <= qry_len ? qry_len : ref_len;
  const bool ref_rows_if_short = ref_len <= qry_len;
  if (!wide && shorter <= kFast16MaxRows && shorter < b.g64_min_rows && longer <= kFastMaxCols) {
    cls = kClassFast16 + fast16_class_of(shorter);
  } else if (!wide && shorter <= kFast64MaxRows && longer <= kFastMaxCols) {
    const int P = (shorter + 127) / 128;
    cls = kClassFast64 + P - 1;
  } else {
    cls = kClassPlain;
  }
  if (pass >= kPassRev) {  // a small job's reverse rectangles, and pass kPassRev2's: one class per group size
    const int one16 = pass == kPassRev2 ? b.rev2_one16 : b.rev_one16;
    const int one64 = pass == kPassRev2 ? b.rev2_one64 : b.rev_one64;
    if (one16 > 0 && cls >= kClassFast16 && cls < kClassFast16 + kMaxP) cls = kClassFast16 + one16 - 1;
    if (one64 > 0 && cls >= kClassFast64 && cls < kClassFast64 + kMaxP) cls = kClassFast64 + one64 - 1;
  }
  bool rows_ref;
  if (cls == kClassPlain) {
    rows_ref = !ref_rows_if_short;  // rows = the longer sequence, columns (LDS row) = the shorter
    if (ref_len == qry_len) rows_ref = true;
  } else {
    rows_ref = ref_rows_if_short;  // rows = the shorter sequence
  }
  if (rows_ref) {
    t.flags |= kTaskRowsAreRef;
    t.row_start = ref_start;
    t.row_len = ref_len;
    t.col_start = qry_start;
    t.col_len = qry_len;
  } else {
    t.row_start = qry_start;
    t.row_len = qry_len;
    t.col_start = ref_start;
    t.col_len = ref_len;
  }
  int bn = t.col_len >> b.bin_shift;
  bin = bn < kNumBins ? bn : kNumBins - 1;
  return true;
}

// At the start of a job: error words, counters (reverse cells, a probe flag) and the
// scheduling state of its passes (b.sched is the first of kNumPasses).
__global__ void clear_job_kernel(BatchView b) {
  if (blockIdx.x == 0 && threadIdx.x < 4) b.errors[threadIdx.x] = 0;
  if (blockIdx.x == 0 && threadIdx.x < 2) b.counters[threadIdx.x] = 0;
  int32_t* w = reinterpret_cast<int32_t*>(b.sched);
  const int n = (int)(kNumPasses * sizeof(PassSched) / sizeof(int32_t));
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) w[i] = 0;
}

// Each 256-thread block classifies a contiguous run of kClassifyPerThread*256
// pairs: ranks inside the block come from LDS atomics, then one global
// atomic per used bucket reserves the block's range (far less contention on
// the few buckets most pairs fall into).
constexpr int kClassifyPerThread = 8;

__global__ __launch_bounds__(256) void classify_kernel(BatchView b, int pass) {
  // pass kPassRev2 checks the probes of pass kPassRev: with none issued it has
  // no task, and its scatter_kernel reads none of the buckets left unwritten
  if (pass == kPassRev2 && b.counters[1] == 0) return;
  __shared__ int32_t hist[kNumBuckets];
  __shared__ unsigned long long blk_cells;  // reverse cells of the block's pairs: one global atomic per block
  __shared__ int blk_probe;                 // a probe among the block's pairs: one global store per block
  for (int i = threadIdx.x; i < kNumBuckets; i += blockDim.x) hist[i] = 0;
  if (threadIdx.x == 0) {
    blk_cells = 0;
    blk_probe = 0;
  }
  __syncthreads();
  unsigned long long my_cells = 0;
  const int64_t p0 = (int64_t)blockIdx.x * blockDim.x * kClassifyPerThread;
  int32_t bk[kClassifyPerThread], lr[kClassifyPerThread];
#pragma unroll
  for (int k = 0; k < kClassifyPerThread; k++) {
    const int64_t p = p0 + (int64_t)k * blockDim.x + threadIdx.x;
    bk[k] = -1;
    lr[k] = 0;
    if (p >= b.npairs) continue;
    SwTask t;
    int cls = 0, bin = 0;
    PairEnd direct{0, 0, 0, 0};
    bool probe = false;
    const bool task = make_task(b, pass, p, t, cls, bin, &direct, &probe);
    if (probe) blk_probe = 1;
    if (!task) {
      b.bucket[p] = -1;
      if (pass == kPassFwd) {
        b.fwd[p] = PairEnd{0, -1, 0, 0};
        if (b.score2) b.s2[p] = 0;
      }
      if (pass == kPassRev) b.rev[p] = direct;
      continue;
    }
    bk[k] = cls * kNumBins + bin;
    lr[k] = atomicAdd(&hist[bk[k]], 1);
    if (pass == kPassRerun) atomicAdd(&b.errors[1], 1);
    if (pass >= kPassRev) {
      // cells computed: the whole rectangle, or for a band task its 2D+1 diagonals
      int64_t w = t.col_len;
      if (cls >= kClassBand) {
        const int64_t bw = 2 * (int64_t)((t.flags >> kTaskBandShift) & 0x1F) + 1;
        w = w < bw ? w : bw;
      }
      my_cells += (unsigned long long)t.row_len * (unsigned long long)w;
    }
  }
  if (my_cells) atomicAdd(&blk_cells, my_cells);
  __syncthreads();
  if (threadIdx.x == 0 && blk_cells) atomicAdd(&b.counters[0], blk_cells);
  if (threadIdx.x == 0 && blk_probe) b.counters[1] = 1;
  for (int i = threadIdx.x; i < kNumBuckets; i += blockDim.x) {
    const int32_t n = hist[i];
    hist[i] = n ? atomicAdd(&b.sched->bucket_count[i], n) : 0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kClassifyPerThread; k++) {
    const int64_t p = p0 + (int64_t)k * blockDim.x + threadIdx.x;
    if (p >= b.npairs || bk[k] < 0) continue;
    b.bucket[p] = bk[k];
    b.rank[p] = hist[bk[k]] + lr[k];
  }
}

// single block: class totals, slot bases (bins in descending order inside a
// class, so the longest columns are taken first), padding slots
__global__ void schedule_kernel(BatchView b) {
  PassSched* s = b.sched;
  const int c = threadIdx.x;
  __shared__ int32_t slots[kNumClasses];
  if (c < kNumClasses) {
    int32_t tot = 0;
    for (int bn = kNumBins - 1; bn >= 0; bn--) tot += s->bucket_count[c * kNumBins + bn];
    s->class_count[c] = tot;
    int32_t padded = tot;
    if (c < kClassPlain) padded = (tot + 63) & ~63;       // up to 64 groups per wavefront
    if (c >= kClassBand) padded = (tot + 63) & ~63;       // one pair per lane
    if (c == kClassBand2) padded = (tot + 15) & ~15;      // four lanes per pair
    slots[c] = padded;
    s->class_slots[c] = padded;
    s->class_next[c] = 0;
  }
  __syncthreads();
  if (c < kNumClasses) {
    int32_t base = 0;
    for (int k = 0; k < c; k++) base += slots[k];
    s->class_base[c] = base;
    int32_t run = base;
    for (int bn = kNumBins - 1; bn >= 0; bn--) {
      s->bucket_base[c * kNumBins + bn] = run;
      run += s->bucket_count[c * kNumBins + bn];
    }
    for (int32_t k = run; k < base + slots[c]; k++) {
      SwTask e;
      e.pair = -1;
      e.flags = 0;
      e.row_start = kPadBases;
      e.col_start = kPadBases;
      e.row_len = 0;
      e.col_len = 0;
      b.tasks[k] = e;
    }
  }
}

__global__ void scatter_kernel(BatchView b, int pass) {
  if (pass == kPassRev2 && b.counters[1] == 0) return;  // no probes: no tasks (classify_kernel)
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < b.npairs; p += (int64_t)gridDim.x * blockDim.x) {
    const int bk = b.bucket[p];
    if (bk < 0) continue;
    SwTask t;
    int cls = 0, bin = 0;
    make_task(b, pass, p, t, cls, bin);
    b.tasks[b.sched->bucket_base[bk] + b.rank[p]] = t;
  }
}

#endif  // MHMSW_FAST_PART

// ----------------------------------------------------------------------------
// fast packed 16-bit kernel
// ----------------------------------------------------------------------------
template <int G, int P>
struct FastState {
  uint32_t rb[P];      // row codes (selector form), packed
  uint32_t ho[P];      // 16*H(r, c_last), packed
  uint32_t hg[P];      // 16*(F(r, c_last + 1) + gO), packed (gap along the columns), clamped >= 0
  uint32_t best[P];    // per-row running max of 16*H + tag, packed u16
  uint32_t rec_h[P];   // per-row record: 16*H | 15 of the row's maximum so far, packed u16
  uint32_t rec_s[P];   // and the first step at which it was reached, packed u16
  uint32_t xa, xb;     // exchanged 16*H of the row above, 2 steps / 1 step ago
  uint32_t vg;         // exchanged 16*(E + gO) of the first row, clamped >= 0
  uint32_t xs, vs;     // DPP shift results; lane 0 of a group never receives and keeps its boundary (G >= 16)
  uint32_t xsel;       // G < 16: v_perm selector joining the received and own halves (see fast_cells)
  uint32_t q;          // column codes (low: this step, high: last step)
  uint32_t w;          // next 8 column codes
  uint32_t dcur, dnext;  // code dwords feeding the next refill
};

// Substitution forms of a step's packed cells.
//   kSubCodes: rb[p] and q hold codes; sp = perm(thi, tlo, subst_sel(rb[p], q))
//              over the scoring table (2 ops per packed cell).
//   kSubPad:   kSubCodes, then score2's padding columns (edge chunks of
//              score2 tasks): the npad columns after the last one score 0
//              against every row, as SSW's query padding does, so their cells
//              get s' = 16*(0 + gO) in place of the kill value.
//   kSubTab:   pairs without N.  The step turns its column's code into that
//              column's table of s' against A, C, G, T: one v_alignbyte over
//              the scoring window wh:wl (bytes x x x m x x x x), which the
//              code's low two bits shift, so byte 3 - r is m when r is the
//              column's base and x otherwise.  rb[p] holds byte selectors
//              (low half: 3 - r over this column's table; high half: 7 - r
//              over the last column's table, kept in q; rows past the end:
//              0x0C, which reads 0) and sp = perm(q, tc, rb[p]) (1 op).
constexpr int kSubCodes = 0, kSubPad = 1, kSubTab = 2;

// The packed cells of one step (substitutions from t1:t0 as above) and the
// exchange with the neighbouring lanes.
template <int G, int P, int S, int SUB, int T0>
__device__ __forceinline__ void fast_cells(FastState<G, P>& st, uint32_t t1, uint32_t t0, uint32_t keep, uint32_t pm,
                                           uint32_t go, uint32_t ge) {
  constexpr uint32_t tag = (uint32_t)(T0 - S) * 0x00010001u;
  uint32_t diag = st.xa;
  uint32_t vg = st.vg;
#pragma unroll
  for (int p = 0; p < P; p++) {
    uint32_t sp;
    if constexpr (SUB == kSubTab) {
      sp = __builtin_amdgcn_perm(t1, t0, st.rb[p]);
    } else {
      sp = __builtin_amdgcn_perm(t1, t0, subst_sel(st.rb[p], st.q));
    }
    // score2 reads every row's maximum over the real and padding columns, so
    // columns past the padding get s' = 0 (d = Hdiag - gO): below the row's
    // own cells, where N-like values (d = Hdiag - ambiguity) might not be
    if constexpr (SUB == kSubPad) sp = (sp & keep) | (go & pm);
    // d, E and F carry a bias of 16*gO: d = 16*(Hdiag + s + gO) (sp holds
    // 16*(s + gO) >= 0), E' = 16*(E + gO) >= 0, F' likewise.  Removing the bias
    // with an unsigned saturating subtract is H = max(0, d, E, F).
    const uint32_t d = pk_addu_sat(diag, sp);
    const uint32_t hb = pk_max3_u16f(d, st.hg[p], vg);
    const uint32_t h = pk_subu_sat(hb, go);
    diag = st.ho[p];
    st.ho[p] = h;
    // E' and F' update: max(E' - gE, H) is 16*(max(E - gE, H - gO) + gO)
    st.hg[p] = pk_max(pk_subu_sat(st.hg[p], ge), h);
    vg = pk_max(pk_subu_sat(vg, ge), h);
    // The row maxima take two steps at a time: at an odd step, diag still
    // holds the register's H of the step before (tag one higher), so one
    // three-way max covers both (16*H + tag stays below 0x7C00).
    if constexpr ((S & 1) != 0) {
      constexpr uint32_t tag_prev = (uint32_t)(T0 + 1 - S) * 0x00010001u;
      st.best[p] = pk_max3_u16f(st.best[p], h | tag, diag | tag_prev);
    }
  }
  // The next step's values above the lane's two first rows: the low half
  // (row 2Pl - 1) is the high half of the lane below's last register, the high
  // half (row 2Pl + P - 1) the low half of its own.  Lane 0 of a group reads
  // the boundary (0) instead.
  const uint32_t last = st.ho[P - 1];
  uint32_t xnew, vnew;
  if constexpr (G < 16 || G == 32) {
    // row_shr:1 (wave_shr:1 for 32-lane groups) crosses group edges; the perm
    // selector of a group's lane 0 writes zeros in place of what it received
    constexpr int shr = G == 32 ? 0x138 : 0x111;
    const uint32_t xs = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)last, shr, 0xF, 0xF, true);
    const uint32_t vs = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)vg, shr, 0xF, 0xF, true);
    xnew = __builtin_amdgcn_perm(last, xs, st.xsel);
    vnew = __builtin_amdgcn_perm(vg, vs, st.xsel);
  } else {
    st.xs = shift_in<G>(st.xs, last);
    st.vs = shift_in<G>(st.vs, vg);
    xnew = __builtin_amdgcn_alignbit(last, st.xs, 16);
    vnew = __builtin_amdgcn_alignbit(vg, st.vs, 16);
  }
  st.xa = st.xb;
  st.xb = xnew;
  st.vg = vnew;
}

// Scoring constants of a launch: the substitution table (tlo, thi), the
// kSubTab window (wl, wh) and the all-ambiguity column table (amb4).
struct StepConst {
  uint32_t tlo, thi, go, ge, wl, wh, amb4;
};

template <int G, int P, bool EDGE, int SUB, int S, int T0>
__device__ __forceinline__ void fast_step(FastState<G, P>& st, int t, int l, int col_len, int npad,
                                          const StepConst& k) {
  if constexpr (SUB == kSubTab) {
    uint32_t tc = __builtin_amdgcn_alignbyte(k.wh, k.wl, st.w);
    st.w >>= 4;
    if constexpr (EDGE) {
      const bool kill = (uint32_t)(t - 2 * l) >= (uint32_t)col_len;
      tc = kill ? k.amb4 : tc;
    }
    fast_cells<G, P, S, kSubTab, T0>(st, st.q, tc, 0, 0, k.go, k.ge);
    st.q = tc;
  } else {
    uint32_t code = st.w & 0xFu;
    st.w >>= 4;
    if constexpr (EDGE) {
      const bool kill = (uint32_t)(t - 2 * l) >= (uint32_t)col_len;
      code = kill ? kKillColN : code;
    }
    uint32_t pm = 0;    // halves whose column is a padding column
    uint32_t keep = 0;  // halves whose column is a real column
    if constexpr (SUB == kSubPad) {
      const uint32_t clo = (uint32_t)(t - 2 * l - col_len);
      pm = (clo < (uint32_t)npad ? 0x0000FFFFu : 0u) | (clo - 1u < (uint32_t)npad ? 0xFFFF0000u : 0u);
      const uint32_t rlo = (uint32_t)(t - 2 * l);
      keep = (rlo < (uint32_t)col_len ? 0x0000FFFFu : 0u) | (rlo - 1u < (uint32_t)col_len ? 0xFFFF0000u : 0u);
    }
    st.q = (st.q << 16) | code;
    fast_cells<G, P, S, SUB, T0>(st, k.thi, k.tlo, keep, pm, k.go, k.ge);
  }
}

template <int G, int P, bool EDGE, int SUB, bool REV, int TB, int T0>
__device__ __forceinline__ void fast_chunk(FastState<G, P>& st, int t0, int l, int col_len, int npad,
                                           const uint32_t* __restrict__ codes, int64_t& dw, uint32_t sh,
                                           const StepConst& k, bool fold, bool half = false) {
  // Forward: w takes the 8 codes from dcur:dnext (loaded one refill
  // earlier); the dword after them is requested now and used at the next
  // refill.  Reverse (REV): the window of the next 8 codes lies one dword
  // lower each time; dcur:dnext hold it (dcur the lower dword), its nibbles
  // reversed give the codes in column order, and the dword below is
  // requested now.
#define MHMSW_REFILL()                                                        \
  {                                                                           \
    if constexpr (REV) {                                                      \
      st.w = nib_reverse(__builtin_amdgcn_alignbit(st.dnext, st.dcur, sh));   \
      st.dnext = st.dcur;                                                     \
      dw--;                                                                   \
      st.dcur = codes[dw];                                                    \
    } else {                                                                  \
      st.w = __builtin_amdgcn_alignbit(st.dnext, st.dcur, sh);                \
      st.dcur = st.dnext;                                                     \
      dw++;                                                                   \
      st.dnext = codes[dw];                                                   \
    }                                                                         \
  }
  MHMSW_REFILL();
  fast_step<G, P, EDGE, SUB, 0, T0>(st, t0 + 0, l, col_len, npad, k);
  fast_step<G, P, EDGE, SUB, 1, T0>(st, t0 + 1, l, col_len, npad, k);
  fast_step<G, P, EDGE, SUB, 2, T0>(st, t0 + 2, l, col_len, npad, k);
  fast_step<G, P, EDGE, SUB, 3, T0>(st, t0 + 3, l, col_len, npad, k);
  fast_step<G, P, EDGE, SUB, 4, T0>(st, t0 + 4, l, col_len, npad, k);
  fast_step<G, P, EDGE, SUB, 5, T0>(st, t0 + 5, l, col_len, npad, k);
  fast_step<G, P, EDGE, SUB, 6, T0>(st, t0 + 6, l, col_len, npad, k);
  fast_step<G, P, EDGE, SUB, 7, T0>(st, t0 + 7, l, col_len, npad, k);
  // a pair's last chunk (an edge chunk) may end after 8 steps (half; uniform)
  if (!EDGE || !half) {
    MHMSW_REFILL();
    fast_step<G, P, EDGE, SUB, 8, T0>(st, t0 + 8, l, col_len, npad, k);
    fast_step<G, P, EDGE, SUB, 9, T0>(st, t0 + 9, l, col_len, npad, k);
    fast_step<G, P, EDGE, SUB, 10, T0>(st, t0 + 10, l, col_len, npad, k);
    fast_step<G, P, EDGE, SUB, 11, T0>(st, t0 + 11, l, col_len, npad, k);
    fast_step<G, P, EDGE, SUB, 12, T0>(st, t0 + 12, l, col_len, npad, k);
    fast_step<G, P, EDGE, SUB, 13, T0>(st, t0 + 13, l, col_len, npad, k);
    fast_step<G, P, EDGE, SUB, 14, T0>(st, t0 + 14, l, col_len, npad, k);
    fast_step<G, P, EDGE, SUB, 15, T0>(st, t0 + 15, l, col_len, npad, k);
  }
#undef MHMSW_REFILL
  // Fold the per-row maxima into the records (at the end of a tag period, or
  // of the pair): a row whose maximum H' + tag (H' = H << TB) has a higher H
  // than its record takes the step of the tag, the period's first step plus
  // (2^TB - 1 - tag), i.e. its first step | (tag ^ (2^TB - 1)).  best is never
  // cleared, so it never falls: a maximum carried over from an earlier period
  // is the record itself, and the new record is best with its tag bits set.
  if (fold) {
    constexpr uint32_t low = (uint32_t)((1 << TB) - 1) * 0x00010001u;
    const uint32_t cstep = (uint32_t)(t0 & -(1 << TB)) * 0x00010001u | low;
#pragma unroll
    for (int p = 0; p < P; p++) {
      const uint32_t m = st.best[p];
      // 0xFFFF in the halves where m > rec_h = H'rec | low, i.e. where m's H
      // is above Hrec (values below 0x7C00: the signed difference is exact)
      uint32_t beat = u32(vi(pk_sub(st.rec_h[p], m)) >> (short)15);
      // (opaque to the compiler, which would otherwise turn the bit select
      // below into a compare and a select per half)
      __asm__("" : "+v"(beat));
      const uint32_t ms = (m & low) ^ cstep;
      st.rec_s[p] = (ms & beat) | (st.rec_s[p] & ~beat);
      st.rec_h[p] = m | low;
    }
  }
}

// A pair's chunks: edge chunks (column checks, substitution EDGE_SUB) while
// some lane may be left of its column 0 (t0 < 2G), middle chunks while every
// lane is inside its pair (MID_SUB), edge chunks to the end.  Three loops, so
// that no chunk chooses its form at run time.
//   TB = 4: H is kept as 16 H; each chunk's 16 steps carry tags 15..0 and the
//   chunk folds.
//   TB = 5 (G <= 8, so 2G <= 16): 32 H; a tag period is two chunks (tags
//   31..16, then 15..0) and folds after the second, or after the first when
//   that is the pair's last.  The first period is edge chunks, so that the
//   middle loop runs whole periods.
template <int G, int P, int EDGE_SUB, int MID_SUB, bool REV, int TB>
__device__ __forceinline__ void fast_run(FastState<G, P>& st, int nsteps, int mincols, int l, int col_len, int npad,
                                         const uint32_t* __restrict__ codes, int64_t& dw, uint32_t sh,
                                         const StepConst& k) {
  int t0 = 0;
  if constexpr (TB == 4) {
    const int head = nsteps < 2 * G ? nsteps : 2 * G;
    for (; t0 < head; t0 += 16)
      fast_chunk<G, P, true, EDGE_SUB, REV, 4, 15>(st, t0, l, col_len, npad, codes, dw, sh, k, true, t0 + 8 >= nsteps);
    for (; t0 < nsteps && t0 + 16 <= mincols; t0 += 16)
      fast_chunk<G, P, false, MID_SUB, REV, 4, 15>(st, t0, l, col_len, npad, codes, dw, sh, k, true);
    for (; t0 < nsteps; t0 += 16)
      fast_chunk<G, P, true, EDGE_SUB, REV, 4, 15>(st, t0, l, col_len, npad, codes, dw, sh, k, true, t0 + 8 >= nsteps);
  } else {
    static_assert(2 * G <= 16, "the first chunk holds every lane's column 0");
    for (; t0 < nsteps && t0 < 32; t0 += 32) {
      fast_chunk<G, P, true, EDGE_SUB, REV, 5, 31>(st, t0, l, col_len, npad, codes, dw, sh, k, t0 + 16 >= nsteps,
                                                   t0 + 8 >= nsteps);
      if (t0 + 16 < nsteps)
        fast_chunk<G, P, true, EDGE_SUB, REV, 5, 15>(st, t0 + 16, l, col_len, npad, codes, dw, sh, k, true,
                                                     t0 + 24 >= nsteps);
    }
    for (; t0 + 32 <= nsteps && t0 + 32 <= mincols; t0 += 32) {
      fast_chunk<G, P, false, MID_SUB, REV, 5, 31>(st, t0, l, col_len, npad, codes, dw, sh, k, false);
      fast_chunk<G, P, false, MID_SUB, REV, 5, 15>(st, t0 + 16, l, col_len, npad, codes, dw, sh, k, true);
    }
    for (; t0 < nsteps; t0 += 32) {
      fast_chunk<G, P, true, EDGE_SUB, REV, 5, 31>(st, t0, l, col_len, npad, codes, dw, sh, k, t0 + 16 >= nsteps,
                                                   t0 + 8 >= nsteps);
      if (t0 + 16 < nsteps)
        fast_chunk<G, P, true, EDGE_SUB, REV, 5, 15>(st, t0 + 16, l, col_len, npad, codes, dw, sh, k, true,
                                                     t0 + 24 >= nsteps);
    }
  }
}

// One task of every group of the wavefront.  TB: H is kept as H << TB (4,
// or 5 for forward launches whose scoring and rows allow it: see the kernel).
template <int G, int P, int TB>
__device__ __forceinline__ void fast_pair(const BatchView& b, const SwTask& task, int l) {
  const bool active = task.pair >= 0;
  const int32_t row_len = active ? task.row_len : 0;
  const int32_t col_len = active ? task.col_len : 0;
  // score2 (forward tasks): SSW's 8-bit-pass query padding after the last column
  const bool s2task = b.score2 && !(task.flags & kTaskReverse);
  const int npad = (active && s2task) ? ssw_query_pad(col_len, 0, 0) : 0;
  // per-group values are uniform within a group: combine the groups' lane 0
  int nsteps_need = 0, mincols = 0x7FFFFFFF;
  {
    const int need = active ? col_len + npad + 2 * G - 1 : 0;
    const int cmin = active ? col_len : 0x7FFFFFFF;
#pragma unroll
    for (int g = 0; g < 64 / G; g++) {
      const int a = __builtin_amdgcn_readlane(need, g * G);
      const int c = __builtin_amdgcn_readlane(cmin, g * G);
      nsteps_need = a > nsteps_need ? a : nsteps_need;
      mincols = c < mincols ? c : mincols;
    }
  }
  if (nsteps_need <= 0) return;
  // whole 8-step halves of a chunk (the column stream refills 8 codes at a
  // time); every chunk but a pair's last runs 16 steps
  const int nsteps = (nsteps_need + 7) & ~7;

  const uint32_t* __restrict__ codes = b.codes;
  StepConst k;
  // the host's table holds 16*(s + gO) per byte; at TB = 5 every byte is
  // below 0x80 (match + gO <= 7), so doubling the words doubles each byte
  k.tlo = b.sc.perm_lo << (TB - 4);
  k.thi = b.sc.perm_hi << (TB - 4);
  k.go = (uint32_t)(b.sc.gap_open << TB) * 0x00010001u;
  k.ge = (uint32_t)(b.sc.gap_ext << TB) * 0x00010001u;
  {
    // table bytes: 0 match, 1-3 mismatch, 4-7 ambiguity (all s' = (s + gO) << TB)
    const uint32_t m = k.tlo & 0xFFu, x = (k.tlo >> 8) & 0xFFu, a = k.thi & 0xFFu;
    k.wl = x * 0x00010101u | (m << 24);
    k.wh = x * 0x01010101u;
    k.amb4 = a * 0x01010101u;
  }

  // uniform: a launch holds the tasks of one pass (padding slots carry no
  // flags); reverse tasks read the forward codes backwards.  (TB = 5 launches
  // are forward launches.)
  const bool rev = TB == 4 && __ballot(task.flags & kTaskReverse) != 0;
  FastState<G, P> st;
  // this lane's 2P row codes are consecutive nibbles (descending for a
  // reverse task): NRW + 1 dword loads, then nibble m of the rw words is the
  // row code m (forward) or 2P-1-m (reverse)
  constexpr int NRW = (2 * P + 7) / 8;
  const int64_t r0 = rev ? task.row_start - 2 * P * l - (2 * P - 1) : task.row_start + 2 * P * l;
  const uint32_t rsh = (uint32_t)(r0 & 7) * 4u;
  uint32_t rw[NRW];
  {
    uint32_t ra[NRW + 1];
#pragma unroll
    for (int i = 0; i <= NRW; i++) ra[i] = codes[(r0 >> 3) + i];
#pragma unroll
    for (int i = 0; i < NRW; i++) rw[i] = __builtin_amdgcn_alignbit(ra[i + 1], ra[i], rsh);
  }
  bool has_n = false;  // an N among this lane's rows, or its group's columns
  int32_t row0 = 2 * P * l;  // (not hoisted out of the task loop, as below)
  __asm__ volatile("" : "+v"(row0));
#pragma unroll
  for (int p = 0; p < P; p++) {
    const int rlo = row0 + p, rhi = row0 + P + p;
    const int mlo = rev ? 2 * P - 1 - p : p, mhi = rev ? P - 1 - p : P + p;
    const uint32_t nlo = (rw[mlo >> 3] >> (4 * (mlo & 7))) & 0xFu;
    const uint32_t nhi = (rw[mhi >> 3] >> (4 * (mhi & 7))) & 0xFu;
    const uint32_t clo = rlo < row_len ? (nlo | kRowHi) : kKillRowN;
    const uint32_t chi = rhi < row_len ? (nhi | kRowHi) : kKillRowN;
    has_n = has_n || (rlo < row_len && nlo == (uint32_t)kCodeN) || (rhi < row_len && nhi == (uint32_t)kCodeN);
    st.rb[p] = clo | (chi << 16);
    st.ho[p] = 0;
    st.hg[p] = 0;
    st.best[p] = 0;
    st.rec_h[p] = (uint32_t)((1 << TB) - 1) * 0x00010001u;  // H = 0
    st.rec_s[p] = 0;
  }
  st.xa = 0;  // boundary row: H = 0
  st.xb = 0;
  st.vg = 0;  // boundary E' (E = -gO: never above 0, so never part of an H)
  st.xs = 0;
  st.vs = 0;
  st.xsel = l == 0 ? 0x05040C0Cu : 0x05040302u;  // (own low half << 16) | received high half
  st.q = (kKillColN << 16) | kKillColN;
  // column stream: at step t the low half reads column t - 2l, i.e. code
  // index cb + t (forward) or cb - t (reverse).  Forward: dw indexes
  // st.dnext.  Reverse: the first refill's window starts at cb - 7 and dw
  // indexes st.dcur, its lower dword.  Either way the stream stays within a
  // few dwords of the sequence plus its padding (>= kPadBases - 2G - 24 >= 0).
  const int64_t cb = rev ? task.col_start + 2 * l : task.col_start - 2 * l;
  const int64_t wa = rev ? cb - 7 : cb;
  const uint32_t sh = (uint32_t)(wa & 7) * 4u;
  int64_t dw = rev ? (wa >> 3) : (wa >> 3) + 1;
  st.dcur = rev ? codes[dw] : codes[dw - 1];
  st.dnext = rev ? codes[dw + 1] : codes[dw];
  st.w = 0;

  // N among the columns: the group's lanes scan the column sequence's code
  // dwords (codes [c0, c0 + col_len); nibbles outside it are masked off)
  if (col_len > 0) {
    const int64_t c0 = rev ? task.col_start - col_len + 1 : task.col_start;
    const int64_t c1 = c0 + col_len;
    for (int64_t d = (c0 >> 3) + l; d <= ((c1 - 1) >> 3); d += G) {
      uint32_t v = codes[d] & 0x44444444u;  // bit 2 of a nibble: N (codes 0..4)
      const int64_t n0 = d * 8;
      if (n0 < c0) v &= 0xFFFFFFFFu << (uint32_t)(4 * (c0 - n0));
      if (n0 + 8 > c1) v &= 0xFFFFFFFFu >> (uint32_t)(4 * (n0 + 8 - c1));
      has_n = has_n || v != 0;
    }
  }
  const bool pads = b.score2 != 0 && !rev;
  // the table form serves wavefronts whose pairs hold no N (and no score2
  // padding): row codes become byte selectors, the last column's table starts
  // as the kill table
  const bool tab = !pads && __ballot(has_n) == 0;
  if (tab) {
#pragma unroll
    for (int p = 0; p < P; p++) {
      const uint32_t c = st.rb[p];
      const uint32_t nlo = c & 0xFu, nhi = (c >> 16) & 0xFu;
      const uint32_t slo = nlo == (uint32_t)kCodeN ? 0x0C0Cu : ((3u - nlo) | 0x0C00u);
      const uint32_t shi = nhi == (uint32_t)kCodeN ? 0x0C0Cu : ((7u - nhi) | 0x0C00u);
      st.rb[p] = slo | (shi << 16);
    }
    st.q = k.amb4;
  }

  if (tab) {
    if (rev) {
      fast_run<G, P, kSubTab, kSubTab, true, TB>(st, nsteps, mincols, l, col_len, npad, codes, dw, sh, k);
    } else {
      fast_run<G, P, kSubTab, kSubTab, false, TB>(st, nsteps, mincols, l, col_len, npad, codes, dw, sh, k);
    }
  } else if (rev) {
    fast_run<G, P, kSubCodes, kSubCodes, true, TB>(st, nsteps, mincols, l, col_len, npad, codes, dw, sh, k);
  } else if (pads) {
    fast_run<G, P, kSubPad, kSubCodes, false, TB>(st, nsteps, mincols, l, col_len, npad, codes, dw, sh, k);
  } else {
    fast_run<G, P, kSubCodes, kSubCodes, false, TB>(st, nsteps, mincols, l, col_len, npad, codes, dw, sh, k);
  }

  // per-lane best (score, i, j), then the group's.  The lane's first row
  // passes through an empty asm so that the compiler does not hoist the 2P
  // rows' key parts out of the task loop (20 registers held across it).
  int32_t lrow = 2 * P * l;
  __asm__ volatile("" : "+v"(lrow));
  uint64_t bestk = 0;
  int32_t hmax = 0;
  const bool rows_ref = (task.flags & kTaskRowsAreRef) != 0;
#pragma unroll
  for (int p = 0; p < P; p++) {
#pragma unroll
    for (int half = 0; half < 2; half++) {
      const int32_t h = (int32_t)(((st.rec_h[p] >> (16 * half)) & 0xFFFFu) >> TB);
      hmax = h > hmax ? h : hmax;
      const int32_t tf = (int32_t)((st.rec_s[p] >> (16 * half)) & 0xFFFFu);
      const int32_t r = lrow + half * P + p;
      const int32_t col = tf - (2 * l + half);
      if (r < row_len && h > 0 && col >= 0 && col < col_len) {
        const uint64_t k = rows_ref ? end_key(h, r, col) : end_key(h, col, r);
        bestk = k > bestk ? k : bestk;
      }
    }
  }
#pragma unroll
  for (int o = G / 2; o >= 1; o >>= 1) {
    const uint64_t other = shfl_xor_u64(bestk, o, G);
    bestk = other > bestk ? other : bestk;
    const int32_t ho2 = __shfl_xor(hmax, o, G);
    hmax = ho2 > hmax ? ho2 : hmax;
  }
  // score2: the largest per-row maximum (padding columns included, as SSW's
  // maxColumn) over the SSW-ref rows outside the mask window around the end
  // (ssw_core.cpp:512-536, :822-842); rows are SSW-ref for these tasks.
  int32_t s2v = 0;
  bool s2_rerun = false;
  if (pads) {
    int32_t es = 0, ei = -1, ej = 0;
    if (bestk != 0) key_decode(bestk, es, ei, ej);
    const bool word = es + b.sc.bias >= 255;
    const int32_t mask = row_len / 2 > 15 ? row_len / 2 : 15;
    const int32_t lo = ei - mask, hi = ei + mask - (word ? 1 : 0);
    int32_t m = 0;
#pragma unroll
    for (int p = 0; p < P; p++) {
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const int32_t r = 2 * P * l + half * P + p;
        const int32_t h = (int32_t)(((st.rec_h[p] >> (16 * half)) & 0xFFFFu) >> TB);
        if (r < row_len && (r < lo || r > hi)) m = h > m ? h : m;
      }
    }
#pragma unroll
    for (int o = G / 2; o >= 1; o >>= 1) {
      const int32_t mo = __shfl_xor(m, o, G);
      m = mo > m ? mo : m;
    }
    s2v = m;
    // the 16-bit pass pads to 8 positions: recompute when that differs
    s2_rerun = ssw_query_pad(col_len, es, b.sc.bias) != npad;
  }
  if (active && l == 0) {
    PairEnd e;
    if (bestk == 0) {
      e = PairEnd{0, -1, 0, 0};
    } else {
      key_decode(bestk, e.score, e.i, e.j);
      e.flags = 0;
    }
    // values may have left the exact range: rerun
    if (hmax >= (TB == 5 ? kFastOverflowScore32 : kFastOverflowScore)) e.flags |= kEndOverflow;
    // rows past the group's 2GP (a trimmed shape given a taller pair, which
    // the host's class row bound rules out): rerun exactly
    if (row_len > 2 * G * P) e.flags |= kEndOverflow;
    if (s2task) {
      if (s2_rerun) {
        e.flags |= kEndScore2;
      } else {
        b.s2[task.pair] = s2v;
      }
    }
    if (task.flags & kTaskReverse) {
      b.rev[task.pair] = e;
    } else {
      b.fwd[task.pair] = e;
    }
  }
}

// Waves per SIMD the compiler must leave room for (168 VGPRs for 3)
template <int G, int P>
constexpr int fast_min_waves() {
  return P > 8 ? 2 : P >= 7 ? 3 : 1;
}

// The class a shape of 2GP rows serves.  The 16-lane class Pc
// (kClassFast16 + Pc - 1, rows up to 32 Pc or the fifth class's 152) runs G
// lanes per pair with P packed registers each: 16 Pc / G, or fewer (the
// fifth class's 19, a trimmed shape for a class whose rows are all short
// enough), 2GP rows still above the class below's 32 (Pc - 1).  The 64-lane
// class Pc (kClassFast64 + Pc - 1, rows up to 128 Pc) runs 64 lanes of Pc
// registers, or (forward, rows above 256) 16 or 32 lanes of 2GP = 128 Pc rows.
template <int G, int P>
constexpr int fast_class_of() {
  return G == 64 ? kClassFast64 + P - 1
         : 2 * G * P > 32 * kMaxP ? kClassFast64 + (2 * G * P + 127) / 128 - 1
                                  : kClassFast16 + (G * P + 15) / 16 - 1;
}

// The class's wave tasks from its work counter.  The next task's number is
// taken before the current one is aligned, and (shapes of at most kMaxP
// registers) its record loaded; the shapes of more registers load it
// afterwards.
template <int G, int P, int TB>
__device__ __forceinline__ void fast_tasks(const BatchView& b, int cls, int32_t base, int32_t nwt, int lane, int l,
                                           int grp) {
  constexpr int GPW = 64 / G;
  PassSched* s = b.sched;
  int w = next_wave_task(&s->class_next[cls], lane);
  SwTask task;
  if (w < nwt) task = b.tasks[base + w * GPW + grp];
  while (w < nwt) {
    const int wn = next_wave_task(&s->class_next[cls], lane);
    if constexpr (G >= 16 && P <= kMaxP) {
      SwTask next;
      if (wn < nwt) next = b.tasks[base + wn * GPW + grp];
      fast_pair<G, P, TB>(b, task, l);
      task = next;
    } else {
      fast_pair<G, P, TB>(b, task, l);
      if (wn < nwt) task = b.tasks[base + wn * GPW + grp];
    }
    w = wn;
  }
}

template <int G, int P>
__global__ __launch_bounds__(256, (fast_min_waves<G, P>())) void sw_fast_kernel(BatchView b) {
  constexpr int GPW = 64 / G;
  static_assert(G >= 2 && 64 % G == 0 && 2 * G * P <= 128 * kMaxP, "a packed class's shape");
  const int cls = fast_class_of<G, P>();
  const int lane = threadIdx.x & 63;
  const int l = lane & (G - 1);
  const int grp = lane / G;
  PassSched* s = b.sched;
  const int32_t nwt = s->class_slots[cls] / GPW;
  // wavefronts numbered past the class's work leave before touching the shared
  // counter (the rest still take every task: each loops until it runs out)
  if ((int32_t)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64) >= nwt) return;
  const int32_t base = s->class_base[cls];
  // Narrow groups (G <= 8) of forward launches keep H as 32 H when the host
  // allows it (b.fold32: match + gap open <= 7 and the class's rows times
  // match below kFastOverflowScore32), which folds the row maxima every 32
  // steps instead of 16.
  if constexpr (G <= 8) {
    if (b.fold32) {
      fast_tasks<G, P, 5>(b, cls, base, nwt, lane, l, grp);
      return;
    }
  }
  fast_tasks<G, P, 4>(b, cls, base, nwt, lane, l, grp);
}

#ifndef MHMSW_FAST_PART

// ----------------------------------------------------------------------------
// banded reverse kernel: one lane per pair
// ----------------------------------------------------------------------------
// The reverse rectangle's rows k index the reversed SSW-ref prefix, its
// columns l the reversed SSW-query prefix, and only diagonals c = l - k + D in
// [0, 2D] can hold the score.  A lane keeps 2P slots (P packed registers; low
// halves are slots 0..P-1, high halves P..2P-1).  At u-step u, slot j works on
// row k = u - j, first (even step) on diagonal c = 2j, then (odd step) on
// c = 2j + 1, so 4P diagonals are covered.  Every input of a cell was produced
// one step earlier:
//   even step, c = 2j:   H(k-1,l-1) own slot, last even step
//                        E from (k-1,l) = c+1: own slot, last odd step
//                        F from (k,l-1) = c-1: slot j-1, last odd step
//   odd step, c = 2j+1:  H(k-1,l-1) own slot, last odd step
//                        E from (k-1,l) = c+1: slot j+1, this u's even step
//                        F from (k,l-1) = c-1: own slot, this u's even step
// so a u-step has no chain across slots and no cross-lane traffic.  Row codes
// move up one slot per u-step (slot j needs R'[u-j]) and column codes down one
// (slot j needs Q'[u+j-D] and Q'[u+j+1-D]); both are register rings indexed
// with the step number, so a move costs one insertion per u-step.
// Cells off the sequences are killed like the packed kernel's: left of column
// 0 and above row 0 they hold the boundary (H = E = F = 0); right of the last
// column and below the last row they only feed each other and stay below the
// score.  The answer is the least (k, l) with H = score, which SSW's reverse
// pass selects (ssw_core.cpp:1162-1178).
template <int P>
struct BandState {
  uint32_t rr[P];          // row codes (selector form), slot p at rr[(p - s) mod P]
  uint32_t qr[P];          // even-step column codes, slot p at qr[(p + s) mod P]
  uint32_t hoe[P], hoo[P]; // 16*H of the slot's last even / odd cell
  uint32_t eo[P], fo[P];   // 16*(E + gO), 16*(F + gO) leaving the last odd step (clamped >= 0)
  uint32_t be[P], bo[P];   // max of 16*H + tag over this chunk's even / odd steps
};

// Code streams of the band kernels.  A lane streams its reversed row prefix
// and reversed column prefix: one dword window per chunk of 8 u-steps, one
// dword lower each chunk.  The narrow classes (P <= 2) stage the streams
// through LDS: per block of S chunks the wavefront loads, for each of its 64
// lanes and each stream, the segment of S dwords those chunks read (the
// window's upper dword of the block's first chunk is the previous segment's
// first dword, kept in a register).  Each load instruction covers 64 / S
// lanes' segments of S consecutive dwords, so a segment costs one coalesced
// request instead of one request per dword per lane; segments are S + 1
// dwords apart in LDS so that the lanes' reads hit distinct banks.  The wide
// classes spend many more VALU cycles per chunk and hide the per-lane loads,
// one chunk ahead, behind them; there the segment loads, waited on at once,
// measured slower, and S = 0 selects the per-lane loads.
// A segment of 32 dwords is a 128-byte line: each lane's line is consumed
// whole by the chunks of one block, before the L2 (shared by about 130k
// concurrent streams on the device) evicts it; 8-dword segments re-fetched
// each line up to four times.
template <int P>
constexpr int band_seg() {
  return P <= 2 ? 32 : 0;
}
template <int P>
constexpr int band_lds_wave() {  // dwords per wavefront: row and column streams
  return 2 * 64 * (band_seg<P>() + 1);
}
using band_lds_t = __attribute__((address_space(3))) uint32_t;

// Load segment [base_s, base_s + S) of every lane s's row stream (clamped
// below at that lane's first dword, first_s) into lds[s * (S + 1) ..], and
// likewise its column stream into lds[(64 + s) * (S + 1) ..].  Every load is
// issued before the first LDS write, so the wavefront waits for memory once
// per block of chunks rather than once per few loads (with one or two
// wavefronts per SIMD nothing else hides that latency).
template <int S>
__device__ __forceinline__ void band_load_segments(band_lds_t* lds, const uint32_t* __restrict__ codes, int lane,
                                                   int32_t rbase, int32_t rfirst, int32_t qbase, int32_t qfirst) {
  constexpr int per = 64 / S;  // lanes' segments per load instruction
  const int i = lane % S;
  uint32_t rv[S], qv[S];
#pragma unroll
  for (int j = 0; j < S; j++) {
    const int sl = per * j + lane / S;
    const int32_t rb = __shfl(rbase, sl, 64), rf = __shfl(rfirst, sl, 64);
    const int32_t qb = __shfl(qbase, sl, 64), qf = __shfl(qfirst, sl, 64);
    rv[j] = codes[rb + i > rf ? rb + i : rf];
    qv[j] = codes[qb + i > qf ? qb + i : qf];
  }
#pragma unroll
  for (int j = 0; j < S; j++) {
    const int sl = per * j + lane / S;
    lds[sl * (S + 1) + i] = rv[j];
    lds[(64 + sl) * (S + 1) + i] = qv[j];
  }
}

// One packed cell update in the forward kernel's biased form: h_slot holds
// 16*H of the slot's diagonal predecessor (the slot's cell one u-step
// earlier), ein/fin 16*(E + gO) and 16*(F + gO).  The chunk maxima take two
// u-steps at a time: at an odd u-step (TWO) one three-way max adds this cell
// (tagged tag) and the predecessor (tagged tag + 1).
template <bool TWO>
__device__ __forceinline__ void band_cell(uint32_t r, uint32_t q, uint32_t& h_slot, uint32_t ein, uint32_t fin,
                                          uint32_t& eout, uint32_t& fout, uint32_t& best, uint32_t tag, uint32_t tlo,
                                          uint32_t thi, uint32_t go, uint32_t ge) {
  const uint32_t sel = subst_sel(r, q);
  const uint32_t sp = __builtin_amdgcn_perm(thi, tlo, sel);
  const uint32_t prev = h_slot;
  const uint32_t d = pk_addu_sat(prev, sp);
  const uint32_t h = pk_subu_sat(pk_max3_u16f(d, ein, fin), go);
  h_slot = h;
  eout = pk_max(pk_subu_sat(ein, ge), h);
  fout = pk_max(pk_subu_sat(fin, ge), h);
  if constexpr (TWO) best = pk_max3_u16f(best, h | tag, prev | (tag + 0x00010001u));
}

// One chunk of 8 u-steps (16 cell steps).  rw holds row codes R'[u0+1 ..
// u0+8] and qw column codes Q'[u0+2P-D .. +7], lowest nibble first: the
// reversed prefixes are the forward codes read backwards, so each chunk's
// window lies one dword below the last one and its nibbles are reversed.
// NL lanes per pair (NL = 2 or 4): lane g of the pair holds the pair's slots
// 2P*g .. 2P*g + 2P - 1.  The cells' cross-slot inputs at adjacent lanes'
// common edges (F into a lane's slot 0 from the lane below, E into its top
// slot from the lane above) and the codes moving between the lanes' rings
// come from the neighbouring lane by a quad_perm DPP move; the lowest lane
// takes the row stream, the highest the column stream.
template <int NL>
__device__ __forceinline__ uint32_t from_lower_lane(uint32_t v, int g) {  // 0 for the lowest lane
  constexpr int ctrl = NL == 2 ? 0xA0 : 0x90;  // quad_perm 0,0,2,2 / 0,0,1,2
  const uint32_t x = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xF, 0xF, false);
  return g ? x : 0u;
}
template <int NL>
__device__ __forceinline__ uint32_t from_upper_lane(uint32_t v, int g) {  // 0 for the highest lane
  constexpr int ctrl = NL == 2 ? 0xF5 : 0xF9;  // quad_perm 1,1,3,3 / 1,2,3,3
  const uint32_t x = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xF, 0xF, false);
  return g == NL - 1 ? 0u : x;
}

template <int P, int NL, bool EDGE>
__device__ __forceinline__ void band_chunk(BandState<P>& st, int g, int u0, int K, int L, int D, uint32_t rw,
                                           uint32_t qw, uint32_t tlo, uint32_t thi, uint32_t go, uint32_t ge) {
#pragma unroll
  for (int p = 0; p < P; p++) {
    st.be[p] = 0;
    st.bo[p] = 0;
  }
#pragma unroll
  for (int s = 0; s < 8; s++) {
    const uint32_t tag = (uint32_t)(7 - s) * 0x00010001u;
    uint32_t ee[P], fe[P];
    // F into slot 0 comes from slot -1: the lane below's top slot
    uint32_t f0 = st.fo[P - 1] << 16;
    if constexpr (NL > 1) f0 |= from_lower_lane<NL>(st.fo[P - 1], g) >> 16;
#pragma unroll
    for (int p = 0; p < P; p++) {
      const uint32_t fin = p ? st.fo[p - 1] : f0;
      if (s & 1) {  // (s is a constant after unrolling)
        band_cell<true>(st.rr[(p - s + 8 * P) % P], st.qr[(p + s) % P], st.hoe[p], st.eo[p], fin, ee[p], fe[p],
                        st.be[p], tag, tlo, thi, go, ge);
      } else {
        band_cell<false>(st.rr[(p - s + 8 * P) % P], st.qr[(p + s) % P], st.hoe[p], st.eo[p], fin, ee[p], fe[p],
                         st.be[p], tag, tlo, thi, go, ge);
      }
    }
    // column code entering the top slot for the odd step: from the stream (the
    // highest lane of a pair), or the code leaving the next lane's slot 0
    uint32_t qn = (qw >> (4 * s)) & 0xFu;
    if constexpr (EDGE) qn = (u0 + s + 2 * P * NL - D < L) ? qn : kKillColN;
    if constexpr (NL > 1) {
      const uint32_t down = from_upper_lane<NL>(st.qr[s % P] & 0xFFFFu, g);
      qn = g == NL - 1 ? qn : down;
    }
    st.qr[s % P] = __builtin_amdgcn_alignbit(qn, st.qr[s % P], 16);
    // E into the top slot comes from slot 2P: the lane above's slot 0
    uint32_t etop = ee[0] >> 16;
    if constexpr (NL > 1) etop |= from_upper_lane<NL>(ee[0], g) << 16;
#pragma unroll
    for (int p = 0; p < P; p++) {
      const uint32_t ein = p < P - 1 ? ee[p + 1] : etop;
      if (s & 1) {
        band_cell<true>(st.rr[(p - s + 8 * P) % P], st.qr[(p + s + 1) % P], st.hoo[p], ein, fe[p], st.eo[p],
                        st.fo[p], st.bo[p], tag, tlo, thi, go, ge);
      } else {
        band_cell<false>(st.rr[(p - s + 8 * P) % P], st.qr[(p + s + 1) % P], st.hoo[p], ein, fe[p], st.eo[p],
                         st.fo[p], st.bo[p], tag, tlo, thi, go, ge);
      }
    }
    // row code entering slot 0 for the next u-step: from the stream (the
    // lowest lane of a pair), or the code leaving the lane below's top slot
    uint32_t rn = ((rw >> (4 * s)) & 0xFu) | kRowHi;
    if constexpr (EDGE) rn = (u0 + s + 1 < K) ? rn : kKillRowN;
    const int ri = (P - 1 - s + 8 * P) % P;
    if constexpr (NL > 1) {
      const uint32_t up = from_lower_lane<NL>(st.rr[ri] >> 16, g);
      rn = g ? up : rn;
    }
    st.rr[ri] = (st.rr[ri] << 16) | rn;
  }
}

// Least (k, l) among this chunk's cells holding the score (s16 = 16 * score).
template <int P>
__device__ __forceinline__ void band_hits(const BandState<P>& st, int g, int u0, int K, int L, int D, uint32_t s16,
                                          uint32_t& keymin) {
#pragma unroll
  for (int par = 0; par < 2; par++) {
#pragma unroll
    for (int p = 0; p < P; p++) {
#pragma unroll
      for (int half = 0; half < 2; half++) {
        const uint32_t v = ((par ? st.bo[p] : st.be[p]) >> (16 * half)) & 0xFFFFu;
        if ((v & 0xFFF0u) == s16) {
          const int j = 2 * P * g + half * P + p;  // the pair's slot
          const int u = u0 + 7 - (int)(v & 0xFu);
          const int k = u - j;
          const int l = k + 2 * j + par - D;
          if (k >= 0 && k < K && l >= 0 && l < L) {
            const uint32_t key = ((uint32_t)k << 16) | (uint32_t)l;
            keymin = key < keymin ? key : keymin;
          }
        }
      }
    }
  }
}

template <int P, int NL>
__device__ void band_pair(const BatchView& b, const SwTask& task, band_lds_t* lds) {
  static_assert(NL == 1 || NL == 2 || NL == 4, "one, two or four lanes per pair");
  static_assert(NL == 1 || band_seg<P>() == 0, "multi-lane pairs load their streams per lane");
  const int g = (threadIdx.x & 63) & (NL - 1);  // the lane's place in its pair
  const bool active = task.pair >= 0;
  const int K = active ? task.row_len : 0;
  const int L = active ? task.col_len : 0;
  const int D = active ? ((task.flags >> kTaskBandShift) & 0x1F) : 0;
  const int score = active ? (task.flags >> kTaskScoreShift) : 0;
  // wave-uniform loop length and the chunks that need no end-of-sequence checks
  const int nU = uni(wave_max_i32(active ? K + D : 0));
  if (nU <= 0) return;
  const int kmin = uni(wave_min_i32(active ? K : 0x7FFFFFFF));
  const int qmin = uni(wave_min_i32(active ? L - 2 * P * NL + D : 0x7FFFFFFF));

  const uint32_t* __restrict__ codes = b.codes;
  const uint32_t tlo = b.sc.perm_lo, thi = b.sc.perm_hi;
  const uint32_t go = (uint32_t)(16 * b.sc.gap_open) * 0x00010001u;
  const uint32_t ge = (uint32_t)(16 * b.sc.gap_ext) * 0x00010001u;
  const uint32_t s16 = (uint32_t)(16 * score);

  BandState<P> st;
  // u = 0: slot j holds R'[-j] (only slot 0 is on the sequence) and
  // Q'[j - D] for the even step (R'[k] = codes[row_start - k], Q'[l] =
  // codes[col_start - l]); j is the pair's slot, 2P*g + the lane's slot
  const uint32_t r0 = K > 0 && g == 0 ? (code_at(codes, task.row_start) | kRowHi) : kKillRowN;
#pragma unroll
  for (int p = 0; p < P; p++) {
    st.rr[p] = p == 0 ? (r0 | (kKillRowN << 16)) : (kKillRowN | (kKillRowN << 16));
    const int ilo = 2 * P * g + p - D, ihi = 2 * P * g + P + p - D;
    const uint32_t qlo = (ilo >= 0 && ilo < L) ? code_at(codes, task.col_start - ilo) : kKillColN;
    const uint32_t qhi = (ihi >= 0 && ihi < L) ? code_at(codes, task.col_start - ihi) : kKillColN;
    st.qr[p] = qlo | (qhi << 16);
    st.hoe[p] = 0;
    st.hoo[p] = 0;
    st.eo[p] = 0;
    st.fo[p] = 0;
  }
  // code streams: rows from R'[1], columns from Q'[2P - D] (2P - D >= 1);
  // chunk m's window (8 codes) starts 8m codes below the first one (ra, qa),
  // so it is dwords rd0 - m and rd0 - m + 1 (qd0 likewise).  Dword indices
  // are 32-bit (the job's code buffer holds fewer than 2^31 dwords).
  const int64_t ra = task.row_start - 8, qa = task.col_start - (2 * P * NL - D) - 7;
  const int32_t rd0 = (int32_t)(ra >> 3), qd0 = (int32_t)(qa >> 3);
  const uint32_t rsh = (uint32_t)(ra & 7) * 4u, qsh = (uint32_t)(qa & 7) * 4u;
  // stream reads stay within the lane's own sequence (less one dword): its
  // first code is R'[K-1] = codes[row_start - K + 1]
  const int32_t rfirst = (int32_t)((task.row_start - K + 1) >> 3) - 1;
  const int32_t qfirst = (int32_t)((task.col_start - L + 1) >> 3) - 1;
  constexpr int S = band_seg<P>();
  const int lane = threadIdx.x & 63;
  band_lds_t* rseg = lds + lane * (S + 1);         // this lane's row segment
  band_lds_t* qseg = lds + (64 + lane) * (S + 1);  // and column segment
  // upper dwords of the first chunk's windows (then: of each LDS block's
  // first chunk, or of the next chunk's window with per-lane loads)
  uint32_t rcarry = codes[rd0 + 1], qcarry = codes[qd0 + 1];
  // per-lane loads: lower dwords of the next chunk's windows
  uint32_t rnow = 0, qnow = 0;
  if constexpr (S == 0) {
    rnow = codes[rd0 > rfirst ? rd0 : rfirst];
    qnow = codes[qd0 > qfirst ? qd0 : qfirst];
  }

  uint32_t keymin = 0xFFFFFFFFu;
  for (int u0 = 0; u0 < nU; u0 += 8) {
    const int mc = u0 >> 3, c = S > 0 ? mc & (S - 1) : 0;
    uint32_t rlo, rhi, qlo, qhi;
    if constexpr (S == 0) {
      rlo = rnow;
      rhi = rcarry;
      qlo = qnow;
      qhi = qcarry;
      rcarry = rnow;
      qcarry = qnow;
      const int32_t rn = rd0 - mc - 1, qn = qd0 - mc - 1;
      rnow = codes[rn > rfirst ? rn : rfirst];
      qnow = codes[qn > qfirst ? qn : qfirst];
    } else if (c == 0) {
      // block of chunks mc .. mc + S - 1: dwords d0 - mc - (S - 1) .. d0 - mc
      // (the LDS reads of the previous block came before these writes: a
      // wavefront's LDS operations run in order)
      __builtin_amdgcn_wave_barrier();
      band_load_segments<S>(lds, codes, lane, rd0 - mc - (S - 1), rfirst, qd0 - mc - (S - 1), qfirst);
      __builtin_amdgcn_wave_barrier();
      rhi = rcarry;
      qhi = qcarry;
      rcarry = rseg[0];
      qcarry = qseg[0];
      rlo = rseg[S - 1];
      qlo = qseg[S - 1];
    } else {
      rhi = rseg[S - c];
      qhi = qseg[S - c];
      rlo = rseg[S - 1 - c];
      qlo = qseg[S - 1 - c];
    }
    const uint32_t rw = nib_reverse(__builtin_amdgcn_alignbit(rhi, rlo, rsh));
    const uint32_t qw = nib_reverse(__builtin_amdgcn_alignbit(qhi, qlo, qsh));
    const bool middle = u0 + 8 < kmin && u0 + 8 <= qmin;
    if (middle) {
      band_chunk<P, NL, false>(st, g, u0, K, L, D, rw, qw, tlo, thi, go, ge);
    } else {
      band_chunk<P, NL, true>(st, g, u0, K, L, D, rw, qw, tlo, thi, go, ge);
    }
    // scores never exceed the forward score: a hit is a chunk maximum equal to it
    uint32_t m = 0;
#pragma unroll
    for (int p = 0; p < P; p++) m = pk_maxu(m, pk_maxu(st.be[p], st.bo[p]));
    const uint32_t mh = (m & 0xFFFFu) > (m >> 16) ? (m & 0xFFFFu) : (m >> 16);
    if (active && score > 0 && (mh & 0xFFF0u) == s16) band_hits<P>(st, g, u0, K, L, D, s16, keymin);
  }
#pragma unroll
  for (int m = 1; m < NL; m <<= 1) {  // the pair's least hit over its lanes, written by its lowest lane
    const uint32_t other = (uint32_t)__shfl_xor((int)keymin, m, 64);
    keymin = other < keymin ? other : keymin;
  }
  if (active && g == 0) {
    PairEnd e{0, 0, 0, 0};
    if (keymin != 0xFFFFFFFFu) e = PairEnd{score, (int32_t)(keymin >> 16), (int32_t)(keymin & 0xFFFFu), 0};
    b.rev[task.pair] = e;
  }
}

// band class of P packed registers per lane (4P diagonals): P = 1, 2, 4, 8
template <int P, int NL>
constexpr int band_class_of() {  // the class a (P, NL) layout serves: 4P * NL diagonals
  return kClassBand + (4 * P * NL == 4 ? 0 : 4 * P * NL == 8 ? 1 : 4 * P * NL == 16 ? 2 : 4 * P * NL == 32 ? 3 : 4);
}

// Take the class's wave tasks from its work counter until they run out.
// Wavefronts numbered past the class's work leave before touching the shared
// counter (the rest still take every task: each loops until it runs out).
template <int P, int NL = 1>
__device__ __forceinline__ void band_class(const BatchView& b, int lane, int32_t wave, band_lds_t* lds) {
  constexpr int cls = band_class_of<P, NL>();
  constexpr int per = 64 / NL;  // pairs per wavefront
  PassSched* s = b.sched;
  const int32_t nwt = s->class_slots[cls] / per;
  if (wave >= nwt) return;
  const int32_t base = s->class_base[cls];
  int w = next_wave_task(&s->class_next[cls], lane);
  SwTask task;
  if (w < nwt) task = b.tasks[base + w * per + lane / NL];
  while (w < nwt) {
    const int wn = next_wave_task(&s->class_next[cls], lane);
    SwTask next;
    if (wn < nwt) next = b.tasks[base + wn * per + lane / NL];
    band_pair<P, NL>(b, task, lds);
    task = next;
    w = wn;
  }
}

// Reverse pass of a pair whose band is the single diagonal (D = 0: no
// alignment with a gap can reach the score).  With the cells off the
// diagonal absent, the band's cells are H'(k, k) = max(0, H'(k-1, k-1) +
// s(R'[k], Q'[k])), and the answer is the least k holding the score.  One
// lane per pair walks the diagonal 8 positions per chunk: the chunk's 8 row
// and column codes make 8 selector bytes (even and odd positions) and two
// v_perm turn them into the scoring table's 16*(s + gap_open) bytes; then per
// position one extract, one add and one max update 16*H, and a shift-or and
// a max keep the chunk's first maximum.  The code streams are staged through
// LDS in the narrow band classes' segments, and a wavefront stops when every
// lane has its answer.
__device__ void diag_pair(const BatchView& b, const SwTask& task, band_lds_t* lds) {
  constexpr int S = band_seg<1>();  // the P = 1 segment layout
  static_assert(band_lds_wave<1>() == 2 * 64 * (S + 1), "the diagonal class uses the P = 1 segment layout");
  const bool active = task.pair >= 0;
  const int K = active ? task.row_len : 0;  // rows and columns: with D = 0 the rectangle is K x K
  const int score = active ? (task.flags >> kTaskScoreShift) : 0;
  const int nU = uni(wave_max_i32(K));
  if (nU <= 0) return;
  const int kmin = uni(wave_min_i32(active ? K : 0x7FFFFFFF));
  const uint32_t* __restrict__ codes = b.codes;
  const uint32_t tlo = b.sc.perm_lo, thi = b.sc.perm_hi;
  const int32_t go = 16 * b.sc.gap_open;
  const uint32_t s16 = (uint32_t)(16 * score);
  // streams: R'[k] = codes[row_start - k], Q'[k] = codes[col_start - k]; chunk
  // m's window of 8 codes starts at row_start - 8m - 7, i.e. dwords rd0 - m and
  // rd0 - m + 1 (qd0 likewise)
  const int64_t ra = task.row_start - 7, qa = task.col_start - 7;
  const int32_t rd0 = (int32_t)(ra >> 3), qd0 = (int32_t)(qa >> 3);
  const uint32_t rsh = (uint32_t)(ra & 7) * 4u, qsh = (uint32_t)(qa & 7) * 4u;
  const int32_t rfirst = (int32_t)((task.row_start - K + 1) >> 3) - 1;
  const int32_t qfirst = (int32_t)((task.col_start - K + 1) >> 3) - 1;
  const int lane = threadIdx.x & 63;
  band_lds_t* rseg = lds + lane * (S + 1);
  band_lds_t* qseg = lds + (64 + lane) * (S + 1);
  uint32_t rcarry = codes[rd0 + 1], qcarry = codes[qd0 + 1];
  int32_t h = 0;     // 16 * H'(k - 1, k - 1)
  int32_t hit = -1;  // least k with H' = score
  for (int u0 = 0; u0 < nU; u0 += 8) {
    const int mc = u0 >> 3, c = mc & (S - 1);
    uint32_t rlo, rhi, qlo, qhi;
    if (c == 0) {
      __builtin_amdgcn_wave_barrier();
      band_load_segments<S>(lds, codes, lane, rd0 - mc - (S - 1), rfirst, qd0 - mc - (S - 1), qfirst);
      __builtin_amdgcn_wave_barrier();
      rhi = rcarry;
      qhi = qcarry;
      rcarry = rseg[0];
      qcarry = qseg[0];
      rlo = rseg[S - 1];
      qlo = qseg[S - 1];
    } else {
      rhi = rseg[S - c];
      qhi = qseg[S - c];
      rlo = rseg[S - 1 - c];
      qlo = qseg[S - 1 - c];
    }
    const uint32_t rw = nib_reverse(__builtin_amdgcn_alignbit(rhi, rlo, rsh));
    const uint32_t qw = nib_reverse(__builtin_amdgcn_alignbit(qhi, qlo, qsh));
    // selector bytes: bits 0-1 r ^ q, bits 2-3 r | q (the N bit); even
    // positions from the low nibbles, odd ones from the high nibbles.
    // Positions past the pair (k >= K) get selector 12, which v_perm turns
    // into 0 (s = -gap_open): there H only falls, so the chunk's maximum
    // never passes the score on codes beyond the pair.
    uint32_t sle = __builtin_amdgcn_bitop3_b32(rw & 0x0F0F0F0Fu, qw & 0x0F0F0F0Fu, 0x03030303u, 0x7C);
    uint32_t slo = __builtin_amdgcn_bitop3_b32((rw >> 4) & 0x0F0F0F0Fu, (qw >> 4) & 0x0F0F0F0Fu, 0x03030303u, 0x7C);
    if (u0 + 8 > kmin) {  // some lane's pair ends in this chunk
      const int lim = K - u0 < 0 ? 0 : K - u0;  // positions kept: i < lim
      const int je = (lim + 1) >> 1, jo = lim >> 1;  // first killed byte of each half
      const uint32_t me = je >= 4 ? 0u : 0xFFFFFFFFu << (8 * je);
      const uint32_t mo = jo >= 4 ? 0u : 0xFFFFFFFFu << (8 * jo);
      sle = (sle & ~me) | (0x0C0C0C0Cu & me);
      slo = (slo & ~mo) | (0x0C0C0C0Cu & mo);
    }
    const uint32_t se = __builtin_amdgcn_perm(thi, tlo, sle);
    const uint32_t so = __builtin_amdgcn_perm(thi, tlo, slo);
    uint32_t best = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int32_t sp = (int32_t)__builtin_amdgcn_ubfe((i & 1) ? so : se, 8 * (i >> 1), 8);
      h = max(h + sp - go, 0);
      best = max(best, ((uint32_t)h << 4) | (uint32_t)(7 - i));
    }
    if (hit < 0 && (best >> 4) == s16 && score > 0) {
      const int k = u0 + 7 - (int)(best & 15u);
      if (k < K) hit = k;
    }
    if (__ballot(active && hit < 0 && u0 + 8 < K) == 0) break;
  }
  if (active) b.rev[task.pair] = hit >= 0 ? PairEnd{score, hit, hit, 0} : PairEnd{0, 0, 0, 0};
}

__device__ __forceinline__ void diag_class(const BatchView& b, int lane, int32_t wave, band_lds_t* lds) {
  constexpr int cls = kClassDiag;
  PassSched* s = b.sched;
  const int32_t nwt = s->class_slots[cls] / 64;
  if (wave >= nwt) return;
  const int32_t base = s->class_base[cls];
  int w = next_wave_task(&s->class_next[cls], lane);
  SwTask task;
  if (w < nwt) task = b.tasks[base + w * 64 + lane];
  while (w < nwt) {
    const int wn = next_wave_task(&s->class_next[cls], lane);
    SwTask next;
    if (wn < nwt) next = b.tasks[base + wn * 64 + lane];
    diag_pair(b, task, lds);
    task = next;
    w = wn;
  }
}

// this wavefront's part of the block's band LDS (dw dwords per wavefront)
__device__ __forceinline__ band_lds_t* band_wave_lds(uint32_t* block_lds, int dw) {
  return (band_lds_t*)block_lds + (threadIdx.x / 64) * dw;
}

template <int P, int NL = 1>
__global__ __launch_bounds__(256) void sw_band_kernel(BatchView b) {
  const int32_t wave = (int32_t)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
  if constexpr (band_seg<P>() > 0) {
    __shared__ uint32_t lds[4 * band_lds_wave<P>()];
    band_class<P, NL>(b, threadIdx.x & 63, wave, band_wave_lds(lds, band_lds_wave<P>()));
  } else {
    band_class<P, NL>(b, threadIdx.x & 63, wave, nullptr);
  }
}

// Every band class in one launch, for small jobs whose classes each fill only
// a few wavefronts: each wavefront takes the classes from the widest down, so
// they run side by side instead of one launch after the other.  (Large jobs
// keep one launch per class: this kernel holds the widest class's registers,
// which would cost the narrower classes waves per SIMD.)
__global__ __launch_bounds__(256) void sw_band_all_kernel(BatchView b) {
  static_assert(band_seg<1>() == band_seg<2>() && band_seg<4>() == 0 && band_seg<8>() == 0, "");
  __shared__ uint32_t lds[4 * band_lds_wave<1>()];  // the LDS-staged classes' segments
  band_lds_t* wl = band_wave_lds(lds, band_lds_wave<1>());
  const int lane = threadIdx.x & 63;
  const int32_t wave = (int32_t)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
  const int32_t nwaves = (int32_t)(gridDim.x * (blockDim.x / 64));
  band_class<8>(b, lane, wave, wl);
  band_class<4>(b, lane, nwaves - 1 - wave, wl);
  band_class<2>(b, lane, wave, wl);
  band_class<1>(b, lane, nwaves - 1 - wave, wl);
  diag_class(b, lane, wave, wl);
}

__global__ __launch_bounds__(256) void sw_diag_kernel(BatchView b) {
  __shared__ uint32_t lds[4 * band_lds_wave<1>()];
  const int32_t wave = (int32_t)(blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64);
  diag_class(b, threadIdx.x & 63, wave, band_wave_lds(lds, band_lds_wave<1>()));
}

// ----------------------------------------------------------------------------
// plain 32-bit kernel: one wavefront per pair, rows in tiles of 64
// ----------------------------------------------------------------------------
// One sweep of the 32-bit DP over a pair's rows in tiles of 64 (wave-uniform
// task).  Columns [cols, cols + npad) score 0 against every row: SSW's query
// padding, used only for score2.  s2 = false: returns the lexicographic end
// key over the real columns.  s2 = true: returns the largest row maximum
// (padding columns included) over the rows outside [lo, hi].
__device__ uint64_t plain_sweep(const BatchView& b, const SwTask& task, int k, int npad, bool s2, int32_t lo,
                                int32_t hi, int2* bnd) {
  const int32_t M = b.sc.match, X = b.sc.mismatch, A = b.sc.ambiguity, gO = b.sc.gap_open, gE = b.sc.gap_ext;
  const int32_t NEG = -(1 << 28);
  const uint32_t* __restrict__ codes = b.codes;
  const bool rows_ref = (task.flags & kTaskRowsAreRef) != 0;
  // reverse tasks read the forward codes backwards: row r at row_start - r,
  // column c at col_start - c
  const int64_t dir = (task.flags & kTaskReverse) ? -1 : 1;
  const int32_t rows = task.row_len, cols = task.col_len, ncols = cols + npad;
  uint64_t pair_best = 0;
  int32_t s2best = 0;
  const int ntiles = (rows + 63) / 64;
  for (int tile = 0; tile < ntiles; tile++) {
    const int32_t r = tile * 64 + k;
    const bool row_ok = r < rows;
    const uint32_t rc = row_ok ? code_at(codes, task.row_start + dir * r) : 0u;
    int32_t hprev = 0, hg = NEG, best = 0, bestc = 0, rowall = 0;
    int32_t up_h_prev = 0;  // H(r-1, c-1)
    int32_t out_h = 0, out_e = NEG;
    const int64_t cb = task.col_start - dir * k;  // column c = t - k is code cb + dir * t
    for (int t = 0; t < ncols + 63; t++) {
      const int32_t c = t - k;
      const bool col_ok = c >= 0 && c < ncols;
      const bool real = c < cols;
      // inputs from the row above at column c
      int32_t top_h = 0, top_e = NEG;
      if (k == 0 && tile > 0 && col_ok) {
        const int2 v = bnd[c];
        top_h = v.x;
        top_e = v.y;
      }
      const int32_t up_h = __builtin_amdgcn_update_dpp(top_h, out_h, 0x138, 0xF, 0xF, false);
      const int32_t up_e = __builtin_amdgcn_update_dpp(top_e, out_e, 0x138, 0xF, 0xF, false);
      const uint32_t q = (col_ok && real) ? code_at(codes, cb + dir * t) : 0u;
      const int32_t sub = real ? (((rc | q) & 4u) ? -A : (rc == q ? M : -X)) : 0;
      const int32_t e = max(up_e - gE, up_h - gO);     // gap along the rows
      const int32_t f = max(hg - gE, hprev - gO);      // gap along the columns
      int32_t h = max(max(up_h_prev + sub, e), max(f, 0));
      const bool ok = row_ok && col_ok;
      h = ok ? h : 0;
      const int32_t f_keep = ok ? f : NEG;
      const int32_t e_out = ok ? e : NEG;
      if (ok && real && h > best) {
        best = h;
        bestc = c;
      }
      rowall = h > rowall ? h : rowall;
      if (k == 63 && col_ok) bnd[c] = make_int2(h, e_out);
      up_h_prev = up_h;
      hprev = h;
      hg = f_keep;
      out_h = h;
      out_e = e_out;
    }
    if (s2) {
      const int32_t v = (row_ok && (r < lo || r > hi)) ? rowall : 0;
      const int32_t m = wave_max_i32(v);
      s2best = m > s2best ? m : s2best;
    } else {
      uint64_t key = 0;
      if (row_ok && best > 0) key = rows_ref ? end_key(best, r, bestc) : end_key(best, bestc, r);
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) {
        const uint64_t other = shfl_xor_u64(key, o, 64);
        key = other > key ? other : key;
      }
      pair_best = key > pair_best ? key : pair_best;
    }
    __syncthreads();
  }
  return s2 ? (uint64_t)(uint32_t)s2best : pair_best;
}

__global__ __launch_bounds__(64) void sw_plain_kernel(BatchView b) {
  extern __shared__ int2 bnd[];  // last row of the previous tile: (H, E) per column
  const int k = threadIdx.x;
  PassSched* s = b.sched;
  const int32_t n = s->class_count[kClassPlain];
  if ((int32_t)blockIdx.x >= n) return;  // one wavefront per block; as in sw_fast_kernel
  const int32_t base = s->class_base[kClassPlain];
  for (;;) {
    const int w = from_lane0(atomicAdd(&s->class_next[kClassPlain], k == 0 ? 1 : 0));
    if (w >= n) break;
    const SwTask* tp = &b.tasks[base + w];
    SwTask task;
    task.pair = uni(tp->pair);
    task.flags = uni(tp->flags);
    task.row_start = uni64(tp->row_start);
    task.col_start = uni64(tp->col_start);
    task.row_len = uni(tp->row_len);
    task.col_len = uni(tp->col_len);
    const uint64_t pair_best = plain_sweep(b, task, k, 0, false, 0, 0, bnd);
    PairEnd e;
    if (pair_best == 0) {
      e = PairEnd{0, -1, 0, kEndWide};
    } else {
      key_decode(pair_best, e.score, e.i, e.j);
      e.flags = kEndWide;
    }
    if (b.score2 && !(task.flags & kTaskReverse)) {
      // score2 (rows are SSW-ref): a second sweep with the query padding of
      // the pass SSW would use for this score, masked around the end row
      const bool word = e.score + b.sc.bias >= 255;
      const int32_t mask = task.row_len / 2 > 15 ? task.row_len / 2 : 15;
      const int32_t lo = e.i - mask, hi = e.i + mask - (word ? 1 : 0);
      const int npad = ssw_query_pad(task.col_len, e.score, b.sc.bias);
      const int32_t v = (int32_t)plain_sweep(b, task, k, npad, true, lo, hi, bnd);
      if (k == 0) b.s2[task.pair] = v;
    }
    if (k == 0) {
      if (task.flags & kTaskReverse) {
        b.rev[task.pair] = e;
      } else {
        b.fwd[task.pair] = e;
      }
    }
  }
}

// ----------------------------------------------------------------------------
// results in klign's terms
// ----------------------------------------------------------------------------
__global__ void finalize_kernel(BatchView b) {
  for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < b.npairs; p += (int64_t)gridDim.x * blockDim.x) {
    const PairEnd f = b.fwd[p];
    int32_t rb = -1, re = -1, qb = 0, qe = 0, score = 0;
    if (f.score > 0) {
      const PairEnd r = b.rev[p];
      score = f.score;
      re = f.i;
      qe = f.j;
      rb = f.i - r.i;
      qb = f.j - r.j;
      if (r.score != f.score) atomicAdd(&b.errors[0], 1);
    }
    AlnResult o;
    o.score = score;
    if (b.sc.read_is_ref) {
      o.read_begin = rb;
      o.read_end = re;
      o.ctg_begin = qb;
      o.ctg_end = qe;
    } else {
      o.read_begin = qb;
      o.read_end = qe;
      o.ctg_begin = rb;
      o.ctg_end = re;
    }
    b.results[p] = o;
  }
}

// ----------------------------------------------------------------------------
// host-side launchers
// ----------------------------------------------------------------------------
static int grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (int)g;
}

hipError_t launch_clear_job(const BatchView& b, hipStream_t st) {
  hipLaunchKernelGGL(clear_job_kernel, dim3(16), dim3(256), 0, st, b);
  return hipGetLastError();
}

hipError_t launch_unpack(const BatchView& b, hipStream_t st) {
  hipLaunchKernelGGL(unpack_kernel, dim3(grid_for(b.total >> 5, 256)), dim3(256), 0, st, b);
  return hipGetLastError();
}

hipError_t launch_extract_windows(const BatchView& b, hipStream_t st) {
  if (b.win_ctg && b.npairs > 0)
    hipLaunchKernelGGL(extract_windows_kernel, dim3(grid_for(b.npairs, 4)), dim3(256), 0, st, b);
  return hipGetLastError();
}


hipError_t launch_schedule(const BatchView& b, int pass, hipStream_t st) {
  const BatchView v = with_pass(b, pass);  // the pass's own scheduling state, cleared by clear_job_kernel
  {
    const int64_t per_block = 256 * kClassifyPerThread;
    const unsigned nblk = (unsigned)std::max<int64_t>(1, (b.npairs + per_block - 1) / per_block);
    hipLaunchKernelGGL(classify_kernel, dim3(nblk), dim3(256), 0, st, v, pass);
  }
  hipLaunchKernelGGL(schedule_kernel, dim3(1), dim3(64), 0, st, v);
  hipLaunchKernelGGL(scatter_kernel, dim3(grid_for(b.npairs, 256)), dim3(256), 0, st, v, pass);
  return hipGetLastError();
}

// Forward pass, 16-lane class P (rows <= 32P, fast16_class_of): G lanes of
// 16P/G packed registers per pair (the fifth class, rows <= 152: 19), the
// fewest lanes whose registers fit two wavefronts per SIMD.  Fewer lanes
// per pair mean fewer pipeline fill steps per pair (2G - 1) and the per-step
// exchange spread over more cells: on the length-skewed 4M-pair batch the
// 150-row class (4 x 20) runs 11% faster than as 16 x 5, the 250-row class
// (8 x 16) 7% faster than as 16 x 8.  A class too small to give every SIMD
// two such wavefronts keeps 16-lane groups (four times the wavefronts): a
// 10k-pair batch of 150-row reads ran its forward pass in 0.35 ms as 4 x 20
// and 0.23 ms as 16 x 5.  The reverse pass's rectangles (shorter and less
// even) keep 16-lane groups, which measured faster there.
// A class whose rows all fit the 2G * kTrim rows of its trimmed shape runs
// that shape: 100/101-base reads as 4 x 13 (104 rows) instead of 4 x 16 (128).
// The 64-lane classes of rows above 256 run forward as 16 lanes of 4P
// registers (P = 3..5) or 32 lanes of 2P (P = 6..8): a pipeline fill of 31
// or 63 steps in place of 127, the same rows.
FastShape fast_shape(int group, int P, bool forward, int64_t pairs, int cus, int max_rows) {
  if (!forward || (group == 64 && P <= 2)) return FastShape{group, P};
  if (group == 64) {
    const int G = P <= 5 ? 16 : 32;
    if (pairs * G < (int64_t)64 * 2 * 4 * cus) return FastShape{64, P};
    return FastShape{G, 64 * P / G};
  }
  static constexpr int kLanes[kMaxP] = {2, 2, 4, 4, 4, 8, 8, 8};
  static constexpr int kRegs[kMaxP] = {8, 16, 12, 16, 19, 12, 14, 16};
  static constexpr int kTrim[kMaxP] = {0, 0, 0, 13, 0, 0, 0, 0};
  const int G = kLanes[P - 1];
  if (pairs * G < (int64_t)64 * 2 * 4 * cus) return FastShape{16, P};  // 2 wavefronts on each of 4 SIMDs
  if (kTrim[P - 1] > 0 && max_rows <= 2 * G * kTrim[P - 1]) return FastShape{G, kTrim[P - 1]};
  return FastShape{G, kRegs[P - 1]};
}

hipError_t launch_fast(const BatchView& b, FastShape s, int grid, hipStream_t st) {
  switch (s.lanes) {
    case 2: return launch_fast_2_8(b, s.regs, grid, st);
    case 4:
      return s.regs <= 13   ? launch_fast_4_12(b, s.regs, grid, st)
             : s.regs <= 16 ? launch_fast_4_16(b, s.regs, grid, st)
                            : launch_fast_4_19(b, s.regs, grid, st);
    case 8: return s.regs <= 12 ? launch_fast_8_12(b, s.regs, grid, st) : launch_fast_8_14(b, s.regs, grid, st);
    case 16:
      return s.regs <= 4    ? launch_fast_16_1(b, s.regs, grid, st)
             : s.regs <= 8  ? launch_fast_16_5(b, s.regs, grid, st)
             : s.regs <= 12 ? launch_fast_16_12(b, s.regs, grid, st)
             : s.regs <= 16 ? launch_fast_16_16(b, s.regs, grid, st)
                            : launch_fast_16_20(b, s.regs, grid, st);
    case 32: return s.regs <= 12 ? launch_fast_32_12(b, s.regs, grid, st) : launch_fast_32_14(b, s.regs, grid, st);
    case 64: return s.regs <= 4 ? launch_fast_64_1(b, s.regs, grid, st) : launch_fast_64_5(b, s.regs, grid, st);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_band(const BatchView& b, int P, int grid, hipStream_t st) {
  switch (P) {
    case 0: hipLaunchKernelGGL(sw_band_all_kernel, dim3(grid), dim3(256), 0, st, b); break;
    case 1: hipLaunchKernelGGL(sw_band_kernel<1>, dim3(grid), dim3(256), 0, st, b); break;
    case 2: hipLaunchKernelGGL(sw_band_kernel<2>, dim3(grid), dim3(256), 0, st, b); break;
    case 4: hipLaunchKernelGGL(sw_band_kernel<4>, dim3(grid), dim3(256), 0, st, b); break;
    case 8: hipLaunchKernelGGL(sw_band_kernel<8>, dim3(grid), dim3(256), 0, st, b); break;
    case 16: hipLaunchKernelGGL((sw_band_kernel<4, 4>), dim3(grid), dim3(256), 0, st, b); break;
    case kBandDiag: hipLaunchKernelGGL(sw_diag_kernel, dim3(grid), dim3(256), 0, st, b); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_plain(const BatchView& b, int grid, int max_cols, hipStream_t st) {
  const size_t lds = (size_t)(max_cols > 0 ? max_cols : 1) * sizeof(int2);
  hipLaunchKernelGGL(sw_plain_kernel, dim3(grid), dim3(64), lds, st, b);
  return hipGetLastError();
}

hipError_t launch_finalize(const BatchView& b, hipStream_t st) {
  hipLaunchKernelGGL(finalize_kernel, dim3(grid_for(b.npairs, 256)), dim3(256), 0, st, b);
  return hipGetLastError();
}

template <class K>
static hipError_t occ(K kernel, int* blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, kernel, 256, 0);
}

hipError_t fast_kernel_occupancy(FastShape s, int* blocks_per_cu) {
  switch (s.lanes) {
    case 2: return occupancy_fast_2_8(s.regs, blocks_per_cu);
    case 4:
      return s.regs <= 13   ? occupancy_fast_4_12(s.regs, blocks_per_cu)
             : s.regs <= 16 ? occupancy_fast_4_16(s.regs, blocks_per_cu)
                            : occupancy_fast_4_19(s.regs, blocks_per_cu);
    case 8: return s.regs <= 12 ? occupancy_fast_8_12(s.regs, blocks_per_cu) : occupancy_fast_8_14(s.regs, blocks_per_cu);
    case 16:
      return s.regs <= 4    ? occupancy_fast_16_1(s.regs, blocks_per_cu)
             : s.regs <= 8  ? occupancy_fast_16_5(s.regs, blocks_per_cu)
             : s.regs <= 12 ? occupancy_fast_16_12(s.regs, blocks_per_cu)
             : s.regs <= 16 ? occupancy_fast_16_16(s.regs, blocks_per_cu)
                            : occupancy_fast_16_20(s.regs, blocks_per_cu);
    case 32: return s.regs <= 12 ? occupancy_fast_32_12(s.regs, blocks_per_cu) : occupancy_fast_32_14(s.regs, blocks_per_cu);
    case 64: return s.regs <= 4 ? occupancy_fast_64_1(s.regs, blocks_per_cu) : occupancy_fast_64_5(s.regs, blocks_per_cu);
  }
  return hipErrorInvalidValue;
}

hipError_t band_kernel_occupancy(int P, int* blocks_per_cu) {
  switch (P) {
    case 0: return occ(sw_band_all_kernel, blocks_per_cu);
    case 1: return occ(sw_band_kernel<1>, blocks_per_cu);
    case 2: return occ(sw_band_kernel<2>, blocks_per_cu);
    case 4: return occ(sw_band_kernel<4>, blocks_per_cu);
    case 8: return occ(sw_band_kernel<8>, blocks_per_cu);
    case 16: return occ(sw_band_kernel<4, 4>, blocks_per_cu);
    case kBandDiag: return occ(sw_diag_kernel, blocks_per_cu);
  }
  return hipErrorInvalidValue;
}

#else  // MHMSW_FAST_PART: one part of the packed kernels' instantiations

template <int G, int Pk>
static bool launch_fast_one(int P, const BatchView& b, int grid, hipStream_t st, hipError_t& e) {
  if (P != Pk) return false;
  hipLaunchKernelGGL((sw_fast_kernel<G, Pk>), dim3(grid), dim3(256), 0, st, b);
  e = hipGetLastError();
  return true;
}

template <int G, int Pk>
static bool occupancy_fast_one(int P, int* blocks_per_cu, hipError_t& e) {
  if (P != Pk) return false;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, sw_fast_kernel<G, Pk>, 256, 0);
  return true;
}

template <int G, int... Ps>
static hipError_t launch_fast_list(const BatchView& b, int P, int grid, hipStream_t st) {
  hipError_t e = hipErrorInvalidValue;
  (void)(launch_fast_one<G, Ps>(P, b, grid, st, e) || ...);
  return e;
}

template <int G, int... Ps>
static hipError_t occupancy_fast_list(int P, int* blocks_per_cu) {
  hipError_t e = hipErrorInvalidValue;
  (void)(occupancy_fast_one<G, Ps>(P, blocks_per_cu, e) || ...);
  return e;
}

#define MHMSW_PART_NAME_(f, g, p) f##_##g##_##p
#define MHMSW_PART_NAME(f, g, p) MHMSW_PART_NAME_(f, g, p)
hipError_t MHMSW_PART_NAME(launch_fast, MHMSW_FAST_G, MHMSW_FAST_P0)(const BatchView& b, int P, int grid,
                                                                     hipStream_t st) {
  return launch_fast_list<MHMSW_FAST_G, MHMSW_FAST_PS>(b, P, grid, st);
}
hipError_t MHMSW_PART_NAME(occupancy_fast, MHMSW_FAST_G, MHMSW_FAST_P0)(int P, int* blocks_per_cu) {
  return occupancy_fast_list<MHMSW_FAST_G, MHMSW_FAST_PS>(P, blocks_per_cu);
}
#undef MHMSW_PART_NAME
#undef MHMSW_PART_NAME_

#endif  // MHMSW_FAST_PART

}  // namespace mhmsw
