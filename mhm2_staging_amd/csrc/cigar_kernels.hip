// HIP kernels of the CIGAR pass: SSW's CIGAR and mismatch count for every
// pair of a finished job, as klign gets them when it reports CIGARs
// (src/ssw/ssw_core.cpp:1180-1192 banded_sw, src/ssw/ssw.cpp:57-93 and
// :129-254; klign.cpp:589-599 keeps such batches off ADEPT's GPU path).
// The per-pair arithmetic of the narrow bands lives in cigar_core.h.
//
//  * cigar_classify_kernel: the rectangle of each pair from its result;
//    trivial pairs (empty query, distance filter, score 0) written at once;
//    the rest appended to the list of their band class (half-width
//    |columns - rows| + 1: 1..7 narrow, wider -- or taller than
//    kCigNarrowMaxRows -- wide), block-aggregated.
//  * Narrow bands, one lane per pair, the band row (2*BW+1 diagonals of H
//    and E) in registers, 4 direction bits per cell stored in the
//    wavefront's scratch slot in blocks of 8 rows per lane (16-byte stores
//    and loads): cigar_bw1_kernel for half-width 1 (most pairs),
//    cigar_narrow_kernel for half-widths 2..7.  A pair whose band misses the
//    score doubles its half-width in the wide kernel.  A counting walk over
//    the directions ('=' / 'X' from the
//    equality bits) that keeps up to 16 ops besides the clips in registers,
//    one atomic per wavefront to reserve the ops, and a writing walk only
//    for longer CIGARs.
//  * cigar_wide_kernel: one wavefront per pair.  Lane l holds diagonals
//    l + 64c; E and the diagonal come from the previous row (lane shifts),
//    F from a prefix maximum over the row: with gap_open >= gap_ext,
//    F(d) = max(0, max_{k<d} (H0(k) - gO - (d-1-k)*gE)) where H0 = max(diag + s,
//    E, 0), because a gap extended from an F cell costs more than opening it
//    from that cell's H0.  Directions as one byte per cell; the traceback is
//    walked by all 64 lanes in step over blocks of rows copied to LDS.  Bands
//    wider than 1024 diagonals take a serial path on lane 0.
//  Scratch that a pass finds too small is counted; the host reruns the pass
//  with what it asked for.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cigar_core.h"
#include "sw_launch.h"

namespace mhmsw {

// layout of CigarView::list_count: entries of each class, then the work
// counters of the launches (half-width 1, half-widths 2..7, wide)
constexpr int kCigWork = kCigClasses;
constexpr int kWideBlkBytes = 8192;                  // LDS rows of a wide traceback
constexpr int kWideMaxChunks = 16;                   // wide kernel: up to 1024 diagonals in lanes

namespace {

__device__ __forceinline__ int32_t lane0(int32_t v) { return __builtin_amdgcn_readlane(v, 0); }

__device__ __forceinline__ int64_t lane0_64(int64_t v) {
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)v, 0);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)((uint64_t)v >> 32), 0);
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ int32_t* list_of(const CigarView& v, int cls) {
  return v.lists + (int64_t)cls * v.npairs;
}

// Rectangle of pair p in SSW naming (rows: SSW-query, columns: SSW-ref).
__device__ __forceinline__ CigarGeo pair_geo(const CigarView& v, int64_t p) {
  const AlnResult r = v.results[p];
  CigarGeo g;
  const int64_t rd = v.seq_base[2 * p], ct = v.seq_base[2 * p + 1];
  if (v.sc.read_is_ref) {
    g.rbase = rd;
    g.qbase = ct;
    g.rb = r.read_begin;
    g.re = r.read_end;
    g.qb = r.ctg_begin;
    g.qe = r.ctg_end;
    g.n = v.seq_len[2 * p + 1];
  } else {
    g.rbase = ct;
    g.qbase = rd;
    g.rb = r.ctg_begin;
    g.re = r.ctg_end;
    g.qb = r.read_begin;
    g.qe = r.read_end;
    g.n = v.seq_len[2 * p];
  }
  g.score = r.score;
  return g;
}

// Exclusive offset of this lane's n inside the wavefront, the wavefront's
// base from one atomic on ctr; every lane of the wavefront must call it.
__device__ __forceinline__ int64_t wave_reserve(unsigned long long* ctr, int32_t n) {
  const int lane = threadIdx.x & 63;
  int32_t x = n;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int32_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  const int32_t total = __builtin_amdgcn_readlane(x, 63);
  unsigned long long base = 0;
  if (lane == 0 && total > 0) base = atomicAdd(ctr, (unsigned long long)total);
  return lane0_64((int64_t)base) + (x - n);
}

// Work index of the wavefront: one atomic from lane 0, broadcast.
__device__ __forceinline__ int32_t next_work(int32_t* ctr) {
  int32_t w = 0;
  if ((threadIdx.x & 63) == 0) w = atomicAdd(ctr, 1);
  return lane0(w);
}

// Atomic add from lane 0 only, result broadcast.
__device__ __forceinline__ int64_t lane0_add(unsigned long long* ctr, int64_t n) {
  unsigned long long r = 0;
  if ((threadIdx.x & 63) == 0) r = atomicAdd(ctr, (unsigned long long)n);
  return lane0_64((int64_t)r);
}

// Whether SSW-query[qb] equals SSW-ref[rb] (the traceback's last M).
__device__ __forceinline__ bool first_eq(const CigarView& v, const CigarGeo& g) {
  return nib_at(v.codes, g.qbase + g.qb) == nib_at(v.codes, g.rbase + g.rb);
}

// A pair whose band of half-width BW missed the score retries with twice
// the half-width.  Half-width 1 retries go to the half-width-2 list, which
// the narrow kernel (launched after this one) reads; later retries go to the
// wide class (one wavefront per pair: the few pairs that retry from the
// narrow kernel would otherwise need another narrow launch).  Rare: per-lane
// atomic.
template <int BW>
__device__ __forceinline__ void retry(const CigarView& v, int64_t p) {
  v.band[p] = 2 * BW;
  const int cls = BW == 1 ? 1 : kCigClassWide;
  const int32_t at = atomicAdd(&v.list_count[cls], 1);
  list_of(v, cls)[at] = (int32_t)p;
}

// One narrow pair per lane (p < 0: idle lane); all lanes call it.  DS is
// the words per row of the wavefront's slot.  The counting walk keeps the
// CIGAR in registers when it has at most 16 ops between the clips; the
// wavefront then reserves its ops with one atomic, and only pairs with more
// ops (or whose traceback stops off column 0) walk again to write them.
template <int BW, int DS>
__device__ __forceinline__ void narrow_pair(const CigarView& v, int64_t p, uint32_t* dirs) {
  constexpr int DW = DS;
  bool done = false, eq00 = false, queued = false;
  int32_t jf = -1, count = 0, mism = 0;
  CigarGeo g;
  CigarEmitQ q;
  if (p >= 0) {
    g = pair_geo(v, p);
    const int32_t corner = cigar_band_dp<BW, DS>(v.codes, g, v.sc.match, v.sc.mismatch, v.sc.ambiguity,
                                                 v.sc.gap_open, v.sc.gap_ext, dirs);
    if (corner == g.score) {
      done = true;
      eq00 = first_eq(v, g);
      jf = cigar_walk_rows<DW>(g, BW, dirs, eq00, q);
      if (jf > 0) {  // traceback stopped off column 0: count again comparing shifted codes
        CigarEmit em;
        em.out = nullptr;
        em.nops = 0;
        jf = cigar_walk(v.codes, g, BW, dirs, DW, true, jf, eq00, em);
        count = em.count;
        mism = em.mism;
      } else {
        count = q.count;
        mism = q.mism;
        queued = q.fits();
      }
    } else {
      retry<BW>(v, p);
    }
  }
  const bool ok = done && jf >= 0;
  const int64_t base = wave_reserve(&v.used[0], ok ? count : 0);
  if (!done) return;
  if (!ok) {
    atomicAdd(&v.used[3], 1ull);  // traceback left the band: reported as an internal error
    v.n_ops[p] = -3;
    v.mism[p] = 0;
    v.off[p] = 0;
    return;
  }
  const bool fits = base + count <= v.ops_cap;
  if (fits && count > 0) {
    uint32_t* out = v.ops + base;
    if (queued) {
      q.write(out);
    } else {
      CigarEmit w;
      w.out = out;
      w.nops = count;
      if (jf == 0) {
        cigar_walk_rows<DW>(g, BW, dirs, eq00, w);
      } else {
        cigar_walk(v.codes, g, BW, dirs, DW, true, jf, eq00, w);
      }
    }
  }
  v.n_ops[p] = fits ? count : -1;
  v.mism[p] = mism;
  v.off[p] = base;
}

// The lane's first block in its wavefront's scratch slot (rows of the
// tallest narrow rectangle, rounded up to 8, x DW words x 64 lanes, blocked
// layout), or null when the scratch has no slot for the wavefront.
__device__ __forceinline__ uint32_t* narrow_slot(const CigarView& v, int DW) {
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t slot_dwords = (int64_t)((lane0(v.max_rows[0]) + 7) & ~7) * DW * 64;
  if (slot_dwords <= 0 || (wave + 1) * slot_dwords > v.dirs_dwords) return nullptr;
  return v.dirs + wave * slot_dwords + (int64_t)lane * 8 * DW;
}

}  // namespace

__global__ void cigar_clear_kernel(CigarView v) {
  const int t = threadIdx.x;
  if (t < kCigWork + 3) v.list_count[t] = 0;
  if (t < 8) v.used[t] = 0ull;
  if (t == 0) v.max_rows[0] = 0;
}

// Each 256-thread block classifies chunks of kCigClassifyPerThread*256
// pairs: ranks inside a class from LDS atomics, then one global atomic per
// class used by the chunk reserves its range of the class list.
constexpr int kCigClassifyPerThread = 8;
constexpr int kCigChunk = 256 * kCigClassifyPerThread;

__global__ __launch_bounds__(256) void cigar_classify_kernel(CigarView v) {
  __shared__ int32_t cnt[kCigClasses], base[kCigClasses], rows_max;
  const int t = threadIdx.x;
  for (int64_t chunk = blockIdx.x; chunk * kCigChunk < v.npairs; chunk += gridDim.x) {
    if (t < kCigClasses) cnt[t] = 0;
    if (t == 0) rows_max = 0;
    __syncthreads();
    int cls[kCigClassifyPerThread];
    int32_t rank[kCigClassifyPerThread];
#pragma unroll
    for (int k = 0; k < kCigClassifyPerThread; k++) {
      const int64_t p = chunk * kCigChunk + k * 256 + t;
      const bool live = p < v.npairs;
      cls[k] = -1;
      rank[k] = 0;
      int32_t trivial_n = 0;
      CigarGeo g;
      CigarEmit em;
      em.out = nullptr;
      em.nops = 0;
      if (live) {
        g = pair_geo(v, p);
        if (cigar_trivial(v.codes, g, em)) {
          trivial_n = em.count;
        } else {
          const int32_t bw = (g.rl() > g.ql() ? g.rl() - g.ql() : g.ql() - g.rl()) + 1;
          cls[k] = cigar_class(bw, g.ql());
          v.band[p] = bw;
          rank[k] = atomicAdd(&cnt[cls[k]], 1);
          if (cls[k] != kCigClassWide) atomicMax(&rows_max, g.ql());
        }
      }
      // trivial pairs (empty query, distance filter, score 0): reserve and write now
      const bool triv = live && cls[k] < 0;
      const int64_t at = wave_reserve(&v.used[0], triv ? trivial_n : 0);
      if (triv) {
        const bool fits = at + trivial_n <= v.ops_cap;
        if (fits && trivial_n > 0) {
          CigarEmit w;
          w.out = v.ops + at;
          w.nops = trivial_n;
          cigar_trivial(v.codes, g, w);
        }
        v.n_ops[p] = fits ? trivial_n : -1;
        v.mism[p] = em.mism;
        v.off[p] = at;
      }
    }
    __syncthreads();
    if (t < kCigClasses && cnt[t] > 0) base[t] = atomicAdd(&v.list_count[t], cnt[t]);
    if (t == 0 && rows_max > 0) atomicMax(&v.max_rows[0], rows_max);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kCigClassifyPerThread; k++) {
      if (cls[k] >= 0) {
        const int64_t p = chunk * kCigChunk + k * 256 + t;
        list_of(v, cls[k])[base[cls[k]] + rank[k]] = (int32_t)p;
      }
    }
    __syncthreads();
  }
}

// Half-width 1.
__global__ __launch_bounds__(256) void cigar_bw1_kernel(CigarView v) {
  uint32_t* dirs = narrow_slot(v, 1);
  if (!dirs) return;
  const int lane = threadIdx.x & 63;
  const int32_t count = lane0(v.list_count[0]);
  const int32_t* list = list_of(v, 0);
  for (;;) {
    const int32_t w = next_work(&v.list_count[kCigWork + 0]);
    if ((int64_t)w * 64 >= count) break;
    const int32_t idx = w * 64 + lane;
    narrow_pair<1, 1>(v, idx < count ? list[idx] : -1, dirs);
  }
}

// Half-widths 2..7, one class per wavefront iteration.
__global__ __launch_bounds__(256) void cigar_narrow_kernel(CigarView v) {
  uint32_t* dirs = narrow_slot(v, 2);
  if (!dirs) return;
  const int lane = threadIdx.x & 63;
  int32_t items[kCigExactMaxBw], cnt[kCigExactMaxBw];
  int32_t total = 0;
  items[0] = 0;
  cnt[0] = 0;
#pragma unroll
  for (int c = 1; c < kCigExactMaxBw; c++) {
    cnt[c] = lane0(v.list_count[c]);
    total += (cnt[c] + 63) / 64;
    items[c] = total;  // items of classes 1..c
  }
  for (;;) {
    const int32_t w = next_work(&v.list_count[kCigWork + 1]);
    if (w >= total) break;
    int c = 1, first = 0;
#pragma unroll
    for (int k = 1; k < kCigExactMaxBw - 1; k++) {
      if (w >= items[k]) {
        c = k + 1;
        first = items[k];
      }
    }
    const int32_t idx = (w - first) * 64 + lane;
    int32_t n = cnt[1];
#pragma unroll
    for (int k = 2; k < kCigExactMaxBw; k++) n = c == k ? cnt[k] : n;
    const int64_t p = idx < n ? list_of(v, c)[idx] : -1;
    switch (c) {
      case 1: narrow_pair<2, 2>(v, p, dirs); break;
      case 2: narrow_pair<3, 2>(v, p, dirs); break;
      case 3: narrow_pair<4, 2>(v, p, dirs); break;
      case 4: narrow_pair<5, 2>(v, p, dirs); break;
      case 5: narrow_pair<6, 2>(v, p, dirs); break;
      default: narrow_pair<7, 2>(v, p, dirs); break;
    }
  }
}

// ---------------------------------------------------------------------------
// wide bands: one wavefront per pair
// ---------------------------------------------------------------------------
namespace {

// Lane shifts and scans over the wavefront with DPP (no LDS round trip).
// from_next: lane l gets lane l+1's v, lane 63 gets last.  from_prev: lane l
// gets lane l-1's v, lane 0 gets first.
__device__ __forceinline__ int32_t from_next(int32_t v, int32_t last) {
  return __builtin_amdgcn_update_dpp(last, v, 0x130, 0xF, 0xF, false);  // wave_shl:1
}
__device__ __forceinline__ int32_t from_prev(int32_t v, int32_t first) {
  return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
}
// inclusive maximum over lanes 0..l: row_shr 1, 2, 4, 8 inside rows of 16
// lanes, then row_bcast 15 and 31 across rows
__device__ __forceinline__ int32_t wave_prefix_max(int32_t x, int32_t neg) {
  x = imax(x, __builtin_amdgcn_update_dpp(neg, x, 0x111, 0xF, 0xF, false));
  x = imax(x, __builtin_amdgcn_update_dpp(neg, x, 0x112, 0xF, 0xF, false));
  x = imax(x, __builtin_amdgcn_update_dpp(neg, x, 0x114, 0xF, 0xF, false));
  x = imax(x, __builtin_amdgcn_update_dpp(neg, x, 0x118, 0xF, 0xF, false));
  x = imax(x, __builtin_amdgcn_update_dpp(neg, x, 0x142, 0xA, 0xF, false));
  x = imax(x, __builtin_amdgcn_update_dpp(neg, x, 0x143, 0xC, 0xF, false));
  return x;
}

// Band DP of half-width bw with C chunks of 64 diagonals per lane; writes one
// direction byte per cell at dirs[i * 64C + d].  Returns H of the last cell
// (the same in every lane).  The SSW-ref codes read up to bw + 8 positions
// before and 64C + 8 after the rectangle's columns, inside the slack around
// every sequence of the code buffer for bands of up to 1024 diagonals;
// codes off the rectangle are not used.
template <int C>
__device__ int32_t wide_dp(const CigarView& v, const CigarGeo& g, int32_t bw, uint8_t* dirs) {
  constexpr int32_t NEG = -(1 << 29);
  const int lane = threadIdx.x & 63;
  const int32_t W = 2 * bw + 1, rl = g.rl(), ql = g.ql();
  const int32_t gO = v.sc.gap_open, gE = v.sc.gap_ext;
  int32_t H[C], E[C];
#pragma unroll
  for (int c = 0; c < C; c++) {
    H[c] = 0;
    E[c] = 0;
  }
  // codes in blocks of 8 rows, loaded one block ahead: the SSW-query code of
  // each row, and for each chunk the SSW-ref codes of the lane's diagonal
  // (column i - bw + d moves one to the right per row)
  const int64_t qa = g.qbase + g.qb, ra = g.rbase + g.rb - bw + lane;
  uint32_t qn = load8(v.codes, qa);
  uint32_t rn[C];
#pragma unroll
  for (int c = 0; c < C; c++) rn[c] = load8(v.codes, ra + 64 * c);
  for (int32_t i0 = 0; i0 < ql; i0 += 8) {
    const uint32_t qw = qn;
    uint32_t rw[C];
#pragma unroll
    for (int c = 0; c < C; c++) rw[c] = rn[c];
    qn = load8(v.codes, qa + i0 + 8);
#pragma unroll
    for (int c = 0; c < C; c++) rn[c] = load8(v.codes, ra + 64 * c + i0 + 8);
    for (int r = 0; r < 8 && i0 + r < ql; r++) {
      const int32_t i = i0 + r;
      const uint32_t qc = (qw >> (4 * r)) & 0xFu;
      uint32_t tlo, thi;
      subst_table(qc, v.sc.match, v.sc.mismatch, v.sc.ambiguity, tlo, thi);
      int32_t carry = NEG;         // max X over earlier chunks
      int32_t hl63 = 0, fl63 = 0;  // H and F of the diagonal left of this chunk's lane 0
      uint8_t* row = dirs + (int64_t)i * (64 * C);
#pragma unroll
      for (int c = 0; c < C; c++) {
        const int32_t d = lane + 64 * c;
        const int32_t j = i - bw + d;
        const bool valid = d < W;
        const uint32_t rc = (valid && j >= 0 && j < rl) ? (rw[c] >> (4 * r)) & 0xFu : kCigKill;
        // E from (i-1, d+1): the next lane, or lane 0 of the next chunk
        const int32_t hnext = c + 1 < C ? __builtin_amdgcn_readlane(H[(c + 1) % C], 0) : 0;
        const int32_t enext = c + 1 < C ? __builtin_amdgcn_readlane(E[(c + 1) % C], 0) : 0;
        int32_t hu = from_next(H[c], hnext), eu = from_next(E[c], enext);
        if (d + 1 >= W) {
          hu = 0;
          eu = 0;
        }
        const int32_t eo = hu - gO, ee = eu - gE;
        const int32_t e = imax3(eo, ee, 0);
        const int32_t t2 = H[c] + subst_lookup(tlo, thi, rc);
        const int32_t h0 = imax3(e, t2, 0);
        // F(d) = max(0, max_{k<d} X(k) - (d-1)*gE), X(k) = H0(k) - gO + k*gE
        const int32_t incl = wave_prefix_max(valid ? h0 - gO + d * gE : NEG, NEG);
        const int32_t excl = from_prev(incl, NEG);
        const int32_t f = imax(0, imax(excl, carry) - (d - 1) * gE);
        carry = imax(carry, __builtin_amdgcn_readlane(incl, 63));
        const int32_t h = imax(h0, f);
        // H and F of (i, d-1) for F's open/extend bit
        const int32_t hl = from_prev(h, hl63), fl = from_prev(f, fl63);
        hl63 = __builtin_amdgcn_readlane(h, 63);
        fl63 = __builtin_amdgcn_readlane(f, 63);
        const int32_t t1 = imax(e, f);
        if (valid) row[d] = (uint8_t)cell_bits(t1, t2, e, f, rc == qc, eo > ee, hl - gO > fl - gE);
        H[c] = valid ? h : 0;
        E[c] = valid ? e : 0;
      }
    }
  }
  const int32_t dc = (rl - 1) - (ql - 1) + bw;
  int32_t corner = -1;
#pragma unroll
  for (int c = 0; c < C; c++) {
    const int32_t x = __builtin_amdgcn_readlane(H[c], dc & 63);
    corner = (dc >> 6) == c ? x : corner;
  }
  return corner;
}

// Traceback of a wide band (bytes, Wp per row), walked by all lanes in step
// over blocks of rows copied to LDS; ops written from lane 0.  As cigar_walk.
template <class Emit>
__device__ int32_t wide_walk(const CigarView& v, const CigarGeo& g, int32_t bw, int32_t Wp, const uint8_t* dirs,
                             uint32_t* blk, int32_t shift, bool eq00, Emit& em) {
  const int lane = threadIdx.x & 63;
  const int32_t R = kWideBlkBytes / Wp;
  const uint8_t* b8 = reinterpret_cast<const uint8_t*>(blk);
  NibCursor rcur, qcur;
  em.push(kOpS, (uint32_t)(g.n - 1 - g.qe));
  int32_t i = g.ql() - 1, j = g.rl() - 1;
  int32_t d = j - i + bw;
  int state = 2;
  int32_t lo = i + 1;
  while (i > 0) {
    if (d < 0 || d > 2 * bw || j < 0) return -1;
    if (i < lo) {
      lo = i - R + 1 > 0 ? i - R + 1 : 0;
      const int32_t ndw = (i - lo + 1) * (Wp >> 2);
      const uint32_t* src = reinterpret_cast<const uint32_t*>(dirs + (int64_t)lo * Wp);
      __syncthreads();
      for (int32_t k = lane; k < ndw; k += 64) blk[k] = src[k];
      __syncthreads();
    }
    uint32_t op = step_op(b8[(i - lo) * Wp + d] & 0xFu, state);
    if (op == kOpEq || op == kOpX) {
      if (shift != 0) {
        const uint32_t a = qcur.get(v.codes, g.qbase + g.qb + i);
        const uint32_t b = rcur.get(v.codes, g.rbase + g.rb + j - shift);
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
  if (shift != 0) last_eq = nib_at(v.codes, g.qbase + g.qb) == nib_at(v.codes, g.rbase + g.rb + j - shift);
  em.push(last_eq ? kOpEq : kOpX, 1);
  em.push(kOpS, (uint32_t)g.qb);
  em.flush();
  return j;
}

}  // namespace

__global__ __launch_bounds__(64) void cigar_wide_kernel(CigarView v) {
  __shared__ uint32_t blk[kWideBlkBytes / 4];
  const int lane = threadIdx.x & 63;
  const int32_t count = lane0(v.list_count[kCigClassWide]);
  const int32_t* list = list_of(v, kCigClassWide);
  for (;;) {
    const int32_t w = next_work(&v.list_count[kCigWork + 2]);
    if (w >= count) break;
    const int64_t p = lane0(list[w]);
    const CigarGeo g = pair_geo(v, p);
    const bool eq00 = first_eq(v, g);
    int32_t bw = v.band[p];
    int status = 0;  // 1 done (lanes), 2 done (serial, lane 0), 3 deferred, 4 error
    int32_t Wp = 0, sDW = 0;
    uint32_t* mem = nullptr;
    int32_t corner = -1;
    for (;;) {
      const int32_t W = 2 * bw + 1;
      const int32_t C = (W + 63) / 64;
      const bool lanes = C <= kWideMaxChunks;
      const int32_t Cp = C <= 1 ? 1 : C <= 2 ? 2 : C <= 4 ? 4 : C <= 8 ? 8 : 16;
      Wp = 64 * Cp;
      sDW = (4 * W + 31) / 32;
      const int64_t need = lanes ? ((int64_t)g.ql() * Wp + 3) / 4 : 2 * (int64_t)W + (int64_t)g.ql() * sDW;
      const int64_t at = lane0_add(&v.used[1], need);
      if (at + need > v.gen_dwords) {
        status = 3;
        break;
      }
      mem = v.gen + at;
      if (lanes) {
        uint8_t* dirs = reinterpret_cast<uint8_t*>(mem);
        switch (Cp) {
          case 1: corner = wide_dp<1>(v, g, bw, dirs); break;
          case 2: corner = wide_dp<2>(v, g, bw, dirs); break;
          case 4: corner = wide_dp<4>(v, g, bw, dirs); break;
          case 8: corner = wide_dp<8>(v, g, bw, dirs); break;
          default: corner = wide_dp<16>(v, g, bw, dirs); break;
        }
      } else {
        // bands beyond the lanes: one cell at a time on lane 0
        int32_t cr = -1;
        if (lane == 0)
          cr = cigar_band_dp_any(v.codes, g, bw, v.sc.match, v.sc.mismatch, v.sc.ambiguity, v.sc.gap_open,
                                 v.sc.gap_ext, (int32_t*)mem, (int32_t*)mem + W, mem + 2 * W);
        corner = lane0(cr);
      }
      if (corner == g.score) {
        status = lanes ? 1 : 2;
        break;
      }
      if (bw > g.rl() + g.ql() + 2) {  // a band this wide holds the whole rectangle
        status = 4;
        break;
      }
      bw *= 2;
    }
    if (status == 1) {
      // every lane walks the same traceback; the CIGAR stays in registers
      // when it is short, else a second walk writes it from lane 0
      const uint8_t* dirs = reinterpret_cast<const uint8_t*>(mem);
      CigarEmitQ q;
      int32_t jf = wide_walk(v, g, bw, Wp, dirs, blk, 0, eq00, q);
      int32_t count = q.count, mism = q.mism;
      bool queued = jf == 0 && q.fits();
      if (jf > 0) {
        CigarEmit em;
        em.out = nullptr;
        em.nops = 0;
        jf = wide_walk(v, g, bw, Wp, dirs, blk, jf, eq00, em);
        count = em.count;
        mism = em.mism;
      }
      if (jf < 0) {
        status = 4;
      } else {
        const int64_t base = lane0_add(&v.used[0], count);
        const bool fits = base + count <= v.ops_cap;
        if (fits && count > 0) {
          if (queued) {
            if (lane == 0) q.write(v.ops + base);
          } else {
            CigarEmit wr;
            wr.out = lane == 0 ? v.ops + base : nullptr;
            wr.nops = count;
            wide_walk(v, g, bw, Wp, dirs, blk, jf, eq00, wr);
          }
        }
        if (lane == 0) {
          v.n_ops[p] = fits ? count : -1;
          v.mism[p] = mism;
          v.off[p] = base;
        }
      }
    } else if (status == 2 && lane == 0) {
      const uint32_t* dirs = mem + 2 * (2 * bw + 1);
      CigarEmit em;
      em.out = nullptr;
      em.nops = 0;
      int32_t jf = cigar_walk(v.codes, g, bw, dirs, sDW, false, 0, eq00, em);
      if (jf > 0) {
        em = CigarEmit();
        em.out = nullptr;
        em.nops = 0;
        jf = cigar_walk(v.codes, g, bw, dirs, sDW, false, jf, eq00, em);
      }
      if (jf < 0) {
        atomicAdd(&v.used[3], 1ull);
        v.n_ops[p] = -3;
      } else {
        const int64_t base = (int64_t)atomicAdd(&v.used[0], (unsigned long long)em.count);
        const bool fits = base + em.count <= v.ops_cap;
        if (fits && em.count > 0) {
          CigarEmit wr;
          wr.out = v.ops + base;
          wr.nops = em.count;
          cigar_walk(v.codes, g, bw, dirs, sDW, false, jf, eq00, wr);
        }
        v.n_ops[p] = fits ? em.count : -1;
        v.mism[p] = em.mism;
        v.off[p] = base;
      }
    }
    if (status == 3 && lane == 0) {
      atomicAdd(&v.used[2], 1ull);  // deferred: the host grows the pool and reruns the pass
      v.n_ops[p] = -2;
    }
    if (status == 4 && lane == 0) {
      atomicAdd(&v.used[3], 1ull);
      v.n_ops[p] = -3;
    }
  }
}

static int cig_grid(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (int)g;
}

hipError_t launch_cigar(const CigarView& v, int narrow_waves, int wide_waves, hipStream_t st) {
  hipLaunchKernelGGL(cigar_clear_kernel, dim3(1), dim3(64), 0, st, v);
  if (v.npairs <= 0) return hipGetLastError();
  hipLaunchKernelGGL(cigar_classify_kernel, dim3(cig_grid(v.npairs, kCigChunk)), dim3(256), 0, st, v);
  const int blocks = std::max(1, narrow_waves / 4);
  hipLaunchKernelGGL(cigar_bw1_kernel, dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL(cigar_narrow_kernel, dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL(cigar_wide_kernel, dim3(std::max(1, wide_waves)), dim3(64), 0, st, v);
  return hipGetLastError();
}

}  // namespace mhmsw
