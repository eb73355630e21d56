// Local assembly walks on gfx950: one wavefront per contig side.
//
// The reference (src/gpu_loc_assem/kernel.cpp:391-545) gives each contig a
// warp whose lanes count mers into a string-keyed hash table in global
// memory (MurmurHash of the mer bytes at every probe, kernel.cpp:36-136,
// 197-241), after which one lane walks the table, re-hashing the growing
// walk at every step and checking a second table for repeats
// (kernel.cpp:243-307).  Here, per mer length:
//
//   1. count: the side's mer occurrences are split evenly over the 64 lanes;
//      each lane rolls a polynomial hash along its run of occurrences and
//      inserts into an open-addressing table whose 64-bit slot word holds a
//      32-bit fingerprint and the position of one occurrence (claimed with
//      one 64-bit CAS, so the position is never read before it is written);
//      a fingerprint match is confirmed by comparing the bytes;
//   2. decide: every occupied slot gets its extension (the reference's
//      set_ext pass, kernel.cpp:376-381) and, when the extension is a base,
//      the slot of the mer that follows it (the mer the walk would look up
//      next), found with the rolled hash of the slot's own occurrence;
//   3. walk: one lane follows successor links from the slot of the contig's
//      last mer.  A step reads the slot's info word (extension and successor)
//      and marks the slot visited: a visited slot is a repeat (the
//      reference's repeat table holds exactly the mers of the walk so far,
//      each of which is a slot).
//
// Slot words and counts change through atomics during a phase, so reads of
// them in that phase are relaxed atomic loads.
// A side whose reads fit the workgroup's LDS keeps them there, and a table
// whose info words fit beside them keeps those there too.  Pointers carry
// their address space in their type (side_walk<AS>), so LDS data is read
// with ds_read and not with flat loads, which go through the texture
// address path like global loads.
#include "locassm_launch.h"

namespace mhmla {
namespace {

constexpr int kLanes = 64;
constexpr int kU = 2;           // mer occurrences per lane per count batch
constexpr int kD = 2;           // slots per lane per decide batch
constexpr int kLoadNum = 16;    // table slots per 16 mer occurrences, before a resize
constexpr int kMaxProbes = 96;  // longest probe run in a table below its allocated size

constexpr int kGlobalAS = 1;
constexpr int kLdsAS = 3;
template <int AS> using cu8 = const __attribute__((address_space(AS))) uint8_t;
template <int AS> using cu32 = const __attribute__((address_space(AS))) uint32_t;
using lds_u32 = __attribute__((address_space(3))) uint32_t;
using lds_u16 = __attribute__((address_space(3))) uint16_t;
using glb_u32 = __attribute__((address_space(1))) uint32_t;

extern __shared__ __align__(16) uint8_t lds_side[];  // dynamic LDS: reads, then info words

// A side is one workgroup of one wavefront, and every access to its scratch
// and LDS comes from that wavefront, on one CU: ordering at workgroup scope
// is all the phases need.  (Agent scope would write back and invalidate the
// L2 at every fence.)  Atomics are performed at L2 through the CU's L1, which
// keeps the L1 coherent with them for the workgroup.
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

__device__ __forceinline__ uint64_t ld64(uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Bytes a[0..n) and b[0..n) equal?  Aligned dword loads funnel-shifted with
// v_alignbyte, C dwords per round trip, no branch inside a round; reads up to
// 7 bytes past either window (every buffer read here is padded).
template <int C, int AA, int AB>
__device__ __forceinline__ bool same_window(cu8<AA>* a, cu8<AB>* b, int n) {
  cu32<AA>* pa = (cu32<AA>*)((uintptr_t)a & ~(uintptr_t)3);
  cu32<AB>* pb = (cu32<AB>*)((uintptr_t)b & ~(uintptr_t)3);
  const uint32_t sa = (uint32_t)((uintptr_t)a & 3), sb = (uint32_t)((uintptr_t)b & 3);
  const int nw = (n + 3) >> 2;
  for (int g = 0; g < nw; g += C) {
    uint32_t A[C + 1], B[C + 1];
#pragma unroll
    for (int j = 0; j <= C; j++) {
      const bool in = g + j <= nw;
      A[j] = in ? pa[g + j] : 0u;
      B[j] = in ? pb[g + j] : 0u;
    }
    uint32_t diff = 0;
#pragma unroll
    for (int j = 0; j < C; j++) {
      const int rem = n - 4 * (g + j);
      if (rem > 0) {
        const uint32_t m = rem >= 4 ? 0xFFFFFFFFu : (1u << (8 * rem)) - 1u;
        diff |= (__builtin_amdgcn_alignbyte(A[j + 1], A[j], sa) ^ __builtin_amdgcn_alignbyte(B[j + 1], B[j], sb)) & m;
      }
    }
    if (diff) return false;
  }
  return true;
}

// polynomial hash of s[0..k) (bytes in order), from aligned dword loads
template <int AS>
__device__ __forceinline__ uint64_t hash_window(cu8<AS>* s, int k) {
  cu32<AS>* p = (cu32<AS>*)((uintptr_t)s & ~(uintptr_t)3);
  const uint32_t sh = (uint32_t)((uintptr_t)s & 3);
  const int nw = (k + 3) >> 2;
  uint64_t h = 0;
  uint32_t prev = p[0];
  for (int j = 0; j < nw; j++) {
    const uint32_t nxt = p[j + 1];
    const uint32_t w = __builtin_amdgcn_alignbyte(nxt, prev, sh);
    prev = nxt;
    const int m = min(4, k - 4 * j);
    for (int b = 0; b < m; b++) h = h * kHashBase + ((w >> (8 * b)) & 0xFFu);
  }
  return h;
}

// One side's data; RAS is the address space of its read bases.
template <int RAS>
struct Side {
  cu8<RAS>* seq;        // the side's read bases
  const uint8_t* qual;  // the side's read quality bytes (global)
  uint64_t* key;
  uint32_t* cnt;        // per slot: 4 words of 16-bit vote pairs, or 8 words (wide)
  glb_u32* info_g;      // info words in the scratch
  lds_u32* info_l;      // info words in LDS, when info_lds
  bool info_lds;
  int64_t* roff;        // offsets of the side's reads
  int32_t* rlen;
  uint32_t* cum;        // occurrences before read i, at the current mer length
  uint8_t* walk[2];
  uint8_t* smer;        // the side's last mer (reverse-complemented on the left)
  int cap;              // slots in use at the current mer length
  int cap_alloc;        // slots the scratch holds
  int nr;
  bool wide;            // 32-bit vote counters (sides with 65535 mers or more)

  __device__ cu8<RAS>* occ(uint64_t w) const { return seq + ((uint32_t)w & kOffMask); }
  __device__ uint32_t info_get(int s) const { return info_lds ? info_l[s] : info_g[s]; }
  __device__ void info_put(int s, uint32_t x) const {
    if (info_lds) info_l[s] = x;
    else info_g[s] = x;
  }
};

// First probe slot and fingerprint of a mer with rolled hash h.
template <int RAS>
__device__ __forceinline__ int home_slot(const Side<RAS>& S, uint64_t h, uint32_t* fp) {
  const uint64_t m = mix64(h);
  *fp = (uint32_t)(m >> 32);
  return (int)(((uint64_t)(uint32_t)m * (uint64_t)S.cap) >> 32);
}

// Probe from slot s for the slot whose mer equals the bytes 'same' accepts,
// at most 'probes' slots; insert 'mine' (fingerprint | byte offset of the
// occurrence) into the first empty slot when insert is set.
// Returns the slot, -1 when absent and not inserted, -2 if the table is full.
template <int RAS, class Same>
__device__ int probe_from(const Side<RAS>& S, int s, int probes, uint32_t fp, uint64_t mine, bool insert, bool* won,
                          Same same) {
  for (; probes > 0; probes--) {
    // read first: most probes find an occupied slot (a mer seen before), and
    // loads of one hot slot by many lanes do not serialize as atomics do
    uint64_t w = ld64(&S.key[s]);
    if (w == kEmpty) {
      if (!insert) return -1;
      w = atomicCAS((unsigned long long*)&S.key[s], (unsigned long long)kEmpty, (unsigned long long)mine);
      if (w == kEmpty) {
        *won = true;
        return s;
      }
    }
    if ((uint32_t)(w >> 32) == fp && same(S.occ(w))) return s;
    s = s + 1 == S.cap ? 0 : s + 1;
  }
  return -2;
}

// Find (and with insert, claim) the slot of the mer with rolled hash h.
template <int RAS, class Same>
__device__ int find_slot(const Side<RAS>& S, uint64_t h, uint64_t mine_low, bool insert, bool* won, Same same) {
  uint32_t fp;
  const int s = home_slot(S, h, &fp);
  return probe_from(S, s, S.cap, fp, ((uint64_t)fp << 32) | mine_low, insert, won, same);
}

// A decided info word in 16 bits, for a walk in LDS (tables of at most 8190
// slots): successor slot << 2 | base, where successor 8190 means a base with no
// successor and 8191 no base (low bits 1: a fork, 0: a dead end); bit 15 marks
// a visited slot.
constexpr uint32_t kW16NoSucc = 8190u, kW16End = 8191u, kW16Visited = 0x8000u;
__device__ __forceinline__ uint16_t pack_info16(uint32_t x) {
  const uint32_t ext = x & 7u, succ = (x >> 3) & kNoSlot;
  if (ext >= 4) return (uint16_t)((kW16End << 2) | (ext == (uint32_t)kExtF ? 1u : 0u));
  return (uint16_t)((((succ == kNoSlot ? kW16NoSucc : succ) & 8191u) << 2) | ext);
}

__device__ __forceinline__ uint8_t side_byte(const uint8_t* ctg, int L, bool right, int i) {
  return right ? ctg[i] : complement(ctg[L - 1 - i]);
}

// Every mer length the schedule visits for one contig side, then the side's
// output.  RAS = kLdsAS: the reads are staged into lds_side first.
template <int RAS>
__device__ __forceinline__ void side_walk(const WalkView& v, int slot, int job, int nr, int64_t f0, int64_t base,
                                          int64_t side_bytes) {
  __shared__ int sh_t, sh_res, sh_over, sh_e;
  const int lane = threadIdx.x;
  const int c = job >> 1;
  const bool right = (job & 1) != 0;
  const int L = v.ctg_len[c];
  const uint8_t* ctg = v.ctg_seq + v.ctg_off[c];
  const int W = v.walk_len_limit;

  Side<RAS> S;
  S.cap_alloc = v.cap[slot] & kCapMask;
  S.cap = S.cap_alloc;
  S.wide = (v.cap[slot] & kWideCounts) != 0;
  S.nr = nr;
  {
    uint8_t* sb = v.scratch + v.scratch_off[slot];
    const JobLayout lay = job_layout(S.cap_alloc, S.wide, nr, W, v.max_kmer_len);
    S.key = (uint64_t*)(sb + lay.key);
    S.cnt = (uint32_t*)(sb + lay.cnt);
    S.info_g = (glb_u32*)(sb + lay.info);
    S.roff = (int64_t*)(sb + lay.roff);
    S.rlen = (int32_t*)(sb + lay.rlen);
    S.cum = (uint32_t*)(sb + lay.cum);
    S.walk[0] = sb + lay.walk0;
    S.walk[1] = sb + lay.walk1;
    S.smer = sb + lay.smer;
  }
  // the side's reads are contiguous from base; offsets below are relative to it
  for (int i = lane; i < nr; i += kLanes) {
    S.roff[i] = v.read_off[f0 + i] - base;
    S.rlen[i] = v.read_len[f0 + i];
  }
  S.qual = v.read_qual + base;
  int64_t info_at = 0;  // LDS bytes taken by the reads
  // small side: its bases in LDS (every compare and hash below reads them)
  auto stage_reads = [&]() {
    const uint32_t* src = (const uint32_t*)((uintptr_t)(v.read_seq + base) & ~(uintptr_t)3);
    const int sh = (int)((uintptr_t)(v.read_seq + base) & 3);
    const int nw = (int)((side_bytes + 3) >> 2);
    lds_u32* dst = (lds_u32*)lds_side;
    for (int j = lane; j < nw; j += kLanes) dst[j] = __builtin_amdgcn_alignbyte(src[j + 1], src[j], sh);
    for (int j = nw + lane; j < nw + 16; j += kLanes) dst[j] = 0;
  };
  if constexpr (RAS == kLdsAS) {
    stage_reads();
    S.seq = (cu8<RAS>*)lds_side;
    info_at = (side_bytes + 64 + 15) & ~(int64_t)15;
  } else {
    S.seq = (cu8<RAS>*)(v.read_seq + base);
  }
  S.info_l = (lds_u32*)(lds_side + info_at);
  if (lane == 0) sh_over = 0;
  wave_fence();
  __syncthreads();

  const int depth = v.depth[c];
  const int max_len = min(v.max_kmer_len, L);
  int longest = 0, best = 0, cur = 0, shift = 0;
  int term_x = 0, term_f = 0, term_r = 0;
  unsigned long long mers = 0, steps = 0, walks = 0, resizes = 0;
  unsigned long long cyc[4] = {0, 0, 0, 0};  // clear, count, decide, walk (shader clock, lane 0)
  long long t0 = clock64();

  for (int k = v.kmer_len; k >= v.min_kmer_len && k <= max_len; k += shift) {
    // ---- occurrences per read at this mer length ----
    uint32_t carry = 0;
    if (lane == 0) S.cum[0] = 0;
    for (int i0 = 0; i0 < nr; i0 += kLanes) {
      const int i = i0 + lane;
      uint32_t x = i < nr ? (uint32_t)max(0, S.rlen[i] - k) : 0u;
      for (int d = 1; d < kLanes; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, kLanes);
        if (lane >= d) x += y;
      }
      if (i < nr) S.cum[i + 1] = carry + x;
      carry += __shfl(x, kLanes - 1, kLanes);
    }
    const uint32_t N = carry;
    const uint64_t Bk1 = pow_base(k - 1);
    // Table size for this mer length: kLoadNum/16 slots per occurrence (the
    // distinct mers are fewer: sequencing errors aside, reads overlap).  A
    // table that proves too small (a probe run longer than kMaxProbes, or
    // full) is cleared and the count redone at the allocated size, which
    // holds every occurrence at most 80% full.
    S.cap = (int)min((int64_t)S.cap_alloc, ((int64_t)N * kLoadNum / 16 + 64 + 63) & ~(int64_t)63);
    for (;;) {
      const bool bounded = S.cap < S.cap_alloc;
      const int probes = bounded ? min(kMaxProbes, S.cap - 1) : S.cap - 1;
      // info words in LDS when they fit beside the reads: the walk chases them
      S.info_lds = info_at + 4 * (int64_t)S.cap <= v.lds_bytes;
      for (int s = lane; s < S.cap; s += kLanes) S.key[s] = kEmpty;
      {
        uint4* cp = (uint4*)S.cnt;
        const int nq = S.wide ? 2 * S.cap : S.cap;
        for (int s = lane; s < nq; s += kLanes) cp[s] = make_uint4(0, 0, 0, 0);
      }
      wave_fence();
      __syncthreads();
      long long t1 = clock64();
      cyc[0] += (unsigned long long)(t1 - t0);
      t0 = t1;

      // ---- count: lane takes occurrences [q0, q1) of the side's flattened list ----
      const uint32_t per = (N + kLanes - 1) / kLanes;
      const uint32_t q0 = lane * per, q1 = min(N, q0 + per);
      if (q0 < q1) {
        int lo = 0, hi = nr - 1;
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (S.cum[mid] <= q0) lo = mid; else hi = mid - 1;
        }
        int i = lo, pos = (int)(q0 - S.cum[i]);
        int run = S.rlen[i] - k;            // occurrences in read i
        uint32_t ro = (uint32_t)S.roff[i];  // side-local offset of read i
        uint64_t h = hash_window<RAS>(S.seq + ro + pos, k);
        int won_prev = -1;  // slot this lane claimed for the previous occurrence, if its next is this one
        // kU occurrences at a time: their first probes, following bases and
        // qualities load together, then their empty first probes are claimed
        // together, so a lane waits on one round trip per batch, not per mer
        for (uint32_t q = q0; q < q1; q += kU) {
          // the table is too small at this size: every lane stops, to start again
          if (__hip_atomic_load(&sh_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
          const int nb = (int)min((uint32_t)kU, q1 - q);
          uint32_t off[kU], fp[kU];
          uint64_t w[kU];
          int s[kU];
          bool nh[kU], won[kU];
          uint8_t eb[kU], qb[kU];
#pragma unroll
          for (int u = 0; u < kU; u++) {
            if (u >= nb) break;
            off[u] = ro + (uint32_t)pos;
            // the successor hint is written by this lane, so the flag is set
            // only when this lane also takes the next occurrence of the read
            const bool more = q + u + 1 < q1;
            nh[u] = pos + 1 < run && more;
            s[u] = home_slot(S, h, &fp[u]);
            if (!more) break;
            if (++pos >= run) {
              do { i++; } while (S.rlen[i] - k <= 0);
              run = S.rlen[i] - k;
              ro = (uint32_t)S.roff[i];
              pos = 0;
              h = hash_window<RAS>(S.seq + ro, k);
            } else {
              cu8<RAS>* rd = S.seq + ro + pos;
              h = (h - (uint64_t)rd[-1] * Bk1) * kHashBase + rd[k - 1];
            }
          }
#pragma unroll
          for (int u = 0; u < kU; u++) {
            if (u >= nb) break;
            w[u] = ld64(&S.key[s[u]]);
            eb[u] = S.seq[off[u] + k];
            qb[u] = S.qual[off[u] + k];
          }
#pragma unroll
          for (int u = 0; u < kU; u++) {
            if (u >= nb) break;
            const uint64_t mine = ((uint64_t)fp[u] << 32) | off[u] | (nh[u] ? kHasNext : 0ull);
            won[u] = false;
            if (w[u] == kEmpty) {
              const uint64_t old =
                  atomicCAS((unsigned long long*)&S.key[s[u]], (unsigned long long)kEmpty, (unsigned long long)mine);
              won[u] = old == kEmpty;
              w[u] = won[u] ? mine : old;
            }
          }
#pragma unroll
          for (int u = 0; u < kU; u++) {
            if (u >= nb) break;
            cu8<RAS>* mine_at = S.seq + off[u];
            auto same = [&](cu8<RAS>* o) { return same_window<12, RAS, RAS>(o, mine_at, k); };
            int sl = s[u];
            if (!won[u] && !((uint32_t)(w[u] >> 32) == fp[u] && same(S.occ(w[u])))) {
              const uint64_t mine = ((uint64_t)fp[u] << 32) | off[u] | (nh[u] ? kHasNext : 0ull);
              sl = probe_from(S, s[u] + 1 == S.cap ? 0 : s[u] + 1, probes, fp[u], mine, true, &won[u], same);
            }
            if (sl == -2) {
              if (!bounded) {
                atomicAdd(&v.counters[3], 1ull);
              } else {
                __hip_atomic_store(&sh_over, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                break;
              }
            }
            if (won_prev >= 0 && sl >= 0) S.info_put(won_prev, (uint32_t)sl);
            won_prev = (won[u] && nh[u]) ? sl : -1;
            const int b = vote_base(eb[u]);
            if (sl >= 0 && b >= 0) {
              const int qd = (int)(int8_t)qb[u] - v.qual_offset;
              if (S.wide) {
                if (qd >= v.min_qual) atomicAdd(&S.cnt[8 * (int64_t)sl + b], 1u);
                if (qd >= v.min_hi_qual) atomicAdd(&S.cnt[8 * (int64_t)sl + 4 + b], 1u);
              } else if (qd >= v.min_qual) {
                // one word per base: votes in the low half, high-quality votes in the high half
                atomicAdd(&S.cnt[4 * (int64_t)sl + b], qd >= v.min_hi_qual ? 0x10001u : 1u);
              } else if (qd >= v.min_hi_qual) {
                atomicAdd(&S.cnt[4 * (int64_t)sl + b], 0x10000u);
              }
            }
          }
        }
      }
      wave_fence();
      __syncthreads();
      t1 = clock64();
      cyc[1] += (unsigned long long)(t1 - t0);
      t0 = t1;
      const bool over = sh_over != 0;
      __syncthreads();  // every lane has read the flag before it is cleared
      if (!over) break;
      if (lane == 0) sh_over = 0;
      S.cap = S.cap_alloc;
      resizes++;
    }
    mers += N;

    // ---- decide: extension and successor of every occupied slot ----
    // Keys and votes no longer change: plain (vector) loads.
    // kD slots per lane at a time: keys, votes and hints load together
    for (int s0 = lane; s0 < S.cap; s0 += kLanes * kD) {
      uint64_t w[kD];
      uint4 va[kD], vb[kD];
      uint32_t hint[kD];
      int ext[kD];
      uint8_t nextc[kD];
#pragma unroll
      for (int u = 0; u < kD; u++) {
        const int s = s0 + u * kLanes;
        w[u] = kEmpty;
        if (s < S.cap) {
          w[u] = S.key[s];
          hint[u] = S.info_get(s);
          if (S.wide) {
            va[u] = *(const uint4*)(S.cnt + 8 * (int64_t)s);
            vb[u] = *(const uint4*)(S.cnt + 8 * (int64_t)s + 4);
          } else {
            va[u] = *(const uint4*)(S.cnt + 4 * (int64_t)s);
          }
        }
      }
#pragma unroll
      for (int u = 0; u < kD; u++) {
        ext[u] = kExtX;
        if (w[u] == kEmpty) continue;
        uint32_t low[4], hiq[4];
        if (S.wide) {
          low[0] = va[u].x;
          low[1] = va[u].y;
          low[2] = va[u].z;
          low[3] = va[u].w;
          hiq[0] = vb[u].x;
          hiq[1] = vb[u].y;
          hiq[2] = vb[u].z;
          hiq[3] = vb[u].w;
        } else {
          const uint32_t wv[4] = {va[u].x, va[u].y, va[u].z, va[u].w};
          for (int b = 0; b < 4; b++) {
            low[b] = wv[b] & 0xFFFFu;
            hiq[b] = wv[b] >> 16;
          }
        }
        ext[u] = choose_ext(low, hiq, depth, v.rules);
        // the base after the stored occurrence, where its read continues
        if (ext[u] < 4 && (w[u] & kHasNext)) nextc[u] = S.occ(w[u])[k];
      }
#pragma unroll
      for (int u = 0; u < kD; u++) {
        if (w[u] == kEmpty) continue;
        uint32_t succ = kNoSlot;
        if (ext[u] < 4) {
          cu8<RAS>* rep = S.occ(w[u]);
          const uint8_t ec = (uint8_t)ext_char(ext[u]);
          if ((w[u] & kHasNext) && nextc[u] == ec) {
            // the stored occurrence continues with the chosen base: the next
            // occurrence in its read is the successor mer, and the lane that
            // stored it left that occurrence's slot in the info word
            succ = hint[u];
          } else {
            const uint64_t hs = hash_window<RAS>(rep + 1, k - 1) * kHashBase + ec;
            bool won = false;
            const int t = find_slot(S, hs, 0, false, &won, [&](cu8<RAS>* o) {
              return same_window<12, RAS, RAS>(o, rep + 1, k - 1) && o[k - 1] == ec;
            });
            if (t >= 0) succ = (uint32_t)t;
          }
        }
        S.info_put(s0 + u * kLanes, (uint32_t)ext[u] | (succ << 3));
      }
    }
    wave_fence();
    __syncthreads();
    long long t1 = clock64();
    cyc[2] += (unsigned long long)(t1 - t0);
    t0 = t1;

    // ---- walk from the contig's last mer ----
    for (int j = lane; j < k; j += kLanes) S.smer[j] = side_byte(ctg, L, right, L - k + j);
    wave_fence();
    __syncthreads();
    if (lane == 0) {
      cu8<kGlobalAS>* smer = (cu8<kGlobalAS>*)S.smer;
      const uint64_t h0 = hash_window<kGlobalAS>(smer, k);
      bool won = false;
      sh_e = find_slot(S, h0, 0, false, &won,
                       [&](cu8<RAS>* o) { return same_window<12, RAS, kGlobalAS>(o, smer, k); });
    }
    // A table whose info words are in HBM but fit LDS as 16-bit entries is
    // walked there: the walk needs no read bases once its first slot is found.
    const bool walk16 = !S.info_lds && S.cap <= 8190 && 2 * (int64_t)S.cap <= v.lds_bytes;
    lds_u16* w16 = (lds_u16*)lds_side;
    if (walk16) {
      __syncthreads();  // the first slot is found before the reads are overwritten
      for (int s = lane; s < S.cap; s += kLanes) w16[s] = pack_info16(*(S.info_g + s));
      wave_fence();
      __syncthreads();
    }
    if (lane == 0) {
      int e = sh_e;
      uint8_t* wb = cur ? S.walk[1] : S.walk[0];
      int t = 0, res = 0;  // 0 X, 1 F, 2 R
      while (t < W) {
        if (e < 0) {
          res = 0;
          break;
        }
        // only this lane touches the info words now: read, then mark visited
        if (walk16) {
          const uint32_t x = w16[e];
          w16[e] = (uint16_t)(x | kW16Visited);
          if (x & kW16Visited) {
            res = 2;
            break;
          }
          const uint32_t f = (x >> 2) & 8191u, b = x & 3u;
          if (f == kW16End) {
            res = b == 1u ? 1 : 0;
            break;
          }
          wb[t++] = (uint8_t)ext_char((int)b);
          e = f == kW16NoSucc ? -1 : (int)f;
          continue;
        }
        uint32_t old;
        if (S.info_lds) {
          old = S.info_l[e];
          S.info_l[e] = old | kVisited;
        } else {
          old = __hip_atomic_fetch_or(S.info_g + e, kVisited, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        if (old & kVisited) {
          res = 2;
          break;
        }
        const int ext = (int)(old & 7u);
        if (ext >= 4) {
          res = ext == kExtF ? 1 : 0;
          break;
        }
        wb[t++] = (uint8_t)ext_char(ext);
        const uint32_t nx = (old >> 3) & kNoSlot;
        e = nx == kNoSlot ? -1 : (int)nx;
      }
      sh_t = t;
      sh_res = res;
    }
    wave_fence();
    __syncthreads();
    // the next mer length counts from the reads again
    if constexpr (RAS == kLdsAS) {
      if (walk16) stage_reads();
    }
    const int t = sh_t, res = sh_res;
    t1 = clock64();
    cyc[3] += (unsigned long long)(t1 - t0);
    t0 = t1;
    steps += (unsigned long long)t;
    walks++;
    term_x += res == 0;
    term_f += res == 1;
    term_r += res == 2;
    if (t > longest) {
      longest = t;
      best = cur;
      cur ^= 1;
    }
    __syncthreads();  // sh_t / sh_res are rewritten by the next walk
    if (res == 0) {
      if (shift == v.shift) break;
      shift = -v.shift;
    } else {
      if (shift == -v.shift) break;
      if (k > L) break;
      shift = v.shift;
    }
  }

  // ---- the longest walk, reverse-complemented on the left side ----
  uint8_t* out = v.out_ext + (int64_t)job * W;
  const uint8_t* src = best ? S.walk[1] : S.walk[0];
  for (int j = lane; j < longest; j += kLanes) out[j] = right ? src[j] : complement(src[longest - 1 - j]);
  if (lane == 0) {
    v.out_len[job] = longest;
    v.out_terms[3 * job] = term_x;
    v.out_terms[3 * job + 1] = term_f;
    v.out_terms[3 * job + 2] = term_r;
    atomicAdd(&v.counters[0], mers);
    atomicAdd(&v.counters[1], steps);
    atomicAdd(&v.counters[2], walks);
    atomicAdd(&v.counters[4], cyc[0]);
    atomicAdd(&v.counters[5], cyc[1]);
    atomicAdd(&v.counters[6], cyc[2]);
    atomicAdd(&v.counters[7], cyc[3]);
    atomicAdd(&v.counters[8], resizes);
  }
}

__global__ __launch_bounds__(kLanes) void walk_kernel(WalkView v) {
  const int slot = blockIdx.x;
  const int job = v.jobs[slot];
  const int64_t f0 = v.rfirst[job];
  const int nr = (int)min(v.rfirst[job + 1] - f0, (int64_t)v.max_reads);
  const int64_t base = nr > 0 ? v.read_off[f0] : 0;
  const int64_t side_bytes = nr > 0 ? v.read_off[f0 + nr - 1] + v.read_len[f0 + nr - 1] - base : 0;
  if (side_bytes + 64 <= v.lds_bytes) side_walk<kLdsAS>(v, slot, job, nr, f0, base, side_bytes);
  else side_walk<kGlobalAS>(v, slot, job, nr, f0, base, side_bytes);
}

}  // namespace

void launch_walks(const WalkView& v, int64_t njobs, hipStream_t st) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(walk_kernel, dim3((unsigned)njobs), dim3(kLanes), (size_t)v.lds_bytes, st, v);
}

}  // namespace mhmla
