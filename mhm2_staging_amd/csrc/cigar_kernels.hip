// HIP kernels of the CIGAR pass: SSW's CIGAR and mismatch count for every
// pair of a finished job, as klign gets them when it reports CIGARs
// (src/ssw/ssw_core.cpp:1180-1192 banded_sw, src/ssw/ssw.cpp:57-93 and
// :129-254; klign.cpp:589-599 keeps such batches off ADEPT's GPU path).
// The per-pair arithmetic lives in cigar_core.h.
//
//  * cigar_classify_kernel: the rectangle of each pair from its result,
//    trivial pairs (empty query, distance filter, score 0) written at once,
//    the rest appended to the list of their band class (band half-width
//    |columns - rows| + 1: 1..7 exact, 8..15 masked in a 31-wide band,
//    wider or taller than kCigMaxRows generic), wavefront-aggregated.
//  * cigar_band_kernel<BW, MASKED>: persistent, one lane per pair, the band
//    row (2*BW+1 diagonals of H and E) in registers, 4 direction bits per
//    cell stored lane-interleaved in the wavefront's scratch slot (coalesced
//    row stores).  A pair whose band misses the score doubles its half-width
//    and moves to that class's list (launched later).  Then two walks over
//    the directions count the ops (a third when the traceback stops off
//    column 0), the wavefront reserves its ops with one atomic, and a last
//    walk writes them.
//  * cigar_generic_kernel: the same for any half-width, rows of H/E and the
//    directions in a bump-allocated pool; pairs that do not fit are counted
//    and the host reruns the pass with a larger pool.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cigar_core.h"
#include "sw_launch.h"

namespace mhmsw {

namespace {

__device__ __forceinline__ int32_t lane0(int32_t v) { return __builtin_amdgcn_readlane(v, 0); }

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
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)base, 0);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int32_t)(uint32_t)(base >> 32), 0);
  return (int64_t)(((uint64_t)hi << 32) | lo) + (x - n);
}

// Appends p to the list of class cls (rare path: per-lane atomic).
__device__ __forceinline__ void enqueue(const CigarView& v, int cls, int64_t p, int32_t bw) {
  v.band[p] = bw;
  const int32_t at = atomicAdd(&v.list_count[cls], 1);
  v.lists[(int64_t)cls * v.npairs + at] = (int32_t)p;
}

// Work index of the wavefront: one atomic from lane 0, broadcast.
__device__ __forceinline__ int32_t next_work(int32_t* ctr) {
  int32_t w = 0;
  if ((threadIdx.x & 63) == 0) w = atomicAdd(ctr, 1);
  return lane0(w);
}

// Whether SSW-query[qb] equals SSW-ref[rb] (the traceback's last M).
__device__ __forceinline__ bool first_eq(const CigarView& v, const CigarGeo& g) {
  return nib_at(v.codes, g.qbase + g.qb) == nib_at(v.codes, g.rbase + g.rb);
}

// Counting walks over a register-class band: rows fetched in groups and
// '=' / 'X' from the equality bits; when the traceback stops off column 0
// the count is redone comparing shifted codes.  Returns the shift of the
// writing walk (-1: the traceback left the band).
template <int DW>
__device__ __forceinline__ int32_t count_band(const CigarView& v, const CigarGeo& g, int32_t bw, const uint32_t* dirs,
                                              bool eq00, CigarEmit& em) {
  em = CigarEmit();
  em.out = nullptr;
  em.nops = 0;
  int32_t jf = cigar_walk_rows<DW>(g, bw, dirs, eq00, em);
  if (jf > 0) {
    em = CigarEmit();
    em.out = nullptr;
    em.nops = 0;
    jf = cigar_walk(v.codes, g, bw, dirs, DW, 64, jf, eq00, em);
  }
  return jf;
}

// Reserves the ops of every lane's finished pair and writes them (all lanes
// of the wavefront call it).  BAND: dirs is a register-class slot.
template <int DW, bool BAND>
__device__ __forceinline__ void finish_pairs(const CigarView& v, int64_t p, bool done, int32_t jf, const CigarGeo& g,
                                             int32_t layout_bw, const uint32_t* dirs, int dwn, int64_t stride,
                                             bool eq00, const CigarEmit& em) {
  const bool ok = done && jf >= 0;
  const int64_t base = wave_reserve(&v.used[0], ok ? em.count : 0);
  if (!done) return;
  if (!ok) {
    atomicAdd(&v.used[3], 1ull);  // traceback left the band: reported as an internal error
    v.n_ops[p] = -3;
    v.mism[p] = 0;
    v.off[p] = 0;
    return;
  }
  const bool fits = base + em.count <= v.ops_cap;
  if (fits && em.count > 0) {
    CigarEmit w;
    w.out = v.ops + base;
    w.nops = em.count;
    if (BAND && jf == 0) {
      cigar_walk_rows<DW>(g, layout_bw, dirs, eq00, w);
    } else {
      cigar_walk(v.codes, g, layout_bw, dirs, dwn, stride, jf, eq00, w);
    }
  }
  v.n_ops[p] = fits ? em.count : -1;
  v.mism[p] = em.mism;
  v.off[p] = base;
}

}  // namespace

__global__ void cigar_clear_kernel(CigarView v) {
  const int t = threadIdx.x;
  if (t < 2 * kCigClasses) v.list_count[t] = 0;
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
          if (cls[k] != kCigClassGeneric) atomicMax(&rows_max, g.ql());
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
        v.lists[(int64_t)cls[k] * v.npairs + base[cls[k]] + rank[k]] = (int32_t)p;
      }
    }
    __syncthreads();
  }
}

template <int BW, bool MASKED>
__global__ __launch_bounds__(256) void cigar_band_kernel(CigarView v) {
  constexpr int W = 2 * BW + 1;
  constexpr int DW = (4 * W + 31) / 32;
  constexpr int cls = MASKED ? kCigClassMasked : BW - 1;
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  // one scratch slot per wavefront: rows of the tallest rectangle x DW dwords x 64 lanes
  const int64_t slot_dwords = (int64_t)lane0(v.max_rows[0]) * DW * 64;
  if (slot_dwords <= 0 || (wave + 1) * slot_dwords > v.dirs_dwords) return;
  uint32_t* dirs = v.dirs + wave * slot_dwords + lane;
  const int32_t count = lane0(v.list_count[cls]);
  int32_t* next = &v.list_count[kCigClasses + cls];
  for (;;) {
    const int32_t w = next_work(next);
    if ((int64_t)w * 64 >= count) break;
    const int32_t idx = w * 64 + lane;
    const int64_t p = idx < count ? v.lists[(int64_t)cls * v.npairs + idx] : -1;
    bool done = false, eq00 = false;
    int32_t jf = -1;
    CigarGeo g;
    CigarEmit em;
    if (p >= 0) {
      g = pair_geo(v, p);
      const int32_t bw = v.band[p];
      const int32_t corner = cigar_band_dp<BW, MASKED>(v.codes, g, bw, v.sc.match, v.sc.mismatch, v.sc.ambiguity,
                                                       v.sc.gap_open, v.sc.gap_ext, dirs, 64);
      if (corner == g.score) {
        done = true;
        eq00 = first_eq(v, g);
        jf = count_band<DW>(v, g, BW, dirs, eq00, em);
      } else {
        enqueue(v, cigar_class(2 * bw, g.ql()), p, 2 * bw);
      }
    }
    finish_pairs<DW, true>(v, p, done, jf, g, BW, dirs, DW, 64, eq00, em);
  }
}

__global__ __launch_bounds__(64) void cigar_generic_kernel(CigarView v) {
  const int lane = threadIdx.x & 63;
  const int32_t count = lane0(v.list_count[kCigClassGeneric]);
  int32_t* next = &v.list_count[kCigClasses + kCigClassGeneric];
  for (;;) {
    const int32_t w = next_work(next);
    if ((int64_t)w * 64 >= count) break;
    const int32_t idx = w * 64 + lane;
    const int64_t p = idx < count ? v.lists[(int64_t)kCigClassGeneric * v.npairs + idx] : -1;
    bool done = false, eq00 = false;
    int32_t jf = -1, bw = 0, DW = 0;
    const uint32_t* dirs = nullptr;
    CigarGeo g;
    CigarEmit em;
    if (p >= 0) {
      g = pair_geo(v, p);
      bw = v.band[p];
      for (;;) {
        const int32_t W = 2 * bw + 1;
        DW = (4 * W + 31) / 32;
        const int64_t need = 2 * (int64_t)W + (int64_t)g.ql() * DW;
        const int64_t at = (int64_t)atomicAdd(&v.used[1], (unsigned long long)need);
        if (at + need > v.gen_dwords) {
          atomicAdd(&v.used[2], 1ull);  // deferred: the host grows the pool and reruns
          v.n_ops[p] = -2;
          break;
        }
        uint32_t* mem = v.gen + at;
        const int32_t corner = cigar_band_dp_any(v.codes, g, bw, v.sc.match, v.sc.mismatch, v.sc.ambiguity,
                                                 v.sc.gap_open, v.sc.gap_ext, (int32_t*)mem, (int32_t*)mem + W,
                                                 mem + 2 * W);
        if (corner == g.score) {
          dirs = mem + 2 * W;
          done = true;
          eq00 = first_eq(v, g);
          em = CigarEmit();
          em.out = nullptr;
          em.nops = 0;
          jf = cigar_walk(v.codes, g, bw, dirs, DW, 1, 0, eq00, em);
          if (jf > 0) {
            em = CigarEmit();
            em.out = nullptr;
            em.nops = 0;
            jf = cigar_walk(v.codes, g, bw, dirs, DW, 1, jf, eq00, em);
          }
          break;
        }
        if (bw > g.rl() + g.ql() + 2) {  // a band this wide holds the whole rectangle
          done = true;
          jf = -1;
          break;
        }
        bw *= 2;
      }
    }
    finish_pairs<1, false>(v, p, done, jf, g, bw, dirs, DW, 1, eq00, em);
  }
}

static int cig_grid(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 65535) g = 65535;
  return (int)g;
}

hipError_t launch_cigar(const CigarView& v, int band_waves, int generic_waves, hipStream_t st) {
  hipLaunchKernelGGL(cigar_clear_kernel, dim3(1), dim3(64), 0, st, v);
  if (v.npairs <= 0) return hipGetLastError();
  hipLaunchKernelGGL(cigar_classify_kernel, dim3(cig_grid(v.npairs, kCigChunk)), dim3(256), 0, st, v);
  const int blocks = std::max(1, band_waves / 4);
  hipLaunchKernelGGL((cigar_band_kernel<1, false>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL((cigar_band_kernel<2, false>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL((cigar_band_kernel<3, false>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL((cigar_band_kernel<4, false>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL((cigar_band_kernel<5, false>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL((cigar_band_kernel<6, false>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL((cigar_band_kernel<7, false>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL((cigar_band_kernel<kCigMaskedBw, true>), dim3(blocks), dim3(256), 0, st, v);
  hipLaunchKernelGGL(cigar_generic_kernel, dim3(std::max(1, generic_waves)), dim3(64), 0, st, v);
  return hipGetLastError();
}

}  // namespace mhmsw
