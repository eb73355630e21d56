// Host-side 2-bit packing of ASCII sequences (no HIP dependency).
//
// Codes follow SSW's translation table (src/ssw/ssw.cpp:13-26): A/a and U/u
// -> 0, C/c -> 1, G/g -> 2, T/t -> 3; every other byte is N, stored as code 0
// with its bit set in a separate N bitmap.  A sequence is packed into whole
// dwords from a base that is a multiple of 32: 16 bases per 2-bit dword,
// 32 per bitmap dword, base k at bits 2k / k.
#pragma once
#include <immintrin.h>
#include <stdint.h>

namespace mhmsw {
namespace pack2 {

inline uint32_t base_code(unsigned char c) {
  switch (c) {
    case 'A': case 'a': case 'U': case 'u': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
  }
}

struct CodeTables {
  uint8_t two[256];
  uint8_t isn[256];
  CodeTables() {
    for (int c = 0; c < 256; c++) {
      const uint32_t k = base_code((unsigned char)c);
      two[c] = (uint8_t)(k == 4 ? 0 : k);
      isn[c] = (uint8_t)(k == 4 ? 1 : 0);
    }
  }
};
inline const CodeTables& code_tables() {
  static const CodeTables t;
  return t;
}

// Scalar form: bases [x0, x0 + m) of a sequence, m <= 32, x0 a multiple of 32.
// Writes the block's two code dwords and its flag dword whole (zero past m).
inline void pack_scalar(const unsigned char* u, int32_t x0, int32_t m, uint32_t* two, uint32_t* nb) {
  const CodeTables& t = code_tables();
  uint32_t w0 = 0, w1 = 0, n = 0;
  for (int32_t k = 0; k < m; k++) {
    const unsigned char ch = u[x0 + k];
    const uint32_t code = t.two[ch];
    if (k < 16) {
      w0 |= code << (2 * k);
    } else {
      w1 |= code << (2 * (k - 16));
    }
    n |= (uint32_t)t.isn[ch] << k;
  }
  two[x0 >> 4] = w0;
  two[(x0 >> 4) + 1] = w1;
  nb[x0 >> 5] = n;
}

// 32 bases -> two dwords of 2-bit codes and one dword of N flags, 16 bases per
// SSSE3 step.  The low nibble of A/a, C/c, G/g, T/t, U/u is 1, 3, 7, 4, 5
// (all distinct), so one byte shuffle maps it to the 2-bit code and another to
// the upper-case letter it must be; a byte is N unless (byte & 0xDF) equals
// that letter (0xFF for the other nibbles can never match), and N bytes get code 0.
__attribute__((target("ssse3"))) inline void pack32_simd(const unsigned char* u, uint32_t* two, uint32_t* nb) {
  const __m128i code_lut = _mm_setr_epi8(0, 0, 0, 1, 3, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0);
  const __m128i want_lut = _mm_setr_epi8((char)0xFF, 'A', (char)0xFF, 'C', 'T', 'U', (char)0xFF, 'G', (char)0xFF,
                                         (char)0xFF, (char)0xFF, (char)0xFF, (char)0xFF, (char)0xFF, (char)0xFF,
                                         (char)0xFF);
  const __m128i lo4 = _mm_set1_epi8(0x0F), up = _mm_set1_epi8((char)0xDF);
  const __m128i w14 = _mm_set1_epi16(0x0401);       // bytes (1, 4): c0 + 4 c1
  const __m128i w116 = _mm_set1_epi32(0x00100001);  // words (1, 16)
  uint32_t nmask = 0;
  for (int h = 0; h < 2; h++) {
    const __m128i x = _mm_loadu_si128((const __m128i*)(u + 16 * h));
    const __m128i nib = _mm_and_si128(x, lo4);
    const __m128i ok = _mm_cmpeq_epi8(_mm_and_si128(x, up), _mm_shuffle_epi8(want_lut, nib));
    const __m128i code = _mm_and_si128(_mm_shuffle_epi8(code_lut, nib), ok);  // N -> code 0
    nmask |= (uint32_t)(~_mm_movemask_epi8(ok) & 0xFFFF) << (16 * h);
    const __m128i c2 = _mm_madd_epi16(_mm_maddubs_epi16(code, w14), w116);  // one byte per 4 bases
    const __m128i b = _mm_packus_epi16(_mm_packs_epi32(c2, c2), c2);
    two[h] = (uint32_t)_mm_cvtsi128_si32(b);
  }
  *nb = nmask;
}

// Up to 64 bases (m of them, 1..64) -> the 2-bit and N-flag dwords of their
// 32-base blocks in one AVX-512 step: the same two nibble shuffles as
// pack32_simd on all four 16-byte lanes, the comparison giving the N flags as
// a 64-bit mask, and one byte per four bases narrowed out of the dwords with
// vpmovdb.  Bytes past m are neither read (a masked load) nor flagged; the
// dwords of the ceil(m / 32) blocks are written whole, as pack_scalar does.
__attribute__((target("avx512f,avx512bw,avx512vl"))) inline void pack64_avx512(const unsigned char* u, int32_t m,
                                                                              uint32_t* two, uint32_t* nb) {
  const __m512i code_lut = _mm512_broadcast_i32x4(_mm_setr_epi8(0, 0, 0, 1, 3, 0, 0, 2, 0, 0, 0, 0, 0, 0, 0, 0));
  const __m512i want_lut = _mm512_broadcast_i32x4(
      _mm_setr_epi8((char)0xFF, 'A', (char)0xFF, 'C', 'T', 'U', (char)0xFF, 'G', (char)0xFF, (char)0xFF, (char)0xFF,
                    (char)0xFF, (char)0xFF, (char)0xFF, (char)0xFF, (char)0xFF));
  const __mmask64 in = m >= 64 ? ~(__mmask64)0 : (((__mmask64)1 << m) - 1);
  const __m512i x = _mm512_maskz_loadu_epi8(in, (const void*)u);
  const __m512i nib = _mm512_and_si512(x, _mm512_set1_epi8(0x0F));
  const __mmask64 ok = _mm512_mask_cmpeq_epi8_mask(in, _mm512_and_si512(x, _mm512_set1_epi8((char)0xDF)),
                                                   _mm512_shuffle_epi8(want_lut, nib));
  const __m512i code = _mm512_maskz_shuffle_epi8(ok, code_lut, nib);  // N and bytes past m -> code 0
  const __m512i c2 =
      _mm512_madd_epi16(_mm512_maddubs_epi16(code, _mm512_set1_epi16(0x0401)), _mm512_set1_epi32(0x00100001));
  const uint64_t n = (uint64_t)(in & ~ok);
  if (m > 32) {
    _mm_storeu_si128((__m128i*)two, _mm512_cvtepi32_epi8(c2));
    nb[0] = (uint32_t)n;
    nb[1] = (uint32_t)(n >> 32);
  } else {
    _mm_storel_epi64((__m128i*)two, _mm512_cvtepi32_epi8(c2));
    nb[0] = (uint32_t)n;
  }
}

__attribute__((target("avx512f,avx512bw,avx512vl"))) inline void pack_seq_avx512(const unsigned char* u, int32_t len,
                                                                                uint32_t* two, uint32_t* nb) {
  for (int32_t x0 = 0; x0 < len; x0 += 64) pack64_avx512(u + x0, len - x0, two + (x0 >> 4), nb + (x0 >> 5));
}

inline void pack_seq_ssse3(const unsigned char* u, int32_t len, uint32_t* two, uint32_t* nb) {
  const int32_t full = len & ~31;
  for (int32_t x0 = 0; x0 < full; x0 += 32) pack32_simd(u + x0, two + (x0 >> 4), nb + (x0 >> 5));
  if (full < len) pack_scalar(u, full, len - full, two, nb);
}

// One sequence of len bases into whole dwords: every dword of its 32-base
// blocks is written (two: 2 per block, nb: 1 per block; codes and flags past
// len are 0), with AVX-512 where the CPU has it.
inline void pack_seq(const char* src, int32_t len, uint32_t* two, uint32_t* nb) {
  static const bool avx512 = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
                             __builtin_cpu_supports("avx512vl");
  const unsigned char* u = (const unsigned char*)src;
  if (avx512) {
    pack_seq_avx512(u, len, two, nb);
  } else {
    pack_seq_ssse3(u, len, two, nb);
  }
}

}  // namespace pack2
}  // namespace mhmsw
