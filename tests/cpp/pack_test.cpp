// The SSSE3 and AVX-512 2-bit packers against the scalar per-base packer (SSW's
// translation table, src/ssw/ssw.cpp:13-26): random bytes of every value,
// ACGTU/acgtu text with N, every length 0..300, and the same dwords written
// (the buffers start as a fill pattern).  The AVX-512 form is checked where the
// CPU has it.  Prints "ok <cases> <bases/s pack_seq> <bases/s scalar> <avx512 0|1>".
#include <chrono>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "../../mhm2_staging_amd/csrc/pack2bit.h"

using namespace mhmsw::pack2;

static void pack_ref(const std::string& s, std::vector<uint32_t>& two, std::vector<uint32_t>& nb) {
  const int32_t len = (int32_t)s.size();
  two.assign((size_t)(len + 31) / 32 * 2 + 1, 0xA5A5A5A5u);
  nb.assign((size_t)(len + 31) / 32 + 1, 0xA5A5A5A5u);
  for (int32_t x0 = 0; x0 < len; x0 += 32)
    pack_scalar((const unsigned char*)s.data(), x0, std::min<int32_t>(32, len - x0), two.data(), nb.data());
}

int main() {
  const bool avx512 = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
                      __builtin_cpu_supports("avx512vl");
  std::mt19937_64 rng(7);
  const char alpha[] = "ACGTUacgtuNnRY-*.\\0";
  long cases = 0;
  for (int kind = 0; kind < 3; kind++) {
    for (int len = 0; len <= 300; len++) {
      for (int rep = 0; rep < 20; rep++) {
        std::string s((size_t)len, 'A');
        for (auto& c : s) {
          if (kind == 0) c = (char)(rng() & 0xFF);
          else if (kind == 1) c = alpha[rng() % (sizeof(alpha) - 1)];
          else c = "ACGT"[rng() & 3];
        }
        std::vector<uint32_t> t1, n1, t2, n2;
        pack_ref(s, t1, n1);
        for (int form = 0; form < (avx512 ? 3 : 2); form++) {
          t2.assign(t1.size(), 0xA5A5A5A5u);
          n2.assign(n1.size(), 0xA5A5A5A5u);
          const unsigned char* u = (const unsigned char*)s.data();
          if (form == 0) pack_seq(s.data(), len, t2.data(), n2.data());
          if (form == 1) pack_seq_ssse3(u, len, t2.data(), n2.data());
          if (form == 2) pack_seq_avx512(u, len, t2.data(), n2.data());
          if (t1 != t2 || n1 != n2) {
            std::printf("MISMATCH form %d kind %d len %d rep %d\n", form, kind, len, rep);
            return 1;
          }
        }
        cases++;
      }
    }
  }
  // throughput on 64 MB of ACGT text
  std::string big((size_t)1 << 26, 'A');
  for (auto& c : big) c = "ACGT"[rng() & 3];
  std::vector<uint32_t> two(big.size() / 16 + 2), nb(big.size() / 32 + 2);
  auto t0 = std::chrono::steady_clock::now();
  pack_seq(big.data(), (int32_t)big.size(), two.data(), nb.data());
  double ts = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  t0 = std::chrono::steady_clock::now();
  for (int32_t x0 = 0; x0 < (int32_t)big.size(); x0 += 32)
    pack_scalar((const unsigned char*)big.data(), x0, 32, two.data(), nb.data());
  double tr = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::printf("ok %ld %.3g %.3g %d\n", cases, big.size() / ts, big.size() / tr, avx512 ? 1 : 0);
  return 0;
}
