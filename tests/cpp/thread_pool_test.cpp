// CPU check of the host worker pool (mhm2_staging_amd/csrc/thread_pool.h):
// every index of every region runs exactly once, for region widths 1..16,
// with regions issued from several threads at once (a busy pool makes the
// caller run its region on threads of its own); in claimed regions every
// chunk of work taken from the region's counter is done exactly once, and
// each index runs at most once.
#include <atomic>
#include <cstdio>
#include <thread>
#include <vector>

#include "../../mhm2_staging_amd/csrc/thread_pool.h"

int main() {
  std::atomic<long> bad(0), regions(0);
  auto caller = [&](int seed) {
    for (int r = 0; r < 400; r++) {
      const int nth = 1 + (seed * 7 + r) % 16;
      std::vector<std::atomic<int>> hits(nth);
      for (auto& h : hits) h = 0;
      hostpool::pool().run(nth, [&](int t) {
        if (t < 0 || t >= nth) {
          bad++;
          return;
        }
        hits[t]++;
      });
      for (int t = 0; t < nth; t++)
        if (hits[t] != 1) bad++;
      // a claimed region: 64 chunks shared out through a counter
      std::vector<std::atomic<int>> done(64), ran(nth);
      for (auto& d : done) d = 0;
      for (auto& h : ran) h = 0;
      std::atomic<int> next(0);
      hostpool::pool().run_claimed(nth, [&](int t) {
        if (t < 0 || t >= nth) {
          bad++;
          return;
        }
        ran[t]++;
        for (int k; (k = next.fetch_add(1)) < 64;) done[k]++;
      });
      for (int k = 0; k < 64; k++)
        if (done[k] != 1) bad++;
      for (int t = 0; t < nth; t++)
        if (ran[t] > 1 || (t == 0 && ran[t] != 1)) bad++;
      regions++;
    }
  };
  std::vector<std::thread> callers;
  for (int k = 0; k < 4; k++) callers.emplace_back(caller, k);
  caller(4);
  for (auto& c : callers) c.join();
  std::printf("regions %ld bad %ld\n", regions.load(), bad.load());
  return bad == 0 && regions == 2000 ? 0 : 1;
}
