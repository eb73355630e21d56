// CPU check of the extension rule the local assembly kernel runs
// (mhm2_staging_amd/csrc/locassm_core.h, mhmla::choose_ext) against the
// oracle's restatement of the reference's set_ext with its exchange sort
// (oracle/locassm_oracle.c, oracle_la_ext): every vote combination of a small
// range at several depths, then random larger counts.  Prints "ok <cases>"
// or the first difference (exit status 1).
#include <cstdio>
#include <random>

#include "../../mhm2_staging_amd/csrc/locassm_core.h"
#include "../../oracle/locassm_oracle.h"

static int check(const oracle_la_params& p, const uint32_t low[4], const uint32_t hi[4], int depth) {
  const mhmla::Rules r{p.min_viable_depth, p.min_expected_depth, p.rating_thres};
  const int got = mhmla::ext_char(mhmla::choose_ext(low, hi, depth, r));
  const int want = oracle_la_ext(&p, low, hi, depth, nullptr);
  if (got != want) {
    std::printf("depth %d low %u %u %u %u hi %u %u %u %u: got %c want %c\n", depth, low[0], low[1], low[2], low[3],
                hi[0], hi[1], hi[2], hi[3], got, want);
    return 1;
  }
  return 0;
}

int main() {
  oracle_la_params p;
  oracle_la_default_params(&p);
  long cases = 0;
  const int depths[] = {0, 1, 5, 10, 12, 13, 19, 20, 21, 33, 47, 100};
  // (low, hi) pairs with hi <= low <= 4 for each of the four bases
  int pl[16], ph[16], np = 0;
  for (int l = 0; l <= 4; l++)
    for (int h = 0; h <= l; h++) {
      pl[np] = l;
      ph[np] = h;
      np++;
    }
  for (int d : depths)
    for (int a = 0; a < np; a++)
      for (int b = 0; b < np; b++)
        for (int c = 0; c < np; c++)
          for (int e = 0; e < np; e++) {
            const uint32_t low[4] = {(uint32_t)pl[a], (uint32_t)pl[b], (uint32_t)pl[c], (uint32_t)pl[e]};
            const uint32_t hi[4] = {(uint32_t)ph[a], (uint32_t)ph[b], (uint32_t)ph[c], (uint32_t)ph[e]};
            if (check(p, low, hi, d)) return 1;
            cases++;
          }
  std::mt19937_64 rng(7);
  for (int t = 0; t < 2000000; t++) {
    const int d = (int)(rng() % 120);
    const uint32_t span = (t & 1) ? 40 : 3000;
    uint32_t low[4], hi[4];
    for (int i = 0; i < 4; i++) {
      low[i] = (uint32_t)(rng() % span);
      hi[i] = (t % 7 == 0) ? (uint32_t)(rng() % span) : (low[i] ? (uint32_t)(rng() % (low[i] + 1)) : 0);
      if (rng() % 4 == 0) low[i] = low[(i + 1) & 3];  // ties
      if (rng() % 4 == 0) hi[i] = hi[(i + 3) & 3];
    }
    if (check(p, low, hi, d)) return 1;
    cases++;
  }
  // other thresholds
  p.min_viable_depth = 0.15;
  p.min_expected_depth = 0.45;
  p.rating_thres = 2;
  for (int t = 0; t < 500000; t++) {
    const int d = (int)(rng() % 80);
    uint32_t low[4], hi[4];
    for (int i = 0; i < 4; i++) {
      low[i] = (uint32_t)(rng() % 30);
      hi[i] = low[i] ? (uint32_t)(rng() % (low[i] + 1)) : 0;
    }
    if (check(p, low, hi, d)) return 1;
    cases++;
  }
  std::printf("ok %ld\n", cases);
  return 0;
}
