// span_check.cpp -- test-only: the per-window lane spans (two divisions per window,
// es::WindowSpan) against the per-lane formula (es::lane_span) on 3 million random
// windows and lanes inside the scanner's domain (hi <= the stage-E ceiling).
//
// usage: span_check     prints "span ok" or the first mismatch.
#include <cstdio>
#include <random>

#include "es_plan.h"

using namespace es;

int main() {
  std::mt19937_64 g(3);
  const u64 mps[] = {840, 9240, 120120, 185640, 2042040};
  for (int t = 0; t < 3000000; ++t) {
    const u64 Mp = mps[g() % 5];
    u64 lo = g() >> (g() % 64), hi = lo + (g() >> (g() % 64));
    if (hi < lo || hi > kDirectPMax) hi = kDirectPMax;
    if (lo > hi) lo = hi;
    if (t % 7 == 0) {               // windows near the origin and shorter than a lane period
      lo = g() % (3 * Mp);
      hi = lo + g() % (3 * Mp);
    }
    const u64 c = g() % Mp;
    u64 a, b, x, y;
    lane_span(lo, hi, c, Mp, a, b);
    WindowSpan(lo, hi, Mp).lane(c, x, y);
    if (b != y || (b && a != x)) {
      printf("mismatch lo=%llu hi=%llu c=%llu Mp=%llu\n", (unsigned long long)lo, (unsigned long long)hi,
             (unsigned long long)c, (unsigned long long)Mp);
      return 1;
    }
  }
  printf("span ok\n");
  return 0;
}
