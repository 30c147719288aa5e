// format_check.cpp -- test-only: es::format_rung (table-driven decimal digits, 128-bit
// witnesses split at 10^19) against printf-style formatting with digit-by-digit 128-bit
// division, on edge values and 2 million random lines.
//
// usage: format_check     prints "format ok" or the first mismatch.
#include <cstdint>
#include <cstdio>
#include <random>
#include <string>

#include "es_tail.h"

typedef unsigned __int128 u128;

static std::string slow(uint64_t p, uint64_t A, unsigned r, u128 u) {
  char h[96];
  snprintf(h, sizeof h, "RUNG %llu %llu %u ", (unsigned long long)p, (unsigned long long)A, r);
  std::string s = h;
  char b[48];
  int n = 0;
  do { b[n++] = (char)('0' + (int)(u % 10)); u /= 10; } while (u);
  while (n--) s.push_back(b[n]);
  s.push_back('\n');
  return s;
}

int main() {
  std::mt19937_64 g(1);
  std::string got;
  const u128 ten19 = (u128)10000000000000000000ull;
  const u128 edge[] = {0, 1, 9, 10, 99, 100, 101, (u128)~0ull, (u128)~0ull + 1, ten19 - 1, ten19, ten19 * 10 - 1,
                       ten19 * ten19 - 1, ten19 * ten19, ~(u128)0, ((u128)1 << 100) + 12345};
  for (u128 u : edge)
    for (uint64_t p : {0ull, 9ull, 10ull, 18446744073709551615ull}) {
      es::format_rung(p, p / 4, 255, u, got);
      if (got != slow(p, p / 4, 255, u)) { printf("edge mismatch: %s", got.c_str()); return 1; }
    }
  for (int i = 0; i < 2000000; ++i) {
    const uint64_t p = g() >> (g() % 64), A = g() >> (g() % 64);
    const unsigned r = (unsigned)(g() % 256);
    const u128 u = (((u128)g() << 64) | g()) >> (g() % 128);
    es::format_rung(p, A, r, u, got);
    if (got != slow(p, A, r, u)) { printf("mismatch: %s vs %s", got.c_str(), slow(p, A, r, u).c_str()); return 1; }
  }
  printf("format ok\n");
  return 0;
}
