// kernel_replay.cpp -- test-only CPU replay of k_sieve's per-tile arithmetic over the
// product's own layout tables (es_plan), so layout and tiling mistakes show up on the
// CPU. It restates the kernel step by step: 32-bit magic-reciprocal remainders, the
// 64-bit remainder split, the periodic wheel-mask window (alignbit), the byte map marked
// one filter per thread, the three popcounts. Output: "S p" per survivor in ascending p,
// then "C positions after_covers survivors".
//
// usage: kernel_replay LO HI WHEEL MMAX SIEVE MP MAXBITS [FILTERFILE]
#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "es_plan.h"

using namespace es;

static u32 mod32(u32 x, u32 s, u32 mg) {
  const u32 q = (u32)(((u64)x * mg) >> 32);
  const u32 r = x - q * s;
  return r >= s ? r - s : r;
}
static u32 mod64(u64 x, u32 s, u32 mg, u32 c32) {
  const u32 h = mod32((u32)(x >> 32), s, mg), l = mod32((u32)x, s, mg);
  const u32 t = mod32(h * c32, s, mg) + l;
  return t >= s ? t - s : t;
}
static u32 alignbit(u32 hi, u32 lo, u32 sh) { return (u32)((((u64)hi << 32) | lo) >> (sh & 31)); }

int main(int argc, char **argv) {
  if (argc < 8) { fprintf(stderr, "usage\n"); return 1; }
  const u64 lo = strtoull(argv[1], 0, 10), hi = strtoull(argv[2], 0, 10), M = strtoull(argv[3], 0, 10),
            mmax = strtoull(argv[4], 0, 10), sieve = strtoull(argv[5], 0, 10), Mp = strtoull(argv[6], 0, 10);
  const u32 maxb = (u32)strtoul(argv[7], 0, 10);
  const char *ff = argc > 8 ? argv[8] : nullptr;
  Plan plan;
  std::string msg;
  if (build_plan(M, mmax, plan, msg) || plan_filters(plan, ff, sieve, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 1; }
  Layout L;
  if (!build_layout(M, Mp, plan.wheel.classes, plan.filters, L, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 1; }
  const bool has_sieve = plan.filters.mod.size() != plan.filters.ncov;
  const u32 P = (u32)L.P, Pmg = P == 1 ? 0xFFFFFFFFu : (u32)((1ull << 32) / P), Pc32 = (u32)((1ull << 32) % P);
  std::vector<u64> surv;
  u64 pos = 0, aft = 0, sur = 0;
  std::vector<u32> bits;
  for (u32 lane = 0; lane < L.lanes.size(); ++lane) {
    u64 k0l, nbl;
    lane_span(lo, hi, L.lanes[lane], Mp, k0l, nbl);
    for (u64 b = 0; b < nbl; b += maxb) {
      const u64 k0 = k0l + b;
      const u32 nb = (u32)std::min<u64>(maxb, nbl - b), nw = (nb + 31) / 32;
      bits.assign(nw, 0);
      const u32 *mrow = L.mask.data() + (size_t)lane * L.mask_words;
      const u16 *offrow = L.off.data() + (size_t)lane * L.nf;
      const u32 kP = P == 1 ? 0 : mod64(k0, P, Pmg, Pc32);
      for (u32 w = 0; w < nw; ++w) {
        u32 word = 0xFFFFFFFFu;
        if (P != 1) {
          const u32 bb = mod32(kP + 32u * w, P, Pmg);
          word = alignbit(mrow[(bb >> 5) + 1], mrow[bb >> 5], bb & 31);
        }
        if (w == nw - 1 && (nb & 31)) word &= (1u << (nb & 31)) - 1;
        pos += __builtin_popcount(word);
        bits[w] = word;
      }
      // Byte-map marking as the kernel does it: one filter per thread, from its first hit
      // in the tile, stride s.
      std::vector<uint8_t> map(nw * 32, 0);
      for (u32 w = 0; w < nw; ++w)
        for (u32 b = 0; b < 32; ++b) map[w * 32 + b] = (bits[w] >> b) & 1;
      auto first_hit = [&](u32 j, u32 o) {
        const u32 s = L.step[j];
        const u32 km = mod64(k0, s, L.magic[j], L.c32[j]);
        return o >= km ? o - km : o + s - km;
      };
      auto mark = [&](u32 from, u32 to) {
        for (u32 j = from; j < to; ++j) {
          const u32 o = offrow[j];
          if (o == kNoHit) continue;
          const u32 s = L.step[j];
          for (u32 i = first_hit(j, o); i < nb; i += s) map[i] = 0;
        }
      };
      auto pack = [&]() {
        for (u32 w = 0; w < nw; ++w) {
          u32 x = 0;
          for (u32 b = 0; b < 32; ++b) x |= (u32)(map[w * 32 + b] != 0) << b;
          bits[w] = x;
        }
      };
      mark(0, L.ncov);
      pack();
      u64 a = 0;
      for (u32 w = 0; w < nw; ++w) a += __builtin_popcount(bits[w]);
      if (has_sieve) {
        mark(L.ncov, L.nf);
        pack();
      }
      u64 s = 0;
      for (u32 w = 0; w < nw; ++w) {
        u32 x = bits[w];
        s += __builtin_popcount(x);
        while (x) {
          const u32 i = __builtin_ctz(x);
          x &= x - 1;
          surv.push_back(L.lanes[lane] + Mp * (k0 + 32ull * w + i));
        }
      }
      aft += has_sieve ? a : s;
      sur += s;
    }
  }
  std::sort(surv.begin(), surv.end());
  for (u64 p : surv) printf("S %" PRIu64 "\n", p);
  printf("C %" PRIu64 " %" PRIu64 " %" PRIu64 "\n", pos, aft, sur);
  return 0;
}
