// es_prune.cpp -- chooses a certified filter table for a wheel under a device cost model
// (the role of sieve/prune_filters.py --device, with MI355X constants), and writes it in
// the filter-file format rung_scan3 and es_sieve read (FILTER mod res kind lines).
//
// Soundness does not depend on the choice: every candidate is a certified family member
// (an F1 cover m = 4acd-1 <= mmax, or an F2 uniform rung certificate), so a table can
// only leave more survivors than the full family, never certify a wrong class. The
// choice decides speed only.
//
// Domain (as prune_filters.py:1-88 defines it): every lane of the wheel, `npos` positions
// per lane, anchored at lo_ref (position J of lane c is the J-th p >= lo_ref with
// p == c mod M). Exact per-position bookkeeping in bitsets: which positions are still
// alive, and which alive ones are prime (sieved exactly below sqrt of the domain's top).
//
// Cost per lane-position of keeping filter i (ns):
//   c_mark * (its marks on the domain) / positions  +  c_pair / tile_positions
// (the marking kernel pays a per-(tile, filter) setup whether or not the filter hits).
// Benefit per lane-position: (alive positions it kills) / positions *
//   (c_comp + phi * (c_prime - c_comp)),
// phi = the live prime fraction of the alive positions, measured on the domain and kept
// non-increasing (prune_filters.py:422-552 explains why the aggregate fraction, not each
// filter's own prime kills, is the right weight). Lazy greedy (CELF): repeatedly take the
// filter with the best benefit/cost, stop when the best no longer pays for itself.
//
// usage: es_prune [--wheel M] [--mmax N] [--lo-ref X] [--npos N] [--threads T]
//                 [--ns-mark c] [--ns-pair c] [--tile-positions n] [--ns-comp c]
//                 [--ns-prime c] [--tag NAME] [--cost-form ref] --out FILE
// --cost-form ref charges (1/step + 1/262144) marks per position for every filter, as
// prune_filters.py:503-511 does, instead of its exact marks on the domain; with the
// reference's L4 constants it reproduces the reference's published GPU table.
#include "es_plan.h"

#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <queue>
#include <string>
#include <thread>
#include <omp.h>
#include <vector>

namespace {

using namespace es;

// ---------------------------------------------------------------- SHA-256 (FIPS 180-4)
struct Sha256 {
  uint32_t h[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                   0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
  uint8_t buf[64];
  size_t blen = 0;
  uint64_t total = 0;
  static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }
  void block(const uint8_t *p) {
    static const uint32_t k[64] = {
        0x428a2f98u, 0x71374491u, 0xb5c0fbcfu, 0xe9b5dba5u, 0x3956c25bu, 0x59f111f1u, 0x923f82a4u, 0xab1c5ed5u,
        0xd807aa98u, 0x12835b01u, 0x243185beu, 0x550c7dc3u, 0x72be5d74u, 0x80deb1feu, 0x9bdc06a7u, 0xc19bf174u,
        0xe49b69c1u, 0xefbe4786u, 0x0fc19dc6u, 0x240ca1ccu, 0x2de92c6fu, 0x4a7484aau, 0x5cb0a9dcu, 0x76f988dau,
        0x983e5152u, 0xa831c66du, 0xb00327c8u, 0xbf597fc7u, 0xc6e00bf3u, 0xd5a79147u, 0x06ca6351u, 0x14292967u,
        0x27b70a85u, 0x2e1b2138u, 0x4d2c6dfcu, 0x53380d13u, 0x650a7354u, 0x766a0abbu, 0x81c2c92eu, 0x92722c85u,
        0xa2bfe8a1u, 0xa81a664bu, 0xc24b8b70u, 0xc76c51a3u, 0xd192e819u, 0xd6990624u, 0xf40e3585u, 0x106aa070u,
        0x19a4c116u, 0x1e376c08u, 0x2748774cu, 0x34b0bcb5u, 0x391c0cb3u, 0x4ed8aa4au, 0x5b9cca4fu, 0x682e6ff3u,
        0x748f82eeu, 0x78a5636fu, 0x84c87814u, 0x8cc70208u, 0x90befffau, 0xa4506cebu, 0xbef9a3f7u, 0xc67178f2u};
    uint32_t w[64];
    for (int i = 0; i < 16; ++i)
      w[i] = (uint32_t)p[4 * i] << 24 | (uint32_t)p[4 * i + 1] << 16 | (uint32_t)p[4 * i + 2] << 8 | p[4 * i + 3];
    for (int i = 16; i < 64; ++i) {
      const uint32_t s0 = rotr(w[i - 15], 7) ^ rotr(w[i - 15], 18) ^ (w[i - 15] >> 3);
      const uint32_t s1 = rotr(w[i - 2], 17) ^ rotr(w[i - 2], 19) ^ (w[i - 2] >> 10);
      w[i] = w[i - 16] + s0 + w[i - 7] + s1;
    }
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3], e = h[4], f = h[5], g = h[6], hh = h[7];
    for (int i = 0; i < 64; ++i) {
      const uint32_t t1 = hh + (rotr(e, 6) ^ rotr(e, 11) ^ rotr(e, 25)) + ((e & f) ^ (~e & g)) + k[i] + w[i];
      const uint32_t t2 = (rotr(a, 2) ^ rotr(a, 13) ^ rotr(a, 22)) + ((a & b) ^ (a & c) ^ (b & c));
      hh = g; g = f; f = e; e = d + t1; d = c; c = b; b = a; a = t1 + t2;
    }
    h[0] += a; h[1] += b; h[2] += c; h[3] += d; h[4] += e; h[5] += f; h[6] += g; h[7] += hh;
  }
  void update(const std::string &s) {
    for (unsigned char ch : s) {
      buf[blen++] = ch;
      ++total;
      if (blen == 64) { block(buf); blen = 0; }
    }
  }
  std::string hex() {
    const uint64_t bits = total * 8;
    uint8_t pad = 0x80;
    std::string tail(1, (char)pad);
    while ((total + tail.size()) % 64 != 56) tail.push_back('\0');
    for (int i = 7; i >= 0; --i) tail.push_back((char)(bits >> (8 * i)));
    for (unsigned char ch : tail) {
      buf[blen++] = ch;
      if (blen == 64) { block(buf); blen = 0; }
    }
    char out[65];
    for (int i = 0; i < 8; ++i) snprintf(out + 8 * i, 9, "%08x", h[i]);
    return std::string(out, 64);
  }
};

u64 gcd64(u64 a, u64 b) {
  while (b) { const u64 t = a % b; a = b; b = t; }
  return a;
}

// a^{-1} mod m (gcd(a, m) == 1), extended Euclid.
u64 inv_mod(u64 a, u64 m) {
  long long t = 0, nt = 1, r = (long long)m, nr = (long long)(a % m);
  while (nr) {
    const long long q = r / nr;
    long long x = t - q * nt; t = nt; nt = x;
    x = r - q * nr; r = nr; nr = x;
  }
  return (u64)(t < 0 ? t + (long long)m : t);
}

struct Cand {
  u64 mod, res;
  const char *kind;
  u64 g, step, inv;   // gcd(M, mod), mod / g, ((M/g) mod step)^{-1} mod step
  u64 marks = 0;      // hits on the whole domain
};

struct Opts {
  u64 M = 2042040, mmax = 2000, lo_ref = 1000000000000ull;
  u32 npos = 65536, threads = 0;
  double ns_mark = 0, ns_pair = 0, tile_positions = 8192, ns_comp = 0, ns_prime = 0;
  std::string out, tag = "mi355x";
  bool ref_cost = false;   // the reference pruner's cost form: (1/step + 1/seg) marks per position
  double seg = 262144;
};

// Parallel sum of f(i) over [0, n) on T threads (OpenMP: one pool for the whole run).
template <class F>
u64 par_sum(size_t n, unsigned T, F f) {
  u64 s = 0;
#pragma omp parallel for num_threads(T) reduction(+ : s) schedule(static)
  for (long long i = 0; i < (long long)n; ++i) s += f((size_t)i);
  return s;
}

int run(const Opts &o) {
  Plan p;
  std::string msg;
  if (build_plan(o.M, 10000, p, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 1; }
  const std::vector<u64> &lanes = p.wheel.classes;
  const size_t L = lanes.size();
  const u32 npos = o.npos, W = (npos + 63) / 64;
  const unsigned T = o.threads ? o.threads : std::max(1u, std::thread::hardware_concurrency());
  const double pos = (double)L * npos;

  // Candidates: F1 covers with m <= mmax and m not dividing M, F2 certificates with L not
  // dividing M (those dividing M are already spent on the wheel).
  std::vector<Cand> cand;
  for (const Cover &c : covers_upto(o.mmax))
    if (o.M % c.m) cand.push_back(Cand{c.m, c.res, "F1", 0, 0, 0});
  const size_t nf1 = cand.size();
  for (const Ucert &u : uniform_certs(kUcertRmax, kUcertUmax))
    if (o.M % u.L) cand.push_back(Cand{u.L, u.res, "F2", 0, 0, 0});
  for (Cand &c : cand) {
    c.g = gcd64(o.M, c.mod);
    c.step = c.mod / c.g;
    c.inv = c.step == 1 ? 0 : inv_mod((o.M / c.g) % c.step, c.step);
  }
  const size_t N = cand.size();
  fprintf(stderr, "candidates: %zu F1 (mmax %" PRIu64 ") + %zu F2, %zu lanes x %u positions from %" PRIu64 "\n",
          nf1, o.mmax, N - nf1, L, npos, o.lo_ref);

  // First position of each lane at or above lo_ref, and each candidate's first hit per
  // lane (0xFFFFFFFF: the lane misses it).
  std::vector<u64> first(L);
  for (size_t l = 0; l < L; ++l) first[l] = o.lo_ref + ((lanes[l] + o.M - o.lo_ref % o.M) % o.M);
  std::vector<uint32_t> j0((size_t)N * L);
  par_sum(N, T, [&](size_t i) {
    const Cand &c = cand[i];
    u64 marks = 0;
    for (size_t l = 0; l < L; ++l) {
      const u64 rhs = (c.res % c.mod + c.mod - first[l] % c.mod) % c.mod;
      uint32_t j = 0xFFFFFFFFu;
      if (rhs % c.g == 0) {
        j = (uint32_t)((u128)(rhs / c.g) * c.inv % c.step);
        if (j < npos) marks += (npos - 1 - j) / c.step + 1;
        else j = 0xFFFFFFFFu;
      }
      j0[i * L + l] = j;
    }
    cand[i].marks = marks;
    return (u64)0;
  });

  // Alive positions, and the alive ones that are prime: sieve every position by the
  // primes q <= sqrt(top) not dividing M (a position divisible by such q is composite,
  // since every position exceeds lo_ref > q).
  std::vector<u64> alive((size_t)L * W, ~0ull), prime((size_t)L * W, ~0ull);
  if (npos % 64) {
    const u64 last = (1ull << (npos % 64)) - 1;
    for (size_t l = 0; l < L; ++l) alive[l * W + W - 1] = prime[l * W + W - 1] = last;
  }
  const u64 top = o.lo_ref + o.M * (u64)npos;
  u64 qmax = (u64)std::sqrt((double)top);
  while ((u128)qmax * qmax > top) --qmax;
  while ((u128)(qmax + 1) * (qmax + 1) <= top) ++qmax;
  std::vector<uint8_t> comp(qmax + 1, 0);
  std::vector<u64> qs;
  for (u64 q = 2; q <= qmax; ++q) {
    if (comp[q]) continue;
    if (o.M % q) qs.push_back(q);
    for (u64 j = q * q; j <= qmax; j += q) comp[j] = 1;
  }
  std::vector<u64> minv(qs.size());
  for (size_t k = 0; k < qs.size(); ++k) minv[k] = inv_mod(o.M % qs[k], qs[k]);
  par_sum(L, T, [&](size_t l) {
    u64 *pr = &prime[l * W];
    for (size_t k = 0; k < qs.size(); ++k) {
      const u64 q = qs[k];
      const u64 j = (u64)((u128)((q - first[l] % q) % q) * minv[k] % q);
      for (u64 jj = j; jj < npos; jj += q) pr[jj >> 6] &= ~(1ull << (jj & 63));
    }
    return (u64)0;
  });
  u64 alive_n = (u64)L * npos;
  u64 prime_n = par_sum(L, T, [&](size_t l) {
    u64 s = 0;
    for (u32 w = 0; w < W; ++w) s += __builtin_popcountll(prime[l * W + w]);
    return s;
  });
  fprintf(stderr, "domain: %.0f positions, %" PRIu64 " prime (%.4f)\n", pos, prime_n, prime_n / pos);

  auto gain = [&](size_t i) {
    const Cand &c = cand[i];
    return par_sum(L, T, [&](size_t l) {
      const uint32_t j = j0[i * L + l];
      if (j == 0xFFFFFFFFu) return (u64)0;
      const u64 *a = &alive[l * W];
      u64 s = 0;
      for (u64 jj = j; jj < npos; jj += c.step) s += (a[jj >> 6] >> (jj & 63)) & 1;
      return s;
    });
  };
  double phi = prime_n / (double)alive_n;
  auto per_kill = [&]() { return o.ns_comp + phi * (o.ns_prime - o.ns_comp); };
  auto cost = [&](size_t i) {
    if (o.ref_cost) return (1.0 / cand[i].step + 1.0 / o.seg) * o.ns_mark + o.ns_pair / o.tile_positions;
    return o.ns_mark * cand[i].marks / pos + o.ns_pair / o.tile_positions;
  };
  auto score = [&](size_t i) { return gain(i) * per_kill() / pos / cost(i); };

  // Lazy greedy: a stale heap entry over-estimates (gains only shrink and phi is kept
  // non-increasing), so re-scoring the top until it is fresh picks the true best.
  typedef std::pair<double, std::pair<size_t, u64>> Entry;   // score, (candidate, stamp)
  std::priority_queue<Entry> heap;
  for (size_t i = 0; i < N; ++i)
    if (cand[i].marks) heap.push(Entry{score(i), {i, 0}});
  std::vector<size_t> chosen;
  u64 stamp = 0, marks = 0;
  while (!heap.empty()) {
    const Entry e = heap.top();
    heap.pop();
    const size_t i = e.second.first;
    if (e.second.second != stamp) {
      heap.push(Entry{score(i), {i, stamp}});
      continue;
    }
    if (e.first <= 1.0) break;   // benefit no longer exceeds cost
    chosen.push_back(i);
    marks += cand[i].marks;
    const Cand &c = cand[i];
    const u64 killed = par_sum(L, T, [&](size_t l) {
      const uint32_t j = j0[i * L + l];
      if (j == 0xFFFFFFFFu) return (u64)0;
      u64 *a = &alive[l * W], *pr = &prime[l * W];
      u64 k = 0;
      for (u64 jj = j; jj < npos; jj += c.step) {
        const u64 bit = 1ull << (jj & 63);
        k += (a[jj >> 6] & bit) != 0;
        a[jj >> 6] &= ~bit;
        pr[jj >> 6] &= ~bit;
      }
      return k;
    });
    alive_n -= killed;
    prime_n = par_sum(L, T, [&](size_t l) {
      u64 s = 0;
      for (u32 w = 0; w < W; ++w) s += __builtin_popcountll(prime[l * W + w]);
      return s;
    });
    phi = std::min(phi, alive_n ? prime_n / (double)alive_n : 0.0);
    ++stamp;
    if (chosen.size() % 250 == 0)
      fprintf(stderr, "  %zu filters: %.3f marks/pos, survivor density %.3e, prime density %.3e\n", chosen.size(),
              marks / pos, alive_n / pos, prime_n / pos);
  }
  size_t c1 = 0;
  for (size_t i : chosen) c1 += cand[i].kind[1] == '1';
  fprintf(stderr, "chosen: %zu filters (%zu F1 + %zu F2), %.4f marks/pos, survivor density %.6e, "
                  "prime density %.6e\n", chosen.size(), c1, chosen.size() - c1, marks / pos, alive_n / pos,
          prime_n / pos);

  // The table: header comments, then the filters in choice order; the sha256 of that text
  // is inserted as its own comment line before the summary line.
  char line[512];
  std::string head, body;
  head += "# certified filter table chosen by es_prune (lazy greedy weighted set cover) under an MI355X\n";
  head += "# cost model. Every filter is a certified F1 cover or F2 uniform rung certificate, so the\n";
  head += "# choice changes speed only, never which classes are certified.\n";
  snprintf(line, sizeof line, "# M=%" PRIu64 " mmax=%" PRIu64 " npos=%u lanes=%zu lo_ref=%" PRIu64 "\n", o.M, o.mmax,
           npos, L, o.lo_ref);
  head += line;
  snprintf(line, sizeof line,
           "# ns_per_mark=%g ns_per_tile_filter=%g tile_positions=%g ns_per_survivor_composite=%g "
           "ns_per_survivor_prime=%g cost_form=%s device=%s\n",
           o.ns_mark, o.ns_pair, o.tile_positions, o.ns_comp, o.ns_prime, o.ref_cost ? "per_step" : "exact_marks",
           o.tag.c_str());
  head += line;
  std::string summary;
  snprintf(line, sizeof line, "# filters=%zu marks_per_pos=%.4f survivor_density=%.6e prime_density=%.6e\n",
           chosen.size(), marks / pos, alive_n / pos, prime_n / pos);
  summary = line;
  for (size_t i : chosen) {
    snprintf(line, sizeof line, "FILTER %" PRIu64 " %" PRIu64 " %s\n", cand[i].mod, cand[i].res, cand[i].kind);
    body += line;
  }
  Sha256 sh;
  sh.update(head + summary + body);
  const std::string digest = sh.hex();
  FILE *f = fopen(o.out.c_str(), "w");
  if (!f) { fprintf(stderr, "cannot write %s\n", o.out.c_str()); return 1; }
  fputs(head.c_str(), f);
  fprintf(f, "# sha256=%s\n", digest.c_str());
  fputs(summary.c_str(), f);
  fputs(body.c_str(), f);
  fclose(f);
  printf("wrote %s sha256=%s\n", o.out.c_str(), digest.c_str());
  return 0;
}

}  // namespace

#ifndef ES_PRUNE_NO_MAIN
int main(int argc, char **argv) {
  Opts o;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](void) -> const char * {
      if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(1); }
      return argv[++i];
    };
    if (a == "--wheel") o.M = strtoull(next(), 0, 10);
    else if (a == "--mmax") o.mmax = strtoull(next(), 0, 10);
    else if (a == "--lo-ref") o.lo_ref = strtoull(next(), 0, 10);
    else if (a == "--npos") o.npos = (u32)strtoul(next(), 0, 10);
    else if (a == "--threads") o.threads = (u32)strtoul(next(), 0, 10);
    else if (a == "--ns-mark") o.ns_mark = atof(next());
    else if (a == "--ns-pair") o.ns_pair = atof(next());
    else if (a == "--tile-positions") o.tile_positions = atof(next());
    else if (a == "--ns-comp") o.ns_comp = atof(next());
    else if (a == "--ns-prime") o.ns_prime = atof(next());
    else if (a == "--tag") o.tag = next();
    else if (a == "--cost-form") o.ref_cost = std::string(next()) == "ref";
    else if (a == "--out") o.out = next();
    else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 1; }
  }
  if (o.out.empty() || o.ns_mark <= 0 || o.ns_prime <= 0) {
    fprintf(stderr, "usage: es_prune --ns-mark c --ns-pair c --ns-comp c --ns-prime c [--wheel M] [--mmax N] "
                    "[--lo-ref X] [--npos N] [--tile-positions n] [--threads T] [--tag NAME] --out FILE\n");
    return 1;
  }
  return run(o);
}
#endif
