// es_plan.cpp -- certificate families, wheel, filter set and sub-wheel layouts.
// Follows sieve/rung_scan3.cpp (line references on each function); all of it runs once
// per context on the host and feeds the tables the device kernels read.
#include "es_plan.h"

#include <algorithm>
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>

namespace es {

namespace {

u64 gcd64(u64 a, u64 b) {
  while (b) { const u64 t = a % b; a = b; b = t; }
  return a;
}

// a^{-1} mod m for gcd(a, m) == 1 (extended Euclid; strides are often composite).
bool inverse_mod(u64 a, u64 m, u64 &inv) {
  if (m == 1) { inv = 0; return true; }
  int64_t r0 = (int64_t)m, r1 = (int64_t)(a % m), t0 = 0, t1 = 1;
  while (r1) {
    const int64_t q = r0 / r1;
    int64_t t = r0 - q * r1; r0 = r1; r1 = t;
    t = t0 - q * t1; t0 = t1; t1 = t;
  }
  if (r0 != 1) return false;
  inv = (u64)(t0 < 0 ? t0 + (int64_t)m : t0);
  return true;
}

// Keeps, per (m, res), the cover with the smallest pmin; the first seen wins ties.
struct CoverTable {
  std::map<std::pair<u64, u64>, Cover> best;
  void offer(u64 a, u64 c, u64 d) {
    const u128 m = (u128)4 * a * c * d - 1, b = (u128)4 * a * a * d;
    const u64 res = (u64)((m - b % m) % m);
    const u64 pmin = (u64)(m * (a / c + 1) - b);
    const auto key = std::make_pair((u64)m, res);
    auto it = best.find(key);
    if (it == best.end() || pmin < it->second.pmin)
      best[key] = Cover{(u64)m, res, pmin, (u32)a, (u32)c, (u32)d};
  }
  std::vector<Cover> sorted() const {
    std::vector<Cover> v;
    v.reserve(best.size());
    for (const auto &kv : best) v.push_back(kv.second);
    return v;
  }
};

}  // namespace

// rung_scan3.cpp:93-112 -- every (a,c,d) with 4acd-1 <= mmax; loop order a, d, c.
std::vector<Cover> covers_upto(u64 mmax) {
  CoverTable t;
  const u64 n = (mmax + 1) / 4;
  for (u64 a = 1; a <= n; ++a)
    for (u64 d = 1; a * d <= n; ++d)
      for (u64 c = 1; a * d * c <= n; ++c) t.offer(a, c, d);
  return t.sorted();
}

// rung_scan3.cpp:126-165 -- moduli m | M with m == 3 (mod 4), acd = (m+1)/4, a and c
// ascending over the divisors.
std::vector<Cover> covers_dividing(u64 M) {
  std::vector<u64> divs;
  for (u64 x = 1; x * x <= M; ++x)
    if (M % x == 0) {
      divs.push_back(x);
      if (x * x != M) divs.push_back(M / x);
    }
  std::sort(divs.begin(), divs.end());
  CoverTable t;
  for (u64 m : divs) {
    if (m % 4 != 3) continue;
    const u64 acd = (m + 1) / 4;
    for (u64 a = 1; a <= acd; ++a) {
      if (acd % a) continue;
      const u64 cd = acd / a;
      for (u64 c = 1; c <= cd; ++c)
        if (cd % c == 0) t.offer(a, c, cd / c);
    }
  }
  return t.sorted();
}

// rung_scan3.cpp:177-202 -- for coprime (r, u), L = 4ru; class rho qualifies when
// r == -rho (mod 4), u | (pA)^2 and r | pA + u at p = rho and p = rho + L. Smallest
// (r, u) per (L, rho).
std::vector<Ucert> uniform_certs(u32 rmax, u32 umax) {
  std::map<std::pair<u64, u64>, Ucert> best;
  for (u32 r = 1; r <= rmax; ++r)
    for (u32 u = 1; u <= umax; ++u) {
      if (gcd64(r, u) != 1) continue;
      const u64 L = 4ull * r * u;
      for (u64 rho = (4 - r % 4) % 4; rho < L; rho += 4) {
        bool ok = true;
        for (u64 lift = 0; ok && lift < 2; ++lift) {
          const u64 p = rho + lift * L, A = (p + r) / 4, q = p * A, qu = q % u;
          ok = (qu * qu) % u == 0 && (q + u) % r == 0;
        }
        if (!ok) continue;
        const auto key = std::make_pair(L, rho);
        auto it = best.find(key);
        if (it == best.end() || r < it->second.r || (r == it->second.r && u < it->second.u))
          best[key] = Ucert{L, rho, r, u};
      }
    }
  std::vector<Ucert> v;
  v.reserve(best.size());
  for (const auto &kv : best) v.push_back(kv.second);
  return v;
}

// rung_scan3.cpp:211-226
Wheel build_wheel(u64 M, const std::vector<Cover> &cov, const std::vector<Ucert> &uc) {
  Wheel w;
  std::vector<uint8_t> dead(M, 0);
  for (const Cover &f : cov) {
    if (M % f.m) continue;
    w.max_pmin = std::max(w.max_pmin, f.pmin);
    for (u64 x = f.res; x < M; x += f.m) dead[x] = 1;
  }
  std::vector<uint8_t> unit(M, 0);
  for (u64 x = 1; x < M; ++x) unit[x] = gcd64(x, M) == 1;
  for (u64 x = 1; x < M; ++x) w.f1_only += unit[x] && !dead[x];
  for (const Ucert &c : uc) {
    if (M % c.L) continue;
    for (u64 x = c.res; x < M; x += c.L) dead[x] = 1;
  }
  for (u64 x = 1; x < M; ++x)
    if (unit[x] && !dead[x]) w.classes.push_back(x);
  return w;
}

// rung_scan3.cpp:228-236
u64 fnv64_classes(const std::vector<u64> &cls) {
  u64 h = 14695981039346656037ull;
  char buf[32];
  for (size_t i = 0; i < cls.size(); ++i) {
    const int n = snprintf(buf, sizeof buf, i ? ",%" PRIu64 : "%" PRIu64, cls[i]);
    for (int j = 0; j < n; ++j) { h ^= (unsigned char)buf[j]; h *= 1099511628211ull; }
  }
  return h;
}

// rung_scan3.cpp:57-67
const ClassPin *class_pin(u64 M) {
  static const ClassPin pins[] = {
      {120120, 0xdbc4ab9b16d73f76ull,
       "154e82672d9da66cff97d48079047b8b0b0a31fa049ac62fb05b5c82ad540f60"},
      {2042040, 0xb9b5d1408da3428eull,
       "a77b2187201ff96ff3ecf822150f4152d841941deeb0cf616238bcbb0a648bbc"},
  };
  for (const ClassPin &p : pins)
    if (p.M == M) return &p;
  return nullptr;
}

// rung_scan3.cpp:1085-1097 -- "FILTER mod res ..." lines; everything else ignored.
bool load_filter_file(const char *path, FilterSet &fs, std::string &msg) {
  FILE *fp = fopen(path, "r");
  if (!fp) { msg = std::string("FATAL: cannot open ") + path; return false; }
  char line[256];
  while (fgets(line, sizeof line, fp)) {
    unsigned long long m, r;
    if (sscanf(line, "FILTER %llu %llu", &m, &r) == 2) { fs.mod.push_back(m); fs.res.push_back(r); }
  }
  fclose(fp);
  if (fs.mod.empty()) { msg = std::string("FATAL: no FILTER lines in ") + path; return false; }
  fs.ncov = fs.mod.size();
  return true;
}

// rung_scan3.cpp:1098-1102 -- generated covers whose modulus does not divide M.
void generated_filters(u64 M, const std::vector<Cover> &covers, FilterSet &fs) {
  for (const Cover &f : covers)
    if (M % f.m) { fs.mod.push_back(f.m); fs.res.push_back(f.res); }
  fs.ncov = fs.mod.size();
}

// rung_scan3.cpp:664-674 -- primes q <= P with q not dividing M, as (q, 0).
void append_small_primes(u64 P, u64 M, FilterSet &fs) {
  std::vector<uint8_t> comp(P + 1, 0);
  for (u64 i = 2; i <= P; ++i) {
    if (comp[i]) continue;
    if (M % i) { fs.mod.push_back(i); fs.res.push_back(0); }
    for (u64 j = i * i; j <= P; j += i) comp[j] = 1;
  }
}

void finish_filter_digest(FilterSet &fs) {
  u64 h = 14695981039346656037ull;
  char b[48];
  for (size_t i = 0; i < fs.ncov; ++i) {
    const int n = snprintf(b, sizeof b, "%" PRIu64 ",%" PRIu64 ";", fs.mod[i], fs.res[i]);
    for (int j = 0; j < n; ++j) { h ^= (unsigned char)b[j]; h *= 1099511628211ull; }
  }
  fs.fnv = h;
}

std::vector<u64> layout_candidates(u64 M) {
  std::vector<u64> out;
  if (M % 840 == 0) {
    const u64 q = M / 840;
    for (u64 d = 1; d <= q; ++d)
      if (q % d == 0) out.push_back(M / d);
  } else {
    out.push_back(M);
  }
  std::sort(out.begin(), out.end());
  return out;
}

bool build_layout(u64 M, u64 Mp, const std::vector<u64> &classes, const FilterSet &fs,
                  Layout &L, std::string &msg) {
  if (Mp == 0 || M % Mp) { msg = "layout modulus must divide M"; return false; }
  L = Layout();
  L.Mp = Mp;
  L.P = M / Mp;
  if (L.P >= 65536) { msg = "layout period too large"; return false; }
  for (u64 c : classes) L.lanes.push_back(c % Mp);
  std::sort(L.lanes.begin(), L.lanes.end());
  L.lanes.erase(std::unique(L.lanes.begin(), L.lanes.end()), L.lanes.end());
  const size_t nl = L.lanes.size();

  // Wheel mask: bit k of lane c' says whether c' + Mp*k (mod M) is a wheel class. The
  // pattern is stored for k in [0, P + 64) so any 32-bit window starting below P reads
  // whole words.
  std::vector<uint8_t> in_wheel(M, 0);
  for (u64 c : classes) in_wheel[c] = 1;
  L.mask_words = (u32)((L.P + 64 + 31) / 32 + 1);
  L.mask.assign(nl * L.mask_words, 0);
  for (size_t i = 0; i < nl; ++i)
    for (u64 k = 0; k < L.P + 64; ++k)
      if (in_wheel[(L.lanes[i] + Mp * (k % L.P)) % M])
        L.mask[i * L.mask_words + k / 32] |= 1u << (k % 32);

  // Marking order: covers sorted by stride, then primes sorted by stride. The order is
  // free (clearing is idempotent); sorted strides balance the threads of a wave. A filter
  // of stride b < kSplitStride is marked as k = ceil(kSplitStride / b) pieces of stride
  // s = k*b, piece t taking the hits o + t*b (mod s): the same positions, in runs as long
  // as those of the larger strides they are sorted among. Short strides are few, so a bin
  // of 64 filters would otherwise mix strides whose run lengths differ by half or more,
  // and the wave would wait on the longest.
  struct F { u32 idx, k, t; u64 g, b, s; };
  std::vector<F> cov, pr;
  for (size_t i = 0; i < fs.mod.size(); ++i) {
    const u64 m = fs.mod[i];
    if (m == 0) { msg = "zero filter modulus"; return false; }
    const u64 g = gcd64(m, Mp), b = m / g;
    const u32 k = b < kSplitStride ? (u32)((kSplitStride + b - 1) / b) : 1u;
    if (k * b >= 65535) { msg = "filter stride does not fit the 16-bit offset tables"; return false; }
    for (u32 t = 0; t < k; ++t) (i < fs.ncov ? cov : pr).push_back(F{(u32)i, k, t, g, b, k * b});
  }
  auto by_s = [](const F &a, const F &b) { return a.s != b.s ? a.s < b.s : a.idx != b.idx ? a.idx < b.idx : a.t < b.t; };
  std::sort(cov.begin(), cov.end(), by_s);
  std::sort(pr.begin(), pr.end(), by_s);
  std::vector<F> all = cov;
  all.insert(all.end(), pr.begin(), pr.end());
  L.nf = (u32)all.size();
  L.ncov = (u32)cov.size();
  L.order.resize(L.nf);
  L.step.resize(L.nf);
  L.magic.resize(L.nf);
  L.c32.resize(L.nf);
  std::vector<u64> inv(L.nf);
  for (u32 j = 0; j < L.nf; ++j) {
    const F &f = all[j];
    L.order[j] = f.idx;
    L.step[j] = (u32)f.s;
    L.magic[j] = f.s == 1 ? 0xFFFFFFFFu : (u32)((1ull << 32) / f.s);
    L.c32[j] = (u32)((1ull << 32) % f.s);
    if (!inverse_mod((Mp / f.g) % f.b, f.b, inv[j])) { msg = "non-invertible layout stride"; return false; }
  }
  L.off.assign(nl * (size_t)L.nf, kNoHit);
  double marks = 0;
  for (size_t i = 0; i < nl; ++i) {
    const u64 c = L.lanes[i];
    for (u32 j = 0; j < L.nf; ++j) {
      const F &f = all[j];
      const u64 m = fs.mod[f.idx];
      const u64 rhs = (fs.res[f.idx] % m + m - c % m) % m;   // Mp*k == res - c' (mod m)
      if (rhs % f.g) continue;
      const u64 o = f.b == 1 ? 0 : (u64)((u128)(rhs / f.g) * inv[j] % f.b);
      L.off[i * L.nf + j] = (u16)(o + f.t * f.b);
      marks += 1.0 / (double)f.s;
      ++L.hit_pairs;
    }
  }
  L.marks_per_int = marks / (double)Mp;
  L.bits_per_int = (double)nl / (double)Mp;
  return true;
}

int build_plan(u64 M, u64 mmax, Plan &p, std::string &msg) {
  p = Plan();
  p.M = M;
  p.mmax = mmax;
  if (M < 2) { msg = "FATAL: wheel modulus must be at least 2"; return 1; }
  const std::vector<Cover> covers = covers_upto(mmax);
  p.wheel = build_wheel(M, covers_dividing(M), uniform_certs(kUcertRmax, kUcertUmax));
  for (const Cover &f : covers) p.cover_pmax = std::max(p.cover_pmax, f.pmin);
  p.pin = class_pin(M);
  if (p.pin && fnv64_classes(p.wheel.classes) != p.pin->fnv64) {
    char b[256];
    snprintf(b, sizeof b, "FATAL: class table digest mismatch at M=%" PRIu64 " -- derived fnv64 %016" PRIx64
             ", expected %016" PRIx64, M, fnv64_classes(p.wheel.classes), p.pin->fnv64);
    msg = b;
    return 1;
  }
  return 0;
}

int plan_filters(Plan &p, const char *filter_file, u64 sieve, std::string &msg) {
  p.sieve = sieve;
  p.filters = FilterSet();
  if (filter_file) {
    if (!load_filter_file(filter_file, p.filters, msg)) return 1;
    p.filter_source = filter_file;
  } else {
    generated_filters(p.M, covers_upto(p.mmax), p.filters);
    p.filter_source = "generated";
  }
  if (sieve) append_small_primes(sieve, p.M, p.filters);
  finish_filter_digest(p.filters);
  return 0;
}

int check_interval(const Plan &p, u64 lo, u64 hi, std::string &msg) {
  char b[256];
  if (hi > kDirectPMax) {       // rung_scan3.cpp:1057-1062
    snprintf(b, sizeof b, "FATAL: HI must not exceed %" PRIu64 ", the stage-E ceiling", kDirectPMax);
    msg = b;
    return 1;
  }
  const u64 fl = std::max(p.wheel.max_pmin, p.cover_pmax);
  if (lo <= fl) {               // rung_scan3.cpp:1064-1069
    snprintf(b, sizeof b, "FATAL: LO must exceed max pmin of the filter table (%" PRIu64 "); "
             "the region below is settled by the existing censuses", fl);
    msg = b;
    return 1;
  }
  return 0;
}

}  // namespace es
