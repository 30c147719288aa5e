// es_oracle.cpp -- CPU restatement of the reference verification scanner.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker the GPU product is compared
// against; it is built into oracle/es_oracle and used by tests/, by smoke() and by the
// bench's CPU-baseline leg. The product (erdos-strauss_amd/) never links, loads or runs it.
//
// It restates sieve/rung_scan3.cpp (with the primitives it takes from
// census/rung_scan.cpp) function by function, in deliberately plain arithmetic: every
// modular product is a u128 `%`, factorization is trial division plus Brent's rho on
// that product. It shares no code with the product, whose host tail is Montgomery-based,
// so agreement between the two is a differential check rather than a tautology.
//
// Pinning (what makes this restatement trustworthy, see tests/test_oracle.py):
//   * records byte-identical to the reference binary built from source (oracle/_ref)
//     on the benchmark intervals and on random windows;
//   * the reference's own golden records data/rederive/decades/d6..d19 and the RUNG
//     lines of w19/w64/s0/s69 (each prime re-solved to the identical line);
//   * --dump-classes byte-identical to tables/class_table_{120120,2042040}.txt;
//   * the in-binary KATs of rung_scan3.cpp:821-897 (--verify);
//   * primality against a sieve of Eratosthenes and strong-pseudoprime lists.
//
// CLI: identical to rung_scan3 (rung_scan3.cpp:972-1016), plus two checker modes:
//   --dump-survivors   after the normal scan options: print "S p isprime" per stage-C
//                      survivor in lane order instead of the record (kernel parity)
//   --solve P...       print the stage D/E line for each given prime (golden lines)
//   --solve-direct P.. print the stage E line alone (the reference's --check-direct rows)
//   --batch FILE       one record per "lo hi" line of FILE, tables built once (the
//                      serving trace's golden digests)
#include <algorithm>
#include <cinttypes>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef uint64_t u64;
typedef uint32_t u32;
typedef unsigned __int128 u128;
using std::string;
using std::vector;

namespace {

// ---- constants (rung_scan3.cpp:23, :42-43, :57-63; census/rung_scan.cpp:64, :175)
const u32 kRungCap = 255;                 // RCAP
const u64 kTrialBound = 65536;            // TD_BOUND
const u32 kUcertR = 63, kUcertU = 64;     // UCERT_RMAX / UCERT_UMAX
const u64 kDirectNMax = 18446744065119617025ull;   // (2^32-1)^2
const u64 kDirectPMax = kDirectNMax - 2048 - 1;

struct Pin { u64 M; u64 fnv; const char *sha; };
const Pin kPins[] = {
  {120120, 0xdbc4ab9b16d73f76ull, "154e82672d9da66cff97d48079047b8b0b0a31fa049ac62fb05b5c82ad540f60"},
  {2042040, 0xb9b5d1408da3428eull, "a77b2187201ff96ff3ecf822150f4152d841941deeb0cf616238bcbb0a648bbc"},
};
const Pin *find_pin(u64 M) {
  for (const Pin &p : kPins) if (p.M == M) return &p;
  return nullptr;
}

// ---- plain modular arithmetic (census/rung_scan.cpp:68-74)
u64 mulm(u64 a, u64 b, u64 m) { return (u64)((u128)a * b % m); }
u64 powm(u64 b, u64 e, u64 m) {
  u64 acc = 1 % m;
  b %= m;
  for (; e; e >>= 1) {
    if (e & 1) acc = mulm(acc, b, m);
    b = mulm(b, b, m);
  }
  return acc;
}
u64 gcd_u64(u64 a, u64 b) { while (b) { u64 t = a % b; a = b; b = t; } return a; }

const u64 kSmallPrimes[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};

// One Miller-Rabin round: n odd > 2, n-1 = d*2^s. True when n is a strong probable
// prime to base a.
bool mr_round(u64 n, u64 a, u64 d, int s) {
  u64 x = powm(a, d, n);
  if (x == 1 || x == n - 1) return true;
  for (int i = 1; i < s; ++i) {
    x = mulm(x, x, n);
    if (x == n - 1) return true;
  }
  return false;
}

// census/rung_scan.cpp:76-92: trial prologue by the first twelve primes, then the
// twelve-base test; deterministic for every u64.
bool prime_q(u64 n) {
  if (n < 2) return false;
  for (u64 q : kSmallPrimes) {
    if (n == q) return true;
    if (n % q == 0) return false;
  }
  u64 d = n - 1; int s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  for (u64 a : kSmallPrimes)
    if (!mr_round(n, a, d, s)) return false;
  return true;
}

// rung_scan3.cpp:680-689: base-2 strong probable prime (the cheap first test).
bool sprp_base2(u64 n) {
  if (n < 2) return false;
  if (!(n & 1)) return n == 2;
  u64 d = n - 1; int s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  return mr_round(n, 2, d, s);
}

// ---- factorization. The reference's factorA_fast (rung_scan3.cpp:410-431) and the
// frozen factorA (census/rung_scan.cpp:194-204) both return THE factorization; so does
// this, by trial division to 65536 and Brent's rho on the cofactor. Entry order differs
// from both and cannot reach any output (minimums and sorted divisor lists only).
struct Factors {
  u64 p[24]; int e[24]; int k = 0;
  void add(u64 q, int x) {
    for (int i = 0; i < k; ++i) if (p[i] == q) { e[i] += x; return; }
    if (k >= 24) { fprintf(stderr, "oracle: too many factors\n"); abort(); }
    p[k] = q; e[k] = x; ++k;
  }
};

vector<u64> g_trial;   // odd primes 3..65536
void init_trial() {
  if (!g_trial.empty()) return;
  vector<char> comp(kTrialBound + 1, 0);
  for (u64 i = 3; i <= kTrialBound; i += 2) {
    if (comp[i]) continue;
    g_trial.push_back(i);
    for (u64 j = i * i; j <= kTrialBound; j += 2 * i) comp[j] = 1;
  }
}

u64 isqrt_exact(u64 n) {
  u64 s = (u64)std::sqrt((long double)n);
  if (s > 0xFFFFFFFFull) s = 0xFFFFFFFFull;
  while (s && s * s > n) --s;
  while (s < 0xFFFFFFFFull && (s + 1) * (s + 1) <= n) ++s;
  return s;
}

// Brent's cycle-finding rho (census/rung_scan.cpp:94-128 structure) with u128 `%`.
u64 rho_split(u64 n) {
  if (!(n & 1)) return 2;
  for (u64 c = 1;; ++c) {
    auto f = [&](u64 v) { return (u64)(((u128)v * v + c) % n); };
    u64 y = 2, x = 2, ys = 2, q = 1, g = 1, r = 1;
    const u64 m = 128;
    while (g == 1) {
      x = y;
      for (u64 i = 0; i < r; ++i) y = f(y);
      for (u64 k = 0; k < r && g == 1; k += m) {
        ys = y;
        const u64 lim = std::min(m, r - k);
        for (u64 i = 0; i < lim; ++i) {
          y = f(y);
          q = mulm(q, x > y ? x - y : y - x, n);
        }
        g = gcd_u64(q, n);
      }
      r <<= 1;
      if (r > (1ull << 26)) break;
    }
    if (g == n || g == 0) {
      g = 1;
      for (int guard = 0; g == 1 && guard < (1 << 27); ++guard) {
        ys = f(ys);
        g = gcd_u64(x > ys ? x - ys : ys - x, n);
      }
    }
    if (g != 1 && g != n && g != 0) return g;
  }
}

void split_into(u64 n, Factors &f) {
  if (n == 1) return;
  if (prime_q(n)) { f.add(n, 1); return; }
  const u64 s = isqrt_exact(n);
  if (s * s == n) {
    Factors g; split_into(s, g);
    for (int i = 0; i < g.k; ++i) f.add(g.p[i], 2 * g.e[i]);
    return;
  }
  const u64 d = rho_split(n);
  split_into(d, f);
  split_into(n / d, f);
}

void factorize(u64 n, Factors &f) {
  init_trial();
  f.k = 0;
  if (n == 0) return;
  int e2 = 0;
  while (!(n & 1)) { n >>= 1; ++e2; }
  if (e2) f.add(2, e2);
  for (u64 q : g_trial) {
    if (q * q > n) break;
    if (n % q) continue;
    int e = 0;
    while (n % q == 0) { n /= q; ++e; }
    f.add(q, e);
  }
  if (n > 1) {
    // Past the loop every prime factor of n exceeds min(sqrt(n), 65536); below 2^32 that
    // leaves n prime.
    if (n <= kTrialBound * kTrialBound) f.add(n, 1);
    else split_into(n, f);
  }
}

// ---- F1 covers (rung_scan3.cpp:91-112): m = 4acd-1, res = -4a^2 d mod m,
// pmin = m(floor(a/c)+1) - 4a^2 d; one entry per (m, res), smallest pmin first seen.
struct Cover { u64 m, res, pmin; u32 a, c, d; };

void keep_cover(std::map<std::pair<u64, u64>, Cover> &best, u64 a, u64 c, u64 d) {
  const u128 mm = (u128)4 * a * c * d - 1, bb = (u128)4 * a * a * d;
  const u64 m = (u64)mm;
  const u64 res = (u64)((mm - bb % mm) % mm);
  const u64 pmin = (u64)(mm * (a / c + 1) - bb);
  auto key = std::make_pair(m, res);
  auto it = best.find(key);
  if (it == best.end() || pmin < it->second.pmin) best[key] = Cover{m, res, pmin, (u32)a, (u32)c, (u32)d};
}

vector<Cover> flatten(const std::map<std::pair<u64, u64>, Cover> &best) {
  vector<Cover> v;
  for (auto &kv : best) v.push_back(kv.second);   // map order == ascending (m, res)
  return v;
}

// Every cover with 4acd-1 <= mmax, enumerated a, then d, then c (rung_scan3.cpp:93-98).
vector<Cover> covers_upto(u64 mmax) {
  std::map<std::pair<u64, u64>, Cover> best;
  const u64 t = (mmax + 1) / 4;
  for (u64 a = 1; a <= t; ++a)
    for (u64 d = 1; a * d <= t; ++d)
      for (u64 c = 1; a * d * c <= t; ++c) keep_cover(best, a, c, d);
  return flatten(best);
}

// Covers whose modulus divides M: moduli from the divisors of M that are 3 mod 4, then
// a | t and c | t/a in ascending order (rung_scan3.cpp:126-165).
vector<Cover> covers_dividing(u64 M) {
  vector<u64> divs;
  for (u64 x = 1; x * x <= M; ++x)
    if (M % x == 0) { divs.push_back(x); if (x != M / x) divs.push_back(M / x); }
  std::sort(divs.begin(), divs.end());
  std::map<std::pair<u64, u64>, Cover> best;
  for (u64 m : divs) {
    if (m % 4 != 3) continue;
    const u64 t = (m + 1) / 4;
    for (u64 a = 1; a <= t; ++a) {
      if (t % a) continue;
      for (u64 c = 1; c <= t / a; ++c) {
        if ((t / a) % c) continue;
        keep_cover(best, a, c, t / a / c);
      }
    }
  }
  return flatten(best);
}

// ---- F2 uniform rung certificates (rung_scan3.cpp:175-202).
struct Ucert { u64 L, res; u32 r, u; };

vector<Ucert> uniform_certs(u32 rmax, u32 umax) {
  std::map<std::pair<u64, u64>, Ucert> best;
  for (u32 r = 1; r <= rmax; ++r)
    for (u32 u = 1; u <= umax; ++u) {
      if (gcd_u64(u, r) != 1) continue;
      const u64 L = 4ull * r * u;                 // lcm(r,u) = r*u when coprime
      for (u64 rho = 0; rho < L; ++rho) {
        if ((rho + r) % 4) continue;
        bool good = true;
        for (u64 lift = 0; lift < 2 && good; ++lift) {
          const u64 p = rho + lift * L, A = (p + r) / 4, q = p * A;
          const u64 qu = q % u;
          if ((qu * qu) % u != 0 || (q + u) % r != 0) good = false;
        }
        if (!good) continue;
        auto key = std::make_pair(L, rho);
        auto it = best.find(key);
        if (it == best.end() || r < it->second.r || (r == it->second.r && u < it->second.u))
          best[key] = Ucert{L, rho, r, u};
      }
    }
  vector<Ucert> v;
  for (auto &kv : best) v.push_back(kv.second);
  return v;
}

// ---- wheel (rung_scan3.cpp:206-226)
struct Wheel { vector<u64> classes; u64 max_pmin = 0; u64 f1_only = 0; };

Wheel make_wheel(u64 M, const vector<Cover> &cov, const vector<Ucert> &uc) {
  Wheel w;
  vector<char> dead(M, 0);
  for (const Cover &f : cov) {
    if (M % f.m) continue;
    w.max_pmin = std::max(w.max_pmin, f.pmin);
    for (u64 x = f.res; x < M; x += f.m) dead[x] = 1;
  }
  for (u64 x = 1; x < M; ++x) if (!dead[x] && gcd_u64(x, M) == 1) ++w.f1_only;
  for (const Ucert &c : uc) {
    if (M % c.L) continue;
    for (u64 x = c.res; x < M; x += c.L) dead[x] = 1;
  }
  for (u64 x = 1; x < M; ++x) if (!dead[x] && gcd_u64(x, M) == 1) w.classes.push_back(x);
  return w;
}

// FNV-1a 64 over the comma-joined decimal class list (rung_scan3.cpp:228-236).
u64 fnv_classes(const vector<u64> &cls) {
  u64 h = 14695981039346656037ull;
  for (size_t i = 0; i < cls.size(); ++i) {
    string s = (i ? "," : "") + std::to_string(cls[i]);
    for (unsigned char ch : s) { h ^= ch; h *= 1099511628211ull; }
  }
  return h;
}

// ---- stage D (rung_scan3.cpp:544-603)
// Least u | q^2 with u == target (mod r) and u <= q, or 0: minimum per residue class
// carried prime power by prime power over q^2 = p^2 * prod(pr^(2e)).
u128 least_witness(u64 p, const Factors &fa, u128 q, u64 r, u64 target) {
  const u128 NONE = ~(u128)0;
  vector<u128> best(r, NONE), next(r, NONE);
  best[1 % r] = 1;
  auto absorb = [&](u64 pr, int e) {
    std::fill(next.begin(), next.end(), NONE);
    for (u64 s = 0; s < r; ++s) {
      if (best[s] == NONE) continue;
      u128 v = best[s];
      u64 res = s;
      for (int j = 0;; ++j) {
        if (v < next[res]) next[res] = v;
        if (j == e || v > q / pr) break;
        v *= pr;
        res = (u64)((u128)res * (pr % r) % r);
      }
    }
    best.swap(next);
  };
  absorb(p, 2);
  for (int i = 0; i < fa.k; ++i) absorb(fa.p[i], 2 * fa.e[i]);
  return best[target] == NONE ? 0 : best[target];
}

string u128_dec(u128 v) {
  if (!v) return "0";
  string s;
  while (v) { s.push_back((char)('0' + (int)(v % 10))); v /= 10; }
  std::reverse(s.begin(), s.end());
  return s;
}

bool rung_line(u64 p, string &out) {
  u64 r = (4 - p % 4) % 4;
  if (r == 0) r = 4;
  for (; r <= kRungCap; r += 4) {
    const u64 A = (p + r) / 4;
    if (A == 0) continue;
    const u128 q = (u128)p * A;
    const u64 qm = (u64)(q % r);
    if (r > 1 && gcd_u64(qm, r) != 1) continue;
    const u64 target = qm ? r - qm : 0;
    Factors fa; factorize(A, fa);
    const u128 u = least_witness(p, fa, q, r, target);
    if (!u) continue;
    out = "RUNG " + std::to_string(p) + " " + std::to_string(A) + " " + std::to_string(r) +
          " " + u128_dec(u) + "\n";
    return true;
  }
  return false;
}

// ---- stage E (rung_scan3.cpp:607-647)
bool direct_line(u64 p, string &out) {
  for (u64 a = 1; a <= 8; ++a)
    for (u64 d = 1; d <= 8; ++d) {
      const u128 nn = (u128)p + (u128)4 * a * a * d;
      if (nn > kDirectNMax) continue;
      const u64 n = (u64)nn, q4 = 4 * a * d;
      Factors f; factorize(n, f);
      vector<u64> divs{1};
      for (int i = 0; i < f.k; ++i) {
        const size_t base = divs.size();
        u64 pe = 1;
        for (int e = 1; e <= f.e[i]; ++e) {
          pe *= f.p[i];
          for (size_t j = 0; j < base; ++j) divs.push_back(divs[j] * pe);
        }
      }
      std::sort(divs.begin(), divs.end());
      for (u64 m : divs) {
        if (m % q4 != q4 - 1) continue;
        const u64 c = (m + 1) / q4, k = n / m;
        if (!c || !k || (u128)c * k <= a) continue;
        out = "SOLVED " + std::to_string(p) + " " + std::to_string(a) + " " + std::to_string(c) +
              " " + std::to_string(d) + " " + std::to_string(k) + "\n";
        return true;
      }
    }
  return false;
}

// ---- lane sweep (rung_scan3.cpp:711-784)
struct Stats {
  u64 positions = 0, covered = 0, sieved = 0, mr = 0, composite = 0, rung = 0, direct = 0,
      survivors = 0;
  void add(const Stats &o) {
    positions += o.positions; covered += o.covered; sieved += o.sieved; mr += o.mr;
    composite += o.composite; rung += o.rung; direct += o.direct; survivors += o.survivors;
  }
};

// Inverse of a modulo m (gcd 1), by the extended Euclidean algorithm.
u64 inverse_mod(u64 a, u64 m) {
  if (m == 1) return 0;
  int64_t r0 = (int64_t)m, r1 = (int64_t)(a % m), t0 = 0, t1 = 1;
  while (r1) {
    const int64_t qq = r0 / r1;
    int64_t tmp = r0 - qq * r1; r0 = r1; r1 = tmp;
    tmp = t0 - qq * t1; t0 = t1; t1 = tmp;
  }
  if (r0 != 1) { fprintf(stderr, "oracle: inverse_mod(%" PRIu64 ",%" PRIu64 ")\n", a, m); abort(); }
  return (u64)(t0 < 0 ? t0 + (int64_t)m : t0);
}

struct FilterPlan { u64 g, s, inv; };   // per filter, independent of the lane

struct Options {
  u64 lo = 0, hi = 0, M = 2042040, mmax = 10000, seg = 262144, shard = 0, nshard = 1,
      progress = 0, sieve = 0;
  int threads = 0;
  const char *filters = nullptr;
  bool emit_surv = true, dump_surv = false;
};

void sweep(u64 rho, const Options &o, const vector<u64> &mods, const vector<u64> &ress,
           const vector<FilterPlan> &plan, size_t ncov, Stats &st, vector<string> &out) {
  const u64 M = o.M;
  const u64 first = o.lo + (rho + M - o.lo % M) % M;
  if (first >= o.hi) return;
  const u64 npos = (o.hi - first + M - 1) / M;
  const size_t nall = mods.size();
  vector<u64> off(nall), step(nall);
  for (size_t i = 0; i < nall; ++i) {
    const u64 m = mods[i];
    const u64 rhs = (ress[i] + m - first % m) % m;   // J with first + M J == res (mod m)
    if (rhs % plan[i].g) { step[i] = 0; continue; }
    step[i] = plan[i].s;
    off[i] = plan[i].s == 1 ? 0 : (u64)((u128)(rhs / plan[i].g) * plan[i].inv % plan[i].s);
  }
  vector<char> alive;
  for (u64 base = 0; base < npos; base += o.seg) {
    const u64 n = std::min(o.seg, npos - base);
    alive.assign(n, 1);
    auto mark = [&](size_t from, size_t to) {
      for (size_t i = from; i < to; ++i) {
        if (!step[i]) continue;
        u64 j = off[i];
        for (; j < n; j += step[i]) alive[j] = 0;
        off[i] = j - n;
      }
    };
    mark(0, ncov);
    const u64 after_cov = (u64)std::count(alive.begin(), alive.end(), 1);
    mark(ncov, nall);
    u64 left = 0;
    for (u64 j = 0; j < n; ++j) left += alive[j];
    st.positions += n;
    if (ncov == nall) st.covered += n - left;
    else { st.covered += n - after_cov; st.sieved += after_cov - left; }
    for (u64 j = 0; j < n; ++j) {
      if (!alive[j]) continue;
      const u64 p = first + M * (base + j);
      ++st.mr;
      const bool pr = sprp_base2(p) && prime_q(p);
      if (o.dump_surv) {
        out.push_back("S " + std::to_string(p) + " " + (pr ? "1" : "0") + "\n");
        if (!pr) ++st.composite;
        continue;
      }
      if (!pr) { ++st.composite; continue; }
      string line;
      if (rung_line(p, line)) { ++st.rung; out.push_back(line); continue; }
      if (direct_line(p, line)) { ++st.direct; out.push_back(line); continue; }
      ++st.survivors;
      out.push_back("SURVIVOR " + std::to_string(p) + "\n");
    }
  }
}

// ---- self-test (rung_scan3.cpp:821-897)
int self_test() {
  int fails = 0;
  auto check = [&](const char *what, u64 got, u64 want) {
    printf("  %s %-34s %" PRIu64 "\n", got == want ? "ok  " : "FAIL", what, got);
    if (got != want) ++fails;
  };
  {
    vector<Cover> slow;
    for (const Cover &f : covers_upto(120120)) if (120120 % f.m == 0) slow.push_back(f);
    vector<Cover> fast = covers_dividing(120120);
    bool same = slow.size() == fast.size();
    for (size_t i = 0; same && i < slow.size(); ++i)
      same = slow[i].m == fast[i].m && slow[i].res == fast[i].res && slow[i].pmin == fast[i].pmin &&
             slow[i].a == fast[i].a && slow[i].c == fast[i].c && slow[i].d == fast[i].d;
    check("wheel covers: divisor == full", same ? 0 : 1, 0);
  }
  const vector<Ucert> uc = uniform_certs(kUcertR, kUcertU);
  const Wheel w840 = make_wheel(840, covers_upto(840), uc);
  check("wheel(840) F1 only", w840.f1_only, 24);
  check("wheel(840) classes", w840.classes.size(), 6);
  check("wheel(840) max pmin", w840.max_pmin, 34);
  const vector<u64> hard = {1, 121, 169, 289, 361, 529};
  check("wheel(840) == hard classes", w840.classes == hard ? 0 : 1, 0);
  const Wheel w120 = make_wheel(120120, covers_upto(120120), uc);
  check("wheel(120120) F1 only", w120.f1_only, 920);
  check("wheel(120120) classes", w120.classes.size(), 220);
  check("wheel(120120) max pmin", w120.max_pmin, 15014);
  check("wheel(120120) fnv64 pinned", fnv_classes(w120.classes) == find_pin(120120)->fnv ? 0 : 1, 0);
  const Wheel w204 = make_wheel(2042040, covers_dividing(2042040), uc);
  check("wheel(2042040) classes", w204.classes.size(), 2308);
  check("wheel(2042040) fnv64 pinned", fnv_classes(w204.classes) == find_pin(2042040)->fnv ? 0 : 1, 0);
  u64 strays = 0;
  for (const Wheel *w : {&w120, &w204})
    for (u64 c : w->classes)
      if (c % 24 != 1 || std::find(hard.begin(), hard.end(), c % 840) == hard.end()) ++strays;
  check("classes outside the hard class", strays, 0);
  printf(fails ? "\nSELF-TEST FAILED (%d checks)\n" : "\nSELF-TEST PASSED\n", fails);
  return fails ? 1 : 0;
}

// ---- class table (rung_scan3.cpp:904-953)
int dump_classes(u64 M) {
  const vector<Cover> cov = covers_dividing(M);
  const vector<Ucert> uc = uniform_certs(kUcertR, kUcertU);
  const Wheel w = make_wheel(M, cov, uc);
  u64 units = 0;
  for (u64 x = 1; x < M; ++x) if (gcd_u64(x, M) == 1) ++units;
  const Pin *pin = find_pin(M);
  printf("# certified class table mod %" PRIu64 ": every unit class is killed by a named\n"
         "# certificate or survives.  F1: covers m=4acd-1 | M, kill rho \xe2\x89\xa1 -4a^2d (mod m).\n"
         "# F2: uniform rung certificates (r,u), r <= %u u <= %u, kill rho \xe2\x89\xa1 res (mod L).\n",
         M, kUcertR, kUcertU);
  printf("# unit_classes=%" PRIu64 " classes=%zu kills=%" PRIu64 " fnv64=%016" PRIx64 " sha256=%s\n",
         units, w.classes.size(), units - w.classes.size(), fnv_classes(w.classes),
         pin ? pin->sha : "unpinned");
  for (u64 c : w.classes) printf("CLASS %" PRIu64 "\n", c);
  // First killer in canonical order: F1 ascending (m,res), then F2 ascending (L,res).
  vector<int64_t> who(M, -1);
  for (size_t i = 0; i < cov.size(); ++i)
    for (u64 x = cov[i].res; x < M; x += cov[i].m) if (who[x] < 0) who[x] = (int64_t)i;
  const int64_t f2 = (int64_t)cov.size();
  for (size_t i = 0; i < uc.size(); ++i) {
    if (M % uc[i].L) continue;
    for (u64 x = uc[i].res; x < M; x += uc[i].L) if (who[x] < 0) who[x] = f2 + (int64_t)i;
  }
  for (u64 x = 1; x < M; ++x) {
    if (gcd_u64(x, M) != 1 || who[x] < 0) continue;
    if (who[x] < f2) {
      const Cover &f = cov[who[x]];
      printf("KILL %" PRIu64 " F1 %" PRIu64 " %" PRIu64 " %u %u %u\n", x, f.m, f.res, f.a, f.c, f.d);
    } else {
      const Ucert &c = uc[who[x] - f2];
      printf("KILL %" PRIu64 " F2 %" PRIu64 " %" PRIu64 " %u %u\n", x, c.L, c.res, c.r, c.u);
    }
  }
  return 0;
}

u64 small_primes_append(u64 P, u64 M, vector<u64> &mods, vector<u64> &ress) {
  vector<char> comp(P + 1, 0);
  u64 n = 0;
  for (u64 i = 2; i <= P; ++i) {
    if (comp[i]) continue;
    if (M % i) { mods.push_back(i); ress.push_back(0); ++n; }
    for (u64 j = i * i; j <= P; j += i) comp[j] = 1;
  }
  return n;
}

// One record per interval. Without a batch: the interval in o (after --shard), exactly as
// rung_scan3 refuses and scans it. With a batch (--batch FILE): every "lo hi" line of the
// file in turn, with the tables built once; a refused interval's record is its FATAL line.
int scan(Options o, const vector<std::pair<u64, u64>> *batch) {
  if (!o.seg) o.seg = 262144;
  if (o.nshard == 0) o.nshard = 1;
  if (o.nshard > 1) {                                   // rung_scan3.cpp:1022-1027
    const u64 span = (o.hi - o.lo + o.nshard - 1) / o.nshard;
    const u64 a = o.lo + o.shard * span, b = std::min(o.hi, a + span);
    o.lo = a; o.hi = std::max(a, b);
  }
  const vector<Cover> covers = covers_upto(o.mmax);
  const Wheel w = make_wheel(o.M, covers_dividing(o.M), uniform_certs(kUcertR, kUcertU));
  u64 cover_pmax = 0;
  for (const Cover &f : covers) cover_pmax = std::max(cover_pmax, f.pmin);
  const Pin *pin = find_pin(o.M);
  if (pin) {
    if (fnv_classes(w.classes) != pin->fnv) {
      fprintf(stderr, "FATAL: class table digest mismatch at M=%" PRIu64 "\n", o.M);
      return 1;
    }
  } else {
    fprintf(stderr, "[es_oracle] WARNING: no pinned digest for M=%" PRIu64 " -- running UNPINNED\n", o.M);
  }
  const u64 floor_ = std::max(w.max_pmin, cover_pmax);
  // rung_scan3.cpp:1057-1069; in a batch the message goes to stdout as the record.
  auto refused = [&](u64 lo, u64 hi, FILE *to) {
    if (hi > kDirectPMax) {
      fprintf(to, "FATAL: HI must not exceed %" PRIu64 ", the stage-E ceiling\n", kDirectPMax);
      return true;
    }
    if (lo <= floor_) {
      fprintf(to, "FATAL: LO must exceed max pmin of the filter table (%" PRIu64 ")\n", floor_);
      return true;
    }
    return false;
  };
  if (!batch && refused(o.lo, o.hi, stderr)) return 1;
  vector<u64> mods, ress;
  if (o.filters) {
    FILE *fp = fopen(o.filters, "r");
    if (!fp) { fprintf(stderr, "FATAL: cannot open %s\n", o.filters); return 1; }
    char line[256];
    while (fgets(line, sizeof line, fp)) {
      unsigned long long m, r;
      if (sscanf(line, "FILTER %llu %llu", &m, &r) == 2) { mods.push_back(m); ress.push_back(r); }
    }
    fclose(fp);
    if (mods.empty()) { fprintf(stderr, "FATAL: no FILTER lines in %s\n", o.filters); return 1; }
  } else {
    for (const Cover &f : covers) if (o.M % f.m) { mods.push_back(f.m); ress.push_back(f.res); }
  }
  const size_t ncov = mods.size();
  if (o.sieve) small_primes_append(o.sieve, o.M, mods, ress);
  vector<FilterPlan> plan(mods.size());
  for (size_t i = 0; i < mods.size(); ++i) {
    const u64 g = gcd_u64(o.M, mods[i]), s = mods[i] / g;
    plan[i] = FilterPlan{g, s, s == 1 ? 0 : inverse_mod((o.M / g) % s, s)};
  }
  fprintf(stderr, "[es_oracle] wheel M=%" PRIu64 " lanes=%zu filters=%zu (+%zu primes)\n", o.M,
          w.classes.size(), ncov, mods.size() - ncov);
  init_trial();                                         // before the threads start
#ifdef _OPENMP
  if (o.threads > 0) omp_set_num_threads(o.threads);
#endif
  const size_t L = w.classes.size();
  auto one = [&](const Options &oi) -> int {
    vector<Stats> lst(L);
    vector<vector<string>> lout(L);
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (size_t i = 0; i < L; ++i) sweep(w.classes[i], oi, mods, ress, plan, ncov, lst[i], lout[i]);
    Stats st;
    for (size_t i = 0; i < L; ++i) {
      st.add(lst[i]);
      for (const string &s : lout[i])
        if (oi.emit_surv || s.compare(0, 9, "SURVIVOR ") != 0) fputs(s.c_str(), stdout);
    }
    if (oi.dump_surv) return 0;
    printf("SUMMARY lo=%" PRIu64 " hi=%" PRIu64 " wheelM=%" PRIu64 " lanes=%zu mmax=%" PRIu64
           " ucert_rmax=%u ucert_umax=%u wheel_pmin=%" PRIu64 " classes_sha256=%s positions=%" PRIu64
           " covered=%" PRIu64 " sieved=%" PRIu64 " mr=%" PRIu64 " composite=%" PRIu64 " rung=%" PRIu64
           " direct=%" PRIu64 " survivors=%" PRIu64 "\n",
           oi.lo, oi.hi, oi.M, L, oi.mmax, kUcertR, kUcertU, w.max_pmin, pin ? pin->sha : "unpinned",
           st.positions, st.covered, st.sieved, st.mr, st.composite, st.rung, st.direct, st.survivors);
    return st.survivors ? 2 : 0;
  };
  if (!batch) return one(o);
  for (const auto &iv : *batch) {
    Options oi = o;
    oi.lo = iv.first;
    oi.hi = iv.second;
    if (!refused(oi.lo, oi.hi, stdout)) one(oi);
  }
  return 0;
}

}  // namespace

// ---- C entry points for tests (ctypes). Checker use only.
extern "C" {
int es_oracle_is_prime(u64 n) { return prime_q(n) ? 1 : 0; }
int es_oracle_sprp2(u64 n) { return sprp_base2(n) ? 1 : 0; }
// Writes the stage D line (or stage E, or "SURVIVOR p") for a prime p into buf.
int es_oracle_solve(u64 p, char *buf, int cap) {
  init_trial();
  string line;
  int kind = 0;
  if (rung_line(p, line)) kind = 1;
  else if (direct_line(p, line)) kind = 2;
  else line = "SURVIVOR " + std::to_string(p) + "\n";
  snprintf(buf, (size_t)cap, "%s", line.c_str());
  return kind;
}
// Canonical factorization: primes ascending with exponents; returns the count.
int es_oracle_factor(u64 n, u64 *pr, int *ex) {
  Factors f; factorize(n, f);
  for (int i = 1; i < f.k; ++i)
    for (int j = i; j > 0 && f.p[j - 1] > f.p[j]; --j) { std::swap(f.p[j - 1], f.p[j]); std::swap(f.e[j - 1], f.e[j]); }
  for (int i = 0; i < f.k; ++i) { pr[i] = f.p[i]; ex[i] = f.e[i]; }
  return f.k;
}
}

#ifndef ES_ORACLE_NO_MAIN
int main(int argc, char **argv) {
  if (argc >= 2 && !strcmp(argv[1], "--verify")) return self_test();
  if (argc >= 2 && !strcmp(argv[1], "--dump-classes")) {
    u64 M = 2042040;
    for (int i = 2; i + 1 < argc; ++i) if (!strcmp(argv[i], "--wheel")) M = strtoull(argv[i + 1], 0, 10);
    return dump_classes(M);
  }
  if (argc >= 2 && !strcmp(argv[1], "--solve-direct")) {   // stage E alone, as --check-direct
    init_trial();
    for (int i = 2; i < argc; ++i) {
      string line;
      if (direct_line(strtoull(argv[i], 0, 10), line)) fputs(line.c_str(), stdout);
      else printf("DECLINED %s\n", argv[i]);
    }
    return 0;
  }
  if (argc >= 2 && !strcmp(argv[1], "--solve")) {
    for (int i = 2; i < argc; ++i) {
      char buf[256];
      es_oracle_solve(strtoull(argv[i], 0, 10), buf, sizeof buf);
      fputs(buf, stdout);
    }
    return 0;
  }
  vector<std::pair<u64, u64>> batch;
  const bool batch_mode = argc >= 3 && !strcmp(argv[1], "--batch");
  if (batch_mode) {
    FILE *fp = fopen(argv[2], "r");
    if (!fp) { fprintf(stderr, "FATAL: cannot open %s\n", argv[2]); return 1; }
    unsigned long long a, b;
    char line[256];
    while (fgets(line, sizeof line, fp))
      if (sscanf(line, "%llu %llu", &a, &b) == 2) batch.emplace_back(a, b);
    fclose(fp);
  }
  if (argc < 3) {
    fprintf(stderr, "usage: %s LO HI [--mmax N] [--wheel M] [--seg N] [--threads N] [--shard i/n]\n"
                    "          [--progress N] [--filters FILE] [--sieve P] [--no-emit-survivors]\n"
                    "          [--dump-survivors]\n"
                    "       %s --batch FILE [the same options]\n"
                    "       %s --verify | --dump-classes [--wheel M] | --solve P...\n", argv[0], argv[0], argv[0]);
    return 1;
  }
  Options o;
  if (!batch_mode) {
    o.lo = strtoull(argv[1], 0, 10);
    o.hi = strtoull(argv[2], 0, 10);
  }
  for (int i = 3; i < argc; ++i) {
    const char *a = argv[i];
    const bool has = i + 1 < argc;
    if (!strcmp(a, "--mmax") && has) o.mmax = strtoull(argv[++i], 0, 10);
    else if (!strcmp(a, "--wheel") && has) o.M = strtoull(argv[++i], 0, 10);
    else if (!strcmp(a, "--seg") && has) o.seg = strtoull(argv[++i], 0, 10);
    else if (!strcmp(a, "--progress") && has) o.progress = strtoull(argv[++i], 0, 10);
    else if (!strcmp(a, "--threads") && has) o.threads = atoi(argv[++i]);
    else if (!strcmp(a, "--filters") && has) o.filters = argv[++i];
    else if (!strcmp(a, "--sieve") && has) o.sieve = strtoull(argv[++i], 0, 10);
    else if (!strcmp(a, "--shard") && has) {
      unsigned long long x = 0, y = 1;
      if (sscanf(argv[++i], "%llu/%llu", &x, &y) == 2) { o.shard = x; o.nshard = y; }
    } else if (!strcmp(a, "--no-emit-survivors")) o.emit_surv = false;
    else if (!strcmp(a, "--dump-survivors")) o.dump_surv = true;
    else { fprintf(stderr, "FATAL: unknown option '%s'\n", a); return 1; }
  }
  return scan(o, batch_mode ? &batch : nullptr);
}
#endif
