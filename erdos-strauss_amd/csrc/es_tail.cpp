// es_tail.cpp -- stage D/E on the host. Montgomery multiplication modulo odd n < 2^64
// (REDC with the explicit high-half carry, so n >= 2^63 is exact), magic-inverse trial
// division, Pollard's rho with Brent's cycle finding, and the per-residue minimum over
// divisors of q^2.
#include "es_tail.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace es {

// Decimal digits of v written backwards ending at `end`, two at a time from a table
// (64-bit division by a constant compiles to a multiply); returns the first digit.
char *put_u64(char *end, u64 v) {
  static const char kPairs[] =
      "00010203040506070809101112131415161718192021222324252627282930313233343536373839"
      "40414243444546474849505152535455565758596061626364656667686970717273747576777879"
      "8081828384858687888990919293949596979899";
  while (v >= 100) {
    const u64 q = v / 100;
    const unsigned r = (unsigned)(v - q * 100);
    end -= 2;
    memcpy(end, kPairs + 2 * r, 2);
    v = q;
  }
  if (v >= 10) {
    end -= 2;
    memcpy(end, kPairs + 2 * v, 2);
  } else {
    *--end = (char)('0' + v);
  }
  return end;
}

namespace {

constexpr u64 kTrialBound = 65536;   // census/rung_scan.cpp:175
constexpr u64 kDirectNMax = 18446744065119617025ull;
constexpr unsigned kRungCap = 255;

u64 gcd64(u64 a, u64 b) {
  while (b) { const u64 t = a % b; a = b; b = t; }
  return a;
}

// Montgomery form modulo an odd n with R = 2^64.
struct Mont {
  u64 n, nneg_inv, r1, r2;
  static u64 inv2_64(u64 a) {            // a^{-1} mod 2^64 for odd a (Newton, 6 steps)
    u64 x = a;
    for (int i = 0; i < 6; ++i) x *= 2 - a * x;
    return x;
  }
  explicit Mont(u64 n_) : n(n_) {
    nneg_inv = 0 - inv2_64(n_);
    r1 = (u64)(((u128)1 << 64) % n_);
    r2 = (u64)((u128)r1 * r1 % n_);
  }
  // t < n * 2^64: returns t / 2^64 mod n. The low half of t + m n is zero, so the sum
  // of the high halves plus the carry out of the low half is the result; it can reach
  // 2n - 1 >= 2^64 when n >= 2^63, hence the u128 accumulation.
  u64 redc(u128 t) const {
    const u64 m = (u64)t * nneg_inv;
    const u128 s = (t >> 64) + (((u128)m * n) >> 64) + ((u64)t != 0);
    return (u64)(s >= n ? s - n : s);
  }
  u64 mul(u64 a, u64 b) const { return redc((u128)a * b); }
  u64 to(u64 a) const { return mul(a % n, r2); }
  u64 add(u64 a, u64 b) const {
    const u64 s = a + b;
    return (s < a || s >= n) ? s - n : s;
  }
};

const u64 kBases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};

// n odd > 37; d odd with n-1 = d 2^s; a in Montgomery form.
bool strong_probable(const Mont &M, u64 a_m, u64 d, int s) {
  const u64 one = M.r1, minus_one = M.n - M.r1;
  u64 x = one, b = a_m;
  for (u64 e = d; e; e >>= 1) {
    if (e & 1) x = M.mul(x, b);
    b = M.mul(b, b);
  }
  if (x == one || x == minus_one) return true;
  for (int i = 1; i < s; ++i) {
    x = M.mul(x, x);
    if (x == minus_one) return true;
  }
  return false;
}

// Trial-division tables: odd primes 3..65536 with q^{-1} mod 2^64 and floor((2^64-1)/q):
// q | n  <=>  n * q^{-1} mod 2^64 <= floor((2^64-1)/q).
struct TrialTables {
  std::vector<u64> q, inv, lim, q2;
  TrialTables() {
    std::vector<uint8_t> comp(kTrialBound + 1, 0);
    for (u64 i = 3; i <= kTrialBound; i += 2) {
      if (comp[i]) continue;
      q.push_back(i);
      inv.push_back(Mont::inv2_64(i));
      lim.push_back(~0ull / i);
      q2.push_back(i * i);
      for (u64 j = i * i; j <= kTrialBound; j += 2 * i) comp[j] = 1;
    }
  }
};
const TrialTables &trial() {
  static const TrialTables t;
  return t;
}

struct Fac {
  u64 p[16];
  int e[16];
  int k = 0;
  bool overflow = false;
  void add(u64 q, int x) {
    for (int i = 0; i < k; ++i)
      if (p[i] == q) { e[i] += x; return; }
    if (k >= 16) { overflow = true; return; }
    p[k] = q; e[k] = x; ++k;
  }
};

// floor(sqrt(n)) by Newton's iteration from above: x0 = 2^ceil(bits/2) exceeds sqrt(n),
// and x -> (x + n/x) / 2 decreases strictly until it reaches floor(sqrt(n)).
u64 isqrt64(u64 n) {
  if (n < 2) return n;
  u64 x = 1ull << ((65 - __builtin_clzll(n)) / 2);
  for (;;) {
    const u64 y = (x + n / x) / 2;
    if (y >= x) return x;
    x = y;
  }
}

// One rho walk x -> x^2 + c (Montgomery form) from x0, with Brent's cycle finding (R. P.
// Brent, "An improved Monte Carlo factorization algorithm", BIT 20 (1980) 176-184): the
// saved point moves to the walk's position at every power-of-two length, and the walk's
// differences from it are multiplied together kRhoBatch at a time, one gcd per batch.
// A batch whose gcd is n is retraced one step and one gcd at a time. Returns a proper
// factor, or 0 when the walk closed its cycle modulo n itself.
constexpr u64 kRhoBatch = 96;

// gcd(a, n) for odd n by Stein's binary method (shifts and subtractions, no division).
u64 gcd_odd(u64 a, u64 n) {
  if (!a) return n;
  a >>= __builtin_ctzll(a);
  while (a != n) {
    if (a > n) std::swap(a, n);
    n -= a;
    n >>= __builtin_ctzll(n);
  }
  return a;
}

u64 rho_walk(const Mont &M, u64 cm, u64 x0) {
  const u64 n = M.n;
  auto next = [&](u64 v) { return M.add(M.mul(v, v), cm); };
  auto dist = [](u64 u, u64 v) { return u > v ? u - v : v - u; };
  u64 x = x0;
  for (u64 len = 1;; len <<= 1) {
    const u64 saved = x;
    for (u64 walked = 0; walked < len;) {
      const u64 run = std::min(kRhoBatch, len - walked), from = x;
      u64 prod = M.r1;
      for (u64 i = 0; i < run; ++i) {
        x = next(x);
        prod = M.mul(prod, dist(saved, x));
      }
      u64 g = gcd_odd(prod, n);
      if (g == n) {
        g = 1;
        for (u64 y = from, i = 0; i < run && g == 1; ++i) {
          y = next(y);
          g = gcd_odd(dist(saved, y), n);
        }
      }
      if (g == n) return 0;
      if (g != 1) return g;
      walked += run;
    }
  }
}

// A proper factor of odd composite n that is not a perfect square: walks with c = 1, 3,
// 5, ... from pseudo-random starting points until one splits n (each walk ends, since its
// sequence modulo n is eventually periodic).
u64 rho_factor(u64 n) {
  const Mont M(n);
  u64 state = n ^ 0x243F6A8885A308D3ull;
  for (u64 c = 1;; c += 2) {
    state = state * 6364136223846793005ull + 1442695040888963407ull;   // Knuth's MMIX LCG
    const u64 g = rho_walk(M, M.to(c), M.to(state >> 1));
    if (g) return g;
  }
}

void split(u64 n, Fac &f) {
  if (n == 1) return;
  if (is_prime_u64(n)) { f.add(n, 1); return; }
  const u64 s = isqrt64(n);
  if (s * s == n) {
    Fac g;
    split(s, g);
    for (int i = 0; i < g.k; ++i) f.add(g.p[i], 2 * g.e[i]);
    f.overflow |= g.overflow;
    return;
  }
  const u64 d = rho_factor(n);
  split(d, f);
  split(n / d, f);
}

void factor(u64 n, Fac &f) {
  f.k = 0;
  f.overflow = false;
  if (n == 0) return;
  if (!(n & 1)) {
    const int e = __builtin_ctzll(n);
    n >>= e;
    f.add(2, e);
  }
  const TrialTables &t = trial();
  for (size_t i = 0; i < t.q.size(); ++i) {
    if (t.q2[i] > n) break;
    if (n * t.inv[i] > t.lim[i]) continue;
    int e = 0;
    do { n /= t.q[i]; ++e; } while (n * t.inv[i] <= t.lim[i]);
    f.add(t.q[i], e);
  }
  if (n > 1) {
    if (n <= kTrialBound * kTrialBound) f.add(n, 1);   // all prime factors exceed 65536
    else split(n, f);
  }
}

// Least u | q^2 with u == target (mod r) and u <= q (0 if none). best[s] is the least
// divisor built so far in residue class s; multiplying by a prime power preserves the
// order within a class, so the per-class minimum is all that needs carrying, and
// nothing above q needs to exist (rung_scan3.cpp:534-565).
u128 least_witness(u64 p, const Fac &fa, u128 q, unsigned r, unsigned target) {
  const u128 NONE = ~(u128)0;
  u128 best[kRungCap + 1], next[kRungCap + 1];
  for (unsigned s = 0; s < r; ++s) best[s] = NONE;
  best[1 % r] = 1;
  auto absorb = [&](u64 pr, int e) {
    for (unsigned s = 0; s < r; ++s) next[s] = NONE;
    const unsigned pm = (unsigned)(pr % r);
    const u128 cap = q / pr;
    for (unsigned s = 0; s < r; ++s) {
      if (best[s] == NONE) continue;
      u128 v = best[s];
      unsigned res = s;
      for (int j = 0;; ++j) {
        if (v < next[res]) next[res] = v;
        if (j == e || v > cap) break;
        v *= pr;
        res = (res * pm) % r;
      }
    }
    for (unsigned s = 0; s < r; ++s) best[s] = next[s];
  };
  absorb(p, 2);
  for (int i = 0; i < fa.k; ++i) absorb(fa.p[i], 2 * fa.e[i]);
  return best[target] == NONE ? 0 : best[target];
}

// As put_u64 for 128-bit v: one 128-bit division splits off the low 19 digits.
char *put_u128(char *end, u128 v) {
  if (!(v >> 64)) return put_u64(end, (u64)v);
  const u64 kTen19 = 10000000000000000000ull;
  const u64 lo = (u64)(v % kTen19);
  char *start = put_u64(end, lo);
  while (start > end - 19) *--start = '0';          // zero-pad the low part to 19 digits
  return put_u128(start, v / kTen19);
}

void append_u128(std::string &s, u128 v) {
  char buf[48];
  char *b = put_u128(buf + sizeof buf, v);
  s.append(b, buf + sizeof buf - b);
}

}  // namespace

bool is_prime_u64(u64 n) {
  if (n < 2) return false;
  for (u64 q : kBases) {
    if (n == q) return true;
    if (n % q == 0) return false;
  }
  const Mont M(n);
  u64 d = n - 1;
  const int s = __builtin_ctzll(d);
  d >>= s;
  for (u64 a : kBases)
    if (!strong_probable(M, M.to(a), d, s)) return false;
  return true;
}

bool sprp2_u64(u64 n) {
  if (n < 2) return false;
  if (!(n & 1)) return n == 2;
  if (n == 3) return true;
  const Mont M(n);
  u64 d = n - 1;
  const int s = __builtin_ctzll(d);
  d >>= s;
  return strong_probable(M, M.to(2), d, s);
}

int factor_u64(u64 n, u64 *primes, int *exps) {
  Fac f;
  factor(n, f);
  if (f.overflow) return -1;
  for (int i = 1; i < f.k; ++i)
    for (int j = i; j > 0 && f.p[j - 1] > f.p[j]; --j) {
      std::swap(f.p[j - 1], f.p[j]);
      std::swap(f.e[j - 1], f.e[j]);
    }
  for (int i = 0; i < f.k; ++i) { primes[i] = f.p[i]; exps[i] = f.e[i]; }
  return f.k;
}

// rung_scan3.cpp:574-603: rungs r == -p (mod 4) from the smallest, skipping those with
// gcd(pA mod r, r) > 1; the first with a witness wins.
bool solve_rung(u64 p, std::string &line) {
  unsigned r = (unsigned)((4 - p % 4) % 4);
  if (r == 0) r = 4;
  for (; r <= kRungCap; r += 4) {
    const u64 A = (p + r) / 4;
    if (!A) continue;
    const u128 q = (u128)p * A;
    const unsigned qm = (unsigned)(q % r);
    if (r > 1 && gcd64(qm, r) != 1) continue;
    const unsigned target = qm ? r - qm : 0;
    Fac fa;
    factor(A, fa);
    if (fa.overflow) {
      fprintf(stderr, "FATAL: more than 16 distinct prime factors\n");
      abort();
    }
    const u128 u = least_witness(p, fa, q, r, target);
    if (!u) continue;
    format_rung(p, A, r, u, line);
    return true;
  }
  return false;
}

// Stage E (rung_scan3.cpp:607-647): the pairs (a, d) in 1..8 x 1..8, a major. For
// n = p + 4a^2 d (at most the ceiling kDirectNMax), the certificate is the least divisor
// m of n with m == -1 (mod 4ad) such that c = (m + 1) / (4ad) and k = n / m have
// c k > a, printed as "SOLVED p a c d k". The first pair with such a divisor wins.
bool solve_direct(u64 p, std::string &line) {
  for (unsigned pair = 0; pair < 64; ++pair) {
    const u64 a = pair / 8 + 1, d = pair % 8 + 1, q4 = 4 * a * d;
    const u128 wide = (u128)p + (u128)(4 * a * a * d);
    if (wide > kDirectNMax) continue;
    const u64 n = (u64)wide;
    Fac f;
    factor(n, f);
    if (f.overflow) {
      fprintf(stderr, "FATAL: more than 16 distinct prime factors\n");
      abort();
    }
    // every divisor, each prime power times those built from the earlier primes
    std::vector<u64> divs(1, 1);
    for (int i = 0; i < f.k; ++i) {
      const size_t have = divs.size();
      u64 pw = 1;
      for (int e = 0; e < f.e[i]; ++e) {
        pw *= f.p[i];
        for (size_t j = 0; j < have; ++j) divs.push_back(divs[j] * pw);
      }
    }
    std::sort(divs.begin(), divs.end());
    for (const u64 m : divs) {
      if ((m + 1) % q4) continue;                   // m == -1 (mod 4ad), so c >= 1
      const u64 c = (m + 1) / q4, k = n / m;        // m | n and m <= n, so k >= 1
      if ((u128)c * k > a) {
        format_direct(p, a, c, d, k, line);
        return true;
      }
    }
  }
  return false;
}

void format_direct(u64 p, u64 a, u64 c, u64 d, u64 k, std::string &line) {
  char buf[160];
  snprintf(buf, sizeof buf, "SOLVED %llu %llu %llu %llu %llu\n", (unsigned long long)p, (unsigned long long)a,
           (unsigned long long)c, (unsigned long long)d, (unsigned long long)k);
  line = buf;
}

void format_rung(u64 p, u64 A, unsigned r, u128 u, std::string &line) {
  char buf[128];
  char *end = buf + sizeof buf, *b = end;
  *--b = '\n';
  b = put_u128(b, u);
  *--b = ' ';
  b = put_u64(b, r);
  *--b = ' ';
  b = put_u64(b, A);
  *--b = ' ';
  b = put_u64(b, p);
  b -= 5;
  memcpy(b, "RUNG ", 5);
  line.assign(b, end - b);
}

int certify_prime(u64 p, std::string &line) {
  if (solve_rung(p, line)) return 0;
  if (solve_direct(p, line)) return 1;
  line = "SURVIVOR " + std::to_string(p) + "\n";
  return 2;
}

}  // namespace es
