// es_sieve_main.cpp -- `es_sieve`, the drop-in for rung_scan3 (and sieve_cuda) under a
// shard driver: the same argv, the same stdout (certificate lines then one SUMMARY),
// the same exit codes (0 clean, 2 survivors, 1 refusal/config error). The funnel and
// stage D run on the GPU; stage E, and any prime the device declines, on the host.
//
//   es_sieve LO HI [--mmax N] [--wheel M] [--seg N] [--threads N] [--shard i/n]
//                  [--progress N] [--filters FILE] [--sieve P] [--no-emit-survivors]
//            GPU options (additive, output-invariant):
//                  [--device D] [--block B] [--tile-bits N] [--layout Mp|auto]
//                  [--plain] [--chunk N] [--order lane|round] [--profile] [--quiet-time]
//                  [--host-tail] [--graph]
//   es_sieve --verify | --dump-classes [--wheel M] | --check-direct LO N
//            | --check-factor [N] | --check-primality LO N [--device D]
//
// rung_scan3.cpp:960-1163 (main), sieve/cuda/sieve_cuda.cu:177-274 (GPU flags).
#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "es_plan.h"
#include "es_scan.h"
#include "es_tail.h"

using namespace es;

namespace {

double wall_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

u64 gcd64(u64 a, u64 b) {
  while (b) { const u64 t = a % b; a = b; b = t; }
  return a;
}

// ---------------------------------------------------------------- --verify
// The wheel KATs of rung_scan3.cpp:821-897 on this build's own generators.
int self_test() {
  int fails = 0;
  auto check = [&](const char *what, u64 got, u64 want) {
    printf("  %s %-34s %" PRIu64 "\n", got == want ? "ok  " : "FAIL", what, got);
    if (got != want) ++fails;
  };
  {
    std::vector<Cover> slow;
    for (const Cover &f : covers_upto(120120)) if (120120 % f.m == 0) slow.push_back(f);
    const std::vector<Cover> fast = covers_dividing(120120);
    bool same = slow.size() == fast.size();
    for (size_t i = 0; same && i < slow.size(); ++i)
      same = slow[i].m == fast[i].m && slow[i].res == fast[i].res && slow[i].pmin == fast[i].pmin &&
             slow[i].a == fast[i].a && slow[i].c == fast[i].c && slow[i].d == fast[i].d;
    check("wheel covers: divisor == full", same ? 0 : 1, 0);
  }
  const std::vector<Ucert> uc = uniform_certs(kUcertRmax, kUcertUmax);
  const Wheel w840 = build_wheel(840, covers_upto(840), uc);
  check("wheel(840) F1 only", w840.f1_only, 24);
  check("wheel(840) classes", w840.classes.size(), 6);
  check("wheel(840) max pmin", w840.max_pmin, 34);
  const std::vector<u64> hard = {1, 121, 169, 289, 361, 529};
  check("wheel(840) == hard classes", w840.classes == hard ? 0 : 1, 0);
  const Wheel w120 = build_wheel(120120, covers_upto(120120), uc);
  check("wheel(120120) F1 only", w120.f1_only, 920);
  check("wheel(120120) classes", w120.classes.size(), 220);
  check("wheel(120120) max pmin", w120.max_pmin, 15014);
  check("wheel(120120) fnv64 pinned", fnv64_classes(w120.classes) == class_pin(120120)->fnv64 ? 0 : 1, 0);
  const Wheel w204 = build_wheel(2042040, covers_dividing(2042040), uc);
  check("wheel(2042040) classes", w204.classes.size(), 2308);
  check("wheel(2042040) fnv64 pinned", fnv64_classes(w204.classes) == class_pin(2042040)->fnv64 ? 0 : 1, 0);
  u64 strays = 0;
  for (const Wheel *w : {&w120, &w204})
    for (u64 c : w->classes)
      if (c % 24 != 1 || std::find(hard.begin(), hard.end(), c % 840) == hard.end()) ++strays;
  check("classes outside the hard class", strays, 0);
  printf(fails ? "\nSELF-TEST FAILED (%d checks)\n" : "\nSELF-TEST PASSED\n", fails);
  return fails ? 1 : 0;
}

// ---------------------------------------------------------------- --dump-classes
// rung_scan3.cpp:904-953: surviving classes, then the first killer of every other unit
// class (F1 ascending (m,res), then F2 ascending (L,res)).
int dump_classes(u64 M) {
  const std::vector<Cover> cov = covers_dividing(M);
  const std::vector<Ucert> uc = uniform_certs(kUcertRmax, kUcertUmax);
  const Wheel w = build_wheel(M, cov, uc);
  u64 units = 0;
  for (u64 x = 1; x < M; ++x) units += gcd64(x, M) == 1;
  const ClassPin *pin = class_pin(M);
  printf("# certified class table mod %" PRIu64 ": every unit class is killed by a named\n"
         "# certificate or survives.  F1: covers m=4acd-1 | M, kill rho \xe2\x89\xa1 -4a^2d (mod m).\n"
         "# F2: uniform rung certificates (r,u), r <= %u u <= %u, kill rho \xe2\x89\xa1 res (mod L).\n",
         M, kUcertRmax, kUcertUmax);
  printf("# unit_classes=%" PRIu64 " classes=%zu kills=%" PRIu64 " fnv64=%016" PRIx64 " sha256=%s\n", units,
         w.classes.size(), units - w.classes.size(), fnv64_classes(w.classes), pin ? pin->sha256 : "unpinned");
  for (u64 c : w.classes) printf("CLASS %" PRIu64 "\n", c);
  std::vector<int32_t> who(M, -1);
  for (size_t i = 0; i < cov.size(); ++i)
    for (u64 x = cov[i].res; x < M; x += cov[i].m) if (who[x] < 0) who[x] = (int32_t)i;
  const int32_t f2 = (int32_t)cov.size();
  for (size_t i = 0; i < uc.size(); ++i) {
    if (M % uc[i].L) continue;
    for (u64 x = uc[i].res; x < M; x += uc[i].L) if (who[x] < 0) who[x] = f2 + (int32_t)i;
  }
  for (u64 x = 1; x < M; ++x) {
    if (who[x] < 0 || gcd64(x, M) != 1) continue;
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

// ---------------------------------------------------------------- --check-direct
// rung_scan3.cpp:798-817: stage E alone on the first N primes >= LO; SOLVED lines to
// stdout for an exact-arithmetic verifier, status to stderr.
int check_direct(u64 lo, u64 N) {
  u64 found = 0, solved = 0;
  for (u64 p = lo | 1ull; found < N; p += 2) {
    if (!is_prime_u64(p)) continue;
    ++found;
    std::string line;
    if (solve_direct(p, line)) { ++solved; fputs(line.c_str(), stdout); }
    else fprintf(stderr, "  stage E DECLINED p=%" PRIu64 "\n", p);
  }
  fprintf(stderr, "%s check-direct %" PRIu64 "/%" PRIu64 " primes at or above %" PRIu64 " solved by stage E\n",
          solved == N ? "ok  " : "FAIL", solved, N, lo);
  return solved == N ? 0 : 1;
}

// ---------------------------------------------------------------- --check-factor
// rung_scan3.cpp:459-530: the Montgomery factorization against plain trial division
// plus u128 rho, on adversarial shapes; products and factor primality checked too.
u64 plain_rho(u64 n) {
  for (u64 c = 1;; ++c) {
    u64 x = 2, y = 2, g = 1;
    auto f = [&](u64 v) { return (u64)(((u128)v * v + c) % n); };
    while (g == 1) {
      x = f(x);
      y = f(f(y));
      g = gcd64(x > y ? x - y : y - x, n);
    }
    if (g != n) return g;
  }
}
bool plain_prime(u64 n) {
  if (n < 2) return false;
  for (u64 q = 2; q < 1000 && q * q <= n; ++q) if (n % q == 0) return false;
  if (n < 1000000) return true;
  u64 d = n - 1; int s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  for (u64 a : {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37}) {
    u64 x = 1, b = a % n;
    for (u64 e = d; e; e >>= 1) { if (e & 1) x = (u64)((u128)x * b % n); b = (u64)((u128)b * b % n); }
    if (x == 1 || x == n - 1) continue;
    bool comp = true;
    for (int i = 1; i < s && comp; ++i) { x = (u64)((u128)x * x % n); comp = x != n - 1; }
    if (comp) return false;
  }
  return true;
}
void plain_factor(u64 n, std::vector<std::pair<u64, int>> &out) {
  std::vector<u64> stack;
  for (u64 q = 2; q <= 65536 && q * q <= n; ++q)
    while (n % q == 0) { out.push_back({q, 1}); n /= q; }
  if (n > 1) stack.push_back(n);
  while (!stack.empty()) {
    const u64 v = stack.back();
    stack.pop_back();
    if (plain_prime(v)) { out.push_back({v, 1}); continue; }
    const u64 d = plain_rho(v);
    stack.push_back(d);
    stack.push_back(v / d);
  }
  std::sort(out.begin(), out.end());
  std::vector<std::pair<u64, int>> merged;
  for (auto &pe : out) {
    if (!merged.empty() && merged.back().first == pe.first) merged.back().second += pe.second;
    else merged.push_back(pe);
  }
  out.swap(merged);
}
int check_factor(u64 nper) {
  std::vector<u64> vals;
  u64 seed = 0x9e3779b97f4a7c15ull;
  auto rnd = [&]() { seed ^= seed << 13; seed ^= seed >> 7; seed ^= seed << 17; return seed; };
  for (u64 dec = 1000000ull; dec < kDirectNMax / 10; dec *= 10)
    for (u64 i = 0; i < nper; ++i) vals.push_back(dec + rnd() % (9 * dec));
  for (u64 i = 0; i < nper; ++i) {
    u64 a = 3037000000ull + rnd() % 1257000000ull, b = 2147483648ull + rnd() % 2147000000ull;
    while (!is_prime_u64(a)) ++a;
    while (!is_prime_u64(b)) ++b;
    if ((u128)a * b <= (u128)kDirectNMax) vals.push_back(a * b);
    u64 c = 2147483648ull + rnd() % 2147000000ull;
    while (!is_prime_u64(c)) ++c;
    vals.push_back(c * c);
    u64 v = 1;
    while (v < kDirectNMax / 65536) v *= 2 + rnd() % 65535;
    vals.push_back(v);
  }
  const u64 edge[] = {65535ull * 65535ull, 65536ull * 65536ull - 1, 65536ull * 65536ull,
                      65536ull * 65536ull + 1, (1ull << 63) - 1, 1ull << 63, (1ull << 63) + 1,
                      kDirectNMax - 2, kDirectNMax - 1, kDirectNMax, 1ull, 2ull};
  for (u64 v : edge) vals.push_back(v);
  u64 bad = 0;
  for (u64 n : vals) {
    u64 pr[16];
    int ex[16];
    const int k = factor_u64(n, pr, ex);
    std::vector<std::pair<u64, int>> ref;
    plain_factor(n, ref);
    bool ok = k == (int)ref.size();
    u128 prod = 1;
    for (int i = 0; ok && i < k; ++i) {
      ok = pr[i] == ref[i].first && ex[i] == ref[i].second && is_prime_u64(pr[i]);
      for (int e = 0; e < ex[i]; ++e) prod *= pr[i];
    }
    if (!ok || (n && prod != n)) {
      if (++bad <= 5) fprintf(stderr, "  MISMATCH n=%" PRIu64 "\n", n);
    }
  }
  printf("%s check-factor %zu values -- mismatches: %" PRIu64 "\n", bad ? "FAIL" : "ok  ", vals.size(), bad);
  return bad ? 1 : 0;
}

// ---------------------------------------------------------------- --check-primality
// sieve/cuda/sieve_cuda.cu:89-102: device Miller-Rabin against the host test on N odd
// integers from LO.
int check_primality(u64 lo, u64 N, int device) {
  Scanner sc;
  std::string msg;
  if (sc.open_device(device, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 1; }
  std::vector<u64> v(N);
  for (u64 i = 0; i < N; ++i) v[i] = (lo | 1ull) + 2 * i;
  std::vector<uint8_t> f(N);
  if (sc.device().primality(v.data(), N, f.data())) {
    fprintf(stderr, "FATAL: %s\n", sc.device().last_error().c_str());
    return 1;
  }
  u64 bad = 0, primes = 0;
  for (u64 i = 0; i < N; ++i) {
    const bool h = is_prime_u64(v[i]);
    primes += h;
    if ((f[i] != 0) != h && ++bad <= 5) fprintf(stderr, "  MISMATCH n=%" PRIu64 "\n", v[i]);
  }
  printf("%s check-primality %" PRIu64 " odd values from %" PRIu64 " -- %" PRIu64 " primes, %" PRIu64
         " mismatches\n", bad ? "FAIL" : "ok  ", N, lo, primes, bad);
  return bad ? 1 : 0;
}

struct Cli {
  u64 lo = 0, hi = 0, M = 2042040, mmax = 10000, seg = 262144, shard = 0, nshard = 1, progress = 0, sieve = 0;
  int threads = 0, device = 0;
  const char *filters = nullptr;
  bool emit_surv = true, plain = false, profile = false, quiet_time = false, host_tail = false, graph = false;
  u32 block = 256, tile_bits = 32768;
  u64 layout = 0, chunk = 1ull << 40;
  Order order = Order::Lane;
};

}  // namespace

int main(int argc, char **argv) {
  if (argc >= 2 && !strcmp(argv[1], "--verify")) return self_test();
  if (argc >= 4 && !strcmp(argv[1], "--check-direct"))
    return check_direct(strtoull(argv[2], 0, 10), strtoull(argv[3], 0, 10));
  if (argc >= 2 && !strcmp(argv[1], "--check-factor"))
    return check_factor(argc >= 3 ? strtoull(argv[2], 0, 10) : 2000);
  if (argc >= 4 && !strcmp(argv[1], "--check-primality")) {
    int dev = 0;
    for (int i = 4; i + 1 < argc; ++i) if (!strcmp(argv[i], "--device")) dev = atoi(argv[i + 1]);
    return check_primality(strtoull(argv[2], 0, 10), strtoull(argv[3], 0, 10), dev);
  }
  if (argc >= 2 && !strcmp(argv[1], "--dump-classes")) {
    u64 M = 2042040;
    for (int i = 2; i + 1 < argc; ++i) if (!strcmp(argv[i], "--wheel")) M = strtoull(argv[i + 1], 0, 10);
    return dump_classes(M);
  }
  if (argc < 3) {
    fprintf(stderr,
            "usage: %s LO HI [--mmax N] [--wheel M] [--seg N] [--threads N]\n"
            "                 [--shard i/n] [--progress N] [--filters FILE] [--sieve P]\n"
            "                 [--no-emit-survivors] [--device D] [--block B] [--tile-bits N]\n"
            "                 [--layout Mp|auto] [--plain] [--chunk N] [--order lane|round]\n"
            "                 [--profile] [--quiet-time] [--host-tail] [--graph]\n"
            "       %s --verify | --dump-classes [--wheel M] | --check-direct LO N\n"
            "          | --check-factor [N] | --check-primality LO N [--device D]\n"
            "Verification sieve on an MI355X: certifies 4/p solvable for every prime in [LO, HI).\n"
            "Output is byte-identical to rung_scan3 for the same flags.\n",
            argv[0], argv[0]);
    return 1;
  }
  Cli o;
  o.lo = strtoull(argv[1], 0, 10);
  o.hi = strtoull(argv[2], 0, 10);
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
    else if (!strcmp(a, "--device") && has) o.device = atoi(argv[++i]);
    else if (!strcmp(a, "--block") && has) o.block = (u32)strtoul(argv[++i], 0, 10);
    else if (!strcmp(a, "--tile-bits") && has) o.tile_bits = (u32)strtoul(argv[++i], 0, 10);
    else if (!strcmp(a, "--layout") && has) {
      ++i;
      o.layout = strcmp(argv[i], "auto") ? strtoull(argv[i], 0, 10) : 0;
    } else if (!strcmp(a, "--plain")) o.plain = true;
    else if (!strcmp(a, "--chunk") && has) o.chunk = strtoull(argv[++i], 0, 10);
    else if (!strcmp(a, "--order") && has) {
      ++i;
      if (!strcmp(argv[i], "lane")) o.order = Order::Lane;
      else if (!strcmp(argv[i], "round")) o.order = Order::Round;
      else { fprintf(stderr, "FATAL: --order takes lane or round\n"); return 1; }
    } else if (!strcmp(a, "--profile")) o.profile = true;
    else if (!strcmp(a, "--host-tail")) o.host_tail = true;
    else if (!strcmp(a, "--graph")) o.graph = true;
    else if (!strcmp(a, "--quiet-time")) o.quiet_time = true;
    else {
      fprintf(stderr, "FATAL: unknown option '%s'\n", a);
      return 1;
    }
  }
  if (!o.seg) o.seg = 262144;
  if (o.block == 0 || o.block > 1024 || o.block % 64 || o.tile_bits == 0 || o.tile_bits % 32 ||
      o.tile_bits > kMaxTileBits) {
    fprintf(stderr, "FATAL: --block must be a multiple of 64 up to 1024, --tile-bits a multiple of 32 up to 61440\n");
    return 1;
  }
  if (o.chunk == 0) o.chunk = 1ull << 40;
  const double t_start = wall_s();
  shard_interval(o.lo, o.hi, o.shard, o.nshard);

  Scanner sc;
  std::string msg;
  if (sc.init_wheel(o.M, o.mmax, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 1; }
  if (!sc.plan().pin)
    fprintf(stderr, "[es_sieve] WARNING: no pinned digest for M=%" PRIu64 " -- running UNPINNED\n", o.M);
  if (check_interval(sc.plan(), o.lo, o.hi, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 1; }
  if (sc.init_filters(o.filters, o.sieve, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 1; }
  if (o.layout && o.M % o.layout) { fprintf(stderr, "FATAL: --layout must divide the wheel modulus\n"); return 1; }
  const Plan &pl = sc.plan();
  fprintf(stderr,
          "[es_sieve] wheel M=%" PRIu64 " lanes=%zu  covers=%zu (mmax=%" PRIu64 ")  wheel_pmin=%" PRIu64 "\n"
          "[es_sieve] filters=%zu fnv64=%016" PRIx64 " source=%s sieve=%" PRIu64 " (+%zu primes)\n",
          o.M, pl.wheel.classes.size(), pl.filters.ncov, o.mmax, pl.wheel.max_pmin, pl.filters.ncov,
          pl.filters.fnv, pl.filter_source.c_str(), o.sieve, pl.filters.mod.size() - pl.filters.ncov);
  sc.threads = o.threads;
  sc.order = o.order;
  sc.round_seg = o.seg;
  sc.tuning().fixed_Mp = o.layout;
  sc.tuning().launch.persistent = !o.plain;
  sc.tuning().launch.block = o.block;
  sc.tuning().launch.max_bits = o.tile_bits;
  sc.tuning().launch.device_tail = !o.host_tail;
  sc.tuning().launch.graph = o.graph;
  if (o.hi > o.lo && sc.open_device(o.device, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 3; }
  if (o.hi > o.lo) fprintf(stderr, "[es_sieve] device %s\n", sc.device().device_name().c_str());
  const double t_ready = wall_s();

  // Chunks of at most o.chunk integers; counters add, lines are merged by the emission
  // key of the whole interval. In round order the integers [lo + r*seg*M, lo +
  // (r+1)*seg*M) hold exactly round r's positions of every lane, so with chunks cut at
  // those boundaries no later chunk has a smaller key: each chunk's lines are sorted and
  // written at once and memory stays bounded by one chunk. Lane order interleaves every
  // chunk in every lane, so its lines are held to the end (as rung_scan3 holds them).
  WindowRecord total;
  total.lo = o.lo;
  total.hi = o.hi;
  struct Line { u64 k1, k2, p; std::string text; };
  std::vector<Line> all;
  ScanTiming acc;
  const bool stream = o.order == Order::Round;
  if (stream) {
    const u128 round_w = (u128)o.seg * o.M;
    const u128 want = (u128)o.chunk < round_w ? round_w : ((u128)o.chunk / round_w) * round_w;
    o.chunk = want > (u128)~0ull ? ~0ull : (u64)want;
  }
  auto flush_lines = [&]() {
    std::sort(all.begin(), all.end(), [](const Line &x, const Line &y) {
      if (x.k1 != y.k1) return x.k1 < y.k1;
      if (x.k2 != y.k2) return x.k2 < y.k2;
      return x.p < y.p;
    });
    for (const Line &l : all)
      if (o.emit_surv || l.text.compare(0, 9, "SURVIVOR ") != 0) fwrite(l.text.data(), 1, l.text.size(), stdout);
    all.clear();
  };
  // Chunks k+1 .. k+kMaxInFlight-1 are submitted before chunk k is collected, so the host
  // tail of one chunk overlaps the device work of the next ones (sieve/cuda/scan.cu:1-6).
  std::deque<Interval> inflight;
  u64 next = o.lo;
  auto submit_next = [&]() -> bool {
    const u64 b = o.hi - next > o.chunk ? next + o.chunk : o.hi;
    if (sc.submit({Interval{next, b}}, msg)) return false;
    inflight.push_back(Interval{next, b});
    next = b;
    return true;
  };
  while (next < o.hi && inflight.size() + 1 < (size_t)DeviceCtx::kMaxInFlight)
    if (!submit_next()) { fprintf(stderr, "%s\n", msg.c_str()); return 3; }
  while (!inflight.empty()) {
    if (next < o.hi && !submit_next()) { fprintf(stderr, "%s\n", msg.c_str()); return 3; }
    const u64 a = inflight.front().lo, b = inflight.front().hi;
    inflight.pop_front();
    std::vector<WindowRecord> recs;
    ScanTiming tm;
    if (sc.collect(recs, tm, msg)) { fprintf(stderr, "%s\n", msg.c_str()); return 3; }
    const WindowRecord &r = recs[0];
    total.positions += r.positions; total.covered += r.covered; total.sieved += r.sieved; total.mr += r.mr;
    total.composite += r.composite; total.rung += r.rung; total.direct += r.direct;
    total.survivors += r.survivors;
    for (size_t i = 0; i < r.lines.size(); ++i) {
      Line L{0, 0, r.line_p[i], r.lines[i]};
      sc.emit_key(L.p, o.lo, L.k1, L.k2);   // the key of the whole interval, not the chunk
      all.push_back(std::move(L));
    }
    acc.plan_ms += tm.plan_ms; acc.batch_ms += tm.batch_ms; acc.device_ms += tm.device_ms;
    acc.sieve_ms += tm.sieve_ms; acc.tail_ms += tm.tail_ms; acc.total_ms += tm.total_ms;
    acc.tiles += tm.tiles; acc.primes += tm.primes; acc.survivors += tm.survivors; acc.replays += tm.replays;
    if (o.profile)
      fprintf(stderr, "[es_sieve] chunk [%" PRIu64 ",%" PRIu64 ") layout=%" PRIu64 " tiles=%" PRIu64
              " survivors=%" PRIu64 " primes=%" PRIu64 "\n", a, b, r.layout_Mp, tm.tiles, tm.survivors, tm.primes);
    if (stream) flush_lines();
  }
  flush_lines();
  const std::string text = sc.record_text(total, o.emit_surv);   // no lines left: the SUMMARY
  fwrite(text.data(), 1, text.size(), stdout);
  fflush(stdout);
  const double t_end = wall_s();
  if (o.profile)
    fprintf(stderr, "[es_sieve] profile: startup %.3f s  plan %.3f ms  batch %.3f ms (device %.3f ms, sieve %.3f ms)"
            "  tail %.3f ms  scan %.3f ms  tiles %" PRIu64 " survivors %" PRIu64 " primes %" PRIu64 " replays %d\n",
            t_ready - t_start, acc.plan_ms, acc.batch_ms, acc.device_ms, acc.sieve_ms, acc.tail_ms, acc.total_ms,
            acc.tiles, acc.survivors, acc.primes, acc.replays);
  if (o.quiet_time) fprintf(stderr, "%.6f\n", t_end - t_start);
  return total.survivors ? 2 : 0;
}
