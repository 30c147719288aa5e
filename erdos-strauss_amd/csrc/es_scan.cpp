// es_scan.cpp -- windows -> tiles -> device batch -> host tail -> records.
#include "es_scan.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>
#include <condition_variable>
#include <mutex>
#include <thread>

#include "es_tail.h"

namespace es {

namespace {

double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

u64 gcd64(u64 a, u64 b) {
  while (b) { const u64 t = a % b; a = b; b = t; }
  return a;
}

}  // namespace

namespace {

// Workers that live for the process: a parallel_for hands them an index range and
// waits; spawning threads per call would cost more than the work it spreads.
class Pool {
 public:
  explicit Pool(int n) {
    for (int i = 0; i < n; ++i) workers_.emplace_back([this] { loop(); });
  }
  ~Pool() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_.notify_all();
    for (std::thread &t : workers_) t.join();
  }
  int size() const { return (int)workers_.size(); }
  // Runs fn(i) for i < n on up to `width` workers plus the caller.
  void run(size_t n, int width, const std::function<void(size_t)> &fn) {
    std::unique_lock<std::mutex> call(call_mu_);   // one parallel_for at a time
    {
      std::lock_guard<std::mutex> g(mu_);
      fn_ = &fn;
      n_ = n;
      next_.store(0);
      active_ = std::min(width, (int)workers_.size());
      pending_ = active_;
      ++gen_;
    }
    cv_.notify_all();
    drain();
    std::unique_lock<std::mutex> g(mu_);
    done_.wait(g, [this] { return pending_ == 0; });
    fn_ = nullptr;
  }

 private:
  void drain() {
    for (size_t i; (i = next_.fetch_add(1)) < n_;) (*fn_)(i);
  }
  void loop() {
    u64 seen = 0;
    for (;;) {
      int slot;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_.wait(g, [&] { return stop_ || gen_ != seen; });
        if (stop_) return;
        seen = gen_;
        if (active_ <= 0) continue;
        slot = active_--;
      }
      (void)slot;
      drain();
      std::lock_guard<std::mutex> g(mu_);
      if (--pending_ == 0) done_.notify_all();
    }
  }
  std::vector<std::thread> workers_;
  std::mutex mu_, call_mu_;
  std::condition_variable cv_, done_;
  const std::function<void(size_t)> *fn_ = nullptr;
  size_t n_ = 0;
  std::atomic<size_t> next_{0};
  int active_ = 0, pending_ = 0;
  u64 gen_ = 0;
  bool stop_ = false;
};

Pool &pool() {
  static Pool p((int)std::max(1u, std::thread::hardware_concurrency()) - 1);
  return p;
}

}  // namespace

void parallel_for(size_t n, int threads, const std::function<void(size_t)> &fn) {
  if (n == 0) return;
  const int T = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  if (T <= 1 || n == 1) {
    for (size_t i = 0; i < n; ++i) fn(i);
    return;
  }
  pool().run(n, (int)std::min<size_t>((size_t)T - 1, n - 1), fn);
}

void shard_interval(u64 &lo, u64 &hi, u64 shard, u64 nshard) {
  if (nshard == 0) nshard = 1;
  if (nshard > 1) {
    const u64 span = (hi - lo + nshard - 1) / nshard;
    const u64 a = lo + shard * span, b = std::min(hi, a + span);
    lo = a;
    hi = std::max(a, b);
  }
}

int Scanner::init_wheel(u64 M, u64 mmax, std::string &msg) {
  if (build_plan(M, mmax, plan_, msg)) return 1;
  char b[320];
  snprintf(b, sizeof b,
           " wheelM=%" PRIu64 " lanes=%zu mmax=%" PRIu64 " ucert_rmax=%u ucert_umax=%u wheel_pmin=%" PRIu64
           " classes_sha256=%s",
           plan_.M, plan_.wheel.classes.size(), plan_.mmax, kUcertRmax, kUcertUmax, plan_.wheel.max_pmin,
           plan_.pin ? plan_.pin->sha256 : "unpinned");
  summary_mid_ = b;
  return 0;
}

int Scanner::init_filters(const char *filter_file, u64 sieve, std::string &msg) {
  if (plan_filters(plan_, filter_file, sieve, msg)) return 1;
  // Per candidate sub-wheel: lanes, and expected marks per integer from how many lanes
  // each filter hits (lanes c' with c' == res mod gcd(m, Mp)).
  est_.clear();
  const FilterSet &fs = plan_.filters;
  for (u64 Mp : layout_candidates(plan_.M)) {
    std::vector<u64> lanes;
    for (u64 c : plan_.wheel.classes) lanes.push_back(c % Mp);
    std::sort(lanes.begin(), lanes.end());
    lanes.erase(std::unique(lanes.begin(), lanes.end()), lanes.end());
    std::map<u64, std::map<u64, u64>> hist;   // g -> residue -> lanes
    double marks = 0;
    for (size_t i = 0; i < fs.mod.size(); ++i) {
      const u64 m = fs.mod[i], g = gcd64(m, Mp), s = m / g;
      auto it = hist.find(g);
      if (it == hist.end()) {
        std::map<u64, u64> h;
        for (u64 c : lanes) ++h[c % g];
        it = hist.emplace(g, std::move(h)).first;
      }
      auto jt = it->second.find(fs.res[i] % m % g);
      if (jt != it->second.end()) marks += (double)jt->second / (double)s;
    }
    est_.push_back(Estimate{Mp, (double)lanes.size(), marks / (double)Mp, (double)lanes.size() / (double)Mp});
  }
  return 0;
}

int Scanner::open_device(int device, std::string &msg) {
  if (dev_.open(device)) { msg = "FATAL: HIP device: " + dev_.last_error(); return -1; }
  dev_open_ = true;
  return 0;
}

u64 Scanner::choose_layout(u64 w) const {
  if (tune_.fixed_Mp) return tune_.fixed_Mp;
  if (!tune_.table.empty()) {
    const double lw = std::log2((double)std::max<u64>(w, 1));
    u64 best = tune_.table[0].second;
    double gap = 1e300;
    for (const auto &e : tune_.table)
      if (std::fabs(e.first - lw) < gap) { gap = std::fabs(e.first - lw); best = e.second; }
    return best;
  }
  const double nf = (double)plan_.filters.mod.size();
  const double maxb = (double)tune_.launch.max_bits;
  u64 best = plan_.M;
  double best_cost = 1e300;
  for (const Estimate &e : est_) {
    const double per_lane = (double)w / (double)e.Mp;
    const double tiles = e.nl * (per_lane <= 1.0 ? per_lane : std::ceil(per_lane / maxb));
    const double cost = tiles * nf * tune_.c_pair +
                        (double)w * (e.marks_per_int * tune_.c_mark + e.bits_per_int * tune_.c_bit);
    if (cost < best_cost) { best_cost = cost; best = e.Mp; }
  }
  return best;
}

int Scanner::autotune(const std::vector<Interval> &sample, std::string &report, std::string &msg) {
  if (!dev_open_) { msg = "FATAL: HIP device not open"; return -1; }
  const Tuning saved = tune_;
  char line[256];
  // Reference height: the median window start of the sample (10^15 without one), kept
  // clear of the ceiling so every probe window is valid.
  std::vector<u64> los;
  for (const Interval &w : sample) if (w.hi > w.lo) los.push_back(w.lo);
  std::sort(los.begin(), los.end());
  u64 base = los.empty() ? 1000000000000000ull : los[los.size() / 2];
  const u64 floor_ = std::max(plan_.wheel.max_pmin, plan_.cover_pmax) + 1;
  base = std::max(base, floor_);
  if (base > kDirectPMax - (1ull << 40)) base = kDirectPMax - (1ull << 40);
  auto time_batch = [&](const std::vector<Interval> &wins, double &sieve_ms) -> int {
    std::vector<WindowRecord> recs;
    BatchResult br;
    ScanTiming tm;
    if (device_pass(wins, recs, br, tm, msg)) return -1;      // warm (uploads the layout)
    std::vector<double> t;
    for (int r = 0; r < 5; ++r) {
      if (device_pass(wins, recs, br, tm, msg)) return -1;
      t.push_back(tm.sieve_ms > 0 ? tm.sieve_ms : tm.device_ms);
    }
    std::sort(t.begin(), t.end());
    sieve_ms = t[2];
    return 0;
  };
  // Every launch shape on the sample, then the closest few timed again in interleaved
  // rounds: single medians of five passes differ by a few percent from one call to the
  // next (clocks, the device's other work), enough to rank close shapes the wrong way.
  // Interleaving spreads any drift over all the finalists alike.
  auto sweep_launch = [&](const char *tag, LaunchCfg &best_cfg, double &best_ms) -> int {
    std::vector<std::pair<double, LaunchCfg>> timed;
    for (u32 block : {128u, 256u, 512u, 1024u})
      for (u32 bits : {8192u, 16384u, 32768u})
        for (u32 waves : {16u, 24u, 0u})
          for (u32 group : {1u, 2u, 4u}) {
            if (waves && waves < block / 64) continue;
            tune_.launch.block = block;
            tune_.launch.max_bits = bits;
            tune_.launch.waves_per_cu = waves;
            tune_.launch.group = group;
            double ms;
            if (time_batch(sample, ms)) return -1;
            snprintf(line, sizeof line, "%s block=%u tile_bits=%u waves_per_cu=%u group=%u sieve_ms=%.4f\n", tag,
                     block, bits, waves, group, ms);
            report += line;
            timed.emplace_back(ms, tune_.launch);
          }
    std::sort(timed.begin(), timed.end(), [](const std::pair<double, LaunchCfg> &a,
                                             const std::pair<double, LaunchCfg> &b) { return a.first < b.first; });
    const size_t nf = std::min<size_t>(4, timed.size());
    std::vector<std::vector<double>> again(nf);
    for (int round = 0; round < 5; ++round)
      for (size_t i = 0; i < nf; ++i) {
        tune_.launch = timed[i].second;
        double ms;
        if (time_batch(sample, ms)) return -1;
        again[i].push_back(ms);
      }
    for (size_t i = 0; i < nf; ++i) {
      std::sort(again[i].begin(), again[i].end());
      const double ms = again[i][2];
      const LaunchCfg &c = timed[i].second;
      snprintf(line, sizeof line, "%s-final block=%u tile_bits=%u waves_per_cu=%u group=%u sieve_ms=%.4f\n", tag,
               c.block, c.max_bits, c.waves_per_cu, c.group, ms);
      report += line;
      if (ms < best_ms) { best_ms = ms; best_cfg = c; }
    }
    return 0;
  };
  tune_.table.clear();
  tune_.launch.graph = false;
  std::vector<std::pair<double, u64>> table;
  struct Row { std::vector<Interval> wins; std::vector<std::pair<double, u64>> timed; };
  std::vector<Row> rows;
  for (int lw = 16; lw <= 34; lw += 2) {
    const u64 W = 1ull << lw;
    const u64 n = std::max<u64>(1, std::min<u64>(256, (1ull << 30) / W));
    std::vector<Interval> wins;
    for (u64 i = 0; i < n; ++i) wins.push_back(Interval{base + i * W, base + (i + 1) * W});
    u64 best = 0;
    double best_ms = 1e300;
    std::vector<std::pair<double, u64>> timed;
    for (const Estimate &e : est_) {
      tune_.fixed_Mp = e.Mp;
      double ms;
      if (time_batch(wins, ms)) { tune_ = saved; return -1; }
      snprintf(line, sizeof line, "layout W=2^%d x%llu Mp=%llu sieve_ms=%.4f\n", lw, (unsigned long long)n,
               (unsigned long long)e.Mp, ms);
      report += line;
      if (ms < best_ms) { best_ms = ms; best = e.Mp; }
      timed.emplace_back(ms, e.Mp);
    }
    table.emplace_back((double)lw, best);
    rows.push_back(Row{wins, timed});
  }
  tune_.fixed_Mp = saved.fixed_Mp;
  tune_.table = table;
  // Launch shape on the sample batch itself.
  if (!sample.empty()) {
    LaunchCfg best_cfg = tune_.launch;
    double best_ms = 1e300;
    if (sweep_launch("launch", best_cfg, best_ms)) { tune_ = saved; return -1; }
    tune_.launch = best_cfg;
    // The layouts above were timed at the default launch shape, and close ones can rank
    // the other way at the tuned shape (one 2^34 window at 10^15: wheel 2042040 is 14%
    // ahead of 120120 at block 256, 37% at block 512 with 16384-position tiles). So every
    // layout within 1.5x of its row's best is timed again at the tuned shape, and the
    // fastest there takes the row.
    for (size_t i = 0; i < rows.size(); ++i) {
      const Row &row = rows[i];
      double row_best = 1e300;
      for (const auto &c : row.timed) row_best = std::min(row_best, c.first);
      u64 best = table[i].second;
      double best_ms = 1e300;
      int n = 0;
      for (const auto &c : row.timed) {
        if (c.first > 1.5 * row_best) continue;
        ++n;
        tune_.fixed_Mp = c.second;
        double ms;
        if (time_batch(row.wins, ms)) { tune_ = saved; return -1; }
        if (ms < best_ms) { best_ms = ms; best = c.second; }
      }
      tune_.fixed_Mp = saved.fixed_Mp;
      if (n > 1) {
        snprintf(line, sizeof line, "refine W=2^%.0f Mp=%llu sieve_ms=%.4f (%d layouts)\n", table[i].first,
                 (unsigned long long)best, best_ms, n);
        report += line;
      }
      tune_.table[i].second = best;
    }
    // Those rows were timed on batches of one window size, and for the larger sizes a few
    // windows make too few tiles to fill the grid, so layouts of many short lanes win there
    // on parallelism alone. In a batch of many windows the total work decides instead (the
    // 512-window batch from 1e9 to 1.8e19: wheel 840 for its 2^26 windows runs the sieve in
    // 83 us against 87 with 10920). So every row the sample uses is tried on the sample
    // itself with each layout (rows in order; the two fastest of a row timed again in
    // interleaved rounds), keeping what is faster there, and the launch shape is chosen
    // again for the layouts that result.
    double cur_ms;
    if (time_batch(sample, cur_ms)) { tune_ = saved; return -1; }
    auto row_of = [&](u64 w) {
      const double lw = std::log2((double)std::max<u64>(w, 1));
      size_t r = 0;
      for (size_t k = 1; k < tune_.table.size(); ++k)
        if (std::fabs(tune_.table[k].first - lw) < std::fabs(tune_.table[r].first - lw)) r = k;
      return r;
    };
    std::vector<bool> used(tune_.table.size(), false);
    for (const Interval &w : sample) if (w.hi > w.lo) used[row_of(w.hi - w.lo)] = true;
    bool changed = false;
    for (size_t i = 0; i < tune_.table.size(); ++i) {
      if (!used[i]) continue;
      const u64 was = tune_.table[i].second;
      std::vector<std::pair<double, u64>> cand{{cur_ms, was}};
      for (const Estimate &e : est_) {
        if (e.Mp == was) continue;
        tune_.table[i].second = e.Mp;
        double ms;
        if (time_batch(sample, ms)) { tune_ = saved; return -1; }
        cand.emplace_back(ms, e.Mp);
      }
      std::sort(cand.begin(), cand.end());
      u64 keep = cand[0].second;
      double keep_ms = cand[0].first;
      if (cand.size() > 1) {   // the two fastest again, in interleaved rounds
        std::vector<double> t0, t1;
        for (int round = 0; round < 3; ++round) {
          double ms;
          tune_.table[i].second = cand[0].second;
          if (time_batch(sample, ms)) { tune_ = saved; return -1; }
          t0.push_back(ms);
          tune_.table[i].second = cand[1].second;
          if (time_batch(sample, ms)) { tune_ = saved; return -1; }
          t1.push_back(ms);
        }
        std::sort(t0.begin(), t0.end());
        std::sort(t1.begin(), t1.end());
        if (t1[1] < t0[1]) { keep = cand[1].second; keep_ms = t1[1]; }
        else keep_ms = t0[1];
      }
      tune_.table[i].second = keep;
      cur_ms = keep_ms;
      if (keep != was) {
        snprintf(line, sizeof line, "sample W=2^%.0f Mp=%llu sieve_ms=%.4f\n", tune_.table[i].first,
                 (unsigned long long)keep, cur_ms);
        report += line;
        changed = true;
      }
    }
    if (changed) {
      LaunchCfg re_cfg = tune_.launch;
      double re_ms = 1e300;
      if (sweep_launch("relaunch", re_cfg, re_ms)) { tune_ = saved; return -1; }
      tune_.launch = re_cfg;
    }
  }
  tune_.launch.graph = saved.launch.graph;
  for (const auto &e : tune_.table) {
    snprintf(line, sizeof line, "choice W=2^%.0f Mp=%llu\n", e.first, (unsigned long long)e.second);
    report += line;
  }
  snprintf(line, sizeof line, "choice block=%u tile_bits=%u waves_per_cu=%u group=%u\n", tune_.launch.block,
           tune_.launch.max_bits, tune_.launch.waves_per_cu, tune_.launch.group);
  report += line;
  return 0;
}

int Scanner::mark_count(u64 lo, u64 hi, u64 Mp, u64 &marks, u64 &bits, u64 &pairs, u64 &tiles,
                        std::string &msg) {
  marks = bits = pairs = tiles = 0;
  if (hi <= lo) return 0;
  if (!Mp) Mp = choose_layout(hi - lo);
  const Layout *L = host_layout(Mp, msg);
  if (!L) return -1;
  const u64 maxb = tune_.launch.max_bits;
  for (size_t i = 0; i < L->lanes.size(); ++i) {
    u64 k0, nb;
    lane_span(lo, hi, L->lanes[i], Mp, k0, nb);
    for (u64 b = 0; b < nb; b += maxb) {
      const u64 kt = k0 + b, n = std::min(maxb, nb - b);
      ++tiles;
      bits += n;
      const u16 *row = L->off.data() + i * L->nf;
      for (u32 j = 0; j < L->nf; ++j) {
        if (row[j] == kNoHit) continue;
        const u64 s = L->step[j], o = row[j], km = kt % s;
        const u64 first = o >= km ? o - km : o + s - km;
        ++pairs;
        if (first < n) marks += (n - 1 - first) / s + 1;
      }
    }
  }
  return 0;
}

const Layout *Scanner::layout(u64 Mp) const {
  auto it = layouts_.find(Mp);
  return it == layouts_.end() ? nullptr : it->second.get();
}

const Layout *Scanner::host_layout(u64 Mp, std::string &msg) {
  auto it = layouts_.find(Mp);
  if (it != layouts_.end()) return it->second.get();
  std::unique_ptr<Layout> L(new Layout);
  if (!build_layout(plan_.M, Mp, plan_.wheel.classes, plan_.filters, *L, msg)) {
    msg = "FATAL: layout " + std::to_string(Mp) + ": " + msg;
    return nullptr;
  }
  const Layout *p = L.get();
  layouts_[Mp] = std::move(L);
  return p;
}

int Scanner::ensure_layout(u64 Mp, std::string &msg) {
  auto it = layout_id_.find(Mp);
  if (it != layout_id_.end()) return it->second;
  if (!dev_open_) { msg = "FATAL: HIP device not open"; return -1; }
  const Layout *L = host_layout(Mp, msg);
  if (!L) return -1;
  const int id = dev_.add_layout(*L);
  if (id < 0) { msg = "FATAL: HIP device: " + dev_.last_error(); return -1; }
  layout_id_[Mp] = id;
  return id;
}

int Scanner::device_submit(const std::vector<Interval> &wins, std::string &msg) {
  pending_.emplace_back();
  Pending &pb = pending_.back();
  pb.wins = wins;
  if (plan_and_launch(pb, msg)) {
    pending_.pop_back();
    return -1;
  }
  pb.launched = true;
  return 0;
}

// Plans pb's windows into tiles and launches its device batch. Touches only pb, the
// layouts and the device context, so the submission thread can run it while the caller
// collects older batches.
int Scanner::plan_and_launch(Pending &pb, std::string &msg) {
  const double t0 = now_ms();
  const std::vector<Interval> &wins = pb.wins;
  ScanTiming &tm = pb.tm;
  std::vector<WindowRecord> &out = pb.out;
  out.assign(wins.size(), WindowRecord());
  auto fail = [&]() { return -1; };
  const u32 maxb = tune_.launch.max_bits;
  // Tiles go out longest first, so the persistent queue ends with the short ones. Only
  // load balance depends on that order (never a result), so tiles are placed straight
  // into buckets of nbits / 64 (counted in a first pass) instead of being sorted.
  const u32 nbk = (maxb >> 6) + 1;
  auto bucket = [&](u32 nb) { return nbk - 1 - (nb >> 6); };
  struct LaneSpan { u64 k0, nbits; };
  std::vector<LaneSpan> spans;
  std::vector<int> ids(wins.size(), -1);
  std::vector<u32> start(nbk + 1, 0);
  for (size_t w = 0; w < wins.size(); ++w) {
    WindowRecord &r = out[w];
    r.lo = wins[w].lo;
    r.hi = wins[w].hi;
    if (r.hi <= r.lo) continue;
    const u64 Mp = choose_layout(r.hi - r.lo);
    const int id = ensure_layout(Mp, msg);
    if (id < 0) return fail();
    ids[w] = id;
    r.layout_Mp = Mp;
    const Layout &L = *layouts_[Mp];
    const WindowSpan span(r.lo, r.hi, Mp);
    for (u32 i = 0; i < (u32)L.lanes.size(); ++i) {
      LaneSpan ls;
      span.lane(L.lanes[i], ls.k0, ls.nbits);
      spans.push_back(ls);
      for (u64 b = 0; b < ls.nbits; b += maxb) ++start[bucket((u32)std::min<u64>(maxb, ls.nbits - b)) + 1];
    }
  }
  for (u32 b = 0; b < nbk; ++b) start[b + 1] += start[b];
  std::vector<TileDesc> tiles(start[nbk]);
  size_t si = 0;
  for (size_t w = 0; w < wins.size(); ++w) {
    if (ids[w] < 0) continue;
    const u32 nl = (u32)layouts_[out[w].layout_Mp]->lanes.size();
    for (u32 i = 0; i < nl; ++i, ++si) {
      const LaneSpan &ls = spans[si];
      for (u64 b = 0; b < ls.nbits; b += maxb) {
        const u32 nb = (u32)std::min<u64>(maxb, ls.nbits - b);
        tiles[start[bucket(nb)]++] = TileDesc{ls.k0 + b, nb, i, (u32)w, (u32)ids[w]};
      }
    }
  }
  tm.tiles = tiles.size();
  const double t1 = now_ms();
  tm.plan_ms = t1 - t0;

  if (!dev_open_) { msg = "FATAL: HIP device not open"; return fail(); }
  if (dev_.launch(tiles, (u32)wins.size(), plan_.filters.mod.size() != plan_.filters.ncov, tune_.launch)) {
    msg = "FATAL: HIP device: " + dev_.last_error();
    return fail();
  }
  tm.batch_ms = now_ms() - t1;
  tm.total_ms = now_ms() - t0;
  return 0;
}

int Scanner::device_collect(std::vector<Interval> &wins, std::vector<WindowRecord> &out, BatchResult &br,
                            ScanTiming &tm, std::string &msg) {
  Pending pb;
  {
    std::unique_lock<std::mutex> lk(mu_);
    if (pending_.empty()) { msg = "FATAL: no batch in flight"; return -1; }
    cv_launched_.wait(lk, [&] { return pending_.front().launched; });   // the submission thread's turn
    pb = std::move(pending_.front());
    pending_.pop_front();
  }
  if (pb.rc) { msg = pb.msg; return -1; }
  const double t1 = now_ms();
  wins = std::move(pb.wins);
  out = std::move(pb.out);
  tm = pb.tm;
  if (dev_.finish(br)) {
    msg = "FATAL: HIP device: " + dev_.last_error();
    return -1;
  }
  const double t2 = now_ms();
  tm.batch_ms += t2 - t1;
  tm.device_ms = br.device_ms;
  tm.sieve_ms = br.sieve_ms;
  tm.mr_ms = br.mr_ms;
  tm.rung_ms = br.tail_kernel_ms;
  tm.sieve_event_ms = br.sieve_event_ms;
  tm.mr_event_ms = br.mr_event_ms;
  tm.rung_event_ms = br.tail_event_ms;
  tm.replays = br.replays;
  tm.survivors = br.surv_p.size();

  // Counters (rung_scan3.cpp:764-772): covered counts cover kills only.
  const bool sieve = plan_.filters.mod.size() != plan_.filters.ncov;
  for (size_t w = 0; w < wins.size(); ++w) {
    WindowRecord &r = out[w];
    const WinCounts &c = br.counts[w];
    r.positions = c.positions;
    r.mr = c.survivors;
    if (!sieve) {
      r.covered = c.positions - c.survivors;
    } else {
      r.covered = c.positions - c.after_covers;
      r.sieved = c.after_covers - c.survivors;
    }
  }
  for (size_t i = 0; i < br.surv_p.size(); ++i)
    if (!br.surv_prime[i]) ++out[br.surv_w[i]].composite;
  tm.total_ms += now_ms() - t1;
  return 0;
}

int Scanner::device_pass(const std::vector<Interval> &wins, std::vector<WindowRecord> &out, BatchResult &br,
                         ScanTiming &tm, std::string &msg) {
  if (in_flight()) { msg = "FATAL: batches are in flight: collect them first"; return -1; }
  std::vector<Interval> w;
  if (device_submit(wins, msg)) return -1;
  return device_collect(w, out, br, tm, msg);
}

int Scanner::survivors(const std::vector<Interval> &wins, std::vector<WindowRecord> &out, std::vector<u64> &p,
                       std::vector<uint8_t> &prime, ScanTiming &tm, std::string &msg) {
  BatchResult br;
  if (device_pass(wins, out, br, tm, msg)) return -1;
  std::vector<size_t> idx(br.surv_p.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    return br.surv_w[a] != br.surv_w[b] ? br.surv_w[a] < br.surv_w[b] : br.surv_p[a] < br.surv_p[b];
  });
  p.resize(idx.size());
  prime.resize(idx.size());
  for (size_t i = 0; i < idx.size(); ++i) {
    p[i] = br.surv_p[idx[i]];
    prime[i] = br.surv_prime[idx[i]];
  }
  return 0;
}

int Scanner::scan(const std::vector<Interval> &wins, std::vector<WindowRecord> &out, ScanTiming &tm,
                  std::string &msg) {
  BatchResult br;
  if (device_pass(wins, out, br, tm, msg)) return -1;
  finish_records(wins, out, br, tm);
  return 0;
}

int Scanner::submit(const std::vector<Interval> &wins, std::string &msg) {
  std::lock_guard<std::mutex> lk(mu_);
  if (pending_.size() >= (size_t)DeviceCtx::kMaxInFlight) { msg = "FATAL: every batch slot is in flight"; return -1; }
  if (!worker_.joinable()) worker_ = std::thread([this] { submission_loop(); });
  pending_.emplace_back();
  pending_.back().wins = wins;
  ++unlaunched_;
  cv_work_.notify_one();
  return 0;
}

// The submission thread: plans and launches submitted batches in order, so the caller's
// collect (records of batch i) runs while batches i+1.. are being planned and launched.
// Only this thread launches while batches are in flight; the caller then waits on the
// batches' pinned completion flags and makes no HIP call meanwhile.
void Scanner::submission_loop() {
  std::unique_lock<std::mutex> lk(mu_);
  for (;;) {
    cv_work_.wait(lk, [&] { return stop_ || unlaunched_ > 0; });
    if (stop_ && unlaunched_ == 0) return;
    Pending &pb = pending_[pending_.size() - unlaunched_];   // references survive push_back/pop_front
    lk.unlock();
    std::string msg;
    const int rc = plan_and_launch(pb, msg);
    lk.lock();
    pb.rc = rc;
    pb.msg = msg;
    pb.launched = true;
    --unlaunched_;
    cv_launched_.notify_all();
  }
}

Scanner::~Scanner() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  cv_work_.notify_one();
  if (worker_.joinable()) worker_.join();
}

int Scanner::collect(std::vector<WindowRecord> &out, ScanTiming &tm, std::string &msg) {
  BatchResult br;
  std::vector<Interval> wins;
  if (device_collect(wins, out, br, tm, msg)) return -1;
  finish_records(wins, out, br, tm);
  return 0;
}

void Scanner::finish_records(const std::vector<Interval> &wins, std::vector<WindowRecord> &out, BatchResult &br,
                             ScanTiming &tm) {
  const double t2 = now_ms();
  std::vector<size_t> prime_idx;
  for (size_t i = 0; i < br.surv_p.size(); ++i)
    if (br.surv_prime[i]) prime_idx.push_back(i);
  tm.primes = prime_idx.size();

  // Stage D results from the device become lines here; what the device declined (and
  // stage E) runs on the host, in parallel, into per-prime slots.
  std::vector<std::string> lines(prime_idx.size());
  std::vector<int> kinds(prime_idx.size(), -1);
  std::vector<size_t> host_j;
  for (size_t j = 0; j < prime_idx.size(); ++j) {
    const size_t i = prime_idx[j];
    if (!br.rung.empty() && br.rung[i].status == 1) {
      const RungResult &g = br.rung[i];
      format_rung(br.surv_p[i], g.A, g.r, ((u128)g.u_hi << 64) | g.u_lo, lines[j]);
      kinds[j] = 0;
    } else if (!br.rung.empty() && br.rung[i].status == 4) {   // stage E on the device
      const RungResult &g = br.rung[i];
      format_direct(br.surv_p[i], g.r & 0xFFu, g.A, g.r >> 8, g.u_lo, lines[j]);
      kinds[j] = 1;
    } else {
      host_j.push_back(j);
    }
  }
  tm.host_primes = host_j.size();
  parallel_for(host_j.size(), threads, [&](size_t h) {
    const size_t j = host_j[h];
    kinds[j] = certify_prime(br.surv_p[prime_idx[j]], lines[j]);
  });
  const double t3 = now_ms();
  tm.tail_ms = t3 - t2;

  // Emission order.
  struct Keyed { u64 k1, k2, p; size_t j; };
  std::vector<std::vector<Keyed>> per(wins.size());
  for (size_t j = 0; j < prime_idx.size(); ++j) {
    const size_t i = prime_idx[j];
    const u64 p = br.surv_p[i];
    const u32 w = br.surv_w[i];
    u64 k1, k2;
    emit_key(p, wins[w].lo, k1, k2);
    per[w].push_back(Keyed{k1, k2, p, j});
  }
  for (size_t w = 0; w < wins.size(); ++w) {
    std::vector<Keyed> &v = per[w];
    std::sort(v.begin(), v.end(), [](const Keyed &a, const Keyed &b) {
      if (a.k1 != b.k1) return a.k1 < b.k1;
      if (a.k2 != b.k2) return a.k2 < b.k2;
      return a.p < b.p;
    });
    WindowRecord &r = out[w];
    for (const Keyed &k : v) {
      const int kind = kinds[k.j];
      if (kind == 0) ++r.rung;
      else if (kind == 1) ++r.direct;
      else ++r.survivors;
      r.lines.push_back(std::move(lines[k.j]));
      r.line_p.push_back(k.p);
    }
  }
  tm.total_ms += now_ms() - t2;
}

void Scanner::emit_key(u64 p, u64 lo, u64 &k1, u64 &k2) const {
  const u64 M = plan_.M;
  const std::vector<u64> &cls = plan_.wheel.classes;
  const u64 rho = p % M;
  const u64 rank = (u64)(std::lower_bound(cls.begin(), cls.end(), rho) - cls.begin());
  k1 = rank;
  k2 = 0;
  if (order == Order::Round) {
    const u64 first = lo + (rho + M - lo % M) % M;
    k1 = ((p - first) / M) / round_seg;
    k2 = rank;
  }
}

int Scanner::warm_layouts(std::string &msg) {
  std::vector<u64> want;
  if (tune_.fixed_Mp) want.push_back(tune_.fixed_Mp);
  for (const auto &e : tune_.table) want.push_back(e.second);
  if (want.empty())
    for (int lw = 10; lw <= 40; ++lw) want.push_back(choose_layout(1ull << lw));
  for (u64 Mp : want)
    if (ensure_layout(Mp, msg) < 0) return -1;
  return 0;
}

void Scanner::append_record(std::string &s, const WindowRecord &r, bool emit_survivors) const {
  for (const std::string &l : r.lines)
    if (emit_survivors || l.compare(0, 9, "SURVIVOR ") != 0) s += l;
  // SUMMARY lo=.. hi=.. <the scanner's constant fields> positions=.. ... survivors=..
  // (rung_scan3.cpp:1153-1161): the same text as the printf format, written into a stack
  // buffer with put_u64 instead of parsing a format string per record.
  char b[640];
  char *w = b;
  auto lit = [&](const char *k, size_t n) {
    memcpy(w, k, n);
    w += n;
  };
  auto num = [&](u64 v) {
    char t[24];
    char *e = t + sizeof t;
    const char *d = put_u64(e, v);
    memcpy(w, d, e - d);
    w += e - d;
  };
  lit("SUMMARY lo=", 11);
  num(r.lo);
  lit(" hi=", 4);
  num(r.hi);
  lit(summary_mid_.data(), summary_mid_.size());
  lit(" positions=", 11);
  num(r.positions);
  lit(" covered=", 9);
  num(r.covered);
  lit(" sieved=", 8);
  num(r.sieved);
  lit(" mr=", 4);
  num(r.mr);
  lit(" composite=", 11);
  num(r.composite);
  lit(" rung=", 6);
  num(r.rung);
  lit(" direct=", 8);
  num(r.direct);
  lit(" survivors=", 11);
  num(r.survivors);
  *w++ = '\n';
  s.append(b, w - b);
}

std::string Scanner::record_text(const WindowRecord &r, bool emit_survivors) const {
  std::string s;
  append_record(s, r, emit_survivors);
  return s;
}

}  // namespace es
