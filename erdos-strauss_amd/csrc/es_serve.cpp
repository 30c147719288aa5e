// es_serve.cpp -- continuous-batching request server (see es_serve.h).
#include "es_serve.h"

#include <algorithm>
#include <chrono>

namespace es {

namespace {
double clock_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}
// Device bytes per survivor entry (p, 2^64 mod p, window, flag, stage-D result) and per
// tile descriptor, as the batch buffers hold them.
constexpr u64 kSurvivorBytes = 8 + 8 + 4 + 1 + 32;
constexpr u64 kTileBytes = 24;
}  // namespace

Server::Server(Scanner &sc, const ServeConfig &cfg) : sc_(sc), cfg_(cfg) {}

Server::~Server() { stop(); }

double Server::now_ms() const { return clock_ms() - t0_; }

// Upper bound of the device memory [lo, hi) can occupy: every position a survivor, every
// lane of the widest layout cut once per slice, plus its window counters.
u64 Server::footprint(u64 lo, u64 hi) const {
  const Plan &P = sc_.plan();
  const u64 lanes = P.wheel.classes.size();
  const u64 len = hi - lo;
  const u64 slices = (len + cfg_.slice - 1) / cfg_.slice;
  const u64 positions = (u64)((double)len * (double)lanes / (double)P.M) + lanes * slices;
  const u64 tiles = positions / 1024 + lanes * slices;
  return positions * kSurvivorBytes + tiles * kTileBytes + slices * 24;
}

int Server::start(std::string &msg) {
  if (started_) return 0;
  if (cfg_.slice == 0 || cfg_.batch_ints < cfg_.slice || cfg_.batch_windows == 0 || cfg_.slots == 0) {
    msg = "bad serving configuration";
    return -1;
  }
  if (sc_.warm_layouts(msg)) return -1;
  size_t free_b = 0, total_b = 0;
  if (sc_.device().mem_info(free_b, total_b)) {
    msg = "FATAL: HIP device: " + sc_.device().last_error();
    return -1;
  }
  st_.budget = cfg_.hbm_budget ? cfg_.hbm_budget : (u64)(cfg_.hbm_fraction * (double)total_b);
  st_.resident = sc_.device().resident_bytes();
  avail_ = st_.budget > st_.resident ? st_.budget - st_.resident : 0;
  t0_ = clock_ms();
  stop_ = false;
  started_ = true;
  worker_ = std::thread([this] { loop(); });
  return 0;
}

u64 Server::submit(u64 lo, u64 hi) {
  std::lock_guard<std::mutex> lk(mu_);
  const u64 id = next_id_++;
  Req &r = reqs_[id];
  r.lo = lo;
  r.hi = std::max(lo, hi);
  r.next = r.lo;
  r.info.arrival_ms = now_ms();
  ++st_.submitted;
  std::string msg;
  if (check_interval(sc_.plan(), lo, hi, msg)) {
    // rung_scan3 refuses before scanning (rung_scan3.cpp:1057-1069): exit 1, no record.
    r.record = msg;
    r.info.exit_code = 1;
    r.info.admit_ms = r.info.done_ms = r.info.arrival_ms;
    r.done = true;
    ++st_.refused;
    ++st_.completed;
    cv_done_.notify_all();
    return id;
  }
  r.info.reserved = footprint(r.lo, r.hi);
  queue_.push_back(id);
  st_.max_queued = std::max<u32>(st_.max_queued, (u32)queue_.size());
  cv_work_.notify_one();
  return id;
}

int Server::wait(u64 id, double timeout_ms, std::string &record, RequestInfo &info) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = reqs_.find(id);
  if (it == reqs_.end()) return -1;
  auto ready = [&] { return it->second.done || st_.error; };
  if (timeout_ms < 0) {
    cv_done_.wait(lk, ready);
  } else if (!cv_done_.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), ready)) {
    return 1;
  }
  if (!it->second.done) return -1;
  record.swap(it->second.record);
  info = it->second.info;
  reqs_.erase(it);
  return 0;
}

ServeStats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  return st_;
}

void Server::stop() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (!started_) return;
    stop_ = true;
  }
  cv_work_.notify_all();
  if (worker_.joinable()) worker_.join();
  started_ = false;
}

// Merges the slice records of r into its record (called with mu_ held).
void Server::finish(Req &r) {
  WindowRecord m;
  m.lo = r.lo;
  m.hi = r.hi;
  struct Keyed { u64 k1, k2, p; size_t part, line; };
  std::vector<Keyed> keys;
  for (size_t i = 0; i < r.parts.size(); ++i) {
    const WindowRecord &w = r.parts[i];
    m.positions += w.positions;
    m.covered += w.covered;
    m.sieved += w.sieved;
    m.mr += w.mr;
    m.composite += w.composite;
    m.rung += w.rung;
    m.direct += w.direct;
    m.survivors += w.survivors;
    for (size_t j = 0; j < w.lines.size(); ++j) {
      u64 k1, k2;
      sc_.emit_key(w.line_p[j], r.lo, k1, k2);
      keys.push_back(Keyed{k1, k2, w.line_p[j], i, j});
    }
  }
  std::sort(keys.begin(), keys.end(), [](const Keyed &a, const Keyed &b) {
    if (a.k1 != b.k1) return a.k1 < b.k1;
    if (a.k2 != b.k2) return a.k2 < b.k2;
    return a.p < b.p;
  });
  for (const Keyed &k : keys) {
    m.lines.push_back(std::move(r.parts[k.part].lines[k.line]));
    m.line_p.push_back(k.p);
  }
  r.parts.clear();
  r.record = sc_.record_text(m, cfg_.emit_survivors);
  r.info.exit_code = m.survivors ? 2 : 0;
  r.info.done_ms = now_ms();
  r.done = true;
}

void Server::loop() {
  // Up to DeviceCtx::kMaxInFlight batches are on the device at once: the worker builds and
  // submits the next batch (admitting whatever arrived meanwhile) before it collects the
  // oldest, so the slices of one request run in consecutive batches side by side.
  struct Flight {
    std::vector<Interval> wins;
    std::vector<u64> owner;
    u64 ints = 0;
    double t0 = 0;
  };
  std::unique_lock<std::mutex> lk(mu_);
  std::deque<Flight> flight;
  std::vector<WindowRecord> recs;
  for (;;) {
    // Requests with nothing left in flight or to issue (empty ones included) complete,
    // which frees their slots and budget for the admission below.
    bool any = false;
    for (size_t i = 0; i < active_.size();) {
      Req &r = reqs_[active_[i]];
      if (r.done_ints == r.hi - r.lo) {
        finish(r);
        st_.reserved -= r.info.reserved;
        ++st_.completed;
        active_.erase(active_.begin() + i);
        any = true;
      } else {
        ++i;
      }
    }
    if (any) cv_done_.notify_all();
    // Admission: FIFO while a slot and the budget allow (one request always fits, so an
    // oversized one runs alone rather than never).
    while (!queue_.empty() && active_.size() < cfg_.slots) {
      Req &r = reqs_[queue_.front()];
      if (!active_.empty() && st_.reserved + r.info.reserved > avail_) break;
      st_.reserved += r.info.reserved;
      st_.peak_reserved = std::max(st_.peak_reserved, st_.reserved);
      r.info.admit_ms = now_ms();
      active_.push_back(queue_.front());
      queue_.pop_front();
    }
    st_.max_active = std::max<u32>(st_.max_active, (u32)active_.size());
    // An admitted empty request completes at once.
    for (size_t i = 0; i < active_.size();) {
      Req &r = reqs_[active_[i]];
      if (r.hi == r.lo) {
        finish(r);
        st_.reserved -= r.info.reserved;
        ++st_.completed;
        active_.erase(active_.begin() + i);
        cv_done_.notify_all();
      } else {
        ++i;
      }
    }
    bool issuable = false;
    for (u64 id : active_) issuable |= reqs_[id].next < reqs_[id].hi;
    if (issuable && flight.size() < (size_t)DeviceCtx::kMaxInFlight) {
      // One batch: round robin over the admitted requests, one slice each per pass.
      Flight f;
      for (bool more = true; more && f.wins.size() < cfg_.batch_windows;) {
        more = false;
        for (u64 id : active_) {
          Req &r = reqs_[id];
          if (r.next >= r.hi) continue;
          const u64 len = std::min(cfg_.slice, r.hi - r.next);
          if (!f.wins.empty() && f.ints + len > cfg_.batch_ints) continue;
          f.wins.push_back(Interval{r.next, r.next + len});
          f.owner.push_back(id);
          r.next += len;
          f.ints += len;
          more = true;
          if (f.wins.size() >= cfg_.batch_windows) break;
        }
      }
      std::vector<u64> in_batch(f.owner);
      std::sort(in_batch.begin(), in_batch.end());
      in_batch.erase(std::unique(in_batch.begin(), in_batch.end()), in_batch.end());
      st_.max_batch_requests = std::max<u32>(st_.max_batch_requests, (u32)in_batch.size());
      for (u64 id : in_batch) ++reqs_[id].info.batches;
      lk.unlock();
      std::string msg;
      f.t0 = clock_ms();
      const int rc = sc_.submit(f.wins, msg);
      lk.lock();
      if (rc) {
        st_.error = 1;
        st_.error_msg = msg;
        cv_done_.notify_all();
        break;
      }
      flight.push_back(std::move(f));
      continue;   // admit and submit again before collecting, while slots remain
    }
    if (flight.empty()) {
      // Nothing on the device and nothing to issue: every admitted request has completed
      // above, so the loop sleeps until a request arrives (or stop), unless one is queued.
      if (active_.empty() && queue_.empty()) {
        if (stop_ || st_.error) break;
        cv_work_.wait(lk);
      }
      continue;
    }
    // Collect the oldest batch in flight.
    Flight f = std::move(flight.front());
    flight.pop_front();
    lk.unlock();
    ScanTiming tm;
    std::string msg;
    recs.clear();
    const int rc = sc_.collect(recs, tm, msg);
    const double b1 = clock_ms();
    lk.lock();
    st_.busy_ms += b1 - f.t0;
    if (rc) {
      st_.error = 1;
      st_.error_msg = msg;
      cv_done_.notify_all();
      break;
    }
    ++st_.batches;
    st_.slices += f.wins.size();
    st_.ints += f.ints;
    for (size_t i = 0; i < f.wins.size(); ++i) {
      Req &r = reqs_[f.owner[i]];
      r.done_ints += f.wins[i].hi - f.wins[i].lo;
      r.parts.push_back(std::move(recs[i]));
    }
  }
  // A failure leaves batches in flight; collect them so the scanner is usable again.
  lk.unlock();
  for (size_t k = 0; k < flight.size(); ++k) {
    ScanTiming tm;
    std::string msg;
    recs.clear();
    (void)sc_.collect(recs, tm, msg);
  }
}

}  // namespace es
