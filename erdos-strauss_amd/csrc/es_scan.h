// es_scan.h -- the funnel end to end for a batch of intervals ("windows"): pick a
// layout per window, cut windows into tiles, run the device batch, certify the primes
// on the host, and assemble each window's record exactly as rung_scan3 prints it.
#pragma once
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "es_device.h"
#include "es_plan.h"

namespace es {

struct Interval { u64 lo, hi; };

struct WindowRecord {
  u64 lo = 0, hi = 0;
  u64 positions = 0, covered = 0, sieved = 0, mr = 0, composite = 0, rung = 0, direct = 0,
      survivors = 0;
  u64 layout_Mp = 0;                 // the sub-wheel the window was sieved in
  std::vector<std::string> lines;    // certificate lines in emission order
  std::vector<u64> line_p;           // the prime each line certifies
};

struct ScanTiming {
  double plan_ms = 0, device_ms = 0, sieve_ms = 0, batch_ms = 0, tail_ms = 0, total_ms = 0;
  double mr_ms = 0, rung_ms = 0;   // base-2 MR kernel, stage D kernel (device-clock stamps, like sieve_ms)
  // The three kernel times from HIP events (eager launches only; -1 from a graph replay).
  double sieve_event_ms = 0, mr_event_ms = 0, rung_event_ms = 0;
  u64 tiles = 0, primes = 0, survivors = 0;
  u64 host_primes = 0;        // primes whose stage D/E ran on the host
  int replays = 0;
};

// Emission order of certificate lines inside a record.
enum class Order {
  Lane,    // rung_scan3: (rank of p mod M in the wheel, p)   rung_scan3.cpp:1146-1151
  Round,   // sieve_cuda: (floor(J / seg), rank, p)            sieve/cuda/scan.cu:295-309
};

struct Tuning {
  // Cost model in VALU-op units per window of W integers in layout Mp, used until the
  // autotuner has measured the device:
  //   tiles * nf * c_pair  +  W * (marks_per_int * c_mark + bits_per_int * c_bit)
  double c_pair = 24.0, c_mark = 7.0, c_bit = 0.15;
  u64 fixed_Mp = 0;           // nonzero: always this layout (must divide M)
  // Measured: the fastest layout per window size (log2 W -> Mp), from autotune().
  std::vector<std::pair<double, u64>> table;
  LaunchCfg launch;
};

class Scanner {
 public:
  Scanner() = default;
  ~Scanner();
  Scanner(const Scanner &) = delete;
  Scanner &operator=(const Scanner &) = delete;
  // Phase 1: wheel + pin (and default tuning). Returns 0 or 1 with msg.
  int init_wheel(u64 M, u64 mmax, std::string &msg);
  // Phase 2: filter set. Returns 0 or 1 with msg.
  int init_filters(const char *filter_file, u64 sieve, std::string &msg);
  // Phase 3: the device (HIP). Returns 0 or -1 with msg.
  int open_device(int device, std::string &msg);

  // The funnel over a batch of windows; each window gets its own record. Returns 0, or
  // -1 on a device error (msg).
  int scan(const std::vector<Interval> &wins, std::vector<WindowRecord> &out, ScanTiming &tm,
           std::string &msg);
  // Pipelined scan: submit plans a batch and launches it without waiting (at most
  // DeviceCtx::kMaxInFlight in flight); collect finishes the oldest submitted batch into the records
  // scan would have produced. The host tail of one batch then runs while the device
  // works on the next (sieve/cuda/scan.cu:1-6, 197-207). scan refuses while a batch is
  // in flight.
  int submit(const std::vector<Interval> &wins, std::string &msg);
  int collect(std::vector<WindowRecord> &out, ScanTiming &tm, std::string &msg);
  size_t in_flight() const {
    std::lock_guard<std::mutex> lk(mu_);
    return pending_.size();
  }
  // Device stages only: each window's counters (positions..composite) and all stage-C
  // survivors of the batch, ascending by (window, p), with their primality.
  int survivors(const std::vector<Interval> &wins, std::vector<WindowRecord> &out,
                std::vector<u64> &p, std::vector<uint8_t> &prime, ScanTiming &tm, std::string &msg);

  // Exact work of [lo, hi) in layout Mp (0 = the tuner's pick), as the kernel issues it:
  // bit clears (marks), bitmap bits, and (tile, filter) setups. Host arithmetic only.
  int mark_count(u64 lo, u64 hi, u64 Mp, u64 &marks, u64 &bits, u64 &pairs, u64 &tiles, std::string &msg);

  // Layout the tuner picks for a window of w integers.
  u64 choose_layout(u64 w) const;
  // Measures the device: kernel time of every candidate layout for window sizes 2^16
  // .. 2^34 (at the sample's median height) -> the size-to-layout table; then workgroup
  // size x tile size on the sample batch itself. Returns 0, or -1 (msg) on a device
  // error; `report` gets one line per measurement.
  int autotune(const std::vector<Interval> &sample, std::string &report, std::string &msg);
  // Makes sure the layout for Mp exists on the device; returns its device id or -1.
  int ensure_layout(u64 Mp, std::string &msg);
  // The host tables of layout Mp, built on first use (no device needed); null on error.
  const Layout *host_layout(u64 Mp, std::string &msg);
  const Layout *layout(u64 Mp) const;

  // Emission key (k1, k2, then p) of the certificate line of prime p in a record that
  // starts at lo, for the current order.
  void emit_key(u64 p, u64 lo, u64 &k1, u64 &k2) const;
  // Uploads every layout the tuner can pick, so no later batch pays for an upload.
  int warm_layouts(std::string &msg);

  // The full record text: lines (SURVIVOR lines dropped if !emit_survivors) + SUMMARY.
  std::string record_text(const WindowRecord &r, bool emit_survivors) const;
  // The same text appended to s.
  void append_record(std::string &s, const WindowRecord &r, bool emit_survivors) const;

  const Plan &plan() const { return plan_; }
  Tuning &tuning() { return tune_; }
  DeviceCtx &device() { return dev_; }
  int threads = 0;            // host tail threads (0 = hardware concurrency)
  Order order = Order::Lane;
  u64 round_seg = 262144;     // positions per lane per round for Order::Round

  struct Estimate { u64 Mp; double nl, marks_per_int, bits_per_int; };
  const std::vector<Estimate> &estimates() const { return est_; }

 private:
  // Tiles + device batch + counters; composite counts filled, primes left in br.
  int device_pass(const std::vector<Interval> &wins, std::vector<WindowRecord> &out, BatchResult &br,
                  ScanTiming &tm, std::string &msg);
  // device_pass in two halves: plan and launch, then finish and count.
  int device_submit(const std::vector<Interval> &wins, std::string &msg);
  int device_collect(std::vector<Interval> &wins, std::vector<WindowRecord> &out, BatchResult &br,
                     ScanTiming &tm, std::string &msg);
  // Stage D leftovers on the host, then the emission order, into out's lines.
  void finish_records(const std::vector<Interval> &wins, std::vector<WindowRecord> &out, BatchResult &br,
                      ScanTiming &tm);
  struct Pending {
    std::vector<Interval> wins;
    std::vector<WindowRecord> out;   // lo, hi and layout filled at submission
    ScanTiming tm;                   // plan and launch times so far
    bool launched = false;           // planned and launched (or failed: rc, msg)
    int rc = 0;
    std::string msg;
  };
  int plan_and_launch(Pending &pb, std::string &msg);
  void submission_loop();
  // Batches submitted and not collected, oldest first. The submission thread launches
  // the last unlaunched_ of them; mu_ guards the deque and the counters.
  std::deque<Pending> pending_;
  mutable std::mutex mu_;
  std::condition_variable cv_work_, cv_launched_;
  std::thread worker_;
  size_t unlaunched_ = 0;
  bool stop_ = false;
  Plan plan_;
  Tuning tune_;
  DeviceCtx dev_;
  bool dev_open_ = false;
  std::vector<Estimate> est_;
  std::map<u64, std::unique_ptr<Layout>> layouts_;
  std::map<u64, int> layout_id_;
  std::string summary_mid_;   // SUMMARY's fields that do not depend on the window
};

// Splits [lo, hi) into shard i of n the way rung_scan3.cpp:1022-1027 does.
void shard_interval(u64 &lo, u64 &hi, u64 shard, u64 nshard);

// Runs fn(i) for i in [0, n) on `threads` host threads (dynamic assignment).
void parallel_for(size_t n, int threads, const std::function<void(size_t)> &fn);

}  // namespace es
