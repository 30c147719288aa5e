// es_serve.h -- a continuous-batching request server over one Scanner.
//
// Requests are intervals [lo, hi) whose answer is exactly the record rung_scan3 prints
// for `rung_scan3 lo hi` with the scanner's wheel and filters (rung_scan3.cpp:1128-1162).
// A worker thread owns the scanner and keeps up to eight device batches in flight:
//   * admission: FIFO, bounded by request slots and by an HBM budget (a fraction of the
//     device's memory, minus what the scanner keeps resident); each request reserves an
//     upper bound of the device memory its slices occupy until it completes;
//   * batching: every batch takes the next slice (<= `slice` integers) of each admitted
//     request, round robin, up to `batch_ints` integers / `batch_windows` slices, and
//     runs them as one persistent launch -- requests that arrive while batches are on the
//     device join the next one, so short requests never wait behind long ones, and the
//     slices of a long request run in consecutive batches side by side;
//   * completion: a request's slice records merge into one record (lines in the
//     scanner's emission order over the whole request, counters summed), which is what
//     the same request gives alone.
#pragma once
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "es_scan.h"

namespace es {

struct ServeConfig {
  u64 slice = 1ull << 32;          // integers per slice of a request
  u64 batch_ints = 1ull << 35;     // integers per device batch (one slice always fits)
  u32 batch_windows = 4096;        // slices per device batch
  u32 slots = 256;                 // requests admitted at once
  double hbm_fraction = 0.8;       // admission budget as a fraction of device memory
  u64 hbm_budget = 0;              // bytes; 0 = hbm_fraction * total device memory
  bool emit_survivors = true;
};

struct RequestInfo {
  double arrival_ms = 0, admit_ms = 0, done_ms = 0;   // on the server's clock
  u64 batches = 0;                                     // device batches it was part of
  u64 reserved = 0;                                    // bytes it held while admitted
  int exit_code = 0;                                   // 0, 2 (survivors), 1 (refused)
};

struct ServeStats {
  u64 submitted = 0, completed = 0, refused = 0, batches = 0, slices = 0, ints = 0;
  u64 budget = 0, resident = 0, reserved = 0, peak_reserved = 0;
  u32 max_active = 0, max_batch_requests = 0, max_queued = 0;
  double busy_ms = 0;              // summed submit-to-collect time of the device batches
  int error = 0;                   // nonzero after a device failure (message in error_msg)
  std::string error_msg;
};

class Server {
 public:
  Server(Scanner &sc, const ServeConfig &cfg);
  ~Server();
  // Uploads the layouts, sizes the budget and starts the worker. 0 or -1 (msg).
  int start(std::string &msg);
  // Enqueues [lo, hi) arriving now; returns its id. Interval refusals complete at once
  // with exit code 1 and the refusal message as the record.
  u64 submit(u64 lo, u64 hi);
  // Waits up to timeout_ms (< 0: forever) for request id: 0 done (record, info filled),
  // 1 timed out, -1 unknown id or server failure. A finished request is handed out once.
  int wait(u64 id, double timeout_ms, std::string &record, RequestInfo &info);
  ServeStats stats();
  double now_ms() const;
  // Finishes the admitted and queued requests, then stops the worker.
  void stop();

 private:
  struct Req {
    u64 lo = 0, hi = 0, next = 0, done_ints = 0;
    RequestInfo info;
    std::vector<WindowRecord> parts;
    std::string record;
    bool done = false;
  };
  u64 footprint(u64 lo, u64 hi) const;
  void finish(Req &r);
  void loop();

  Scanner &sc_;
  ServeConfig cfg_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::map<u64, Req> reqs_;
  std::deque<u64> queue_;
  std::vector<u64> active_;
  u64 next_id_ = 1;
  u64 avail_ = 0;
  ServeStats st_;
  bool stop_ = false, started_ = false;
  std::thread worker_;
  double t0_ = 0;
};

}  // namespace es
