// es_device.h -- the device side of the funnel, seen from host C++: upload layouts,
// run a batch of tiles through marking + compaction + Miller-Rabin, read results back.
// Only es_device.hip includes HIP headers; everything here is plain C++.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "es_plan.h"

namespace es {

// One unit of device work: bits [k0, k0 + nbits) of one lane of one layout, for one
// window of the batch. p = lanes[lane] + Mp * k for bit k.
struct TileDesc {
  u64 k0;
  u32 nbits;
  u32 lane;
  u32 win;
  u32 layout;
};
static_assert(sizeof(TileDesc) == 24, "TileDesc is shared with the device");

// Per-window device counters (sums over the window's tiles).
struct WinCounts {
  u64 positions;      // wheel positions in the window
  u64 after_covers;   // still alive after stage B
  u64 survivors;      // alive after stage B' (== after_covers without a sieve)
};

// Stage D computed on the device for one survivor. status: 1 = RUNG (r, A, u), 2 = no
// rung up to 255 has a witness, 3 = the device declined (the host finishes), 0 = not prime.
struct RungResult {
  u64 A, u_lo, u_hi;
  u32 r, status;
};

struct BatchResult {
  std::vector<WinCounts> counts;      // [nwin]
  std::vector<u64> surv_p;            // stage-C survivors (any order)
  std::vector<u32> surv_w;            // their window index
  std::vector<uint8_t> surv_prime;    // 1 if prime (deterministic MR on the device)
  std::vector<RungResult> rung;       // per survivor, when stage D ran on the device
  double device_ms = 0;               // sieve + MR kernels, HIP events
  // Kernel times from the kernels' own device-clock stamps (first workgroup's start to
  // last wave's end), graph replays included.
  double sieve_ms = 0;                // sieve kernel alone
  double mr_ms = 0;                   // base-2 Miller-Rabin kernel alone
  double tail_kernel_ms = 0;          // stage D kernel (or the other MR bases) alone
  // The same three from HIP events on the eager dispatches (-1 from a graph replay): an
  // eager dispatch that carries events also carries the marker gap before it.
  double sieve_event_ms = 0, mr_event_ms = 0, tail_event_ms = 0;
  int replays = 0;                    // survivor-buffer overflows that forced a rerun
};

// Largest tile: one LDS byte per position, within the 64 KiB a workgroup gets by
// default (the rest of the kernel's LDS is a few hundred bytes).
constexpr u32 kMaxTileBits = 61440;

// Values one call of the allocation-free primality entry point may test: its scratch
// (a count and one u64 per value) is a fixed device array of 8 MiB per GPU.
constexpr size_t kNoAllocMaxValues = size_t(1) << 20;

struct LaunchCfg {
  bool persistent = true;   // work-queue persistent kernel vs one workgroup per tile
  bool graph = false;       // replay a captured hipGraph (persistent form only)
  u32 block = 256;          // threads per workgroup
  u32 max_bits = 32768;     // largest tile in positions (LDS bytes of the byte map)
  u32 waves_per_cu = 16;    // resident waves per CU the persistent grid is sized for (0: all that fit)
  bool device_tail = true;  // stage D on the device (k_rung); the host finishes what it declines
  u32 group = 4;            // most tiles one k_sieve work item groups (1: every tile alone)
};

class DeviceCtx {
 public:
  DeviceCtx();
  ~DeviceCtx();
  DeviceCtx(const DeviceCtx &) = delete;
  DeviceCtx &operator=(const DeviceCtx &) = delete;

  // Selects the device and creates the stream. 0 or -1 with last_error().
  int open(int device);
  // Uploads a layout's tables; returns its id (>= 0) or -1.
  int add_layout(const Layout &L);
  // Runs one batch (launch then finish; no batch may be in flight). Tiles must reference
  // valid layouts/lanes, nbits <= cfg.max_bits.
  int run(const std::vector<TileDesc> &tiles, u32 nwin, bool has_sieve, const LaunchCfg &cfg,
          BatchResult &out);
  // Pipelined form: launch enqueues a batch and returns at once (at most kMaxInFlight);
  // finish waits for the oldest batch in flight and returns its results. The host can
  // finish one batch while the device runs the next.
  int launch(const std::vector<TileDesc> &tiles, u32 nwin, bool has_sieve, const LaunchCfg &cfg);
  int finish(BatchResult &out);
  int in_flight() const;
  static constexpr int kMaxInFlight = 8;
  // Deterministic MR of arbitrary values on the device (parity checks).
  int primality(const u64 *vals, size_t n, uint8_t *flags);
  // The same test on caller-owned device buffers, enqueued on `stream` without waiting and
  // without allocating (state in the library's fixed device scratch). -2 for an argument
  // error (too many values, not device memory), -1 for a HIP failure.
  int primality_noalloc(const u64 *d_vals, size_t n, uint8_t *d_flags, void *stream);
  // Stage D on the device for primes vals[0..n) (one RungResult each), then stage E for
  // those no rung certifies; direct_only skips stage D (every prime goes to stage E).
  // status: 1 RUNG (A, r, u), 2 neither stage certified p, 3 declined (the host
  // finishes p), 4 SOLVED (A = c, u_lo = k, r = a | d << 8).
  int rung(const u64 *vals, size_t n, RungResult *out, bool direct_only = false);
  // Measured LDS atomic-AND rate of the whole device (32-bit operations per second).
  int lds_peak(double &ops_per_s);
  int lds_scatter_peak(double &stores_per_s);   // scattered byte stores per second
  const std::string &last_error() const { return err_; }
  int device() const { return device_; }
  size_t layout_bytes() const { return layout_bytes_; }
  // Device memory: free and total bytes (hipMemGetInfo) and what this context holds.
  int mem_info(size_t &free_b, size_t &total_b);
  size_t resident_bytes() const;
  std::string device_name() const;
  int cu_count() const { return cus_; }

 private:
  struct Impl;
  Impl *impl_ = nullptr;
  int grow_surv(int slot, size_t need);
  int enqueue_batch(int slot);
  int read_batch(int slot, BatchResult &out);
  int wait_slot(int slot);
  std::string err_;
  int device_ = -1;
  int cus_ = 0;
  size_t lds_limit_ = 0;   // dynamic LDS one workgroup may allocate (bytes)
  size_t layout_bytes_ = 0;
  void *scratch_ = nullptr;   // this device's copy of the scratch array (device address)
  double clock_ms_per_tick_ = 1e-5;   // the device's constant clock (s_memrealtime)
};

}  // namespace es
