/* es_sieve.h -- C ABI of the MI355X verification-sieve funnel.
 *
 * Plain C types only: integers, pointers, sizes. Every entry point returns 0 on success
 * and a negative status on failure (es_last_error() then says why); nothing aborts
 * across this boundary. A scanner owns its device memory (tables resident in HBM) and
 * is used from one host thread at a time.
 *
 * Each entry point names the reference interface it replaces (paths relative to the
 * reference repository):
 *   sieve/rung_scan3.cpp        the CPU scanner (CLI, record, stage D/E)
 *   sieve/cuda/scan.cuh:6-20    ScanOpts / scan_run, the GPU driver
 *   sieve/cuda/ref_host.h       the host accessors the GPU driver binds
 *   sieve/cuda/mark.cuh,
 *   sieve/cuda/survivors.cuh    the device stages (k_mark, k_compact, k_primality)
 */
#ifndef ES_SIEVE_H
#define ES_SIEVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ES_OK 0
#define ES_ERR_CONFIG (-1)   /* a refusal rung_scan3 makes with exit code 1 */
#define ES_ERR_DEVICE (-2)   /* HIP device missing or a HIP call failed */
#define ES_ERR_ARG (-3)      /* bad argument to this API */

typedef struct es_scanner es_scanner;

/* Scanner configuration: the flags of rung_scan3.cpp:984-1016 that shape the filter
 * set, plus the device ordinal and the host-tail thread count. */
typedef struct es_config {
  uint64_t wheel_M;          /* --wheel, default 2042040 */
  uint64_t mmax;             /* --mmax, default 10000 */
  uint64_t sieve_P;          /* --sieve, default 0 (stage B' off) */
  const char *filter_file;   /* --filters FILE, or NULL for generated covers */
  int threads;               /* host tail threads, 0 = all */
  int device;                /* HIP device ordinal; -1 = host-side planning only (every
                                call that needs the device then fails ES_ERR_DEVICE) */
} es_config;

/* SUMMARY counters of one window (rung_scan3.cpp:651-660, :1153-1161). */
typedef struct es_counts {
  uint64_t positions, covered, sieved, mr, composite, rung, direct, survivors;
} es_counts;

/* Timing of the last scan call, milliseconds (scan.cu:150-162's --profile split).
 * sieve_ms, mr_ms and rung_ms time the marking, base-2 Miller-Rabin and stage-D kernels
 * alone (HIP events on their dispatches); a batch replayed from a graph reports -1. */
typedef struct es_timing {
  double plan_ms, batch_ms, device_ms, sieve_ms, tail_ms, total_ms;
  uint64_t tiles, survivors, primes, host_primes;
  int replays;
  double mr_ms, rung_ms;
  /* The kernel times above (sieve_ms, mr_ms, rung_ms) come from the kernels' own stamps on
   * the device's constant clock, graph replays included; these three are the same times
   * from HIP events on eager launches (-1 when the batch was replayed from a graph). */
  double sieve_event_ms, mr_event_ms, rung_event_ms;
} es_timing;

/* Fills *cfg with rung_scan3's defaults. */
void es_config_default(es_config *cfg);

/* Builds the wheel (pin-checked), the filter set and the device context.
 * Replaces ref3_build_wheel + ref3_filters (sieve/cuda/ref_host.h:39-50) and the
 * setup half of scan_run (sieve/cuda/scan.cu:55-192). ES_ERR_CONFIG on a pin mismatch
 * or an unreadable / empty filter file (rung_scan3.cpp:1041-1048, :1085-1097). */
int es_scanner_open(const es_config *cfg, es_scanner **out);
void es_scanner_close(es_scanner *s);

/* Message of the last failure on this thread. */
const char *es_last_error(void);

/* Interval refusals of rung_scan3.cpp:1057-1069 (HI above the stage-E ceiling, LO not
 * above the filter table's pmin): ES_OK or ES_ERR_CONFIG. */
int es_check_interval(const es_scanner *s, uint64_t lo, uint64_t hi);

/* The funnel over nwin windows [lo[i], hi[i]) in one device batch, producing each
 * window's record exactly as rung_scan3 prints it for that interval: certificate
 * lines in lane order, then SUMMARY. Records are concatenated into *text (free with
 * es_free); rec_len[i] (optional) receives each record's byte length and exit_code[i]
 * (optional) 2 if the window has survivors, else 0. Replaces rung_scan3's main scan
 * (rung_scan3.cpp:1128-1162) and scan_run (sieve/cuda/scan.cu:55-360). */
int es_scan_records(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin,
                    int emit_survivors, char **text, size_t *text_len, size_t *rec_len,
                    int *exit_code);

/* Pipelined es_scan_records. es_scan_submit plans the windows and launches their batch
 * without waiting; es_scan_collect finishes the oldest submitted batch and returns its
 * records exactly as es_scan_records would have (rec_len and exit_code need room for that
 * batch's window count, which *nwin receives). Up to ES_MAX_IN_FLIGHT batches are in
 * flight (a further submit returns ES_ERR_ARG), so the host tail of one (stage-D
 * leftovers, record text) runs while the device works on the next: the overlap of
 * sieve_cuda's driver, which keeps round s on the host tail while round s+1 runs on the
 * device (sieve/cuda/scan.cu:1-6, 197-207). es_scan_records, es_scan_counts,
 * es_survivors and es_autotune refuse (ES_ERR_ARG) while a batch is in flight. */
#define ES_MAX_IN_FLIGHT 8
int es_scan_submit(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin);
int es_scan_collect(es_scanner *s, int emit_survivors, char **text, size_t *text_len,
                    size_t *rec_len, int *exit_code, size_t *nwin);
/* Batches submitted and not yet collected (0 .. ES_MAX_IN_FLIGHT). */
int es_scan_in_flight(const es_scanner *s);

/* As es_scan_records, counters only. */
int es_scan_counts(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin,
                   es_counts *out);

/* Stage-C survivors of [lo, hi) (positions no filter removed) with their primality,
 * in ascending p, plus the window's counters. The parity hook of sieve_cuda's
 * --check-mark / --check-compact (sieve/cuda/sieve_cuda.cu:30-88). *n receives the
 * survivor count even when it exceeds cap (then nothing beyond cap is written). */
int es_survivors(es_scanner *s, uint64_t lo, uint64_t hi, uint64_t *p, uint8_t *is_prime,
                 size_t cap, size_t *n, es_counts *counts);

/* Deterministic Miller-Rabin on the device for arbitrary values (flags[i] = 1 iff
 * vals[i] is prime). Replaces k_primality (sieve/cuda/survivors.cuh:47-48) as used by
 * --check-primality (sieve/cuda/sieve_cuda.cu:89-102). */
int es_primality(es_scanner *s, const uint64_t *vals, size_t n, uint8_t *flags);

/* es_primality for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). It has no counterpart in the
 * reference. d_vals (n values) and d_flags (n verdicts out, 1 iff prime, the same bytes
 * es_primality writes) are device memory of the scanner's device, allocated by the
 * caller; the kernels are enqueued on `stream` (a hipStream_t; NULL = the null stream)
 * and the call returns without waiting. Nothing is allocated: the count and a 64-bit word
 * per value live in a fixed device array of the library, one copy per GPU held until the
 * process ends, so n may not exceed ES_NOALLOC_MAX_VALUES (ES_ERR_ARG beyond it, or for a
 * pointer that is not device memory). Every call on one GPU shares that array: calls that
 * overlap on one GPU (other streams, other threads) race; calls in order on one stream
 * do not. */
#define ES_NOALLOC_MAX_VALUES (1u << 20)
int es_primality_noalloc(es_scanner *s, const uint64_t *d_vals, size_t n, uint8_t *d_flags, void *stream);

/* Stage D then stage E for one prime p, on the host: writes "RUNG p A r u\n",
 * "SOLVED p a c d k\n" or "SURVIVOR p\n" into buf and returns 0, 1 or 2 (negative if
 * buf is too small). Replaces ref3_solve_rung / ref3_solve_direct
 * (sieve/cuda/ref_host.h:20-21; rung_scan3.cpp:578-647). */
int es_certify(uint64_t p, char *buf, size_t cap);

/* Stage D on the device for n primes: one wave per prime (k_rung), lines formatted on
 * the host exactly as es_certify writes them; a prime the device declines is finished
 * by the host. Lines are concatenated into *text (free with es_free); *on_host
 * (optional) receives how many primes the host finished. The parity hook of the device
 * tail (the reference keeps stage D on the host, sieve/cuda/ref_host.h:20-21). */
int es_certify_batch(es_scanner *s, const uint64_t *p, size_t n, char **text, size_t *on_host);

/* Stage E alone on the device for n primes (k_rung with no rungs: one workgroup per
 * prime, pairs (a, d) taken in order by its waves): "SOLVED p a c d k\n" per prime stage E
 * certifies, "SURVIVOR p\n" per prime it cannot, in input order, into *text (free with
 * es_free); a prime the device declines is finished by the host (*on_host counts them).
 * Pairs whose n = p + 4a^2 d exceeds (2^32-1)^2 are skipped, as the reference's guard
 * skips them. Replaces solve_direct3 as
 * --check-direct drives it (rung_scan3.cpp:624-647, :798-817). In a scan, stage E runs
 * on the device the same way for every prime stage D leaves uncertified. */
int es_direct_batch(es_scanner *s, const uint64_t *p, size_t n, char **text, size_t *on_host);

/* Stage D placement for the following scans: 1 = on the device (default), 0 = host. */
int es_set_device_tail(es_scanner *s, int on);

/* Emission order of certificate lines in the following records: 0 = lane-major
 * (rung_scan3: rank of p mod M in the wheel, then p; rung_scan3.cpp:1146-1151), 1 =
 * round-major (sieve_cuda: floor(J / seg), rank, p, with J = (p - first lane position)
 * / M and seg positions per lane per round; sieve/cuda/scan.cu:295-309). seg 0 keeps
 * the current value (default 262144, rung_scan3's --seg default). */
int es_set_order(es_scanner *s, int order, uint64_t seg);

/* Sub-wheel layout for the following scans: Mp must divide the wheel modulus; 0
 * restores the autotuner. No counterpart in the reference (whose only layout is
 * Mp = M, rung_scan3.cpp:711-716); output does not depend on it. */
int es_set_layout(es_scanner *s, uint64_t Mp);

/* Launch shape: persistent work-queue kernel (1) or one workgroup per tile (0),
 * threads per workgroup (multiple of 64, <= 1024), bits per tile (multiple of 32).
 * Replaces ScanOpts::block (sieve/cuda/scan.cuh:14) and --seg's role as the bitmap
 * size (rung_scan3.cpp:999). */
int es_set_launch(es_scanner *s, int persistent, uint32_t block, uint32_t max_bits);

/* Resident waves per CU that the persistent grid is sized for (default 16; 0 = as many
 * workgroups as fit). Fewer resident waves than fit can mark faster. No counterpart in
 * the reference (one grid of ceil(tiles / block) workgroups, sieve/cuda/scan.cu:212-221);
 * output does not depend on it. */
int es_set_residency(es_scanner *s, uint32_t waves_per_cu);

/* Most tiles one work item of the marking kernel groups (1 to 16, default 4; 1 = every
 * tile alone). Tiles shorter than half the tile size are grouped, up to this many and one
 * tile's positions, so they share one pass of the per-tile init, counts and barriers.
 * autotune sweeps it. No counterpart in the reference (one thread block per segment,
 * sieve/cuda/scan.cu:212-221); output does not depend on it. */
int es_set_grouping(es_scanner *s, uint32_t max_tiles);

/* Replays the persistent form's device pipeline (counter reset, marking, Miller-Rabin)
 * from one captured hipGraph (1) or launches it kernel by kernel (0, the default).
 * No counterpart in the reference, which launches per round (sieve/cuda/scan.cu:212-221). */
int es_set_graph(es_scanner *s, int on);

/* Cost-model coefficients of the layout autotuner (see es_set_layout). */
int es_set_tuning(es_scanner *s, double c_pair, double c_mark, double c_bit);

/* The project's autotuner: times every candidate sub-wheel layout on the device for
 * window sizes 2^16..2^34 (at the sample's median height) and keeps the fastest per
 * size, then picks the workgroup size, tile size and resident waves per CU that run
 * the sample batch (nwin windows [lo[i], hi[i]), may be 0) fastest. *report (free with
 * es_free, optional) receives the measurements. Output never depends on the choices. No
 * counterpart in the reference, whose launch shape is fixed by --seg/--block
 * (sieve/cuda/scan.cuh:9-14). */
int es_autotune(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin, char **report);

/* Layout the autotuner picks for a window of w integers. */
uint64_t es_choose_layout(const es_scanner *s, uint64_t w);

/* Timing of the last es_scan_* / es_survivors call. */
int es_last_timing(const es_scanner *s, es_timing *t);

/* Wheel facts: lanes (classes), wheel_pmin, filter count. Replaces ref3_wheel_size /
 * ref3_wheel_pmin (sieve/cuda/ref_host.h:32-34). */
int es_wheel_info(const es_scanner *s, uint64_t *lanes, uint64_t *wheel_pmin, uint64_t *nfilters,
                  uint64_t *ncov);

/* Exact work the marking kernel issues for [lo, hi) in layout Mp (0 = the autotuner's
 * pick): bit clears, bitmap bits, (tile, filter) setups and tiles. Host arithmetic
 * only. Replaces the exact `marks` count of sieve_cuda --profile
 * (sieve/cuda/scan.cu:150-162); it is the numerator of the marking roofline. */
int es_mark_count(es_scanner *s, uint64_t lo, uint64_t hi, uint64_t Mp, uint64_t *marks,
                  uint64_t *bits, uint64_t *pairs, uint64_t *tiles);

/* Measured LDS atomic-AND rate of the device (32-bit operations per second), the
 * denominator of the marking roofline. Counterpart of sieve_cuda --bench-mark
 * (sieve/cuda/mark.cu:197-263), which times the marking kernel itself. */
int es_lds_peak(es_scanner *s, double *ops_per_s);

/* Measured LDS rate of scattered byte stores (stores per second), at the marking kernel's
 * occupancy: the rate the marking primitive can reach when its addresses share banks at
 * random. Second denominator for the same --bench-mark comparison (mark.cu:197-263). */
int es_lds_scatter_peak(es_scanner *s, double *stores_per_s);

/* The wheel classes (ascending), for lane-order merging of window records. Writes
 * min(cap, lanes) values and returns ES_OK; *n receives the lane count. Replaces
 * ref3_wheel (sieve/cuda/ref_host.h:33). */
int es_wheel_classes(const es_scanner *s, uint64_t *out, size_t cap, size_t *n);

/* ---------------------------------------------------------------- serving
 * A continuous-batching request server over one scanner. The reference has no request
 * server; each request's answer is exactly rung_scan3's stdout for `rung_scan3 LO HI`
 * with the scanner's flags (rung_scan3.cpp:1128-1162), alone or in any mix. Admission is
 * FIFO, bounded by request slots and an HBM budget (hbm_fraction of device memory, minus
 * the scanner's resident tables); every device batch takes the next slice of each
 * admitted request, so requests that arrive while a batch runs join the next one. While
 * a server is open it owns the scanner: make no other call on that scanner until
 * es_server_close. */
typedef struct es_server es_server;

typedef struct es_serve_config {
  uint64_t slice;            /* integers per slice of a request (default 2^32) */
  uint64_t batch_ints;       /* integers per device batch (default 2^35, >= slice) */
  uint32_t batch_windows;    /* slices per device batch (default 4096) */
  uint32_t slots;            /* requests admitted at once (default 256) */
  double hbm_fraction;       /* admission budget / device memory (default 0.8) */
  uint64_t hbm_budget;       /* bytes; nonzero overrides hbm_fraction */
  int emit_survivors;        /* 0 = --no-emit-survivors (default 1) */
} es_serve_config;

typedef struct es_request_info {
  double arrival_ms, admit_ms, done_ms;   /* server clock: submitted, admitted, record done */
  uint64_t batches;                       /* device batches the request took part in */
  uint64_t reserved;                      /* HBM bytes it held while admitted */
  int exit_code;                          /* 0, 2 (survivors) or 1 (refused) */
} es_request_info;

typedef struct es_serve_stats {
  uint64_t submitted, completed, refused, batches, slices, ints;
  uint64_t budget, resident, peak_reserved;
  uint32_t max_active, max_batch_requests, max_queued;
  double busy_ms;            /* time inside device batches */
  double wall_ms;            /* es_serve_trace: first arrival to last completion */
} es_serve_stats;

void es_serve_config_default(es_serve_config *cfg);

/* Starts a server on s (uploads every layout the tuner can pick, sizes the budget). */
int es_server_open(es_scanner *s, const es_serve_config *cfg, es_server **out);

/* Submits [lo, hi), arriving now; *id identifies it. A refused interval
 * (rung_scan3.cpp:1057-1069) completes at once with exit code 1 and the refusal message
 * as its record. */
int es_server_submit(es_server *v, uint64_t lo, uint64_t hi, uint64_t *id);

/* Waits up to timeout_ms (negative: no limit) for request id. Returns ES_OK with the
 * record in *record (free with es_free; *len its length) and *info filled, 1 on timeout,
 * or a negative status. A finished request is handed out once. */
int es_server_wait(es_server *v, uint64_t id, double timeout_ms, char **record, size_t *len,
                   es_request_info *info);

int es_server_stats(es_server *v, es_serve_stats *st);

/* Finishes every submitted request, then stops the server. */
void es_server_close(es_server *v);

/* Replays a request trace through a server: request i is submitted at arrival_s[i]
 * seconds after the start (realtime = 1) or all at once (realtime = 0). Records are
 * concatenated in request order into *text (free with es_free), rec_len[i] and info[i]
 * per request, stats (optional) for the run. */
int es_serve_trace(es_scanner *s, const es_serve_config *cfg, const double *arrival_s, const uint64_t *lo,
                   const uint64_t *hi, size_t n, int realtime, char **text, size_t *rec_len,
                   es_request_info *info, es_serve_stats *stats);

/* Frees memory returned by this library. */
void es_free(void *p);

#ifdef __cplusplus
}
#endif

#endif /* ES_SIEVE_H */
