// es_capi.cpp -- the extern "C" boundary over es::Scanner (include/es_sieve.h).
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/es_sieve.h"
#include "es_scan.h"
#include "es_serve.h"
#include "es_tail.h"

#include <chrono>
#include <thread>

struct es_scanner {
  es::Scanner sc;
  es::ScanTiming last;
};

namespace {
thread_local std::string g_err;
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}

// The ES_MAX_IN_FLIGHT batch slots overlap on the device only if each slot's stream has a
// hardware queue of its own. HIP gives a process GPU_MAX_HW_QUEUES queues (four by default)
// and streams beyond that share them in order, so a batch waits behind the stage-D kernel of
// the batch four slots back (the 512-window batch eight deep: 0.125 ms per batch with four
// queues, 0.113 with eight or sixteen). The library raises the setting to the slot count
// when it is lower, when it is loaded: the HIP runtime reads it at its first call in the
// process, so a process that used HIP before loading the library keeps what it had.
// Priority 101 runs this before the constructors that register the device code.
__attribute__((constructor(101))) void raise_hw_queues() {
  const char *v = std::getenv("GPU_MAX_HW_QUEUES");
  if (!v || std::atoi(v) < ES_MAX_IN_FLIGHT) setenv("GPU_MAX_HW_QUEUES", std::to_string(ES_MAX_IN_FLIGHT).c_str(), 1);
}
}  // namespace

extern "C" {

void es_config_default(es_config *c) {
  if (!c) return;
  c->wheel_M = 2042040;
  c->mmax = 10000;
  c->sieve_P = 0;
  c->filter_file = nullptr;
  c->threads = 0;
  c->device = 0;
}

const char *es_last_error(void) { return g_err.c_str(); }

int es_scanner_open(const es_config *cfg, es_scanner **out) {
  if (!cfg || !out) return fail(ES_ERR_ARG, "null argument");
  *out = nullptr;
  es_scanner *s = new es_scanner;
  std::string msg;
  if (s->sc.init_wheel(cfg->wheel_M, cfg->mmax, msg) ||
      s->sc.init_filters(cfg->filter_file, cfg->sieve_P, msg)) {
    delete s;
    return fail(ES_ERR_CONFIG, msg);
  }
  s->sc.threads = cfg->threads;
  if (cfg->device >= 0 && s->sc.open_device(cfg->device, msg)) {
    delete s;
    return fail(ES_ERR_DEVICE, msg);
  }
  *out = s;
  return ES_OK;
}

void es_scanner_close(es_scanner *s) { delete s; }

int es_check_interval(const es_scanner *s, uint64_t lo, uint64_t hi) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  std::string msg;
  if (es::check_interval(s->sc.plan(), lo, hi, msg)) return fail(ES_ERR_CONFIG, msg);
  return ES_OK;
}

static int run_windows(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin,
                       std::vector<es::WindowRecord> &recs) {
  if (!s || (nwin && (!lo || !hi))) return fail(ES_ERR_ARG, "null argument");
  if (s->sc.in_flight()) return fail(ES_ERR_ARG, "submitted batches are in flight: collect them first");
  std::vector<es::Interval> w(nwin);
  std::string msg;
  for (size_t i = 0; i < nwin; ++i) {
    w[i] = es::Interval{lo[i], hi[i]};
    if (es::check_interval(s->sc.plan(), lo[i], hi[i], msg))
      return fail(ES_ERR_CONFIG, msg);
  }
  if (s->sc.scan(w, recs, s->last, msg)) return fail(ES_ERR_DEVICE, msg);
  return ES_OK;
}

// The records of one batch as one malloc'd text, with each record's length and exit code.
static int records_text(es_scanner *s, const std::vector<es::WindowRecord> &recs, int emit_survivors, char **text,
                        size_t *text_len, size_t *rec_len, int *exit_code) {
  const size_t nwin = recs.size();
  std::string all;
  size_t guess = 0;
  for (const es::WindowRecord &r : recs) guess += 480 + r.lines.size() * 64;
  all.reserve(guess);
  for (size_t i = 0; i < nwin; ++i) {
    const size_t before = all.size();
    s->sc.append_record(all, recs[i], emit_survivors != 0);
    if (rec_len) rec_len[i] = all.size() - before;
    if (exit_code) exit_code[i] = recs[i].survivors ? 2 : 0;
  }
  const size_t total = all.size();
  char *buf = (char *)malloc(total + 1);
  if (!buf) return fail(ES_ERR_ARG, "out of host memory");
  memcpy(buf, all.data(), total);
  buf[total] = 0;
  *text = buf;
  if (text_len) *text_len = total;
  return ES_OK;
}

int es_scan_records(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin, int emit_survivors,
                    char **text, size_t *text_len, size_t *rec_len, int *exit_code) {
  if (!text) return fail(ES_ERR_ARG, "null text");
  *text = nullptr;
  std::vector<es::WindowRecord> recs;
  const int rc = run_windows(s, lo, hi, nwin, recs);
  if (rc) return rc;
  return records_text(s, recs, emit_survivors, text, text_len, rec_len, exit_code);
}

static_assert(ES_MAX_IN_FLIGHT == es::DeviceCtx::kMaxInFlight, "header and device agree");

int es_scan_submit(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin) {
  if (!s || (nwin && (!lo || !hi))) return fail(ES_ERR_ARG, "null argument");
  if (s->sc.in_flight() >= ES_MAX_IN_FLIGHT) return fail(ES_ERR_ARG, "every batch slot is in flight: collect one first");
  std::vector<es::Interval> w(nwin);
  std::string msg;
  for (size_t i = 0; i < nwin; ++i) {
    w[i] = es::Interval{lo[i], hi[i]};
    if (es::check_interval(s->sc.plan(), lo[i], hi[i], msg)) return fail(ES_ERR_CONFIG, msg);
  }
  if (s->sc.submit(w, msg)) return fail(ES_ERR_DEVICE, msg);
  return ES_OK;
}

int es_scan_collect(es_scanner *s, int emit_survivors, char **text, size_t *text_len, size_t *rec_len,
                    int *exit_code, size_t *nwin) {
  if (!s || !text) return fail(ES_ERR_ARG, "null argument");
  *text = nullptr;
  if (s->sc.in_flight() == 0) return fail(ES_ERR_ARG, "no batch in flight");
  std::vector<es::WindowRecord> recs;
  std::string msg;
  if (s->sc.collect(recs, s->last, msg)) return fail(ES_ERR_DEVICE, msg);
  if (nwin) *nwin = recs.size();
  return records_text(s, recs, emit_survivors, text, text_len, rec_len, exit_code);
}

int es_scan_in_flight(const es_scanner *s) { return s ? (int)s->sc.in_flight() : 0; }

int es_scan_counts(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin, es_counts *out) {
  if (!out && nwin) return fail(ES_ERR_ARG, "null output");
  std::vector<es::WindowRecord> recs;
  const int rc = run_windows(s, lo, hi, nwin, recs);
  if (rc) return rc;
  for (size_t i = 0; i < nwin; ++i) {
    const es::WindowRecord &r = recs[i];
    out[i] = es_counts{r.positions, r.covered, r.sieved, r.mr, r.composite, r.rung, r.direct, r.survivors};
  }
  return ES_OK;
}

int es_survivors(es_scanner *s, uint64_t lo, uint64_t hi, uint64_t *p, uint8_t *is_prime, size_t cap,
                 size_t *n, es_counts *counts) {
  if (!s || !n) return fail(ES_ERR_ARG, "null argument");
  if (s->sc.in_flight()) return fail(ES_ERR_ARG, "submitted batches are in flight: collect them first");
  std::vector<es::Interval> w{es::Interval{lo, hi}};
  std::vector<es::u64> sp;
  std::vector<uint8_t> fl;
  std::vector<es::WindowRecord> recs;
  std::string msg;
  if (es::check_interval(s->sc.plan(), lo, hi, msg)) return fail(ES_ERR_CONFIG, msg);
  if (s->sc.survivors(w, recs, sp, fl, s->last, msg)) return fail(ES_ERR_DEVICE, msg);
  *n = sp.size();
  for (size_t i = 0; i < sp.size() && i < cap; ++i) {
    if (p) p[i] = sp[i];
    if (is_prime) is_prime[i] = fl[i];
  }
  if (counts) {
    const es::WindowRecord &r = recs[0];
    *counts = es_counts{r.positions, r.covered, r.sieved, r.mr, r.composite, r.rung, r.direct, r.survivors};
  }
  return ES_OK;
}

int es_primality(es_scanner *s, const uint64_t *vals, size_t n, uint8_t *flags) {
  if (!s || (n && (!vals || !flags))) return fail(ES_ERR_ARG, "null argument");
  if (s->sc.device().primality(vals, n, flags)) return fail(ES_ERR_DEVICE, s->sc.device().last_error());
  return ES_OK;
}

static_assert(ES_NOALLOC_MAX_VALUES == es::kNoAllocMaxValues, "header and device agree");

int es_primality_noalloc(es_scanner *s, const uint64_t *d_vals, size_t n, uint8_t *d_flags, void *stream) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  const int rc = s->sc.device().primality_noalloc(d_vals, n, d_flags, stream);
  if (rc == -2) return fail(ES_ERR_ARG, s->sc.device().last_error());
  if (rc) return fail(ES_ERR_DEVICE, s->sc.device().last_error());
  return ES_OK;
}

int es_certify(uint64_t p, char *buf, size_t cap) {
  std::string line;
  const int kind = es::certify_prime(p, line);
  if (!buf || line.size() + 1 > cap) return fail(ES_ERR_ARG, "buffer too small");
  memcpy(buf, line.c_str(), line.size() + 1);
  return kind;
}

int es_certify_batch(es_scanner *s, const uint64_t *p, size_t n, char **text, size_t *on_host) {
  if (!s || !text || (n && !p)) return fail(ES_ERR_ARG, "null argument");
  std::vector<es::RungResult> rr(n);
  if (s->sc.device().rung(p, n, rr.data())) return fail(ES_ERR_DEVICE, s->sc.device().last_error());
  std::string all, line;
  size_t host = 0;
  for (size_t i = 0; i < n; ++i) {
    if (rr[i].status == 1) {
      es::format_rung(p[i], rr[i].A, rr[i].r, ((es::u128)rr[i].u_hi << 64) | rr[i].u_lo, line);
    } else if (rr[i].status == 4) {
      es::format_direct(p[i], rr[i].r & 0xFFu, rr[i].A, rr[i].r >> 8, rr[i].u_lo, line);
    } else {
      ++host;
      es::certify_prime(p[i], line);
    }
    all += line;
  }
  *text = (char *)malloc(all.size() + 1);
  if (!*text) return fail(ES_ERR_ARG, "out of host memory");
  memcpy(*text, all.c_str(), all.size() + 1);
  if (on_host) *on_host = host;
  return ES_OK;
}

int es_direct_batch(es_scanner *s, const uint64_t *p, size_t n, char **text, size_t *on_host) {
  if (!s || !text || (n && !p)) return fail(ES_ERR_ARG, "null argument");
  std::vector<es::RungResult> rr(n);
  if (s->sc.device().rung(p, n, rr.data(), true)) return fail(ES_ERR_DEVICE, s->sc.device().last_error());
  std::string all, line;
  size_t host = 0;
  for (size_t i = 0; i < n; ++i) {
    if (rr[i].status == 4) {
      es::format_direct(p[i], rr[i].r & 0xFFu, rr[i].A, rr[i].r >> 8, rr[i].u_lo, line);
    } else if (rr[i].status == 2) {
      line = "SURVIVOR " + std::to_string(p[i]) + "\n";
    } else {
      ++host;
      if (!es::solve_direct(p[i], line)) line = "SURVIVOR " + std::to_string(p[i]) + "\n";
    }
    all += line;
  }
  *text = (char *)malloc(all.size() + 1);
  if (!*text) return fail(ES_ERR_ARG, "out of host memory");
  memcpy(*text, all.c_str(), all.size() + 1);
  if (on_host) *on_host = host;
  return ES_OK;
}

int es_set_device_tail(es_scanner *s, int on) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  s->sc.tuning().launch.device_tail = on != 0;
  return ES_OK;
}

int es_set_order(es_scanner *s, int order, uint64_t seg) {
  if (!s || (order != 0 && order != 1)) return fail(ES_ERR_ARG, "order must be 0 (lane) or 1 (round)");
  s->sc.order = order ? es::Order::Round : es::Order::Lane;
  if (seg) s->sc.round_seg = seg;
  return ES_OK;
}

int es_set_layout(es_scanner *s, uint64_t Mp) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  if (Mp && s->sc.plan().M % Mp) return fail(ES_ERR_ARG, "layout modulus must divide the wheel modulus");
  s->sc.tuning().fixed_Mp = Mp;
  return ES_OK;
}

int es_set_launch(es_scanner *s, int persistent, uint32_t block, uint32_t max_bits) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  if (block == 0 || block > 1024 || block % 64 || max_bits == 0 || max_bits % 32 || max_bits > es::kMaxTileBits)
    return fail(ES_ERR_ARG, "bad launch shape");
  es::LaunchCfg &c = s->sc.tuning().launch;
  c.persistent = persistent != 0;
  c.block = block;
  c.max_bits = max_bits;
  return ES_OK;
}

int es_set_residency(es_scanner *s, uint32_t waves_per_cu) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  if (waves_per_cu > 64) return fail(ES_ERR_ARG, "more resident waves per CU than a CU holds");
  s->sc.tuning().launch.waves_per_cu = waves_per_cu;
  return ES_OK;
}

int es_set_grouping(es_scanner *s, uint32_t max_tiles) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  if (max_tiles == 0 || max_tiles > 16) return fail(ES_ERR_ARG, "a work item groups 1 to 16 tiles");
  s->sc.tuning().launch.group = max_tiles;
  return ES_OK;
}

int es_set_graph(es_scanner *s, int on) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  s->sc.tuning().launch.graph = on != 0;
  return ES_OK;
}

int es_set_tuning(es_scanner *s, double c_pair, double c_mark, double c_bit) {
  if (!s || c_pair < 0 || c_mark < 0 || c_bit < 0) return fail(ES_ERR_ARG, "bad tuning");
  es::Tuning &t = s->sc.tuning();
  t.c_pair = c_pair;
  t.c_mark = c_mark;
  t.c_bit = c_bit;
  return ES_OK;
}

int es_autotune(es_scanner *s, const uint64_t *lo, const uint64_t *hi, size_t nwin, char **report) {
  if (!s || (nwin && (!lo || !hi))) return fail(ES_ERR_ARG, "null argument");
  if (s->sc.in_flight()) return fail(ES_ERR_ARG, "submitted batches are in flight: collect them first");
  std::vector<es::Interval> w(nwin);
  std::string msg, rep;
  for (size_t i = 0; i < nwin; ++i) {
    w[i] = es::Interval{lo[i], hi[i]};
    if (es::check_interval(s->sc.plan(), lo[i], hi[i], msg)) return fail(ES_ERR_CONFIG, msg);
  }
  if (s->sc.autotune(w, rep, msg)) return fail(ES_ERR_DEVICE, msg);
  if (report) {
    *report = (char *)malloc(rep.size() + 1);
    if (!*report) return fail(ES_ERR_ARG, "out of host memory");
    memcpy(*report, rep.c_str(), rep.size() + 1);
  }
  return ES_OK;
}

uint64_t es_choose_layout(const es_scanner *s, uint64_t w) {
  return s ? s->sc.choose_layout(w) : 0;
}

int es_last_timing(const es_scanner *s, es_timing *t) {
  if (!s || !t) return fail(ES_ERR_ARG, "null argument");
  const es::ScanTiming &m = s->last;
  *t = es_timing{m.plan_ms, m.batch_ms, m.device_ms, m.sieve_ms, m.tail_ms, m.total_ms,
                 m.tiles, m.survivors, m.primes, m.host_primes, m.replays, m.mr_ms, m.rung_ms,
                 m.sieve_event_ms, m.mr_event_ms, m.rung_event_ms};
  return ES_OK;
}

int es_wheel_info(const es_scanner *s, uint64_t *lanes, uint64_t *pmin, uint64_t *nf, uint64_t *ncov) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  const es::Plan &p = s->sc.plan();
  if (lanes) *lanes = p.wheel.classes.size();
  if (pmin) *pmin = p.wheel.max_pmin;
  if (nf) *nf = p.filters.mod.size();
  if (ncov) *ncov = p.filters.ncov;
  return ES_OK;
}

int es_mark_count(es_scanner *s, uint64_t lo, uint64_t hi, uint64_t Mp, uint64_t *marks, uint64_t *bits,
                  uint64_t *pairs, uint64_t *tiles) {
  if (!s) return fail(ES_ERR_ARG, "null scanner");
  if (Mp && s->sc.plan().M % Mp) return fail(ES_ERR_ARG, "layout modulus must divide the wheel modulus");
  es::u64 a, b, c, d;
  std::string msg;
  if (s->sc.mark_count(lo, hi, Mp, a, b, c, d, msg)) return fail(ES_ERR_DEVICE, msg);
  if (marks) *marks = a;
  if (bits) *bits = b;
  if (pairs) *pairs = c;
  if (tiles) *tiles = d;
  return ES_OK;
}

int es_lds_peak(es_scanner *s, double *ops_per_s) {
  if (!s || !ops_per_s) return fail(ES_ERR_ARG, "null argument");
  if (s->sc.device().lds_peak(*ops_per_s)) return fail(ES_ERR_DEVICE, s->sc.device().last_error());
  return ES_OK;
}

int es_lds_scatter_peak(es_scanner *s, double *stores_per_s) {
  if (!s || !stores_per_s) return fail(ES_ERR_ARG, "null argument");
  if (s->sc.device().lds_scatter_peak(*stores_per_s)) return fail(ES_ERR_DEVICE, s->sc.device().last_error());
  return ES_OK;
}

int es_wheel_classes(const es_scanner *s, uint64_t *out, size_t cap, size_t *n) {
  if (!s || !n) return fail(ES_ERR_ARG, "null argument");
  const std::vector<es::u64> &c = s->sc.plan().wheel.classes;
  *n = c.size();
  for (size_t i = 0; i < c.size() && i < cap && out; ++i) out[i] = c[i];
  return ES_OK;
}

void es_serve_config_default(es_serve_config *c) {
  if (!c) return;
  const es::ServeConfig d;
  c->slice = d.slice;
  c->batch_ints = d.batch_ints;
  c->batch_windows = d.batch_windows;
  c->slots = d.slots;
  c->hbm_fraction = d.hbm_fraction;
  c->hbm_budget = d.hbm_budget;
  c->emit_survivors = d.emit_survivors ? 1 : 0;
}

struct es_server {
  es::Server srv;
  es_server(es::Scanner &sc, const es::ServeConfig &cfg) : srv(sc, cfg) {}
};

static es::ServeConfig serve_config(const es_serve_config *c) {
  es::ServeConfig k;
  if (c) {
    k.slice = c->slice;
    k.batch_ints = c->batch_ints;
    k.batch_windows = c->batch_windows;
    k.slots = c->slots;
    k.hbm_fraction = c->hbm_fraction;
    k.hbm_budget = c->hbm_budget;
    k.emit_survivors = c->emit_survivors != 0;
  }
  return k;
}

static void fill_info(const es::RequestInfo &i, es_request_info *o) {
  if (!o) return;
  o->arrival_ms = i.arrival_ms;
  o->admit_ms = i.admit_ms;
  o->done_ms = i.done_ms;
  o->batches = i.batches;
  o->reserved = i.reserved;
  o->exit_code = i.exit_code;
}

static void fill_stats(const es::ServeStats &s, es_serve_stats *o) {
  if (!o) return;
  o->submitted = s.submitted;
  o->completed = s.completed;
  o->refused = s.refused;
  o->batches = s.batches;
  o->slices = s.slices;
  o->ints = s.ints;
  o->budget = s.budget;
  o->resident = s.resident;
  o->peak_reserved = s.peak_reserved;
  o->max_active = s.max_active;
  o->max_batch_requests = s.max_batch_requests;
  o->max_queued = s.max_queued;
  o->busy_ms = s.busy_ms;
  o->wall_ms = 0;
}

int es_server_open(es_scanner *s, const es_serve_config *cfg, es_server **out) {
  if (!s || !out) return fail(ES_ERR_ARG, "null argument");
  *out = nullptr;
  es_server *v = new es_server(s->sc, serve_config(cfg));
  std::string msg;
  if (v->srv.start(msg)) {
    delete v;
    return fail(ES_ERR_DEVICE, msg);
  }
  *out = v;
  return ES_OK;
}

int es_server_submit(es_server *v, uint64_t lo, uint64_t hi, uint64_t *id) {
  if (!v || !id) return fail(ES_ERR_ARG, "null argument");
  *id = v->srv.submit(lo, hi);
  return ES_OK;
}

int es_server_wait(es_server *v, uint64_t id, double timeout_ms, char **record, size_t *len,
                   es_request_info *info) {
  if (!v || !record) return fail(ES_ERR_ARG, "null argument");
  *record = nullptr;
  std::string rec;
  es::RequestInfo ri;
  const int rc = v->srv.wait(id, timeout_ms, rec, ri);
  if (rc == 1) return 1;
  if (rc < 0) {
    const es::ServeStats st = v->srv.stats();
    return fail(st.error ? ES_ERR_DEVICE : ES_ERR_ARG, st.error ? st.error_msg : "unknown request id");
  }
  *record = (char *)malloc(rec.size() + 1);
  if (!*record) return fail(ES_ERR_ARG, "out of host memory");
  memcpy(*record, rec.data(), rec.size());
  (*record)[rec.size()] = 0;
  if (len) *len = rec.size();
  fill_info(ri, info);
  return ES_OK;
}

int es_server_stats(es_server *v, es_serve_stats *st) {
  if (!v || !st) return fail(ES_ERR_ARG, "null argument");
  fill_stats(v->srv.stats(), st);
  return ES_OK;
}

void es_server_close(es_server *v) { delete v; }

int es_serve_trace(es_scanner *s, const es_serve_config *cfg, const double *arrival_s, const uint64_t *lo,
                   const uint64_t *hi, size_t n, int realtime, char **text, size_t *rec_len,
                   es_request_info *info, es_serve_stats *stats) {
  if (!s || !text || (n && (!lo || !hi || (realtime && !arrival_s)))) return fail(ES_ERR_ARG, "null argument");
  *text = nullptr;
  es::Server srv(s->sc, serve_config(cfg));
  std::string msg;
  if (srv.start(msg)) return fail(ES_ERR_DEVICE, msg);
  std::vector<uint64_t> ids(n);
  const double t_start = srv.now_ms();
  for (size_t i = 0; i < n; ++i) {
    if (realtime) {
      const double due = t_start + arrival_s[i] * 1e3;
      const double wait = due - srv.now_ms();
      if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double, std::milli>(wait));
    }
    ids[i] = srv.submit(lo[i], hi[i]);
  }
  std::string all;
  double first = 1e300, last = 0;
  for (size_t i = 0; i < n; ++i) {
    std::string rec;
    es::RequestInfo ri;
    if (srv.wait(ids[i], -1, rec, ri) != 0) {
      const es::ServeStats st = srv.stats();
      srv.stop();
      return fail(ES_ERR_DEVICE, st.error ? st.error_msg : "request lost");
    }
    if (rec_len) rec_len[i] = rec.size();
    all += rec;
    fill_info(ri, info ? info + i : nullptr);
    first = std::min(first, ri.arrival_ms);
    last = std::max(last, ri.done_ms);
  }
  srv.stop();
  if (stats) {
    fill_stats(srv.stats(), stats);
    stats->wall_ms = n ? last - first : 0;
  }
  *text = (char *)malloc(all.size() + 1);
  if (!*text) return fail(ES_ERR_ARG, "out of host memory");
  memcpy(*text, all.data(), all.size());
  (*text)[all.size()] = 0;
  return ES_OK;
}

void es_free(void *p) { free(p); }

}  // extern "C"
