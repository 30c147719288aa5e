#!/usr/bin/env python3
"""Benchmark of the verification-sieve funnel on one MI355X.

Headline workload (one "step"): the batched configuration -- 512 windows, seed 0, start
log-uniform over [1e9, 1.8e19], length log-uniform over [2^16, 2^26] integers
(tests/golden/c4_windows.txt) -- scanned end to end into one record per window. Each
window's record must equal the golden record for the wheel in use
(tests/golden/c4/records_w<M>.txt) at every step. For each wheel (120120 with its runtime
table, 2042040 with the published GPU table) the autotuner chooses the sub-wheel layout
per window size and the launch shape on the batch itself; the faster wheel is the one
reported. Inputs (tables, layouts) are resident in HBM before the timed region.

Forms timed on the same batch:
  batched   all 512 windows in one persistent work-queue launch replayed from a hipGraph,
            up to eight batches in flight (one stream each): the headline `value`
  one at a time   the same, nothing overlapped: the per-batch latency p50/p99
  plain     one launch (and one host round trip) per window

Metric: verified integers/s = sum of window lengths / wall seconds over the timed steps.

Also reported, in the same JSON line:
  roofline     the marking kernel (k_sieve) on the headline batch against the LDS
               roofline (exact marks from the layout tables, kernel time from HIP events on
               its dispatch), the stage-D kernel (k_rung) as VALU issue against its peak,
               and the compaction's HBM bytes
  legs         the plain configurations: 1e8 integers from 255255 (wheel 120120, 96 windows
               of 2^20 in the plain form), 2^34 integers from 1e15 (4096 windows of 2^22 in
               the plain form) and the top 2^34 integers below the stage-E ceiling, each one
               at a time and eight in flight; and the MI355X-tuned filter table
  c5           the serving trace (Poisson 200 req/s, seed 1) replayed in real time through
               the continuous-batching server, every request's record checked against its
               committed sha256; then the whole trace at 10x, 40x and 80x its arrival rate, the
               server near saturation
  cpu_baseline the reference scanner built from its own sources (oracle/_ref/rung_scan3)
               on a bounded sample of the batch's windows, on this host's cores, startup
               and work separated; cpu_baseline_port: the oracle over all 512 windows

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-legs]
                       [--no-tuned] [--c5-seconds S] [--calib-copy MIB] [--only WORKLOAD]
"""
import argparse
import hashlib
import importlib.util
import json
import os
import platform
import subprocess
import sys
import tempfile
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "erdos-strauss_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")
TOP = 18446744065119614976                     # the stage-E ceiling (DIRECT_PMAX)
NOMINAL_LDS_BPS = 256 * 128 * 2.4e9            # 256 CUs x 128 B/clk x 2.4 GHz = 78.6 TB/s
# Wave64 VALU instructions per second: the published FP32 vector rate (157.3 TFLOP/s) over
# 64 lanes x 2 flops of an FMA, i.e. two wave-instructions per CU per clock at 2.4 GHz.
VALU_ISSUE_PEAK = 157.3e12 / 128
es_mod = None                                  # the binding (erdos-strauss_amd/es_sieve.py), loaded by main()


def load(name):
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def pct(vals, q):
    v = sorted(vals)
    k = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[k]


def med(vals):
    return sorted(vals)[len(vals) // 2]


def split_records(text):
    out, cur = [], []
    for line in text.splitlines(keepends=True):
        cur.append(line)
        if line.startswith("SUMMARY "):
            out.append("".join(cur))
            cur = []
    return out


def c4_windows():
    return [tuple(map(int, l.split())) for l in open(os.path.join(GOLDEN, "c4_windows.txt"))
            if not l.startswith("#")]


def profile_value(name, key):
    """A figure from a committed rocprofv3 summary under profiles/ (None if absent)."""
    path = os.path.join(REPO, "profiles", name)
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1].split()[0])
    return None


def tuned_launch(report):
    """The autotuner's launch choice: {"block", "tile_bits", "waves_per_cu", "group"}."""
    out = {}
    for l in report.splitlines():
        if l.startswith("choice block="):
            out = {k: int(v) for k, v in (x.split("=") for x in l.split()[1:])}
    out["tile_bits"] = out.pop("tile_bits", 32768)
    return out


def blob(records):
    """Records as the library returns a batch: one byte string and each window's length."""
    return "".join(records).encode("ascii"), [len(r) for r in records]


def pipelined(sc, batch, steps, want=None, depth=None):
    """`steps` passes over one batch with up to `depth` in flight (pass i+depth-1 is
    submitted before pass i is collected): the wall time and whether every pass's records
    equal `want` (when given), compared byte for byte, window by window."""
    ok = True
    depth = depth or es_mod.MAX_IN_FLIGHT
    b = es_mod.Batch(batch)
    wraw, wlens = blob(want) if want is not None else (None, None)
    t0 = time.perf_counter()
    for _ in range(min(depth - 1, steps)):
        sc.submit(b)
    for i in range(steps):
        if i + depth - 1 < steps:
            sc.submit(b)
        raw, lens, _ = sc.collect_raw()
        if want is not None:
            ok = ok and raw == wraw and lens == wlens
    return time.perf_counter() - t0, ok


def one_at_a_time(sc, batch, steps, want):
    lat, ok = [], True
    b = es_mod.Batch(batch)
    wraw, wlens = blob(want)
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        raw, lens, _ = sc.records_raw(b)
        lat.append((time.perf_counter() - s) * 1e3)
        ok = ok and raw == wraw and lens == wlens
    return time.perf_counter() - t0, lat, ok


def kernel_split(sc, batch, want, reps=5):
    """Kernel times of one batch: from the kernels' own device-clock stamps over `reps`
    graph-replayed passes, one at a time (the headline form), and from HIP events over
    `reps` eager passes (each eager dispatch that carries events also waits out a marker
    before it, which the events include)."""
    rows, ok = [], True
    for _ in range(reps):
        recs, _ = sc.records(batch)
        ok = ok and recs == want
        rows.append(sc.timing())
    sc.set_graph(False)
    erows = []
    for _ in range(reps):
        recs, _ = sc.records(batch)
        ok = ok and recs == want
        erows.append(sc.timing())
    sc.set_graph(True)
    keys = ("device_ms", "sieve_ms", "mr_ms", "rung_ms", "tail_ms", "plan_ms")
    out = {k: med([r[k] for r in rows]) for k in keys}
    out["eager"] = {k: med([r[k] for r in erows])
                    for k in ("sieve_ms", "mr_ms", "rung_ms", "sieve_event_ms", "mr_event_ms", "rung_event_ms")}
    out.update(tiles=rows[-1]["tiles"], survivors=rows[-1]["survivors"], primes=rows[-1]["primes"],
               host_primes=rows[-1]["host_primes"])
    return out, ok


def marking_roofline(sc, batch, sieve_ms, layouts=None, wheel=None):
    """Algorithmic LDS bytes of k_sieve over a batch (8 B per 32-bit read-modify-write mark,
    8 B per 32 bitmap positions for the init and the count) against the nominal LDS rate.
    `marks` counts the marks the chosen layouts execute; with `wheel`, `marks_lane` counts
    those of the full-wheel layout (one lane per wheel class), the layout-independent useful
    marks: a sub-wheel layout also marks positions its periodic wheel mask has cleared."""
    marks = bits = pairs = tiles = lane_marks = 0
    for lo, hi in batch:
        w = sc.mark_count(lo, hi, 0)
        marks += w["marks"]
        bits += w["bits"]
        pairs += w["pairs"]
        tiles += w["tiles"]
        if wheel:
            lane_marks += sc.mark_count(lo, hi, wheel)["marks"]
    lds_bytes = 8 * marks + 8 * (bits // 32)
    achieved = lds_bytes / (sieve_ms * 1e-3) / 1e9
    return {"bound": "lds", "kernel": "k_sieve", "achieved": achieved, "peak": NOMINAL_LDS_BPS / 1e9,
            "unit": "GB/s", "frac": achieved / (NOMINAL_LDS_BPS / 1e9), "algorithmic_bytes": lds_bytes,
            "marks": marks, "bits": bits, "pairs": pairs, "tiles": tiles, "kernel_ms": sieve_ms,
            "kernel_ms_from": "the kernel's own device-clock stamps (first workgroup's start to last "
                              "workgroup's end), median of 5 graph replays one at a time",
            "marks_per_s": marks / (sieve_ms * 1e-3),
            **({"marks_lane": lane_marks, "marks_lane_over_marks": lane_marks / marks} if wheel and marks else {})}


def c4_wheel(es, M, F, steps, warmup):
    wins = c4_windows()
    total = sum(b - a for a, b in wins)
    gold = split_records(open(os.path.join(GOLDEN, "c4", "records_w%d.txt" % M)).read())
    sc = es.Scanner(wheel=M, filters=F, device=0)
    report = sc.autotune(wins)
    launch = tuned_launch(report)
    sc.set_graph(True)
    ok = True
    for _ in range(warmup):
        recs, _ = sc.records(wins)
        ok = ok and recs == gold
    seq_wall, lat, ok1 = one_at_a_time(sc, wins, steps, gold)
    pipelined(sc, wins, es.MAX_IN_FLIGHT, gold)                 # every slot's graph captured
    wall, ok2 = pipelined(sc, wins, steps, gold)
    split, ok3 = kernel_split(sc, wins, gold)
    roof = marking_roofline(sc, wins, split["sieve_ms"], wheel=M)
    roof["kernel_ms_hip_events_eager"] = split["eager"]["sieve_event_ms"]
    sc.set_graph(False)
    sc.set_launch(False, launch["block"], launch["tile_bits"])
    t = time.perf_counter()
    precs = [sc.records([w])[0][0] for w in wins]
    plain_s = time.perf_counter() - t
    ok_plain = precs == gold
    layouts = sorted({sc.choose_layout(hi - lo) for lo, hi in wins})
    sc.close()
    return {"value": total * steps / wall, "ms_per_step": wall * 1e3 / steps,
            "one_at_a_time": {"value": total * steps / seq_wall, "ms_per_step": seq_wall * 1e3 / steps},
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
            "plain": {"value": total / plain_s, "windows": len(wins), "s_per_pass": plain_s,
                      "record_identical": ok_plain},
            "record_identical": ok and ok1 and ok2 and ok3, "split_ms": split, "roofline": roof,
            "layouts_Mp": layouts, "launch": launch, "total_integers": total}


def plain_leg(es, records_mod, name, M, F, lo, hi, steps, plain_window):
    """One plain configuration: the whole interval as one batched step (one at a time and
    eight in flight, graph replayed) and, in the plain form, windows of plain_window
    integers one launch each, merged in lane order. Records compared every step."""
    gold = open(os.path.join(GOLDEN, "records", name + ".txt")).read()
    sc = es.Scanner(wheel=M, filters=F, device=0)
    report = sc.autotune([(lo, hi)])
    launch = tuned_launch(report)
    sc.set_graph(True)
    batch = [(lo, hi)]
    want = [gold]
    for _ in range(3):
        if sc.records(batch)[0] != want:
            raise SystemExit("%s: batched record differs from the golden record" % name)
    seq_wall, lat, ok1 = one_at_a_time(sc, batch, steps, want)
    pipelined(sc, batch, es.MAX_IN_FLIGHT, want)
    wall, ok2 = pipelined(sc, batch, steps, want)
    split, ok3 = kernel_split(sc, batch, want)
    roof = marking_roofline(sc, batch, split["sieve_ms"])
    roof["kernel_ms_hip_events_eager"] = split["eager"]["sieve_event_ms"]
    out = {"interval": [lo, hi], "wheel": M, "table": os.path.relpath(F, REPO),
           "value": (hi - lo) * steps / wall, "ms_per_step": wall * 1e3 / steps,
           "one_at_a_time": {"value": (hi - lo) * steps / seq_wall, "ms_per_step": seq_wall * 1e3 / steps},
           "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
           "split_ms": split, "roofline": roof, "launch": launch,
           "layout_Mp": sc.choose_layout(hi - lo), "record_identical": ok1 and ok2 and ok3}
    if plain_window:
        sc.set_graph(False)
        sc.set_launch(False, launch["block"], launch["tile_bits"])
        wins = [(a, min(hi, a + plain_window)) for a in range(lo, hi, plain_window)]
        sc.records([wins[0]])
        t = time.perf_counter()
        recs = [sc.records([w])[0][0] for w in wins]
        plain_s = time.perf_counter() - t
        ok_plain = records_mod.merge(recs, sc.wheel_classes(), lo, hi) == gold
        out["plain"] = {"value": (hi - lo) / plain_s, "window": plain_window, "windows": len(wins),
                        "s_per_pass": plain_s, "record_identical": ok_plain}
        out["batched_over_plain"] = out["value"] / out["plain"]["value"]
    sc.close()
    return out


def tuned_table_bench(es, steps):
    """The MI355X-tuned filter table (tables/filters/f_M2042040_mmax2000_mi355x.txt, made by
    bin/es_prune) beside the reference's GPU table on 32 consecutive 2^34 windows from
    1e15 (eight in flight, the first window's record compared with the reference scanner's
    record for that table), and the 1e15 window alone."""
    lo, hi = 10 ** 15, 10 ** 15 + 2 ** 34
    gold = open(os.path.join(GOLDEN, "records", "c2_mi355x.txt")).read()
    wins = [(lo + i * (hi - lo), lo + (i + 1) * (hi - lo)) for i in range(32)]
    out = {"table": os.path.relpath(es.F204_MI355X, REPO)}
    for line in open(es.F204_MI355X):
        if line.startswith("# sha256="):
            out["sha256"] = line.split("=", 1)[1].strip()
        elif line.startswith("# filters="):
            out["table_stats"] = dict(kv.split("=") for kv in line[2:].split())
    for name, F in (("mi355x", es.F204_MI355X), ("gpu", es.F204)):
        sc = es.Scanner(wheel=2042040, filters=F, device=0)
        sc.autotune([(lo, hi)])
        sc.set_graph(True)
        r = {}
        if name == "mi355x":
            pipelined(sc, [(lo, hi)], es.MAX_IN_FLIGHT)
            wall, ok = pipelined(sc, [(lo, hi)], steps, [gold])
            r["c2"] = {"value": (hi - lo) * steps / wall, "ms_per_step": wall * 1e3 / steps, "record_identical": ok}
        recs, _ = sc.records(wins)
        first_ok = name != "mi355x" or recs[0] == gold
        pipelined(sc, wins, es.MAX_IN_FLIGHT)
        wall, ok = pipelined(sc, wins, 6, recs)
        wall /= 6
        r["batch32"] = {"value": (wins[-1][1] - wins[0][0]) / wall, "ms_per_batch": wall * 1e3,
                        "first_record_identical": first_ok and ok}
        out[name] = r
        sc.close()
    out["batch32_over_gpu_table"] = out["mi355x"]["batch32"]["value"] / out["gpu"]["batch32"]["value"]
    return out


def c5_compressed(es, factor):
    """The whole trace with its arrival times divided by `factor` (factor x the offered
    load): the server near or past saturation. Every record is checked against its digest."""
    rows = [l.split() for l in open(os.path.join(GOLDEN, "c5_trace.txt")) if not l.startswith("#")]
    digests = [l.split()[0] for l in open(os.path.join(GOLDEN, "c5", "records_sha256.txt")) if not l.startswith("#")]
    reqs = [(float(a) / factor, int(lo), int(hi)) for a, lo, hi in rows]
    sc = es.Scanner(wheel=2042040, filters=es.F204, device=0)
    sc.autotune([(lo, min(hi, lo + (1 << 32))) for _, lo, hi in reqs[:64]])
    recs, infos, st = es.serve_trace(sc, reqs, realtime=True, hbm_fraction=0.8)
    sc.close()
    bad = sum(1 for i, r in enumerate(recs) if hashlib.sha256(r.encode()).hexdigest() != digests[i])
    lat = [i["done_ms"] - i["arrival_ms"] for i in infos]
    total = sum(hi - lo for _, lo, hi in reqs)
    return {"arrival_rate_req_s": 200.0 * factor, "offered": total / (reqs[-1][0] or 1e-9),
            "sustained": total / (st["wall_ms"] * 1e-3), "requests": len(recs), "records_identical": bad == 0,
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat)},
            "device_busy_frac": st["busy_ms"] / st["wall_ms"], "max_active": st["max_active"],
            "max_batch_requests": st["max_batch_requests"], "batches": st["batches"], "wall_ms": st["wall_ms"]}


def c5_bench(es, seconds):
    """The serving trace replayed in real time for its first `seconds` through the
    continuous-batching server (wheel 2042040, the GPU filter table, admission budget 80%
    of device memory). Latency is arrival to record complete. Every request's record is
    hashed and compared with the committed digest of the oracle's record."""
    rows = [l.split() for l in open(os.path.join(GOLDEN, "c5_trace.txt")) if not l.startswith("#")]
    digests = [l.split()[0] for l in open(os.path.join(GOLDEN, "c5", "records_sha256.txt")) if not l.startswith("#")]
    reqs = [(float(a), int(lo), int(hi)) for a, lo, hi in rows if float(a) < seconds]
    sc = es.Scanner(wheel=2042040, filters=es.F204, device=0)
    sc.autotune([(lo, min(hi, lo + (1 << 32))) for _, lo, hi in reqs[:64]])
    recs, infos, st = es.serve_trace(sc, reqs, realtime=True, hbm_fraction=0.8)
    sc.close()
    bad = [i for i, r in enumerate(recs) if hashlib.sha256(r.encode()).hexdigest() != digests[i]]
    lat = [i["done_ms"] - i["arrival_ms"] for i in infos]
    total = sum(hi - lo for _, lo, hi in reqs)
    return {"workload": "Poisson 200 req/s (seed 1), %.0f s of the trace: %d requests, %d integers"
                        % (seconds, len(reqs), total),
            "sustained": total / (st["wall_ms"] * 1e-3), "unit": "integers/s", "offered": total / seconds,
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat)},
            "requests": len(recs), "records_identical": not bad, "records_differing": len(bad),
            "batches": st["batches"], "busy_ms": st["busy_ms"], "wall_ms": st["wall_ms"],
            "device_busy_frac": st["busy_ms"] / st["wall_ms"], "max_active": st["max_active"],
            "max_batch_requests": st["max_batch_requests"], "hbm_budget_gb": st["budget"] / 1e9,
            "resident_gb": st["resident"] / 1e9, "peak_reserved_gb": st["peak_reserved"] / 1e9}


def host_info():
    model, smt = platform.processor() or "?", None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        smt = open("/sys/devices/system/cpu/smt/active").read().strip() == "1"
    except OSError:
        pass
    return model, smt


def cpu_baseline(cores, every=2, hi_cap=10 ** 13):
    """The reference scanner (oracle/_ref/rung_scan3, built from the reference's own
    sources) on every `every`-th batched window with hi < hi_cap, one process per window as
    rung_scan3 runs, --threads cores. Its startup (a base-prime sieve to sqrt(hi), which
    has no effect on output) is timed by the same command on the empty interval [lo, lo)
    and subtracted for the work-only rate; both rates are reported."""
    ref = os.path.join(REPO, "oracle", "_ref", "rung_scan3")
    if not os.path.exists(ref):
        return None
    F = os.path.join(PKG, "tables", "filters", "f_M2042040_mmax2000_gpu.txt")
    gold = split_records(open(os.path.join(GOLDEN, "c4", "records_w2042040.txt")).read())
    wins = c4_windows()
    sample = [(i, w) for i, w in enumerate(wins) if w[1] < hi_cap][::every]
    full = empty = 0.0
    ints = 0
    ok = True
    for i, (lo, hi) in sample:
        argv = ["--wheel", "2042040", "--filters", F, "--threads", str(cores)]
        t = time.perf_counter()
        out = subprocess.run([ref, str(lo), str(hi)] + argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL).stdout
        full += time.perf_counter() - t
        t = time.perf_counter()
        subprocess.run([ref, str(lo), str(lo)] + argv, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        empty += time.perf_counter() - t
        ints += hi - lo
        ok = ok and out.decode() == gold[i]
    model, smt = host_info()
    return {"value": ints / max(full - empty, 1e-9), "unit": "integers/s", "cores": cores, "kind": "reference",
            "sample": "%d of the 512 batched windows (every %d-th with hi < %.0e), %d integers, wheel 2042040, "
                      "GPU table; one rung_scan3 process per window; value = work only (startup of each process "
                      "timed on [lo, lo) and subtracted)" % (len(sample), every, hi_cap, ints),
            "value_wall_incl_startup": ints / full, "wall_s": full, "startup_s": empty, "work_s": full - empty,
            "cpu_model": model, "smt": smt, "records_identical": ok}


def cpu_baseline_port(cores):
    """The oracle (the build's CPU restatement of rung_scan3, oracle/es_oracle) over all 512
    batched windows in one process (--batch: tables built once), --threads cores."""
    exe = os.path.join(REPO, "oracle", "es_oracle")
    if not os.path.exists(exe):
        return None
    F = os.path.join(PKG, "tables", "filters", "f_M2042040_mmax2000_gpu.txt")
    wins = c4_windows()
    gold = open(os.path.join(GOLDEN, "c4", "records_w2042040.txt"), "rb").read()
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        for lo, hi in wins:
            f.write("%d %d\n" % (lo, hi))
        path = f.name
    with tempfile.NamedTemporaryFile("w", suffix=".txt", delete=False) as f:
        empty_path = f.name
    argv = ["--wheel", "2042040", "--filters", F, "--threads", str(cores)]
    t = time.perf_counter()
    out = subprocess.run([exe, "--batch", path] + argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL).stdout
    wall = time.perf_counter() - t
    t = time.perf_counter()
    subprocess.run([exe, "--batch", empty_path] + argv, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
    startup = time.perf_counter() - t
    os.unlink(path)
    os.unlink(empty_path)
    total = sum(b - a for a, b in wins)
    model, smt = host_info()
    return {"value": total / max(wall - startup, 1e-9), "unit": "integers/s", "cores": cores, "kind": "port",
            "sample": "all 512 batched windows in one oracle process (tables built once)",
            "wall_s": wall, "startup_s": startup, "cpu_model": model, "smt": smt, "records_identical": out == gold}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-legs", action="store_true", help="skip the plain configurations")
    ap.add_argument("--no-tuned", action="store_true", help="skip the MI355X-tuned filter table leg")
    ap.add_argument("--c5-seconds", type=float, default=60.0,
                    help="seconds of the serving trace to replay in real time (0 = skip)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device to device (a kernel of known size for scaling HBM counters)")
    ap.add_argument("--only", choices=("c4w2042040", "c4w120120", "c1", "c2", "c3"),
                    help="run one workload alone and print its figures (for profiling runs), not the bench line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the sieve shards into independent intervals with no "
              "exchange step, and only the one-GPU configuration is built here.", file=sys.stderr)
        sys.exit(2)
    if args.calib_copy:
        import torch
        src = torch.ones(args.calib_copy << 18, dtype=torch.int32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    global es_mod
    es = es_mod = load("es_sieve")
    records_mod = load("records")
    if args.only:
        legs = {"c4w2042040": lambda: c4_wheel(es, 2042040, es.F204, args.steps, args.warmup),
                "c4w120120": lambda: c4_wheel(es, 120120, es.F120, args.steps, args.warmup),
                "c1": lambda: plain_leg(es, records_mod, "c1", 120120, es.F120, 255255, 255255 + 10 ** 8, args.steps, 0),
                "c2": lambda: plain_leg(es, records_mod, "c2", 2042040, es.F204, 10 ** 15, 10 ** 15 + 2 ** 34,
                                        args.steps, 0),
                "c3": lambda: plain_leg(es, records_mod, "c3", 2042040, es.F204, TOP - 2 ** 34, TOP, args.steps, 0)}
        print(json.dumps({"only": args.only, "figures": legs[args.only]()}))
        return

    by_wheel = {M: c4_wheel(es, M, F, args.steps, args.warmup)
                for M, F in ((2042040, es.F204), (120120, es.F120))}
    wheel = max(by_wheel, key=lambda m: by_wheel[m]["value"])
    h = by_wheel[wheel]
    table = "f_M2042040_mmax2000_gpu.txt" if wheel == 2042040 else "f_M120120_mmax2000_runtime.txt"
    roof = dict(h["roofline"])
    traffic = profile_value("k_sieve_c4_hbm_traffic.txt", "traffic_bytes_per_launch_w%d" % wheel)
    roof["traffic"] = int(traffic) if traffic is not None else None
    sc = es.Scanner(wheel=wheel, filters=es.F204 if wheel == 2042040 else es.F120, device=0)
    lds_ops, scatter = sc.lds_peak(), sc.lds_scatter_peak()
    sc.close()
    roof["peak_measured_lds_and"] = lds_ops * 8 / 1e9
    roof["frac_of_measured"] = roof["achieved"] / roof["peak_measured_lds_and"]
    roof["scattered_byte_stores_per_s"] = scatter
    roof["frac_of_scattered_store_rate"] = roof["marks_per_s"] / scatter
    valu = profile_value("k_sieve_rung_sq_counters.txt", "valu_wave_insts_per_dispatch_w%d" % wheel)
    rung_ms = h["split_ms"]["rung_ms"]
    roof["k_rung"] = {"bound": "valu", "kernel": "k_rung (stage D)", "unit": "wave-instructions/s",
                      "peak": VALU_ISSUE_PEAK,
                      "peak_from": "FP32 vector 157.3 TFLOP/s / (64 lanes x 2 flops)", "kernel_ms": rung_ms, "valu_wave_insts_per_dispatch": valu,
                      "achieved": valu / (rung_ms * 1e-3) if valu else None,
                      "frac": valu / (rung_ms * 1e-3) / VALU_ISSUE_PEAK if valu else None}
    surv, tiles = h["split_ms"]["survivors"], h["split_ms"]["tiles"]
    comp_bytes = 12 * surv + 16 * tiles          # survivor value + window index; tile counters
    roof["compaction"] = {"bound": "hbm", "unit": "GB/s", "hbm_bytes": comp_bytes, "peak": 8000.0,
                          "achieved": comp_bytes / (h["split_ms"]["sieve_ms"] * 1e-3) / 1e9,
                          "note": "bytes k_sieve writes to HBM for the survivor lists and tile counters, "
                                  "over the k_sieve time"}
    out = {
        "metric": "verified integers/s on 1 MI355X, batched vs plain, bit-exact; p50/p99 latency",
        "value": h["value"],
        "unit": "integers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: the committed list of 512 intervals (numpy PCG64 seed 0); no dataset",
        "config": {"workload": "batched: 512 windows, seed 0, start log-uniform over [1e9, 1.8e19], length "
                               "log-uniform over [2^16, 2^26] (%d integers)" % h["total_integers"],
                   "wheel": wheel, "table": table,
                   "form": "all 512 windows in one persistent work-queue launch replayed from a hipGraph, "
                           "layout per window size and launch shape from the autotuner, up to eight batches in "
                           "flight (one stream each)",
                   "layouts_Mp": h["layouts_Mp"], "launch": h["launch"]},
        "record_identical": h["record_identical"],
        "latency_ms": h["latency_ms"],
        "one_at_a_time": h["one_at_a_time"],
        "plain": h["plain"],
        "batched_over_plain": h["value"] / h["plain"]["value"],
        "split_ms": h["split_ms"],
        "roofline": roof,
        "by_wheel": {str(m): {k: v for k, v in r.items() if k != "roofline"} for m, r in by_wheel.items()},
    }
    if not args.no_legs:
        out["legs"] = {
            "c1": plain_leg(es, records_mod, "c1", 120120, es.F120, 255255, 255255 + 10 ** 8, args.steps, 2 ** 20),
            "c2": plain_leg(es, records_mod, "c2", 2042040, es.F204, 10 ** 15, 10 ** 15 + 2 ** 34, args.steps, 2 ** 22),
            "c3": plain_leg(es, records_mod, "c3", 2042040, es.F204, TOP - 2 ** 34, TOP, args.steps, 0),
        }
        if not args.no_tuned:
            out["legs"]["tuned_table"] = tuned_table_bench(es, args.steps)
    if args.c5_seconds > 0:
        out["c5"] = c5_bench(es, args.c5_seconds)
        out["c5"]["compressed"] = [c5_compressed(es, f) for f in (10.0, 40.0, 80.0)]
    if not args.no_cpu:
        cores = min(os.cpu_count() or 1, 16)
        out["cpu_baseline"] = cpu_baseline(cores)
        out["cpu_baseline_port"] = cpu_baseline_port(cores)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
