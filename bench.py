#!/usr/bin/env python3
"""Benchmark of the verification-sieve funnel on one MI355X.

Workload (one "step"): the plain configuration at height 10^15 -- every integer of
[10^15, 10^15 + 2^34), wheel 2042040, the published GPU filter table
f_M2042040_mmax2000_gpu.txt -- scanned end to end into its record (certificate lines in
lane order + SUMMARY), which must be byte-identical to the reference scanner's record
(tests/golden/records/c2.txt). Inputs (tables, layouts) are resident in HBM before the
timed region; the record is produced inside it.

Two forms are timed:
  batched  the interval cut into tiles (layout and launch shape from the autotuner) and
           run as one persistent work-queue launch replayed from a hipGraph (the
           headline `value`)
  plain    4096 windows of 2^22 integers, one launch (and one host round trip) per
           window, records merged by the lane key

Metric: verified integers/s = (hi - lo) / wall seconds over the timed steps.
Also reported: the marking kernel against the LDS roofline (exact marks from the
layout tables, kernel time from HIP events on its own stream), and the reference
scanner built from source (oracle/_ref/rung_scan3) timed on this host's cores.

Two more configurations ride along (skippable): the batched 512-window set (--no-c4) and
a real-time replay of the serving trace (--c5-seconds 0 to skip).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--plain-steps P] [--no-cpu]
                       [--no-c4] [--c5-seconds S]
"""
import argparse
import importlib.util
import json
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "erdos-strauss_amd")


def load(name):
    spec = importlib.util.spec_from_file_location(name, os.path.join(PKG, name + ".py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


LO, HI = 10 ** 15, 10 ** 15 + 2 ** 34
WINDOW = 2 ** 22
NOMINAL_LDS_BPS = 256 * 128 * 2.4e9        # 256 CUs x 128 B/clk x 2.4 GHz = 78.6 TB/s


def profiled_traffic():
    """HBM bytes per k_sieve launch on this workload, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (copy-calibrated); None if it is absent."""
    path = os.path.join(REPO, "profiles", "k_sieve_c2_hbm_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def pct(vals, q):
    v = sorted(vals)
    k = min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))
    return v[k]


def cpu_baseline(cores):
    """The reference scanner (built from its own sources) on the full workload."""
    ref = os.path.join(REPO, "oracle", "_ref", "rung_scan3")
    kind = "reference"
    if not os.path.exists(ref):
        ref = os.path.join(REPO, "oracle", "es_oracle")
        kind = "port"
    if not os.path.exists(ref):
        return None
    argv = [ref, str(LO), str(HI), "--wheel", "2042040", "--filters",
            os.path.join(PKG, "tables", "filters", "f_M2042040_mmax2000_gpu.txt"), "--threads", str(cores)]
    best = None
    for _ in range(2):
        t = time.perf_counter()
        out = subprocess.run(argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL).stdout
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    gold = open(os.path.join(REPO, "tests", "golden", "records", "c2.txt"), "rb").read()
    return {"value": (HI - LO) / best, "unit": "integers/s", "cores": cores, "kind": kind,
            "sample": "the full workload [1e15, 1e15+2^34) once per run, best of 2 runs, wall incl. "
                      "process startup (%.3f s)" % best,
            "record_identical": out == gold}


def tuned_launch(report):
    """(block, tile_bits) the autotuner chose, from its report."""
    block, bits = 256, 32768
    for l in report.splitlines():
        if l.startswith("choice block="):
            block, bits = [int(x.split("=")[1]) for x in l.split()[1:3]]
    return block, bits


def split_records(text):
    out, cur = [], []
    for line in text.splitlines(keepends=True):
        cur.append(line)
        if line.startswith("SUMMARY "):
            out.append("".join(cur))
            cur = []
    return out


def c4_bench(es, reps=5):
    """Batched configuration: 512 windows (seed 0, start log-uniform over [1e9, 1.8e19],
    length log-uniform over [2^16, 2^26]). For each wheel the autotuner measures the
    device on a 64-window sample; the batched form (one persistent launch replayed from
    a hipGraph) and the plain form (one launch per window) are timed; every window's
    record must equal the golden record for that wheel. The faster wheel is the choice."""
    gdir = os.path.join(REPO, "tests", "golden")
    wins = [tuple(map(int, l.split())) for l in open(os.path.join(gdir, "c4_windows.txt"))
            if not l.startswith("#")]
    total = sum(b - a for a, b in wins)
    res = {}
    for M, F in ((2042040, es.F204), (120120, es.F120)):
        gold = split_records(open(os.path.join(gdir, "c4", "records_w%d.txt" % M)).read())
        sc = es.Scanner(wheel=M, filters=F, device=0)
        report = sc.autotune(wins[:64])
        sc.set_graph(True)
        recs, _ = sc.records(wins)
        ok = recs == gold
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            recs, _ = sc.records(wins)
            ts.append(time.perf_counter() - t)
            ok = ok and recs == gold
        tm = sc.timing()
        # The same batch repeated with up to four in flight (the host records of one batch
        # beside the device work of the next).
        pipelined(sc, wins, 4)
        pwall, pok = pipelined(sc, wins, 2 * reps, gold[0])
        ok = ok and pok
        sc.set_graph(False)
        block, bits = tuned_launch(report)
        sc.set_launch(False, block, bits)
        t = time.perf_counter()
        precs = [sc.records([w])[0][0] for w in wins]
        plain_s = time.perf_counter() - t
        ok_plain = precs == gold
        bt = sorted(ts)[len(ts) // 2]
        res[M] = {"batched": total / bt, "batched_ms": bt * 1e3,
                  "batched_pipelined": total * 2 * reps / pwall, "batched_pipelined_ms": pwall * 1e3 / (2 * reps),
                  "plain": total / plain_s,
                  "plain_ms": plain_s * 1e3, "batched_over_plain": plain_s / bt,
                  "records_identical": ok and ok_plain, "tiles": tm["tiles"], "block": block,
                  "tile_bits": bits, "device_ms": tm["device_ms"], "tail_ms": tm["tail_ms"],
                  "layouts": sorted({l.split()[2] for l in report.splitlines() if l.startswith("choice W=")})}
        sc.close()
    wheel = max(res, key=lambda m: res[m]["batched"])
    return {"workload": "512 windows, seed 0, lengths 2^16..2^26, %d integers" % total, "wheel_choice": wheel,
            "by_wheel": res}


def pipelined(sc, batch, steps, want=None, depth=4):
    """steps passes over one batch with up to `depth` in flight (pass i+depth-1 is
    submitted before pass i is collected): the wall time, and whether every pass's first
    record equals want (when given)."""
    ok = True
    t0 = time.perf_counter()
    for i in range(min(depth - 1, steps)):
        sc.submit(batch)
    for i in range(steps):
        if i + depth - 1 < steps:
            sc.submit(batch)
        recs, _ = sc.collect()
        if want is not None:
            ok = ok and recs[0] == want
    return time.perf_counter() - t0, ok


def tuned_table_bench(es, block, bits, steps):
    """The MI355X-tuned filter table (tables/filters/f_M2042040_mmax2000_mi355x.txt, made by
    bin/es_prune) beside the reference's GPU table: the 1e15 window [1e15, 1e15+2^34) as a
    graph step (record compared with the reference scanner's record for that table every
    step), and a 32-window batch of consecutive 2^34 windows from 1e15 (throughput; the
    first window's record compared). Same launch shape for both tables."""
    gold = open(os.path.join(REPO, "tests", "golden", "records", "c2_mi355x.txt")).read()
    wins = [(LO + i * (HI - LO), LO + (i + 1) * (HI - LO)) for i in range(32)]
    out = {"table": os.path.relpath(es.F204_MI355X, REPO)}
    for line in open(es.F204_MI355X):
        if line.startswith("# sha256="):
            out["sha256"] = line.split("=", 1)[1].strip()
        elif line.startswith("# filters="):
            out["table_stats"] = dict(kv.split("=") for kv in line[2:].split())
    for name, F in (("mi355x", es.F204_MI355X), ("gpu", es.F204)):
        sc = es.Scanner(wheel=2042040, filters=F, device=0)
        sc.set_launch(True, block, bits)
        sc.set_graph(True)
        r = {}
        if name == "mi355x":
            pipelined(sc, [(LO, HI)], 4)
            wall, ok = pipelined(sc, [(LO, HI)], steps, gold)
            r["c2"] = {"value": (HI - LO) * steps / wall, "ms_per_step": wall * 1e3 / steps, "record_identical": ok,
                       "device_ms": sc.timing()["device_ms"]}
        recs, _ = sc.records(wins)
        first_ok = name != "mi355x" or recs[0] == gold
        pipelined(sc, wins, 4)
        wall, ok = pipelined(sc, wins, 6, recs[0])
        wall /= 6
        r["batch32"] = {"value": (wins[-1][1] - wins[0][0]) / wall, "ms_per_batch": wall * 1e3,
                        "device_ms": sc.timing()["device_ms"], "first_record_identical": first_ok and ok}
        out[name] = r
        sc.close()
    out["batch32_over_gpu_table"] = out["mi355x"]["batch32"]["value"] / out["gpu"]["batch32"]["value"]
    return out


def c5_bench(es, seconds):
    """Serving configuration: the Poisson request trace (200 req/s, seed 1, start
    log-uniform over [1e6, 1.8e19], length log-uniform over [1e4, 1e10]) replayed in real
    time for its first `seconds` through the continuous-batching server (wheel 2042040,
    the GPU filter table, admission budget 80% of device memory). Latency is arrival to
    record complete; the first requests' records must equal the golden records."""
    gdir = os.path.join(REPO, "tests", "golden")
    rows = [l.split() for l in open(os.path.join(gdir, "c5_trace.txt")) if not l.startswith("#")]
    reqs = [(float(a), int(lo), int(hi)) for a, lo, hi in rows if float(a) < seconds]
    gold = split_records(open(os.path.join(gdir, "c5", "records_first400.txt")).read())
    sc = es.Scanner(wheel=2042040, filters=es.F204, device=0)
    sample = [(lo, min(hi, lo + (1 << 32))) for _, lo, hi in reqs[:64]]
    sc.autotune(sample)
    recs, infos, st = es.serve_trace(sc, reqs, realtime=True, hbm_fraction=0.8)
    sc.close()
    lat = [i["done_ms"] - i["arrival_ms"] for i in infos]
    k = min(len(gold), len(recs))
    total = sum(hi - lo for _, lo, hi in reqs)
    return {"workload": "Poisson 200 req/s (seed 1), first %.0f s of the trace: %d requests, %d integers"
                        % (seconds, len(reqs), total),
            "sustained": total / (st["wall_ms"] * 1e-3), "unit": "integers/s",
            "offered": total / seconds,
            "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99), "max": max(lat)},
            "records_identical_first": k, "records_identical": recs[:k] == gold[:k],
            "batches": st["batches"], "busy_ms": st["busy_ms"], "wall_ms": st["wall_ms"],
            "max_active": st["max_active"], "max_batch_requests": st["max_batch_requests"],
            "hbm_budget_gb": st["budget"] / 1e9, "resident_gb": st["resident"] / 1e9,
            "peak_reserved_gb": st["peak_reserved"] / 1e9}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c4", action="store_true", help="skip the batched 512-window configuration")
    ap.add_argument("--no-tuned", action="store_true", help="skip the MI355X-tuned filter table leg")
    ap.add_argument("--c5-seconds", type=float, default=10.0,
                    help="seconds of the serving trace to replay in real time (0 = skip)")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device (a known-size kernel for scaling HBM counters)")
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
    es = load("es_sieve")
    records = load("records")
    gold = open(os.path.join(REPO, "tests", "golden", "records", "c2.txt")).read()
    sc = es.Scanner(wheel=2042040, filters=es.F204, device=0)

    # ---- batched form: the autotuner measures layouts and the launch shape on the
    # workload itself (untimed setup), then every step replays the captured graph
    report = sc.autotune([(LO, HI)])
    block, bits = tuned_launch(report)
    sc.set_graph(True)
    lat = []
    for i in range(args.warmup):
        (rec,), _ = sc.records([(LO, HI)])
        if rec != gold:
            raise SystemExit("batched record differs from the golden record")
    # One step at a time: the request latency (p50/p99) and the device/host split.
    device_ms = []
    ok_batched = True
    t0 = time.perf_counter()
    for i in range(args.steps):
        s = time.perf_counter()
        (rec,), _ = sc.records([(LO, HI)])
        lat.append((time.perf_counter() - s) * 1e3)
        device_ms.append(sc.timing()["device_ms"])
        ok_batched = ok_batched and rec == gold
    seq_wall = time.perf_counter() - t0
    timing = sc.timing()
    # Steady state (the headline): up to four steps in flight, the host tail and record
    # of step i overlapping the device work of the next steps, as the CLI runs its chunks.
    # Warmup captures every buffer slot's graph.
    if not pipelined(sc, [(LO, HI)], max(es.MAX_IN_FLIGHT, args.warmup), gold)[1]:
        raise SystemExit("pipelined record differs from the golden record")
    wall, ok = pipelined(sc, [(LO, HI)], args.steps, gold)
    ok_batched = ok_batched and ok
    value = (HI - LO) * args.steps / wall
    # The marking kernel's own time: the same launch shape run eagerly, so HIP events
    # bracket k_sieve alone on the scanner's stream.
    sc.set_graph(False)
    sieve_ms = []
    for i in range(max(5, args.steps // 2)):
        (r2,), _ = sc.records([(LO, HI)])
        ok_batched = ok_batched and r2 == gold
        sieve_ms.append(sc.timing()["sieve_ms"])
    layout = sc.choose_layout(HI - LO)
    work = sc.mark_count(LO, HI, layout)

    # ---- plain form: one launch per 2^22-integer window
    wins = [(a, min(HI, a + WINDOW)) for a in range(LO, HI, WINDOW)]
    classes = sc.wheel_classes()
    plain = None
    if args.plain_steps > 0:
        sc.set_launch(False, block, bits)
        (r0,), _ = sc.records([wins[0]])          # warm the window layout
        plain_wall = []
        ok_plain = True
        for _ in range(args.plain_steps):
            t = time.perf_counter()
            recs = [sc.records([w])[0][0] for w in wins]
            plain_wall.append(time.perf_counter() - t)
            ok_plain = ok_plain and records.merge(recs, classes, LO, HI) == gold
        sc.set_launch(True, block, bits)
        plain_value = (HI - LO) * len(plain_wall) / sum(plain_wall)
        plain = {"value": plain_value, "window": WINDOW, "windows": len(wins),
                 "s_per_pass": sum(plain_wall) / len(plain_wall), "record_identical": ok_plain}

    # ---- roofline of the marking kernel (k_sieve)
    lds_ops = sc.lds_peak()
    scatter = sc.lds_scatter_peak()
    kernel_s = sorted(sieve_ms)[len(sieve_ms) // 2] * 1e-3
    lds_bytes = 8 * work["marks"] + 8 * (work["bits"] // 32)   # 32-bit RMW per mark; init + scan per word
    achieved = lds_bytes / kernel_s / 1e9
    out = {
        "metric": "verified integers/s on 1 MI355X, batched vs plain, bit-exact; p50/p99 latency",
        "value": value,
        "unit": "integers/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic: the integer interval itself (no dataset)",
        "config": {"workload": "plain sieve at height 1e15: [1e15, 1e15+2^34), wheel 2042040, "
                               "f_M2042040_mmax2000_gpu.txt (3745 filters)",
                   "form": "batched: one persistent work-queue launch per step replayed from a hipGraph, "
                           "layout and launch shape autotuned; up to four steps in flight, one stream "
                           "each (the host tail of step i overlaps the device work of the next steps)", "layout_Mp": layout,
                   "tiles": work["tiles"], "block": block, "tile_bits": bits},
        "record_identical": ok_batched,
        "latency_ms": {"p50": pct(lat, 50), "p99": pct(lat, 99)},
        "one_at_a_time": {"value": (HI - LO) * args.steps / seq_wall, "ms_per_step": seq_wall * 1e3 / args.steps},
        "split_ms": {"device": sorted(device_ms)[len(device_ms) // 2], "sieve_kernel": kernel_s * 1e3,
                     "host_tail": timing["tail_ms"], "plan": timing["plan_ms"], "batch": timing["batch_ms"]},
        "plain": plain,
        "batched_over_plain": value / plain["value"] if plain else None,
        "roofline": {"bound": "lds", "achieved": achieved, "peak": NOMINAL_LDS_BPS / 1e9, "unit": "GB/s",
                     "frac": achieved / (NOMINAL_LDS_BPS / 1e9),
                     "peak_measured": lds_ops * 8 / 1e9,
                     "frac_of_measured": achieved / (lds_ops * 8 / 1e9),
                     "marks": work["marks"], "bits": work["bits"], "pairs": work["pairs"],
                     "marks_per_s": work["marks"] / kernel_s,
                     "scattered_byte_stores_per_s": scatter,
                     "frac_of_scattered_store_rate": work["marks"] / kernel_s / scatter,
                     "algorithmic_bytes": lds_bytes, "kernel_ms": kernel_s * 1e3,
                     "traffic": profiled_traffic(),
                     "traffic_source": "profiles/k_sieve_c2_hbm_traffic.txt: rocprofv3 FETCH_SIZE + WRITE_SIZE "
                                       "per k_sieve dispatch on this workload, scaled by a 1 GiB copy"},
    }
    if not args.no_tuned:
        out["tuned_table"] = tuned_table_bench(es, block, bits, args.steps)
    if not args.no_c4:
        out["c4"] = c4_bench(es)
    if args.c5_seconds > 0:
        out["c5"] = c5_bench(es, args.c5_seconds)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(min(os.cpu_count() or 1, 16))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
