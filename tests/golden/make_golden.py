#!/usr/bin/env python3
"""Regenerate the golden fixtures under tests/golden/.

Runs only in the development container, where the reference tree is mounted at
REF_ROOT (default /root/reference) and its scanner has been built from source by
`make -C oracle ref` into oracle/_ref/rung_scan3. Nothing here runs where the tests run
on a GPU: the tests read the committed outputs.

What it writes:
  records/<name>.txt     complete stdout of the reference scanner (rung_scan3) on a
                         fixed interval and flag set; <name>.args holds the argv used
  rederive/...           data files the reference itself keeps as golden records
                         (data/rederive: decades d6..d19, w19 certificates, the
                         head/tail edges of w64, s0, s69, and their SUMMARY lines)
  direct/<lo>.txt        the reference's --check-direct rows (stage E lines)
  class_tables.sha256    sha256 of the reference's certified class tables
  c4_windows.txt         512 batched windows, numpy PCG64 seed 0
  c5_trace.txt           Poisson request trace, 200 req/s, numpy PCG64 seed 1
  c4/records_w<M>.txt    per-window records of the 512 batched windows (--c4-only)
  c5/records_first400.txt  records of the trace's first 400 requests (--c5-only)
  c5/records_sha256.txt  sha256 of every request's record over the whole trace (--c5-digests)
  shards/lane/shard_*.meta shard metas of a 7-shard run of the reference binary (--shards-only)

Usage: python3 tests/golden/make_golden.py [--skip-slow | --c4-only | --c5-only | --c5-digests |
                                            --shards-only | --only NAME[,NAME...]]
  --skip-slow  leave out the top-of-range interval, whose reference startup
               (a base-prime sieve to sqrt(hi)) takes minutes and ~9 GB.
"""
import hashlib
import math
import os
import shutil
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REF_ROOT = os.environ.get("REF_ROOT", "/root/reference")
REF_BIN = os.path.join(REPO, "oracle", "_ref", "rung_scan3")
FILT = os.path.join(REPO, "erdos-strauss_amd", "tables", "filters")
F120 = os.path.join(FILT, "f_M120120_mmax2000_runtime.txt")
F204 = os.path.join(FILT, "f_M2042040_mmax2000_gpu.txt")
F204R = os.path.join(FILT, "f_M2042040_mmax2000_runtime.txt")
F204M = os.path.join(FILT, "f_M2042040_mmax2000_mi355x.txt")   # the MI355X-tuned table (es_prune)
DIRECT_PMAX = 18446744065119614976
TOP = DIRECT_PMAX

# (name, lo, hi, extra args). Filter paths are written relative to the repo root in the
# .args file so the tests can rerun the same command anywhere.
RECORDS = [
    # the three plain configurations
    ("c1", 255255, 255255 + 10**8, ["--wheel", "120120", "--filters", F120]),
    ("c2", 10**15, 10**15 + 2**34, ["--wheel", "2042040", "--filters", F204]),
    ("c3", TOP - 2**34, TOP, ["--wheel", "2042040", "--filters", F204]),
    # survivor-heavy rows (small generated cover sets leave many primes for stage D)
    ("mmax200_w120", 10**12, 10**12 + 2 * 10**8, ["--wheel", "120120", "--mmax", "200"]),
    ("mmax200_w204", 10**12, 10**12 + 2 * 10**8, ["--wheel", "2042040", "--mmax", "200"]),
    ("mmax2000_w120", 10**12, 10**12 + 10**7, ["--wheel", "120120", "--mmax", "2000"]),
    # stage B' on: covered unchanged, sieved > 0
    ("sieve1000", 10**12, 10**12 + 10**9, ["--filters", F204, "--sieve", "1000"]),
    ("sieve1000_w120", 10**12, 10**12 + 10**9, ["--wheel", "120120", "--filters", F120, "--sieve", "1000"]),
    # segment carry: many segments per lane
    ("seg4096", 10**12, 10**12 + 2 * 10**10, ["--filters", F204, "--seg", "4096"]),
    # runtime table and shard arithmetic (uneven and empty shards)
    ("runtime_tbl", 10**13, 10**13 + 10**10, ["--filters", F204R]),
    ("shard3of7", 10**13, 10**13 + 10**10, ["--filters", F204, "--shard", "3/7"]),
    ("shard6of7", 10**13, 10**13 + 10**10, ["--filters", F204, "--shard", "6/7"]),
    ("shard_empty", 10**13, 10**13 + 50, ["--filters", F204, "--shard", "63/64"]),
    ("tiny", 10**13, 10**13 + 1, ["--filters", F204]),
    ("empty", 10**13, 10**13, ["--filters", F204]),
    # survivors present: small mmax with few filters and SURVIVOR lines suppressed or not
    ("mmax60_w120", 10**9, 10**9 + 10**8, ["--wheel", "120120", "--mmax", "60"]),
    ("mmax60_w120_nosurv", 10**9, 10**9 + 10**8, ["--wheel", "120120", "--mmax", "60", "--no-emit-survivors"]),
    # above 2^63 at moderate cost
    ("above63", 2**63 + 10**9, 2**63 + 10**9 + 4 * 10**9, ["--filters", F204]),
    # the MI355X-tuned table, read by the reference scanner itself
    ("c2_mi355x", 10**15, 10**15 + 2**34, ["--wheel", "2042040", "--filters", F204M]),
    ("d12_mi355x", 10**12, 10**12 + 10**10, ["--filters", F204M]),
    ("d18_mi355x", 10**18, 10**18 + 4 * 10**9, ["--filters", F204M]),
    ("above63_mi355x", 2**63 + 10**9, 2**63 + 10**9 + 4 * 10**9, ["--filters", F204M]),
    ("top_mi355x", TOP - 2**30, TOP, ["--filters", F204M]),
]
SLOW = {"c3"}

DIRECT_ROWS = [1000003, 9223372036854770000, 9223372036854775808, 9999999999999990000,
               18446744065119000000]


def rel(a):
    return os.path.relpath(a, REPO) if a.startswith(REPO) else a


def run_ref(name, lo, hi, extra):
    out = os.path.join(HERE, "records", name + ".txt")
    args = [str(lo), str(hi)] + extra
    with open(out, "w") as f:
        p = subprocess.run([REF_BIN] + args, stdout=f, stderr=subprocess.PIPE, text=True)
    with open(os.path.join(HERE, "records", name + ".args"), "w") as f:
        f.write(" ".join(rel(a) for a in args) + "\n")
        f.write("rc=%d\n" % p.returncode)
    print("record", name, "rc", p.returncode, flush=True)


def copy_rederive():
    src = os.path.join(REF_ROOT, "data", "rederive")
    dst = os.path.join(HERE, "rederive")
    os.makedirs(dst, exist_ok=True)
    for k in range(6, 20):
        shutil.copy(os.path.join(src, "decades", "d%d.txt" % k), os.path.join(dst, "d%d.txt" % k))
    pairs = [("windows/w19_certificates.txt", "w19_certificates.txt"),
             ("windows/w64_head.txt", "w64_head.txt"), ("windows/w64_tail.txt", "w64_tail.txt"),
             ("windows/w64_summary.txt", "w64_summary.txt"), ("windows/w64.sha256", "w64.sha256"),
             ("s0/head.txt", "s0_head.txt"), ("s0/tail.txt", "s0_tail.txt"),
             ("s0/summary.txt", "s0_summary.txt"), ("s0/info.txt", "s0_info.txt"),
             ("s69/head.txt", "s69_head.txt"), ("s69/tail.txt", "s69_tail.txt"),
             ("s69/summary.txt", "s69_summary.txt"), ("s69/info.txt", "s69_info.txt")]
    for a, b in pairs:
        shutil.copy(os.path.join(src, a), os.path.join(dst, b))
    for root in ("data/c19/shard_0.meta", "data/c19/shard_1.meta"):
        shutil.copy(os.path.join(REF_ROOT, root), os.path.join(dst, "c19_" + os.path.basename(root)))


def direct_rows():
    d = os.path.join(HERE, "direct")
    os.makedirs(d, exist_ok=True)
    for lo in DIRECT_ROWS:
        with open(os.path.join(d, "%d.txt" % lo), "w") as f:
            subprocess.run([REF_BIN, "--check-direct", str(lo), "40"], stdout=f,
                           stderr=subprocess.DEVNULL, check=True)


def class_table_digests():
    with open(os.path.join(HERE, "class_tables.sha256"), "w") as f:
        for M in (120120, 2042040):
            p = os.path.join(REF_ROOT, "tables", "class_table_%d.txt" % M)
            f.write("%s  %d\n" % (hashlib.sha256(open(p, "rb").read()).hexdigest(), M))


def c4_windows():
    """512 windows: start log-uniform over [1e9, 1.8e19], length log-uniform over
    [2^16, 2^26] integers, drawn as (u_start, u_len) per window from PCG64 seed 0."""
    rng = np.random.default_rng(0)
    with open(os.path.join(HERE, "c4_windows.txt"), "w") as f:
        f.write("# lo hi   (512 windows, numpy default_rng(0), (u_start, u_len) per window)\n")
        for _ in range(512):
            us, ul = rng.random(), rng.random()
            start = int(math.exp(math.log(1e9) + us * (math.log(1.8e19) - math.log(1e9))))
            length = int(2.0 ** (16 + 10 * ul))
            f.write("%d %d\n" % (start, start + length))


def c5_trace(seconds=60.0):
    """Poisson arrivals at 200 req/s (exponential gaps), start log-uniform over
    [1e6, 1.8e19], length log-uniform over [1e4, 1e10]; PCG64 seed 1, drawn as
    (gap, u_start, u_len) per request."""
    rng = np.random.default_rng(1)
    t = 0.0
    with open(os.path.join(HERE, "c5_trace.txt"), "w") as f:
        f.write("# arrival_s lo hi   (Poisson 200/s, numpy default_rng(1), (gap, u_start, u_len))\n")
        while True:
            gap, us, ul = rng.exponential(1 / 200.0), rng.random(), rng.random()
            t += gap
            if t >= seconds:
                break
            start = int(math.exp(math.log(1e6) + us * (math.log(1.8e19) - math.log(1e6))))
            length = int(math.exp(math.log(1e4) + ul * (math.log(1e10) - math.log(1e4))))
            f.write("%.6f %d %d\n" % (t, start, start + length))


ORACLE_BIN = os.path.join(REPO, "oracle", "es_oracle")


def c4_records():
    """Per-window records of the 512 batched windows for both wheels, from the oracle
    (itself pinned to the reference binary and its golden records by tests/test_oracle.py);
    windows with hi < 1e15, where the reference binary's startup is cheap, are also run
    through the reference binary and must agree byte for byte."""
    wins = [tuple(map(int, l.split())) for l in open(os.path.join(HERE, "c4_windows.txt"))
            if not l.startswith("#")]
    d = os.path.join(HERE, "c4")
    os.makedirs(d, exist_ok=True)
    for M, F in ((2042040, F204), (120120, F120)):
        out = []
        checked = 0
        for lo, hi in wins:
            argv = [str(lo), str(hi), "--wheel", str(M), "--filters", F]
            rec = subprocess.run([ORACLE_BIN] + argv + ["--threads", "8"], stdout=subprocess.PIPE,
                                 stderr=subprocess.DEVNULL).stdout
            if hi < 10 ** 15:
                ref = subprocess.run([REF_BIN] + argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL).stdout
                if ref != rec:
                    raise SystemExit("oracle and reference disagree on window %d %d" % (lo, hi))
                checked += 1
            out.append(rec)
        with open(os.path.join(d, "records_w%d.txt" % M), "wb") as f:
            f.write(b"".join(out))
        print("c4 wheel", M, "records", len(out), "cross-checked against the reference:", checked, flush=True)


C5_GOLDEN = 400


def c5_records():
    """Records of the first C5_GOLDEN requests of the serving trace (wheel 2042040, the GPU
    filter table), from the oracle; requests with hi < 1e15 are also run through the
    reference binary and must agree byte for byte."""
    rows = [l.split() for l in open(os.path.join(HERE, "c5_trace.txt")) if not l.startswith("#")]
    d = os.path.join(HERE, "c5")
    os.makedirs(d, exist_ok=True)
    out, checked = [], 0
    for _, lo, hi in rows[:C5_GOLDEN]:
        argv = [lo, hi, "--wheel", "2042040", "--filters", F204]
        rec = subprocess.run([ORACLE_BIN] + argv + ["--threads", "8"], stdout=subprocess.PIPE,
                             stderr=subprocess.DEVNULL).stdout
        if int(hi) < 10 ** 15:
            ref = subprocess.run([REF_BIN] + argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL).stdout
            if ref != rec:
                raise SystemExit("oracle and reference disagree on request %s %s" % (lo, hi))
            checked += 1
        out.append(rec)
    with open(os.path.join(d, "records_first%d.txt" % C5_GOLDEN), "wb") as f:
        f.write(b"".join(out))
    print("c5 records", len(out), "cross-checked against the reference:", checked, flush=True)


def split_records(data):
    out, cur = [], []
    for line in data.splitlines(keepends=True):
        cur.append(line)
        if line.startswith(b"SUMMARY ") or line.startswith(b"FATAL"):
            out.append(b"".join(cur))
            cur = []
    return out


def c5_digests(ref_every=8):
    """sha256 of the record of every request of the serving trace (wheel 2042040, the GPU
    filter table), from the oracle in batch mode. The first C5_GOLDEN must equal the
    committed records, and every ref_every-th request with hi < 1e15 (where the reference
    binary's startup is cheap) is also run through the reference binary and must agree
    byte for byte."""
    rows = [l.split() for l in open(os.path.join(HERE, "c5_trace.txt")) if not l.startswith("#")]
    tmp = os.path.join(REPO, "_work", "c5_batch.txt")
    os.makedirs(os.path.dirname(tmp), exist_ok=True)
    with open(tmp, "w") as f:
        for _, lo, hi in rows:
            f.write("%s %s\n" % (lo, hi))
    data = subprocess.run([ORACLE_BIN, "--batch", tmp, "--wheel", "2042040", "--filters", F204, "--threads", "8"],
                          stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, check=True).stdout
    recs = split_records(data)
    if len(recs) != len(rows):
        raise SystemExit("oracle batch gave %d records for %d requests" % (len(recs), len(rows)))
    first = open(os.path.join(HERE, "c5", "records_first%d.txt" % C5_GOLDEN), "rb").read()
    if b"".join(recs[:C5_GOLDEN]) != first:
        raise SystemExit("oracle batch disagrees with the committed first records")
    checked = 0
    cheap = [i for i, (_, lo, hi) in enumerate(rows) if int(hi) < 10 ** 15]
    for i in cheap[::ref_every]:
        _, lo, hi = rows[i]
        ref = subprocess.run([REF_BIN, lo, hi, "--wheel", "2042040", "--filters", F204], stdout=subprocess.PIPE,
                             stderr=subprocess.DEVNULL).stdout
        if ref != recs[i]:
            raise SystemExit("oracle and reference disagree on request %d: %s %s" % (i, lo, hi))
        checked += 1
    with open(os.path.join(HERE, "c5", "records_sha256.txt"), "w") as f:
        f.write("# sha256(record) lo hi per request of c5_trace.txt, in trace order (oracle, batch mode; "
                "%d requests cross-checked against the reference binary)\n" % checked)
        for (_, lo, hi), r in zip(rows, recs):
            f.write("%s %s %s\n" % (hashlib.sha256(r).hexdigest(), lo, hi))
    print("c5 digests", len(recs), "cross-checked against the reference:", checked, flush=True)


SHARD_LO, SHARD_HI, SHARD_N = 10 ** 12 + 7, 10 ** 12 + 4 * 10 ** 10 + 3, 7


def shard_metas():
    """Shard metas of [SHARD_LO, SHARD_HI) in SHARD_N shards, written by the production
    driver (erdos-strauss_amd/run_shards.py) running the reference binary itself, in its
    own (lane) order: the deterministic fields (certs, cert_sha256, SURVIVOR, SUMMARY)
    are what a run of the drop-in must reproduce."""
    out = os.path.join(HERE, "shards", "lane")
    shutil.rmtree(out, ignore_errors=True)
    subprocess.run([sys.executable, os.path.join(REPO, "erdos-strauss_amd", "run_shards.py"), str(SHARD_LO),
                    str(SHARD_HI), str(SHARD_N), out, "--bin", REF_BIN, "--table", F204, "--order", "lane",
                    "--no-gpu-check", "--threads", "8", "--retry-sleep", "0"], check=True)
    for f in os.listdir(out):
        if not f.endswith(".meta"):
            os.remove(os.path.join(out, f))


def main():
    skip_slow = "--skip-slow" in sys.argv
    if "--shards-only" in sys.argv:
        shard_metas()
        return
    if "--c4-only" in sys.argv:
        c4_records()
        return
    if "--c5-only" in sys.argv:
        c5_records()
        return
    if "--c5-digests" in sys.argv:
        c5_digests()
        return
    if "--only" in sys.argv:
        names = set(sys.argv[sys.argv.index("--only") + 1].split(","))
        for name, lo, hi, extra in RECORDS:
            if name in names:
                run_ref(name, lo, hi, extra)
        return
    os.makedirs(os.path.join(HERE, "records"), exist_ok=True)
    c4_windows()
    c5_trace()
    class_table_digests()
    copy_rederive()
    direct_rows()
    for name, lo, hi, extra in RECORDS:
        if skip_slow and name in SLOW:
            continue
        run_ref(name, lo, hi, extra)


if __name__ == "__main__":
    main()
