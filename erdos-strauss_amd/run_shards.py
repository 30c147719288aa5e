#!/usr/bin/env python3
"""Production shard driver for the verification sieve on an AMD GPU, restartable.

The counterpart of the reference's tools/run19.sh (tools/run19.sh:1-140) for the MI355X
drop-in: the same positional arguments, the same shard arithmetic (the scanner's own
--shard i/N), the same per-shard meta file, the same restart and retry rules.

    run_shards.py [LO HI N [OUTDIR]] [--bin es_sieve] [--table FILE] [--threads 12]
                  [--order round|lane] [--attempts 3] [--retry-sleep 30] [--device D]
                  [--no-gpu-check] [--only I[,J...]]

Each shard's certificate stream is hashed while the scanner writes it and then dropped
(at 10^19 the stream is terabytes); the meta keeps what makes the shard re-derivable
and checkable (tools/run19.sh:70-77):

    # shard I/N  rc=R  wall=Ws  attempt=A  YYYY-MM-DD HH:MM:SS
    # binary_commit=... table_sha256=... threads=...
    certs=<RUNG + SOLVED lines>
    cert_sha256=<sha256 of every stdout line but SUMMARY>
    SURVIVOR ...            (verbatim, zero or more)
    SUMMARY lo=... hi=...   (last: its presence marks the meta complete)

Rules kept from the reference: a shard whose meta has a SUMMARY line is never run again;
exit 0 (clean) and exit 2 (survivors present -- a result, never retried) complete a
shard; any other exit, or a run without its SUMMARY line, is a crash and is retried up
to --attempts times, after which shard_I.FAILED is written and the pass moves on; a
lock file guards against two drivers on one directory. The default emission order is
round-major, the order of the GPU scanner the published metas came from
(sieve/cuda/scan.cu:295-309), so cert_sha256 values compare with theirs.
"""
import argparse
import datetime
import fcntl
import glob
import hashlib
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def utc(fmt="%Y-%m-%d %H:%M:%S"):
    return datetime.datetime.now(datetime.timezone.utc).strftime(fmt)


class Run:
    def __init__(self, a):
        self.a = a
        self.out = a.outdir
        self.logf = os.path.join(self.out, "run.log")

    def log(self, msg):
        with open(self.logf, "a") as f:
            f.write("[%s] %s\n" % (utc(), msg))

    def fatal(self, msg):
        print("run_shards: FATAL " + msg)
        self.log("FATAL " + msg)
        return 1


def amd_gpu_visible():
    """A KFD device node and at least one render node: what a HIP process needs."""
    return os.path.exists("/dev/kfd") and bool(glob.glob("/dev/dri/renderD*"))


def commit_id():
    p = os.path.join(HERE, "..", ".git-commit")
    if os.path.exists(p):
        return open(p).read().strip() or "unknown"
    try:
        return subprocess.run(["git", "-C", HERE, "rev-parse", "--short", "HEAD"], stdout=subprocess.PIPE,
                              stderr=subprocess.DEVNULL, timeout=10).stdout.decode().strip() or "unknown"
    except (OSError, subprocess.SubprocessError):
        return "unknown"


def scan_shard(argv, err_path):
    """Runs one shard, hashing its stdout as it arrives: (rc, certs, sha256, survivor
    lines, SUMMARY line or None)."""
    h = hashlib.sha256()
    certs, survivors, summary, buf = 0, [], None, b""
    with open(err_path, "wb") as err:
        p = subprocess.Popen(argv, stdout=subprocess.PIPE, stderr=err)
        while True:
            chunk = p.stdout.read(1 << 20)
            if not chunk:
                break
            parts = (buf + chunk).split(b"\n")
            buf = parts.pop()
            for line in parts:
                if line.startswith(b"SUMMARY "):
                    summary = line.decode()
                    continue
                h.update(line + b"\n")
                if line.startswith(b"RUNG ") or line.startswith(b"SOLVED "):
                    certs += 1
                elif line.startswith(b"SURVIVOR "):
                    survivors.append(line.decode())
        rc = p.wait()
    if buf:                                  # a final line without its newline
        if buf.startswith(b"SUMMARY "):
            summary = buf.decode()
        else:
            h.update(buf)
    return rc, certs, h.hexdigest(), survivors, summary


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("pos", nargs="*", help="LO HI N [OUTDIR]")
    ap.add_argument("--bin", default=os.path.join(HERE, "bin", "es_sieve"))
    ap.add_argument("--table", default=os.path.join(HERE, "tables", "filters", "f_M2042040_mmax2000_gpu.txt"))
    ap.add_argument("--threads", type=int, default=12)
    ap.add_argument("--order", choices=("round", "lane"), default="round")
    ap.add_argument("--attempts", type=int, default=3)
    ap.add_argument("--retry-sleep", type=float, default=30.0)
    ap.add_argument("--device", type=int, default=None)
    ap.add_argument("--no-gpu-check", action="store_true",
                    help="skip the AMD GPU check (for a CPU scanner with the same command line)")
    ap.add_argument("--only", default=None, metavar="I[,J...]",
                    help="run only these shard indices (e.g. to re-derive published shards one at a time)")
    a = ap.parse_args(argv)
    pos = a.pos + [None] * (4 - len(a.pos))
    lo = int(pos[0]) if pos[0] else 255255
    hi = int(pos[1]) if pos[1] else 10 ** 19
    n = int(pos[2]) if pos[2] else 250
    a.outdir = pos[3] or os.path.join(os.path.expanduser("~"), "run19")
    os.makedirs(a.outdir, exist_ok=True)
    run = Run(a)

    lock = open(os.path.join(a.outdir, ".lock"), "w")
    try:
        fcntl.flock(lock, fcntl.LOCK_EX | fcntl.LOCK_NB)
    except OSError:
        print("run_shards: already running")
        return 0

    # Refuse on missing configuration rather than burn every shard's retries on it.
    if not os.access(a.bin, os.X_OK):
        return run.fatal("scanner missing: %s" % a.bin)
    if not os.access(a.table, os.R_OK):
        return run.fatal("filter table missing: %s" % a.table)
    if not a.no_gpu_check and not amd_gpu_visible():
        return run.fatal("no AMD GPU visible")
    tsha = hashlib.sha256(open(a.table, "rb").read()).hexdigest()
    gitc = commit_id()
    run.log("start: [%d,%d) x%d table=%s commit=%s" % (lo, hi, n, tsha, gitc))

    only = None if a.only is None else {int(x) for x in a.only.split(",") if x != ""}
    if only is not None and any(not 0 <= i < n for i in only):
        return run.fatal("--only: shard index out of range [0, %d)" % n)
    work_s, done_now = 0.0, 0
    for i in range(n):
        if only is not None and i not in only:
            continue
        meta = os.path.join(a.outdir, "shard_%d.meta" % i)
        if os.path.exists(meta) and any(l.startswith("SUMMARY ") for l in open(meta)):
            continue
        if os.path.exists(os.path.join(a.outdir, "shard_%d.FAILED" % i)):
            continue
        cmd = [a.bin, str(lo), str(hi), "--shard", "%d/%d" % (i, n), "--filters", a.table, "--threads",
               str(a.threads)]
        if a.order == "round":               # lane order is rung_scan3's default: no flag needed
            cmd += ["--order", "round"]
        if a.device is not None:
            cmd += ["--device", str(a.device)]
        ok = False
        for attempt in range(1, a.attempts + 1):
            err = os.path.join(a.outdir, "shard_%d.err" % i)
            t0 = time.time()
            rc, certs, sha, survivors, summary = scan_shard(cmd, err)
            wall = int(time.time() - t0)
            if rc in (0, 2):
                # Completion is judged by content: the SUMMARY line is printed last.
                if summary is None:
                    run.log("shard %d attempt %d: rc=%d but no SUMMARY -- treating as crash" % (i, attempt, rc))
                else:
                    body = ["# shard %d/%d  rc=%d  wall=%ds  attempt=%d  %s" % (i, n, rc, wall, attempt, utc()),
                            "# binary_commit=%s table_sha256=%s threads=%d" % (gitc, tsha, a.threads),
                            "certs=%d" % certs, "cert_sha256=%s" % sha] + survivors + [summary]
                    with open(meta + ".tmp", "w") as f:
                        f.write("\n".join(body) + "\n")
                    os.replace(meta + ".tmp", meta)
                    if rc == 2:
                        run.log("shard %d: *** SURVIVORS PRESENT *** see %s" % (i, meta))
                    work_s += wall
                    done_now += 1
                    left = sum(1 for j in range(i + 1, n)
                               if not (os.path.exists(os.path.join(a.outdir, "shard_%d.meta" % j))))
                    eta = "%dh" % (left * (work_s / done_now) / 3600) if done_now else "?"
                    run.log("shard %d done rc=%d wall=%ds certs=%d left=%d eta~%s" % (i, rc, wall, certs, left, eta))
                    ok = True
            else:
                tail = open(err, "rb").read()[-200:].decode(errors="replace").replace("\n", " ")
                run.log("shard %d attempt %d: rc=%d -- %s" % (i, attempt, rc, tail))
            if ok:
                break
            if attempt < a.attempts:
                time.sleep(a.retry_sleep)
        if not ok:
            with open(os.path.join(a.outdir, "shard_%d.FAILED" % i), "w") as f:
                f.write(utc() + "\n")
            run.log("shard %d FAILED after %d attempts -- skipping, re-run manually" % (i, a.attempts))

    nfail = len(glob.glob(os.path.join(a.outdir, "shard_*.FAILED")))
    ndone = sum(1 for m in glob.glob(os.path.join(a.outdir, "shard_*.meta"))
                if any(l.startswith("SUMMARY ") for l in open(m)))
    run.log("pass complete: %d/%d shards done, %d FAILED" % (ndone, n, nfail))
    return 0


if __name__ == "__main__":
    sys.exit(main())
