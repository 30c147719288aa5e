#!/usr/bin/env python3
"""Audit a sieve run from its shard metas; exit 0 iff every check passes.

    audit_shards.py DIR [DIR ...] [--band] [--expect-lo LO] [--expect-hi HI]
                    [--expect-shards N] [--against REFDIR]

The counterpart of the reference's sieve/verify_shards.py (:29-130) with its meta
grammar (sieve/shard_meta.py:8-16), for runs of run_shards.py or of the reference's
tools/run19.sh. Checks, per directory (or over all of them as one band with --band):
  * every meta complete (its SUMMARY line present), shard indices exactly 0..N-1;
  * every shard rc=0;
  * shard boundaries, sorted by lo, tile the band exactly (no gap, no overlap) -- the
    check summed counters cannot replace: a span computed with floor instead of ceiling
    drops (HI-LO) mod N integers that almost never hold a prime;
  * no survivor anywhere, no SURVIVOR line;
  * per shard: mr - composite = rung + direct + survivors;
  * provenance uniform across shards (table, class digest, wheel, lanes, mmax,
    certificate bounds, wheel_pmin);
  * no two shards with the same certificate-stream digest.
--against REFDIR also requires every shard's certs, cert_sha256, SURVIVOR lines and
SUMMARY to equal the same shard's in REFDIR (a re-derivation check).
"""
import argparse
import glob
import os
import re
import sys

HEADER = re.compile(r"# shard (\d+)/(\d+)\s+rc=(-?\d+)\s+wall=(\d+)s\s+attempt=(\d+)\s*(.*)")
UNIFORM = ("table_sha256", "classes_sha256", "wheelM", "lanes", "mmax", "ucert_rmax", "ucert_umax",
           "wheel_pmin")


def parse_meta(path):
    """One meta file -> dict (integers as int)."""
    d = {"path": path, "survivor_lines": [], "complete": False}
    for raw in open(path):
        line = raw.strip()
        if not line:
            continue
        m = HEADER.match(line)
        if m:
            d.update(idx=int(m.group(1)), nshards=int(m.group(2)), rc=int(m.group(3)), wall=int(m.group(4)),
                     attempt=int(m.group(5)))
            continue
        if line.startswith("# "):
            for tok in line[2:].split():
                if "=" in tok:
                    k, v = tok.split("=", 1)
                    d[k] = int(v) if v.isdigit() else v
            continue
        if line.startswith("SURVIVOR "):
            d["survivor_lines"].append(line)
            continue
        if line.startswith("SUMMARY "):
            d["summary_line"] = line
            for tok in line.split()[1:]:
                k, v = tok.split("=", 1)
                d[k] = int(v) if v.isdigit() else v
            d["complete"] = True
            continue
        if "=" in line:
            k, v = line.split("=", 1)
            d[k] = int(v) if v.isdigit() else v
    if d["complete"]:
        d["primes"] = d.get("mr", 0) - d.get("composite", 0)
    return d


def load(dirs):
    out = []
    for d in dirs:
        for p in glob.glob(os.path.join(d, "shard_*.meta")):
            m = parse_meta(p)
            m["box"] = d
            out.append(m)
    out.sort(key=lambda m: (m.get("lo", -1), m.get("idx", -1)))
    return out


def check(ds, label, expect_lo=None, expect_hi=None, expect_n=None, against=None):
    ok = True

    def fail(msg):
        nonlocal ok
        ok = False
        print("  FAIL " + msg)

    def good(msg):
        print("  ok   " + msg)

    print("== %s  (%d shard metas)" % (label, len(ds)))
    if not ds:
        fail("no shard metas found")
        return False
    bad = [m.get("idx") for m in ds if not m["complete"]]
    fail("metas without a SUMMARY line: %s" % bad) if bad else good("every meta complete (SUMMARY present)")
    boxes = sorted({m["box"] for m in ds})
    for box in boxes:
        sub = [m for m in ds if m["box"] == box]
        n = sub[0].get("nshards")
        idx = sorted(m.get("idx", -1) for m in sub)
        missing = sorted(set(range(n or 0)) - set(idx))
        tag = (box + ": ") if len(boxes) > 1 else ""
        if n is None or missing or len(idx) != len(set(idx)):
            fail("%sshard indices not exactly 0..%s (missing %s, duplicates %s)" %
                 (tag, None if n is None else n - 1, missing[:10], len(idx) != len(set(idx))))
        else:
            good("%sshard indices 0..%d complete, no duplicates" % (tag, n - 1))
    rcs = [(m.get("idx"), m.get("rc")) for m in ds if m.get("rc") != 0]
    fail("shards with rc != 0: %s" % rcs[:10]) if rcs else good("every shard rc=0")
    complete = [m for m in ds if m["complete"]]
    gaps = [(a["hi"], b["lo"]) for a, b in zip(complete, complete[1:]) if a["hi"] != b["lo"]]
    if gaps:
        fail("%d boundary gap(s)/overlap(s), first: hi=%d then lo=%d (delta %+d)" %
             (len(gaps), gaps[0][0], gaps[0][1], gaps[0][1] - gaps[0][0]))
    elif complete:
        good("boundaries tile exactly: [%d, %d)" % (complete[0]["lo"], complete[-1]["hi"]))
    if complete and expect_lo is not None and complete[0]["lo"] != expect_lo:
        fail("band starts at %d, expected %d" % (complete[0]["lo"], expect_lo))
    if complete and expect_hi is not None and complete[-1]["hi"] != expect_hi:
        fail("band ends at %d, expected %d" % (complete[-1]["hi"], expect_hi))
    if expect_n is not None and len(ds) != expect_n:
        fail("%d shards, expected %d" % (len(ds), expect_n))
    surv = sum(m.get("survivors", 0) for m in complete)
    lines = [l for m in ds for l in m["survivor_lines"]]
    if surv or lines:
        fail("survivors=%d, %d SURVIVOR line(s) -- an unsolved prime" % (surv, len(lines)))
        for l in lines[:20]:
            print("         " + l)
    else:
        good("survivors=0 across the band, no SURVIVOR lines")
    acc = [m.get("idx") for m in complete
           if m["primes"] != m.get("rung", 0) + m.get("direct", 0) + m.get("survivors", 0)]
    if acc:
        fail("mr - composite != rung + direct + survivors in shards %s" % acc[:10])
    else:
        good("accounting identity holds per shard: mr - composite = rung + direct + survivors")
    for k in UNIFORM:
        vals = {str(m.get(k)) for m in ds}
        if len(vals) != 1:
            fail("%s not uniform across shards: %s" % (k, sorted(vals)[:5]))
    digests = [m.get("cert_sha256") for m in ds]
    if len(set(digests)) != len(digests):
        fail("duplicate cert-stream sha256 -- two shards produced identical output")
    else:
        good("%d distinct cert-stream digests" % len(digests))
    if against is not None:
        ref = {m.get("idx"): m for m in load([against])}
        diff = [m.get("idx") for m in ds if m.get("idx") not in ref or any(
            m.get(k) != ref[m.get("idx")].get(k) for k in ("certs", "cert_sha256", "summary_line")) or
            m["survivor_lines"] != ref[m.get("idx")]["survivor_lines"]]
        if diff:
            fail("%d shard(s) differ from %s: %s" % (len(diff), against, diff[:10]))
        else:
            good("every shard's certs, cert_sha256, SURVIVOR lines and SUMMARY equal %s's" % against)
    commits = sorted({str(m.get("binary_commit")) for m in ds})
    print("  --   binary_commit=%s  retries=%d" % (",".join(commits), sum(m.get("attempt", 1) - 1 for m in ds)))
    print("  ==> " + ("PASS" if ok else "FAIL"))
    return ok


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("dirs", nargs="+")
    ap.add_argument("--band", action="store_true", help="all directories as one contiguous band")
    ap.add_argument("--expect-lo", type=int)
    ap.add_argument("--expect-hi", type=int)
    ap.add_argument("--expect-shards", type=int)
    ap.add_argument("--against", help="directory of reference metas that every shard must equal")
    a = ap.parse_args(argv)
    if a.band:
        ok = check(load(a.dirs), " + ".join(a.dirs), a.expect_lo, a.expect_hi, a.expect_shards, a.against)
    else:
        ok = all([check(load([d]), d, a.expect_lo, a.expect_hi, a.expect_shards, a.against) for d in a.dirs])
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
