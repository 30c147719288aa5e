"""Record text helpers: parse a rung_scan3 record, and merge the records of adjacent
windows into the record of their union (counters add; certificate lines re-sorted by
the lane key (rank of p mod M in the wheel, p), rung_scan3.cpp:1146-1151)."""
import bisect

COUNTERS = ("positions", "covered", "sieved", "mr", "composite", "rung", "direct", "survivors")


def parse_summary(line):
    out = {}
    for tok in line.split()[1:]:
        k, _, v = tok.partition("=")
        out[k] = int(v) if v.isdigit() else v
    return out


def split_record(rec):
    lines = rec.splitlines(keepends=True)
    assert lines and lines[-1].startswith("SUMMARY "), "record must end with SUMMARY"
    return lines[:-1], parse_summary(lines[-1])


def format_summary(s):
    keys = ("lo", "hi", "wheelM", "lanes", "mmax", "ucert_rmax", "ucert_umax", "wheel_pmin", "classes_sha256") + COUNTERS
    return "SUMMARY " + " ".join("%s=%s" % (k, s[k]) for k in keys) + "\n"


def merge(records, classes, lo, hi):
    """Records of windows tiling [lo, hi) -> the record of [lo, hi)."""
    M = None
    keyed = []
    total = None
    for rec in records:
        lines, s = split_record(rec)
        M = s["wheelM"]
        if total is None:
            total = dict(s)
            for c in COUNTERS:
                total[c] = 0
        for c in COUNTERS:
            total[c] += s[c]
        for l in lines:
            p = int(l.split()[1])
            keyed.append((bisect.bisect_left(classes, p % M), p, l))
    keyed.sort()
    total["lo"], total["hi"] = lo, hi
    return "".join(l for _, _, l in keyed) + format_summary(total)
