"""GPU parity: the HIP funnel against the oracle and the reference's records.

Bar: bit-exact. Survivor sets and primality flags must equal the oracle's element for
element; records (certificate lines in lane order + SUMMARY) and exit codes must be
byte-identical to the reference scanner's golden records. Every sub-wheel layout and
launch shape must give the same bytes.
"""
import importlib.util
import os
import subprocess

import pytest

from tests import paths

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
es = importlib.util.module_from_spec(spec)
spec.loader.exec_module(es)

F120, F204 = es.F120, es.F204
TOP = 18446744065119614976


def parse_args(argv):
    """rung_scan3 argv -> (scanner kwargs, lo, hi, emit_survivors), shard applied."""
    lo, hi = int(argv[0]), int(argv[1])
    kw = {"wheel": 2042040, "mmax": 10000, "sieve": 0, "filters": None}
    shard, nshard, emit = 0, 1, True
    i = 2
    while i < len(argv):
        a = argv[i]
        if a == "--wheel": kw["wheel"] = int(argv[i + 1]); i += 2
        elif a == "--mmax": kw["mmax"] = int(argv[i + 1]); i += 2
        elif a == "--sieve": kw["sieve"] = int(argv[i + 1]); i += 2
        elif a == "--filters": kw["filters"] = os.path.join(paths.REPO, argv[i + 1]); i += 2
        elif a == "--seg": i += 2
        elif a == "--shard":
            shard, nshard = map(int, argv[i + 1].split("/")); i += 2
        elif a == "--no-emit-survivors": emit = False; i += 1
        else: raise ValueError(a)
    if nshard > 1:
        span = (hi - lo + nshard - 1) // nshard
        a = lo + shard * span
        lo, hi = a, max(a, min(hi, a + span))
    return kw, lo, hi, emit


_scanners = {}


def scanner(**kw):
    key = tuple(sorted(kw.items()))
    if key not in _scanners:
        _scanners[key] = es.Scanner(device=0, **kw)
    sc = _scanners[key]
    sc.set_layout(0)
    sc.set_launch(True, 256, 32768)
    return sc


def oracle_survivors(lo, hi, kw):
    argv = [paths.ORACLE_BIN, str(lo), str(hi), "--wheel", str(kw["wheel"]), "--mmax", str(kw["mmax"]),
            "--dump-survivors", "--threads", "8"]
    if kw.get("filters"):
        argv += ["--filters", kw["filters"]]
    if kw.get("sieve"):
        argv += ["--sieve", str(kw["sieve"])]
    out = subprocess.run(argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, check=True).stdout.decode()
    pairs = sorted((int(t[1]), int(t[2])) for t in (l.split() for l in out.splitlines()))
    return [p for p, _ in pairs], [f for _, f in pairs]


def oracle_record(lo, hi, kw, emit=True):
    argv = [paths.ORACLE_BIN, str(lo), str(hi), "--wheel", str(kw["wheel"]), "--mmax", str(kw["mmax"]),
            "--threads", "8"]
    if kw.get("filters"):
        argv += ["--filters", kw["filters"]]
    if kw.get("sieve"):
        argv += ["--sieve", str(kw["sieve"])]
    if not emit:
        argv.append("--no-emit-survivors")
    p = subprocess.run(argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL)
    return p.stdout.decode(), p.returncode


def summary_counts(rec):
    last = rec.strip().splitlines()[-1].split()
    return {t.split("=")[0]: int(t.split("=")[1]) for t in last[1:] if "=" in t and t.split("=")[1].isdigit()}


# ---------------------------------------------------------------- records

RECORDS = paths.record_cases()


@pytest.mark.parametrize("name,argv,rc", RECORDS, ids=[c[0] for c in RECORDS])
def test_record_via_capi(name, argv, rc):
    """Every committed reference record, byte for byte, through the C ABI."""
    kw, lo, hi, emit = parse_args(argv)
    print("record", name, lo, hi, kw, flush=True)
    sc = scanner(**kw)
    (rec,), (code,) = sc.records([(lo, hi)], emit_survivors=emit)
    assert rec == paths.golden_record(name).decode()
    assert code == rc


@pytest.mark.parametrize("name", ["c1", "c2", "c3", "sieve1000", "shard3of7", "mmax60_w120_nosurv", "empty"])
def test_record_via_cli(name):
    """The drop-in binary: same argv, same stdout bytes, same exit code."""
    argv, rc = [(a, r) for n, a, r in RECORDS if n == name][0]
    print("cli", name, argv, flush=True)
    got_rc, out, err = paths.run(paths.SIEVE_BIN, argv, timeout=300)
    assert got_rc == rc, err.decode()
    assert out == paths.golden_record(name)


def test_cli_default_filters():
    """The binary with no flags at all: covers generated to mmax 10000 (the reference's
    default), whose offset rows are too long to stage beside long tiles. Same stdout
    and exit code as the oracle's record."""
    lo, hi = 10 ** 9, 10 ** 9 + 10 ** 10      # 11.3 M positions, 13,223 composite survivors
    print("cli default", lo, hi, flush=True)
    want, want_rc = oracle_record(lo, hi, dict(wheel=2042040, mmax=10000, sieve=0, filters=None))
    got_rc, out, err = paths.run(paths.SIEVE_BIN, [str(lo), str(hi)], timeout=300)
    assert got_rc == want_rc, err.decode()
    assert out.decode() == want


@pytest.mark.parametrize("k", [6, 9, 13, 16, 19])
def test_decade_records(k):
    """The reference's own records data/rederive/decades/d<k>.txt."""
    want = open(os.path.join(paths.GOLDEN, "rederive", "d%d.txt" % k)).read()
    s = summary_counts(want)
    print("decade", k, flush=True)
    sc = scanner(filters=F204)
    (rec,), _ = sc.records([(s["lo"], s["hi"])])
    assert rec == want


# ---------------------------------------------------------------- survivors (kernel parity)

SURV_CASES = [
    ("c1win", 255255, 255255 + 2 ** 20, dict(wheel=120120, filters=F120)),
    ("w120_1e12", 10 ** 12, 10 ** 12 + 3 * 10 ** 8, dict(wheel=120120, filters=F120)),
    ("w204_1e12", 10 ** 12, 10 ** 12 + 10 ** 9, dict(filters=F204)),
    ("w204_top", TOP - 2 ** 31, TOP, dict(filters=F204)),
    ("mmax200", 10 ** 12, 10 ** 12 + 10 ** 8, dict(mmax=200)),
    # The command line's default filter set (covers generated to mmax 10000, ~77,000
    # marking units): a 154 KB offset row, staged in LDS beside short tiles (wheel
    # 2042040) and read from global memory beside long ones (the small sub-wheels).
    ("generated", 10 ** 12, 10 ** 12 + 10 ** 8, dict()),
    ("sieve", 10 ** 12, 10 ** 12 + 10 ** 9, dict(filters=F204, sieve=1000)),
    ("sieve_w120", 10 ** 11, 10 ** 11 + 10 ** 8, dict(wheel=120120, filters=F120, sieve=3000)),
]


def candidates(M):
    q = M // 840
    return sorted(M // d for d in range(1, q + 1) if q % d == 0)


@pytest.mark.parametrize("name,lo,hi,kw", SURV_CASES, ids=[c[0] for c in SURV_CASES])
def test_survivors_every_layout(name, lo, hi, kw):
    """Stage-C survivor set, primality flags and counters equal the oracle's in every
    sub-wheel layout (k_mark/k_compact/k_primality parity, sieve_cuda --check-mark)."""
    full = dict(wheel=2042040, mmax=10000, sieve=0, filters=None)
    full.update(kw)
    want_p, want_f = oracle_survivors(lo, hi, full)
    want_rec, _ = oracle_record(lo, hi, full)
    want_c = summary_counts(want_rec)
    sc = scanner(**full)
    for Mp in [0] + candidates(full["wheel"]):
        print(name, "layout", Mp, flush=True)
        sc.set_layout(Mp)
        p, f, c = sc.survivors(lo, hi)
        assert p == want_p, (name, Mp)
        assert f == want_f, (name, Mp)
        for key in ("positions", "covered", "sieved", "mr", "composite"):
            assert c[key] == want_c[key], (name, Mp, key)


@pytest.mark.parametrize("persistent,block,bits,waves", [(True, 64, 32, 1), (False, 128, 1024, 16), (True, 512, 4096, 0),
                                                           (False, 1024, 61440, 16), (True, 256, 32768, 24),
                                                           (True, 1024, 16384, 16)])
def test_launch_shapes(persistent, block, bits, waves):
    """Workgroup size, tile size, resident waves per CU and the persistent queue do not
    change a byte."""
    lo, hi = 10 ** 13, 10 ** 13 + 2 * 10 ** 9
    kw = dict(wheel=2042040, mmax=10000, sieve=1000, filters=F204)
    want, _ = oracle_record(lo, hi, kw)
    sc = scanner(**kw)
    print("shape", persistent, block, bits, waves, flush=True)
    for Mp in (840, 2042040):
        sc.set_layout(Mp)
        sc.set_launch(persistent, block, bits)
        sc.set_residency(waves)
        (rec,), _ = sc.records([(lo, hi)])
        assert rec == want, (persistent, block, bits, waves, Mp)


def test_graph_replays_match_eager():
    """The captured hipGraph of reset + sieve + MR, replayed over different batches with
    eager work in between, gives the eager records every time, and those are the
    oracle's."""
    kw = dict(wheel=2042040, mmax=10000, sieve=0, filters=F204)
    sc = scanner(**kw)
    batches = [[(10 ** 13, 10 ** 13 + 10 ** 9)], [(10 ** 15, 10 ** 15 + 2 ** 30), (10 ** 16, 10 ** 16 + 2 ** 24)],
               [(TOP - 2 ** 28, TOP)], [(10 ** 13, 10 ** 13 + 10 ** 9)]]
    eager = [sc.records(b)[0] for b in batches]
    for b, recs in zip(batches, eager):                 # the eager records are the oracle's
        assert recs == [oracle_record(lo, hi, kw)[0] for lo, hi in b]
    sc.set_graph(True)
    try:
        for rep in range(3):
            for b, want in zip(batches, eager):
                print("graph replay", rep, b, flush=True)
                assert sc.records(b)[0] == want
                sc.set_graph(False)
                assert sc.records(b)[0] == want          # eager work between replays
                sc.set_graph(True)
    finally:
        sc.set_graph(False)


@pytest.mark.parametrize("graph,depth", [(False, 2), (True, 2), (False, 4), (True, 4)])
def test_pipelined_batches_match_records(graph, depth):
    """submit/collect with up to `depth` batches in flight (the host tail of one batch
    overlapping the device work of the next ones, each slot on its own stream) returns,
    batch by batch and in submission order, the records records() gives one batch at a
    time (and the oracle gives): uneven batches, a batch with an empty window, an empty
    batch, each slot reused several times, graph on and off."""
    kw = dict(wheel=2042040, mmax=10000, sieve=0, filters=F204)
    sc = scanner(**kw)
    batches = [[(10 ** 15, 10 ** 15 + 2 ** 30)], [(10 ** 13, 10 ** 13 + 10 ** 9), (10 ** 16, 10 ** 16)],
               [(TOP - 2 ** 28, TOP)], [], [(10 ** 16, 10 ** 16 + 2 ** 24), (10 ** 12, 10 ** 12 + 2 ** 20)],
               [(10 ** 15, 10 ** 15 + 2 ** 30)]]
    want = [sc.records(b) for b in batches]
    for b, (recs, codes) in zip(batches, want):         # what they are compared with is the oracle's
        assert list(zip(recs, codes)) == [oracle_record(lo, hi, kw) for lo, hi in b]
    sc.set_graph(graph)
    try:
        for rep in range(2):
            print("pipelined", graph, depth, rep, flush=True)
            for b in batches[:depth - 1]:
                sc.submit(b)
            for i in range(len(batches)):
                if i + depth - 1 < len(batches):
                    sc.submit(batches[i + depth - 1])
                assert sc.in_flight() == min(depth, len(batches) - i)
                assert sc.collect() == want[i], (rep, i)
            assert sc.in_flight() == 0
    finally:
        sc.set_graph(False)


def test_pipelined_refusals():
    """At most MAX_IN_FLIGHT batches in flight; no collect without a submission; the
    one-shot entry points refuse while a batch is in flight and work again once it is
    collected."""
    sc = scanner(wheel=2042040, mmax=10000, sieve=0, filters=F204)
    w = [(10 ** 13, 10 ** 13 + 2 ** 24)]
    want = sc.records(w)
    with pytest.raises(es.EsError):
        sc.collect()
    for _ in range(es.MAX_IN_FLIGHT):
        sc.submit(w)
    with pytest.raises(es.EsError):
        sc.submit(w)
    with pytest.raises(es.EsError):
        sc.records(w)
    for _ in range(es.MAX_IN_FLIGHT):
        assert sc.collect() == want
    assert sc.records(w) == want


def test_pipelined_survivor_overflow_replay():
    """A batch that overflows its fresh survivor list while the next batch is already in
    flight replays on its own slot; both batches come back as records() gives them."""
    big = [(10 ** 9, 10 ** 9 + 4 * 10 ** 8)]
    small = [(10 ** 12, 10 ** 12 + 2 ** 24)]
    kw = dict(wheel=120120, mmax=60, sieve=0, filters=None)
    want_big = oracle_record(big[0][0], big[0][1], kw)
    with es.Scanner(device=0, **kw) as sc:
        sc.submit(big)
        sc.submit(small)
        (rec,), (code,) = sc.collect()
        assert sc.timing()["replays"] >= 1 and sc.timing()["survivors"] > 65536
        assert (rec, code) == want_big
        got_small = sc.collect()
        assert got_small == sc.records(small)


def test_survivor_overflow_replay():
    """More survivors than the initial device list holds: counts are complete, the
    batch replays with a larger list (scan.cu:223-242), output unchanged."""
    lo, hi = 10 ** 9, 10 ** 9 + 4 * 10 ** 8
    kw = dict(wheel=120120, mmax=60, sieve=0, filters=None)
    want, rc = oracle_record(lo, hi, kw)
    sc = scanner(**kw)
    (rec,), (code,) = sc.records([(lo, hi)])
    t = sc.timing()
    assert t["survivors"] > 65536
    assert t["replays"] >= 1
    assert rec == want and code == rc


def test_batch_equals_alone():
    """A batch of uneven windows gives each window the record it gets alone, and the
    oracle's record."""
    wins = [tuple(map(int, l.split())) for l in open(os.path.join(paths.GOLDEN, "c4_windows.txt"))
            if not l.startswith("#")][:24]
    kw = dict(wheel=2042040, mmax=10000, sieve=0, filters=F204)
    sc = scanner(**kw)
    recs, codes = sc.records(wins)
    for (lo, hi), rec in zip(wins, recs):
        print("window", lo, hi, flush=True)
        (alone,), _ = sc.records([(lo, hi)])
        assert rec == alone
        want, _ = oracle_record(lo, hi, kw)
        assert rec == want


@pytest.mark.parametrize("persistent,sieve,group", [(True, 0, 4), (True, 1000, 2), (False, 1000, 16), (True, 0, 16)])
def test_grouped_items_equal_single_tiles(persistent, sieve, group):
    """Short tiles grouped into shared work items give the records of one item per tile,
    and the oracle's: windows of 2^10 to 2^22 integers from 1e9 (32-bit tile starts) to
    1e16 (64-bit), with and without stage B', persistent or not, groups of 2 to 16."""
    kw = dict(wheel=2042040, mmax=10000, sieve=sieve, filters=F204)
    wins = []
    for i, lo in enumerate((10 ** 9 + 7, 10 ** 12 + 11, 10 ** 13 + 5, 10 ** 16 + 3)):
        for e in (10, 13, 16, 19, 22):
            base = lo + i * 2 ** 23 + e * 2 ** 23
            wins.append((base, base + 2 ** e + 17 * e))
    sc = scanner(**kw)
    sc.set_launch(persistent, 512, 32768)
    print("one tile per item", persistent, sieve, flush=True)
    sc.set_grouping(1)
    alone, codes0 = sc.records(wins)
    print("groups of up to", group, flush=True)
    sc.set_grouping(group)
    packed, codes1 = sc.records(wins)
    assert packed == alone and codes1 == codes0
    for (lo, hi), rec in zip(wins, packed):
        want, _ = oracle_record(lo, hi, kw)
        assert rec == want, (lo, hi)


# ---------------------------------------------------------------- stage D on the device

GOLDEN_LINE_FILES = ["w19_certificates.txt", "w64_head.txt", "w64_tail.txt", "s0_head.txt", "s0_tail.txt",
                     "s69_head.txt", "s69_tail.txt"]


@pytest.mark.parametrize("fn", GOLDEN_LINE_FILES)
def test_device_stage_d_reproduces_golden_lines(fn):
    """k_rung (one wave per prime) re-derives every certificate line the reference
    published, byte for byte: smallest rung, then smallest witness."""
    want = [l for l in open(os.path.join(paths.GOLDEN, "rederive", fn)) if l.startswith("RUNG")]
    primes = [int(l.split()[1]) for l in want]
    sc = scanner(filters=F204)
    got, on_host = sc.certify_batch(primes)
    print(fn, len(primes), "primes,", on_host, "finished on the host", flush=True)
    assert got == want
    assert on_host <= len(primes) // 10


def test_device_tail_on_and_off_agree():
    """The same records whether stage D runs on the device or the host."""
    kw = dict(wheel=120120, mmax=60, sieve=0, filters=None)
    lo, hi = 10 ** 9, 10 ** 9 + 10 ** 8
    sc = scanner(**kw)
    want = paths.golden_record("mmax60_w120").decode()
    try:
        for tail in (True, False):
            sc.set_device_tail(tail)
            (rec,), _ = sc.records([(lo, hi)])
            assert rec == want, tail
            print("tail", tail, sc.timing(), flush=True)
    finally:
        sc.set_device_tail(True)


# ---------------------------------------------------------------- Miller-Rabin

def test_primality_dense_and_kats():
    """Device MR against the oracle's test, exhaustively over dense odd ranges at 10^12,
    2^63 and the ceiling, plus strong pseudoprimes (sieve_cuda --check-primality)."""
    import ctypes
    olib = ctypes.CDLL(paths.ORACLE_LIB)
    olib.es_oracle_is_prime.argtypes = [ctypes.c_uint64]
    sc = scanner(filters=F204)
    vals = []
    for base in (10 ** 12 + 1, 2 ** 63 - 100001, TOP - 200001, 3):
        vals += [base + 2 * i for i in range(100000)]
    vals += [2047, 3277, 4033, 4681, 8321, 3215031751, 2152302898747, 3474749660383, 341550071728321,
             3825123056546413051, 18446744073709551557, 18446744065119614953, 2, 1, 0, 4, 9, 25]
    got = sc.primality(vals)
    want = [olib.es_oracle_is_prime(v) for v in vals]
    bad = [v for v, g, w in zip(vals, got, want) if g != w]
    assert not bad, bad[:10]


W19 = os.path.join(paths.GOLDEN, "rederive", "w19_certificates.txt")
W19_LO = 5 * 10 ** 18
W19_HI = W19_LO + 10 ** 12


def test_round_major_order_reproduces_published_window():
    """[5e18, 5e18 + 1e12) in round-major order (seg 262144): the record (8413 certificate
    lines in order, then SUMMARY) equals the published stream of that window byte for
    byte. The window spans two rounds, so the round key is exercised, not only the lane
    key; lane order gives the same lines in another order."""
    want = open(W19).read()
    with es.Scanner(wheel=2042040, filters=F204, device=0) as sc:
        sc.set_order("round", 262144)
        (rec,), (ec,) = sc.records([(W19_LO, W19_HI)])
        assert rec == want and ec == 0
        sc.set_order("lane")
        (lane,), _ = sc.records([(W19_LO, W19_HI)])
        assert lane != want and sorted(lane.splitlines()) == sorted(want.splitlines())


SEG = 262144
M204 = 2042040
SHARD_EDGES = [  # published re-runs of production shards: (name, lo, hi)
    ("s0", 255255, 40000000000254234),
    ("s69", 15590940119293459447, 15631162710079743316),
]


@pytest.mark.parametrize("name,lo,hi", SHARD_EDGES, ids=[e[0] for e in SHARD_EDGES])
def test_published_shard_edges(name, lo, hi):
    """The first and last 2,000 certificate lines of two published production shards
    (round-major, seg 262144, as sieve_cuda emitted them). Round r holds positions
    J in [r*seg, (r+1)*seg) of every lane, i.e. the integers [lo + r*seg*M, lo + (r+1)*seg*M),
    so the head is round 0 scanned alone and the tail is the last two rounds scanned from
    their boundary: the sieve, stage D and the round key all have to agree with the
    reference for the lines to come out identical."""
    head = open(os.path.join(paths.GOLDEN, "rederive", "%s_head.txt" % name)).read().splitlines(keepends=True)
    tail = open(os.path.join(paths.GOLDEN, "rederive", "%s_tail.txt" % name)).read().splitlines(keepends=True)
    last_round = ((hi - lo + M204 - 1) // M204 - 1) // SEG
    with es.Scanner(wheel=M204, filters=F204, device=0) as sc:
        sc.set_order("round", SEG)
        print(name, "head", flush=True)
        (rec,), _ = sc.records([(lo, min(hi, lo + SEG * M204))])
        assert rec.splitlines(keepends=True)[:2000] == head
        start = lo + max(0, last_round - 1) * SEG * M204
        print(name, "tail from", start, flush=True)
        (rec,), _ = sc.records([(start, hi)])
        lines = rec.splitlines(keepends=True)
        assert lines[-1].startswith("SUMMARY ")
        assert lines[-2001:-1] == tail


def test_cli_round_order_across_chunks():
    """The CLI cuts the window into chunks and merges by the key of the whole interval."""
    argv = [str(W19_LO), str(W19_HI), "--filters", os.path.relpath(F204, paths.REPO), "--order", "round",
            "--chunk", str(1 << 37)]
    rc, out, err = paths.run(paths.SIEVE_BIN, argv, timeout=300)
    assert rc == 0, err.decode()[-2000:]
    assert out.decode() == open(W19).read()


def test_cli_lane_order_across_chunks():
    """Lane order holds every chunk's lines to the end: the chunked CLI run (eight
    chunks, pipelined) gives the record of the whole window in one batch."""
    with es.Scanner(wheel=2042040, filters=F204, device=0) as sc:
        (want,), (ec,) = sc.records([(W19_LO, W19_HI)])
    argv = [str(W19_LO), str(W19_HI), "--filters", os.path.relpath(F204, paths.REPO), "--chunk", str(1 << 37)]
    rc, out, err = paths.run(paths.SIEVE_BIN, argv, timeout=300)
    assert rc == ec, err.decode()[-2000:]
    assert out.decode() == want


def test_top_of_range_window_reproduces_published_digests():
    """[E - 1e15, E), E = the stage-E ceiling: the drop-in binary in round-major order
    re-derives the published window -- 8,284,620 certificates (Montgomery carry path:
    every survivor exceeds 2^63) whose stdout sha256, certificate-only sha256, SUMMARY
    and 2000-line head and tail all match the reference's files. Streams the output
    through the hashes rather than holding it."""
    import hashlib
    e = 18446744065119614976
    argv = [paths.SIEVE_BIN, str(e - 10 ** 15), str(e), "--filters", os.path.relpath(F204, paths.REPO),
            "--order", "round"]
    want = dict(l.split()[::-1] for l in open(os.path.join(paths.GOLDEN, "rederive", "w64.sha256")))
    head_want = open(os.path.join(paths.GOLDEN, "rederive", "w64_head.txt")).readlines()
    tail_want = open(os.path.join(paths.GOLDEN, "rederive", "w64_tail.txt")).readlines()
    summary_want = open(os.path.join(paths.GOLDEN, "rederive", "w64_summary.txt")).readline()
    p = subprocess.Popen(argv, cwd=paths.REPO, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL)
    full, certs = hashlib.sha256(), hashlib.sha256()
    head, tail, summary, buf, n = [], [], None, b"", 0
    while True:
        chunk = p.stdout.read(1 << 20)
        if not chunk:
            break
        full.update(chunk)
        parts = (buf + chunk).split(b"\n")
        buf = parts.pop()
        for l in parts:
            l = l.decode() + "\n"
            if l.startswith("SUMMARY "):
                summary = l
                continue
            certs.update(l.encode())
            n += 1
            if len(head) < len(head_want):
                head.append(l)
            tail.append(l)
            if len(tail) > 2 * len(tail_want):
                del tail[:len(tail_want)]
    assert p.wait(timeout=600) == 0
    assert n == 8284620 and summary == summary_want
    assert head == head_want and tail[-len(tail_want):] == tail_want
    assert full.hexdigest() == want["w64_full_stdout.txt"]
    assert certs.hexdigest() == want["w64_certificates_only"]
