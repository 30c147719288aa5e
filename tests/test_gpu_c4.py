"""GPU parity of the batched configuration exactly as it is timed.

512 windows (seed 0, start log-uniform over [1e9, 1.8e19], length log-uniform over
[2^16, 2^26]) for each wheel: the autotuner chooses the sub-wheel layout per window size
and the launch shape (workgroup and tile size) on a sample of the batch; all 512 windows
then go through one persistent work-queue launch replayed from a captured hipGraph, and
through submit/collect with eight batches in flight, each slot's graph replayed several
times. Every window's record must equal the golden record committed for that wheel
(tests/golden/c4/records_w<M>.txt: the oracle's records, cross-checked against the
reference binary on every window below 1e15), byte for byte, with its exit code.
"""
import importlib.util
import os

import pytest

from tests import paths

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
es = importlib.util.module_from_spec(spec)
spec.loader.exec_module(es)


def split_records(text):
    out, cur = [], []
    for line in text.splitlines(keepends=True):
        cur.append(line)
        if line.startswith("SUMMARY "):
            out.append("".join(cur))
            cur = []
    return out


def windows():
    return [tuple(map(int, l.split())) for l in open(os.path.join(paths.GOLDEN, "c4_windows.txt"))
            if not l.startswith("#")]


def expected_code(rec):
    return 0 if rec.rstrip().endswith(" survivors=0") else 2


@pytest.mark.parametrize("M", [2042040, 120120])
def test_c4_all_windows_autotuned_graph_in_flight(M):
    wins = windows()
    assert len(wins) == 512
    gold = split_records(open(os.path.join(paths.GOLDEN, "c4", "records_w%d.txt" % M)).read())
    assert len(gold) == 512
    codes_want = [expected_code(r) for r in gold]
    F = es.F204 if M == 2042040 else es.F120
    with es.Scanner(wheel=M, filters=F, device=0) as sc:
        report = sc.autotune(wins[:64])
        choices = [l for l in report.splitlines() if l.startswith("choice ")]
        print("wheel", M, "autotuner:", " | ".join(choices), flush=True)
        assert any(l.startswith("choice block=") for l in choices)
        layouts = {sc.choose_layout(hi - lo) for lo, hi in wins}
        print("wheel", M, "layouts used by the 512 windows:", sorted(layouts), flush=True)
        sc.set_graph(True)
        try:
            # One persistent launch for the whole batch (the first call captures the graph,
            # the second replays it).
            for rep in range(2):
                recs, codes = sc.records(wins)
                bad = [i for i in range(512) if recs[i] != gold[i]]
                assert not bad, ("records", M, rep, bad[:5])
                assert codes == codes_want
            t = sc.timing()
            print("wheel", M, "one batch:", t, flush=True)
            assert t["tiles"] > 0 and t["replays"] == 0
            # Eight batches in flight, each slot's graph replayed: sixteen passes.
            depth, passes = es.MAX_IN_FLIGHT, 2 * es.MAX_IN_FLIGHT
            for _ in range(depth - 1):
                sc.submit(wins)
            for i in range(passes):
                if i + depth - 1 < passes:
                    sc.submit(wins)
                assert sc.in_flight() == min(depth, passes - i)
                recs, codes = sc.collect()
                bad = [k for k in range(512) if recs[k] != gold[k]]
                assert not bad, ("in flight", M, i, bad[:5])
                assert codes == codes_want
            assert sc.in_flight() == 0
        finally:
            sc.set_graph(False)


def test_c4_results_complete_when_posted_over_many_passes():
    """Regression: a batch's completion post must follow every result the host reads.
    k_pack's mirrors were once pinned memory the device could cache in L2; each XCD has
    its own L2 and the posting workgroup's release wrote back only its own, so about one
    pipelined pass in 200 handed the host a stale survivor flag or stage-D result. With
    coherent mirrors and a system-scope release per workgroup, 1,200 passes eight in flight
    and 300 one at a time must all equal the golden records."""
    wins = windows()
    gold = split_records(open(os.path.join(paths.GOLDEN, "c4", "records_w2042040.txt")).read())
    with es.Scanner(wheel=2042040, filters=es.F204, device=0) as sc:
        sc.autotune(wins[:64])
        sc.set_graph(True)
        depth, passes, bad = es.MAX_IN_FLIGHT, 1200, []
        for _ in range(depth - 1):
            sc.submit(wins)
        for i in range(passes):
            if i + depth - 1 < passes:
                sc.submit(wins)
            recs, _ = sc.collect()
            if recs != gold:
                bad.append(("in flight", i, [k for k in range(512) if recs[k] != gold[k]][:3]))
        for i in range(300):
            recs, _ = sc.records(wins)
            if recs != gold:
                bad.append(("one at a time", i, [k for k in range(512) if recs[k] != gold[k]][:3]))
        print("passes differing:", len(bad), flush=True)
        assert not bad, bad[:5]
