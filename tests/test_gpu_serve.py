"""GPU: the continuous-batching request server.

Bar: bit-exact. Every request's record must equal the golden record of the same
interval (the first requests of the Poisson request trace, from the oracle, itself
cross-checked against the reference scanner), whether the request runs alone, in the
mix of an offline batch, under a real-time replay of its arrival times, behind an
admission budget that forces queueing, or cut into slices that span many batches.
"""
import importlib.util
import os

import pytest

from tests import paths

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
es = importlib.util.module_from_spec(spec)
spec.loader.exec_module(es)

GOLDEN = os.path.join(paths.GOLDEN, "c5", "records_first400.txt")
N = 160          # requests per mix (about 0.8 s of the trace, ~1.2e11 integers)


def split_records(text):
    out, cur = [], []
    for line in text.splitlines(keepends=True):
        cur.append(line)
        if line.startswith("SUMMARY "):
            out.append("".join(cur))
            cur = []
    return out


def trace(n):
    rows = [l.split() for l in open(os.path.join(paths.GOLDEN, "c5_trace.txt")) if not l.startswith("#")]
    return [(float(a), int(lo), int(hi)) for a, lo, hi in rows[:n]]


@pytest.fixture(scope="module")
def golden():
    return split_records(open(GOLDEN).read())


@pytest.fixture(scope="module")
def sc():
    s = es.Scanner(wheel=2042040, filters=es.F204, device=0)
    yield s
    s.close()


def test_each_request_alone_matches_golden(sc, golden):
    reqs = trace(N)
    recs, codes = sc.records([(lo, hi) for _, lo, hi in reqs])
    assert recs == golden[:N]
    for r, c in zip(recs, codes):
        assert c == (2 if "\nSURVIVOR " in "\n" + r else 0)
    for k, (_, lo, hi) in enumerate(reqs[:24]):     # and one request per call
        (rec,), _ = sc.records([(lo, hi)])
        assert rec == golden[k]


def test_offline_mix_matches_golden(sc, golden):
    recs, infos, st = es.serve_trace(sc, trace(N), realtime=False)
    assert recs == golden[:N]
    assert st["completed"] == N and st["refused"] == 0
    assert st["max_batch_requests"] > 1          # requests shared device batches
    assert all(i["exit_code"] in (0, 2) and i["done_ms"] >= i["arrival_ms"] for i in infos)
    assert st["peak_reserved"] <= st["budget"] - st["resident"]


def test_realtime_replay_matches_golden(sc, golden):
    reqs = trace(N)
    recs, infos, st = es.serve_trace(sc, reqs, realtime=True)
    assert recs == golden[:N]
    # Arrivals follow the trace's clock (submission is paced, not batched up front).
    span = infos[-1]["arrival_ms"] - infos[0]["arrival_ms"]
    assert span >= 0.9 * (reqs[-1][0] - reqs[0][0]) * 1e3
    assert st["wall_ms"] > 0 and st["ints"] == sum(hi - lo for _, lo, hi in reqs)


def test_admission_budget_forces_queueing(sc, golden):
    # A one-byte budget admits one request at a time: the queue builds up behind it and
    # every record is still the golden one.
    recs, infos, st = es.serve_trace(sc, trace(40), realtime=False, hbm_budget=1)
    assert recs == golden[:40]
    assert st["max_active"] == 1 and st["max_queued"] > 1
    assert max(i["admit_ms"] - i["arrival_ms"] for i in infos) > 0


def test_slices_spanning_many_batches(sc, golden):
    recs, infos, st = es.serve_trace(sc, trace(60), realtime=False, slice=1 << 22, batch_ints=1 << 26,
                                     batch_windows=64)
    assert recs == golden[:60]
    assert max(i["batches"] for i in infos) > 1
    assert st["slices"] > 60


def test_streaming_api_refusal_and_empty(sc, golden):
    reqs = trace(12)
    with es.Server(sc) as srv:
        ids = [srv.submit(lo, hi) for _, lo, hi in reqs]
        bad = srv.submit(18446744065119614976 - 10, 18446744065119614977)   # above the stage-E ceiling
        empty = srv.submit(10 ** 12, 10 ** 12)
        for k, rid in enumerate(ids):
            rec, info = srv.wait(rid)
            assert rec == golden[k]
        rec, info = srv.wait(bad)
        assert info["exit_code"] == 1 and "FATAL" in rec
        rec, info = srv.wait(empty)
        assert info["exit_code"] == 0 and rec.startswith("SUMMARY lo=1000000000000 hi=1000000000000 ")
        assert "positions=0 " in rec
        with pytest.raises(es.EsError):
            srv.wait(10 ** 9, timeout_ms=1)             # unknown id
        st = srv.stats()
        assert st["completed"] == len(ids) + 2 and st["refused"] == 1


def test_realtime_first_20s_every_record_matches_digest(sc):
    """The first 20 s of the trace (about 3,900 requests, 2.6e12 integers) replayed in real
    time through the continuous-batching server: every request's record hashes to the
    committed sha256 of the oracle's record for that request (tests/golden/c5/
    records_sha256.txt, 1,018 of whose requests were also run through the reference
    scanner), with requests overlapping on the device."""
    import hashlib
    rows = [l.split() for l in open(os.path.join(paths.GOLDEN, "c5_trace.txt")) if not l.startswith("#")]
    digests = [l.split() for l in open(os.path.join(paths.GOLDEN, "c5", "records_sha256.txt"))
               if not l.startswith("#")]
    reqs = [(float(a), int(lo), int(hi)) for a, lo, hi in rows if float(a) < 20.0]
    assert len(reqs) > 3000
    print("replaying", len(reqs), "requests in real time", flush=True)
    recs, infos, st = es.serve_trace(sc, reqs, realtime=True, hbm_fraction=0.8)
    bad = [i for i, r in enumerate(recs)
           if hashlib.sha256(r.encode()).hexdigest() != digests[i][0] or int(digests[i][1]) != reqs[i][1]]
    print("requests", len(recs), "differing", len(bad), "batches", st["batches"], "max active", st["max_active"],
          flush=True)
    assert not bad, bad[:5]
    assert st["completed"] == len(reqs) and st["max_batch_requests"] > 1
