"""Every certificate the HIP path emits holds in exact arithmetic.

The records come from the GPU (device stage D included) and go, unchanged, through the
build's exact verifier (erdos-strauss_amd/verify_certs.py, the validity rules of the
reference's sieve/verify_covers.py:246-275) as a separate process with --lo/--hi, which
also checks that each certified p is prime, lies in the window and is 1 mod 24. The
verifier reads the records as text, so it checks what a user of the scanner receives.
Windows: the height-1e15 and top-of-range benchmark intervals, the 5e18 window in
round-major order, the first 2^36 integers of the top-of-range window [E - 1e15, E), and
the 512 batched windows on both wheels.
"""
import importlib.util
import os
import subprocess
import sys

import pytest

from tests import paths

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
es = importlib.util.module_from_spec(spec)
spec.loader.exec_module(es)

TOP = 18446744065119614976
VERIFY = os.path.join(paths.PKG, "verify_certs.py")

CASES = [
    ("c2", 10 ** 15, 10 ** 15 + 2 ** 34, "lane"),
    ("c3", TOP - 2 ** 34, TOP, "lane"),
    ("w19_round", 5 * 10 ** 18, 5 * 10 ** 18 + 10 ** 12, "round"),
    ("top_head", TOP - 10 ** 15, TOP - 10 ** 15 + 2 ** 36, "lane"),
]


def verify(path, lo, hi):
    p = subprocess.run([sys.executable, VERIFY, "--lo", str(lo), "--hi", str(hi), path], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=600)
    return p.returncode, p.stdout.decode()


def summary(rec):
    last = rec.strip().splitlines()[-1].split()
    return {t.split("=")[0]: t.split("=")[1] for t in last[1:] if "=" in t}


@pytest.mark.parametrize("name,lo,hi,order", CASES, ids=[c[0] for c in CASES])
def test_gpu_certificates_verify_exactly(tmp_path, name, lo, hi, order):
    with es.Scanner(wheel=2042040, filters=es.F204, device=0) as sc:
        sc.set_order(order, 262144 if order == "round" else 0)
        (rec,), (code,) = sc.records([(lo, hi)])
    s = summary(rec)
    certs = sum(1 for l in rec.splitlines() if l.startswith(("RUNG ", "SOLVED ")))
    assert certs == int(s["rung"]) + int(s["direct"]) > 0
    assert code == 0 and s["survivors"] == "0"
    path = tmp_path / (name + ".txt")
    path.write_text(rec)
    rc, out = verify(str(path), lo, hi)
    print(name, out.strip(), flush=True)
    assert rc == 0, out
    assert out.strip() == "verified: %d RUNG, %s SOLVED; 0 bad, 0 survivors" % (int(s["rung"]), s["direct"])


@pytest.mark.parametrize("M", [2042040, 120120])
def test_gpu_batched_certificates_verify_exactly(tmp_path, M):
    """Each window's lines against its own band (the verifier's functions in this
    process), then all of them at once through the verifier's command line."""
    vspec = importlib.util.spec_from_file_location("verify_certs", VERIFY)
    vc = importlib.util.module_from_spec(vspec)
    vspec.loader.exec_module(vc)
    wins = [tuple(map(int, l.split())) for l in open(os.path.join(paths.GOLDEN, "c4_windows.txt"))
            if not l.startswith("#")]
    F = es.F204 if M == 2042040 else es.F120
    with es.Scanner(wheel=M, filters=F, device=0) as sc:
        recs, _ = sc.records(wins)
    total = 0
    for (lo, hi), rec in zip(wins, recs):
        nr, ns, bad, surv, problems = vc.verify_lines(rec.splitlines(), lo, hi)
        assert bad == 0 and surv == 0, (lo, hi, problems[:3])
        assert nr + ns == int(summary(rec)["rung"]) + int(summary(rec)["direct"])
        total += nr
    path = tmp_path / ("batch_w%d.txt" % M)
    path.write_text("".join(recs))
    rc, out = verify(str(path), min(w[0] for w in wins), max(w[1] for w in wins))
    print("wheel", M, out.strip(), flush=True)
    assert rc == 0, out
    assert out.strip().startswith("verified: %d RUNG," % total)
    gold = open(os.path.join(paths.GOLDEN, "c4", "records_w%d.txt" % M)).read()
    assert total == sum(1 for l in gold.splitlines() if l.startswith("RUNG ")) > 0
