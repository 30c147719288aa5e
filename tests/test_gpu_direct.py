"""GPU: stage E (the direct F1 instance, rung_scan3.cpp:607-647) on the device.

Bar: bit-exact. es_direct_batch runs k_rung with no rungs, so every prime goes through
the workgroup's stage-E phase (the same code a scan runs for a prime no rung certifies).
Its lines must equal the reference's own --check-direct rows (tests/golden/direct/, made
by rung_scan3 from the reference's sources) and the oracle's stage E on primes at every
magnitude up to the ceiling, and pass the exact verifier.
"""
import glob
import importlib.util
import os
import random
import subprocess

import pytest

from tests import paths

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
es = importlib.util.module_from_spec(spec)
spec.loader.exec_module(es)
vspec = importlib.util.spec_from_file_location("verify_certs", os.path.join(paths.PKG, "verify_certs.py"))
vc = importlib.util.module_from_spec(vspec)
vspec.loader.exec_module(vc)

CEIL = 18446744065119617025     # (2^32 - 1)^2
BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n):
    if n < 2:
        return False
    for q in BASES:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in BASES:
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def primes_from(lo, count):
    out, p = [], lo | 1
    while len(out) < count:
        if is_prime(p):
            out.append(p)
        p += 2
    return out


def oracle_lines(ps):
    out = []
    for i in range(0, len(ps), 256):
        out += subprocess.run([paths.ORACLE_BIN, "--solve-direct"] + [str(p) for p in ps[i:i + 256]],
                              stdout=subprocess.PIPE, check=True).stdout.decode().splitlines(keepends=True)
    return out


@pytest.fixture(scope="module")
def sc():
    s = es.Scanner(wheel=2042040, filters=es.F204, device=0)
    yield s
    s.close()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(paths.GOLDEN, "direct", "*.txt"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_device_stage_e_matches_reference_check_direct_rows(sc, path):
    rows = open(path).read().splitlines(keepends=True)
    ps = [int(l.split()[1]) for l in rows]
    print("stage E on the device:", os.path.basename(path), len(ps), "primes", flush=True)
    lines, host = sc.direct_batch(ps)
    assert lines == rows
    assert host == 0


def test_device_stage_e_matches_oracle_at_every_magnitude(sc):
    rng = random.Random(7)
    starts = [10 ** 6 + 3, 10 ** 9, 2 ** 40 + 11, 10 ** 15, 2 ** 50 - 999, 2 ** 52, 10 ** 18, 2 ** 62 + 1,
              2 ** 63 - 2001, 2 ** 63 + 1, 10 ** 19, CEIL - 100000]
    starts += [rng.randrange(10 ** 7, CEIL - 10 ** 6) for _ in range(8)]
    ps = []
    for lo in starts:
        ps += primes_from(lo, 48)
    # Above the ceiling every pair is skipped: no line certifies these.
    ps += primes_from(CEIL + 2, 4)
    print("stage E on the device:", len(ps), "primes", flush=True)
    lines, host = sc.direct_batch(ps)
    want = [l if not l.startswith("DECLINED") else "SURVIVOR %s\n" % l.split()[1] for l in oracle_lines(ps)]
    assert lines == want
    assert sum(l.startswith("SOLVED") for l in lines) >= len(ps) - 4
    assert all(l.startswith("SURVIVOR") for l in lines[-4:])
    tot = vc.verify_lines([l for l in lines if l.startswith("SOLVED")])
    assert tot[2] == 0, tot[4][:5]


def test_device_stage_e_lines_pass_the_exact_verifier(sc):
    ps = primes_from(10 ** 12, 64) + primes_from(2 ** 63 + 1, 64) + primes_from(CEIL - 10 ** 6, 64)
    lines, _ = sc.direct_batch(ps)
    assert len(lines) == len(ps)
    tot = vc.verify_lines(lines)
    print("verified: %d SOLVED, %d bad, %d survivors" % (tot[1], tot[2], tot[3]), flush=True)
    assert tot[0] == 0 and tot[2] == 0 and tot[1] + tot[3] == len(ps) and tot[1] >= len(ps) - 2
