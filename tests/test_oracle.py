"""Pins the CPU oracle (oracle/es_oracle) before anything is compared against it.

Every check here is against something the oracle did not produce: the reference
scanner's own output (records generated from rung_scan3 built from source), the
reference's committed golden records (data/rederive), its certified class tables,
its --check-direct rows, a sieve of Eratosthenes, known strong pseudoprimes, and the
exact-rational identity every certificate must satisfy.
"""
import ctypes
import hashlib
import os
import random

import pytest

from tests import paths

pytestmark = pytest.mark.skipif(not os.path.exists(paths.ORACLE_BIN),
                                reason="oracle not built (run __graft_entry__.build())")


def _oracle_lib():
    lib = ctypes.CDLL(paths.ORACLE_LIB)
    lib.es_oracle_is_prime.argtypes = [ctypes.c_uint64]
    lib.es_oracle_sprp2.argtypes = [ctypes.c_uint64]
    lib.es_oracle_solve.argtypes = [ctypes.c_uint64, ctypes.c_char_p, ctypes.c_int]
    lib.es_oracle_factor.argtypes = [ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.POINTER(ctypes.c_int)]
    return lib


def test_self_test_kats():
    """rung_scan3.cpp:821-897: wheel(840) = the six hard classes, 920/220/15014 at
    120120, 2308 classes and both pinned fnv64 digests."""
    rc, out, _ = paths.run(paths.ORACLE_BIN, ["--verify"])
    assert rc == 0, out.decode()
    assert b"SELF-TEST PASSED" in out


@pytest.mark.parametrize("M", [120120, 2042040])
def test_class_table_byte_identical(M):
    """--dump-classes reproduces the reference's certified class table (sha256)."""
    want = {}
    with open(os.path.join(paths.GOLDEN, "class_tables.sha256")) as f:
        for line in f:
            h, m = line.split()
            want[int(m)] = h
    rc, out, _ = paths.run(paths.ORACLE_BIN, ["--dump-classes", "--wheel", str(M)])
    assert rc == 0
    assert hashlib.sha256(out).hexdigest() == want[M]


FAST_RECORDS = [c for c in paths.record_cases()]


@pytest.mark.parametrize("name,argv,rc", FAST_RECORDS, ids=[c[0] for c in FAST_RECORDS])
def test_record_matches_reference(name, argv, rc):
    """Byte-identical stdout and exit code against rung_scan3 built from source."""
    got_rc, out, err = paths.run(paths.ORACLE_BIN, argv + ["--threads", "4"])
    assert got_rc == rc, err.decode()
    assert out == paths.golden_record(name)


@pytest.mark.parametrize("k", range(6, 20))
def test_decade_golden_records(k):
    """The reference's own full records data/rederive/decades/d6..d19 (lane order)."""
    want = open(os.path.join(paths.GOLDEN, "rederive", "d%d.txt" % k), "rb").read()
    summ = want.decode().strip().splitlines()[-1].split()
    lo = [t for t in summ if t.startswith("lo=")][0][3:]
    hi = [t for t in summ if t.startswith("hi=")][0][3:]
    rc, out, _ = paths.run(paths.ORACLE_BIN, [lo, hi, "--filters",
                                              os.path.join(paths.FILTERS, "f_M2042040_mmax2000_gpu.txt")])
    assert rc == 0
    assert out == want


GOLDEN_LINE_FILES = ["w19_certificates.txt", "w64_head.txt", "w64_tail.txt", "s0_head.txt",
                     "s0_tail.txt", "s69_head.txt", "s69_tail.txt"]


@pytest.mark.parametrize("fn", GOLDEN_LINE_FILES)
def test_stage_d_reproduces_golden_lines(fn):
    """Every certificate line the reference published re-solves to the identical line:
    smallest rung, then smallest witness (rung_scan3.cpp:574-603)."""
    lib = _oracle_lib()
    buf = ctypes.create_string_buffer(256)
    lines = [l for l in open(os.path.join(paths.GOLDEN, "rederive", fn)) if l.startswith(("RUNG", "SOLVED"))]
    assert lines
    for line in lines:
        p = int(line.split()[1])
        lib.es_oracle_solve(p, buf, 256)
        assert buf.value.decode() == line, line


@pytest.mark.parametrize("lo", [1000003, 9223372036854770000, 9223372036854775808,
                                9999999999999990000, 18446744065119000000])
def test_stage_e_rows(lo):
    """The reference's --check-direct rows either side of 2^63 and under the ceiling."""
    want = open(os.path.join(paths.GOLDEN, "direct", "%d.txt" % lo)).read().splitlines()
    primes = [l.split()[1] for l in want]
    rc, out, _ = paths.run(paths.ORACLE_BIN, ["--solve-direct"] + primes)
    assert rc == 0
    assert out.decode().splitlines() == want


def test_primality_against_sieve():
    """is_prime and sprp2 against Eratosthenes on [0, 300000) (a true function)."""
    lib = _oracle_lib()
    N = 300000
    comp = bytearray(N)
    comp[0] = comp[1] = 1
    for i in range(2, int(N ** 0.5) + 1):
        if not comp[i]:
            comp[i * i::i] = b"\x01" * len(range(i * i, N, i))
    for n in range(N):
        want = not comp[n]
        assert bool(lib.es_oracle_is_prime(n)) == want, n
        if want:
            assert lib.es_oracle_sprp2(n), n


# Composites that fool weaker tests: base-2 strong pseudoprimes, Carmichael numbers, and
# 3825123056546413051 (strong pseudoprime to every prime base up to 23).
PSEUDO = [2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281, 74665,
          80581, 85489, 88357, 90751, 561, 1105, 1729, 2465, 2821, 6601, 8911,
          3215031751, 2152302898747, 3474749660383, 341550071728321, 3825123056546413051]
PRIMES = [2, 3, 37, 41, 65537, 4294967291, 4294967311, 2305843009213693951,
          9223372036854775783, 18446744073709551557, 18446744065119614953]


def test_primality_kats():
    lib = _oracle_lib()
    for n in PSEUDO:
        assert not lib.es_oracle_is_prime(n), n
    for n in PRIMES:
        assert lib.es_oracle_is_prime(n), n
    assert lib.es_oracle_sprp2(2047) and lib.es_oracle_sprp2(3825123056546413051)


def test_factorization_invariants():
    """Product reconstructs n and every factor is prime, on the reference's adversarial
    shapes (rung_scan3.cpp:459-503): hard semiprimes across 2^63, prime squares, smooth
    numbers, the 2^32 and ceiling boundaries."""
    lib = _oracle_lib()
    pr = (ctypes.c_uint64 * 32)()
    ex = (ctypes.c_int * 32)()
    rng = random.Random(0x9E3779B97F4A7C15)

    def next_prime(a):
        while not lib.es_oracle_is_prime(a):
            a += 1
        return a

    vals = [2 ** 32 - 1, 2 ** 32, 2 ** 32 + 1, 2 ** 63 - 1, 2 ** 63, 2 ** 63 + 1,
            18446744065119617023, 18446744065119617024, 18446744065119617025, 1, 2]
    for _ in range(60):
        a = next_prime(3037000000 + rng.randrange(1257000000))
        b = next_prime(2147483648 + rng.randrange(2147000000))
        if a * b < 2 ** 64:
            vals.append(a * b)
        c = next_prime(2147483648 + rng.randrange(2147000000))
        vals.append(c * c)
        v = 1
        while v < 18446744065119617025 // 65536:
            v *= 2 + rng.randrange(65535)
        vals.append(v)
        vals.append(rng.randrange(10 ** 6, 18446744065119617025))
    for n in vals:
        k = lib.es_oracle_factor(n, pr, ex)
        prod = 1
        for i in range(k):
            assert lib.es_oracle_is_prime(pr[i]), (n, pr[i])
            prod *= pr[i] ** ex[i]
        assert prod == n or (n == 0 and k == 0), n


def test_planted_mistake_is_caught():
    """A record comparison must be able to fail: perturb one digit of a golden line and
    one SUMMARY counter, and check the byte comparison and the exact verifier object."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("verify_certs", os.path.join(paths.PKG, "verify_certs.py"))
    vc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(vc)
    good = paths.golden_record("c2").decode().splitlines()
    _, _, bad, surv, _ = vc.verify_lines(good)
    assert bad == 0 and surv == 0
    t = good[0].split()
    t[4] = str(int(t[4]) + 1)                     # a wrong witness u
    _, _, bad, _, _ = vc.verify_lines([" ".join(t)])
    assert bad == 1
    t = good[0].split()
    t[2] = str(int(t[2]) + 1)                     # a wrong A
    _, _, bad, _, _ = vc.verify_lines([" ".join(t)])
    assert bad == 1


def test_exact_verifier_accepts_every_published_certificate():
    """Every RUNG line of the reference's published records (decades d6..d19, the 5e18
    window, the top-of-range edges) and every SOLVED row of its --check-direct battery
    holds as an exact rational identity, with p prime and p == 1 (mod 24) -- the
    build's own verifier against the reference's outputs."""
    import glob
    import importlib.util
    spec = importlib.util.spec_from_file_location("verify_certs", os.path.join(paths.PKG, "verify_certs.py"))
    vc = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(vc)
    files = sorted(glob.glob(os.path.join(paths.GOLDEN, "rederive", "d*.txt")))
    files += [os.path.join(paths.GOLDEN, "rederive", f) for f in
              ("w19_certificates.txt", "w64_head.txt", "w64_tail.txt", "s0_head.txt", "s69_tail.txt")]
    total = 0
    for f in files:
        nr, ns, bad, surv, problems = vc.verify_lines(open(f).read().splitlines(), lo=0, hi=1 << 64)
        assert bad == 0 and surv == 0, (f, problems[:3])
        total += nr
    solved = 0
    for f in glob.glob(os.path.join(paths.GOLDEN, "direct", "*.txt")):
        nr, ns, bad, surv, problems = vc.verify_lines(open(f).read().splitlines())
        assert bad == 0, (f, problems[:3])
        solved += ns
    assert total > 17000 and solved > 100
