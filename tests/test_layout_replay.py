"""CPU replay of the marking kernel's tile arithmetic over the product's layout tables,
compared with the oracle's stage-C survivor set and counters, for every sub-wheel
layout. Catches layout, mask and offset mistakes without a GPU; the GPU parity suite
then checks the kernel itself against the same oracle."""
import os
import subprocess

import pytest

from tests import paths

HERE = os.path.dirname(os.path.abspath(__file__))
REPLAY_SRC = os.path.join(HERE, "replay", "kernel_replay.cpp")
CSRC = os.path.join(paths.PKG, "csrc")
F120 = os.path.join(paths.FILTERS, "f_M120120_mmax2000_runtime.txt")
F204 = os.path.join(paths.FILTERS, "f_M2042040_mmax2000_gpu.txt")
TOP = 18446744065119614976


@pytest.fixture(scope="module")
def replay_bin(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("replay") / "kernel_replay")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, REPLAY_SRC, os.path.join(CSRC, "es_plan.cpp"),
                    "-o", out], check=True)
    return out


def candidates(M):
    q = M // 840
    return sorted(M // d for d in range(1, q + 1) if q % d == 0)


CASES = [
    ("w120_c1", 255255, 255255 + 3 * 2 ** 20, 120120, 10000, 0, F120),
    ("w120_1e12_sieve", 10 ** 12, 10 ** 12 + 10 ** 8, 120120, 10000, 700, F120),
    ("w204_1e15", 10 ** 15, 10 ** 15 + 4 * 10 ** 8, 2042040, 10000, 0, F204),
    ("w204_top", TOP - 3 * 10 ** 8, TOP, 2042040, 10000, 0, F204),
    ("w204_mmax200", 10 ** 12, 10 ** 12 + 2 * 10 ** 7, 2042040, 200, 0, None),
]


def oracle(lo, hi, M, mmax, sieve, ff):
    argv = [paths.ORACLE_BIN, str(lo), str(hi), "--wheel", str(M), "--mmax", str(mmax), "--threads", "4"]
    if ff:
        argv += ["--filters", ff]
    if sieve:
        argv += ["--sieve", str(sieve)]
    surv = subprocess.run(argv + ["--dump-survivors"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL,
                          check=True).stdout.decode().split("\n")
    ps = sorted(int(l.split()[1]) for l in surv if l.startswith("S "))
    rec = subprocess.run(argv, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL).stdout.decode()
    s = dict(t.split("=") for t in rec.strip().splitlines()[-1].split()[1:])
    return ps, {k: int(s[k]) for k in ("positions", "covered", "sieved", "mr")}


@pytest.mark.parametrize("name,lo,hi,M,mmax,sieve,ff", CASES, ids=[c[0] for c in CASES])
def test_replay_matches_oracle_every_layout(replay_bin, name, lo, hi, M, mmax, sieve, ff):
    want_p, want = oracle(lo, hi, M, mmax, sieve, ff)
    for Mp in candidates(M):
        for maxb in (32768, 1024):
            argv = [replay_bin, str(lo), str(hi), str(M), str(mmax), str(sieve), str(Mp), str(maxb)]
            if ff:
                argv.append(ff)
            out = subprocess.run(argv, stdout=subprocess.PIPE, check=True).stdout.decode().splitlines()
            got_p = [int(l.split()[1]) for l in out if l.startswith("S ")]
            pos, aft, sur = map(int, out[-1].split()[1:])
            assert got_p == want_p, (name, Mp, maxb)
            assert pos == want["positions"], (name, Mp)
            assert sur == want["mr"], (name, Mp)
            if sieve:
                assert pos - aft == want["covered"] and aft - sur == want["sieved"], (name, Mp)
            else:
                assert pos - sur == want["covered"], (name, Mp)


def test_stage_d_fp64_mulmod_is_exact(tmp_path):
    """The device's signed FP64 modular multiply (moduli below 2^50) stays within its
    stated bounds and congruent under 128-bit integer arithmetic (edge operands, 2 million
    random ones), and the Miller-Rabin chain run in it with no corrections gives the exact
    test's verdict for every base 2 .. 37 (edge moduli, strong pseudoprimes, 250000 random)."""
    src = os.path.join(HERE, "replay", "mulmod_replay.cpp")
    exe = str(tmp_path / "mulmod_replay")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", src, "-o", exe], check=True)
    out = subprocess.run([exe, "2000000", "7"], stdout=subprocess.PIPE, check=True).stdout.decode()
    assert out.strip() == "ok 2000000", out


def test_certificate_line_formatting(tmp_path):
    """The table-driven decimal formatter of certificate lines against printf-style
    formatting with digit-by-digit 128-bit division (edge values, 2 million random lines)."""
    src = os.path.join(HERE, "replay", "format_check.cpp")
    exe = str(tmp_path / "format_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, src, os.path.join(CSRC, "es_tail.cpp"),
                    os.path.join(CSRC, "es_plan.cpp"), "-o", exe, "-lpthread"], check=True)
    out = subprocess.run([exe], stdout=subprocess.PIPE, check=True).stdout.decode()
    assert out.strip() == "format ok", out


def test_window_lane_spans(tmp_path):
    """Per-window lane spans (two divisions per window) equal the per-lane formula."""
    src = os.path.join(HERE, "replay", "span_check.cpp")
    exe = str(tmp_path / "span_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True)
    out = subprocess.run([exe], stdout=subprocess.PIPE, check=True).stdout.decode()
    assert out.strip() == "span ok", out


def test_device_montgomery_product_is_exact(tmp_path):
    """The kernels' 64-bit Montgomery product (es_montgomery.h, compiled here for the CPU
    from the same source) times 2^64 equals a*b mod n under 128-bit integer arithmetic,
    with the result reduced, for edge moduli up to 2^64 - 1 and 3 million random cases."""
    src = os.path.join(HERE, "replay", "montgomery_replay.cpp")
    exe = str(tmp_path / "montgomery_replay")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, src, "-o", exe], check=True)
    out = subprocess.run([exe, "3000000", "11"], stdout=subprocess.PIPE, check=True).stdout.decode()
    assert out.strip() == "ok 3000000", out
