"""CPU checks of the product's native library: it loads, it exports every entry point
include/es_sieve.h declares, its host-side stage D/E reproduces the reference's golden
certificate lines, and its CLI's host-only modes match the reference's outputs."""
import ctypes
import hashlib
import os
import re

import pytest

from tests import paths

pytestmark = pytest.mark.skipif(not os.path.exists(paths.SIEVE_LIB), reason="libes_sieve.so not built")


def _declared():
    src = open(paths.HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(es_[a-z_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(paths.SIEVE_LIB)
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_open_without_device_fails_loudly():
    """No CPU fallback: opening a scanner on a machine without a HIP device is an error."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
    es = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(es)
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    with pytest.raises(es.EsError):
        es.Scanner(filters=es.F204)


GOLDEN_LINE_FILES = ["w19_certificates.txt", "w64_head.txt", "s69_tail.txt", "s0_head.txt"]


@pytest.mark.parametrize("fn", GOLDEN_LINE_FILES)
def test_host_tail_reproduces_golden_lines(fn):
    """es_certify (Montgomery stage D/E) re-derives the reference's published lines."""
    lib = ctypes.CDLL(paths.SIEVE_LIB)
    lib.es_certify.argtypes = [ctypes.c_uint64, ctypes.c_char_p, ctypes.c_size_t]
    buf = ctypes.create_string_buffer(256)
    lines = [l for l in open(os.path.join(paths.GOLDEN, "rederive", fn)) if l.startswith("RUNG")]
    for line in lines:
        assert lib.es_certify(int(line.split()[1]), buf, 256) == 0
        assert buf.value.decode() == line


def test_cli_self_test_and_class_tables():
    rc, out, _ = paths.run(paths.SIEVE_BIN, ["--verify"])
    assert rc == 0 and b"SELF-TEST PASSED" in out
    want = dict(l.split()[::-1] for l in open(os.path.join(paths.GOLDEN, "class_tables.sha256")))
    for M in ("120120", "2042040"):
        rc, out, _ = paths.run(paths.SIEVE_BIN, ["--dump-classes", "--wheel", M])
        assert rc == 0 and hashlib.sha256(out).hexdigest() == want[M]


@pytest.mark.parametrize("lo", [1000003, 9223372036854775808, 18446744065119000000])
def test_cli_check_direct_rows(lo):
    rc, out, _ = paths.run(paths.SIEVE_BIN, ["--check-direct", str(lo), "40"])
    assert rc == 0
    assert out.decode() == open(os.path.join(paths.GOLDEN, "direct", "%d.txt" % lo)).read()


def test_cli_refusals():
    """Exit code 1 and nothing on stdout for the reference's refusals."""
    F = os.path.join(paths.FILTERS, "f_M2042040_mmax2000_gpu.txt")
    for argv in (["1000", "2000"],                                      # LO below the pmin floor
                 ["1000", "1000"],                                      # the same, empty
                 ["18446744065119614977", "18446744065119614977", "--filters", F],   # HI above, empty
                 ["1000000000000", "18446744065119614977", "--filters", F],   # HI above the ceiling
                 ["1000000000000", "1000000001000", "--bogus"],          # unknown option
                 ["1000000000000", "1000000001000", "--filters", "/nonexistent"]):
        rc, out, _ = paths.run(paths.SIEVE_BIN, argv)
        assert rc == 1 and out == b"", argv


def test_cli_check_factor():
    rc, out, _ = paths.run(paths.SIEVE_BIN, ["--check-factor", "40"])
    assert rc == 0 and out.startswith(b"ok")


BANNER_CASES = [
    ("gpu_table", ["--filters", "erdos-strauss_amd/tables/filters/f_M2042040_mmax2000_gpu.txt"]),
    ("w120_runtime", ["--wheel", "120120", "--filters", "erdos-strauss_amd/tables/filters/f_M120120_mmax2000_runtime.txt"]),
    ("mi355x_table", ["--filters", "erdos-strauss_amd/tables/filters/f_M2042040_mmax2000_mi355x.txt"]),
    ("generated_mmax2000", ["--mmax", "2000"]),
    ("generated_w120_mmax200", ["--wheel", "120120", "--mmax", "200"]),
    ("sieve1000", ["--filters", "erdos-strauss_amd/tables/filters/f_M2042040_mmax2000_gpu.txt", "--sieve", "1000"]),
]


def _digest_line(err, prefix):
    lines = [l for l in err.decode().splitlines() if l.startswith(prefix + " filters=")]
    assert len(lines) == 1, err
    return lines[0][len(prefix):]


@pytest.mark.skipif(not os.path.exists(paths.REF_BIN), reason="reference scanner not built from source")
@pytest.mark.parametrize("name,flags", BANNER_CASES, ids=[c[0] for c in BANNER_CASES])
def test_filter_digest_banner_matches_reference(name, flags):
    """The stderr banner's filter line -- count, FNV-1a digest over "mod,res;" of the
    certified filters (fnv64=), source and sieve bound (rung_scan3.cpp:1113-1126) -- is
    the reference scanner's, for table files, generated covers and --sieve. An empty
    interval needs no device, so this runs anywhere."""
    argv = ["10000000000000", "10000000000000"] + flags
    rc, out, err = paths.run(paths.SIEVE_BIN, argv)
    rrc, rout, rerr = paths.run(paths.REF_BIN, argv)
    assert rc == rrc == 0
    assert out == rout                                   # the empty record
    assert _digest_line(err, "[es_sieve]") == _digest_line(rerr, "[rung_scan3]")


def test_capi_refuses_empty_windows_outside_the_range():
    """Through the C ABI, an empty window is refused exactly as the CLI and rung_scan3
    refuse it (rung_scan3.cpp:1057-1069 run whether or not the interval is empty): LO at
    or below the pmin floor, or HI above the stage-E ceiling. A scanner opened for
    host-side planning only (device -1) makes the check, so this runs without a GPU."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
    es = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(es)
    top = 18446744065119614976
    with es.Scanner(filters=es.F204, device=-1) as sc:
        for lo, hi in [(1000, 1000), (255254, 255254), (1000, 2000), (top + 1, top + 1), (10 ** 13, top + 1)]:
            assert not sc.check_interval(lo, hi)
            with pytest.raises(es.EsError, match="FATAL: (LO|HI) must"):
                sc.records([(lo, hi)])
            with pytest.raises(es.EsError, match="FATAL: (LO|HI) must"):
                sc.submit([(10 ** 13, 10 ** 13 + 1), (lo, hi)])
        assert sc.in_flight() == 0
        # A valid empty window passes the check; scanning it then needs the device.
        assert sc.check_interval(10 ** 13, 10 ** 13) and sc.check_interval(top, top)
        with pytest.raises(es.EsError, match="device"):
            sc.records([(10 ** 13, 10 ** 13)])


@pytest.mark.parametrize("given,want", [(None, "8"), ("4", "8"), ("8", "8"), ("16", "16")])
def test_loading_raises_hardware_queues_to_the_slot_count(given, want):
    """Loading the library raises GPU_MAX_HW_QUEUES to ES_MAX_IN_FLIGHT (8) when it is unset or
    lower, and leaves a larger value alone (checked in a fresh process, without a GPU call)."""
    import subprocess
    import sys
    code = ("import ctypes; ctypes.CDLL(%r); libc = ctypes.CDLL(None); libc.getenv.restype = ctypes.c_char_p; "
            "print(libc.getenv(b'GPU_MAX_HW_QUEUES').decode())" % paths.SIEVE_LIB)
    env = dict(os.environ)
    env.pop("GPU_MAX_HW_QUEUES", None)
    if given is not None:
        env["GPU_MAX_HW_QUEUES"] = given
    out = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE, check=True, timeout=120)
    assert out.stdout.decode().strip() == want
