"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code that runs on the
CPU: the oracle, the product's host-only sources (es_plan: wheel, covers, filters, lane
offsets; es_tail: host stage D/E and line formatting; es_prune: the table chooser) and
the CPU replays that drive them (tests/replay/Makefile, target sanitize). Every program
runs on small inputs with every report fatal (-fno-sanitize-recover=all, a nonzero exit
on any error), and its output is checked as the unsanitized tests check it."""
import os
import subprocess

import pytest

from tests import paths

HERE = os.path.dirname(os.path.abspath(__file__))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")


@pytest.fixture(scope="module")
def san(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("sanitize"))
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "replay"), "sanitize", "OUT=" + out, "-j8"], check=True)
    return out


def run(exe, args, timeout=600):
    p = subprocess.run([exe] + args, stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=ENV, timeout=timeout,
                       cwd=paths.REPO)
    assert "Sanitizer" not in p.stderr.decode(), p.stderr.decode()[-3000:]
    assert "runtime error" not in p.stderr.decode(), p.stderr.decode()[-3000:]
    return p


@pytest.mark.parametrize("name", ["tiny", "empty", "shard_empty", "mmax2000_w120", "mmax60_w120",
                                  "mmax60_w120_nosurv", "mmax200_w204", "c1"])
def test_oracle_under_sanitizers_matches_golden_record(san, name):
    args = open(os.path.join(paths.GOLDEN, "records", name + ".args")).read().split("\n")[0].split()
    want = open(os.path.join(paths.GOLDEN, "records", name + ".txt")).read()
    p = run(os.path.join(san, "es_oracle"), args + ["--threads", "4"])
    assert p.stdout.decode() == want


def test_oracle_sieve_and_batch_under_sanitizers(san, tmp_path):
    """--sieve (stage B') and --batch (many windows in one process) against the
    unsanitized oracle, on windows short enough for a sanitized run."""
    F = os.path.join(paths.FILTERS, "f_M2042040_mmax2000_gpu.txt")
    argv = [str(10 ** 12), str(10 ** 12 + 2 * 10 ** 7), "--filters", F, "--sieve", "1000", "--threads", "4"]
    want = subprocess.run([paths.ORACLE_BIN] + argv, stdout=subprocess.PIPE, check=True).stdout
    assert run(os.path.join(san, "es_oracle"), argv).stdout == want
    wins = [l.split() for l in open(os.path.join(paths.GOLDEN, "c4_windows.txt")) if not l.startswith("#")]
    small = [w for w in wins if int(w[1]) - int(w[0]) < 2 ** 20][:24]
    path = str(tmp_path / "batch.txt")
    with open(path, "w") as f:
        f.writelines("%s %s\n" % (a, b) for a, b in small)
    argv = ["--batch", path, "--filters", F, "--threads", "4"]
    want = subprocess.run([paths.ORACLE_BIN] + argv, stdout=subprocess.PIPE, check=True).stdout
    assert want.count(b"SUMMARY ") == len(small)
    assert run(os.path.join(san, "es_oracle"), argv).stdout == want


def test_oracle_batch_and_stage_e_under_sanitizers(san):
    rows = open(os.path.join(paths.GOLDEN, "direct", "9223372036854775808.txt")).read()
    ps = [l.split()[1] for l in rows.splitlines()]
    p = run(os.path.join(san, "es_oracle"), ["--solve-direct"] + ps)
    assert p.stdout.decode() == rows
    assert run(os.path.join(san, "es_oracle"), ["--verify"]).returncode == 0


def test_kernel_replay_under_sanitizers(san):
    F = os.path.join(paths.FILTERS, "f_M2042040_mmax2000_gpu.txt")
    for Mp in ("840", "9240", "2042040"):
        p = run(os.path.join(san, "kernel_replay"), [str(10 ** 15), str(10 ** 15 + 3 * 10 ** 7), "2042040", "10000",
                                                     "0", Mp, "1024", F])
        assert p.returncode == 0
        last = p.stdout.decode().splitlines()[-1].split()
        assert last[0] == "C" and int(last[1]) > 0


def test_host_tail_formatting_under_sanitizers(san):
    assert run(os.path.join(san, "format_check"), []).stdout.decode().strip() == "format ok"


def test_replays_under_sanitizers(san):
    assert run(os.path.join(san, "span_check"), []).stdout.decode().strip() == "span ok"
    assert run(os.path.join(san, "montgomery_replay"), ["200000", "11"]).stdout.decode().strip() == "ok 200000"
    assert run(os.path.join(san, "mulmod_replay"), ["200000", "7"]).stdout.decode().strip() == "ok 200000"
    F = os.path.join(paths.FILTERS, "f_M2042040_mmax2000_gpu.txt")
    assert run(os.path.join(san, "table_check"), [F, "2042040", "2000"]).stdout.decode().strip() == "ok 3745"


def test_pruner_under_sanitizers(san, tmp_path):
    out = str(tmp_path / "t.txt")
    p = run(os.path.join(san, "es_prune"), ["--cost-form", "ref", "--ns-mark", "0.00001", "--ns-pair", "0.42",
                                            "--tile-positions", "2448532", "--ns-comp", "3.5", "--ns-prime", "140",
                                            "--wheel", "120120", "--mmax", "200", "--npos", "2048", "--threads", "4",
                                            "--out", out])
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    assert any(l.startswith("FILTER") for l in open(out))
