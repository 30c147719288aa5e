"""The production shard driver (run_shards.py) and the run audit (audit_shards.py) on the
CPU: the audit against the reference's own published metas and against planted defects;
the driver's meta format, restart and retry rules with CPU scanners that share the
drop-in's command line (the oracle, and a scripted fake for the failure paths).
"""
import contextlib
import io
import os
import shutil
import subprocess
import sys

import pytest

from tests import paths

sys.path.insert(0, paths.PKG)
import audit_shards  # noqa: E402

DRIVER = os.path.join(paths.PKG, "run_shards.py")
LANE = os.path.join(paths.GOLDEN, "shards", "lane")
F204 = os.path.join(paths.FILTERS, "f_M2042040_mmax2000_gpu.txt")
BAND = (10 ** 12 + 7, 10 ** 12 + 4 * 10 ** 10 + 3, 7)


def audit(*argv):
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        rc = audit_shards.main(list(argv))
    return rc, out.getvalue()


def drive(outdir, *extra, bin_path=None, band=BAND):
    argv = [sys.executable, DRIVER, str(band[0]), str(band[1]), str(band[2]), outdir, "--table", F204,
            "--no-gpu-check", "--retry-sleep", "0", "--threads", "4"] + list(extra)
    if bin_path:
        argv += ["--bin", bin_path]
    return subprocess.run(argv, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=600)


def test_audit_reads_the_published_metas(tmp_path):
    """Shards 0 and 1 of the reference's production run over [255255, 1e19): parsed
    fields, exact tiling between them, the accounting identity."""
    for i in (0, 1):
        shutil.copy(os.path.join(paths.GOLDEN, "rederive", "c19_shard_%d.meta" % i), tmp_path / ("shard_%d.meta" % i))
    m0 = audit_shards.parse_meta(str(tmp_path / "shard_0.meta"))
    assert m0["idx"] == 0 and m0["nshards"] == 250 and m0["certs"] == 394469504
    assert m0["cert_sha256"] == "2a4162859de0d19df809c1b31c6c2b8d2a31823e65d6ab36e6f2e7325545ebe8"
    assert m0["lo"] == 255255 and m0["primes"] == m0["rung"]
    rc, out = audit(str(tmp_path))
    assert rc == 1                                   # 2 of 250 shards: the index set is incomplete
    assert "boundaries tile exactly: [255255, 80000000000253213)" in out
    assert "accounting identity holds" in out and "survivors=0 across the band" in out


def test_audit_passes_reference_run_and_catches_defects(tmp_path):
    rc, out = audit(LANE, "--expect-lo", str(BAND[0]), "--expect-hi", str(BAND[1]), "--expect-shards", "7")
    assert rc == 0, out

    def planted(edit):
        d = tmp_path / edit.__name__
        shutil.copytree(LANE, d)
        edit(d)
        return audit(str(d))[0]

    def gap(d):       # a shard that starts one integer late
        p = d / "shard_3.meta"
        s = p.read_text()
        lo = s.split("SUMMARY lo=")[1].split()[0]
        p.write_text(s.replace("SUMMARY lo=%s" % lo, "SUMMARY lo=%d" % (int(lo) + 1)))

    def accounting(d):
        p = d / "shard_2.meta"
        s = p.read_text()
        c = s.split(" composite=")[1].split()[0]
        p.write_text(s.replace(" composite=%s " % c, " composite=%d " % (int(c) + 1)))

    def missing(d):
        os.remove(d / "shard_5.meta")

    def truncated(d):
        p = d / "shard_1.meta"
        p.write_text("".join(l for l in p.read_text().splitlines(keepends=True) if not l.startswith("SUMMARY")))

    def survivor(d):
        p = d / "shard_6.meta"
        s = p.read_text().splitlines(keepends=True)
        p.write_text("".join(s[:-1]) + "SURVIVOR 1003999999999\n" + s[-1])

    for edit in (gap, accounting, missing, truncated, survivor):
        assert planted(edit) == 1, edit.__name__


def test_driver_metas_from_the_oracle_equal_the_reference_run(tmp_path):
    """The driver running the oracle (same command line as the reference scanner) writes
    metas whose deterministic part equals the reference binary's run; a second pass
    touches nothing."""
    out = str(tmp_path / "run")
    p = drive(out, "--order", "lane", bin_path=paths.ORACLE_BIN)
    assert p.returncode == 0, p.stderr.decode()
    rc, rep = audit(out, "--against", LANE, "--expect-shards", "7")
    assert rc == 0, rep
    before = {f: os.path.getmtime(os.path.join(out, f)) for f in os.listdir(out) if f.endswith(".meta")}
    assert drive(out, "--order", "lane", bin_path=paths.ORACLE_BIN).returncode == 0
    assert before == {f: os.path.getmtime(os.path.join(out, f)) for f in os.listdir(out) if f.endswith(".meta")}
    log = open(os.path.join(out, "run.log")).read()
    assert log.count(" done rc=0 ") == 7 and "pass complete: 7/7 shards done, 0 FAILED" in log


FAKE = r'''#!/usr/bin/env python3
import os, sys
state = os.path.join(os.path.dirname(os.path.abspath(__file__)), "calls")
n = int(open(state).read()) + 1 if os.path.exists(state) else 1
open(state, "w").write(str(n))
mode = os.environ["FAKE_MODE"]
shard = sys.argv[sys.argv.index("--shard") + 1]
if mode == "crash_once" and n == 1:
    print("RUNG 1 2 3 4", flush=True)
    sys.exit(139)
if mode == "always":
    sys.exit(134)
if mode == "no_summary":
    print("RUNG 5 6 7 8")
    sys.exit(0)
print("RUNG 1000000000039 250000000010 1 1")
surv = mode == "survivor"
if surv:
    print("SURVIVOR 1000000000061")
print("SUMMARY lo=1 hi=2 wheelM=2042040 lanes=2308 mmax=10000 ucert_rmax=63 ucert_umax=64 wheel_pmin=255254 "
      "classes_sha256=x positions=9 covered=7 sieved=0 mr=2 composite=0 rung=1 direct=0 survivors=%d" % surv)
sys.exit(2 if surv else 0)
'''


@pytest.mark.parametrize("mode,meta_rc,attempt,failed", [
    ("crash_once", 0, 2, False), ("survivor", 2, 1, False), ("always", None, None, True),
    ("no_summary", None, None, True)])
def test_driver_retry_rules(tmp_path, monkeypatch, mode, meta_rc, attempt, failed):
    """A crash is retried; exit 2 (survivors) is a result and is not; a run without its
    SUMMARY line counts as a crash; a shard that never completes is marked FAILED."""
    fake = tmp_path / "fake_scanner"
    fake.write_text(FAKE)
    fake.chmod(0o755)
    monkeypatch.setenv("FAKE_MODE", mode)
    out = str(tmp_path / "run")
    p = drive(out, "--attempts", "2", bin_path=str(fake), band=(1000, 2000, 1))
    assert p.returncode == 0, p.stderr.decode()
    meta = os.path.join(out, "shard_0.meta")
    assert os.path.exists(os.path.join(out, "shard_0.FAILED")) == failed
    if failed:
        assert not os.path.exists(meta)
        return
    m = audit_shards.parse_meta(meta)
    assert m["rc"] == meta_rc and m["attempt"] == attempt and m["certs"] == 1 and m["complete"]
    import hashlib
    body = "RUNG 1000000000039 250000000010 1 1\n" + ("SURVIVOR 1000000000061\n" if mode == "survivor" else "")
    assert m["cert_sha256"] == hashlib.sha256(body.encode()).hexdigest()
    assert m["survivor_lines"] == (["SURVIVOR 1000000000061"] if mode == "survivor" else [])
    calls = int(open(tmp_path / "calls").read())
    assert calls == attempt


def test_driver_refuses_missing_configuration(tmp_path):
    p = drive(str(tmp_path / "run"), bin_path=str(tmp_path / "nope"))
    assert p.returncode == 1 and b"FATAL scanner missing" in p.stdout


def test_driver_only_runs_the_named_shards(tmp_path):
    """--only I,J runs exactly those shards (each meta equal to the reference run's), and
    refuses an index outside [0, N)."""
    out = str(tmp_path / "run")
    p = drive(out, "--order", "lane", "--only", "1,5", bin_path=paths.ORACLE_BIN)
    assert p.returncode == 0, p.stderr.decode()
    assert sorted(f for f in os.listdir(out) if f.endswith(".meta")) == ["shard_1.meta", "shard_5.meta"]
    for i in (1, 5):
        got = audit_shards.parse_meta(os.path.join(out, "shard_%d.meta" % i))
        want = audit_shards.parse_meta(os.path.join(LANE, "shard_%d.meta" % i))
        for k in ("certs", "cert_sha256", "summary_line", "survivor_lines"):
            assert got[k] == want[k], (i, k)
    p = drive(str(tmp_path / "bad"), "--only", "7", bin_path=paths.ORACLE_BIN)
    assert p.returncode == 1 and b"FATAL --only" in p.stdout
