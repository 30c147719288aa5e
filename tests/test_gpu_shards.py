"""GPU: the production shard driver with the drop-in scanner. A 7-shard run must equal,
shard by shard (certs, cert_sha256, SURVIVOR lines, SUMMARY), the reference binary's run
of the same band through the same driver; the round-major run (the published metas'
order) must tile the band with the same counters."""
import contextlib
import io
import os
import subprocess
import sys

import pytest

from tests import paths

pytestmark = pytest.mark.gpu

sys.path.insert(0, paths.PKG)
import audit_shards  # noqa: E402
import run_shards  # noqa: E402

LANE = os.path.join(paths.GOLDEN, "shards", "lane")
BAND = (10 ** 12 + 7, 10 ** 12 + 4 * 10 ** 10 + 3, 7)


def audit(*argv):
    out = io.StringIO()
    with contextlib.redirect_stdout(out):
        rc = audit_shards.main(list(argv))
    return rc, out.getvalue()


def drive(outdir, order):
    argv = [sys.executable, os.path.join(paths.PKG, "run_shards.py"), str(BAND[0]), str(BAND[1]), str(BAND[2]),
            outdir, "--order", order, "--retry-sleep", "0"]
    return subprocess.run(argv, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=600)


def test_gpu_visible_to_the_driver():
    assert run_shards.amd_gpu_visible()


def test_lane_run_equals_reference_run(tmp_path):
    out = str(tmp_path / "lane")
    p = drive(out, "lane")
    assert p.returncode == 0, p.stdout.decode() + p.stderr.decode()
    rc, rep = audit(out, "--against", LANE, "--expect-lo", str(BAND[0]), "--expect-hi", str(BAND[1]),
                    "--expect-shards", "7")
    assert rc == 0, rep


def test_round_run_tiles_with_the_same_counters(tmp_path):
    out = str(tmp_path / "round")
    p = drive(out, "round")
    assert p.returncode == 0, p.stdout.decode() + p.stderr.decode()
    rc, rep = audit(out, "--expect-lo", str(BAND[0]), "--expect-hi", str(BAND[1]), "--expect-shards", "7")
    assert rc == 0, rep
    for i in range(7):
        r = audit_shards.parse_meta(os.path.join(out, "shard_%d.meta" % i))
        l = audit_shards.parse_meta(os.path.join(LANE, "shard_%d.meta" % i))
        assert r["summary_line"] == l["summary_line"] and r["certs"] == l["certs"]
