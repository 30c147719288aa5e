"""GPU, long (opt-in: ES_LONG_TESTS=1; about 2.5 minutes on one MI355X): a published
production shard re-derived end to end.

Shard 0 of the run over [255255, 10^19) in 250 shards is [255255, 4*10^16): 394,469,504
certificates, an 18.7 GB round-major stream. The production driver runs it with --only 0;
its meta must equal the published one (data/c19/shard_0.meta, committed as
tests/golden/rederive/c19_shard_0.meta): certs, cert_sha256 over every stdout line but
SUMMARY, SURVIVOR lines and SUMMARY.
"""
import os
import subprocess
import sys

import pytest

from tests import paths

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("ES_LONG_TESTS") != "1", reason="long run: set ES_LONG_TESTS=1")]

sys.path.insert(0, paths.PKG)
import audit_shards  # noqa: E402


def test_production_shard0_rederived(tmp_path):
    out = str(tmp_path / "run")
    argv = [sys.executable, os.path.join(paths.PKG, "run_shards.py"), "255255", str(10 ** 19), "250", out,
            "--only", "0", "--attempts", "1"]
    p = subprocess.run(argv, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=900)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    got = audit_shards.parse_meta(os.path.join(out, "shard_0.meta"))
    want = audit_shards.parse_meta(os.path.join(paths.GOLDEN, "rederive", "c19_shard_0.meta"))
    assert got["complete"] and got["rc"] == 0
    for k in ("certs", "cert_sha256", "summary_line", "survivor_lines"):
        assert got[k] == want[k], k
