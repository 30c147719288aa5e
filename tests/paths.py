"""Locations of the built artefacts the tests drive, and small helpers to run them.

The oracle (oracle/es_oracle, oracle/libes_oracle.so) is the checker; the product is
erdos-strauss_amd/bin/es_sieve (rung_scan3's CLI over the HIP path) and
erdos-strauss_amd/lib/libes_sieve.so (the C-ABI). Tests build nothing: `python -c
"import __graft_entry__ as g; g.build()"` does.
"""
import os
import subprocess

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(REPO, "tests", "golden")
ORACLE_BIN = os.path.join(REPO, "oracle", "es_oracle")
ORACLE_LIB = os.path.join(REPO, "oracle", "libes_oracle.so")
REF_BIN = os.path.join(REPO, "oracle", "_ref", "rung_scan3")
PKG = os.path.join(REPO, "erdos-strauss_amd")
SIEVE_BIN = os.path.join(PKG, "bin", "es_sieve")
SIEVE_LIB = os.path.join(PKG, "lib", "libes_sieve.so")
HEADER = os.path.join(REPO, "include", "es_sieve.h")
FILTERS = os.path.join(PKG, "tables", "filters")


def record_cases():
    """(name, argv, rc) for every committed reference record."""
    d = os.path.join(GOLDEN, "records")
    out = []
    for fn in sorted(os.listdir(d)):
        if not fn.endswith(".args"):
            continue
        with open(os.path.join(d, fn)) as f:
            argv = f.readline().split()
            rc = int(f.readline().strip().split("=")[1])
        out.append((fn[:-5], argv, rc))
    return out


def golden_record(name):
    with open(os.path.join(GOLDEN, "records", name + ".txt"), "rb") as f:
        return f.read()


def run(binary, argv, timeout=600, env=None):
    p = subprocess.run([binary] + list(argv), cwd=REPO, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, timeout=timeout, env=env)
    return p.returncode, p.stdout, p.stderr
