"""The allocation-free primality entry point (es_primality_noalloc) against es_primality.

Bar: bit-exact. The same values go through es_primality (host buffers, the library
allocates its device buffers) and through es_primality_noalloc (device buffers the test
allocates, the library's fixed scratch for its own state); every verdict byte must be
equal, on two calls in a row, the second with fewer values, and the oracle's test agrees.
Sizes beyond the scratch maximum and host pointers are refused before any kernel runs.
"""
import ctypes
import importlib.util
import os

import numpy as np
import pytest

from tests import paths

pytestmark = pytest.mark.gpu

spec = importlib.util.spec_from_file_location("es_sieve", os.path.join(paths.PKG, "es_sieve.py"))
es = importlib.util.module_from_spec(spec)
spec.loader.exec_module(es)

TOP = 18446744065119614976
H2D, D2H = 1, 2


def _hip():
    hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return hip


def _values(n, seed):
    """Dense odd runs at 10^12, below 2^63, at the ceiling, small values, strong
    pseudoprimes to several bases, and random u64 values; n of them."""
    rng = np.random.default_rng(seed)
    v = []
    for base in (10 ** 12 + 1, 2 ** 63 - 20001, TOP - 40001, 3):
        v += [base + 2 * i for i in range(n // 8)]
    v += [2047, 3277, 4033, 4681, 8321, 3215031751, 2152302898747, 3474749660383, 341550071728321,
          3825123056546413051, 18446744073709551557, 18446744065119614953, 2, 1, 0, 4, 9, 25]
    v += [int(x) for x in rng.integers(0, 2 ** 63, size=n, dtype=np.uint64) * 2 + 1]
    return np.array(v[:n], dtype=np.uint64)


def _noalloc(sc, hip, d_vals, d_flags, vals):
    n = len(vals)
    assert hip.hipMemcpy(d_vals, vals.ctypes.data, n * 8, H2D) == 0
    rc = es.lib().es_primality_noalloc(sc.h, d_vals, n, d_flags, None)
    assert rc == 0, es.lib().es_last_error().decode()
    assert hip.hipDeviceSynchronize() == 0
    out = np.zeros(n, dtype=np.uint8)
    assert hip.hipMemcpy(out.ctypes.data, d_flags, n, D2H) == 0
    return out


def test_noalloc_primality_matches_es_primality():
    hip = _hip()
    olib = ctypes.CDLL(paths.ORACLE_LIB)
    olib.es_oracle_is_prime.argtypes = [ctypes.c_uint64]
    sc = es.Scanner(wheel=2042040, filters=es.F204, device=0)
    n1, n2 = 240000, 7000
    d_vals, d_flags = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(d_vals), n1 * 8) == 0
    assert hip.hipMalloc(ctypes.byref(d_flags), n1) == 0
    try:
        assert hip.hipMemset(d_flags, 0x5A, n1) == 0
        for n, seed in ((n1, 11), (n2, 12)):
            vals = _values(n, seed)
            print("noalloc call n=%d seed=%d" % (n, seed), flush=True)
            got = _noalloc(sc, hip, d_vals, d_flags, vals)
            want = np.array(sc.primality([int(x) for x in vals]), dtype=np.uint8)
            bad = np.nonzero(got != want)[0]
            assert bad.size == 0, [(int(vals[i]), int(got[i]), int(want[i])) for i in bad[:10]]
            sample = range(0, n, max(1, n // 4000))
            assert all(int(got[i]) == olib.es_oracle_is_prime(int(vals[i])) for i in sample)
        # The second, shorter call wrote its n2 verdicts and nothing beyond them.
        tail = np.zeros(n1 - n2, dtype=np.uint8)
        assert hip.hipMemcpy(tail.ctypes.data, d_flags.value + n2, n1 - n2, D2H) == 0
        first = _values(n1, 11)
        want_tail = np.array(sc.primality([int(x) for x in first[n2:]]), dtype=np.uint8)
        assert (tail == want_tail).all()
        # Refusals happen on the host: too many values, host memory, a null buffer.
        host = np.zeros(8, dtype=np.uint64)
        L = es.lib()
        assert L.es_primality_noalloc(sc.h, d_vals, (1 << 20) + 1, d_flags, None) == -3
        assert b"maximum" in L.es_last_error()
        assert L.es_primality_noalloc(sc.h, host.ctypes.data, 8, d_flags, None) == -3
        assert L.es_primality_noalloc(sc.h, None, 8, d_flags, None) == -3
        assert L.es_primality_noalloc(sc.h, d_vals, 0, d_flags, None) == 0
    finally:
        hip.hipFree(d_vals)
        hip.hipFree(d_flags)
        sc.close()
