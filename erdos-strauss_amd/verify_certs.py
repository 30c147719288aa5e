#!/usr/bin/env python3
"""Exact-arithmetic re-derivation of certificate lines (RUNG / SOLVED / SURVIVOR).

Stand-alone counterpart of the reference's sieve/verify_covers.py:246-275 (validity
rules) and sieve/verify_sample.py:37-94 (independent checks), written from those rules,
standard library only, so it runs wherever the scanner runs, with or without the reference.

A line is accepted only if the Erdos-Straus identity it claims holds in exact rationals:

  RUNG p A r u     A == (p+r)//4, u | (pA)^2, r | (pA + u), and with q = pA,
                   x = (q+u)/r and y = q(q+u)/(r u) integral,
                   4/p == 1/A + 1/x + 1/y
  SOLVED p a c d k m = 4acd-1, e = ck-a >= 1, p == m k - 4a^2 d and
                   4/p == 1/(ade) + 1/(acdp) + 1/(cdep)
  SURVIVOR p       never accepted (an unsolved prime needs escalation)

Independent checks (optional, --lo/--hi): p is prime (deterministic Miller-Rabin for
all 64-bit integers), lo <= p < hi, and p == 1 (mod 24) (every wheel class is).

Usage: verify_certs.py [--lo LO --hi HI] FILE...   exit 0 iff every line verifies.
"""
import sys
from fractions import Fraction

_BASES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37)


def is_prime(n):
    """Deterministic for n < 3.3e24 with the first twelve prime bases."""
    if n < 2:
        return False
    for q in _BASES:
        if n % q == 0:
            return n == q
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in _BASES:
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


def check_rung(p, A, r, u):
    if r <= 0 or u <= 0 or A != (p + r) // 4:
        return False
    q = p * A
    if (q * q) % u or (q + u) % r:
        return False
    x = (q + u) // r
    y, rem = divmod(q * (q + u), r * u)
    if rem or x <= 0 or y <= 0:
        return False
    return Fraction(4, p) == Fraction(1, A) + Fraction(1, x) + Fraction(1, y)


def check_solved(p, a, c, d, k):
    if min(a, c, d, k) <= 0:
        return False
    m = 4 * a * c * d - 1
    e = c * k - a
    if e < 1 or p != m * k - 4 * a * a * d:
        return False
    return Fraction(4, p) == (Fraction(1, a * d * e) + Fraction(1, a * c * d * p)
                              + Fraction(1, c * d * e * p))


def verify_lines(lines, lo=None, hi=None):
    """Returns (n_rung, n_solved, n_bad, n_survivor, problems)."""
    nr = ns = bad = surv = 0
    problems = []
    for ln, line in enumerate(lines, 1):
        t = line.split()
        if not t or t[0] == "SUMMARY":
            continue
        kind = t[0]
        if kind == "SURVIVOR":
            surv += 1
            problems.append((ln, "SURVIVOR " + t[1]))
            continue
        try:
            vals = [int(v) for v in t[1:]]
        except ValueError:
            vals = []
        ok = False
        if kind == "RUNG" and len(vals) == 4:
            nr += 1
            ok = check_rung(*vals)
        elif kind == "SOLVED" and len(vals) == 5:
            ns += 1
            ok = check_solved(*vals)
        if ok:
            p = vals[0]
            if lo is not None and not (lo <= p < hi):
                ok = False
            elif lo is not None and (p % 24 != 1 or not is_prime(p)):
                ok = False
        if not ok:
            bad += 1
            problems.append((ln, line.rstrip("\n")))
    return nr, ns, bad, surv, problems


def main(argv):
    lo = hi = None
    files = []
    i = 0
    while i < len(argv):
        if argv[i] == "--lo":
            lo = int(argv[i + 1]); i += 2
        elif argv[i] == "--hi":
            hi = int(argv[i + 1]); i += 2
        else:
            files.append(argv[i]); i += 1
    tot = [0, 0, 0, 0]
    for path in files:
        with open(path) as f:
            nr, ns, bad, surv, problems = verify_lines(f, lo, hi)
        for ln, what in problems[:20]:
            print("%s:%d: %s" % (path, ln, what))
        for j, v in enumerate((nr, ns, bad, surv)):
            tot[j] += v
    print("verified: %d RUNG, %d SOLVED; %d bad, %d survivors" % tuple(tot))
    return 1 if tot[2] or tot[3] else 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
