// es_tail.h -- per-prime certification on the host: stage D (rung search) and stage E
// (direct F1 instance), with 64-bit Montgomery arithmetic and magic-inverse trial
// division. Certificates are a function of p alone (smallest rung, then smallest
// witness; smallest a, then d, then divisor), so where they are computed cannot change
// them.
#pragma once
#include <cstdint>
#include <string>

namespace es {

typedef uint64_t u64;
typedef unsigned __int128 u128;

// Deterministic primality for every u64 (trial prologue by the first twelve primes,
// then strong-probable-prime tests to those twelve bases; census/rung_scan.cpp:76-92).
bool is_prime_u64(u64 n);
// Base-2 strong probable prime (rung_scan3.cpp:680-689).
bool sprp2_u64(u64 n);

// Canonical factorization (primes ascending). Returns the number of distinct primes;
// at most 16 (census/rung_scan.cpp:144-158), -1 on overflow.
int factor_u64(u64 n, u64 *primes, int *exps);

// Stage D (rung_scan3.cpp:578-603): "RUNG p A r u\n" or false.
bool solve_rung(u64 p, std::string &line);
// Stage E (rung_scan3.cpp:624-647): "SOLVED p a c d k\n" or false.
bool solve_direct(u64 p, std::string &line);

// 0 = RUNG, 1 = SOLVED, 2 = SURVIVOR; line gets the certificate text.
int certify_prime(u64 p, std::string &line);

// Decimal digits of v written backwards so that they end at `end`; returns the first.
char *put_u64(char *end, u64 v);

// "SOLVED p a c d k\n" (rung_scan3.cpp:639-643).
void format_direct(u64 p, u64 a, u64 c, u64 d, u64 k, std::string &line);

// "RUNG p A r u\n" (rung_scan3.cpp:594-600).
void format_rung(u64 p, u64 A, unsigned r, u128 u, std::string &line);

}  // namespace es
