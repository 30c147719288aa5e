// es_montgomery.h -- 64-bit Montgomery multiplication from 32-bit multiplies, shared by
// the device kernels (es_device.hip) and the CPU replay that checks it
// (tests/replay/montgomery_replay.cpp) against 128-bit integer arithmetic.
//
// gfx950 has no 64 x 64 -> 128 multiply; v_mad_u64_u32 (32 x 32 + 64 -> 64) is the
// widest and runs at a quarter of the full VALU rate. Word-by-word Montgomery (CIOS with
// two 32-bit words) needs 8 of them and 2 v_mul_lo_u32 per product: each row adds a_i * b
// and then m_i * n with m_i = t_0 * (-n^{-1}) mod 2^32, which clears the low word. Every
// 64-bit partial sum below is bounded by (2^32 - 1)^2 + 2 (2^32 - 1) = 2^64 - 1, so none
// of them overflows; the 65th bit of the result is carried separately.
#pragma once
#include <cstdint>

#ifdef __HIPCC__
#define ES_HD __host__ __device__ __forceinline__
#else
#define ES_HD inline
#endif

namespace es {
namespace mont {

typedef uint32_t u32;
typedef uint64_t u64;

ES_HD u32 lo32(u64 x) { return (u32)x; }
ES_HD u32 hi32(u64 x) { return (u32)(x >> 32); }

// a * b * 2^-64 mod n for odd n < 2^64 and a, b < n; ni = -n^{-1} mod 2^32 (the low word
// of -n^{-1} mod 2^64). Result in [0, n).
ES_HD u64 mul(u64 a, u64 b, u64 n, u32 ni) {
  const u32 a0 = lo32(a), a1 = hi32(a), b0 = lo32(b), b1 = hi32(b), n0 = lo32(n), n1 = hi32(n);
  // Row 0: a0 * b, then + m0 * n (the low word becomes 0 and is dropped).
  u64 p = (u64)a0 * b0;
  const u32 w0 = lo32(p);
  p = (u64)a0 * b1 + hi32(p);
  const u32 w1 = lo32(p), w2 = hi32(p);
  u32 m = w0 * ni;
  p = (u64)m * n0 + w0;
  p = (u64)m * n1 + ((u64)w1 + hi32(p));
  const u32 v0 = lo32(p);
  const u64 s = (u64)w2 + hi32(p);            // < 2^33: value = v0 + s 2^32
  // Row 1: + a1 * b, then + m1 * n.
  p = (u64)a1 * b0 + v0;
  const u32 x0 = lo32(p);
  p = (u64)a1 * b1 + ((u64)lo32(s) + hi32(p));
  const u32 x1 = lo32(p);
  const u64 x2 = (u64)hi32(p) + hi32(s);      // <= 2^32
  m = x0 * ni;
  p = (u64)m * n0 + x0;
  p = (u64)m * n1 + ((u64)x1 + hi32(p));
  const u64 top = x2 + hi32(p);               // the result is lo32(p) + top 2^32 < 2n
  const u64 r = ((top & 0xFFFFFFFFull) << 32) | lo32(p);
  return (top >> 32) || r >= n ? r - n : r;   // r - n wraps to the true value < n
}

// -n^{-1} mod 2^64 for odd n (Newton's iteration doubles the correct low bits).
ES_HD u64 neg_inv(u64 n) {
  u64 x = n;
  for (int k = 0; k < 6; ++k) x *= 2 - n * x;
  return 0 - x;
}

// x a mod n for x < n (n odd) and 1 <= a < 2^32, with rn = 1/(double)n: two 32 x 32 -> 64
// multiplies form x a in 128 bits; the quotient q = floor(x a / n) < a is estimated as
// trunc((double)x * a * rn), within 2^32 * 5 * 2^-53 < 2^-18 of x a / n (every FP64 step
// rounds by at most 2^-53 relatively), so the estimate is q - 1, q or q + 1; two more
// multiplies form x a - q' n in 128 bits, which lies in [-n, 2n), and one correction
// brings it into [0, n). Multiplying a Montgomery form x R by a plain integer a gives the
// Montgomery form of x a.
ES_HD u64 mul_small(u64 x, u32 a, u64 n, double rn) {
  const u64 p0 = (u64)lo32(x) * a;
  const u64 p1 = (u64)hi32(x) * a + hi32(p0);
  const u64 tl = (p1 << 32) | lo32(p0), th = p1 >> 32;
  const u32 q = (u32)((double)x * (double)a * rn);
  const u64 r0 = (u64)lo32(n) * q;
  const u64 r1 = (u64)hi32(n) * q + hi32(r0);
  const u64 ql = (r1 << 32) | lo32(r0), qh = r1 >> 32;
  const u64 rl = tl - ql;
  const u64 rh = th - qh - (tl < ql ? 1u : 0u);
  if (rh >> 63) return rl + n;               // negative: r + n, whose low word is rl + n
  return (rh || rl >= n) ? rl - n : rl;
}

// Strong probable-prime test of odd n > a to the small base a (2 <= a <= 1625, so that
// a^3 < 2^32), n - 1 = d 2^s with d odd; ni = -n^{-1} mod 2^32, one = 2^64 mod n,
// rn = 1/(double)n. a^d runs left to right in Montgomery form over digits of d: w bits
// per digit, each digit w squarings and one mul_small by a^digit (w = 5 for a = 2, whose
// powers up to 2^31 are shifts; w = 2 otherwise, a^3 < 2^32), about 1.1 to 1.4
// Montgomery products per exponent bit instead of the two of a right-to-left ladder.
ES_HD bool strong_test(u64 n, u32 a, u64 d, u32 s, u32 ni, u64 one, double rn) {
  const u32 w = a == 2 ? 5u : 2u, mask = (1u << w) - 1u;
  const u32 a2 = a * a, pw[4] = {1u, a, a2, a2 * a};
  const int bits = 64 - __builtin_clzll(d);
  u64 x = one;
  bool started = false;   // the top digit (nonzero: it holds d's leading bit) needs no squaring
  for (int sh = ((bits + (int)w - 1) / (int)w - 1) * (int)w; sh >= 0; sh -= (int)w) {
    if (started)
      for (u32 i = 0; i < w; ++i) x = mul(x, x, n, ni);
    started = true;
    const u32 dig = (u32)(d >> sh) & mask;
    if (dig) x = mul_small(x, a == 2 ? 1u << dig : pw[dig], n, rn);
  }
  const u64 minus_one = n - one;
  if (x == one || x == minus_one) return true;
  for (u32 i = 1; i < s; ++i) {
    x = mul(x, x, n, ni);
    if (x == minus_one) return true;
  }
  return false;
}

}  // namespace mont
}  // namespace es
