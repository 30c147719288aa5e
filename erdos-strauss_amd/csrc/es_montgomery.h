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

}  // namespace mont
}  // namespace es
