// es_plan.h -- host-side planning for the verification-sieve funnel: certificate
// families, the wheel, the filter set, and the sub-wheel layouts the device kernels
// sieve in. Pure C++ (no HIP): everything here is computed once per context.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace es {

typedef uint64_t u64;
typedef uint32_t u32;
typedef uint16_t u16;
typedef unsigned __int128 u128;

// Constants of the record format (rung_scan3.cpp:23, :42-43; census/rung_scan.cpp:64).
constexpr u32 kUcertRmax = 63, kUcertUmax = 64;
constexpr u64 kDirectNMax = 18446744065119617025ull;      // (2^32-1)^2
constexpr u64 kDirectPMax = kDirectNMax - 2048 - 1;       // highest allowed HI
constexpr u32 kRungCap = 255;

// F1 cover: p == res (mod m) with p >= pmin is solved by the (a,c,d) identity.
struct Cover { u64 m, res, pmin; u32 a, c, d; };
// F2 uniform rung certificate: p == res (mod L) is solved at rung r with witness u.
struct Ucert { u64 L, res; u32 r, u; };

std::vector<Cover> covers_upto(u64 mmax);                 // rung_scan3.cpp:93-112
std::vector<Cover> covers_dividing(u64 M);                // rung_scan3.cpp:126-165
std::vector<Ucert> uniform_certs(u32 rmax, u32 umax);     // rung_scan3.cpp:177-202

struct Wheel {
  std::vector<u64> classes;   // ascending unit classes mod M that no F1/F2 kills
  u64 max_pmin = 0;           // over the F1 covers with m | M
  u64 f1_only = 0;            // survivors of the F1 pass alone
};
Wheel build_wheel(u64 M, const std::vector<Cover> &cov, const std::vector<Ucert> &uc);  // :211-226
u64 fnv64_classes(const std::vector<u64> &cls);           // rung_scan3.cpp:228-236

struct ClassPin { u64 M; u64 fnv64; const char *sha256; };
const ClassPin *class_pin(u64 M);                         // rung_scan3.cpp:57-67

// The marking set: certified filters [0, ncov) then small primes [ncov, size).
struct FilterSet {
  std::vector<u64> mod, res;
  size_t ncov = 0;
  u64 fnv = 0;               // FNV-1a over "mod,res;" of the certified part (stderr banner)
};
// Loads FILTER lines (rung_scan3.cpp:1079-1097). Returns false with msg on failure.
bool load_filter_file(const char *path, FilterSet &fs, std::string &msg);
void generated_filters(u64 M, const std::vector<Cover> &covers, FilterSet &fs);   // :1098-1102
void append_small_primes(u64 P, u64 M, FilterSet &fs);                            // :664-674, :1106-1108
void finish_filter_digest(FilterSet &fs);                                         // :1113-1118

// ---------------------------------------------------------------- sub-wheel layout
//
// A layout sieves the wheel positions of [lo, hi) on "lanes" of a modulus Mp | M: lane
// c' holds p = c' + Mp*k for k in a window, one bit per k. Lanes are the distinct
// residues mod Mp of the wheel classes. With Mp == M this is the reference's lane
// layout (rung_scan3.cpp:711-716); with Mp == 840 it is the number line of the six
// hard classes. Wheel membership of p depends on k mod P (P = M/Mp) only, so each lane
// carries a periodic P-bit mask. A filter (m, res) hits lane c' at k == o (mod s) with
// g = gcd(m, Mp), s = m/g, o = ((res - c')/g) * inv(Mp/g mod s, s) mod s, iff
// g | (res - c'); o depends on (lane, filter) only, never on the window. A filter with
// s < kSplitStride is marked as k pieces of stride k*s (piece t: offset o + t*s).
struct Layout {
  u64 Mp = 0, P = 1;
  std::vector<u64> lanes;             // c' ascending
  u32 mask_words = 0;                 // per lane: the P-bit pattern extended by 64 bits
  std::vector<u32> mask;              // [lanes][mask_words]
  // Marking units (filters, or pieces of split filters) in marking order: covers by
  // ascending stride, then small primes by ascending stride.
  std::vector<u32> order;             // index into FilterSet
  std::vector<u32> step;              // stride s of each unit
  std::vector<u32> magic;             // floor(2^32 / s) (2^32-1 for s = 1)
  std::vector<u32> c32;               // 2^32 mod s
  u32 nf = 0, ncov = 0;               // marking units; covers among them
  std::vector<u16> off;               // [lanes][nf]: o, or kNoHit
  double marks_per_int = 0;           // expected marks per integer of [lo, hi)
  double bits_per_int = 0;            // bitmap bits per integer
  u64 hit_pairs = 0;
};
constexpr u16 kNoHit = 0xFFFF;
constexpr u64 kSplitStride = 64;      // shorter strides are marked as pieces of about this stride

// Builds a layout; false (with msg) if a stride or P does not fit the 16-bit tables.
bool build_layout(u64 M, u64 Mp, const std::vector<u64> &classes, const FilterSet &fs,
                  Layout &out, std::string &msg);
// Candidate sub-wheel moduli: Mp = M/d for divisors d of M/840 when 840 | M, else {M}.
std::vector<u64> layout_candidates(u64 M);

// First bit index and bit count of lane c' inside [lo, hi) (0 bits if none).
inline void lane_span(u64 lo, u64 hi, u64 c, u64 Mp, u64 &k0, u64 &nbits) {
  if (hi <= c || hi <= lo) { k0 = 0; nbits = 0; return; }
  const u64 kb = lo <= c ? 0 : (lo - c + Mp - 1) / Mp;
  const u64 ke = (hi - c + Mp - 1) / Mp;
  k0 = kb;
  nbits = ke > kb ? ke - kb : 0;
}

// lane_span for every lane of one window with two divisions in all: for 0 <= c < Mp,
// ceil((x - c) / Mp) = floor(x / Mp) + (x mod Mp > c ? 1 : 0), and x <= c gives 0.
struct WindowSpan {
  u64 ql, rl, qh, rh;
  WindowSpan(u64 lo, u64 hi, u64 Mp) : ql(lo / Mp), rl(lo % Mp), qh(hi / Mp), rh(hi % Mp) {}
  void lane(u64 c, u64 &k0, u64 &nbits) const {
    const u64 kb = ql + (rl > c ? 1 : 0), ke = qh + (rh > c ? 1 : 0);
    k0 = kb;
    nbits = ke > kb ? ke - kb : 0;
  }
};

// The whole host-side plan of one scanner configuration.
struct Plan {
  u64 M = 0, mmax = 0, sieve = 0;
  Wheel wheel;
  u64 cover_pmax = 0;          // max pmin over covers_upto(mmax) (the LO floor)
  const ClassPin *pin = nullptr;
  FilterSet filters;
  std::string filter_source;
};
// Three phases in the reference's refusal order (rung_scan3.cpp:1029-1108): the wheel
// and its pin, then the interval refusals, then the filter set. Each returns 0, or 1
// with a FATAL message (exit code 1 in the CLI).
int build_plan(u64 M, u64 mmax, Plan &plan, std::string &msg);
int check_interval(const Plan &plan, u64 lo, u64 hi, std::string &msg);
int plan_filters(Plan &plan, const char *filter_file, u64 sieve, std::string &msg);

}  // namespace es
