// es_device.hip -- gfx950 kernels of the verification-sieve funnel and the host code
// that drives them.
//
//   k_stage   first kernel of a batch: zeroes its counters and queue heads and copies its
//             header and tile descriptors in from mapped pinned host memory.
//   k_sieve   one tile = one LDS byte map of <= max_bits positions of one lane. Init
//             expands the lane's periodic wheel mask (sub-wheel layouts) and the tail
//             mask to one byte per position; every filter then clears its positions
//             with plain byte stores of 0 (idempotent, no atomics). Stage B (certified
//             filters) and stage B' (small primes) are separate phases so the count
//             between them splits `covered` from `sieved` exactly
//             (rung_scan3.cpp:755-772). Survivors are counted per thread, scanned
//             across the wave with shuffles and across the workgroup in LDS, placed
//             with one global atomic per tile, and written as 64-bit integers.
//             Persistent form: workgroups pull tiles from a device queue.
//   k_mr_base2  trial prologue and the base-2 strong test per survivor (FP64 residues
//             below 2^50, 64-bit Montgomery above); k_mr_bases runs the other eleven
//             bases when stage D stays on the host.
//   k_rung    stage D on the device: bases 3..37 for pending values, then the rungs
//             r = 3 (mod 4) up to 255 with a small-prime sieve, least-witness search and
//             Pollard rho, one persistent workgroup per CU taking survivors from a queue.
//   k_pack    survivor records and tile counters into pinned host memory.
//
// Host side: DeviceCtx keeps up to eight batches in flight, each in its own slot (buffers,
// pinned mirrors, captured hipGraph, events) on its own stream.
//
// Reference behaviour: sieve/cuda/mark.cu:22-91 (k_mark), survivors.cu:20-133
// (k_scan_counts, k_compact, k_primality), rung_scan3.cpp:711-784 (sweep_lane).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "es_device.h"
#include "es_montgomery.h"

namespace es {
namespace dev {

typedef unsigned char u8;

struct LayoutDev {
  u64 Mp;
  const u64 *lanes;
  const u32 *mask;
  const u32 *step;
  const u32 *magic;
  const u32 *c32;
  const uint4 *fc;           // per ordered filter: {s, floor(2^32/s), 2^32 mod s, 0}
  const u16 *off;            // per lane a row of ostride offsets (nf used, 0xFFFF after)
  u32 P, Pmagic, Pc32, nl, nf, ncov, mask_words, ostride;
};

// Per-batch scalars: survivor count, tile queue, stage-D queue, survivors stage D has
// finished, workgroups of k_pack done.
constexpr u32 kScalars = 5;
constexpr u32 kStamps = 6;   // (start, end) of k_sieve, k_mr_base2, k_rung / k_mr_bases

struct BatchHdr {
  u32 ntiles, nwin, has_sieve, seq;   // seq: the batch's number, posted back when it completes
  u32 nitems, pad[3];                 // nitems: k_sieve's work items (single tiles and groups)
};

// One work item of k_sieve: `count` consecutive tiles from `first` (head = tiles[first]).
// A tile long enough is an item alone; shorter ones are grouped (up to the launch's group
// size, at most kMaxGroup, and one tile's worth of positions), so they share one pass of
// the per-item work (init, counts, barriers) instead of each paying for it.
struct ItemDesc {
  TileDesc head;
  u32 first, count;
};
static_assert(sizeof(ItemDesc) == 32, "ItemDesc is shared with the host");
constexpr u32 kMaxGroup = 16;

// A sub-tile of a group as k_sieve keeps it in LDS: its descriptor and its first 32-position
// chunk in the group's map (the sub-tiles follow one another, each from a chunk boundary).
struct GroupSub {
  u64 k0;
  u32 nb, lane, win, layout, c0, pad;
};

// x mod s for s < 2^16 with mg = floor(2^32 / s) (2^32 - 1 for s = 1): the quotient
// estimate is low by at most one, so one conditional subtract finishes.
__device__ __forceinline__ u32 mod32(u32 x, u32 s, u32 mg) {
  const u32 q = __umulhi(x, mg);
  const u32 r = x - q * s;
  return r >= s ? r - s : r;
}

// x mod s for 64-bit x: (hi mod s) * (2^32 mod s) + (lo mod s), reduced; every
// intermediate stays below 2^32 because s < 2^16.
__device__ __forceinline__ u32 mod64(u64 x, u32 s, u32 mg, u32 c32) {
  const u32 h = mod32((u32)(x >> 32), s, mg);
  const u32 l = mod32((u32)x, s, mg);
  const u32 t = mod32(h * c32, s, mg) + l;
  return t >= s ? t - s : t;
}

typedef u8 __attribute__((address_space(3))) *LdsU8;

// Clears map bytes i, i + s, i + 2s, ... below nb: four LDS byte addresses advanced by
// one add each per step, then at most three more. i < 2^16 and s < 2^16, so nothing
// wraps. The empty asm makes the four addresses opaque to loop strength reduction,
// which would otherwise rebuild each as base + offset with an extra add per store.
__device__ __forceinline__ void clear_run(u8 *map, u32 i, u32 s, u32 nb) {
  const u32 base = (u32)(size_t)(LdsU8)map;
  const u32 end = base + nb;
  u32 a = base + i, b = a + s, c = b + s, d = c + s;
  const u32 s4 = 4u * s;
  while (d < end) {
    *(LdsU8)(size_t)a = 0;
    *(LdsU8)(size_t)b = 0;
    *(LdsU8)(size_t)c = 0;
    *(LdsU8)(size_t)d = 0;
    a += s4;
    b += s4;
    c += s4;
    d += s4;
    asm volatile("" : "+v"(a), "+v"(b), "+v"(c), "+v"(d));
  }
  if (a < end) *(LdsU8)(size_t)a = 0;
  if (b < end) *(LdsU8)(size_t)b = 0;
  if (c < end) *(LdsU8)(size_t)c = 0;
}

typedef const u32 __attribute__((address_space(1))) *GlobalU32;

// Marks ordered filters [from, to), one filter per thread. Filters are sorted by
// ascending stride, so 64 consecutive ones ("a bin") have similar hit counts. Wave w
// takes bins w, w + nwaves, ... (heaviest first, interleaved, so the waves end close
// together with no shared counter), kMarkGroup bins at a time: it loads the group's
// offsets and constants first, with one wait for all of them, then marks the group's
// bins. (Loading the next bin while marking the current one does not overlap: the wait
// for an LDS read issued before a run of stores is a wait for all of those stores too,
// since LDS operations complete in order, so every bin would drain the wave's stores.)
// kWide: the tile start k0 needs 64 bits; otherwise its remainders are one 32-bit
// reduction.
constexpr u32 kMarkGroup = 4;

template <bool kWide>
__device__ __forceinline__ void mark_range(u8 *map, u32 nb, const LayoutDev *L, const u16 *offrow, u64 k0,
                                           u32 from, u32 to, u32 wave, u32 nwaves, u32 lane64) {
  const u32 nbins = (to - from + 63u) >> 6;
  const u32 klo = (u32)k0;
  const GlobalU32 fc = (GlobalU32)L->fc;    // per filter {s, floor(2^32/s), 2^32 mod s, 0}
  for (u32 g = wave; g < nbins; g += kMarkGroup * nwaves) {
    u32 o[kMarkGroup], fs[kMarkGroup], fm[kMarkGroup], fr[kMarkGroup];
#pragma unroll
    for (u32 q = 0; q < kMarkGroup; ++q) {
      const u32 bq = g + q * nwaves, j = from + (bq << 6) + lane64;
      const bool ok = bq < nbins && j < to;
      const u32 jj = ok ? j : from;        // every load in bounds, so none is conditional
      o[q] = ok ? (u32)offrow[jj] : 0xFFFFu;
      fs[q] = fc[4 * jj];
      fm[q] = fc[4 * jj + 1];
      fr[q] = kWide ? fc[4 * jj + 2] : 0u;
    }
#pragma unroll
    for (u32 q = 0; q < kMarkGroup; ++q) {
      if (o[q] != 0xFFFFu) {
        const u32 km = kWide ? mod64(k0, fs[q], fm[q], fr[q]) : mod32(klo, fs[q], fm[q]);
        clear_run(map, o[q] >= km ? o[q] - km : o[q] + fs[q] - km, fs[q], nb);
      }
    }
  }
}

// Four map bytes (0 or 1 each) as four bits.
__device__ __forceinline__ u32 nibble(u32 x) { return (((x & 0x01010101u) * 0x01020408u) >> 24) & 0xFu; }

// Alive positions of one 32-position chunk as a bit word, from two 16-byte LDS reads.
__device__ __forceinline__ u32 chunk_bits(const u32 *map32, u32 c) {
  const uint4 a = ((const uint4 *)map32)[2 * c], b = ((const uint4 *)map32)[2 * c + 1];
  return nibble(a.x) | nibble(a.y) << 4 | nibble(a.z) << 8 | nibble(a.w) << 12 | nibble(b.x) << 16 |
         nibble(b.y) << 20 | nibble(b.z) << 24 | nibble(b.w) << 28;
}

// Inclusive prefix sum across the 64 lanes of a wave with DPP moves (no LDS): within
// each row of 16 lanes by shifts of 1, 2, 3, then 4 and 8 onto the upper banks, then
// lane 15 of rows 0 and 2 into rows 1 and 3, and lane 31 into rows 2 and 3. A lane the
// row or bank mask leaves out receives 0.
__device__ __forceinline__ u32 wave_scan(u32 x) {
  u32 t = x;
  t += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, true);   // row_shr:1
  t += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, true);   // row_shr:2
  t += (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x113, 0xf, 0xf, true);   // row_shr:3
  t += (u32)__builtin_amdgcn_update_dpp(0, (int)t, 0x114, 0xf, 0xe, false);  // row_shr:4, banks 1-3
  t += (u32)__builtin_amdgcn_update_dpp(0, (int)t, 0x118, 0xf, 0xc, false);  // row_shr:8, banks 2-3
  t += (u32)__builtin_amdgcn_update_dpp(0, (int)t, 0x142, 0xa, 0xf, false);  // row_bcast:15, rows 1, 3
  t += (u32)__builtin_amdgcn_update_dpp(0, (int)t, 0x143, 0xc, 0xf, false);  // row_bcast:31, rows 2, 3
  return t;
}

// Sum over the wave (every lane active), uniform.
__device__ __forceinline__ u32 wave_sum(u32 v) { return (u32)__builtin_amdgcn_readlane((int)wave_scan(v), 63); }

// The tile is a byte map in LDS: byte k says whether position k0 + k is still alive.
// A mark is a plain byte store of 0 -- idempotent, so concurrent stores need no atomics.
// All LDS is one dynamic allocation: the map at offset 0 (so a mark's address is its
// position), then the tile's row of filter offsets, then its lane's periodic wheel mask,
// then a few scalars. lds_words, off_words and mask_words count the map, the offset row
// and the mask row in 32-bit words. The offset rows (one per lane, megabytes in all) do
// not stay in L2, so each tile streams its row into LDS with coalesced 16-byte loads,
// instead of every bin waiting on its own global load. kRowInLds = false (with
// off_words = 0) is the form for a row that does not fit beside the map: it is read from
// global memory.
//
// Counters and survivors are placed wave by wave, with no workgroup-wide step: each wave
// adds its {positions, alive after covers, survivors} to the tile's slot of tile_counts
// (zeroed by k_stage) with atomics whose results nothing waits for, and a wave with
// survivors reserves their places with one atomic on the batch's survivor count. The
// host sums the tile slots per window; the survivor list's order depends on scheduling
// and every output path sorts it.
// Kernel times on the device's constant clock: the first workgroup to start lowers
// stamp[0], the last wave (or workgroup) to finish raises stamp[1]. The host reads the
// pair back with the batch (k_pack copies it), for eager launches and graph replays alike.
// k_stage sets stamp[0] to ~0 and stamp[1] to 0 before every batch; null skips the stamps.
__device__ __forceinline__ void stamp_start(u64 *stamp) {
  if (stamp && threadIdx.x == 0) atomicMin((unsigned long long *)stamp, (unsigned long long)__builtin_amdgcn_s_memrealtime());
}
__device__ __forceinline__ void stamp_end(u64 *stamp, bool writer) {
  if (stamp && writer) atomicMax((unsigned long long *)stamp + 1, (unsigned long long)__builtin_amdgcn_s_memrealtime());
}

// 0, in a register the compiler cannot see through. An atomic on a uniform address has its
// result broadcast to the wave at once (so the wave waits for it right after issuing it);
// offsetting the address by this keeps a claim issued ahead of use asynchronous.
__device__ __forceinline__ u32 opaque_zero() {
  u32 z;
  asm volatile("v_mov_b32 %0, 0" : "=v"(z));
  return z;
}

// uint4 words of the next tile's offset row each thread of k_sieve holds in registers
// between its loads and its LDS stores (a row of up to kRowPrefetch x blockDim uint4;
// a longer row is staged at the start of its tile instead).
constexpr u32 kRowPrefetch = 4;

// Lanes 0..ns-1 of one wave store a group's table from their sub-tile descriptors
// (every lane of the wave calls it: the chunk offsets are a scan across the wave).
__device__ __forceinline__ void stage_sub(GroupSub *tab, const TileDesc &d, u32 ns, u32 lane) {
  const u32 chunks = lane < ns ? (d.nbits + 31u) >> 5 : 0u;
  const u32 incl = wave_scan(chunks);
  if (lane < ns) tab[lane] = GroupSub{d.k0, d.nbits, d.lane, d.win, d.layout, incl - chunks, 0u};
}

// Work item i, loaded as two 16-byte words (a conditional copy of the struct would go
// through scratch memory).
__device__ __forceinline__ ItemDesc load_item(const ItemDesc *items, u32 i) {
  const uint4 a = ((const uint4 *)(items + i))[0], b = ((const uint4 *)(items + i))[1];
  ItemDesc d;
  d.head.k0 = (u64)a.x | ((u64)a.y << 32);
  d.head.nbits = a.z;
  d.head.lane = a.w;
  d.head.win = b.x;
  d.head.layout = b.y;
  d.first = b.z;
  d.count = b.w;
  return d;
}

// The sub-tile of a group's table (ns entries) whose chunks include chunk c.
__device__ __forceinline__ u32 sub_of(const GroupSub *tab, u32 ns, u32 c) {
  u32 lo = 0, hi = ns - 1;
  while (lo < hi) {
    const u32 mid = (lo + hi + 1u) >> 1;
    if (tab[mid].c0 <= c) lo = mid;
    else hi = mid - 1u;
  }
  return lo;
}

// Chunk c of the byte map from its 32 alive bits: one byte (0 or 1) per position.
__device__ __forceinline__ void store_chunk(u32 *map32, u32 c, u32 word) {
  auto bytes = [&](u32 q) { return (((word >> (4 * q)) & 0xFu) * 0x00204081u) & 0x01010101u; };
  ((uint4 *)map32)[2 * c] = make_uint4(bytes(0), bytes(1), bytes(2), bytes(3));
  ((uint4 *)map32)[2 * c + 1] = make_uint4(bytes(4), bytes(5), bytes(6), bytes(7));
}

template <bool kPersistent, bool kRowInLds>
__global__ __launch_bounds__(1024) void k_sieve(const TileDesc *__restrict__ tiles,
                                                const ItemDesc *__restrict__ items,
                                                const BatchHdr *__restrict__ hdr,
                                                const LayoutDev *__restrict__ lays,
                                                uint4 *__restrict__ tile_counts, u64 *__restrict__ surv_p,
                                                u32 *__restrict__ surv_w, u32 cap,
                                                u32 *__restrict__ surv_count, u32 *__restrict__ queue,
                                                u32 lds_words, u32 off_words, u32 mask_words, u64 *stamp) {
  extern __shared__ __attribute__((aligned(16))) u32 lds[];
  // A group's sub-tiles (two tables: the next group's is stored while the current one's
  // compaction may still read its own) and their counters.
  __shared__ GroupSub gsub[2][kMaxGroup];
  __shared__ u32 gcnt[kMaxGroup][3];
  u32 *map32 = lds;
  u8 *map = (u8 *)lds;
  u16 *soff = (u16 *)(lds + lds_words);
  u32 *smask = lds + lds_words + off_words;
  u32 *sc = smask + mask_words;   // [2] the item after next
  const u32 tid = threadIdx.x, bs = blockDim.x;
  const u32 lane64 = tid & 63u;
  const u32 nitems = hdr->nitems;
  const bool sieve = hdr->has_sieve != 0;
  const u32 wave = tid >> 6, nwaves = (bs + 63u) >> 6;
  // Workgroups are dispatched in order of their index, so workgroup 0's start is the
  // kernel's (one atomic instead of one per workgroup at the same instant).
  if (blockIdx.x == 0) stamp_start(stamp);
  // Persistent form: workgroup g takes items g and gridDim.x + g first, and the queue
  // hands out items from 2 gridDim.x on (every workgroup claiming from one counter at
  // the start would serialize on that address). A workgroup loads the next item's
  // descriptor at the top of the current one, claims the item after next once the
  // claiming wave's marking loads are done (the wave then has no load to wait for until
  // the end of the item, so the claim's latency stays hidden) and, once the current
  // item's marking is done, loads what the next item needs into registers while it counts
  // and compacts, storing it into LDS at the end of the item: a single tile's offset and
  // mask rows, or a group's sub-tile table. The claimer is lane 0 of the last wave.
  const u32 claimer = (((bs + 63u) >> 6) - 1u) * 64u;
  u32 t = blockIdx.x, tnext = gridDim.x + blockIdx.x;
  ItemDesc it{};
  if (t < nitems) it = load_item(items, t);
  bool row_ready = false;   // a single tile's offset and mask rows are in LDS already
  bool sub_ready = false;   // a group's sub-tile table is in LDS already
  // Group counters start at zero (ordered before any count by the first item's staging
  // barrier); each group's flush zeroes them again for the next.
  if (tid < kMaxGroup) gcnt[tid][0] = gcnt[tid][1] = gcnt[tid][2] = 0u;
  for (u32 iter = 0;; ++iter) {
    if (t >= nitems) break;
    u32 claim = 0;
    ItemDesc itn{};
    if (kPersistent && tnext < nitems) itn = load_item(items, tnext);
    const u32 cur = iter & 1u;
    const TileDesc td = it.head;
    const bool single = it.count == 1;
    const u32 ns = it.count;

    // ---- staging at the item's top (the first item of a workgroup, or what could not be
    // loaded ahead)
    if (single) {
      if (!row_ready) {
        const LayoutDev *L = lays + td.layout;
        if (kRowInLds) {
          const uint4 *src = (const uint4 *)(L->off + (size_t)td.lane * L->ostride);
          uint4 *dst = (uint4 *)soff;
          for (u32 i = tid, nv = L->ostride >> 3; i < nv; i += bs) dst[i] = src[i];
        }
        const u32 *msrc = L->mask + (size_t)td.lane * L->mask_words;
        for (u32 i = tid; i < L->mask_words; i += bs) smask[i] = msrc[i];
        __syncthreads();
      }
    } else {
      if (!sub_ready) {
        if (tid < 64) {
          TileDesc d{};
          if (tid < ns) d = tiles[it.first + tid];
          stage_sub(gsub[cur], d, ns, tid);
        }
        __syncthreads();
      }
    }

    // ---- init: wheel mask and tail mask, expanded to one byte per position
    u32 my_pos = 0;
    u32 nct;   // chunks of the item's map
    if (single) {
      const LayoutDev *L = lays + td.layout;
      const u32 nb = td.nbits, P = L->P;
      nct = (nb + 31u) >> 5;
      const u32 kP = P == 1 ? 0u : mod64(td.k0, P, L->Pmagic, L->Pc32);
      for (u32 c = tid; c < nct; c += bs) {
        u32 word = 0xFFFFFFFFu;
        if (P != 1) {
          const u32 b = mod32(kP + 32u * c, P, L->Pmagic);
          word = __builtin_amdgcn_alignbit(smask[(b >> 5) + 1], smask[b >> 5], b & 31u);
        }
        if (c == nct - 1 && (nb & 31u)) word &= (1u << (nb & 31u)) - 1u;
        my_pos += __popc(word);
        store_chunk(map32, c, word);
      }
    } else {
      const GroupSub &last = gsub[cur][ns - 1];
      nct = last.c0 + ((last.nb + 31u) >> 5);
      for (u32 c = tid; c < nct; c += bs) {
        const u32 j = sub_of(gsub[cur], ns, c);
        const GroupSub &g = gsub[cur][j];
        const LayoutDev *L = lays + g.layout;
        const u32 P = L->P, cl = c - g.c0, ncj = (g.nb + 31u) >> 5;
        u32 word = 0xFFFFFFFFu;
        if (P != 1) {
          const u32 kP = mod64(g.k0, P, L->Pmagic, L->Pc32);
          const u32 b = mod32(kP + 32u * cl, P, L->Pmagic);
          const u32 *mrow = L->mask + (size_t)g.lane * L->mask_words;
          word = __builtin_amdgcn_alignbit(mrow[(b >> 5) + 1], mrow[b >> 5], b & 31u);
        }
        if (cl == ncj - 1 && (g.nb & 31u)) word &= (1u << (g.nb & 31u)) - 1u;
        atomicAdd(&gcnt[j][0], (u32)__popc(word));
        store_chunk(map32, c, word);
      }
    }
    __syncthreads();

    // ---- stage B: certified filters (a single tile reads its offset row from LDS where
    // it was staged, a group reads its sub-tiles' rows from global memory)
    auto mark = [&](bool covers) {
      if (single) {
        const LayoutDev *L = lays + td.layout;
        const u32 from = covers ? 0u : L->ncov, to = covers ? L->ncov : L->nf;
        const bool wide = (td.k0 >> 32) != 0;
        const u16 *row = kRowInLds ? (const u16 *)soff : L->off + (size_t)td.lane * L->ostride;
        if (wide) mark_range<true>(map, td.nbits, L, row, td.k0, from, to, wave, nwaves, lane64);
        else mark_range<false>(map, td.nbits, L, row, td.k0, from, to, wave, nwaves, lane64);
      } else {
        for (u32 j = 0; j < ns; ++j) {
          const GroupSub &g = gsub[cur][j];
          const LayoutDev *L = lays + g.layout;
          const u32 from = covers ? 0u : L->ncov, to = covers ? L->ncov : L->nf;
          const u16 *row = L->off + (size_t)g.lane * L->ostride;
          if ((g.k0 >> 32) != 0) mark_range<true>(map + 32u * g.c0, g.nb, L, row, g.k0, from, to, wave, nwaves, lane64);
          else mark_range<false>(map + 32u * g.c0, g.nb, L, row, g.k0, from, to, wave, nwaves, lane64);
        }
      }
    };
    mark(true);
    if (kPersistent && !sieve && tid == claimer) claim = atomicAdd(queue + opaque_zero(), 1u);
    __syncthreads();
    u32 my_after = 0;
    if (sieve) {
      for (u32 c = tid; c < nct; c += bs) {
        const u32 n = __popc(chunk_bits(map32, c));
        if (single) my_after += n;
        else if (n) atomicAdd(&gcnt[sub_of(gsub[cur], ns, c)][1], n);
      }
      // Every wave's count of the covers-only map completes before any wave's stage B'
      // store: a wave's chunks are not the bytes it marks.
      __syncthreads();
      // ---- stage B': small primes
      mark(false);
      if (kPersistent && tid == claimer) claim = atomicAdd(queue + opaque_zero(), 1u);
      __syncthreads();
    }
    // ---- marking is over and the rows dead: load what the next item needs into
    // registers now, store it into LDS after the compaction
    u32 nv_next = 0, nm_next = 0, mpre = 0;
    uint4 pre0 = make_uint4(0, 0, 0, 0), pre1 = pre0, pre2 = pre0, pre3 = pre0;
    TileDesc dnext{};
    if (kPersistent) {
      row_ready = false;
      sub_ready = false;
      if (tnext < nitems && itn.count == 1) {
        const TileDesc &tdn = itn.head;
        const LayoutDev *Ln = lays + tdn.layout;
        const u32 nv = kRowInLds ? Ln->ostride >> 3 : 0u, nm = Ln->mask_words;
        if (nv <= kRowPrefetch * bs && nm <= bs) {
          const uint4 *src = (const uint4 *)(Ln->off + (size_t)tdn.lane * Ln->ostride);
          if (tid < nv) pre0 = src[tid];
          if (tid + bs < nv) pre1 = src[tid + bs];
          if (tid + 2 * bs < nv) pre2 = src[tid + 2 * bs];
          if (tid + 3 * bs < nv) pre3 = src[tid + 3 * bs];
          if (tid < nm) mpre = Ln->mask[(size_t)tdn.lane * nm + tid];
          nv_next = nv;
          nm_next = nm;
          row_ready = true;
        }
      } else if (tnext < nitems) {
        if (tid < itn.count) dnext = tiles[itn.first + tid];
        sub_ready = true;
      }
    }

    // ---- counters and survivors, wave by wave (inclusive survivor scan within the wave)
    u32 my_surv = 0;
    for (u32 c = tid; c < nct; c += bs) {
      const u32 n = __popc(chunk_bits(map32, c));
      my_surv += n;
      if (!single && n) atomicAdd(&gcnt[sub_of(gsub[cur], ns, c)][2], n);
    }
    if (!sieve) my_after = my_surv;
    const u32 incl = wave_scan(my_surv);
    const u32 wsur = (u32)__builtin_amdgcn_readlane((int)incl, 63);
    if (single) {
      const u32 wpos = wave_sum(my_pos), waft = sieve ? wave_sum(my_after) : wsur;
      if (lane64 == 0) {
        u32 *slot = (u32 *)(tile_counts + it.first);
        atomicAdd(slot, wpos);
        if (waft) atomicAdd(slot + 1, waft);
        if (wsur) atomicAdd(slot + 2, wsur);
      }
    }
    if (wsur) {
      u32 base = 0;
      if (lane64 == 0) base = atomicAdd(surv_count, wsur);
      base = (u32)__builtin_amdgcn_readfirstlane((int)base);
      if (my_surv) {
        u32 idx = base + incl - my_surv;
        for (u32 ch = tid; ch < nct; ch += bs) {
          u32 x = chunk_bits(map32, ch);
          if (!x) continue;
          u64 c, Mp, k;
          u32 win;
          if (single) {
            const LayoutDev *L = lays + td.layout;
            c = L->lanes[td.lane];
            Mp = L->Mp;
            k = td.k0 + 32ull * ch;
            win = td.win;
          } else {
            const GroupSub &g = gsub[cur][sub_of(gsub[cur], ns, ch)];
            const LayoutDev *L = lays + g.layout;
            c = L->lanes[g.lane];
            Mp = L->Mp;
            k = g.k0 + 32ull * (ch - g.c0);
            win = g.win;
          }
          while (x) {
            const u32 i = __builtin_ctz(x);
            x &= x - 1;
            if (idx < cap) {
              surv_p[idx] = c + Mp * (k + i);
              surv_w[idx] = win;
            }
            ++idx;
          }
        }
      }
    }
    if (!single) {
      // A group's counters are complete once every wave has counted: one store per
      // sub-tile (k_stage zeroed the slots, and each sub-tile belongs to one item). The
      // thread that reads a slot zeroes it; the item's last barrier orders that before the
      // next group's counts.
      __syncthreads();
      if (tid < kMaxGroup) {
        const u32 pos = gcnt[tid][0], aft = gcnt[tid][1], sur = gcnt[tid][2];
        if (tid < ns) tile_counts[it.first + tid] = make_uint4(pos, sieve ? aft : sur, sur, 0u);
        gcnt[tid][0] = gcnt[tid][1] = gcnt[tid][2] = 0u;
      }
    }
    if (!kPersistent) break;
    if (row_ready) {
      uint4 *dst = (uint4 *)soff;
      if (tid < nv_next) dst[tid] = pre0;
      if (tid + bs < nv_next) dst[tid + bs] = pre1;
      if (tid + 2 * bs < nv_next) dst[tid + 2 * bs] = pre2;
      if (tid + 3 * bs < nv_next) dst[tid + 3 * bs] = pre3;
      if (tid < nm_next) smask[tid] = mpre;
    }
    if (sub_ready && tid < 64) stage_sub(gsub[cur ^ 1u], dnext, itn.count, tid);
    if (tid == claimer) sc[2] = 2 * gridDim.x + claim;
    __syncthreads();   // the next item reuses the map and the scalars, and finds its rows or table staged
    t = tnext;
    it = itn;
    tnext = sc[2];
  }
  // Every wave leaves the loop at the same barrier (persistent) or after the item's last
  // barrier, so thread 0's time is the workgroup's.
  stamp_end(stamp, threadIdx.x == 0);
}

// Stages one batch: zeroes the survivor count, the queue heads and the tile counters,
// and copies the batch header, the tile descriptors and k_sieve's work items from pinned host memory (written
// by the host before the launch) into device memory. A kernel rather than a memset and two copies, so a
// captured graph holds no memset or copy node and no copy engine has to be waited for
// between batches. The tile count comes from the header itself, so one captured graph
// serves every batch that fits the buffers; the host keeps ntiles within the buffers'
// capacity. Every other per-batch output is written in full by the kernels.
__global__ void k_stage(const BatchHdr *__restrict__ h_hdr, const uint2 *__restrict__ h_tiles,
                        BatchHdr *__restrict__ hdr, uint2 *__restrict__ tiles, const uint4 *__restrict__ h_items,
                        uint4 *__restrict__ items, uint4 *__restrict__ tile_counts,
                        u32 *__restrict__ scalars, u64 *__restrict__ board_words, u32 nboard_words,
                        u64 *__restrict__ h_sync, u64 *__restrict__ stamps, u32 *__restrict__ active) {
  const u32 g = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  // The header crosses PCIe once per workgroup, not once per wave.
  __shared__ BatchHdr sh;
  if (threadIdx.x == 0) sh = *h_hdr;
  __syncthreads();
  const BatchHdr h = sh;
  if (g == 0) {
    *hdr = h;
    h_sync[1] = __builtin_amdgcn_s_memrealtime();   // the batch's start on the device clock
    // Batches between their k_stage and the end of their k_pack: stage D's board helpers
    // read it to make room for the other batches' kernels.
    atomicAdd(active, 1u);
  }
  if (g < kStamps) stamps[g] = (g & 1) ? 0ull : ~0ull;   // kernel (start, end) pairs
  if (g < kScalars) scalars[g] = 0;   // survivor count, tile queue, stage-D queue, survivors finished,
                                      // workgroups of k_pack done
  for (u32 i = g; i < nboard_words; i += stride) board_words[i] = 0;   // stage D's board of posted splits
  const u32 n = h.ntiles * (u32)(sizeof(TileDesc) / sizeof(uint2));
  for (u32 i = g; i < n; i += stride) tiles[i] = h_tiles[i];
  for (u32 i = g; i < 2 * h.nitems; i += stride) items[i] = h_items[i];
  for (u32 i = g; i < h.ntiles; i += stride) tile_counts[i] = make_uint4(0, 0, 0, 0);   // k_sieve adds per wave
}

// ------------------------------------------------------------------ Miller-Rabin

// a * b * 2^-64 mod n (a, b < n, n odd), ninv = -n^{-1} mod 2^64: word-by-word
// Montgomery with 32-bit multiplies (es_montgomery.h), 10 quarter-rate multiplies.
__device__ __forceinline__ u64 mont_mul(u64 a, u64 b, u64 n, u64 ninv) {
  return mont::mul(a, b, n, (u32)ninv);
}

__device__ __forceinline__ u64 add_mod(u64 a, u64 b, u64 n) {
  const u64 s = a + b;
  return (s < a || s >= n) ? s - n : s;
}

// a * b mod n on FP64 values holding integers, n < 2^50, with signed residues (full-rate
// FMAs instead of the quarter-rate 32-bit multiplies of 64-bit Montgomery): h + l == a*b
// exactly (FMA residual); q = rint(a*b/n) makes h - q*n exact (one FMA), and
// r = a*b - q*n has |r| <= (1/2 + 3*2^-53*|a*b|/n) * n. So |a|, |b| < n + 1024 give
// |r| < 7n/8 (a chain of products stays below n with no correction), and |a| <= 2n,
// |b| < 2n give |r| < 2n (the product of differences stays below 2n with no correction
// either).
__device__ __forceinline__ double mulmod_sgn(double a, double b, double nd, double ninv) {
  const double h = a * b;
  const double l = __fma_rn(a, b, -h);
  const double q = rint(h * ninv);
  return __fma_rn(-q, nd, h) + l;
}

constexpr u64 kF64Mod = 1ull << 50;

// n odd > 37; d odd with n - 1 = d 2^s; base a small; one = 2^64 mod n. The windowed
// left-to-right test of es_montgomery.h (one Montgomery squaring per exponent bit, and a
// multiplication by a small power of a per digit).
__device__ bool strong_probable(u64 n, u64 ninv, u64 one, u64 a, u64 d, u32 s) {
  return mont::strong_test(n, (u32)a, d, s, (u32)ninv, one, 1.0 / (double)n);
}

// The same test for n < 2^50 in signed FP64 residues: every operand of mulmod_sgn stays
// in (-n, n), so square-and-multiply needs no correction, and x == 1, x == -1 (mod n)
// are each one of two representatives in (-n, n).
__device__ bool strong_probable_sgn(u64 n, u32 a, u64 d, u32 s) {
  const double nd = (double)n, ninv = 1.0 / nd, m1 = nd - 1.0, p1 = 1.0 - nd;
  double x = 1.0, b = (double)a;
  for (u64 e = d; e; e >>= 1) {
    if (e & 1) x = mulmod_sgn(x, b, nd, ninv);
    b = mulmod_sgn(b, b, nd, ninv);
  }
  if (x == 1.0 || x == p1 || x == -1.0 || x == m1) return true;
  for (u32 i = 1; i < s; ++i) {
    x = mulmod_sgn(x, x, nd, ninv);
    if (x == -1.0 || x == m1) return true;
  }
  return false;
}

__constant__ u64 c_small_p[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};

// q^{-1} mod 2^64 and floor((2^64-1)/q) for the odd primes above: n is divisible by q
// iff n * q^{-1} (mod 2^64) <= floor((2^64-1)/q), one multiply instead of a division.
struct DivTest { u64 inv, lim; };
constexpr u64 inv_pow2(u64 a) {
  u64 x = a;
  for (int i = 0; i < 6; ++i) x *= 2 - a * x;
  return x;
}
__constant__ DivTest c_div[11] = {
    {inv_pow2(3), ~0ull / 3},   {inv_pow2(5), ~0ull / 5},   {inv_pow2(7), ~0ull / 7},
    {inv_pow2(11), ~0ull / 11}, {inv_pow2(13), ~0ull / 13}, {inv_pow2(17), ~0ull / 17},
    {inv_pow2(19), ~0ull / 19}, {inv_pow2(23), ~0ull / 23}, {inv_pow2(29), ~0ull / 29},
    {inv_pow2(31), ~0ull / 31}, {inv_pow2(37), ~0ull / 37}};

// Verdict codes written by the two Miller-Rabin passes.
constexpr u8 kComposite = 0, kPrime = 1, kPending = 2;

// Pass 1, one thread per value: the trial prologue by the first twelve primes and the
// base-2 test. Values that pass are marked pending and keep 2^64 mod n for pass 2.
__global__ void k_mr_base2(const u64 *__restrict__ vals, const u32 *__restrict__ count, u32 cap,
                           u8 *__restrict__ flags, u64 *__restrict__ one_out, u64 *stamp = nullptr) {
  const u32 n_ = min(*count, cap);
  // Only workgroups with values stamp: a same-address atomic from each of the grid's idle
  // workgroups would serialize at the memory side and hold the kernel's end back.
  const bool works = blockIdx.x == 0 || blockIdx.x * blockDim.x < n_;
  if (works) stamp_start(stamp);
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n_; i += gridDim.x * blockDim.x) {
    const u64 n = vals[i];
    u8 v;
    if (n < 2) v = kComposite;
    else if (!(n & 1)) v = n == 2 ? kPrime : kComposite;
    else if (n <= 37) {
      v = kComposite;
      for (int k = 1; k < 12; ++k) if (n == c_small_p[k]) v = kPrime;
    } else {
      bool divisible = false;
      for (int k = 0; k < 11; ++k) divisible |= n * c_div[k].inv <= c_div[k].lim;
      if (divisible) v = kComposite;
      else if (n < kF64Mod) {
        u64 d = n - 1;
        const u32 s = (u32)__builtin_ctzll(d);
        d >>= s;
        v = strong_probable_sgn(n, 2, d, s) ? kPending : kComposite;
      } else {
        u64 ninv = n;                         // n^{-1} mod 2^64 (Newton), then negate
        for (int k = 0; k < 6; ++k) ninv *= 2 - n * ninv;
        ninv = 0 - ninv;
        const u64 one = (0 - n) % n;          // 2^64 mod n
        u64 d = n - 1;
        const u32 s = (u32)__builtin_ctzll(d);
        d >>= s;
        v = strong_probable(n, ninv, one, 2, d, s) ? kPending : kComposite;
        one_out[i] = one;
      }
    }
    flags[i] = v;
  }
  if (stamp && works) {
    __syncthreads();
    stamp_end(stamp, threadIdx.x == 0);
  }
}

// Pass 2, one thread per (value, base) for the eleven odd prime bases: any failing
// base writes the composite verdict (all writers write the same byte value). Pending
// values that survive every base are prime.
__global__ void k_mr_bases(const u64 *__restrict__ vals, const u32 *__restrict__ count, u32 cap,
                           u8 *__restrict__ flags, const u64 *__restrict__ one_in, u64 *stamp = nullptr) {
  const u32 n_ = min(*count, cap);
  const u32 total = n_ * 11u;
  const bool works = blockIdx.x == 0 || blockIdx.x * blockDim.x < total;   // as in k_mr_base2
  if (works) stamp_start(stamp);
  for (u32 g = blockIdx.x * blockDim.x + threadIdx.x; g < total; g += gridDim.x * blockDim.x) {
    const u32 i = g / 11u, b = g - i * 11u + 1u;
    if (flags[i] != kPending) continue;
    const u64 n = vals[i];
    u64 d = n - 1;
    const u32 s = (u32)__builtin_ctzll(d);
    d >>= s;
    bool ok;
    if (n < kF64Mod) {
      ok = strong_probable_sgn(n, (u32)c_small_p[b], d, s);
    } else {
      u64 ninv = n;
      for (int k = 0; k < 6; ++k) ninv *= 2 - n * ninv;
      ninv = 0 - ninv;
      ok = strong_probable(n, ninv, one_in[i], c_small_p[b], d, s);
    }
    if (!ok) flags[i] = kComposite;
  }
  if (stamp && works) {
    __syncthreads();
    stamp_end(stamp, threadIdx.x == 0);
  }
}

// Scratch of the allocation-free primality entry point, one copy per GPU for the life of
// the process: [0] the value count (u32), [kScratchOne ..) 2^64 mod n per value, written
// by k_mr_base2 before k_mr_bases reads it. Host code reaches it only through the device
// address hipGetSymbolAddress returns for the current device.
constexpr size_t kScratchOne = 32;   // 256 bytes in: the per-value words stay 256-aligned
static __device__ __align__(256) u64 scratch[kScratchOne + kNoAllocMaxValues];

__global__ void k_set_count(u32 *count, u32 n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = n;
}

// Verdicts to the public form: 1 = prime (kPrime, or kPending after every base), 0 = not.
__global__ void k_flags_to_prime(u8 *__restrict__ flags, u32 n) {
  for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    flags[i] = flags[i] != kComposite ? 1 : 0;
}

// ------------------------------------------------------------------ stage D on the device
//
// One workgroup of kRungWaves waves per prime (rung_scan3.cpp:574-603). The candidate
// rungs r = r0, r0 + 4, ... <= 255 (r0 == -p mod 4) have A_j = (p + r0)/4 + j, a run of
// consecutive integers, so one workgroup-wide pass over the odd primes below 2^16
// computes A_0 mod q once per prime and records every q | A_j for all 64 rungs at once
// (a sieve instead of per-rung trial division). Waves then take rungs in order from an
// LDS counter: skip gcd(pA mod r, r) > 1, finish A_j's factorization (exponents of the
// sieved primes by exact division, then Miller-Rabin, a square test and a 64-walk Brent
// rho for the cofactor, whose prime factors all exceed 65521), and search the least
// witness u | (pA)^2 with u == -pA (mod r), u <= pA as a per-residue minimum carried
// prime power by prime power. The least rung with a witness wins; a wave stops as soon
// as a smaller rung is settled. The certificate is a function of p alone, so computing
// it here or on the host gives the same line; a prime the device declines is finished
// on the host.

typedef unsigned __int128 u128;

struct RungOut {
  u64 A, u_lo, u_hi;
  u32 r, status;
};
constexpr u32 kRungFound = 1, kRungNone = 2, kRungHost = 3;
// Stage E's result in a RungOut: A = c, u_lo = k, r = a | d << 8 ("SOLVED p a c d k").
constexpr u32 kRungDirect = 4;
constexpr u32 kDirectPairs = 64;              // (a, d) in 1..8 x 1..8, a major
constexpr u64 kDirectNMaxDev = 18446744065119617025ull;   // (2^32 - 1)^2, as the host's
constexpr u32 kRungWaves = 8;     // waves per prime = rungs in flight (two per SIMD)
constexpr u32 kRungSlots = 64;    // candidate rungs r0 + 4j <= 255
constexpr u32 kSmallFac = 16;     // odd primes < 2^16 dividing one A < 2^62: at most 15
constexpr u32 kSievePer = 15;     // table loads in flight per sieving thread (6541 primes / 448 threads)

struct TrialTab {
  const uint2 *qm;  // odd primes 3 .. 65521 ascending: {q | (2^32 mod q) << 16, floor(2^32/q)}
  const u64 *inv;   // q^{-1} mod 2^64
  const u64 *lim;   // floor((2^64-1)/q)
  u32 n;
};

// Lane-to-lane hand-off through LDS inside one wave: a wave's LDS operations complete in
// issue order, so this only has to stop the compiler from moving accesses across it.
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// a * b <= q (a, b, q < 2^127), exactly, without a 128-bit division.
__device__ __forceinline__ bool mul_le(u128 a, u128 b, u128 q, u128 &out) {
  const u64 ah = (u64)(a >> 64), bh = (u64)(b >> 64);
  if (ah && bh) return false;
  const u128 big = ah ? a : b;
  const u64 small = ah ? (u64)b : (u64)a;
  const u128 t0 = (u128)(u64)big * small;
  const u128 t1 = (u128)(u64)(big >> 64) * small;
  if (t1 >> 64) return false;
  const u128 sum = t0 + (t1 << 64);
  if (sum < t0 || sum > q) return false;
  out = sum;
  return true;
}

__device__ __forceinline__ u64 gcd_u64(u64 a, u64 b) {
  if (!a) return b;
  if (!b) return a;
  const int sh = __builtin_ctzll(a | b);
  a >>= __builtin_ctzll(a);
  do {
    b >>= __builtin_ctzll(b);
    if (a > b) { const u64 t = a; a = b; b = t; }
    b -= a;
  } while (b);
  return a << sh;
}

// n^{-1} mod 2^64 for odd n (Newton: each step doubles the correct low bits).
__device__ __forceinline__ u64 inv64(u64 n) {
  u64 x = n;
  for (int k = 0; k < 6; ++k) x *= 2 - n * x;
  return x;
}

__device__ __forceinline__ u64 neg_inv(u64 n) { return 0 - inv64(n); }

// a * 2^64 mod n for a < 2^32, by double-and-add from one = 2^64 mod n.
__device__ __forceinline__ u64 small_to_mont(u32 a, u64 one, u64 n) {
  u64 am = 0;
  for (int bit = 31; bit >= 0; --bit) {
    am = add_mod(am, am, n);
    if ((a >> bit) & 1) am = add_mod(am, one, n);
  }
  return am;
}

// Deterministic primality of odd n > 37 with the twelve prime bases tested on twelve
// lanes at once; every lane gets the verdict.
__device__ bool wave_is_prime(u64 n, u32 lane) {
  u64 d = n - 1;
  const u32 s = (u32)__builtin_ctzll(d);
  d >>= s;
  bool ok = true;
  if (lane < 12) {
    if (n < kF64Mod) {
      ok = strong_probable_sgn(n, (u32)c_small_p[lane], d, s);
    } else {
      const u64 ninv = neg_inv(n), one = (0 - n) % n;
      ok = strong_probable(n, ninv, one, c_small_p[lane], d, s);
    }
  }
  return __ballot(!ok) == 0;
}

constexpr u64 kRhoFail = 0, kRhoStop = 1;

// Block-wide help with one rho at a time. The wave that needs a cofactor split (the
// owner) publishes n; waves done with their rungs run walks of their own on it and post
// the first factor they find, which the owner checks and takes.
struct RhoShare {
  u64 n, factor;
  u32 gen, busy, done, pad;
};

// Grid-wide posts of long cofactor splits. A wave whose walks have not split n after a
// few Brent windows posts n in a free slot; workgroups that found the survivor queue empty
// run walks of their own on posted values and post the first factor they find, which the
// owner checks (it must divide n) and takes. Correctness never depends on the board: every
// factor is verified, a slot may be reused at any time, and each side only reads single
// 64-bit words with relaxed device-scope atomics.
struct RhoBoard {
  u64 n;        // the value being split; 0 = free
  u64 factor;   // a factor a helper found (0 = none)
  u64 window;   // the owner's current Brent window: helpers join the longest-running split
  u64 pad;
};
constexpr u32 kBoardSlots = 64;           // one per lane of a wave
constexpr u32 kBoardHelpWaves = 4;        // helper waves per workgroup: one per SIMD
// Workgroups whose helpers stay while another batch is on the device (the others leave so
// that its kernels find the CUs free). Every workgroup helps while the batch runs alone.
// Measured on 512 windows near 2^64 / 1e18 with eight batches in flight: helpers in every
// workgroup 0.150 ms per batch, in 96 0.138, in 64 0.125-0.128, in 32 or 48 0.124.
constexpr u32 kBoardHelpGroups = 32;
constexpr u32 kPostWindowMont = 32;       // Brent window at which a Montgomery split is posted
constexpr u32 kPostWindowF64 = 128;       // ... and an FP64 one (its steps are cheaper)

__device__ __forceinline__ u64 load_dev(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u32 load_dev32(const u32 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Lane 0's value for the whole wave. __builtin_amdgcn_readfirstlane returns an int: each
// 32-bit half goes back through u32 before widening, or a half with its top bit set would
// be sign-extended into the other.
__device__ __forceinline__ u32 uniform32(u32 v) { return (u32)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ u64 uniform64(u64 v) {
  return (u64)uniform32((u32)v) | ((u64)uniform32((u32)(v >> 32)) << 32);
}

// How one rho call takes part (helper): 0 = solo or owner of the workgroup's shared split
// (rs), 1 = helper of the workgroup's split `gen`, 2 = helper of board slot `slot`. Walk
// constants start above cbase, so no two waves repeat a walk. An owner posts its split
// on the board (slot >= 0 once posted) after a few windows.
struct RhoRole {
  RhoShare *rs;
  u32 helper, gen, cbase;
  RhoBoard *board;          // null: no grid-wide posts
  const u32 *done;          // board helpers stop once *done >= ndone (every survivor finished)
  u32 ndone;
  int slot;
  const u32 *active;        // non-null: the helper also stops once *active > 1
};

// Whether a board helper leaves: every survivor of its batch is finished, or (a helper
// given `active`, the count of batches on the device) another batch is on the device.
__device__ __forceinline__ bool helper_leaves(const u32 *done, u32 ndone, const u32 *active) {
  return __builtin_amdgcn_readfirstlane(load_dev32(done)) >= ndone ||
         (active && __builtin_amdgcn_readfirstlane(load_dev32(active)) > 1u);
}

// Claims a free board slot for n, or -1 if none is free: each lane reads one slot, and
// lane 0 tries up to four of the free ones, starting from a slot that depends on n so that
// owners posting at the same time try different slots.
__device__ __forceinline__ int board_post(RhoBoard *board, u64 n, u32 lane) {
  u64 free = __ballot(load_dev(&board[lane].n) == 0);
  int slot = -1;
  if (lane == 0) {
    const u32 rot = (u32)(n % kBoardSlots);
    free = (free >> rot) | (rot ? free << (64 - rot) : 0);    // slot rot + bit
    for (int tries = 0; tries < 4 && free && slot < 0; ++tries) {
      const u32 b = (u32)__builtin_ctzll(free);
      free &= free - 1;
      const u32 t = (rot + b) % kBoardSlots;
      if (atomicCAS((unsigned long long *)&board[t].n, 0ull, (unsigned long long)n) == 0ull) slot = (int)t;
    }
  }
  return __builtin_amdgcn_readfirstlane(slot);
}

// The owner's periodic exchange. Returns a factor of n (posted by a helper of the board
// or of the workgroup), kRhoStop for a helper whose split is over, else 0. A post that
// does not divide n (stale, from an earlier value of the slot) is cleared.
__device__ __forceinline__ u64 rho_coop(const RhoRole &ro, u64 n) {
  if (ro.helper == 2)
    return (uniform64(load_dev(&ro.board[ro.slot].n)) != n || helper_leaves(ro.done, ro.ndone, ro.active))
               ? kRhoStop
               : 0;
  if (ro.board && ro.slot >= 0) {
    const u64 g = uniform64(load_dev(&ro.board[ro.slot].factor));
    if (g) {
      if (g > 1 && g < n && n % g == 0) return g;
      atomicCAS((unsigned long long *)&ro.board[ro.slot].factor, (unsigned long long)g, 0ull);
    }
  }
  if (!ro.rs) return 0;
  volatile RhoShare *v = ro.rs;
  if (ro.helper)
    return (__builtin_amdgcn_readfirstlane(v->busy) == 0 || __builtin_amdgcn_readfirstlane(v->gen) != ro.gen)
               ? kRhoStop
               : 0;
  const u64 g = uniform64(v->factor);
  if (g == 0) return 0;
  if (g > 1 && g < n && n % g == 0) return g;
  atomicCAS((unsigned long long *)&ro.rs->factor, (unsigned long long)g, 0ull);
  return 0;
}

// What a walk that found factor g returns: the owner or a solo wave keeps it, a helper
// posts it (first post wins) and stops.
__device__ __forceinline__ u64 rho_found(const RhoRole &ro, u64 n, u64 g, u32 lane) {
  if (!ro.helper) return g;
  if (ro.helper == 2) {   // posted only while the slot still holds n
    if (lane == 0 && load_dev(&ro.board[ro.slot].n) == n)
      atomicCAS((unsigned long long *)&ro.board[ro.slot].factor, 0ull, (unsigned long long)g);
    return kRhoStop;
  }
  if (lane == 0 && ro.rs->gen == ro.gen)
    atomicCAS((unsigned long long *)&ro.rs->factor, 0ull, (unsigned long long)g);
  return kRhoStop;
}

// An owner posts its split once its walks reach Brent window `at` without a factor (and
// tries again at each later window while the board is full), then keeps the slot's window
// current, so helpers can pick the split that has run longest.
__device__ __forceinline__ void maybe_post(RhoRole &ro, u64 n, u32 r, u32 at, u32 lane) {
  if (ro.helper || !ro.board) return;
  if (ro.slot < 0 && r >= at) ro.slot = board_post(ro.board, n, lane);   // again each window if full
  if (ro.slot >= 0 && lane == 0)
    __hip_atomic_store(&ro.board[ro.slot].window, (u64)r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Brent's rho with an independent walk per lane (x -> x^2 + c, c = cbase + lane + 1,
// Montgomery form). Returns a nontrivial factor of the odd composite non-square n,
// kRhoFail when every walk fails inside the budget (the caller then hands the prime to
// the host), or kRhoStop once *stop drops below k (a smaller rung is settled; the result
// is moot) or, for a helper, once its split is over.
__device__ u64 wave_rho(u64 n, u32 lane, const volatile u32 *stop, u32 k, RhoRole &ro) {
  const u64 ninv = neg_inv(n), one = (0 - n) % n;
  const u64 c = small_to_mont(ro.cbase + lane + 1, one, n);
  auto f = [&](u64 v) { return add_mod(mont_mul(v, v, n, ninv), c, n); };
  u64 y = small_to_mont(2, one, n), x = y, ys = y, acc = one;
  bool alive = true;
  for (u32 r = 1; r <= (1u << 16); r <<= 1) {
    maybe_post(ro, n, r, kPostWindowMont, lane);
    x = y;
    // Advance r steps, looking for a stop or a posted factor every 64.
    for (u32 i = 0; i < r; ++i) {
      y = f(y);
      if ((i & 63) == 63) {
        if (stop && *stop < k) return kRhoStop;
        const u64 h = rho_coop(ro, n);
        if (h) return h;
      }
    }
    for (u32 kk = 0; kk < r; kk += 64) {
      ys = y;
      const u32 lim = min(64u, r - kk);
      for (u32 i = 0; i < lim; ++i) {
        y = f(y);
        acc = mont_mul(acc, x > y ? x - y : y - x, n, ninv);
      }
      u64 g = alive ? gcd_u64(acc, n) : 1;
      if (g == n) {                       // the batch overshot: replay it step by step
        g = 1;
        for (u32 i = 0; i < lim && g == 1; ++i) {
          ys = f(ys);
          const u64 dd = x > ys ? x - ys : ys - x;
          g = gcd_u64(dd ? dd : n, n);
        }
        if (g == n) g = 1;
        if (g == 1) alive = false;
      }
      const u64 found = __ballot(g != 1);
      if (found) return rho_found(ro, n, __shfl(g, __ffsll((long long)found) - 1, 64), lane);
      if (__ballot(alive) == 0) return kRhoFail;
      if (stop && *stop < k) return kRhoStop;
      const u64 h = rho_coop(ro, n);
      if (h) return h;
    }
  }
  return kRhoFail;
}


constexpr int kRhoWalks = 4;   // independent walks per lane (4 measured best of 2, 4, 8 on C2)

// wave_rho for n < 2^50 (same contract): 64 x kRhoWalks walks x -> x^2 + c with
// c = cbase + lane + 1 + 64 w, in signed FP64 residues (|x| < n); only the gcd converts
// to integers, and gcd(a, n) does not depend on which representative of a mod n it sees.
__device__ u64 wave_rho_f64(u64 n, u32 lane, const volatile u32 *stop, u32 k, RhoRole &ro) {
  const double nd = (double)n, ninv = 1.0 / nd;
  double c[kRhoWalks], y[kRhoWalks], x[kRhoWalks], ys[kRhoWalks], acc[kRhoWalks];
  bool alive[kRhoWalks];
#pragma unroll
  for (int w = 0; w < kRhoWalks; ++w) {
    c[w] = (double)(ro.cbase + lane + 1 + 64 * w);
    y[w] = x[w] = ys[w] = 2;
    acc[w] = 1;
    alive[w] = true;
  }
  auto f = [&](double v, double cc) { return mulmod_sgn(v, v, nd, ninv) + cc; };
  for (u32 r = 1; r <= (1u << 16); r <<= 1) {
    maybe_post(ro, n, r, kPostWindowF64, lane);
#pragma unroll
    for (int w = 0; w < kRhoWalks; ++w) x[w] = y[w];
    // Advance r steps, looking for a stop or a posted factor every 16.
    for (u32 i = 0; i < r; ++i) {
#pragma unroll
      for (int w = 0; w < kRhoWalks; ++w) y[w] = f(y[w], c[w]);
      if ((i & 15) == 15) {
        if (stop && *stop < k) return kRhoStop;
        const u64 h = rho_coop(ro, n);
        if (h) return h;
      }
    }
    for (u32 kk = 0; kk < r; kk += 64) {
      const u32 lim = min(64u, r - kk);
#pragma unroll
      for (int w = 0; w < kRhoWalks; ++w) ys[w] = y[w];
      for (u32 i = 0; i < lim; ++i) {
#pragma unroll
        for (int w = 0; w < kRhoWalks; ++w) {
          y[w] = f(y[w], c[w]);
          acc[w] = mulmod_sgn(acc[w], alive[w] ? x[w] - y[w] : 1.0, nd, ninv);   // |acc| < 2n
        }
        if ((i & 15) == 15) {              // a stop or a factor another wave posted
          if (stop && *stop < k) return kRhoStop;
          const u64 h = rho_coop(ro, n);
          if (h) return h;
        }
      }
      // One gcd per batch, of the product of every walk's accumulator (the chain of
      // products stays below 2n in magnitude too).
      double all = acc[0];
#pragma unroll
      for (int w = 1; w < kRhoWalks; ++w) all = mulmod_sgn(all, acc[w], nd, ninv);
      u64 g = gcd_u64((u64)fabs(all), n);
      if (g == n) {
        // Some walk overshot (its product is 0 mod n): per-walk gcds, and a step-by-step
        // replay of the batch for each walk that overshot; a walk whose replay finds no
        // proper factor is dead from then on (its accumulator stays 1).
        g = 1;
#pragma unroll
        for (int w = 0; w < kRhoWalks; ++w) {
          u64 gw = alive[w] ? gcd_u64((u64)fabs(acc[w]), n) : 1;
          if (gw == n) {
            gw = 1;
            for (u32 i = 0; i < lim && gw == 1; ++i) {
              ys[w] = f(ys[w], c[w]);
              const u64 dd = (u64)fabs(x[w] - ys[w]);
              gw = gcd_u64(dd ? dd : n, n);
            }
            if (gw == n) gw = 1;
            if (gw == 1) {
              alive[w] = false;
              acc[w] = 1;
            }
          }
          if (g == 1) g = gw;
        }
      }
      bool any = false;
#pragma unroll
      for (int w = 0; w < kRhoWalks; ++w) any |= alive[w];
      const u64 found = __ballot(g != 1);
      if (found) return rho_found(ro, n, __shfl(g, __ffsll((long long)found) - 1, 64), lane);
      if (__ballot(any) == 0) return kRhoFail;
      if (stop && *stop < k) return kRhoStop;
      const u64 h = rho_coop(ro, n);
      if (h) return h;
    }
  }
  return kRhoFail;
}

// Per-wave factor list in LDS.
struct WaveFac {
  u64 p[16];
  u32 e[16];
  u32 k;
  u32 bad;
};

__device__ __forceinline__ void fac_add(WaveFac &F, u64 q, u32 e, u32 lane) {
  if (lane == 0) {
    u32 i = 0;
    while (i < F.k && F.p[i] != q) ++i;
    if (i < F.k) F.e[i] += e;
    else if (F.k < 16) { F.p[F.k] = q; F.e[F.k] = e; ++F.k; }
    else F.bad = 1;
  }
  wave_sync();
}

// floor(sqrt(m)) for every u64 m: the estimate is clamped to 2^32 - 1 first, so no
// square below overflows (stage E's cofactors reach 2^64).
__device__ __forceinline__ u64 isqrt_dev(u64 m) {
  u64 s = min((u64)sqrt((double)m), 0xFFFFFFFFull);
  while (s && s * s > m) --s;
  while (s < 0xFFFFFFFFull && (s + 1) * (s + 1) <= m) ++s;
  return s;
}

constexpr int kFacOk = 0, kFacHost = 1, kFacStop = 2;

// The primes below 2^16 of a (< 2^64) into F, given the indices (into T) of the odd ones
// that divide it (2 by a shift); their exponents on one lane each, and the cofactor, all
// of whose prime factors exceed 65521, by one exact division. kFacHost when F would
// overflow.
__device__ int wave_factor_small(u64 a, const u16 *small, u32 nsmall, const TrialTab &T, WaveFac &F, u32 lane,
                                 u64 &cof) {
  if (lane == 0) { F.k = 0; F.bad = 0; }
  wave_sync();
  u64 n = a;
  const u32 e2 = (u32)__builtin_ctzll(n);
  n >>= e2;
  if (e2) fac_add(F, 2, e2, lane);
  if (F.k + nsmall > 16) return kFacHost;
  u64 pw = 1;
  if (lane < nsmall) {
    const u32 i = small[lane];
    const u32 q = T.qm[i].x & 0xFFFFu;
    const u64 qi = T.inv[i], ql = T.lim[i];
    u64 m = n;
    u32 e = 0;
    do { m *= qi; pw *= q; ++e; } while (m * qi <= ql);
    F.p[F.k + lane] = q;
    F.e[F.k + lane] = e;
  }
  wave_sync();
  if (lane == 0) F.k += nsmall;
  for (int s = 32; s; s >>= 1) pw *= __shfl_xor(pw, s, 64);
  cof = n * inv64(pw);
  wave_sync();
  return kFacOk;
}

// Whether cofactor n (every prime factor above 65521) needs a rho split: it is neither 1,
// nor prime (a part below 65537^2 is), nor the square of a prime.
__device__ __forceinline__ bool needs_split(u64 n, u32 lane) {
  if (n <= 0xFFFFFFFFull || wave_is_prime(n, lane)) return false;
  const u64 s = isqrt_dev(n);
  return s * s != n;
}

// The cofactor's primes into F (every prime factor of n exceeds 65521), splitting
// composite parts with rho. kFacHost if the device declines, kFacStop if a smaller rung
// settled meanwhile.
__device__ int wave_split(u64 n, WaveFac &F, u32 lane, const volatile u32 *stop, u32 k, RhoShare *rs,
                          RhoBoard *board) {
  u64 stack[4];
  u32 top = 0;
  if (n > 1) stack[top++] = n;
  while (top) {
    const u64 m = stack[--top];
    if (m <= 0xFFFFFFFFull || wave_is_prime(m, lane)) { fac_add(F, m, 1, lane); continue; }
    const u64 s = isqrt_dev(m);
    if (s * s == m) { fac_add(F, s, 2, lane); continue; }
    // Become the owner of the block's shared split if nobody holds it.
    u32 own = 0;
    if (lane == 0) own = atomicCAS(&rs->busy, 0u, 1u) == 0u ? 1u : 0u;
    own = __shfl(own, 0, 64);
    if (own && lane == 0) {
      rs->n = m;
      rs->factor = 0;
      __threadfence_block();
      atomicAdd(&rs->gen, 1u);
    }
    wave_sync();
    RhoRole ro{own ? rs : nullptr, 0u, 0u, 0u, board, nullptr, 0u, -1};
    const u64 d = m < kF64Mod ? wave_rho_f64(m, lane, stop, k, ro) : wave_rho(m, lane, stop, k, ro);
    if (own && lane == 0) atomicExch(&rs->busy, 0u);
    if (ro.slot >= 0 && lane == 0) {    // the split is over: free the board slot
      atomicExch((unsigned long long *)&board[ro.slot].factor, 0ull);
      atomicExch((unsigned long long *)&board[ro.slot].n, 0ull);
    }
    if (d == kRhoStop) return kFacStop;
    if (d == kRhoFail || top + 2 > 4) return kFacHost;
    stack[top++] = d;
    stack[top++] = m / d;
  }
  return F.bad == 0 ? kFacOk : kFacHost;
}

// Stage D's witness search can run before a cofactor split: every divisor of (pA)^2
// that the known primes (p, and A's primes below 2^16) do not build alone is a multiple
// of a prime of the unsplit cofactor, so at least 65537. A witness below that over the
// known primes is therefore the rung's least witness, and the split is not needed. (The
// reference factors A completely first, rung_scan3.cpp:578-603; the least witness is
// the same.)
constexpr u128 kWitnessBound = 65537;   // the least prime above 65521

__device__ __forceinline__ u32 small_mod(u64 x, u32 r, u32 mg, u32 c32) { return mod64(x, r, mg, c32); }

// pm^{-1} mod r for a unit pm (every prime power's residue on a rung that is not skipped
// is one): lane x tests x, x + 64, ...; the wave takes the lane that found it.
__device__ __forceinline__ u32 unit_inverse(u32 pm, u32 r, u32 mg, u32 lane) {
  u32 inv = 0;
  const u32 one = r == 1 ? 0u : 1u;
  for (u32 x = lane; x < r; x += 64) if (mod32(pm * x, r, mg) == one) inv = x + 1;
  const u64 bal = __ballot(inv != 0);
  return bal ? __shfl(inv, __ffsll((long long)bal) - 1, 64) - 1 : 0;
}

// Least witness of rung r for q = pA (A factored in F): u | (pA)^2, u == target (mod r),
// u <= q, or ~0 if there is none. best[t] is the least divisor so far (of the prime powers
// absorbed) that is == t (mod r) and <= q; absorbing pr^E makes it min over j <= E of
// best[t / pr^j] * pr^j. Lane l owns residues l, l + 64, l + 128, l + 192: four
// independent chains per lane, with the powers pr^j shared by the wave. A wave runs in
// lockstep, so all its reads of best are done before its writes and best is updated in
// place.
__device__ u128 least_witness(u64 p, const WaveFac &F, u32 r, u32 mg, u32 c32, u32 target, u128 q,
                              u128 *best, u32 lane) {
  const u128 INF = ~(u128)0;
  for (u32 t = lane; t < 256; t += 64) best[t] = t == 1 % r ? (u128)1 : INF;
  wave_sync();
  // Absorb p^2, then every prime power of A^2.
  for (u32 fi = 0; fi <= F.k; ++fi) {
    const u64 pr = fi == 0 ? p : F.p[fi - 1];
    const u32 E = fi == 0 ? 2u : 2u * F.e[fi - 1];
    const u32 pm = small_mod(pr, r, mg, c32);
    const u32 pinv = unit_inverse(pm, r, mg, lane);
    u32 s[4];
    u128 acc[4];
#pragma unroll
    for (u32 w = 0; w < 4; ++w) {
      const u32 t = lane + 64 * w;
      s[w] = t < r ? t : 0u;
      acc[w] = INF;
    }
    u128 pw = 1;
    for (u32 j = 0;; ++j) {
#pragma unroll
      for (u32 w = 0; w < 4; ++w) {
        if (64 * w >= r) break;               // uniform: chains beyond r do not exist
        const u128 b = best[s[w]];
        u128 v;
        if (b != INF && mul_le(b, pw, q, v) && v < acc[w]) acc[w] = v;
      }
      if (j == E) break;
      u128 npw;
      if (!mul_le(pw, (u128)pr, q, npw)) break;   // pr^(j+1) > q: no further term fits
      pw = npw;
#pragma unroll
      for (u32 w = 0; w < 4; ++w)
        if (64 * w < r) s[w] = mod32(s[w] * pinv, r, mg);   // < 255 * 255
    }
    wave_sync();
#pragma unroll
    for (u32 w = 0; w < 4; ++w)
      if (lane + 64 * w < r) best[lane + 64 * w] = acc[w];
    wave_sync();
  }
  const u128 u = best[target];
  wave_sync();
  return u;
}

// Helps the block's cofactor split in progress, if there is one this wave has not helped
// yet, until that split is over; false if there was none.
__device__ bool help_split(RhoShare *rs, u32 &seen, u32 wv, u32 lane) {
  volatile RhoShare *v = rs;
  const u32 gen = __builtin_amdgcn_readfirstlane(v->gen);
  if (!__builtin_amdgcn_readfirstlane(v->busy) || gen == seen) return false;
  seen = gen;
  const u64 n = uniform64(v->n);
  if (__builtin_amdgcn_readfirstlane(v->gen) == gen && n > 1) {
    RhoRole ro{rs, 1u, gen, 64u * kRhoWalks * (wv + 1), nullptr, nullptr, 0u, -1};   // disjoint walks
    (void)(n < kF64Mod ? wave_rho_f64(n, lane, nullptr, 0, ro) : wave_rho(n, lane, nullptr, 0, ro));
  }
  return true;
}

// A wave of a workgroup that found the survivor queue empty: walks on values other
// workgroups posted until every survivor of the batch is finished. `me` numbers the
// helper waves of the grid, so helpers start their slot scan at different slots and no
// two of them (or any owner) share walk constants; each visit takes fresh constants.
// Helpers leave once *done reaches ndone, which every owner's bounded work guarantees, or,
// given `active`, once another batch is on the device.
__device__ void board_help(RhoBoard *board, const u32 *done, u32 ndone, u32 lane, u32 me, const u32 *active) {
  for (u32 visit = 0;; ++visit) {
    if (helper_leaves(done, ndone, active)) return;
    // Each lane reads one slot; the wave joins the open split with the largest Brent
    // window (the one that has run longest), ties going to the first slot at or after `me`.
    const u64 mine = load_dev(&board[lane].n);
    const u64 win = mine ? load_dev(&board[lane].window) : 0;
    const u32 lg = win ? 32u - (u32)__builtin_clz((u32)min(win, (u64)0x80000000u)) : 0u;
    u32 key = mine ? (lg << 6) | (63u - ((lane - me) & 63u)) : 0u;
    u32 best = key;
    for (int o = 32; o > 0; o >>= 1) best = max(best, (u32)__shfl_xor((int)best, o, 64));
    if (!best) {
      __builtin_amdgcn_s_sleep(8);
      continue;
    }
    const u32 slot = (u32)__builtin_ctzll(__ballot(key == best));
    const u64 n = uniform64(__shfl(mine, (int)slot, 64));
    // Owner walks use constants below 64 kRhoWalks (kRungWaves + 1) < 4096; a visit's
    // constants stay below 2^27 < n / 8 (posted values exceed 2^32), which keeps the FP64
    // walk inside the bounds its product is exact for.
    RhoRole ro{nullptr, 2u, 0u, 4096u + ((visit & 255u) * 2048u + (me & 2047u)) * (64u * kRhoWalks), board, done,
               ndone, (int)slot, active};
    (void)(n < kF64Mod ? wave_rho_f64(n, lane, nullptr, 0, ro) : wave_rho(n, lane, nullptr, 0, ro));
  }
}

// ------------------------------------------------------------------ stage E on the device
//
// Stage E (rung_scan3.cpp:607-647) for a prime no rung certifies: the first pair (a, d),
// a then d ascending in 1..8, for which n = p + 4 a^2 d has a divisor m == -1 (mod 4ad)
// with c = (m + 1) / 4ad, k = n / m and c k > a, and for that pair the least such m.
// Waves take pairs in order from an LDS counter, as they take rungs; each factors its n
// completely (trial division by the odd primes below 2^16 across the lanes, then the
// cofactor as in stage D) and searches the divisors across the lanes, keeping the
// least m that qualifies, which is the first one in ascending order.

// Indices (into T) of the odd primes below 2^16 dividing n, into small[0 .. min(count, 16));
// returns the count. Exact divisibility: q | n iff n q^-1 mod 2^64 <= floor((2^64-1)/q).
__device__ u32 wave_trial(u64 n, const TrialTab &T, u16 *small, u32 lane) {
  u32 ns = 0;
  for (u32 i0 = 0; i0 < T.n; i0 += 64) {
    const u32 i = i0 + lane;
    const bool hit = i < T.n && n * T.inv[i] <= T.lim[i];
    const u64 b = __ballot(hit);
    if (hit) {
      const u32 slot = ns + (u32)__popcll(b & ((1ull << lane) - 1ull));
      if (slot < kSmallFac) small[slot] = (u16)i;
    }
    ns += (u32)__popcll(b);
  }
  wave_sync();
  return ns;
}

// The least divisor m of n (factored in F) with m == q4 - 1 (mod q4) and (m+1)/q4 * n/m > a,
// divisors taken across the lanes in mixed radix; false if none qualifies.
__device__ bool wave_direct_divisor(u64 n, u64 q4, u64 a, const WaveFac &F, u32 lane, u64 &c, u64 &k) {
  const u32 nf = F.k;
  u64 D = 1;
  for (u32 i = 0; i < nf; ++i) D *= (u64)(F.e[i] + 1);
  u64 best = ~0ull;
  for (u64 t = lane; t < D; t += 64) {
    u64 m = 1, x = t;
    for (u32 i = 0; i < nf; ++i) {
      const u64 r = (u64)(F.e[i] + 1);
      for (u64 e = x % r; e; --e) m *= F.p[i];
      x /= r;
    }
    if (m < best && m % q4 == q4 - 1) {
      const u64 cc = (m + 1) / q4, kk = n / m;
      if (cc && kk && (u128)cc * kk > a) best = m;
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    const u64 other = (u64)__shfl_xor((unsigned long long)best, o, 64);
    best = min(best, other);
  }
  if (best == ~0ull) return false;
  c = (best + 1) / q4;
  k = n / best;
  return true;
}

// Batches of at most kPendScan survivors are listed: k_rung's queue hands out the values
// k_mr_base2 left pending (at most kPendList of them) instead of every survivor.
constexpr u32 kPendScan = 65536, kPendList = 4096;

__global__ __launch_bounds__(64 * kRungWaves) void k_rung(const u64 *__restrict__ vals,
                                                          const u32 *__restrict__ count, u32 cap,
                                                          u8 *__restrict__ flags, const u64 *__restrict__ one_in,
                                                          TrialTab T, RungOut *__restrict__ out, u32 *queue,
                                                          RhoBoard *board, u32 *done, u64 *stamp = nullptr,
                                                          u32 rung_cap = kRungSlots, const u32 *active = nullptr) {
  __shared__ u128 best[kRungWaves][256];
  __shared__ WaveFac F[kRungWaves];
  __shared__ u16 fac[kRungSlots][kSmallFac];
  __shared__ u32 nfac[kRungSlots];
  __shared__ u32 next_k, found_k, host_k, stop_k;
  __shared__ RungOut wres[kRungWaves];
  __shared__ RhoShare rs;
  __shared__ u32 mr_composite;
  __shared__ u32 cur_idx[2];
  __shared__ u32 help_batch;
  __shared__ u32 pend[kPendList];
  __shared__ u32 npend, wcnt[kRungWaves];
  const u32 tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const u32 n_ = min(*count, cap);
  const volatile u32 *stop = &stop_k;
  if (blockIdx.x == 0) stamp_start(stamp);   // the first workgroup dispatched, as in k_sieve
  // Most survivors are composites k_mr_base2 has already marked. Claiming each of them from
  // the one queue counter (and counting each done) puts hundreds of same-address atomics
  // ahead of the first prime, and they serialize at the memory side. So in a batch of at
  // most kPendScan survivors every workgroup lists the pending ones from one pass over the
  // flags, in index order, so that every workgroup's list is the same and a list position
  // names the same survivor everywhere; the queue hands out list positions. A larger batch,
  // or one with more than kPendList pending values, takes survivors from the queue as they
  // are. Either way `done` counts the items the queue hands out, nq of them.
  bool listed = n_ <= kPendScan;
  if (listed) {
    if (tid == 0) npend = 0;
    __syncthreads();
    for (u32 b0 = 0; b0 < n_; b0 += blockDim.x) {
      const u32 i = b0 + tid;
      const bool pending = i < n_ && flags[i] != kComposite;
      const u64 m = __ballot(pending);
      if (lane == 0) wcnt[wv] = (u32)__popcll(m);
      __syncthreads();                     // every wave's count, and npend, before any update
      u32 base = npend, total = 0;
      for (u32 w = 0; w < kRungWaves; ++w) {
        if (w < wv) base += wcnt[w];
        total += wcnt[w];
      }
      if (pending) {
        const u32 at = base + (u32)__popcll(m & ((1ull << lane) - 1ull));
        if (at < kPendList) pend[at] = i;
      }
      __syncthreads();                     // every read of npend and wcnt before they change
      if (tid == 0) npend += total;
    }
    __syncthreads();
    listed = npend <= kPendList;   // uniform: read after the barrier
  }
  const u32 nq = listed ? npend : n_;
  // A persistent grid takes survivors from a queue, so every prime starts as soon as a
  // workgroup is free (the index slot alternates, so the next claim never overwrites an
  // index a wave has yet to read).
  for (u32 it = 0;; ++it) {
    if (tid == 0) cur_idx[it & 1] = atomicAdd(queue, 1u);
    __syncthreads();
    const u32 q = cur_idx[it & 1];
    if (q >= nq) {
      // The queue is empty: help the splits other workgroups post, one wave per SIMD.
      // Only a batch with a prime p >= 2^52 can need a Montgomery split (A = (p + r) / 4
      // is then at least 2^50); below that every split runs in FP64 and is fast alone,
      // and helpers that stay resident would only crowd the kernels of the next batches.
      if (!board) break;
      if (tid == 0) help_batch = 0;
      __syncthreads();
      bool big = false;
      for (u32 i = tid; i < n_; i += blockDim.x) big |= vals[i] >= (1ull << 52) && flags[i] != kComposite;
      if (big) help_batch = 1;             // any writer writes the same value
      __syncthreads();
      if (help_batch && wv < kBoardHelpWaves)
        board_help(board, done, nq, lane, blockIdx.x * kBoardHelpWaves + wv,
                   blockIdx.x >= kBoardHelpGroups ? active : nullptr);
      break;
    }
    const u32 idx = listed ? pend[q] : q;
    const u8 f0 = flags[idx];
    if (f0 == kComposite) {
      if (tid == 0 && done) atomicAdd(done, 1u);
      continue;
    }
    const u64 p = vals[idx];
    u32 r0 = (u32)((4 - (p & 3)) & 3);
    if (r0 == 0) r0 = 4;
    const u32 nj = min((255 - r0) / 4 + 1, rung_cap);   // rung_cap 0: stage E alone
    const u64 A0 = (p + r0) / 4;
    if (tid < kRungSlots) nfac[tid] = 0;
    if (tid < kRungWaves) wres[tid] = RungOut{0, 0, 0, 0, kRungNone};
    if (tid == 0) {
      next_k = 0;
      found_k = host_k = stop_k = kRungSlots;
      rs = RhoShare{0, 0, 0, 0, 0, 0};
      mr_composite = 0;
    }
    __syncthreads();
    // A value the base-2 pass left pending gets its eleven odd prime bases here, one per
    // lane of wave 0, while the other waves sieve.
    const bool pending = f0 == kPending;
    if (pending && wv == 0) {
      bool ok = true;
      if (lane < 11) {
        u64 d = p - 1;
        const u32 s2 = (u32)__builtin_ctzll(d);
        d >>= s2;
        const u32 a = (u32)c_small_p[lane + 1];
        ok = p < kF64Mod ? strong_probable_sgn(p, a, d, s2)
                         : strong_probable(p, neg_inv(p), one_in[idx], a, d, s2);
      }
      if (__ballot(!ok) && lane == 0) mr_composite = 1;
    }
    // Sieve the 64 consecutive A_j by every odd prime below 2^16. Each sieving thread
    // issues all its table loads before using any (the loop is otherwise one load latency
    // per prime).
    const u32 w0 = pending ? 64u : 0u;
    if (tid >= w0) {
      const u32 me = tid - w0, nw = blockDim.x - w0;
      for (u32 i0 = me; i0 < T.n; i0 += kSievePer * nw) {
        uint2 c[kSievePer];
#pragma unroll
        for (u32 u = 0; u < kSievePer; ++u) {
          const u32 i = i0 + u * nw;
          c[u] = T.qm[min(i, T.n - 1u)];   // unconditional, so every load is in flight at once
        }
        // First A_j each prime divides (64 = none in this slot) ...
        u32 j0[kSievePer];
#pragma unroll
        for (u32 u = 0; u < kSievePer; ++u) {
          const u32 q = c[u].x & 0xFFFFu;
          const u32 x = mod64(A0, q, c[u].y, c[u].x >> 16);
          j0[u] = i0 + u * nw < T.n ? (x ? q - x : 0u) : 64u;
        }
        // ... then every slot claim issued before any is used (a prime above 64 hits at
        // most one A_j, yet some lane of a wave almost always hits one), and the further
        // hits of the primes below 64 last.
        u32 sl[kSievePer];
#pragma unroll
        for (u32 u = 0; u < kSievePer; ++u)
          if (j0[u] < nj) sl[u] = atomicAdd(&nfac[j0[u]], 1u);
#pragma unroll
        for (u32 u = 0; u < kSievePer; ++u)
          if (j0[u] < nj && sl[u] < kSmallFac) fac[j0[u]][sl[u]] = (u16)(i0 + u * nw);
#pragma unroll
        for (u32 u = 0; u < kSievePer; ++u) {
          const u32 q = c[u].x & 0xFFFFu;
          if (j0[u] < nj && q < nj)
            for (u32 j = j0[u] + q; j < nj; j += q) {
              const u32 s = atomicAdd(&nfac[j], 1u);
              if (s < kSmallFac) fac[j][s] = (u16)(i0 + u * nw);
            }
        }
      }
    }
    __syncthreads();
    if (mr_composite) {                    // uniform: read after the barrier
      if (tid == 0) {
        flags[idx] = kComposite;
        if (done) atomicAdd(done, 1u);
      }
      __syncthreads();
      continue;
    }
    // Rungs in order (phase 0). Between rungs a wave first helps a split in progress: that
    // split belongs to a smaller rung than any it could take, so it is on the critical path.
    // A prime no rung certifies (and none declined) goes on to stage E in the same
    // workgroup (phase 1): the pairs (a, d) in order take the rungs' place, through the
    // same factorization code.
    u32 seen = 0;
    bool direct = false;
    for (u32 phase = 0;; ++phase) {
      for (;;) {
        help_split(&rs, seen, wv, lane);
        u32 k = 0;
        if (lane == 0) k = atomicAdd(&next_k, 1u);
        k = __shfl(k, 0, 64);
        if (k >= (phase ? kDirectPairs : nj) || k > *stop) break;
        u64 A;
        const u16 *small;
        u32 ns, r = 0, mg = 0, c32 = 0, qm = 0;
        if (!phase) {
          r = r0 + 4 * k;
          A = A0 + k;
          mg = r == 1 ? 0xFFFFFFFFu : (u32)(0x100000000ull / r);
          c32 = (u32)(0x100000000ull % r);
          qm = (small_mod(p, r, mg, c32) * small_mod(A, r, mg, c32)) % r;
          u32 ga = qm, gb = r;
          while (gb) { const u32 t = ga % gb; ga = gb; gb = t; }
          if (r > 1 && ga != 1) continue;
          ns = nfac[k];
          small = fac[k];
        } else {
          const u64 a = k / 8 + 1, d = k % 8 + 1;
          const u128 nn = (u128)p + (u128)(4 * a * a * d);
          if (nn > kDirectNMaxDev) continue;
          A = (u64)nn;
          ns = wave_trial(A, T, fac[wv], lane);
          small = fac[wv];
        }
        u64 cof = 1;
        int fr = ns > kSmallFac ? kFacHost : wave_factor_small(A, small, ns, T, F[wv], lane, cof);
        // Stage D first tries the witness over the known primes when the cofactor would
        // need a split; the split runs only if that witness is not provably least.
        bool known_only = !phase && fr == kFacOk && needs_split(cof, lane);
        bool hit = false;
        RungOut mine;
        for (;;) {
          if (fr == kFacOk && !known_only && cof > 1) {
            fr = wave_split(cof, F[wv], lane, stop, k, &rs, board);
            cof = 1;
          }
          if (fr != kFacOk) break;
          if (phase) {
            const u64 a = k / 8 + 1, d = k % 8 + 1;
            u64 c = 0, kq = 0;
            hit = wave_direct_divisor(A, 4 * a * d, a, F[wv], lane, c, kq);
            mine = RungOut{c, kq, (u64)k, (u32)(a | (d << 8)), kRungDirect};
            break;
          }
          const u128 u = least_witness(p, F[wv], r, mg, c32, qm ? r - qm : 0u, (u128)p * A, best[wv], lane);
          if (known_only && u >= kWitnessBound) {
            known_only = false;
            continue;
          }
          hit = u != ~(u128)0;
          mine = RungOut{A, (u64)u, (u64)(u >> 64), r, kRungFound};
          break;
        }
        if (fr == kFacStop) break;
        if (fr == kFacHost) {
          if (lane == 0) { atomicMin(&host_k, k); atomicMin(&stop_k, k); }
          break;
        }
        if (hit) {
          if (lane == 0) {
            wres[wv] = mine;
            atomicMin(&found_k, k);
            atomicMin(&stop_k, k);
          }
          break;
        }
      }
      // A wave done with its tasks helps whichever wave still splits a cofactor; all leave
      // once every wave is done (each task loop ends: its rho is bounded).
      if (lane == 0) atomicAdd(&rs.done, 1u);
      while (__builtin_amdgcn_readfirstlane(((volatile RhoShare *)&rs)->done) < kRungWaves)
        if (!help_split(&rs, seen, wv, lane)) __builtin_amdgcn_s_sleep(2);
      __syncthreads();
      if (phase || found_k != kRungSlots || host_k != kRungSlots) break;   // uniform after the barrier
      // ---- on to stage E: every wave has read found_k and host_k before they are reset
      direct = true;
      __syncthreads();
      if (tid == 0) {
        next_k = 0;
        found_k = host_k = stop_k = kDirectPairs;
        rs = RhoShare{0, 0, 0, 0, 0, 0};
      }
      if (tid < kRungWaves) wres[tid] = RungOut{0, 0, 0, 0, kRungNone};
      __syncthreads();
      seen = 0;
    }
    if (tid == 0) {
      RungOut res{0, 0, 0, 0, kRungNone};
      if (host_k < found_k) {
        res.status = kRungHost;
      } else if (direct && found_k < kDirectPairs) {
        for (u32 w = 0; w < kRungWaves; ++w)
          if (wres[w].status == kRungDirect && wres[w].u_hi == found_k)
            res = RungOut{wres[w].A, wres[w].u_lo, 0, wres[w].r, kRungDirect};
      } else if (!direct && found_k < kRungSlots) {
        for (u32 w = 0; w < kRungWaves; ++w)
          if (wres[w].status == kRungFound && wres[w].r == r0 + 4 * found_k) res = wres[w];
      }
      out[idx] = res;
      if (done) atomicAdd(done, 1u);
    }
    __syncthreads();
  }
  stamp_end(stamp, lane == 0);   // per wave: helper waves leave later than the others
}

// Copies one batch's results into pinned host memory with plain stores over the bus
// (zero-copy), so the host needs one stream synchronisation and no staging copies:
// the survivor count, the per-tile counters, and per survivor p, window, primality flag
// and (primes only) the stage-D result.
// The last of its workgroups to finish posts the batch's end time and number into pinned
// host memory (h_sync[2], h_sync[0]) once every mirror store has completed, so the host
// sees the batch complete, with every mirror written, by reading host memory alone.
// h_sync[3 ..) receives the kernels' device-clock stamps. The last workgroup also counts
// the batch out of `active`.
__global__ void k_pack(const BatchHdr *__restrict__ hdr, const uint4 *__restrict__ tile_counts,
                       const u32 *__restrict__ surv_count, u32 cap, const u64 *__restrict__ surv_p,
                       const u32 *__restrict__ surv_w, const u8 *__restrict__ flags,
                       const RungOut *__restrict__ rung, u32 with_rung, u32 *h_count, uint4 *h_tc,
                       u64 *h_p, u32 *h_w, u8 *h_f, RungOut *h_r, u32 *__restrict__ done_blocks,
                       u64 *h_sync, const u64 *__restrict__ stamps, u32 *__restrict__ active) {
  const u32 g = blockIdx.x * blockDim.x + threadIdx.x, stride = gridDim.x * blockDim.x;
  const u32 nt = hdr->ntiles, n = min(*surv_count, cap);
  if (g == 0) h_count[0] = *surv_count;
  if (g < kStamps) h_sync[3 + g] = stamps[g];
  for (u32 i = g; i < nt; i += stride) h_tc[i] = tile_counts[i];
  for (u32 i = g; i < n; i += stride) {
    const u8 f = flags[i];
    h_p[i] = surv_p[i];
    h_w[i] = surv_w[i];
    h_f[i] = f;
    if (with_rung && f) h_r[i] = rung[i];
  }
  // The mirrors are coherent host memory (not cached in any XCD's L2). Every wave waits
  // until its own stores have completed, then one thread per workgroup releases them at
  // system scope before counting the workgroup done; the last workgroup to count itself
  // done acquires the others' releases and posts the end time and, with a system-scope
  // release, the batch's number.
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (atomicAdd(done_blocks, 1u) == gridDim.x - 1) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
      atomicSub(active, 1u);   // the batch leaves the device (k_stage counted it in)
      h_sync[2] = __builtin_amdgcn_s_memrealtime();
      __hip_atomic_store(&h_sync[0], (u64)hdr->seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
  }
}

// LDS atomic-AND throughput probe: every thread clears bits of its own word for `iters`
// rounds (consecutive threads on consecutive banks, so no bank conflicts); the word is
// written out at the end so the work is observable.
__global__ __launch_bounds__(256) void k_lds_peak(u32 iters, u32 *sink) {
  __shared__ u32 w[1024];
  const u32 tid = threadIdx.x;
  for (u32 i = tid; i < 1024; i += blockDim.x) w[i] = 0xFFFFFFFFu;
  __syncthreads();
  u32 slot = tid;
  for (u32 i = 0; i < iters; ++i) {
    atomicAnd(&w[slot], ~(1u << (i & 31u)));
    slot = (slot + 256u) & 1023u;
  }
  __syncthreads();
  if (w[tid] == 0x12345678u) sink[blockIdx.x] = tid;   // never true; keeps the loop live
}

// LDS rate of the marking primitive itself, in its access pattern: byte stores of 0 in
// which every lane walks its own odd stride (19 .. 2001, like the filters of one bin) over
// a 16 KB map, so the lanes of a store land on banks independently of one another;
// 512-thread workgroups three to a CU, as the marking kernel runs.
__global__ __launch_bounds__(512) void k_lds_scatter(u32 iters, u32 *sink) {
  __shared__ u8 map[16384];
  const u32 tid = threadIdx.x;
  for (u32 i = tid; i < 16384 / 4; i += blockDim.x) ((u32 *)map)[i] = 0x01010101u;
  __syncthreads();
  const u32 s = 19u + 2u * ((tid * 7919u + blockIdx.x * 31u) % 992u);
  u32 a = (tid * 2654435761u + blockIdx.x * 97u) & 16383u;
  for (u32 i = 0; i < iters; ++i) {
    map[a] = 0;
    a = (a + s) & 16383u;
  }
  __syncthreads();
  if (map[tid] == 7) sink[blockIdx.x] = tid;   // never true; keeps the stores live
}

}  // namespace dev

// ====================================================================== host side

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) {                                                            \
      err_ = std::string(#expr) + ": " + hipGetErrorString(e_);                        \
      return -1;                                                                       \
    }                                                                                  \
  } while (0)

struct DevLayoutBufs {
  dev::LayoutDev d;
  void *lanes = nullptr, *mask = nullptr, *step = nullptr, *magic = nullptr, *c32 = nullptr,
       *fc = nullptr, *off = nullptr;
};

// One batch's buffers, captured graph and events. Several slots let the host plan and
// launch batch i+1 while the device still runs batch i, then finish batch i (mirrors,
// stage-D leftovers, records) while later batches run: the overlap of the reference driver
// (sieve/cuda/scan.cu:1-6, 197-207), over whole batches. Further slots keep batches
// queued while the host finishes one and submits the next, so the device does not idle
// behind the host. Eight slots, one stream each: a 512-window batch takes about 0.54 ms
// from its staging to its completion post when it shares the device, and the host needs
// about 0.12 ms to collect a batch and launch the next into its slot, so four slots left
// the device short of queued work (measured on that batch: 0.179 ms per batch four deep,
// 0.171 six deep, 0.152 eight deep; more hardware queues than HIP's default four did not
// change it).
constexpr size_t kSyncWords = 3 + dev::kStamps;
// Every pinned buffer the device writes or reads is coherent (fine-grained): the device
// does not keep its lines in L2. Each XCD has its own L2, and a release by the workgroup
// that posts a batch's completion writes back only its own XCD's L2, so mirror stores
// cached in another XCD's L2 could still be on their way when the host sees the post.
constexpr unsigned kPinnedFlags = hipHostMallocMapped | hipHostMallocCoherent;

struct Slot {
  // Each slot has its own stream, so the device may also overlap batches: one batch's
  // stage D (a few primes on a few CUs) with the next batch's staging and sieve.
  // The batches share only read-only tables, and no kernel waits on another workgroup.
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr, ev3 = nullptr;
  hipEvent_t ev4 = nullptr, ev5 = nullptr, ev6 = nullptr, ev7 = nullptr;   // k_mr_base2, stage D
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_ok = false;
  u32 graph_block = 0, graph_grid = 0, graph_off_words = 0;
  size_t graph_lds = 0;
  bool graph_tail = false;
  TileDesc *d_tiles = nullptr;
  dev::ItemDesc *d_items = nullptr;   // k_sieve's work items (at most one per tile)
  size_t tiles_cap = 0;
  size_t ni = 0;                      // work items of the batch
  dev::BatchHdr *d_hdr = nullptr;
  dev::BatchHdr *h_hdr = nullptr;
  uint4 *d_tile_counts = nullptr;   // per tile {positions, alive after covers, survivors, 0}
  u64 *d_surv_p = nullptr;
  u32 *d_surv_w = nullptr;
  uint8_t *d_flags = nullptr;
  u64 *d_one = nullptr;             // 2^64 mod n per survivor, from the base-2 pass
  dev::RungOut *d_rung = nullptr;   // stage D result per survivor
  size_t surv_cap = 0;
  u32 *d_scalars = nullptr;   // [0] survivor count, [1] tile queue head, [2] stage-D queue head,
                              // [3] survivors stage D has finished
  dev::RhoBoard *d_board = nullptr;   // stage D's posted cofactor splits
  // Pinned host memory: tile descriptors staged for the upload, and the mirrors k_pack
  // writes the results into (device-side addresses in the dh_* twins).
  u32 *h_count = nullptr;
  u64 *h_sync = nullptr;            // [0] number of the last batch completed, [1] its start, [2] its end,
                                    // [3, 3 + kStamps) its kernels' (start, end) stamps
  u64 *d_stamps = nullptr;
  u64 *dh_sync = nullptr;
  u32 seq = 0;                      // number of the batch in flight in this slot
  TileDesc *h_tiles = nullptr;
  TileDesc *dh_tiles = nullptr;
  dev::ItemDesc *h_items = nullptr;
  dev::ItemDesc *dh_items = nullptr;
  dev::BatchHdr *dh_hdr = nullptr;
  uint4 *h_tile_counts = nullptr;
  u64 *h_surv_p = nullptr;
  u32 *h_surv_w = nullptr;
  uint8_t *h_flags = nullptr;
  dev::RungOut *h_rung = nullptr;
  u32 *dh_count = nullptr;
  uint4 *dh_tile_counts = nullptr;
  u64 *dh_surv_p = nullptr;
  u32 *dh_surv_w = nullptr;
  uint8_t *dh_flags = nullptr;
  dev::RungOut *dh_rung = nullptr;
  // The batch the slot holds while it is in flight.
  size_t nt = 0;
  u32 nwin = 0;
  LaunchCfg cfg;
  size_t lds = 0;
  u32 lds_words = 0, off_words = 0, grid = 0;
};

struct DeviceCtx::Impl {
  hipStream_t stream = nullptr;     // slot 0's stream; also the one-off calls (rung, primality, probes)
  hipStream_t capture = nullptr;    // graph capture only: nothing ever runs on it
  hipEvent_t pev0 = nullptr, pev1 = nullptr;   // the LDS probes' timing
  std::vector<DevLayoutBufs> layouts;
  dev::LayoutDev *d_lays = nullptr;
  size_t lays_cap = 0;
  dev::TrialTab trial{};            // odd primes below 2^16 with their inverses
  u32 *d_active = nullptr;          // batches between their k_stage and the end of their k_pack
  int occupancy_blocks = 0;
  u32 occ_block = 0;
  size_t occ_lds = 0;
  u32 occ_off_words = 0;
  u32 off_max = 0;                  // longest offset row over the layouts (u16 entries)
  u32 mask_max = 0;                 // longest wheel-mask row over the layouts (u32 words)
  Slot slot[DeviceCtx::kMaxInFlight];
  int head = 0;                     // the oldest batch in flight
  int busy = 0;                     // batches in flight or reserved (0 .. kMaxInFlight)
  // Batches may be launched by one thread while another finishes older ones: slot_mu
  // guards head and busy; launch_mu serialises whatever enqueues work or reallocates
  // shared device buffers (launches, overflow replays, layout uploads).
  std::mutex slot_mu, launch_mu;
};

namespace {
// Device address of dev::scratch per device ordinal, looked up once per device with that
// device current (each GPU holds its own copy of the array at its own address).
std::mutex g_scratch_mu;
std::map<int, void *> g_scratch_addr;
}  // namespace

DeviceCtx::DeviceCtx() : impl_(new Impl) {}

DeviceCtx::~DeviceCtx() {
  if (!impl_) return;
  Impl &m = *impl_;
  if (device_ >= 0) {
    (void)hipSetDevice(device_);
    for (Slot &b : m.slot)
      if (b.stream) (void)hipStreamSynchronize(b.stream);
    for (DevLayoutBufs &b : m.layouts)
      for (void *p : {b.lanes, b.mask, b.step, b.magic, b.c32, b.fc, b.off}) if (p) (void)hipFree(p);
    for (void *p : {(void *)m.d_lays, (void *)m.trial.qm, (void *)m.trial.inv, (void *)m.trial.lim,
                    (void *)m.d_active})
      if (p) (void)hipFree(p);
    for (Slot &b : m.slot) {
      for (void *p : {(void *)b.d_tiles, (void *)b.d_items, (void *)b.d_hdr, (void *)b.d_tile_counts, (void *)b.d_surv_p,
                      (void *)b.d_surv_w, (void *)b.d_flags, (void *)b.d_one, (void *)b.d_rung,
                      (void *)b.d_scalars, (void *)b.d_board, (void *)b.d_stamps})
        if (p) (void)hipFree(p);
      for (void *p : {(void *)b.h_hdr, (void *)b.h_count, (void *)b.h_sync, (void *)b.h_tiles, (void *)b.h_items, (void *)b.h_tile_counts,
                      (void *)b.h_surv_p, (void *)b.h_surv_w, (void *)b.h_flags, (void *)b.h_rung})
        if (p) (void)hipHostFree(p);
      if (b.graph_exec) (void)hipGraphExecDestroy(b.graph_exec);
      if (b.graph) (void)hipGraphDestroy(b.graph);
      for (hipEvent_t e : {b.ev0, b.ev1, b.ev2, b.ev3, b.ev4, b.ev5, b.ev6, b.ev7}) if (e) (void)hipEventDestroy(e);
    }
    for (hipEvent_t e : {m.pev0, m.pev1}) if (e) (void)hipEventDestroy(e);
    if (m.capture) (void)hipStreamDestroy(m.capture);
    for (int i = 1; i < kMaxInFlight; ++i)
      if (m.slot[i].stream) (void)hipStreamDestroy(m.slot[i].stream);
    if (m.stream) (void)hipStreamDestroy(m.stream);   // slot 0's
  }
  delete impl_;
}

std::string DeviceCtx::device_name() const {
  hipDeviceProp_t p;
  if (device_ < 0 || hipGetDeviceProperties(&p, device_) != hipSuccess) return "?";
  return std::string(p.name) + " (" + p.gcnArchName + ")";
}

int DeviceCtx::open(int device) {
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) { err_ = "no such HIP device"; return -1; }
  HIP_TRY(hipSetDevice(device));
  device_ = device;
  {
    std::lock_guard<std::mutex> g(g_scratch_mu);
    auto it = g_scratch_addr.find(device);
    if (it == g_scratch_addr.end()) {
      void *ptr = nullptr;
      HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(dev::scratch)));
      if (!ptr) { err_ = "hipGetSymbolAddress: null address for the scratch array"; return -1; }
      it = g_scratch_addr.emplace(device, ptr).first;
    }
    scratch_ = it->second;
  }
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, device));
  cus_ = p.multiProcessorCount;
  // The dynamic LDS a k_sieve workgroup may take: the device's per-workgroup LDS less the
  // kernel's static arrays (a group's sub-tile tables and counters).
  hipFuncAttributes fa{};
  HIP_TRY(hipFuncGetAttributes(&fa, (const void *)dev::k_sieve<true, true>));
  lds_limit_ = p.sharedMemPerBlock > fa.sharedSizeBytes ? p.sharedMemPerBlock - fa.sharedSizeBytes : 0;
  int khz = 0;
  HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
  clock_ms_per_tick_ = khz > 0 ? 1.0 / khz : 1e-5;
  Impl &m = *impl_;
  HIP_TRY(hipStreamCreateWithFlags(&m.stream, hipStreamNonBlocking));
  m.slot[0].stream = m.stream;
  for (int i = 1; i < kMaxInFlight; ++i) HIP_TRY(hipStreamCreateWithFlags(&m.slot[i].stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&m.capture, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&m.pev0));
  HIP_TRY(hipEventCreate(&m.pev1));
  HIP_TRY(hipMalloc(&m.d_active, sizeof(u32)));
  HIP_TRY(hipMemset(m.d_active, 0, sizeof(u32)));
  for (Slot &b : m.slot) {
    HIP_TRY(hipEventCreate(&b.ev0));
    HIP_TRY(hipEventCreate(&b.ev1));
    HIP_TRY(hipEventCreate(&b.ev2));
    HIP_TRY(hipEventCreate(&b.ev3));
    HIP_TRY(hipEventCreate(&b.ev4));
    HIP_TRY(hipEventCreate(&b.ev5));
    HIP_TRY(hipEventCreate(&b.ev6));
    HIP_TRY(hipEventCreate(&b.ev7));
    HIP_TRY(hipMalloc(&b.d_hdr, sizeof(dev::BatchHdr)));
    HIP_TRY(hipHostMalloc(&b.h_hdr, sizeof(dev::BatchHdr), kPinnedFlags));
    HIP_TRY(hipHostGetDevicePointer((void **)&b.dh_hdr, b.h_hdr, 0));
    HIP_TRY(hipMalloc(&b.d_scalars, dev::kScalars * sizeof(u32)));
    HIP_TRY(hipMalloc(&b.d_stamps, dev::kStamps * sizeof(u64)));
    HIP_TRY(hipHostMalloc(&b.h_sync, kSyncWords * sizeof(u64), kPinnedFlags));
    HIP_TRY(hipHostGetDevicePointer((void **)&b.dh_sync, b.h_sync, 0));
    std::memset(b.h_sync, 0, kSyncWords * sizeof(u64));
    HIP_TRY(hipMalloc(&b.d_board, dev::kBoardSlots * sizeof(dev::RhoBoard)));
    HIP_TRY(hipHostMalloc(&b.h_count, 4 * sizeof(u32), kPinnedFlags));
    HIP_TRY(hipHostGetDevicePointer((void **)&b.dh_count, b.h_count, 0));
  }
  // Sieve table for stage D: odd primes q below 2^16 packed as {q | (2^32 mod q) << 16,
  // floor(2^32/q)} (residues of A by q), and q^{-1} mod 2^64, floor((2^64-1)/q) (exact
  // division by q).
  std::vector<uint8_t> comp(65537, 0);
  std::vector<uint2> tq;
  std::vector<u64> tinv, tlim;
  for (u32 i = 3; i <= 65536; i += 2) {
    if (comp[i]) continue;
    tq.push_back(uint2{i | (u32)(0x100000000ull % i) << 16, (u32)(0x100000000ull / i)});
    u64 x = i;
    for (int k = 0; k < 6; ++k) x *= 2 - (u64)i * x;
    tinv.push_back(x);
    tlim.push_back(~0ull / i);
    for (u32 j = i * i; j <= 65536; j += 2 * i) comp[j] = 1;
  }
  uint2 *dq = nullptr;
  u64 *dinv = nullptr, *dlim = nullptr;
  HIP_TRY(hipMalloc(&dq, tq.size() * sizeof(uint2)));
  HIP_TRY(hipMalloc(&dinv, tinv.size() * 8));
  HIP_TRY(hipMalloc(&dlim, tlim.size() * 8));
  HIP_TRY(hipMemcpy(dq, tq.data(), tq.size() * sizeof(uint2), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dinv, tinv.data(), tinv.size() * 8, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(dlim, tlim.data(), tlim.size() * 8, hipMemcpyHostToDevice));
  m.trial = dev::TrialTab{dq, dinv, dlim, (u32)tq.size()};
  return 0;
}

int DeviceCtx::add_layout(const Layout &L) {
  Impl &m = *impl_;
  if (device_ < 0) { err_ = "device not open"; return -1; }
  std::lock_guard<std::mutex> lk(m.launch_mu);
  HIP_TRY(hipSetDevice(device_));
  DevLayoutBufs b;
  auto up = [&](void **dst, const void *src, size_t bytes) -> int {
    HIP_TRY(hipMalloc(dst, bytes ? bytes : 8));
    if (bytes) HIP_TRY(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    layout_bytes_ += bytes;
    return 0;
  };
  if (up(&b.lanes, L.lanes.data(), L.lanes.size() * 8) || up(&b.mask, L.mask.data(), L.mask.size() * 4) ||
      up(&b.step, L.step.data(), L.step.size() * 4) || up(&b.magic, L.magic.data(), L.magic.size() * 4) ||
      up(&b.c32, L.c32.data(), L.c32.size() * 4))
    return -1;
  // Offset rows padded to a multiple of 8 entries, so each row starts 16-byte aligned.
  const u32 ostride = std::max<u32>(8, (L.nf + 7u) & ~7u);
  const size_t nl = L.lanes.size();
  std::vector<u16> offp(nl * ostride, 0xFFFFu);
  for (size_t l = 0; l < nl; ++l)
    std::memcpy(&offp[l * ostride], &L.off[l * L.nf], L.nf * sizeof(u16));
  if (up(&b.off, offp.data(), offp.size() * 2)) return -1;
  std::vector<uint4> fc(L.nf);
  for (u32 j = 0; j < L.nf; ++j) fc[j] = make_uint4(L.step[j], L.magic[j], L.c32[j], 0u);
  if (up(&b.fc, fc.data(), fc.size() * sizeof(uint4))) return -1;
  dev::LayoutDev &d = b.d;
  d.Mp = L.Mp;
  d.lanes = (const u64 *)b.lanes;
  d.mask = (const u32 *)b.mask;
  d.step = (const u32 *)b.step;
  d.magic = (const u32 *)b.magic;
  d.c32 = (const u32 *)b.c32;
  d.fc = (const uint4 *)b.fc;
  d.off = (const u16 *)b.off;
  d.P = (u32)L.P;
  d.Pmagic = L.P == 1 ? 0xFFFFFFFFu : (u32)((1ull << 32) / L.P);
  d.Pc32 = (u32)((1ull << 32) % L.P);
  d.nl = (u32)L.lanes.size();
  d.nf = L.nf;
  d.ncov = L.ncov;
  d.mask_words = L.mask_words;
  d.ostride = ostride;
  m.layouts.push_back(b);
  if (ostride > m.off_max || L.mask_words > m.mask_max) {
    m.off_max = std::max(m.off_max, ostride);
    m.mask_max = std::max(m.mask_max, (u32)L.mask_words);
    for (Slot &b : m.slot) b.graph_ok = false;
  }
  // Mirror the descriptor array on the device.
  if (m.layouts.size() > m.lays_cap) {
    // Batches in flight read the old array: let them finish before it goes.
    for (Slot &b : m.slot) HIP_TRY(hipStreamSynchronize(b.stream));
    if (m.d_lays) HIP_TRY(hipFree(m.d_lays));
    m.lays_cap = std::max<size_t>(8, m.layouts.size() * 2);
    HIP_TRY(hipMalloc(&m.d_lays, m.lays_cap * sizeof(dev::LayoutDev)));
  }
  std::vector<dev::LayoutDev> all;
  for (const DevLayoutBufs &x : m.layouts) all.push_back(x.d);
  HIP_TRY(hipMemcpy(m.d_lays, all.data(), all.size() * sizeof(dev::LayoutDev), hipMemcpyHostToDevice));
  return (int)m.layouts.size() - 1;
}

// Pinned host buffer with its device-side address (for k_pack's stores).
static hipError_t pinned_alloc(void **h, void **d, size_t bytes) {
  if (*h) {
    const hipError_t e = hipHostFree(*h);
    if (e != hipSuccess) return e;
  }
  *h = nullptr;
  hipError_t e = hipHostMalloc(h, bytes, kPinnedFlags);
  if (e == hipSuccess && d) e = hipHostGetDevicePointer(d, *h, 0);
  return e;
}

// Survivor buffers of slot si, grown to hold need entries. Any reallocation invalidates
// the slot's captured graph.
int DeviceCtx::grow_surv(int si, size_t need) {
  Slot &b = impl_->slot[si];
  if (need <= b.surv_cap) return 0;
  size_t cap = std::max<size_t>(b.surv_cap ? b.surv_cap : (1u << 16), 1u << 16);
  while (cap < need) cap *= 2;
  if (cap > (1ull << 28)) { err_ = "survivor count exceeds 2^28 per batch"; return -1; }
  for (void *p : {(void *)b.d_surv_p, (void *)b.d_surv_w, (void *)b.d_flags, (void *)b.d_one, (void *)b.d_rung})
    if (p) HIP_TRY(hipFree(p));
  HIP_TRY(hipMalloc(&b.d_surv_p, cap * sizeof(u64)));
  HIP_TRY(hipMalloc(&b.d_one, cap * sizeof(u64)));
  HIP_TRY(hipMalloc(&b.d_rung, cap * sizeof(dev::RungOut)));
  HIP_TRY(hipMalloc(&b.d_surv_w, cap * sizeof(u32)));
  HIP_TRY(hipMalloc(&b.d_flags, cap));
  HIP_TRY(pinned_alloc((void **)&b.h_surv_p, (void **)&b.dh_surv_p, cap * sizeof(u64)));
  HIP_TRY(pinned_alloc((void **)&b.h_surv_w, (void **)&b.dh_surv_w, cap * sizeof(u32)));
  HIP_TRY(pinned_alloc((void **)&b.h_flags, (void **)&b.dh_flags, cap));
  HIP_TRY(pinned_alloc((void **)&b.h_rung, (void **)&b.dh_rung, cap * sizeof(dev::RungOut)));
  b.surv_cap = cap;
  b.graph_ok = false;
  return 0;
}

// Enqueues slot si's batch on the stream: staging, sieve, Miller-Rabin, stage D and pack,
// between events ev0 and ev3 (ev1/ev2 time the sieve kernel's own dispatch when it is not
// replayed from a graph). With the graph on, the batch is captured once per slot and launch shape
// (on a stream where nothing runs) and replayed after that. The persistent grid does not
// depend on the batch, which is what lets one graph serve every batch that fits.
int DeviceCtx::enqueue_batch(int si) {
  Impl &m = *impl_;
  Slot &b = m.slot[si];
  const LaunchCfg &cfg = b.cfg;
  auto enqueue = [&](hipStream_t st, bool timed) -> int {
    dev::k_stage<<<std::max(1, cus_), 256, 0, st>>>(b.dh_hdr, (const uint2 *)b.dh_tiles, b.d_hdr,
                                                        (uint2 *)b.d_tiles, (const uint4 *)b.dh_items,
                                                        (uint4 *)b.d_items, b.d_tile_counts, b.d_scalars,
                                                        (u64 *)b.d_board, dev::kBoardSlots * 4u, b.dh_sync,
                                                        b.d_stamps, m.d_active);
    HIP_TRY(hipGetLastError());
    // Timed: ev1/ev2 are taken by the dispatch of k_sieve itself (its start and end),
    // not by separate event packets around it.
    const hipEvent_t e1 = timed ? b.ev1 : nullptr, e2 = timed ? b.ev2 : nullptr;
    if (cfg.persistent) {
      hipExtLaunchKernelGGL(b.off_words ? dev::k_sieve<true, true> : dev::k_sieve<true, false>, dim3(b.grid), dim3(cfg.block), (u32)b.lds, st, e1, e2, 0,
                            (const TileDesc *)b.d_tiles, (const dev::ItemDesc *)b.d_items, (const dev::BatchHdr *)b.d_hdr,
                            (const dev::LayoutDev *)m.d_lays, b.d_tile_counts, b.d_surv_p, b.d_surv_w,
                            (u32)b.surv_cap, b.d_scalars, b.d_scalars + 1, b.lds_words, b.off_words, m.mask_max,
                            b.d_stamps);
    } else {
      hipExtLaunchKernelGGL(b.off_words ? dev::k_sieve<false, true> : dev::k_sieve<false, false>, dim3((u32)b.ni), dim3(cfg.block), (u32)b.lds, st, e1, e2, 0,
                            (const TileDesc *)b.d_tiles, (const dev::ItemDesc *)b.d_items, (const dev::BatchHdr *)b.d_hdr,
                            (const dev::LayoutDev *)m.d_lays, b.d_tile_counts, b.d_surv_p, b.d_surv_w,
                            (u32)b.surv_cap, b.d_scalars, b.d_scalars + 1, b.lds_words, b.off_words, m.mask_max,
                            b.d_stamps);
    }
    HIP_TRY(hipGetLastError());
    hipExtLaunchKernelGGL(dev::k_mr_base2, dim3(std::max(1, cus_ * 2)), dim3(256), 0, st, timed ? b.ev4 : nullptr,
                          timed ? b.ev5 : nullptr, 0, (const u64 *)b.d_surv_p, (const u32 *)b.d_scalars,
                          (u32)b.surv_cap, b.d_flags, b.d_one, b.d_stamps + 2);
    HIP_TRY(hipGetLastError());
    if (cfg.device_tail) {
      // k_rung finishes the Miller-Rabin bases of the pending values itself.
      hipExtLaunchKernelGGL(dev::k_rung, dim3(std::max(1, cus_)), dim3(64 * dev::kRungWaves), 0, st,
                            timed ? b.ev6 : nullptr, timed ? b.ev7 : nullptr, 0, (const u64 *)b.d_surv_p,
                            (const u32 *)b.d_scalars, (u32)b.surv_cap, b.d_flags, (const u64 *)b.d_one, m.trial,
                            b.d_rung, b.d_scalars + 2, b.d_board, b.d_scalars + 3, b.d_stamps + 4,
                            dev::kRungSlots, (const u32 *)m.d_active);
    } else {
      hipExtLaunchKernelGGL(dev::k_mr_bases, dim3(std::max(1, cus_ * 4)), dim3(256), 0, st, timed ? b.ev6 : nullptr,
                            timed ? b.ev7 : nullptr, 0, (const u64 *)b.d_surv_p, (const u32 *)b.d_scalars,
                            (u32)b.surv_cap, b.d_flags, (const u64 *)b.d_one, b.d_stamps + 4);
    }
    HIP_TRY(hipGetLastError());
    dev::k_pack<<<std::max(1, cus_ / 4), 256, 0, st>>>(b.d_hdr, b.d_tile_counts, b.d_scalars, (u32)b.surv_cap,
                                                   b.d_surv_p, b.d_surv_w, b.d_flags, b.d_rung,
                                                   cfg.device_tail ? 1u : 0u, b.dh_count, b.dh_tile_counts,
                                                   b.dh_surv_p, b.dh_surv_w, b.dh_flags, b.dh_rung, b.d_scalars + 4,
                                                   b.dh_sync, (const u64 *)b.d_stamps, m.d_active);
    HIP_TRY(hipGetLastError());
    return 0;
  };
  const bool graphed = cfg.graph && cfg.persistent;
  if (!graphed) HIP_TRY(hipEventRecord(b.ev0, b.stream));
  if (graphed) {
    if (!b.graph_ok || b.graph_block != cfg.block || b.graph_lds != b.lds || b.graph_tail != cfg.device_tail ||
        b.graph_grid != b.grid || b.graph_off_words != b.off_words) {
      if (b.graph_exec) { HIP_TRY(hipGraphExecDestroy(b.graph_exec)); b.graph_exec = nullptr; }
      if (b.graph) { HIP_TRY(hipGraphDestroy(b.graph)); b.graph = nullptr; }
      HIP_TRY(hipStreamBeginCapture(m.capture, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue(m.capture, false);
      hipGraph_t g = nullptr;
      const hipError_t e = hipStreamEndCapture(m.capture, &g);
      if (rc) return rc;
      if (e != hipSuccess) { err_ = std::string("hipStreamEndCapture: ") + hipGetErrorString(e); return -1; }
      b.graph = g;
      HIP_TRY(hipGraphInstantiate(&b.graph_exec, b.graph, nullptr, nullptr, 0));
      b.graph_ok = true;
      b.graph_block = cfg.block;
      b.graph_lds = b.lds;
      b.graph_tail = cfg.device_tail;
      b.graph_grid = b.grid;
      b.graph_off_words = b.off_words;
    }
    HIP_TRY(hipGraphLaunch(b.graph_exec, b.stream));
  } else {
    if (enqueue(b.stream, true)) return -1;
    HIP_TRY(hipEventRecord(b.ev3, b.stream));
  }
  return 0;
}

int DeviceCtx::launch(const std::vector<TileDesc> &tiles, u32 nwin, bool has_sieve, const LaunchCfg &cfg) {
  Impl &m = *impl_;
  if (device_ < 0) { err_ = "device not open"; return -1; }
  // Reserve the slot after the newest one in flight; released again if the launch fails.
  int si;
  {
    std::lock_guard<std::mutex> lk(m.slot_mu);
    if (m.busy == kMaxInFlight) { err_ = "every batch slot is in flight"; return -1; }
    si = (m.head + m.busy) % kMaxInFlight;
    ++m.busy;
  }
  struct Release {
    Impl &m;
    bool keep = false;
    ~Release() {
      if (keep) return;
      std::lock_guard<std::mutex> lk(m.slot_mu);
      if (--m.busy == 0) m.head = 0;
    }
  } release{m};
  std::lock_guard<std::mutex> launch_lk(m.launch_mu);
  HIP_TRY(hipSetDevice(device_));
  const size_t nt = tiles.size();
  if (cfg.block == 0 || cfg.block > 1024 || cfg.block % 64 || cfg.max_bits == 0 || cfg.max_bits % 32 ||
      cfg.max_bits > kMaxTileBits || cfg.group == 0 || cfg.group > dev::kMaxGroup) {
    err_ = "bad launch configuration";
    return -1;
  }
  // Host-side validation of every descriptor against the tables it indexes.
  for (const TileDesc &t : tiles) {
    if (t.layout >= m.layouts.size() || t.lane >= m.layouts[t.layout].d.nl || t.win >= nwin ||
        t.nbits == 0 || t.nbits > cfg.max_bits) {
      err_ = "tile descriptor out of range";
      return -1;
    }
  }
  if (nt > 0xFFFFFFFFull || nwin > 0x55555555u) { err_ = "batch too large"; return -1; }
  Slot &b = m.slot[si];
  b.nt = nt;
  b.nwin = nwin;
  b.cfg = cfg;
  if (nt > 0) {
    // Capacities. Any reallocation invalidates the slot's captured graph.
    if (nt > b.tiles_cap) {
      if (b.d_tiles) HIP_TRY(hipFree(b.d_tiles));
      b.tiles_cap = std::max<size_t>(nt, 2 * b.tiles_cap);
      HIP_TRY(hipMalloc(&b.d_tiles, b.tiles_cap * sizeof(TileDesc)));
      if (b.d_items) HIP_TRY(hipFree(b.d_items));
      HIP_TRY(hipMalloc(&b.d_items, b.tiles_cap * sizeof(dev::ItemDesc)));
      if (b.d_tile_counts) HIP_TRY(hipFree(b.d_tile_counts));
      HIP_TRY(hipMalloc(&b.d_tile_counts, b.tiles_cap * sizeof(uint4)));
      HIP_TRY(pinned_alloc((void **)&b.h_tiles, (void **)&b.dh_tiles, b.tiles_cap * sizeof(TileDesc)));
      HIP_TRY(pinned_alloc((void **)&b.h_items, (void **)&b.dh_items, b.tiles_cap * sizeof(dev::ItemDesc)));
      HIP_TRY(pinned_alloc((void **)&b.h_tile_counts, (void **)&b.dh_tile_counts, b.tiles_cap * sizeof(uint4)));
      b.graph_ok = false;
    }
    if (grow_surv(si, 1u << 16)) return -1;
    // The slot's previous batch has been finished, so its staging buffers are free.
    b.h_hdr->ntiles = (u32)nt;
    b.h_hdr->nwin = nwin;
    b.h_hdr->has_sieve = has_sieve ? 1 : 0;
    b.h_hdr->seq = ++b.seq;
    std::memcpy(b.h_tiles, tiles.data(), nt * sizeof(TileDesc));   // k_stage copies both in

    // Work items: a tile of at least half the largest tile size is an item alone; shorter
    // consecutive tiles (the planner sends them out longest first, so they come last) are
    // grouped, up to cfg.group tiles and max_bits positions (each from a 32-position chunk
    // boundary). A group shares one init, one count and the barriers, but marks its tiles
    // one after another with their offset rows read from global memory, so long groups
    // lose what they save (the 512-window batch from 1e9 to 1.8e19: 97.3 us one tile per
    // item, 77.4 in groups of up to 4, 99.6 of up to 16). The map is sized to the largest
    // item (whole chunks): short items then leave room for more resident workgroups.
    auto padded = [](u32 nb) { return (nb + 31u) & ~31u; };
    size_t ni = 0;
    u32 tile_max = 0;
    for (size_t i = 0; i < nt;) {
      size_t j = i + 1;
      u32 bits = padded(tiles[i].nbits);
      if (2 * tiles[i].nbits < cfg.max_bits)
        while (j < nt && j - i < cfg.group && bits + padded(tiles[j].nbits) <= cfg.max_bits)
          bits += padded(tiles[j++].nbits);
      b.h_items[ni++] = dev::ItemDesc{tiles[i], (u32)i, (u32)(j - i)};
      tile_max = std::max(tile_max, bits);
      i = j;
    }
    b.ni = ni;
    b.h_hdr->nitems = (u32)ni;
    const u32 lds_bytes = tile_max;
    b.lds_words = lds_bytes / 4;
    // The tile's row of filter offsets is staged in LDS beside the map when both fit.
    // Large generated filter sets (about 77,000 marking units at mmax 10000, a 154 KB
    // row) leave no room for a long tile beside it: then k_sieve reads the row from
    // global memory instead (off_words = 0).
    const u32 scalar_bytes = 4 * sizeof(u32), mask_bytes = m.mask_max * 4;
    b.off_words = m.off_max / 2;
    if (lds_bytes + b.off_words * 4 + mask_bytes + scalar_bytes > lds_limit_) b.off_words = 0;
    b.lds = lds_bytes + b.off_words * 4 + mask_bytes + scalar_bytes;
    if (b.lds > lds_limit_) {
      err_ = "tile of " + std::to_string(tile_max) + " positions and " + std::to_string(m.off_max) +
             " filter offsets need " + std::to_string(b.lds) + " B of LDS, above the device's " +
             std::to_string(lds_limit_);
      return -1;
    }
    if (m.occ_block != cfg.block || m.occ_lds != b.lds || m.occ_off_words != b.off_words) {
      if (b.lds > 65536) {
        for (const void *f : {(const void *)dev::k_sieve<true, true>, (const void *)dev::k_sieve<true, false>,
                              (const void *)dev::k_sieve<false, true>, (const void *)dev::k_sieve<false, false>})
          HIP_TRY(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)b.lds));
      }
      int nb = 0;
      if (b.off_words) HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dev::k_sieve<true, true>, cfg.block, b.lds));
      else HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, dev::k_sieve<true, false>, cfg.block, b.lds));
      m.occupancy_blocks = std::max(1, nb);
      m.occ_block = cfg.block;
      m.occ_lds = b.lds;
      m.occ_off_words = b.off_words;
    }
    // Fewer resident waves than fit can mark faster (measured: 16 waves per CU beat 24
    // and 32 for 256-, 512- and 1024-thread workgroups), so the grid is sized to
    // cfg.waves_per_cu when that is set.
    u32 per_cu = (u32)m.occupancy_blocks;
    if (cfg.waves_per_cu) per_cu = std::min(per_cu, std::max(1u, cfg.waves_per_cu / ((cfg.block + 63u) / 64u)));
    b.grid = (u32)cus_ * per_cu;
    if (enqueue_batch(si)) return -1;
  }
  release.keep = true;
  return 0;
}

int DeviceCtx::finish(BatchResult &out) {
  Impl &m = *impl_;
  if (device_ < 0) { err_ = "device not open"; return -1; }
  int si;
  {
    std::lock_guard<std::mutex> lk(m.slot_mu);
    if (m.busy == 0) { err_ = "no batch in flight"; return -1; }
    si = m.head;
  }
  const int rc = read_batch(si, out);
  // The slot is released only now, once its mirrors have been read, and whatever the
  // outcome. With nothing left in flight the next batch takes slot 0 again, so
  // one-at-a-time use keeps to one slot and one graph.
  std::lock_guard<std::mutex> lk(m.slot_mu);
  m.head = (m.head + 1) % kMaxInFlight;
  if (--m.busy == 0) m.head = 0;
  return rc;
}

// Waits until k_pack has posted slot si's batch number into pinned host memory: no HIP
// call while the batch runs, so another thread may launch batches meanwhile. Every
// millisecond the stream is asked for an error, so a failed batch cannot hang the wait.
int DeviceCtx::wait_slot(int si) {
  Slot &b = impl_->slot[si];
  const u64 want = b.seq;
  auto t_next = std::chrono::steady_clock::now() + std::chrono::milliseconds(1);
  for (;;) {
    if (__atomic_load_n(&b.h_sync[0], __ATOMIC_ACQUIRE) == want) return 0;
    __builtin_ia32_pause();
    if (std::chrono::steady_clock::now() < t_next) continue;
    t_next += std::chrono::milliseconds(1);
    const hipError_t e = hipStreamQuery(b.stream);
    if (e == hipSuccess) {
      if (__atomic_load_n(&b.h_sync[0], __ATOMIC_ACQUIRE) == want) return 0;
      err_ = "batch finished without posting its completion";
      return -1;
    }
    if (e != hipErrorNotReady) { err_ = std::string("batch failed: ") + hipGetErrorString(e); return -1; }
  }
}

// Waits for slot si's batch (only that batch: later ones keep running), replays it on a
// survivor-list overflow, and copies its results out of the pinned mirrors.
int DeviceCtx::read_batch(int si, BatchResult &out) {
  Impl &m = *impl_;
  HIP_TRY(hipSetDevice(device_));
  Slot &b = m.slot[si];
  out.counts.assign(b.nwin, WinCounts{0, 0, 0});
  out.surv_p.clear();
  out.surv_w.clear();
  out.surv_prime.clear();
  out.rung.clear();
  out.replays = 0;
  out.device_ms = out.sieve_ms = out.mr_ms = out.tail_kernel_ms = 0;
  out.sieve_event_ms = out.mr_event_ms = out.tail_event_ms = 0;
  if (b.nt == 0) return 0;
  const bool use_graph = b.cfg.graph && b.cfg.persistent;
  for (;;) {
    if (wait_slot(si)) return -1;
    float sv = 0, mr = 0, rg = 0;
    if (!use_graph) {
      // Eager launches time their kernels with events, which may complete a little after
      // the flag is posted.
      HIP_TRY(hipEventSynchronize(b.ev3));
      HIP_TRY(hipEventElapsedTime(&sv, b.ev1, b.ev2));
      HIP_TRY(hipEventElapsedTime(&mr, b.ev4, b.ev5));
      HIP_TRY(hipEventElapsedTime(&rg, b.ev6, b.ev7));
    }
    const u64 *st = b.h_sync + 3;
    auto span = [&](int k) { return st[2 * k + 1] > st[2 * k] ? (double)(st[2 * k + 1] - st[2 * k]) * clock_ms_per_tick_ : 0.0; };
    out.device_ms += (double)(b.h_sync[2] - b.h_sync[1]) * clock_ms_per_tick_;
    out.sieve_ms += span(0);
    out.mr_ms += span(1);
    out.tail_kernel_ms += span(2);
    out.sieve_event_ms += use_graph ? -1.0 : sv;
    out.mr_event_ms += use_graph ? -1.0 : mr;
    out.tail_event_ms += use_graph ? -1.0 : rg;
    const u32 cnt = *b.h_count;
    if (cnt <= b.surv_cap) break;
    // Overflow: the counts are complete, the list is not. Grow and replay the batch
    // (sieve/cuda/scan.cu:223-242 does the same).
    ++out.replays;
    std::lock_guard<std::mutex> launch_lk(m.launch_mu);
    if (grow_surv(si, cnt)) return -1;
    b.h_hdr->seq = ++b.seq;
    if (enqueue_batch(si)) return -1;
  }
  // Everything is in the pinned mirrors now (k_pack posted the batch's number last).
  const u32 cnt = *b.h_count;
  out.surv_p.assign(b.h_surv_p, b.h_surv_p + cnt);
  out.surv_w.assign(b.h_surv_w, b.h_surv_w + cnt);
  out.surv_prime.resize(cnt);
  for (u32 i = 0; i < cnt; ++i) out.surv_prime[i] = b.h_flags[i] != 0;   // pending after every base == prime
  if (b.cfg.device_tail) {
    static_assert(sizeof(dev::RungOut) == sizeof(RungResult), "RungOut mirrors RungResult");
    out.rung.resize(cnt);
    for (u32 i = 0; i < cnt; ++i)
      if (out.surv_prime[i]) std::memcpy(&out.rung[i], &b.h_rung[i], sizeof(RungResult));
  }
  for (size_t i = 0; i < b.nt; ++i) {
    WinCounts &c = out.counts[b.h_tiles[i].win];
    c.positions += b.h_tile_counts[i].x;
    c.after_covers += b.h_tile_counts[i].y;
    c.survivors += b.h_tile_counts[i].z;
  }
  return 0;
}

int DeviceCtx::run(const std::vector<TileDesc> &tiles, u32 nwin, bool has_sieve, const LaunchCfg &cfg,
                   BatchResult &out) {
  if (impl_->busy) { err_ = "batches are in flight: finish them before run"; return -1; }
  if (launch(tiles, nwin, has_sieve, cfg)) return -1;
  return finish(out);
}

int DeviceCtx::in_flight() const { return impl_->busy; }

int DeviceCtx::rung(const u64 *vals, size_t n, RungResult *out, bool direct_only) {
  Impl &m = *impl_;
  if (device_ < 0) { err_ = "device not open"; return -1; }
  if (n == 0) return 0;
  if (n > 0x0FFFFFFFull) { err_ = "too many values"; return -1; }
  HIP_TRY(hipSetDevice(device_));
  u64 *d_v = nullptr;
  uint8_t *d_f = nullptr;
  u32 *d_n = nullptr;
  dev::RungOut *d_o = nullptr;
  dev::RhoBoard *d_b = nullptr;
  HIP_TRY(hipMalloc(&d_v, n * 8));
  HIP_TRY(hipMalloc(&d_f, n));
  HIP_TRY(hipMalloc(&d_n, 16));
  HIP_TRY(hipMalloc(&d_o, n * sizeof(dev::RungOut)));
  HIP_TRY(hipMalloc(&d_b, dev::kBoardSlots * sizeof(dev::RhoBoard)));
  const u32 n32[4] = {(u32)n, 0u, 0u, 0u};   // count, queue head, unused, survivors finished
  std::vector<uint8_t> ones(n, 1);
  HIP_TRY(hipMemcpyAsync(d_v, vals, n * 8, hipMemcpyHostToDevice, m.stream));
  HIP_TRY(hipMemcpyAsync(d_f, ones.data(), n, hipMemcpyHostToDevice, m.stream));
  HIP_TRY(hipMemcpyAsync(d_n, n32, 16, hipMemcpyHostToDevice, m.stream));
  HIP_TRY(hipMemsetAsync(d_b, 0, dev::kBoardSlots * sizeof(dev::RhoBoard), m.stream));
  dev::k_rung<<<std::max(1, cus_), 64 * dev::kRungWaves, 0, m.stream>>>(d_v, d_n, n32[0], d_f, nullptr, m.trial, d_o,
                                                                         d_n + 1, d_b, d_n + 3, nullptr,
                                                                         direct_only ? 0u : dev::kRungSlots);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(out, d_o, n * sizeof(dev::RungOut), hipMemcpyDeviceToHost, m.stream));
  HIP_TRY(hipStreamSynchronize(m.stream));
  for (void *p : {(void *)d_v, (void *)d_f, (void *)d_n, (void *)d_o, (void *)d_b}) HIP_TRY(hipFree(p));
  return 0;
}

int DeviceCtx::mem_info(size_t &free_b, size_t &total_b) {
  if (device_ < 0) { err_ = "device not open"; return -1; }
  HIP_TRY(hipSetDevice(device_));
  HIP_TRY(hipMemGetInfo(&free_b, &total_b));
  return 0;
}

size_t DeviceCtx::resident_bytes() const {
  const Impl &m = *impl_;
  const size_t per_surv = sizeof(u64) * 2 + sizeof(u32) + 1 + sizeof(dev::RungOut);
  size_t batch = 0;
  for (const Slot &b : m.slot) batch += b.tiles_cap * (sizeof(TileDesc) + sizeof(uint4)) + b.surv_cap * per_surv;
  return layout_bytes_ + m.trial.n * (sizeof(uint2) + 16) + batch;
}

int DeviceCtx::lds_scatter_peak(double &stores_per_s) {
  Impl &m = *impl_;
  if (device_ < 0) { err_ = "device not open"; return -1; }
  HIP_TRY(hipSetDevice(device_));
  u32 *sink = nullptr;
  const u32 blocks = (u32)cus_ * 3, iters = 1u << 14;
  HIP_TRY(hipMalloc(&sink, blocks * sizeof(u32)));
  dev::k_lds_scatter<<<blocks, 512, 0, m.stream>>>(64, sink);   // warm-up
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(m.pev0, m.stream));
  dev::k_lds_scatter<<<blocks, 512, 0, m.stream>>>(iters, sink);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(m.pev1, m.stream));
  HIP_TRY(hipStreamSynchronize(m.stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, m.pev0, m.pev1));
  HIP_TRY(hipFree(sink));
  stores_per_s = (double)blocks * 512.0 * iters / (ms * 1e-3);
  return 0;
}

int DeviceCtx::lds_peak(double &ops_per_s) {
  Impl &m = *impl_;
  if (device_ < 0) { err_ = "device not open"; return -1; }
  HIP_TRY(hipSetDevice(device_));
  u32 *sink = nullptr;
  const u32 blocks = (u32)cus_ * 8, iters = 1u << 14;
  HIP_TRY(hipMalloc(&sink, blocks * sizeof(u32)));
  dev::k_lds_peak<<<blocks, 256, 0, m.stream>>>(64, sink);   // warm-up
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(m.pev0, m.stream));
  dev::k_lds_peak<<<blocks, 256, 0, m.stream>>>(iters, sink);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(m.pev1, m.stream));
  HIP_TRY(hipStreamSynchronize(m.stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, m.pev0, m.pev1));
  HIP_TRY(hipFree(sink));
  ops_per_s = (double)blocks * 256.0 * iters / (ms * 1e-3);
  return 0;
}

int DeviceCtx::primality(const u64 *vals, size_t n, uint8_t *flags) {
  Impl &m = *impl_;
  if (device_ < 0) { err_ = "device not open"; return -1; }
  if (n == 0) return 0;
  if (n > 0x0FFFFFFFull) { err_ = "too many values"; return -1; }
  HIP_TRY(hipSetDevice(device_));
  u64 *d_v = nullptr, *d_one = nullptr;
  uint8_t *d_f = nullptr;
  u32 *d_n = nullptr;
  HIP_TRY(hipMalloc(&d_v, n * 8));
  HIP_TRY(hipMalloc(&d_one, n * 8));
  HIP_TRY(hipMalloc(&d_f, n));
  HIP_TRY(hipMalloc(&d_n, 4));
  const u32 n32 = (u32)n;
  HIP_TRY(hipMemcpyAsync(d_v, vals, n * 8, hipMemcpyHostToDevice, m.stream));
  HIP_TRY(hipMemcpyAsync(d_n, &n32, 4, hipMemcpyHostToDevice, m.stream));
  dev::k_mr_base2<<<std::max(1, cus_ * 4), 256, 0, m.stream>>>(d_v, d_n, n32, d_f, d_one);
  HIP_TRY(hipGetLastError());
  dev::k_mr_bases<<<std::max(1, cus_ * 8), 256, 0, m.stream>>>(d_v, d_n, n32, d_f, d_one);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpyAsync(flags, d_f, n, hipMemcpyDeviceToHost, m.stream));
  HIP_TRY(hipStreamSynchronize(m.stream));
  for (size_t i = 0; i < n; ++i) flags[i] = flags[i] != 0;
  for (void *p : {(void *)d_v, (void *)d_one, (void *)d_f, (void *)d_n}) HIP_TRY(hipFree(p));
  return 0;
}

// The kernels of primality() on caller-owned device buffers, with the count and the
// per-value Montgomery ones in this device's copy of dev::scratch: nothing is allocated
// and nothing waits. Pointers are checked to be device memory first, so a host pointer is
// an error here and not a fault in a kernel.
int DeviceCtx::primality_noalloc(const u64 *d_vals, size_t n, uint8_t *d_flags, void *stream) {
  if (device_ < 0 || !scratch_) { err_ = "device not open"; return -1; }
  if (n > kNoAllocMaxValues) {
    err_ = "es_primality_noalloc: " + std::to_string(n) + " values exceed the scratch maximum of " +
           std::to_string(kNoAllocMaxValues);
    return -2;
  }
  if (n == 0) return 0;
  if (!d_vals || !d_flags) { err_ = "es_primality_noalloc: null buffer"; return -2; }
  HIP_TRY(hipSetDevice(device_));
  for (const void *ptr : {(const void *)d_vals, (const void *)d_flags}) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, ptr) != hipSuccess || a.type != hipMemoryTypeDevice || a.device != device_) {
      (void)hipGetLastError();
      err_ = "es_primality_noalloc: buffers must be device memory of the scanner's device";
      return -2;
    }
  }
  hipStream_t st = (hipStream_t)stream;
  u32 *count = (u32 *)scratch_;
  u64 *one = (u64 *)scratch_ + dev::kScratchOne;
  const u32 n32 = (u32)n;
  dev::k_set_count<<<1, 64, 0, st>>>(count, n32);
  HIP_TRY(hipGetLastError());
  dev::k_mr_base2<<<std::max(1, cus_ * 4), 256, 0, st>>>(d_vals, count, n32, d_flags, one);
  HIP_TRY(hipGetLastError());
  dev::k_mr_bases<<<std::max(1, cus_ * 8), 256, 0, st>>>(d_vals, count, n32, d_flags, one);
  HIP_TRY(hipGetLastError());
  dev::k_flags_to_prime<<<std::max(1, cus_), 256, 0, st>>>(d_flags, n32);
  HIP_TRY(hipGetLastError());
  return 0;
}

}  // namespace es
