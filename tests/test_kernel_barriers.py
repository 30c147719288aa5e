"""Source check of the LDS phase barriers in k_sieve and k_rung.

A CPU replay runs a tile's waves one after another, so it cannot see a missing barrier
between two phases that touch the same LDS bytes (a wave's count of chunks it did not
mark, say). This test reads the kernel source instead: for every pair of phases in which
one writes LDS that the other reads or overwrites, a __syncthreads() must lie between
them in program order. The checker itself is tested on copies of the source with one
barrier taken out.
"""
import os
import re

import pytest

from tests import paths

SRC = os.path.join(paths.PKG, "csrc", "es_device.hip")

# (name, a statement that starts the phase), in program order within the tile loop.
SIEVE_PHASES = [
    ("rows staged", "dst[i] = src[i];"),
    ("mask row staged", "smask[i] = msrc[i];"),
    ("group table staged", "stage_sub(gsub[cur], d, ns, tid);"),
    ("mask init", "store_chunk(map32, c, word);"),
    ("group init", "atomicAdd(&gcnt[j][0], (u32)__popc(word));"),
    ("stage B", "mark(true);"),
    ("count after covers", "const u32 n = __popc(chunk_bits(map32, c));"),
    ("stage B'", "mark(false);"),
    ("survivor count", "my_surv += n;"),
    ("compaction", "u32 x = chunk_bits(map32, ch);"),
    ("group counters flushed", "const u32 pos = gcnt[tid][0], aft = gcnt[tid][1], sur = gcnt[tid][2];"),
    ("next rows stored", "if (tid < nv_next) dst[tid] = pre0;"),
    ("next group table stored", "stage_sub(gsub[cur ^ 1u], dnext, itn.count, tid);"),
    ("claim stored", "sc[2] = 2 * gridDim.x + claim;"),
    ("claim read", "tnext = sc[2];"),
    ("next tile", "stamp_end(stamp, threadIdx.x == 0);"),   # after the loop: the next item starts over
]
# Ordered pairs (writer phase, the next phase that conflicts with it) and the LDS they share.
SIEVE_ORDER = [
    ("mask row staged", "mask init"),            # the mask row (and the offset row) of a tile staged at its top
    ("group table staged", "group init"),        # gsub[cur]: a group's sub-tile table staged at its top
    ("mask init", "stage B"),                    # the byte map
    ("group init", "stage B"),                   # the byte map (a group's)
    ("stage B", "count after covers"),           # the byte map
    ("count after covers", "stage B'"),          # the byte map (counted before B' clears it)
    ("stage B'", "survivor count"),              # the byte map
    ("survivor count", "group counters flushed"),  # gcnt: every wave's counts before the flush reads them
    ("claim stored", "claim read"),              # sc[2]: the item after next
    ("stage B'", "next rows stored"),            # the offset and mask rows: read by init and marking, then replaced
    ("next rows stored", "next tile"),           # the next tile's rows, read by its init and marking
    ("next group table stored", "next tile"),    # gsub[cur ^ 1]: the next group's table, read from its init on
    ("group counters flushed", "next tile"),     # gcnt zeroed by the flush, counted into by the next group
    ("compaction", "next tile"),                 # the map and sc[2], reused by the next item
]

RUNG_PHASES = [
    ("pending reset", "if (tid == 0) npend = 0;"),
    ("pending counts", "if (lane == 0) wcnt[wv] = (u32)__popcll(m);"),
    ("pending bases", "u32 base = npend, total = 0;"),
    ("pending list", "if (at < kPendList) pend[at] = i;"),
    ("pending total", "if (tid == 0) npend += total;"),
    ("list length", "listed = npend <= kPendList;"),
    ("index claim", "cur_idx[it & 1] = atomicAdd(queue, 1u);"),
    ("index read", "const u32 q = cur_idx[it & 1];"),
    ("help gate cleared", "if (tid == 0) help_batch = 0;"),
    ("help gate set", "if (big) help_batch = 1;"),
    ("help gate read", "if (help_batch && wv < kBoardHelpWaves)"),
    ("list read", "const u32 idx = listed ? pend[q] : q;"),
    ("prime state init", "if (tid < kRungSlots) nfac[tid] = 0;"),
    ("bases and rung sieve", "const bool pending = f0 == kPending;"),
    ("verdict read", "if (mr_composite) {"),
    ("rungs", "help_split(&rs, seen, wv, lane);"),
    ("all waves done", "if (lane == 0) atomicAdd(&rs.done, 1u);"),
    ("stage E reset", "found_k = host_k = stop_k = kDirectPairs;"),
    ("result", "RungOut res{0, 0, 0, 0, kRungNone};"),
    ("next prime", "stamp_end(stamp, lane == 0);"),
]
RUNG_ORDER = [
    ("pending reset", "pending bases"),            # npend
    ("pending counts", "pending bases"),           # wcnt
    ("pending bases", "pending total"),            # npend and wcnt read before npend changes
    ("pending total", "list length"),              # npend
    ("pending list", "list read"),                 # pend
    ("index claim", "index read"),                 # cur_idx
    ("help gate cleared", "help gate set"),        # help_batch
    ("help gate set", "help gate read"),
    ("prime state init", "bases and rung sieve"),  # nfac, wres, next_k, found_k, rs, mr_composite
    ("bases and rung sieve", "verdict read"),      # mr_composite, nfac, fac
    ("all waves done", "stage E reset"),           # found_k, host_k read before stage E resets them
    ("stage E reset", "result"),                   # next_k, found_k, host_k, stop_k, rs, wres
    ("all waves done", "result"),                  # wres, found_k, host_k
    ("result", "next prime"),                      # every per-prime LDS array
]


def kernel_body(src, name):
    m = re.search(r"void %s\(" % re.escape(name), src)
    assert m, name
    i = src.index("{", src.index(")", m.end()))
    depth = 0
    for j in range(i, len(src)):
        depth += {"{": 1, "}": -1}.get(src[j], 0)
        if depth == 0:
            return src[i:j + 1]
    raise AssertionError("unbalanced " + name)


def strip_comments(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def missing_barriers(body, phases, order):
    body = strip_comments(body)
    pos = {}
    last = -1
    for name, stmt in phases:
        k = body.find(stmt, last + 1)
        assert k >= 0, "phase %r (%s) not found after the previous phase" % (name, stmt)
        pos[name] = k
        last = k
    bad = []
    for a, b in order:
        if "__syncthreads()" not in body[pos[a]:pos[b]]:
            bad.append((a, b))
    return bad


@pytest.fixture(scope="module")
def src():
    return open(SRC).read()


def test_k_sieve_every_lds_phase_pair_has_a_barrier(src):
    assert missing_barriers(kernel_body(src, "k_sieve"), SIEVE_PHASES, SIEVE_ORDER) == []


def test_k_rung_every_lds_phase_pair_has_a_barrier(src):
    assert missing_barriers(kernel_body(src, "k_rung"), RUNG_PHASES, RUNG_ORDER) == []


def test_checker_finds_the_barrier_between_count_and_stage_b_prime(src):
    body = strip_comments(kernel_body(src, "k_sieve"))
    a = body.index("const u32 n = __popc(chunk_bits(map32, c));")
    b = body.index("mark(false);")
    cut = body[:a] + body[a:b].replace("__syncthreads();", "", 1) + body[b:]
    assert ("count after covers", "stage B'") in missing_barriers(cut, SIEVE_PHASES, SIEVE_ORDER)


@pytest.mark.parametrize("pair", SIEVE_ORDER)
def test_checker_flags_each_sieve_barrier_taken_out(src, pair):
    body = strip_comments(kernel_body(src, "k_sieve"))
    stmts = dict(SIEVE_PHASES)
    a = body.index(stmts[pair[0]])
    b = body.index(stmts[pair[1]], a + 1)
    cut = body[:a] + body[a:b].replace("__syncthreads()", "") + body[b:]
    assert pair in missing_barriers(cut, SIEVE_PHASES, SIEVE_ORDER)
