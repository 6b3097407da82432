"""CPU emulation of the fused small-level kernel's per-level logic (bfs_fused_kernel in
blest_amd/csrc/device/bfs.hip), on real BVSS arrays:
  - a level's queue holds the frontier's rsets; their frontier bytes are taken, their
    64-bit frontier words cleared, and the level's discoveries set bits in the same
    bitmap;
  - a row claimed by the visited test gets the level, and is appended (as its rset) to the
    next queue exactly when the bitmap's old value showed no bit of that rset yet;
  - when a run stops (the next queue too big in rsets or slices), it hands back the next
    level's lists, written from the queue in the update kernel's form (an rset's vsets, or
    kChunkVsets-vset chunks beyond that), with the frontier left in the bitmap.
Levels run alternately here: a fused run with small caps (so that runs stop and hand back
at many depths), then one plain top-down level (the dense kernels' result: the frontier
bitmap rewritten whole), then a fused run again from that level's lists.  Checked: the
levels equal the oracle's BFS, and at every hand-back the queue is exactly the set of
rsets with frontier bits and vsets, and the lists name exactly their vsets.  The
sequential loop stands for the kernel's atomics (their order changes neither the claimed
rows nor the appended rsets)."""
import numpy as np
import pytest

CHUNK = 16  # kChunkVsets


class Bvss:
    def __init__(self, blest, g):
        b = blest.BVSS(8, 4).construct_from_csc(g)
        a = b.arrays()
        self.n = g.n
        self.rset_unit = a["real_ptrs"].astype(np.int64)
        self.ssp = a["slice_set_ptrs"].astype(np.int64)
        self.rows = a["row_ids"]
        self.mask = a["masks"].view(np.uint8)
        self.R = len(self.rset_unit) - 1

    def vsets(self, r):
        return range(int(self.rset_unit[r]), int(self.rset_unit[r + 1]))

    def slices(self, r):
        for v in self.vsets(r):
            for sl in range(int(self.ssp[v]), int(self.ssp[v + 1])):
                if self.mask[sl]:
                    yield int(self.rows[sl]), int(self.mask[sl])


def fbyte(F, r):
    return (int(F[r >> 3]) >> (8 * (r & 7))) & 0xFF


def lists_from_queue(bv, queue):
    """The hand-back lists: vset entries, then (first vset, count) chunks."""
    out, chunks = [], []
    for r in queue:
        u0, cnt = int(bv.rset_unit[r]), int(bv.rset_unit[r + 1] - bv.rset_unit[r])
        if cnt > CHUNK:
            chunks += [(u0 + j * CHUNK, min(CHUNK, cnt - j * CHUNK)) for j in range((cnt + CHUNK - 1) // CHUNK)]
        else:
            out += list(range(u0, u0 + cnt))
    return out, chunks


def fused_run(bv, F, visited, levels, queue, L, cap_rsets, cap_slices):
    """Runs levels from L; returns ("done", L) or ("handback", L, queue)."""
    first = True
    while True:
        size = sum(sum(1 for _ in bv.slices(r)) for r in queue if fbyte(F, r))
        if not first and (len(queue) > cap_rsets or size > cap_slices):
            return "handback", L, queue
        fb = {r: fbyte(F, r) for r in queue}
        for r in queue:  # the queue's words take this level's discoveries
            F[r >> 3] = np.uint64(0)
        nxt, news = [], 0
        for r in queue:
            for row, m in bv.slices(r):
                if m & fb[r] and not visited[row]:
                    visited[row] = True
                    levels[row] = L
                    news += 1
                    old = int(F[row >> 6])
                    F[row >> 6] = np.uint64(old | (1 << (row & 63)))
                    if (old >> ((row & 63) & ~7)) & 0xFF == 0:
                        nxt.append(row >> 3)
        if news == 0:
            return "done", L
        assert len(set(nxt)) == len(nxt)  # an rset is appended once
        queue, L, first = nxt, L + 1, False


def dense_level(bv, F, visited, levels, L):
    """One top-down level as the level kernels compute it: the output bitmap written whole."""
    new = np.zeros(len(F), dtype=np.uint64)
    for r in range(bv.R):
        b = fbyte(F, r)
        if not b:
            continue
        for row, m in bv.slices(r):
            if m & b and not visited[row]:
                visited[row] = True
                levels[row] = L
                new[row >> 6] |= np.uint64(1 << (row & 63))
    F[:] = new
    return int(sum(bin(int(x)).count("1") for x in new))


def frontier_rsets(bv, F):
    return {r for r in range(bv.R) if fbyte(F, r) and bv.rset_unit[r + 1] > bv.rset_unit[r]}


def traverse(bv, src, cap_rsets, cap_slices):
    n = bv.n
    W = (n + 63) // 64
    F = np.zeros(W, dtype=np.uint64)
    visited = np.zeros(n, dtype=bool)
    levels = np.full(n, -1, dtype=np.int64)
    F[src >> 6] = np.uint64(1 << (src & 63))
    visited[src] = True
    levels[src] = 0
    queue, L, handbacks = [src >> 3], 1, 0
    traverse.stale_seen = 0  # hand-backs with bits of vset-less rsets in the bitmap
    while True:
        res = fused_run(bv, F, visited, levels, queue, L, cap_rsets, cap_slices)
        if res[0] == "done":
            return levels, handbacks
        _, L, queue = res
        handbacks += 1
        if any(fbyte(F, r) and bv.rset_unit[r + 1] == bv.rset_unit[r] and r not in queue for r in range(bv.R)):
            traverse.stale_seen += 1
        # the queue is complete, and its lists name exactly the frontier rsets' vsets
        want = frontier_rsets(bv, F)
        assert {r for r in queue if bv.rset_unit[r + 1] > bv.rset_unit[r]} == want
        assert all(fbyte(F, r) for r in queue)
        out, chunks = lists_from_queue(bv, queue)
        named = set(out) | {u for u0, c in chunks for u in range(u0, u0 + c)}
        assert named == {u for r in want for u in bv.vsets(r)}
        assert len(out) + sum(c for _, c in chunks) == len(named)  # no vset twice
        # the level kernels take level L, then a fused run goes on from their lists
        if dense_level(bv, F, visited, levels, L) == 0:
            return levels, handbacks
        L += 1
        queue = sorted(frontier_rsets(bv, F))


def graphs(blest):
    rng = np.random.default_rng(11)
    n = 900
    src, dst = rng.integers(0, n, 2500), rng.integers(0, n, 2500)
    return [("rmat10", blest.CSC.rmat(10, 8, seed=5)),
            ("grid_sparse", blest.CSC.grid(30, 0.62, 0.0, 9)),
            ("er_directed", blest.CSC.from_edges(n, src, dst, undirected=False))]


@pytest.mark.parametrize("cap_rsets,cap_slices", [(1 << 30, 1 << 30), (4, 1 << 30), (64, 200), (2, 40)])
def test_fused_levels_match_the_oracle(blest, oracle, cap_rsets, cap_slices):
    total_handbacks = 0
    for name, g in graphs(blest):
        bv = Bvss(blest, g)
        cp, rows = g.arrays()
        for src in g.pick_roots(4, seed=7):
            levels, hb = traverse(bv, int(src), cap_rsets, cap_slices)
            np.testing.assert_array_equal(levels, oracle.bfs(g.n, cp, rows, int(src)), err_msg=f"{name} src={src}")
            total_handbacks += hb
    if cap_rsets < 1 << 30:
        assert total_handbacks > 0  # the small caps did stop runs


def test_stale_bits_stay_harmless(blest, oracle):
    """A frontier vertex whose rset has no vsets (none of its 8 columns has an out-edge)
    is not in the lists a fused run starts from, so its bit stays in the bitmap through
    that run's levels.  Here vertex 64 (rset 8, a word of its own) is a sink found by a
    plain level next to vertex 2, and the fused runs after it (one level each: cap 0)
    start with its bit set; every vertex still gets its oracle level and every hand-back
    queue is exactly the frontier's rsets with vsets."""
    n = 128
    path = [(0, 1), (1, 2), (1, 64)] + [(k, k + 1) for k in range(2, 40)]
    src, dst = np.array([e[0] for e in path]), np.array([e[1] for e in path])
    g = blest.CSC.from_edges(n, src, dst, undirected=False)
    bv = Bvss(blest, g)
    assert bv.rset_unit[9] == bv.rset_unit[8]  # rset 8 (vertices 64..71) has no vsets
    cp, rows = g.arrays()
    levels, hb = traverse(bv, 0, 0, 1 << 30)
    np.testing.assert_array_equal(levels, oracle.bfs(n, cp, rows, 0))
    assert hb > 3 and traverse.stale_seen > 0
