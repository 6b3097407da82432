"""CPU emulation of the single-source bottom-up sweep's per-lane logic
(blest_amd/csrc/device/bfs.hip, bottom-up branch of bfs_expand_kernel): groups of G
bitmap words per wave, every open vertex of a group probed at its first in-neighbour at
once (read from the per-vertex probe array when the lists are sorted by degree, as the
device graph keeps them by default; the same entry), then each lane's remaining vertices
scanned one after another (2, then 4 entries per step), a scan still open after the
budget handed to a whole-wave pass through a bounded queue (a full queue keeps the scan
on its lane).  The emulation runs the same
control flow lane by lane and checks that the vertices it finds are exactly the
unvisited vertices with an in-neighbour in the frontier, and that each is found once,
for both staging depths, tiny budgets and queue capacities that overflow.

This pins the scan's logic (which entries a lane reads, when it stops, what the hand-off
resumes from); the kernel's memory ordering is argued in the design notes."""
import numpy as np
import pytest


def degree_sorted(n, cp, idx):
    """Each vertex's in-neighbour list by degree, highest first, ties by ID (the order
    create_device_graph in blest_amd/csrc/device/dgraph.hip uploads), and the probe array
    (each list's first entry, 0 for an empty list)."""
    deg = np.diff(cp)
    out = idx.copy()
    probe = np.zeros(n, dtype=np.int64)
    for v in range(n):
        nb = idx[cp[v]:cp[v + 1]]
        nb = nb[np.lexsort((nb, -deg[nb]))]
        out[cp[v]:cp[v + 1]] = nb
        if len(nb):
            probe[v] = nb[0]
    return out, probe


def emulate_sweep(n, cp, idx, visited, frontier, G=8, budget=32, batch=4, heavy_cap=1024, words_per_wave=None,
                  probe=None):
    W = (n + 63) // 64
    found, heavy = set(), []
    words = list(range(W))
    # one "wave" takes words w0, w0 + wpb, ...; with wpb = 1 a group is G consecutive words
    for g0 in range(0, W, G):
        group = [w for w in range(g0, min(W, g0 + G))]
        for lane in range(64):
            pb, pe, cand = [], [], 0
            for k, w in enumerate(group):
                u = w * 64 + lane
                open_ = u < n and not visited[u]
                pb.append(cp[u] if open_ else 0)
                pe.append(cp[u + 1] if open_ else 0)
                if pe[k] > pb[k]:
                    cand |= 1 << k
            fk = 0
            # first probe of every candidate (the probe array holds the first entries)
            for k in range(len(group)):
                if (cand >> k) & 1:
                    u = group[k] * 64 + lane
                    hit0 = frontier[probe[u]] if probe is not None else frontier[idx[pb[k]]]
                    if hit0:
                        fk |= 1 << k
                    if hit0 or pb[k] + 1 >= pe[k]:
                        cand &= ~(1 << k)
            k = (cand & -cand).bit_length() - 1 if cand else len(group)
            if k < len(group):
                p, p1 = pb[k] + 1, pe[k]
                bud = p + budget
            step = 2
            while k < len(group):
                lim = min(p + step, p1)
                hit = any(frontier[idx[q]] for q in range(p, lim))
                p = lim
                step = min(2 * step, batch)
                nxt = hit or p >= p1
                if not nxt and p >= bud:
                    if len(heavy) < heavy_cap:
                        heavy.append((group[k] * 64 + lane, p))
                        nxt = True
                    else:
                        bud = p1
                if nxt:
                    if hit:
                        fk |= 1 << k
                    cand &= ~(1 << k)
                    k = (cand & -cand).bit_length() - 1 if cand else len(group)
                    if k < len(group):
                        p, p1 = pb[k] + 1, pe[k]
                        bud = p + budget
                    step = 2
            for k, w in enumerate(group):
                if (fk >> k) & 1:
                    u = w * 64 + lane
                    assert u not in found
                    found.add(u)
        # whole-wave pass over the queue (per workgroup: here per group)
        for u, p0 in heavy:
            if any(frontier[idx[q]] for q in range(p0, cp[u + 1])):
                assert u not in found
                found.add(u)
        heavy = []
    return found


def random_case(seed, n=3000, m=20000, p_vis=0.4, p_front=0.05):
    rng = np.random.default_rng(seed)
    e = rng.integers(0, n, size=(m, 2))
    # a few hubs so that scans run long and overflow the budget
    hubs = rng.integers(0, n, size=5)
    e = np.concatenate([e, np.stack([np.repeat(hubs, 300), rng.integers(0, n, size=1500)], axis=1)])
    e = e[e[:, 0] != e[:, 1]]
    src = np.concatenate([e[:, 0], e[:, 1]])
    dst = np.concatenate([e[:, 1], e[:, 0]])
    order = np.lexsort((src, dst))
    rows, cols = dst[order], src[order]
    keep = np.ones(len(rows), bool)
    keep[1:] = (rows[1:] != rows[:-1]) | (cols[1:] != cols[:-1])
    rows, cols = rows[keep], cols[keep]
    cp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(cp, rows + 1, 1)
    cp = np.cumsum(cp)
    visited = rng.random(n) < p_vis
    frontier = visited & (rng.random(n) < p_front / p_vis)  # the frontier is part of the visited set
    return n, cp, cols, visited, frontier


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("G,budget,heavy_cap", [(8, 32, 1024), (4, 32, 1024), (8, 2, 3), (4, 1, 0)])
@pytest.mark.parametrize("sorted_lists", [True, False])
def test_bottom_up_scan_finds_exactly_the_frontier_neighbours(seed, G, budget, heavy_cap, sorted_lists):
    n, cp, idx, visited, frontier = random_case(seed)
    probe = None
    if sorted_lists:
        idx, probe = degree_sorted(n, cp, idx)
    got = emulate_sweep(n, cp, idx, visited, frontier, G=G, budget=budget, heavy_cap=heavy_cap, probe=probe)
    want = {u for u in range(n) if not visited[u] and any(frontier[idx[q]] for q in range(cp[u], cp[u + 1]))}
    assert got == want


def test_degree_sorted_lists():
    """Hand-made lists: degree descending, ties by ID; every list keeps its entries and the
    probe is the highest-degree in-neighbour."""
    # vertex 0: in-neighbours 1, 2 (degrees 1, 3); vertex 1: 0; vertex 2: 4, 3, 0 (1, 1, 2)
    cp = np.array([0, 2, 3, 6, 7, 8])
    idx = np.array([1, 2, 0, 4, 3, 0, 2, 2])
    srt, pr = degree_sorted(5, cp, idx)
    assert srt.tolist() == [2, 1, 0, 0, 3, 4, 2, 2]
    assert pr.tolist() == [2, 0, 0, 2, 2]
    n, cp, idx, _, _ = random_case(7)
    srt, pr = degree_sorted(n, cp, idx)
    deg = np.diff(cp)
    for v in range(n):
        a, b = idx[cp[v]:cp[v + 1]], srt[cp[v]:cp[v + 1]]
        assert sorted(a.tolist()) == sorted(b.tolist())
        assert (np.diff(deg[b]) <= 0).all()
        if len(b):
            assert pr[v] == b[0] and deg[pr[v]] == deg[a].max()
