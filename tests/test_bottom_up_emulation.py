"""CPU emulation of the single-source bottom-up sweep's per-lane logic
(blest_amd/csrc/device/bfs.hip, bottom-up branch of bfs_expand_kernel): groups of G
bitmap words per wave, every open vertex of a group probed at once at its probe
in-neighbour (the highest-degree one, from the per-vertex probe array the device graph
keeps, or its first list entry without one), then each lane's remaining vertices scanned
one after another (2, then 4 entries per step; from the list start after a probe, from
the second entry after a first-entry probe), a scan still open after the budget handed
to a whole-wave pass through a bounded queue (a full queue keeps the scan on its lane).  The emulation runs the same
control flow lane by lane and checks that the vertices it finds are exactly the
unvisited vertices with an in-neighbour in the frontier, and that each is found once,
for both staging depths, tiny budgets and queue capacities that overflow.

This pins the scan's logic (which entries a lane reads, when it stops, what the hand-off
resumes from); the kernel's memory ordering is argued in the design notes."""
import numpy as np
import pytest


def probe_array(n, cp, idx):
    """Each vertex's highest-degree in-neighbour, the first such in list order (the rule of
    create_device_graph in blest_amd/csrc/device/dgraph.hip); 0 for an empty list."""
    deg = np.diff(cp)
    out = np.zeros(n, dtype=np.int64)
    for v in range(n):
        nb = idx[cp[v]:cp[v + 1]]
        if len(nb):
            out[v] = nb[int(np.argmax(deg[nb]))]  # argmax: the first maximum
    return out


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
            skip = 0 if probe is not None else 1
            # first probe of every candidate
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
                p, p1 = pb[k] + skip, pe[k]
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
                        p, p1 = pb[k] + skip, pe[k]
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
@pytest.mark.parametrize("with_probe", [True, False])
def test_bottom_up_scan_finds_exactly_the_frontier_neighbours(seed, G, budget, heavy_cap, with_probe):
    n, cp, idx, visited, frontier = random_case(seed)
    probe = probe_array(n, cp, idx) if with_probe else None
    got = emulate_sweep(n, cp, idx, visited, frontier, G=G, budget=budget, heavy_cap=heavy_cap, probe=probe)
    want = {u for u in range(n) if not visited[u] and any(frontier[idx[q]] for q in range(cp[u], cp[u + 1]))}
    assert got == want


def test_probe_is_the_highest_degree_in_neighbour():
    """Hand-made lists: ties go to the first in list order, and the probe is an
    in-neighbour of its vertex."""
    # vertex 0: in-neighbours 1, 2 (degrees 1, 3); vertex 1: 0; vertex 2: 0, 3, 4; 3: 2; 4: 2
    cp = np.array([0, 2, 3, 6, 7, 8])
    idx = np.array([1, 2, 0, 0, 3, 4, 2, 2])
    assert probe_array(5, cp, idx).tolist() == [2, 0, 0, 2, 2]
    n, cp, idx, _, _ = random_case(7)
    pr = probe_array(n, cp, idx)
    deg = np.diff(cp)
    for v in range(n):
        nb = idx[cp[v]:cp[v + 1]]
        if len(nb):
            assert pr[v] in nb and deg[pr[v]] == deg[nb].max()
