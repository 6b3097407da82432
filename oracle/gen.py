"""numpy restatement of the seeded graph generators and root picker of libblest
(blest_amd/csrc/common.hpp, csc.cpp).  TEST INFRASTRUCTURE ONLY: pins the C++
generators bit-exactly on small sizes.  Written from the generator definitions:

  mix64(z)         splitmix64 finaliser: z ^= z>>30; z *= 0xBF58476D1CE4E5B9;
                   z ^= z>>27; z *= 0x94D049BB133111EB; z ^= z>>31
  key(seed, s)     mix64(seed + G) ^ mix64(s * 0xD1B54A32D192ED03 + 0x8CB92BA72F3D8DD7)
  draw(key, i)     mix64(key + (i + 1) * G),  G = 0x9E3779B97F4A7C15
"""
from __future__ import annotations

import numpy as np

M64 = (1 << 64) - 1
G = 0x9E3779B97F4A7C15


def mix64_int(z: int) -> int:
    z &= M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def mix64(z: np.ndarray) -> np.ndarray:
    z = z.astype(np.uint64, copy=True)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def key(seed: int, stream: int) -> int:
    return mix64_int((seed + G) & M64) ^ mix64_int((stream * 0xD1B54A32D192ED03 + 0x8CB92BA72F3D8DD7) & M64)


def draw(k: int, idx: np.ndarray) -> np.ndarray:
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(k) + (idx + np.uint64(1)) * np.uint64(G)
    return mix64(z)


def draw_int(k: int, i: int) -> int:
    return mix64_int((k + (i + 1) * G) & M64)


def permutation(n: int, seed: int, stream: int) -> np.ndarray:
    p = np.arange(n, dtype=np.uint32)
    kk = key(seed, stream)
    for i in range(n - 1, 0, -1):
        j = draw_int(kk, i) % (i + 1)
        p[i], p[j] = p[j], p[i]
    return p


def symmetric_csc(n: int, u: np.ndarray, v: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """Both directions of every pair, self-loops dropped, duplicates merged, columns sorted."""
    u = np.asarray(u, dtype=np.int64)
    v = np.asarray(v, dtype=np.int64)
    keep = u != v
    u, v = u[keep], v[keep]
    cols = np.concatenate([u, v])
    rws = np.concatenate([v, u])
    code = np.unique(cols * n + rws)
    cols, rws = code // n, code % n
    cp = np.zeros(n + 1, dtype=np.int64)
    np.add.at(cp, cols + 1, 1)
    return np.cumsum(cp).astype(np.uint32), rws.astype(np.uint32)


def rmat_samples(scale: int, edge_factor: int, a: float, b: float, c: float, seed: int):
    """The raw R-MAT edge samples (u, v) before relabelling and symmetrisation, and the
    per-level quadrant codes (bit 1 = u's bit, bit 0 = v's bit; level 0 = top bit)."""
    m = edge_factor << scale
    kk = key(seed, 1)
    pairs = (scale + 1) // 2
    ks = np.arange(m, dtype=np.uint64)
    su = np.zeros(m, dtype=np.int64)
    sv = np.zeros(m, dtype=np.int64)
    ab, abc = a + b, a + b + c
    h = None
    quads = []
    for lvl in range(scale):
        if lvl % 2 == 0:
            h = draw(kk, ks * np.uint64(pairs) + np.uint64(lvl // 2))
            half = (h >> np.uint64(32)).astype(np.float64)
        else:
            half = (h & np.uint64(0xFFFFFFFF)).astype(np.float64)
        r = half * (1.0 / 4294967296.0)
        bu = (r >= ab).astype(np.int64)                       # quadrants c, d
        bv = (((r >= a) & (r < ab)) | (r >= abc)).astype(np.int64)  # quadrants b, d
        su = (su << 1) | bu
        sv = (sv << 1) | bv
        quads.append((bu << 1 | bv).astype(np.uint8))
    return su, sv, quads


def rmat(scale: int, edge_factor: int, a: float, b: float, c: float, seed: int, permute: bool):
    n = 1 << scale
    su, sv, _ = rmat_samples(scale, edge_factor, a, b, c, seed)
    if permute:
        p = permutation(n, seed, 2).astype(np.int64)
        su, sv = p[su], p[sv]
    return n, *symmetric_csc(n, su, sv)


def grid(side: int, keep: float, shortcut_frac: float, seed: int):
    n = side * side
    keep32 = (1 << 32) if keep >= 1.0 else int(keep * 4294967296.0)
    ids = np.arange(n, dtype=np.uint64)
    kr, kd, ks = key(seed, 3), key(seed, 4), key(seed, 5)
    col = ids % np.uint64(side)
    row = ids // np.uint64(side)
    right = (col + np.uint64(1) < np.uint64(side)) & ((draw(kr, ids) >> np.uint64(32)) < np.uint64(min(keep32, M64)))
    down = (row + np.uint64(1) < np.uint64(side)) & ((draw(kd, ids) >> np.uint64(32)) < np.uint64(min(keep32, M64)))
    u = [ids[right], ids[down]]
    v = [ids[right] + np.uint64(1), ids[down] + np.uint64(side)]
    cnt = int(round(shortcut_frac * n))
    if cnt:
        s = np.arange(cnt, dtype=np.uint64)
        su = draw(ks, np.uint64(2) * s) % np.uint64(n)
        sv = draw(ks, np.uint64(2) * s + np.uint64(1)) % np.uint64(n)
        u.append(su)
        v.append(sv)
    return n, *symmetric_csc(n, np.concatenate(u).astype(np.int64), np.concatenate(v).astype(np.int64))


def erdos_renyi(n: int, avg_degree: float, seed: int):
    m = int(round(n * avg_degree / 2.0))
    kk = key(seed, 6)
    s = np.arange(m, dtype=np.uint64)
    u = draw(kk, np.uint64(2) * s) % np.uint64(n)
    v = draw(kk, np.uint64(2) * s + np.uint64(1)) % np.uint64(n)
    return n, *symmetric_csc(n, u.astype(np.int64), v.astype(np.int64))


def pick_roots(n: int, col_ptrs: np.ndarray, count: int, seed: int) -> np.ndarray:
    deg = np.diff(col_ptrs.astype(np.int64))
    kk = key(seed, 7)
    out, taken, t = [], set(), 0
    while len(out) < count:
        v = draw_int(kk, t) % n
        t += 1
        if deg[v] == 0 or v in taken:
            continue
        taken.add(v)
        out.append(v)
    return np.array(out, dtype=np.uint32)
