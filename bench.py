#!/usr/bin/env python3
"""Headline benchmark: single-source direction-optimising BFS, GTEPS (harmonic mean over
the workload's roots), on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload rmat22_gorder]

A step is one BFS from every root of the workload (the graph already resident in HBM).
Per-root GTEPS = m_c / t, with m_c = half the degree sum of the reached vertices
(undirected edges of the root's component) and t the bfs_kernel launch measured with
HIP events on the query's stream; the reported value is the harmonic mean over all
(step, root) launches.  Before timing, the native autotuner (blest_autotune_bfs) picks the
launch shape and direction thresholds on --autotune-roots of the roots, as a deployment
tunes once per resident graph (--no-autotune keeps the device defaults).  The line also
carries:
  roofline      algorithmic bytes per launch (the reference formulation's B_td / B_bu,
                computed by reference_level_bytes from a sample of launches' own GPU
                levels and direction schedules) / those launches' mean time, against
                8 TB/s;
  cpu_baseline  the oracle's level-synchronous BVSS BFS (the reference algorithm,
                its own direction rule) on all host cores over a bounded root sample;
  cc            connected components of the same resident graph;
  mbfs          batched multi-source BFS (128 roots per launch, 1,024 seeded roots) on
                the same resident graph: aggregate GTEPS, the plain single-source GTEPS
                of the same roots (at the tuned single-source thresholds) and the ratio
                of the two; the batches switch direction at their own thresholds
                (--mbfs-alpha / --mbfs-beta).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ensure_built():
    if not (ROOT / "blest_amd" / "libblest.so").exists():
        import subprocess
        subprocess.run([sys.executable, "-m", "blest_amd.build"], cwd=ROOT, check=True)


def info_sym(dg) -> bool:
    return bool(dg.info()["symmetric"])


def harmonic(values):
    v = np.asarray(values, dtype=np.float64)
    return float(len(v) / np.sum(1.0 / v))


SERVE_RATES = (2000, 5000, 10000, 20000, 50000)


def reference_level_bytes(n, col_ptrs, rows, bv, levels, bu_mask, entry_row=None):
    """Algorithmic bytes per level of one single-source BFS in the reference formulation
    (BFS/BVSSBFSKernel.cuh:96-276; the fixed byte model of the measurement plan), from the
    traversal's own output: levels (reordered IDs, -1 unreached) and the bottom-up mask
    the launch ran.  W = 32-bit bitmap words, s_v = slices of vset v in the BVSS layout,
    e_u = in-neighbours scanned in sorted order up to the first frontier hit (deg if none):
      top-down  B = sum_{v in VF} (5 s_v + 25) + 12 W + 4 |F'| + 4 |VF'|
      bottom-up B = 4 W + n/8 + sum_{u unvisited} (8 + 4 e_u) + 12 W + 4 |F'|
    Levels 1 .. max+1 (the last one finds nothing, as the traversal's last launch).
    entry_row (optional, reused across roots): the CSR row of every entry."""
    lv = np.asarray(levels, dtype=np.int32)
    cp = np.asarray(col_ptrs, dtype=np.int64)
    W = (n + 31) // 32
    rp = bv["real_ptrs"].astype(np.int64)
    ssp = bv["slice_set_ptrs"].astype(np.int64)
    R = len(rp) - 1
    rvs = np.diff(rp)                          # vsets per rset
    rsl = ssp[rp[1:]] - ssp[rp[:-1]]           # slices per rset
    top = int(lv.max()) if (lv >= 0).any() else 0
    reached = lv >= 0
    # per level: distinct rsets holding a vertex of that level
    keys = np.unique(lv[reached].astype(np.int64) * R + (np.nonzero(reached)[0] >> 3))
    kl, kr = keys // R, keys % R
    lvs_vsets = np.bincount(kl, weights=rvs[kr], minlength=top + 2)
    lvs_slices = np.bincount(kl, weights=rsl[kr], minlength=top + 2)
    new_count = np.bincount(lv[reached], minlength=top + 2)
    out = []
    for L in range(1, top + 2):
        bu = L <= 64 and (bu_mask >> (L - 1)) & 1
        newv = int(new_count[L]) if L <= top else 0
        cnt2 = float(lvs_vsets[L]) if L <= top else 0.0
        if not bu:
            b = 5 * lvs_slices[L - 1] + 25 * lvs_vsets[L - 1] + 12 * W + 4 * newv + 4 * cnt2
        else:
            if entry_row is None:
                entry_row = np.repeat(np.arange(n, dtype=np.int32), np.diff(cp))
            open_ = (lv < 0) | (lv >= L)
            unv = int(open_.sum())
            hits = np.flatnonzero(open_[entry_row] & (lv[rows] == L - 1))
            hr = entry_row[hits]
            first = np.ones(len(hr), dtype=bool)
            first[1:] = hr[1:] != hr[:-1]
            # an open row with a hit scans up to its first hit; every other open row its whole list
            scanned = int(np.diff(cp)[open_].sum()) - int((cp[hr[first] + 1] - hits[first] - 1).sum())
            b = 4 * W + (n + 7) // 8 + 8 * unv + 4 * scanned + 12 * W + 4 * newv
        out.append(int(b))
    return out


def serve_bench(a) -> dict:
    """Served traversal: resident R-MAT graphs (scales 18-25, edge factor 16) and two
    perturbed grids (2048^2, 4096^2; keep 0.9, 1% shortcuts), natural order.  An open-loop
    Poisson stream (80% single-root BFS, 20% CC; graph, root and arrival times seeded)
    is submitted to the request server at each rate; results land in device buffers.
    Latency = (submit - scheduled arrival) + queued + service, from the server's own
    timestamps.  The same requests run one at a time through blest_bfs / blest_cc give
    the comparison: their measured service times are replayed as a FIFO single server
    under the same arrivals."""
    import ctypes as C
    import queue
    import threading

    import torch

    import blest_amd as B
    sys.setswitchinterval(1e-4)  # the submitter spins; let the waiter thread run promptly
    lib = B._lib.load()
    scales = [int(x) for x in a.serve_scales.split(",")] if a.serve_scales else list(range(18, 26))
    sides = [int(x) for x in a.serve_grids.split(",")] if a.serve_grids else [2048, 4096]
    graphs = []  # (label, DeviceGraph, n, roots)
    t0 = time.time()
    for s in scales:
        g = B.CSC.rmat(s, 16, 0.57, 0.19, 0.19, seed=1000 + s, permute=True)
        b = B.BVSS(8, 4).construct_from_csc(g, full_padding=not g.is_social_network())
        graphs.append((f"rmat{s}", B.DeviceGraph(b, g, None), g.n, g.pick_roots(256, seed=s)))
        log(f"[serve] rmat{s}: n={g.n} nnz={g.nnz} ({time.time() - t0:.0f}s)")
    for side in sides:
        g = B.CSC.grid(side, 0.9, 0.01, seed=2000 + side)
        b = B.BVSS(8, 4).construct_from_csc(g, full_padding=not g.is_social_network())
        graphs.append((f"grid{side}", B.DeviceGraph(b, g, None), g.n, g.pick_roots(256, seed=side)))
        log(f"[serve] grid{side}: n={g.n} nnz={g.nnz} ({time.time() - t0:.0f}s)")
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(a.serve_seed)
    N = a.serve_requests
    gsel = rng.integers(len(graphs), size=N)
    kinds = np.where(rng.random(N) < 0.8, 0, 1)
    roots = np.array([int(graphs[gi][3][rng.integers(len(graphs[gi][3]))]) for gi in gsel], dtype=np.int64)
    gaps = rng.exponential(1.0, size=N)  # unit-rate Poisson gaps, scaled per rate
    per_graph_bufs = max(2, a.serve_buffers)
    bufs = [[torch.empty(n, dtype=torch.int32, device=dev) for _ in range(per_graph_bufs)]
            for _, _, n, _ in graphs]

    # one-at-a-time reference: plain kernels, one query per graph, same requests
    qs = [B._Query(dg, 1) for _, dg, _, _ in graphs]
    bst, cst = B._lib.BfsStats(), B._lib.CcStats()
    def plain(i):
        gi, k, r = int(gsel[i]), int(kinds[i]), int(roots[i])
        out = C.c_void_p(bufs[gi][0].data_ptr())
        if k == 0:
            B.check(lib.blest_bfs(qs[gi]._h, r, C.cast(out, B._lib.I32P), C.byref(bst)))
            return bst.edges_traversed_component
        B.check(lib.blest_cc(qs[gi]._h, 0, C.cast(out, B._lib.U32P), C.byref(cst)))
        return 0
    for i in range(min(N, 50)):
        plain(i)
    seq_service = np.empty(N)
    mc = np.zeros(N)
    for i in range(N):
        t1 = time.perf_counter()
        mc[i] = plain(i)
        seq_service[i] = time.perf_counter() - t1
    del qs
    log(f"[serve] one-at-a-time: {N} requests in {seq_service.sum():.2f}s")
    for gi, (label, _, _, _) in enumerate(graphs):
        for k, kname in ((0, "bfs"), (1, "cc")):
            sel = (gsel == gi) & (kinds == k)
            if sel.any():
                log(f"[serve]   {label} {kname}: {int(sel.sum())} requests, mean {seq_service[sel].mean() * 1e3:.3f} ms, "
                    f"max {seq_service[sel].max() * 1e3:.3f} ms")

    srv = B.Server([dg for _, dg, _, _ in graphs], workers=a.serve_workers, max_batch=a.serve_batch,
                   capacity=per_graph_bufs * len(graphs), min_multi=a.serve_min_multi)
    free = [queue.SimpleQueue() for _ in graphs]
    for gi, bl in enumerate(bufs):
        for j in range(len(bl)):
            free[gi].put(j)

    def run_rate(rate, count):
        arrivals = np.cumsum(gaps[:count]) / rate
        lat = np.empty(count)
        done_at = np.empty(count)
        batch = np.empty(count)
        pending = queue.SimpleQueue()

        def waiter():
            res = B._lib.RequestResult()
            for _ in range(count):
                i, t, gi, j, tsub = pending.get()
                B.check(lib.blest_server_wait(srv._h, t, 0xFFFFFFFF, C.byref(res)))
                free[gi].put(j)
                svc = (res.queued_ms + res.service_ms) * 1e-3
                lat[i] = (tsub - arrivals[i]) + svc
                done_at[i] = tsub + svc
                batch[i] = res.batch_size

        th = threading.Thread(target=waiter)
        th.start()
        t = C.c_uint64()
        start = time.perf_counter()
        for i in range(count):
            target = start + arrivals[i]
            while True:
                now = time.perf_counter()
                if now >= target:
                    break
                if target - now > 2e-3:
                    time.sleep(target - now - 1e-3)
            gi = int(gsel[i])
            j = free[gi].get()
            tsub = time.perf_counter() - start
            B.check(lib.blest_server_submit(srv._h, gi, int(kinds[i]), int(roots[i]),
                                            C.c_void_p(bufs[gi][j].data_ptr()), C.byref(t)))
            pending.put((i, t.value, gi, j, tsub))
        th.join()
        return lat, float(done_at.max()), batch

    run_rate(SERVE_RATES[0], min(N, 200))  # warm the server's query pool
    rows = []
    for rate in (a.serve_rates or SERVE_RATES):
        lat, wall, batch = run_rate(rate, N)
        # FIFO single server with the measured one-at-a-time service times
        arr = np.cumsum(gaps[:N]) / rate
        fin = np.empty(N)
        tcur = 0.0
        for i in range(N):
            tcur = max(tcur, arr[i]) + seq_service[i]
            fin[i] = tcur
        seq_lat = fin - arr
        row = {"rate": rate, "requests": N,
               "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
               "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
               "throughput_rps": round(N / wall, 1), "bfs_gteps": round(float(mc.sum()) / wall / 1e9, 3),
               "mean_batch": round(float(batch.mean()), 2),
               "one_at_a_time": {"p50_ms": round(float(np.percentile(seq_lat, 50)) * 1e3, 3),
                                 "p99_ms": round(float(np.percentile(seq_lat, 99)) * 1e3, 3),
                                 "throughput_rps": round(N / float(fin[-1]), 1),
                                 "bfs_gteps": round(float(mc.sum()) / float(fin[-1]) / 1e9, 3)}}
        rows.append(row)
        log(f"[serve] {json.dumps(row)}")
    del srv
    head = next((r for r in rows if r["rate"] == 10000), rows[-1])
    return {"metric": f"served p99 latency at {head['rate']} req/s (80% BFS / 20% CC)", "value": head["p99_ms"],
            "unit": "ms", "n_gpus": 1, "steps": len(rows), "warmup": 1, "ms_per_step": None,
            "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic",
            "config": {"workload": "serve", "graphs": [x[0] for x in graphs], "reorder": "natural",
                       "mix": "80% single-root BFS, 20% CC", "requests_per_rate": N,
                       "workers": a.serve_workers, "max_batch": a.serve_batch, "min_multi": a.serve_min_multi,
                       "results": "device buffers"},
            "rates": rows, "roofline": None, "cpu_baseline": None}


def mbfs_bench(a, prep, opts, dg=None, algos=(0,)) -> dict:
    """Batched multi-source BFS against plain single-source runs on the same roots:
    aggregate GTEPS = sum of m_c / sum of kernel time; speedup = batched / plain.
    dg: the resident device graph to use (uploaded with opts when None).  algos: the
    multi-source products timed (0 bit-parallel VALU, 1 i8 MFMA); the first is the
    headline, the others are reported beside it."""
    import ctypes as C

    import blest_amd as B
    lib = B._lib.load()
    if dg is None:
        dg = B.DeviceGraph(prep.bvss, prep.csc, prep.inv_perm, **opts)
    roots = [int(r) for r in prep.roots]
    if a.mbfs_roots > len(roots):
        # more seeded roots of degree >= 1, picked on the resident (reordered) graph and
        # mapped back to original IDs
        new_to_old = np.argsort(prep.inv_perm)
        roots = [int(new_to_old[r]) for r in prep.csc.pick_roots(a.mbfs_roots, seed=prep.workload.root_seed)]
    width = a.mbfs_width
    q1 = B._Query(dg, 1)
    st = B._lib.BfsStats()
    for r in roots[:4]:
        B.check(lib.blest_bfs(q1._h, r, None, C.byref(st)))
    plain_ms, mc = [], []
    for r in roots:
        B.check(lib.blest_bfs(q1._h, r, None, C.byref(st)))
        plain_ms.append(st.kernel_ms)
        mc.append(st.edges_traversed_component)
    del q1
    # The plain leg ran at the graph's (tuned) single-source thresholds; the batched leg
    # switches on (vertex, root) pairs with its own thresholds (--mbfs-alpha/--mbfs-beta).
    dg.set_thresholds(a.mbfs_alpha, a.mbfs_beta, int(opts.get("direction", 0)))
    qm = B._Query(dg, width)
    batches = [np.ascontiguousarray(np.array(roots[i:i + width], dtype=np.uint32)) for i in range(0, len(roots), width)]
    total_mc = float(sum(mc))
    plain = total_mc / (sum(plain_ms) * 1e-3) / 1e9
    per_algo = {}
    for algo in algos:
        o = B._lib.MbfsOpts(algo, width)
        B.check(lib.blest_mbfs(qm._h, B._u32p(batches[0]), len(batches[0]), None, C.byref(o), C.byref(st)))  # warm-up
        ms = []
        for _ in range(a.steps):
            for b in batches:
                B.check(lib.blest_mbfs(qm._h, B._u32p(b), len(b), None, C.byref(o), C.byref(st)))
                ms.append(st.kernel_ms)
        per_algo[algo] = ms
        gt = total_mc * a.steps / (sum(ms) * 1e-3) / 1e9
        log(f"[mbfs] algo {algo}: plain {plain:.2f} GTEPS ({np.mean(plain_ms):.3f} ms/root), batched {gt:.2f} GTEPS "
            f"({np.mean(ms):.3f} ms per {width}-root batch): {gt / plain:.2f}x")
    del qm
    dg.set_thresholds(float(opts.get("alpha", 15.0)), float(opts.get("beta", 18.0)), int(opts.get("direction", 0)))
    batch_ms = per_algo[algos[0]]
    batched = total_mc * a.steps / (sum(batch_ms) * 1e-3) / 1e9
    others = {str(k): {"gteps": round(total_mc * a.steps / (sum(v) * 1e-3) / 1e9, 3),
                       "batch_ms": round(float(np.mean(v)), 4),
                       "speedup_vs_plain": round(total_mc * a.steps / (sum(v) * 1e-3) / 1e9 / plain, 2)}
              for k, v in per_algo.items()}
    info = dg.info()
    return {"metric": f"batched MBFS aggregate GTEPS ({width} roots per launch)", "value": round(batched, 3),
            "unit": "GTEPS", "n_gpus": 1, "steps": a.steps, "warmup": 1,
            "ms_per_step": round(float(sum(batch_ms)) / a.steps, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "u64", "data": "synthetic",
            "config": {"workload": prep.workload.name + "_mbfs", "vertices": info["n"], "directed_edges": info["nnz"],
                       "roots": len(roots), "batch": width, "reorder": prep.workload.reorder.name.lower(),
                       "alpha": a.mbfs_alpha, "beta": a.mbfs_beta},
            "plain_gteps": round(plain, 3), "speedup_vs_plain": round(batched / plain, 2),
            "plain_ms_per_root": round(float(np.mean(plain_ms)), 4),
            "batch_ms": round(float(np.mean(batch_ms)), 4), "by_algo": others, "roofline": None,
            "cpu_baseline": None}


def batch_bench(a) -> dict:
    """Many small graphs in one launch: 4,096 Erdos-Renyi graphs, n ~ U[1024, 65536]
    (seeded), average degree 8, one seeded root each; aggregate GTEPS = sum of m_c over
    graphs / launch time.  Plain = the same graphs one at a time through blest_bfs
    (a seeded sample of graphs, scaled to the batch)."""
    import ctypes as C

    import blest_amd as B
    lib = B._lib.load()
    from blest_amd import workloads as W
    count = a.batch_graphs
    t0 = time.time()
    graphs, roots = W.er_batch(count, a.batch_seed)
    sizes = np.array([g.n for g in graphs], dtype=np.int64)
    rng = np.random.default_rng(a.batch_seed + 1)
    log(f"[batch] {count} graphs, {int(sizes.sum())} vertices generated in {time.time() - t0:.0f}s")
    t0 = time.time()
    gb = B.GraphBatch(graphs)
    log(f"[batch] uploaded in {time.time() - t0:.0f}s")
    levels, _ = gb.bfs(roots)  # warm-up; also gives m_c
    mc = 0
    for g, lv in zip(graphs, levels):
        cp, _ = g.arrays()
        deg = np.diff(cp.astype(np.int64))
        mc += int(deg[lv >= 0].sum()) // 2
    ms = C.c_double()
    times = []
    for _ in range(a.steps):
        B.check(lib.blest_graph_batch_bfs(gb._h, B._u32p(roots), None, C.byref(ms)))
        times.append(ms.value)
    t_batch = float(np.mean(times))
    batched = mc / (t_batch * 1e-3) / 1e9
    # plain: a seeded sample of the graphs one at a time
    sample = rng.choice(count, size=min(count, a.batch_plain_sample), replace=False)
    plain_ms, plain_mc = 0.0, 0
    st = B._lib.BfsStats()
    for i in sample:
        g = graphs[int(i)]
        bv = B.BVSS(8, 4).construct_from_csc(g)
        dg = B.DeviceGraph(bv, g, None)
        qq = B._Query(dg, 1)
        B.check(lib.blest_bfs(qq._h, int(roots[i]), None, C.byref(st)))
        B.check(lib.blest_bfs(qq._h, int(roots[i]), None, C.byref(st)))
        plain_ms += st.kernel_ms
        plain_mc += st.edges_traversed_component
        del qq, dg
    plain = plain_mc / (plain_ms * 1e-3) / 1e9
    log(f"[batch] one launch {t_batch:.3f} ms: {batched:.2f} GTEPS; plain one at a time {plain:.3f} GTEPS "
        f"({len(sample)} graphs sampled)")
    return {"metric": f"graph-batch aggregate GTEPS ({count} graphs, one launch)", "value": round(batched, 3),
            "unit": "GTEPS", "n_gpus": 1, "steps": a.steps, "warmup": 1, "ms_per_step": round(t_batch, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic",
            "config": {"workload": "er_batch", "graphs": count, "n_range": [1024, 65536], "avg_degree": 8,
                       "vertices": int(sizes.sum()), "seed": a.batch_seed},
            "plain_gteps": round(plain, 3), "speedup_vs_plain": round(batched / plain, 2),
            "plain_sample": len(sample), "roofline": None, "cpu_baseline": None}


def copy_calibration(a) -> dict:
    """A plain device copy of known size (torch's copy kernel), run under the same HBM
    byte counters as the traversal to scale them: counter bytes / (2 x copy bytes)."""
    import torch
    nbytes = a.copy_mib << 20
    x = torch.ones(nbytes // 4, dtype=torch.int32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(a.steps):
        y.copy_(x)
    torch.cuda.synchronize()
    return {"metric": "device copy", "bytes_read_per_copy": nbytes, "bytes_written_per_copy": nbytes,
            "copies": a.steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rmat22_gorder")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline leg")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-roofline", action="store_true")
    ap.add_argument("--no-cc", action="store_true", help="skip the connected-components timing")
    ap.add_argument("--roofline-roots", type=int, default=-1,
                    help="roots whose GPU levels feed the byte model (-1: 8; 0: all)")
    ap.add_argument("--block-size", type=int, default=0)
    ap.add_argument("--blocks-per-cu", type=int, default=0)
    ap.add_argument("--alpha", type=float, default=0.0)
    ap.add_argument("--beta", type=float, default=0.0)
    ap.add_argument("--direction", type=int, default=-1, help="0 optimising, 1 top-down only, 2 bottom-up")
    ap.add_argument("--td-quads", type=int, default=0, help="top-down slice tile: quads per lane per step (1, 2, 4)")
    ap.add_argument("--bu-group", type=int, default=0, help="bottom-up staging depth: words probed together (4, 8)")
    ap.add_argument("--bu-probe", type=int, default=0,
                    help="bottom-up first probe: highest-degree in-neighbour (0 default) or list order (< 0)")
    ap.add_argument("--fuse-quads", type=int, default=0,
                    help="quads a small level may stream fused in one workgroup (0 default, < 0 never)")
    ap.add_argument("--autotune", action=argparse.BooleanOptionalAction, default=True,
                    help="pick blocks_per_cu / alpha / beta with the native autotuner on --autotune-roots "
                         "of the workload's roots before timing (default on; --no-autotune keeps the "
                         "device defaults)")
    ap.add_argument("--autotune-roots", type=int, default=32)
    ap.add_argument("--use-graphs", type=int, default=-1,
                    help="replay captured level chunks: 0 after a deep traversal (default), 1 always, 2 never")
    ap.add_argument("--serve-scales", default="", help="serve: comma-separated R-MAT scales (default 18..25)")
    ap.add_argument("--serve-grids", default="", help="serve: comma-separated grid sides (default 2048,4096)")
    ap.add_argument("--serve-requests", type=int, default=2000)
    ap.add_argument("--serve-rates", type=lambda s: [int(x) for x in s.split(",")], default=None)
    ap.add_argument("--serve-workers", type=int, default=4)
    ap.add_argument("--serve-batch", type=int, default=64)
    ap.add_argument("--serve-min-multi", type=int, default=8)
    ap.add_argument("--serve-buffers", type=int, default=48, help="serve: device result buffers per graph")
    ap.add_argument("--serve-seed", type=int, default=5)
    ap.add_argument("--mbfs", action="store_true", help="batched multi-source BFS vs plain on the workload's roots")
    ap.add_argument("--no-mbfs", action="store_true", help="skip the multi-source summary of the default line")
    ap.add_argument("--mbfs-width", type=int, default=128)
    ap.add_argument("--mbfs-algo", type=lambda x: [int(v) for v in x.split(",")], default=[0],
                    help="multi-source products to time: 0 bit-parallel VALU, 1 i8 MFMA (comma list)")
    ap.add_argument("--mbfs-roots", type=int, default=1024)
    ap.add_argument("--mbfs-alpha", type=float, default=15.0, help="multi-source switch threshold alpha")
    ap.add_argument("--mbfs-beta", type=float, default=18.0, help="multi-source switch threshold beta")
    ap.add_argument("--batch-graphs", type=int, default=4096)
    ap.add_argument("--batch-seed", type=int, default=3)
    ap.add_argument("--batch-plain-sample", type=int, default=128)
    ap.add_argument("--copy-mib", type=int, default=1024, help="copy: bytes of the calibration copy")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the multi-GPU sharding is a "
              f"design (rsets partitioned across GPUs, one bitmap OR-exchange per level), not built.",
              file=sys.stderr)
        sys.exit(2)
    ensure_built()
    if a.workload == "serve":
        print("bench.py --workload serve is disabled: earlier runs of this serving benchmark ended in a GPU "
              "fault or queue hang in the multi-threaded server phase, and the cause has not been shown; "
              "it stays off the GPU until it is.", file=sys.stderr)
        sys.exit(3)
    if a.workload == "copy":
        print(json.dumps(copy_calibration(a)), flush=True)
        return
    if a.workload == "er_batch":
        print(json.dumps(batch_bench(a)), flush=True)
        return
    import blest_amd as B
    from blest_amd import workloads as W

    prep = W.prepare(a.workload, log=log, device_bvss=True)
    w = prep.workload
    opts = {}
    if a.block_size:
        opts["block_size"] = a.block_size
    if a.blocks_per_cu:
        opts["blocks_per_cu"] = a.blocks_per_cu
    if a.alpha:
        opts["alpha"] = a.alpha
    if a.beta:
        opts["beta"] = a.beta
    if a.direction >= 0:
        opts["direction"] = a.direction
    if a.use_graphs >= 0:
        opts["use_graphs"] = a.use_graphs
    if a.fuse_quads:
        opts["fuse_quads"] = a.fuse_quads
    if a.bu_probe:
        opts["bu_probe"] = a.bu_probe
    if a.mbfs:
        print(json.dumps(mbfs_bench(a, prep, opts, algos=tuple(a.mbfs_algo))), flush=True)
        return
    tuned = None
    if a.td_quads:
        opts["td_quads"] = a.td_quads
    if a.bu_group:
        opts["bu_group"] = a.bu_group
    if a.autotune and not (a.blocks_per_cu or a.alpha or a.beta or a.td_quads or a.bu_group or a.fuse_quads):
        # launch shape and switch thresholds chosen by the native autotuner on some of the roots
        t0 = time.time()
        best, best_ms, trials = B.autotune_bfs(prep.bvss, prep.csc, prep.inv_perm, prep.roots[:a.autotune_roots],
                                               **opts)
        for k in ("blocks_per_cu", "td_quads", "bu_group", "alpha", "beta", "fuse_quads"):
            opts[k] = best[k]
        tuned = {"blocks_per_cu": best["blocks_per_cu"], "td_quads": best["td_quads"], "bu_group": best["bu_group"],
                 "fuse_quads": best["fuse_quads"],
                 "alpha": best["alpha"], "beta": best["beta"], "trials": trials, "tuning_ms_per_bfs": round(best_ms, 4)}
        log(f"[bench] autotuned in {time.time() - t0:.1f}s: {tuned}")
    t0 = time.time()
    dg = B.DeviceGraph(prep.bvss, prep.csc, prep.inv_perm, **opts)
    log(f"[bench] device graph: {dg.info()['device_bytes'] / 2**20:.0f} MiB uploaded in {time.time() - t0:.1f}s")
    k = B.BVSSBFSKernel(dg)
    lib = B._lib.load()
    q = k.q._h
    roots = [int(r) for r in prep.roots]

    def run_root(r):
        B.check(lib.blest_bfs_async(q, r, None))
        st = B._lib.BfsStats()
        B.check(lib.blest_query_sync(q, st))
        return st

    for _ in range(a.warmup):
        for r in roots:
            run_root(r)
    samples = []  # (root, kernel_ms, m_c, bu_mask, levels)
    t_start = time.perf_counter()
    for _ in range(a.steps):
        for r in roots:
            st = run_root(r)
            samples.append((r, st.kernel_ms, st.edges_traversed_component, st.bottom_up_mask, st.total_levels,
                            st.td_levels, st.bu_levels))
    wall = time.perf_counter() - t_start
    gteps = [m / (ms * 1e-3) / 1e9 for _, ms, m, *_ in samples]
    value = harmonic(gteps)
    ms_per_step = wall * 1e3 / a.steps
    mean_ms = float(np.mean([s[1] for s in samples]))
    log(f"[bench] {len(samples)} launches: harmonic GTEPS {value:.2f}, mean kernel {mean_ms:.3f} ms, "
        f"levels {samples[0][4]} (td {samples[0][5]}, bu {samples[0][6]})")

    roofline = None
    cpu = None
    if a.roofline_roots < 0:
        a.roofline_roots = 8
    bv = None
    if not a.no_roofline or not a.no_cpu_baseline:
        import ctypes as C
        bv = prep.bvss.arrays()
        cp, rows = prep.csc.arrays()  # symmetric: the CSR of the bottom-up pass is the CSC
    if not a.no_roofline:
        # Algorithmic bytes per launch: the reference formulation's per-level bytes,
        # computed from each replayed root's own GPU levels and the direction schedule it
        # ran (reference_level_bytes).  Per direction: each replayed root traced once (a
        # HIP event around every kernel).
        t = time.time()
        replay_roots = roots if a.roofline_roots <= 0 else roots[:a.roofline_roots]
        entry_row = np.repeat(np.arange(prep.csc.n, dtype=np.int32), np.diff(cp.astype(np.int64)))
        tr = (B._lib.LevelTrace * 4096)()
        cnt = C.c_uint32()
        st = B._lib.BfsStats()
        lv = np.empty(prep.csc.n, dtype=np.int32)
        lvr = np.empty_like(lv)
        per_root, per_level = {}, {}
        by_dir = {0: [0, 0.0], 1: [0, 0.0]}  # bytes, seconds
        for r in replay_roots:
            B.check(lib.blest_bfs_trace(q, r, lv.ctypes.data_as(B._lib.I32P), C.byref(st), tr, 4096, C.byref(cnt)))
            lvr[prep.inv_perm] = lv  # levels by reordered ID
            per_level[r] = reference_level_bytes(prep.csc.n, cp, rows, bv, lvr, int(st.bottom_up_mask), entry_row)
            per_root[r] = sum(per_level[r])
            for i, b in enumerate(per_level[r]):
                if i >= cnt.value:
                    break
                d = int(tr[i].bottom_up)
                by_dir[d][0] += b
                by_dir[d][1] += (tr[i].expand_ms + tr[i].update_ms) * 1e-3
        used = [s_ for s_ in samples if s_[0] in per_root]
        total_bytes = sum(per_root[s_[0]] for s_ in used)
        total_s = sum(s_[1] for s_ in used) * 1e-3
        achieved = total_bytes / total_s / 1e9
        directions = {}
        for d, name in ((0, "top_down"), (1, "bottom_up")):
            if by_dir[d][1] > 0:
                ach = by_dir[d][0] / by_dir[d][1] / 1e9
                directions[name] = {"achieved": round(ach, 1), "frac": round(ach / HBM_PEAK_GBS, 4),
                                    "ms_per_launch": round(by_dir[d][1] * 1e3 / len(replay_roots), 4)}
        # HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
        # summary of this workload, used only when it was taken at the schedule this run
        # uses (launch shape, switch thresholds, fused small-level bound); otherwise null
        traffic = None
        tf = ROOT / "profiles" / f"bfs_kernel_{a.workload}_traffic.json"
        if tf.exists():
            tj = json.loads(tf.read_text())
            mine = {k: float(opts.get(k, dflt)) for k, dflt in (("blocks_per_cu", 4), ("td_quads", 2), ("bu_group", 8),
                                                                 ("alpha", 15.0), ("beta", 18.0),
                                                                 ("fuse_quads", 2048))}
            if mine["fuse_quads"] == 0:
                mine["fuse_quads"] = 2048.0
            for entry in tj.get("schedules", []):
                sched = entry.get("schedule") or {}
                if all(abs(float(sched.get(k, -1)) - v) < 1e-9 for k, v in mine.items()):
                    traffic = entry.get("bytes_per_launch")
            if traffic is None:
                log(f"[bench] no traffic summary in {tf.name} taken at this run's schedule {mine}: null")
        used_ms = float(np.mean([s_[1] for s_ in used]))
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": "bfs_kernel", "algorithmic_bytes_per_launch": int(total_bytes / len(used)),
                    "mean_launch_ms": round(used_ms, 4), "replayed_roots": len(replay_roots),
                    "by_direction": directions}
        log(f"[bench] byte model over {len(replay_roots)} roots in {time.time() - t:.1f}s: {achieved:.1f} GB/s")
    if not a.no_cpu_baseline:
        # Reference CPU path: level-synchronous BVSS BFS with the reference's switching
        # rule, all host cores, over as many roots as fit the time budget.
        import oracle
        threads = oracle.max_threads()
        t = time.time()
        cg, cr = [], []
        # BVSSBFSKernel::hostCode picks the switching kernel only for lazy (update
        # divergence > 25000) social graphs (BFS/BVSSBFSKernel.cuh:1292,1317-1340).
        switching = prep.social and prep.bvss.info()["update_divergence"] > 25000
        for r in roots:
            if cr and time.time() - t > a.cpu_seconds:
                break
            t1 = time.perf_counter()
            lv, _, _ = oracle.bvss_bfs(prep.csc.n, bv, cp, rows, int(prep.inv_perm[r]),
                                       allow_bottom_up=switching, threads=threads)
            dt = time.perf_counter() - t1
            m_c = next(s[2] for s in samples if s[0] == r)
            cg.append(m_c / dt / 1e9)
            cr.append(r)
        cpu = {"value": round(harmonic(cg), 4), "unit": "GTEPS", "cores": threads, "kind": "port",
               "sample": f"{len(cr)} of the {len(roots)} roots, oracle BVSS BFS with the reference's "
                         f"direction rule, {threads} OpenMP threads"}
        log(f"[bench] cpu baseline {cpu['value']} GTEPS over {len(cr)} roots")

    cc = None
    if not a.no_cc and info_sym(dg):
        # connected components on the same resident graph (canonical labels stay on the device)
        import ctypes as C
        cs = B._lib.CcStats()
        B.check(lib.blest_cc(q, 0, None, C.byref(cs)))  # warm-up (allocates the label arrays)
        cc_ms = []
        for _ in range(max(1, a.steps)):
            B.check(lib.blest_cc(q, 0, None, C.byref(cs)))
            cc_ms.append(cs.kernel_ms)
        m = dg.info()["nnz"] / 2
        cc = {"ms": round(float(np.mean(cc_ms)), 4), "components": cs.no_components,
              "medges_per_s": round(m / (np.mean(cc_ms) * 1e-3) / 1e6, 1), "algo": "hook+compress"}
        log(f"[bench] cc: {cc}")
    mb = None
    if not a.no_mbfs and info_sym(dg):
        # batched multi-source BFS against plain single-source BFS on the same resident
        # graph: 128-root launches over 1,024 seeded roots
        r = mbfs_bench(a, prep, opts, dg, algos=(0, 1))
        mb = {"gteps": r["value"], "plain_gteps": r["plain_gteps"], "speedup_vs_plain": r["speedup_vs_plain"],
              "roots": r["config"]["roots"], "batch": r["config"]["batch"], "batch_ms": r["batch_ms"],
              "plain_ms_per_root": r["plain_ms_per_root"], "algo": "bit-parallel",
              "alpha": r["config"]["alpha"], "beta": r["config"]["beta"], "i8_mfma": r["by_algo"]["1"]}
    info = dg.info()
    cfg = {"workload": w.name}
    if w.kind == "rmat":
        cfg.update({"graph": f"R-MAT scale {w.scale} edge factor {w.edge_factor} a=0.57 b=0.19 c=0.19 seed {w.seed}",
                    "vertices": info["n"], "directed_edges": info["nnz"]})
    else:
        cfg.update({"graph": f"{w.side}x{w.side} grid keep {w.keep} shortcuts {w.shortcut_frac} seed {w.seed}",
                    "vertices": info["n"], "directed_edges": info["nnz"]})
    if tuned:
        cfg["autotuned"] = tuned
    if a.use_graphs >= 0:
        cfg["use_graphs"] = a.use_graphs
    cfg.update({"reorder": f"{w.reorder.name.lower()} window {w.window}", "roots": len(roots),
                "slices": info["slices"], "vsets": info["vsets"]})
    line = {"metric": f"BFS GTEPS (harmonic mean, {len(roots)} roots)", "value": round(value, 3), "unit": "GTEPS",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_per_step, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic",
            "config": cfg, "roofline": roofline, "cpu_baseline": cpu, "cc": cc, "mbfs": mb}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
