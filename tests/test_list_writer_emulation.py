"""CPU emulation of the block-wide vset list writers (blest_amd/csrc/device/segments.hpp:
block_write_vsets, and the list part of bfs_update_kernel in device/bfs.hip): one thread
per rset of a chunk of blockDim/8 bitmap words, an exclusive block scan of two counts,
then

  * multi-source writer: an rset of fewer than kRunVsets vsets writes its own entries at
    its scan offset; larger ones are queued and written after them as runs at
    total + (their own scan offset), each run by one wave;
  * single-source update: an rset of at most kChunkVsets vsets writes its own entries; a
    larger one emits ceil(cnt / kChunkVsets) chunk entries (first vset, count) at its
    chunk-scan offset.

The emulation writes into an array the size of the segment's list range and checks that
every vset of every active rset appears exactly once, that nothing is written past the
returned count, and that the chunk entries cover each large rset's vsets exactly."""
import numpy as np
import pytest

K_RUN = 32     # kRunVsets
K_CHUNK = 16   # kChunkVsets


def emulate_block_write_vsets(rset_unit, ws, we, active, nb=512, k_run=K_RUN):
    """active[w] = 8-bit mask of active rsets of word w.  Returns the written list."""
    capacity = int(rset_unit[min(8 * we, len(rset_unit) - 1)] - rset_unit[8 * ws])
    out = np.full(capacity + 1, -1, dtype=np.int64)
    written = 0
    for c0 in range(ws, we, nb // 8):
        start = np.zeros(nb, dtype=np.int64)
        cnt = np.zeros(nb, dtype=np.int64)
        for t in range(nb):
            w, b = c0 + t // 8, t % 8
            if w < we and (active[w] >> b) & 1:
                start[t] = rset_unit[8 * w + b]
                cnt[t] = rset_unit[8 * w + b + 1] - start[t]
        run = cnt >= k_run
        x = np.where(run, 0, cnt)
        y = np.where(run, cnt, 0)
        ex = np.concatenate([[0], np.cumsum(x)[:-1]])
        rex = np.concatenate([[0], np.cumsum(y)[:-1]])
        total, rtotal = int(x.sum()), int(y.sum())
        for t in range(nb):
            if not run[t]:
                for j in range(cnt[t]):
                    out[written + ex[t] + j] = start[t] + j
        queue = [t for t in range(nb) if run[t]]
        np.random.default_rng(c0).shuffle(queue)  # the LDS queue's order is free
        nwaves = nb // 64
        for wave in range(nwaves):
            for k in range(wave, len(queue), nwaves):
                t = queue[k]
                for j in range(cnt[t]):
                    out[written + total + rex[t] + j] = start[t] + j
        written += total + rtotal
    assert written <= capacity
    assert (out[written:] == -1).all()
    return out[:written]


def emulate_update_lists(rset_unit, ws, we, active, nb=512, k_chunk=K_CHUNK):
    """The single-source update's list and chunk list for one segment."""
    out, cout = [], []
    written = cwritten = 0
    lst = {}
    chl = {}
    for c0 in range(ws, we, nb // 8):
        start = np.zeros(nb, dtype=np.int64)
        cnt = np.zeros(nb, dtype=np.int64)
        for t in range(nb):
            w, b = c0 + t // 8, t % 8
            if w < we and (active[w] >> b) & 1:
                start[t] = rset_unit[8 * w + b]
                cnt[t] = rset_unit[8 * w + b + 1] - start[t]
        chunks = np.where(cnt > k_chunk, (cnt + k_chunk - 1) // k_chunk, 0)
        x = np.where(chunks > 0, 0, cnt)
        ex = np.concatenate([[0], np.cumsum(x)[:-1]])
        cex = np.concatenate([[0], np.cumsum(chunks)[:-1]])
        for t in range(nb):
            if chunks[t] == 0:
                for j in range(cnt[t]):
                    assert written + ex[t] + j not in lst
                    lst[written + ex[t] + j] = start[t] + j
            for j in range(chunks[t]):
                chl[cwritten + cex[t] + j] = (start[t] + j * k_chunk, min(k_chunk, cnt[t] - j * k_chunk))
        written += int(x.sum())
        cwritten += int(chunks.sum())
    assert sorted(lst) == list(range(written)) and sorted(chl) == list(range(cwritten))
    return [lst[i] for i in range(written)], [chl[i] for i in range(cwritten)]


def random_segment(seed, words=256, hub_every=37):
    rng = np.random.default_rng(seed)
    R = 8 * words
    cnt = rng.integers(1, 4, size=R)
    hubs = rng.choice(R, size=max(1, R // hub_every), replace=False)
    cnt[hubs] = rng.integers(17, 3000, size=len(hubs))  # hub rsets: chunk / run sized
    mid = rng.choice(R, size=max(1, R // 11), replace=False)
    cnt[mid] = rng.integers(5, 80, size=len(mid))  # around both thresholds
    rset_unit = np.concatenate([[0], np.cumsum(cnt)])
    active = rng.integers(0, 256, size=words)
    active[rng.random(words) < 0.3] = 0
    return rset_unit, active


def expected_vsets(rset_unit, ws, we, active):
    want = []
    for w in range(ws, we):
        for b in range(8):
            if (active[w] >> b) & 1:
                want.extend(range(rset_unit[8 * w + b], rset_unit[8 * w + b + 1]))
    return want


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("nb", [512, 256])
def test_multi_source_writer_lists_every_vset_once(seed, nb):
    rset_unit, active = random_segment(seed)
    ws, we = 0, len(active)
    got = emulate_block_write_vsets(rset_unit, ws, we, active, nb=nb)
    assert sorted(got.tolist()) == expected_vsets(rset_unit, ws, we, active)


@pytest.mark.parametrize("seed", range(4))
def test_multi_source_writer_run_threshold_edges(seed):
    """Rsets of exactly kRunVsets - 1 and kRunVsets vsets, and a chunk of only runs."""
    words = 64
    cnt = np.full(8 * words, K_RUN - 1)
    cnt[1::2] = K_RUN
    rset_unit = np.concatenate([[0], np.cumsum(cnt)])
    rng = np.random.default_rng(seed)
    active = rng.integers(1, 256, size=words)
    got = emulate_block_write_vsets(rset_unit, 0, words, active)
    assert sorted(got.tolist()) == expected_vsets(rset_unit, 0, words, active)
    cnt[:] = K_RUN + 3
    rset_unit = np.concatenate([[0], np.cumsum(cnt)])
    got = emulate_block_write_vsets(rset_unit, 0, words, active)
    assert sorted(got.tolist()) == expected_vsets(rset_unit, 0, words, active)


@pytest.mark.parametrize("seed", range(4))
def test_single_source_update_lists_and_chunks(seed):
    rset_unit, active = random_segment(seed)
    ws, we = 0, len(active)
    lst, chl = emulate_update_lists(rset_unit, ws, we, active)
    covered = list(lst)
    for first, n in chl:
        assert 1 <= n <= K_CHUNK
        covered.extend(range(first, first + n))
    assert sorted(covered) == expected_vsets(rset_unit, ws, we, active)
    # entries of rsets listed per vset keep rset order (each rset writes its own run)
    assert lst == sorted(lst)
