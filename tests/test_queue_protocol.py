"""The task FIFO of the queued multi-block launch (glx_run_blocks_queued,
wh_queue_kernel), as an event-driven model on the CPU.

The kernel's protocol: the FIFO starts with every tile's first segment in
slots [0, n_tiles); a workgroup takes slot h = head++ and waits until the
slot holds an entry, runs that task (one tile through one segment), then, if
the tile has a next segment, claims slot = tail++ and writes that segment
into it, and takes its next slot.  There are at most n_tiles workgroups.

Checked over random task durations and shapes: every task runs exactly once;
a tile's segments run in order, each after the previous one finished; the
run never reaches a state where every remaining workgroup waits on a slot
that nothing will fill; and while more tiles are unfinished than there are
workgroups, every slot taken has already been claimed by its writer (the
kernel's only wait then is the few instructions between claim and write)."""
import heapq
import random

import pytest


def run_protocol(n_wg: int, n_tiles: int, n_segs: int, seed: int):
    rng = random.Random(seed)
    n_tasks = n_tiles * n_segs
    ring = [0] * n_tasks
    for i in range(n_tiles):
        ring[i] = i + 1
    head, tail = 0, n_tiles
    done_segs = [0] * n_tiles
    runs = [0] * n_tasks
    events = []
    waiting = {}
    unclaimed_takes = 0

    def take(wg, now):
        nonlocal head, unclaimed_takes
        h = head
        head += 1
        if h >= n_tasks:
            return
        unfinished = sum(1 for d in done_segs if d < n_segs)
        if h >= tail and unfinished > n_wg:
            unclaimed_takes += 1
        if ring[h]:
            start(wg, h, now)
        else:
            waiting[wg] = h

    def start(wg, h, now):
        task = ring[h] - 1
        seg, tile = divmod(task, n_tiles)
        assert done_segs[tile] == seg, "a segment started before its predecessor finished"
        heapq.heappush(events, (now + rng.choice([0.3, 1.0, 3.0]) * rng.uniform(0.8, 1.2), wg, task))

    for wg in range(n_wg):
        take(wg, 0.0)
    while events:
        now, wg, task = heapq.heappop(events)
        seg, tile = divmod(task, n_tiles)
        runs[task] += 1
        done_segs[tile] += 1
        if seg + 1 < n_segs:
            slot = tail
            tail += 1
            ring[slot] = task + n_tiles + 1
            for w2, h in list(waiting.items()):
                if h == slot:
                    del waiting[w2]
                    start(w2, h, now)
        take(wg, now)
    return runs, done_segs, waiting, unclaimed_takes


@pytest.mark.parametrize("seed", range(60))
def test_fifo_protocol_random_shapes(seed):
    r = random.Random(seed)
    n_tiles = r.randint(1, 40)
    n_wg = r.randint(1, n_tiles)
    n_segs = r.randint(1, 12)
    runs, done_segs, waiting, unclaimed = run_protocol(n_wg, n_tiles, n_segs, seed)
    assert all(k == 1 for k in runs)
    assert all(d == n_segs for d in done_segs)
    assert not waiting, "workgroups left waiting on slots nothing fills"
    assert unclaimed == 0


def test_fifo_protocol_benchmark_shape():
    """768 workgroups, 782 tiles (200,000 particles), 10 blocks of 8 segments."""
    runs, done_segs, waiting, unclaimed = run_protocol(768, 782, 80, 1)
    assert all(k == 1 for k in runs) and all(d == 80 for d in done_segs)
    assert not waiting and unclaimed == 0
