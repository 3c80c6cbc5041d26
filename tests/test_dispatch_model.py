"""CPU model of workgroup dispatch with bounded residency.

Why no kernel here waits on another workgroup: a GPU dispatches a grid's
workgroups in blockIdx order into a bounded number of resident slots
(occupancy x CUs) and gives no forward-progress guarantee to workgroups that
are not resident yet. A protocol in which a resident workgroup spins until
another workgroup has produced data (per-XCD work queues with completion
counters and a ring buffer, for instance) hangs as soon as every slot holds a
spinning consumer whose producer has not been dispatched. Splitting a
four-step transform's column and row tiles over one launch is such a
protocol; stream-ordered launches (columns, then rows) and the
one-workgroup-per-transform kernel are not: every workgroup of them finishes
on its own.

The model is deterministic: `slots` workgroups run at once, the dispatcher
fills a free slot with the next blockIdx, a running workgroup completes when
every workgroup it waits on has completed (waits are the only dependencies),
and a step in which nothing completes while workgroups remain is a hang.
"""
import itertools

import pytest


def run_grid(deps, slots):
    """deps[w] = workgroups w waits on. Returns True when the grid drains."""
    n = len(deps)
    done = [False] * n
    resident = []
    nxt = 0
    while True:
        while len(resident) < slots and nxt < n:
            resident.append(nxt)
            nxt += 1
        if not resident:
            return True
        finished = [w for w in resident if all(done[d] for d in deps[w])]
        if not finished:
            return False  # every resident workgroup spins on a non-resident one
        for w in finished:
            done[w] = True
            resident.remove(w)


def queue_grid(transforms, cols, rows, xcds, order):
    """One launch holding the column and row tiles of every transform; a row
    tile waits on all column tiles of its transform. `order` places the
    tiles in blockIdx order: 'interleaved' puts each transform's row tiles
    right after its column tiles per XCD round, 'rows_first' dispatches row
    tiles (consumers) ahead of column tiles (producers) as a per-XCD queue
    that balances by tile cost can."""
    tiles = []
    for t in range(transforms):
        tiles += [("col", t, i) for i in range(cols)]
        tiles += [("row", t, i) for i in range(rows)]
    if order == "rows_first":
        tiles.sort(key=lambda x: (x[0] != "row", x[1], x[2]))
    elif order == "xcd_round_robin":
        # blockIdx b runs on XCD b % xcds; each XCD's queue holds whole
        # transforms, its longer row tiles first (largest-first balancing),
        # so consumers sit at lower blockIdx than the producers they need
        per = [[] for _ in range(xcds)]
        for t in range(transforms):
            per[t % xcds] += [("row", t, i) for i in range(rows)]
        for t in range(transforms):
            per[t % xcds] += [("col", t, i) for i in range(cols)]
        tiles = [tile for group in itertools.zip_longest(*per) for tile in group if tile is not None]
    index = {tile: b for b, tile in enumerate(tiles)}
    deps = []
    for kind, t, i in tiles:
        deps.append([index[("col", t, c)] for c in range(cols)] if kind == "row" else [])
    return deps


def stream_ordered(transforms, cols, rows):
    """Columns in one launch, rows in the next (the library's four-step): no
    workgroup of a launch waits on another workgroup of that launch."""
    return [[] for _ in range(transforms * cols)], [[] for _ in range(transforms * rows)]


@pytest.mark.parametrize("slots", [1, 4, 16, 64, 256])
def test_consumers_ahead_of_producers_hang_under_bounded_residency(slots):
    # enough row tiles ahead of the first column tile to fill every slot
    deps = queue_grid(transforms=4 * slots, cols=4, rows=15, xcds=8, order="rows_first")
    assert not run_grid(deps, slots)


@pytest.mark.parametrize("slots", [8, 32, 256])
def test_per_xcd_queues_hang_when_a_queue_runs_ahead(slots):
    deps = queue_grid(transforms=8 * slots, cols=4, rows=15, xcds=8, order="xcd_round_robin")
    assert not run_grid(deps, slots)


def test_the_same_protocol_drains_only_when_every_workgroup_is_resident():
    # with residency for the whole grid (what a cooperative launch would have
    # to guarantee) the spin-waits end
    deps = queue_grid(transforms=16, cols=4, rows=15, xcds=8, order="rows_first")
    assert run_grid(deps, slots=len(deps))
    assert not run_grid(deps, slots=len(deps) - 16 * 4)


@pytest.mark.parametrize("slots", [1, 3, 16, 256])
def test_stream_ordered_launches_always_drain(slots):
    cols, rows = stream_ordered(transforms=1250, cols=4, rows=15)
    assert run_grid(cols, slots) and run_grid(rows, slots)


@pytest.mark.parametrize("slots", [1, 7, 256])
def test_one_workgroup_per_transform_always_drains(slots):
    # kernels_fs1.hip: each workgroup owns its transforms and scratch slot;
    # its barriers order its own steps and it waits on nobody
    assert run_grid([[] for _ in range(256)], slots)
