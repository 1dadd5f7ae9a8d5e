"""Contractions replayed from a captured hipGraph.

The same sequence of calls runs twice on tensors built from the same
seeds, once with graph launch on and once off: alpha and beta vary
(including beta = 0 and alpha = 0, which select other captured graphs),
and eager work (block writes, xm_add) runs between the replays.  After
every call the results must be bit-identical, and once a plan's tile shape
is decided its calls must replay the graph.
"""
import os

import numpy as np
import pytest

from libxm_amd import workloads, xm

pytestmark = pytest.mark.gpu

CALLS = [(0.5, 1.0), (0.5, 1.0), (0.5, 1.0), (0.5, 1.0), (1.25, 0.0),
         (0.5, 1.0), (0.0, 2.0), (-0.75, 0.5), (0.5, 1.0), (1.25, 0.0)]


def _sequence(lib, w, seed, graph, splits=None):
    """The call sequence; splits (a list) collects each call's count of
    FP64 tiles run as two K halves."""
    lib.set_graph_launch(graph)
    alloc = lib.allocator()
    t, v, c, c2 = (s.build(lib, alloc) for s in (*w.specs, w.c))
    for k, x in enumerate((t, v, c, c2)):
        lib.fill_random(x, seed + k)
    rng = np.random.default_rng(seed)
    tblocks = w.t.canonical_blocks()
    outs, replayed = [], []
    for n, (alpha, beta) in enumerate(CALLS):
        lib.contract(alpha, t, v, beta, c, *w.idx)
        st = lib.stats()
        replayed.append(st["graph"])
        if splits is not None:
            splits.append(st["ksplit_tiles"])
        outs.append(w.c.read(lib, c))
        # eager work between calls: a changed t block, then c += 0.25 c2
        idx = tblocks[rng.integers(len(tblocks))]
        blk = np.asarray(lib.read_block(t, idx), dtype=np.float64)
        lib.write_block(t, idx, blk * 0.5 + 0.125)
        lib.add(1.0, c, 0.25, c2, w.idx[2], w.idx[2])
    for x in (t, v, c, c2):
        lib.tensor_free(x)
    lib.allocator_destroy(alloc)
    lib.set_graph_launch(True)
    return outs, replayed


@pytest.mark.parametrize("make,tile", [
    (lambda: workloads.ladder_d2h(32, 96, 3), None),
    (lambda: workloads.ladder_d2h(32, 96, 3), "64x128"),
    (lambda: workloads.ring_d2h(24, 72, 5, lo=4, hi=12), None),
], ids=["ladder", "ladder-64x128", "ring"])
def test_graph_replays_match_direct_launches(gpulib, make, tile):
    w = make()
    if tile:
        os.environ["XM_GEMM_TILE"] = tile
    try:
        direct, r0 = _sequence(gpulib, w, 31, False)
        splits = []
        graph, r1 = _sequence(gpulib, w, 31, True, splits)
    finally:
        os.environ.pop("XM_GEMM_TILE", None)
    assert not any(r0)
    # the 64 x 128 ladder's replays include the K-split sum kernel
    if tile:
        assert all(splits), splits
    # calls 0-2 tune the tile shape (direct launches; a pinned shape is
    # decided at once); every later call of the plan replays a graph
    assert all(r1[1 if tile else 3:]), r1
    for n, (a, b) in enumerate(zip(direct, graph)):
        for k in a:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), (n, k)


def test_graph_split_precision(gpulib):
    w = workloads.ladder_d2h(32, 96, 5)
    gpulib.set_precision(xm.PRECISION_SPLIT_BF16, 5)
    try:
        direct, _ = _sequence(gpulib, w, 41, False)
        graph, r1 = _sequence(gpulib, w, 41, True)
    finally:
        gpulib.set_precision(xm.PRECISION_NATIVE)
    assert all(r1[1:]), r1      # one tile shape: decided at once
    for a, b in zip(direct, graph):
        for k in a:
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k]))
