"""Tensors beyond the HBM arena: blocks in pinned host memory.

With XM_ARENA_BYTES the arena spills to pinned host memory early, so the
overflow path runs at test sizes.  Contractions that touch host blocks
run in batches of C blocks (XM_STAGE_BYTES makes the batches small here),
each batch's host operand blocks copied into HBM on the side stream while
the previous batch computes; host C blocks go through HBM panels.  Results
must equal the oracle's as for HBM-resident tensors: FP64 to 1e-13
relative, element operations bit for bit.
"""
import numpy as np
import pytest

from libxm_amd import xm, workloads
from libxm_amd.spec import rel_err

from helpers import build_filled

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def fill_hbm(lib, alloc):
    """Allocate 1 MiB blocks until the arena has spilled to host memory,
    so that later blocks land in host memory (or in the HBM remainder)."""
    held = []
    while lib.allocator_usage(alloc)[1] == 0:
        held.append(lib.lib.xm_allocator_allocate(alloc, MiB))
    return held


def read_all(lib, specs, ts):
    return [s.read(lib, t) for s, t in zip(specs, ts)]


@pytest.mark.parametrize("kind", ["ladder", "ring"])
def test_contract_with_host_blocks(kind, gpulib, oracle, monkeypatch):
    if kind == "ladder":
        w = workloads.ladder_small()
    else:
        w = workloads.ring_d2h(24, 60, 3)
    monkeypatch.setenv("XM_ARENA_BYTES", str(64 * MiB))
    alloc = gpulib.allocator()
    monkeypatch.delenv("XM_ARENA_BYTES")
    t = w.specs[0].build(gpulib, alloc)
    held = fill_hbm(gpulib, alloc)
    v, c = (s.build(gpulib, alloc) for s in w.specs[1:])
    for k, x in enumerate((t, v, c)):
        gpulib.fill_random(x, 40 + k)
    dev, host = gpulib.allocator_usage(alloc)
    assert host > 0 and dev > 0
    data = read_all(gpulib, w.specs, (t, v, c))

    oal = oracle.allocator()
    ot = [build_filled(oracle, oal, s, d) for s, d in zip(w.specs, data)]

    monkeypatch.setenv("XM_STAGE_BYTES", str(2 * MiB))
    for al, be in ((0.5, 1.0), (1.5, -0.5), (0.0, 2.0), (0.7, 0.0),
                   (0.5, 1.0)):
        gpulib.contract(al, t, v, be, c, *w.idx)
        st = gpulib.stats()
        oracle.contract(al, ot[0], ot[1], be, ot[2], *w.idx)
        got = w.specs[2].read(gpulib, c)
        want = w.specs[2].read(oracle, ot[2])
        assert rel_err(got, want) <= 1e-13, (al, be)
        assert st["staged_batches"] >= 2, st
        if al != 0:
            assert st["staged_bytes"] > 0, st
    for x in ot:
        oracle.tensor_free(x)
    oracle.allocator_destroy(oal)
    for x in (t, v, c):
        gpulib.tensor_free(x)
    for p in held:
        gpulib.lib.xm_allocator_deallocate(alloc, p)
    gpulib.allocator_destroy(alloc)


def test_element_ops_on_host_blocks(gpulib, oracle, monkeypatch):
    """xm_set, xm_add and xm_copy read and write host blocks in place."""
    w = workloads.ladder_small()
    monkeypatch.setenv("XM_ARENA_BYTES", "0")
    alloc = gpulib.allocator()
    monkeypatch.delenv("XM_ARENA_BYTES")
    a = w.specs[0].build(gpulib, alloc)
    b = w.specs[0].build(gpulib, alloc)
    gpulib.fill_random(a, 1)
    gpulib.fill_random(b, 2)
    assert gpulib.allocator_usage(alloc)[0] == 0
    data = read_all(gpulib, [w.specs[0]] * 2, (a, b))
    oal = oracle.allocator()
    oa, ob = (build_filled(oracle, oal, w.specs[0], d) for d in data)
    outs = []
    for lib, x, y in ((gpulib, a, b), (oracle, oa, ob)):
        seq = []
        lib.add(0.5, x, -1.25, y, "ijab", "jiba")
        seq.append(w.specs[0].read(lib, x))
        lib.copy(y, 2.0, x, "ijab", "ijab")
        seq.append(w.specs[0].read(lib, y))
        lib.set(x, 3.0)
        seq.append(w.specs[0].read(lib, x))
        outs.append(seq)
    for n, (g, r) in enumerate(zip(*outs)):
        for k in r:
            assert np.array_equal(g[k], r[k]), (n, k)
    for x in (oa, ob):
        oracle.tensor_free(x)
    oracle.allocator_destroy(oal)
    for x in (a, b):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)
