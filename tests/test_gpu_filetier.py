"""Allocators created with a path: blocks past HBM and pinned host memory
live in that file.

XM_ARENA_BYTES and XM_HOST_BYTES (read when the allocator is created)
make the tiers small here, so that tensors of test size spill into the
file.  No kernel reads the file: contractions read a batch's file blocks
into pinned buffers, copy the operand ranges into HBM like host blocks and
run the C blocks' permutes on the pinned copies, then write the C blocks
back (XM_STAGE_BYTES makes the batches small); element operations and
the adjoint's folds run on pinned copies of the file blocks they touch.
Results must equal the oracle's as for HBM-resident tensors: FP64
contractions to 1e-13 relative, element operations bit for bit.  The
allocator follows libxm's pagefile (libxm src/alloc.c:147-347): the file
is created by xm_allocator_create (an unusable path makes it fail) and
removed by xm_allocator_destroy.
"""
import ctypes as C
import os

import numpy as np
import pytest

from libxm_amd import xm, workloads
from libxm_amd.spec import rel_err

from cases import ALPHA_BETA
from helpers import build_filled, resolve, round_scalar

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def path_allocator(lib, path, monkeypatch, arena, host):
    """An allocator with a backing file and the given tier caps."""
    monkeypatch.setenv("XM_ARENA_BYTES", str(arena))
    monkeypatch.setenv("XM_HOST_BYTES", str(host))
    alloc = lib.allocator(str(path))
    monkeypatch.delenv("XM_ARENA_BYTES")
    monkeypatch.delenv("XM_HOST_BYTES")
    return alloc


def fill_hbm(lib, alloc):
    """Allocate 1 MiB blocks until the arena has left HBM."""
    held = []
    while (lib.allocator_usage(alloc)[1] == 0 and
           lib.allocator_file_bytes(alloc) == 0):
        held.append(lib.lib.xm_allocator_allocate(alloc, MiB))
    return held


def test_allocator_file(gpulib, tmp_path, monkeypatch):
    path = tmp_path / "pagefile"
    alloc = path_allocator(gpulib, path, monkeypatch, 0, 0)
    assert path.exists()
    lib = gpulib.lib
    rng = np.random.default_rng(5)
    ptrs = []
    for n in (1, 4096, 3 * MiB + 17, 100):
        p = lib.xm_allocator_allocate(alloc, n)
        assert p != xm.NULL_PTR
        data = rng.integers(0, 256, n, dtype=np.uint8)
        lib.xm_allocator_write(alloc, p, data.ctypes.data_as(C.c_void_p), n)
        ptrs.append((p, data))
    assert gpulib.allocator_usage(alloc) == (0, 0)
    assert gpulib.allocator_file_bytes(alloc) > 3 * MiB
    for p, data in ptrs:
        back = np.empty_like(data)
        lib.xm_allocator_read(alloc, p, back.ctypes.data_as(C.c_void_p),
                              data.size)
        assert np.array_equal(back, data)
    # a freed range is reused
    lib.xm_allocator_deallocate(alloc, ptrs[1][0])
    again = lib.xm_allocator_allocate(alloc, 4096)
    assert again == ptrs[1][0]
    for p in [again] + [q for q, _ in ptrs[:1] + ptrs[2:]]:
        lib.xm_allocator_deallocate(alloc, p)
    assert gpulib.allocator_file_bytes(alloc) == 0
    gpulib.allocator_destroy(alloc)
    assert not path.exists()
    with pytest.raises(MemoryError):
        gpulib.allocator(str(tmp_path / "missing" / "pagefile"))


@pytest.mark.parametrize("tiers", ["file", "three"])
@pytest.mark.parametrize("kind", ["ladder", "ring"])
def test_contract_file_blocks(kind, tiers, gpulib, oracle, tmp_path,
                              monkeypatch):
    """Every block in the file, or t in HBM, then pinned host memory up to
    4 MiB, the rest of v and c in the file."""
    w = workloads.ladder_small() if kind == "ladder" else \
        workloads.ring_d2h(24, 60, 3)
    if tiers == "file":
        alloc = path_allocator(gpulib, tmp_path / "pf", monkeypatch, 0, 0)
        held = []
        t = w.specs[0].build(gpulib, alloc)
    else:
        alloc = path_allocator(gpulib, tmp_path / "pf", monkeypatch,
                               64 * MiB, 4 * MiB)
        t = w.specs[0].build(gpulib, alloc)
        held = fill_hbm(gpulib, alloc)
    v, c = (s.build(gpulib, alloc) for s in w.specs[1:])
    for k, x in enumerate((t, v, c)):
        gpulib.fill_random(x, 60 + k)
    dev, host = gpulib.allocator_usage(alloc)
    assert gpulib.allocator_file_bytes(alloc) > 0
    if tiers == "three":
        assert dev > 0 and host > 0
    else:
        assert dev == 0 and host == 0
    data = [s.read(gpulib, x) for s, x in zip(w.specs, (t, v, c))]
    oal = oracle.allocator()
    ot = [build_filled(oracle, oal, s, d) for s, d in zip(w.specs, data)]
    monkeypatch.setenv("XM_STAGE_BYTES", str(2 * MiB))
    rng = np.random.default_rng(3)
    for al, be in ALPHA_BETA + [(0.0, 2.0), (0.7, 0.0), (0.5, 1.0)]:
        al = round_scalar(resolve(al, rng, xm.SCALAR_DOUBLE),
                          xm.SCALAR_DOUBLE).real
        be = round_scalar(resolve(be, rng, xm.SCALAR_DOUBLE),
                          xm.SCALAR_DOUBLE).real
        gpulib.contract(al, t, v, be, c, *w.idx)
        st = gpulib.stats()
        oracle.contract(al, ot[0], ot[1], be, ot[2], *w.idx)
        got = w.specs[2].read(gpulib, c)
        want = w.specs[2].read(oracle, ot[2])
        assert rel_err(got, want) <= 1e-13, (al, be)
        assert st["staged_batches"] >= 2, st
        assert st["file_bytes"] > 0, st
    for x in ot:
        oracle.tensor_free(x)
    oracle.allocator_destroy(oal)
    for x in (t, v, c):
        gpulib.tensor_free(x)
    for p in held:
        gpulib.lib.xm_allocator_deallocate(alloc, p)
    gpulib.allocator_destroy(alloc)


def test_element_ops_file_blocks(gpulib, oracle, tmp_path, monkeypatch):
    """xm_set, xm_add (also a == b), xm_copy, xm_mul, xm_div, xm_dot,
    fill_random and element reads on tensors in the file, in batches of
    64 KiB of staged blocks, against the oracle bit for bit."""
    w = workloads.ladder_small()
    sp = w.specs[0]
    alloc = path_allocator(gpulib, tmp_path / "pf", monkeypatch, 0, 0)
    a, b = sp.build(gpulib, alloc), sp.build(gpulib, alloc)
    monkeypatch.setenv("XM_STAGE_BYTES", str(64 << 10))
    gpulib.fill_random(a, 1)
    gpulib.fill_random(b, 2, 1.0, 2.0)
    assert gpulib.allocator_file_bytes(alloc) > 0
    # fill_random's values do not depend on where the blocks live
    hal = gpulib.allocator()
    h = sp.build(gpulib, hal)
    gpulib.fill_random(h, 1)
    ref = sp.read(gpulib, h)
    gpulib.tensor_free(h)
    gpulib.allocator_destroy(hal)
    data = [sp.read(gpulib, x) for x in (a, b)]
    for k in ref:
        assert np.array_equal(data[0][k], ref[k]), k
    oal = oracle.allocator()
    oa, ob = (build_filled(oracle, oal, sp, d) for d in data)
    outs, dots = [], []
    for lib, x, y in ((gpulib, a, b), (oracle, oa, ob)):
        seq = []
        lib.add(0.5, x, -1.25, y, "ijab", "jiba")
        seq.append(sp.read(lib, x))
        for ia, ib in workloads.antisym_strings(4, w.pairs):
            lib.add(0.5, x, -0.5, x, ia, ib)
        seq.append(sp.read(lib, x))
        lib.mul(x, y, "ijab", "ijab")
        seq.append(sp.read(lib, x))
        lib.div(x, y, "ijab", "jiba")
        seq.append(sp.read(lib, x))
        dots.append(lib.dot(x, y, "ijab", "jiba"))
        lib.copy(y, 2.0, x, "ijab", "ijab")
        seq.append(sp.read(lib, y))
        idx = sp.canonical_blocks()[3]
        off = tuple(o + 1 for o in sp.block_offsets(idx))
        seq.append({"elem": np.array([lib.get_element(y, off)])})
        lib.set(x, 3.0)
        seq.append(sp.read(lib, x))
        outs.append(seq)
    for n, (g, r) in enumerate(zip(*outs)):
        for k in r:
            assert np.array_equal(g[k], r[k]), (n, k)
    assert abs(dots[0] - dots[1]) <= 1e-12 * max(1.0, abs(dots[1]))
    for x in (oa, ob):
        oracle.tensor_free(x)
    oracle.allocator_destroy(oal)
    for x in (a, b):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def test_backward_file_blocks(gpulib, tmp_path, monkeypatch):
    """xm_contract_backward with every tensor in the file (staged
    gradient contractions, folds on pinned copies) against the same call
    on HBM tensors."""
    w = workloads.ladder_small()
    specs = w.specs
    rng = np.random.default_rng(14)
    data = [s.random_data(rng) for s in specs]
    out = []
    for where in ("hbm", "file"):
        if where == "file":
            alloc = path_allocator(gpulib, tmp_path / "pf", monkeypatch, 0, 0)
            monkeypatch.setenv("XM_STAGE_BYTES", str(2 * MiB))
        else:
            alloc = gpulib.allocator()
        a, b, dc = (build_filled(gpulib, alloc, s, d)
                    for s, d in zip(specs, data))
        da, db = specs[0].build(gpulib, alloc), specs[1].build(gpulib, alloc)
        for _ in range(2):      # the second call reuses plans and aux
            gpulib.contract_backward(0.75, a, b, dc, da, db, *w.idx)
        if where == "file":
            assert gpulib.allocator_file_bytes(alloc) > 0
            assert gpulib.allocator_usage(alloc) == (0, 0)
        out.append((specs[0].read(gpulib, da), specs[1].read(gpulib, db)))
        for x in (a, b, dc, da, db):
            gpulib.tensor_free(x)
        gpulib.allocator_destroy(alloc)
    assert rel_err(out[1][0], out[0][0]) <= 1e-13
    assert rel_err(out[1][1], out[0][1]) <= 1e-13
