"""K split of FP64 tiles.

The C blocks at the end of each XCD's tile list run as two halves of their
16-k steps, and a second kernel writes C = beta * C + (half 0 + half 1).
On small contractions every C block of at least 64 steps is split, so
these tests run the split path throughout.  Results must match the oracle
to 1e-13 relative (the two partial sums round differently from one running
sum), keep libxm's alpha == 0 and beta == 0 rules exactly, be bit-identical
across tile shapes, between the DMA-staged and register-staged kernels and
the DMA kernel's persistent form (XM_GEMM_PERSIST=1: resident workgroups
taking tiles from per-XCD queues; every shape cuts a C block at the same
step), and XM_GEMM_KSPLIT=0 turns the split off.
"""
import os

import numpy as np
import pytest

from libxm_amd import workloads
from libxm_amd.spec import Spec, rel_err

from helpers import build_filled
from test_gpu_parity import run_contract

pytestmark = pytest.mark.gpu


def _fill(lib, w, tensors, seed):
    for k, x in enumerate(tensors):
        lib.fill_random(x, seed + k)
        for ia, ib in workloads.antisym_strings(4, w.pairs):
            lib.add(0.5, x, -0.5, x, ia, ib)


def _workload():
    return workloads.ladder_d2h(32, 96, 7)


@pytest.fixture
def env():
    """The K split is a property of the 64 x 128 and 64 x 64 forms (the
    stacked 128 x 128 form, which plans the DMA kernel can run try first,
    runs all of K): the tests pin 64 x 128 unless they choose a shape, and
    keep each C block a GEMM of its own (no dense classes)."""
    keys = ("XM_GEMM_KSPLIT", "XM_GEMM_TILE", "XM_GEMM_DMA",
            "XM_GEMM_PERSIST", "XM_CLASS_GEMM")
    old = {k: os.environ.get(k) for k in keys}
    os.environ["XM_GEMM_TILE"] = "64x128"
    os.environ["XM_CLASS_GEMM"] = "0"
    yield os.environ
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def test_split_matches_oracle(gpulib, oracle, env):
    env.pop("XM_GEMM_KSPLIT", None)
    w = _workload()
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    _fill(gpulib, w, (t, v, c), 31)
    data = [s.read(gpulib, x) for s, x in zip(w.specs, (t, v, c))]
    oal = oracle.allocator()
    ot = [build_filled(oracle, oal, s, d) for s, d in zip(w.specs, data)]
    for al, be in ((0.5, 1.0), (1.5, -0.5), (0.7, 0.0), (-1.0, 2.0)):
        gpulib.contract(al, t, v, be, c, *w.idx)
        st = gpulib.stats()
        assert st["ksplit_tiles"] > 0, st
        oracle.contract(al, ot[0], ot[1], be, ot[2], *w.idx)
        got, want = w.c.read(gpulib, c), w.c.read(oracle, ot[2])
        assert rel_err(got, want) <= 1e-13, (al, be)
    for x in ot:
        oracle.tensor_free(x)
    oracle.allocator_destroy(oal)
    for x in (t, v, c):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def test_split_alpha_beta_zero(gpulib, env):
    """beta == 0 writes the products without reading C (NaN in C
    vanishes); alpha == 0 writes beta * C exactly (signed zeros kept)."""
    env.pop("XM_GEMM_KSPLIT", None)
    w = _workload()
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    _fill(gpulib, w, (t, v, c), 41)
    gpulib.contract(0.5, t, v, 0.0, c, *w.idx)
    ref = w.c.read(gpulib, c)
    blocks = w.c.canonical_blocks()
    for k in blocks:
        x = np.full_like(np.asarray(ref[k]), np.nan)
        gpulib.write_block(c, k, x)
    gpulib.contract(0.5, t, v, 0.0, c, *w.idx)
    assert gpulib.stats()["ksplit_tiles"] > 0
    got = w.c.read(gpulib, c)
    for k in blocks:
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), k
    rng = np.random.default_rng(5)
    before = {}
    for k in blocks:
        x = rng.uniform(-1, 1, np.asarray(ref[k]).shape)
        x[x < -0.5] = -0.0
        gpulib.write_block(c, k, x)
        before[k] = x
    gpulib.contract(0.0, t, v, -1.5, c, *w.idx)
    got = w.c.read(gpulib, c)
    for k in blocks:
        want = -1.5 * before[k]
        g = np.asarray(got[k])
        assert np.array_equal(g, want), k
        assert np.array_equal(np.signbit(g), np.signbit(want)), k
    for x in (t, v, c):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def test_split_identical_across_kernels(gpulib, env):
    env.pop("XM_GEMM_KSPLIT", None)
    w = _workload()
    alloc = gpulib.allocator()
    out = {}
    for name, tile, dma, kernel in (
            ("64x128", "64x128", None, "k_gemm_f64_dma"),
            ("64x128 persistent", "64x128", "persist", "k_gemm_f64_dma_q"),
            ("64x64", "64x64", None, "k_gemm_f64"),
            ("64x128 registers", "64x128", "0", "k_gemm_f64")):
        env["XM_GEMM_TILE"] = tile
        env.pop("XM_GEMM_DMA", None)
        env.pop("XM_GEMM_PERSIST", None)
        if dma == "persist":
            env["XM_GEMM_PERSIST"] = "1"
        elif dma is not None:
            env["XM_GEMM_DMA"] = dma
        t, v, c = (s.build(gpulib, alloc) for s in w.specs)
        _fill(gpulib, w, (t, v, c), 51)
        gpulib.contract(0.75, t, v, 0.5, c, *w.idx)
        st = gpulib.stats()
        assert st["ksplit_tiles"] > 0, (name, st)
        assert "%dx%d" % (st["tile_p"], st["tile_q"]) == tile
        assert st["gemm_kernel"] == kernel, (name, st["gemm_kernel"])
        out[name] = w.c.read(gpulib, c)
        for x in (t, v, c):
            gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)
    first = out["64x128"]
    for name, r in out.items():
        for k in first:
            assert np.array_equal(np.asarray(first[k]), np.asarray(r[k])), \
                (name, k)


def test_split_with_c_panels(gpulib, oracle, env):
    """C(i,a,j,b) = sum_k A(i,k,j) B(k,a,b): C's A and B indices
    interleave, so C is computed in a panel and permuted back; K = 2048 in
    four blocks makes every C block split.  The sum kernel writes the
    panel, the fold the block."""
    env.pop("XM_GEMM_KSPLIT", None)
    kcut = [512, 1024, 1536]
    specs = (Spec((16, 2048, 8), {0: [8], 1: kcut}).all_canonical(),
             Spec((2048, 8, 16), {0: kcut}).all_canonical(),
             Spec((16, 8, 8, 16), {0: [8]}).all_canonical())
    idx = ("ikj", "kab", "iajb")
    rng = np.random.default_rng(81)
    for al, be in ((0.5, 1.0), (1.0, 0.0), (0.0, -2.0)):
        data = [sp.random_data(rng, -1, 1) for sp in specs]
        got, st = run_contract(gpulib, specs, data, al, be, idx, stats=True)
        assert st["ksplit_tiles"] > 0 and st["cout_bytes"] > 0, st
        want = run_contract(oracle, specs, data, al, be, idx)
        assert rel_err(got, want) <= 1e-13, (al, be)
        if al == 0.0:
            for k in want:
                assert np.array_equal(got[k], want[k]), k


@pytest.mark.parametrize("dma", [None, "0"], ids=["dma", "registers"])
def test_split_inside_pairs_of_odd_k(gpulib, oracle, env, dma):
    """K blocks of 300, 417, 555 and 776: the middle step falls inside a
    pair whose k is no multiple of 16, and both halves end in a partial
    step.  The DMA-staged and the register-staged kernel, against the
    oracle."""
    env.pop("XM_GEMM_KSPLIT", None)
    if dma is None:
        env.pop("XM_GEMM_DMA", None)
    else:
        env["XM_GEMM_DMA"] = dma
    kcut = [300, 717, 1272]
    specs = (Spec((16, 8, 2048), {0: [8], 2: kcut}).all_canonical(),
             Spec((2048, 8, 16), {0: kcut}).all_canonical(),
             Spec((16, 8, 8, 16), {0: [8]}).all_canonical())
    idx = ("ijk", "kab", "ijab")
    rng = np.random.default_rng(91)
    for al, be in ((0.75, -1.0), (1.0, 0.0)):
        data = [sp.random_data(rng, -1, 1) for sp in specs]
        got, st = run_contract(gpulib, specs, data, al, be, idx, stats=True)
        assert st["ksplit_tiles"] > 0 and st["cout_bytes"] == 0, st
        want = run_contract(oracle, specs, data, al, be, idx)
        assert rel_err(got, want) <= 1e-13, (al, be)


def test_split_off(gpulib, env):
    w = _workload()
    alloc = gpulib.allocator()
    res = {}
    for flag in ("0", "1"):
        env["XM_GEMM_KSPLIT"] = flag
        t, v, c = (s.build(gpulib, alloc) for s in w.specs)
        _fill(gpulib, w, (t, v, c), 61)
        gpulib.contract(1.25, t, v, -0.5, c, *w.idx)
        st = gpulib.stats()
        assert (st["ksplit_tiles"] > 0) == (flag == "1"), st
        res[flag] = w.c.read(gpulib, c)
        for x in (t, v, c):
            gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)
    assert rel_err(res["1"], res["0"]) <= 1e-13
