"""The stacked 128 x 128 form of the FP64 GEMM (k_gemm_f64_dma_s).

C blocks whose pairs read the same Y panels with the same extents and
factors share a tile's columns, their rows stacked in 32-row tiles; each
output still accumulates the same pairs in the same 16-k steps, so with the
K split off the form's results are bit-identical to the 64 x 128 DMA
kernel's, and they match the oracle to 1e-13.  The ladders below stack
C blocks of 144 to 400 rows (partial 32-row tiles, stacks of 2 and 4); the
C-panel case writes permuted C panels; alpha == 0 and beta == 0 keep
libxm's rules exactly.  The BF16 split form's 128 x 128 tiles stack the
same way (XM_SPLIT_STACK=0 keeps one C block a tile) and the int8 form's
can (XM_CRT_STACK=1); each output's sums are unchanged, so the same bits.
The stacked tiles of the groups that end the XCD lists run their K in two
halves; the 64 x 128 form of the same plan cuts the same C blocks.
"""
import os

import numpy as np
import pytest

from libxm_amd import workloads, xm
from libxm_amd.spec import Spec, rel_err

from helpers import build_filled
from test_gpu_parity import run_contract

pytestmark = pytest.mark.gpu


@pytest.fixture
def env():
    keys = ("XM_GEMM_KSPLIT", "XM_GEMM_TILE", "XM_CRT_STACK",
            "XM_SPLIT_STACK", "XM_CRT_MSPLIT")
    old = {k: os.environ.get(k) for k in keys}
    yield os.environ
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def _fill(lib, w, tensors, seed):
    for k, x in enumerate(tensors):
        lib.fill_random(x, seed + k)
        for ia, ib in workloads.antisym_strings(4, w.pairs):
            lib.add(0.5, x, -0.5, x, ia, ib)


@pytest.mark.parametrize("o,v,seed", [(32, 96, 7), (64, 128, 2)])
def test_stacked_bits_and_oracle(gpulib, oracle, env, o, v, seed):
    env["XM_GEMM_KSPLIT"] = "0"
    w = workloads.ladder_d2h(o, v, seed)
    out = {}
    data = None
    for tile, kernel in (("128x128", "k_gemm_f64_dma_s"),
                         ("64x128", "k_gemm_f64_dma")):
        env["XM_GEMM_TILE"] = tile
        alloc = gpulib.allocator()
        x = [s.build(gpulib, alloc) for s in w.specs]
        _fill(gpulib, w, x, 71)
        if data is None:
            data = [s.read(gpulib, z) for s, z in zip(w.specs, x)]
        gpulib.contract(0.5, x[0], x[1], 1.0, x[2], *w.idx)
        st = gpulib.stats()
        assert st["gemm_kernel"] == kernel, st["gemm_kernel"]
        assert (st["tile_p"], st["tile_q"]) == tuple(
            int(n) for n in tile.split("x"))
        out[tile] = w.c.read(gpulib, x[2])
        for z in x:
            gpulib.tensor_free(z)
        gpulib.allocator_destroy(alloc)
    for k in out["64x128"]:
        assert np.array_equal(out["128x128"][k], out["64x128"][k]), k
    oal = oracle.allocator()
    ot = [build_filled(oracle, oal, s, d) for s, d in zip(w.specs, data)]
    oracle.contract(0.5, ot[0], ot[1], 1.0, ot[2], *w.idx)
    want = w.c.read(oracle, ot[2])
    for z in ot:
        oracle.tensor_free(z)
    oracle.allocator_destroy(oal)
    assert rel_err(out["128x128"], want) <= 1e-13


def test_stacked_alpha_beta_zero(gpulib, env):
    """beta == 0 writes the products without reading C (NaN in C
    vanishes); alpha == 0 writes beta * C exactly (signed zeros kept)."""
    env["XM_GEMM_TILE"] = "128x128"
    w = workloads.ladder_d2h(32, 96, 7)
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    _fill(gpulib, w, (t, v, c), 41)
    gpulib.contract(0.5, t, v, 0.0, c, *w.idx)
    assert gpulib.stats()["gemm_kernel"] == "k_gemm_f64_dma_s"
    ref = w.c.read(gpulib, c)
    blocks = w.c.canonical_blocks()
    for k in blocks:
        gpulib.write_block(c, k, np.full_like(np.asarray(ref[k]), np.nan))
    gpulib.contract(0.5, t, v, 0.0, c, *w.idx)
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


def test_stacked_c_panels(gpulib, oracle, env):
    """C(i,a,j,b) = sum_k A(i,k,j) B(k,a,b): C is computed in panels and
    permuted back; the C blocks of one (a, b) stack."""
    env["XM_GEMM_TILE"] = "128x128"
    kcut = [512, 1024, 1536]
    specs = (Spec((48, 2048, 8), {0: [16, 40], 1: kcut}).all_canonical(),
             Spec((2048, 8, 16), {0: kcut}).all_canonical(),
             Spec((48, 8, 8, 16), {0: [16, 40]}).all_canonical())
    idx = ("ikj", "kab", "iajb")
    rng = np.random.default_rng(81)
    for al, be in ((0.5, 1.0), (1.0, 0.0), (0.0, -2.0)):
        data = [sp.random_data(rng, -1, 1) for sp in specs]
        got, st = run_contract(gpulib, specs, data, al, be, idx, stats=True)
        assert st["gemm_kernel"] == "k_gemm_f64_dma_s", st["gemm_kernel"]
        assert st["cout_bytes"] > 0, st
        want = run_contract(oracle, specs, data, al, be, idx)
        assert rel_err(got, want) <= 1e-13, (al, be)


def test_int8_stacks_identical(gpulib, env):
    """The int8 form with a group's C blocks stacked in shared tiles
    (XM_CRT_STACK=1) against one C block a tile, and with each tile's
    moduli in 1, 2 (default) or 3 parts (XM_CRT_MSPLIT): the residue
    products are exact integers, so the results are the same bits."""
    w = workloads.ladder_d2h(64, 128, 2)
    out = {}
    for flag, parts in (("0", "1"), ("1", "2"), ("0", "2"), ("0", "3")):
        env["XM_CRT_STACK"] = flag
        env["XM_CRT_MSPLIT"] = parts
        flag = flag + parts
        alloc = gpulib.allocator()
        x = [s.build(gpulib, alloc) for s in w.specs]
        _fill(gpulib, w, x, 17)
        gpulib.set_precision(xm.PRECISION_SPLIT_INT8, None, 0)
        gpulib.contract(0.5, x[0], x[1], 1.0, x[2], *w.idx)
        st = gpulib.stats()
        assert st["precision"] == xm.PRECISION_SPLIT_INT8, st["precision"]
        out[flag] = w.c.read(gpulib, x[2])
        for z in x:
            gpulib.tensor_free(z)
        gpulib.allocator_destroy(alloc)
    for f in ("12", "02", "03"):
        for k in out["01"]:
            assert np.array_equal(out["01"][k], out[f][k]), (f, k)


@pytest.mark.parametrize("nslices", [5, 6])
def test_split_stacks_identical(gpulib, oracle, env, nslices):
    """The BF16 split form with stacked tiles against one C block a tile:
    the same bits, within 1e-12 of the oracle; alpha == 0 keeps beta * C
    exactly."""
    w = workloads.ladder_d2h(64, 128, 2)
    out, data = {}, None
    for flag in ("1", "0"):
        env["XM_SPLIT_STACK"] = flag
        alloc = gpulib.allocator()
        x = [s.build(gpulib, alloc) for s in w.specs]
        _fill(gpulib, w, x, 23)
        if data is None:
            data = [s.read(gpulib, z) for s, z in zip(w.specs, x)]
        gpulib.set_precision(xm.PRECISION_SPLIT_BF16, nslices)
        gpulib.contract(0.5, x[0], x[1], 1.0, x[2], *w.idx)
        assert gpulib.stats()["gemm_kernel"] == "k_gemm_split2"
        out[flag] = w.c.read(gpulib, x[2])
        if flag == "1":
            before = w.c.read(gpulib, x[2])
            gpulib.contract(0.0, x[0], x[1], 2.0, x[2], *w.idx)
            got = w.c.read(gpulib, x[2])
            for k in before:
                assert np.array_equal(got[k], 2.0 * before[k]), k
        for z in x:
            gpulib.tensor_free(z)
        gpulib.allocator_destroy(alloc)
    for k in out["0"]:
        assert np.array_equal(out["0"][k], out["1"][k]), k
    oal = oracle.allocator()
    ot = [build_filled(oracle, oal, s, d) for s, d in zip(w.specs, data)]
    oracle.contract(0.5, ot[0], ot[1], 1.0, ot[2], *w.idx)
    want = w.c.read(oracle, ot[2])
    for z in ot:
        oracle.tensor_free(z)
    oracle.allocator_destroy(oal)
    assert rel_err(out["1"], want) <= 1e-12


def test_stacked_ksplit(gpulib, oracle, env):
    """K split of stacked tiles: the groups holding the last tiles of each
    XCD list run as two halves of their 16-k steps, summed by
    k_ksplit_sum_s.  The 64 x 128 candidate of the same plan cuts the same
    C blocks at the same step: a plan's calls on either give the same bits
    (beta = 0), within 1e-13 of the oracle; beta == 0 ignores NaN in C and
    alpha == 0 writes beta * C exactly."""
    env.pop("XM_GEMM_KSPLIT", None)
    env.pop("XM_GEMM_TILE", None)
    w = workloads.ladder_d2h(32, 192, 5)
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    _fill(gpulib, w, (t, v, c), 31)
    data = [s.read(gpulib, z) for s, z in zip(w.specs, (t, v, c))]
    outs, kernels = [], []
    for _ in range(4):
        gpulib.contract(0.5, t, v, 0.0, c, *w.idx)
        st = gpulib.stats()
        assert st["ksplit_tiles"] > 0, st
        kernels.append(st["gemm_kernel"])
        outs.append(w.c.read(gpulib, c))
    assert "k_gemm_f64_dma_s" in kernels and "k_gemm_f64_dma" in kernels, \
        kernels
    for o in outs[1:]:
        for k in outs[0]:
            assert np.array_equal(o[k], outs[0][k]), k
    oal = oracle.allocator()
    ot = [build_filled(oracle, oal, s, d) for s, d in zip(w.specs, data)]
    oracle.contract(0.5, ot[0], ot[1], 0.0, ot[2], *w.idx)
    want = w.c.read(oracle, ot[2])
    assert rel_err(outs[0], want) <= 1e-13
    # beta == 0 with NaN in C; alpha == 0 keeps beta * C
    for k in w.c.canonical_blocks():
        gpulib.write_block(c, k, np.full_like(np.asarray(want[k]), np.nan))
    gpulib.contract(0.5, t, v, 0.0, c, *w.idx)
    assert gpulib.stats()["gemm_kernel"] == kernels[-1]
    got = w.c.read(gpulib, c)
    for k in want:
        assert np.array_equal(got[k], outs[0][k]), k
    gpulib.contract(0.0, t, v, -2.0, c, *w.idx)
    got2 = w.c.read(gpulib, c)
    for k in want:
        assert np.array_equal(got2[k], -2.0 * got[k]), k
    for z in ot:
        oracle.tensor_free(z)
    oracle.allocator_destroy(oal)
    for x in (t, v, c):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)
