"""Tile-shape autotuning of the FP64 grouped GEMM.

A plan's first calls each time one tile-shape candidate (plans the DMA
kernel can run: the stacked 128 x 128 form, then 64 x 128; others: 64 x 128,
then 64 x 64); later calls use the faster.  Every candidate must give
bit-identical results (each output accumulates over the same k sequence in
the same order), a new plan whose C blocks have the same block-size
histogram starts with the decision, and XM_GEMM_TILE pins the shape.
"""
import os

import numpy as np
import pytest

from libxm_amd import workloads, xm

pytestmark = pytest.mark.gpu


def _fill(lib, w, tensors, seed):
    for k, x in enumerate(tensors):
        lib.fill_random(x, seed + k)
        for ia, ib in workloads.antisym_strings(4, getattr(w, "pairs", ())):
            lib.add(0.5, x, -0.5, x, ia, ib)


def _run(lib, w, t, v, c):
    lib.contract(w.alpha, t, v, 0.0, c, *w.idx)
    st = lib.stats()
    return st, w.c.read(lib, c)


@pytest.mark.parametrize("make,cands,classes", [
    (lambda: workloads.ladder_d2h(32, 96, 3), [(128, 128), (64, 128)], ""),
    (lambda: workloads.ring_d2h(24, 72, 5, lo=4, hi=12),
     [(64, 128), (64, 64)], "0"),
    (lambda: workloads.ring_d2h(24, 72, 5, lo=4, hi=12),
     [(128, 128), (64, 128)], "1"),
], ids=["ladder", "ring", "ring-classes"])
def test_tuning_sequence_and_identical_results(gpulib, make, cands, classes,
                                               monkeypatch):
    """The ring with each C block a GEMM of its own (odd extents: the
    register-staged kernel) and as dense classes (padded panels: the DMA
    kernels)."""
    os.environ.pop("XM_GEMM_TILE", None)
    monkeypatch.setenv("XM_CLASS_GEMM", classes)
    w = make()
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    _fill(gpulib, w, (t, v, c), 11)
    gpulib.set_precision(xm.PRECISION_NATIVE)
    runs = [_run(gpulib, w, t, v, c) for _ in range(4)]
    shapes = [(s["tile_p"], s["tile_q"]) for s, _ in runs]
    tuning = [s["tile_tuning"] for s, _ in runs]
    if tuning[0]:
        # a fresh signature: candidate 0 warms untimed, each candidate is
        # timed once, then the decision
        assert tuning == [1, 1, 1, 0], tuning
        assert shapes[:3] == [cands[0], cands[0], cands[1]], shapes
    assert shapes[3] in cands
    r1 = runs[0][1]
    for _, r in runs[1:]:
        for k in r1:
            assert np.array_equal(np.asarray(r1[k]), np.asarray(r[k]))
    shapes = shapes[1:]
    # same block structure, new tensors: a new plan that starts decided
    t2, v2, c2 = (s.build(gpulib, alloc) for s in w.specs)
    _fill(gpulib, w, (t2, v2, c2), 11)
    st4, r4 = _run(gpulib, w, t2, v2, c2)
    assert st4["plan_cached"] == 0 and st4["tile_tuning"] == 0
    assert (st4["tile_p"], st4["tile_q"]) == shapes[2]
    for k in r1:
        assert np.array_equal(np.asarray(r1[k]), np.asarray(r4[k]))
    for x in (t, v, c, t2, v2, c2):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def test_pinned_tile_shape(gpulib):
    """Each pinned shape runs alone; with the K split off (which the
    stacked form never uses) all three give the same bits."""
    w = workloads.ladder_d2h(32, 96, 4)
    alloc = gpulib.allocator()
    out = {}
    kernels = {"64x64": "k_gemm_f64", "64x128": "k_gemm_f64_dma",
               "128x128": "k_gemm_f64_dma_s"}
    try:
        os.environ["XM_GEMM_KSPLIT"] = "0"
        for shape in ("64x64", "64x128", "128x128"):
            os.environ["XM_GEMM_TILE"] = shape
            t, v, c = (s.build(gpulib, alloc) for s in w.specs)
            _fill(gpulib, w, (t, v, c), 21)
            st, r = _run(gpulib, w, t, v, c)
            assert st["tile_tuning"] == 0
            assert "%dx%d" % (st["tile_p"], st["tile_q"]) == shape
            assert st["gemm_kernel"] == kernels[shape], st["gemm_kernel"]
            out[shape] = r
            for x in (t, v, c):
                gpulib.tensor_free(x)
    finally:
        os.environ.pop("XM_GEMM_TILE", None)
        os.environ.pop("XM_GEMM_KSPLIT", None)
    for shape in ("64x128", "128x128"):
        for k in out["64x64"]:
            assert np.array_equal(np.asarray(out["64x64"][k]),
                                  np.asarray(out[shape][k])), (shape, k)
    gpulib.allocator_destroy(alloc)
