"""The DMA-staged GEMM forms against their alternatives.

FP32 and complex plans whose C blocks and panels allow it stage operands
with 16-byte DMAs (16 bytes of a column per lane) instead of dword DMAs;
FP64 plans with even extents run the DMA-staged 64 x 128 kernel instead
of the register-staged one.  Each pair runs the same MFMA sequence per output, so
results must be bit-identical (XM_GEMM_DMA=0 selects the alternative),
and match the oracle to the type's tolerance.  The kernel each call ran
(stats gemm_kernel) is checked too, and a plan with extents that are not
multiples of 4 must fall back to the dword form.
"""
import os

import numpy as np
import pytest

from libxm_amd import workloads, xm
from libxm_amd.spec import rel_err

from helpers import build_filled

pytestmark = pytest.mark.gpu


@pytest.fixture
def no_dma_env():
    old = os.environ.get("XM_GEMM_DMA")
    yield
    if old is None:
        os.environ.pop("XM_GEMM_DMA", None)
    else:
        os.environ["XM_GEMM_DMA"] = old


# the GEMM kernel per (type, tile, XM_GEMM_DMA)
KERNEL = {
    (xm.SCALAR_FLOAT, "64x128", "1"): "k_gemm_mfma_dw<x4>",
    (xm.SCALAR_FLOAT, "64x128", "0"): "k_gemm_mfma_dw",
    (xm.SCALAR_FLOAT, "64x64", "1"): "k_gemm_mfma_dw<x4>",
    (xm.SCALAR_FLOAT, "64x64", "0"): "k_gemm_mfma_dw",
    (xm.SCALAR_DOUBLE, "64x128", "1"): "k_gemm_f64_dma",
    (xm.SCALAR_DOUBLE, "64x128", "0"): "k_gemm_f64",
    (xm.SCALAR_DOUBLE, "64x64", "1"): "k_gemm_f64",
    (xm.SCALAR_DOUBLE, "64x64", "0"): "k_gemm_f64",
}


@pytest.mark.parametrize("t,tol", [(xm.SCALAR_FLOAT, 1e-5),
                                   (xm.SCALAR_DOUBLE, 1e-13)],
                         ids=["f32", "f64"])
@pytest.mark.parametrize("tile", ["64x128", "64x64"])
def test_dma_forms_identical(gpulib, oracle, no_dma_env, t, tol, tile,
                             monkeypatch):
    monkeypatch.setenv("XM_GEMM_TILE", tile)
    w = workloads.ladder_d2h(32, 96, 11, type_=t)
    run_forms(gpulib, oracle, w, t, tol, lambda dma: KERNEL[t, tile, dma])


@pytest.mark.parametrize("t,tol", [(xm.SCALAR_FLOAT_COMPLEX, 1e-5),
                                   (xm.SCALAR_DOUBLE_COMPLEX, 1e-13)],
                         ids=["c64", "c128"])
def test_dma_forms_complex(gpulib, oracle, no_dma_env, t, tol):
    """Complex plans: a 16-byte DMA moves two complex64 rows or one
    complex128 row of a column (64 x 64 tiles)."""
    w = workloads.ladder_d2h(32, 96, 11, type_=t)
    run_forms(gpulib, oracle, w, t, tol, lambda dma: "k_gemm_mfma_dw<x4>"
              if dma == "1" else "k_gemm_mfma_dw")


def test_x4_fallback(gpulib, oracle, no_dma_env, monkeypatch):
    """An FP32 ring with block extents of 4..10 (P or Q not divisible by
    4), each C block a GEMM of its own (a dense class's panels have
    leading dimensions of a multiple of 4): the dword form with DMA on or
    off, still matching the oracle."""
    monkeypatch.setenv("XM_CLASS_GEMM", "0")
    t = xm.SCALAR_FLOAT
    w = workloads.ring_d2h(20, 48, 3, type_=t, lo=4, hi=10)
    run_forms(gpulib, oracle, w, t, 1e-5, lambda dma: "k_gemm_mfma_dw")


def run_forms(gpulib, oracle, w, t, tol, kernel):
    data = None
    out = {}
    for dma in ("1", "0"):
        os.environ["XM_GEMM_DMA"] = dma
        alloc = gpulib.allocator()
        x = [s.build(gpulib, alloc) for s in w.specs]
        for k, z in enumerate(x):
            gpulib.fill_random(z, 90 + k)
        if data is None:
            data = [s.read(gpulib, z) for s, z in zip(w.specs, x)]
        gpulib.contract(0.5, x[0], x[1], 1.0, x[2], *w.idx)
        assert gpulib.stats()["gemm_kernel"] == kernel(dma), dma
        out[dma] = w.c.read(gpulib, x[2])
        for z in x:
            gpulib.tensor_free(z)
        gpulib.allocator_destroy(alloc)
    for k in out["1"]:
        assert np.array_equal(out["1"][k], out["0"][k]), k
    oal = oracle.allocator()
    ot = [build_filled(oracle, oal, s, d) for s, d in zip(w.specs, data)]
    oracle.contract(0.5, ot[0], ot[1], 1.0, ot[2], *w.idx)
    want = w.c.read(oracle, ot[2])
    for z in ot:
        oracle.tensor_free(z)
    oracle.allocator_destroy(oal)
    assert rel_err(out["1"], want) <= tol
