"""Split-precision (BF16 slices) FP64 contraction on the GPU, checked
against the FP64 CPU oracle: relative Frobenius error <= 1e-12 at the
5 slices, on ladders (operands read in place), the ring (operands
and C permuted), and data spread over six decades like the precision
study's inputs (U[-1,1] * 10^U[-6,0])."""
import numpy as np
import pytest

from libxm_amd import xm, workloads
from libxm_amd.spec import rel_err

from test_gpu_parity import run_contract

pytestmark = pytest.mark.gpu


@pytest.fixture
def split(gpulib):
    default = gpulib.lib.xm_get_split_slices()

    def use(n):
        gpulib.set_precision(xm.PRECISION_SPLIT_BF16, n)
    yield use
    gpulib.set_precision(xm.PRECISION_NATIVE, default)


def data_for(specs, seed, wide=False, pairs=None):
    rng = np.random.default_rng(seed)
    out = []
    for s in specs:
        d = s.random_data(rng, -1, 1)
        if wide:
            d = {k: v * 10.0 ** rng.uniform(-6, 0, v.size) for k, v in d.items()}
        if pairs:
            workloads.antisymmetrize_diagonal(s, d, pairs)
        out.append(d)
    return out


@pytest.mark.parametrize("w", [workloads.ladder_small(),
                               workloads.ladder_d2h(32, 96, 1),
                               workloads.ladder_d2h(48, 192, 3)],
                         ids=["ladder-o16-v64", "ladder-o32-v96",
                              "ladder-o48-v192"])
@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_split_ladder(w, beta, gpulib, oracle, split):
    data = data_for(w.specs, 11, pairs=w.pairs)
    want = run_contract(oracle, w.specs, data, w.alpha, beta, w.idx)
    split(5)
    got, st = run_contract(gpulib, w.specs, data, w.alpha, beta, w.idx, True)
    assert st["precision"] == xm.PRECISION_SPLIT_BF16 and st["nslices"] == 5
    assert rel_err(got, want) <= 1e-12


def test_split_wide_exponent_range(gpulib, oracle, split):
    """Entries spread over six decades: each chunk row is scaled by its
    largest entry, so small entries keep fewer bits; 6 slices keep the
    error below 1e-12 (5 slices land near 1e-12 on this data)."""
    w = workloads.ladder_d2h(32, 128, 2)
    data = data_for(w.specs, 12, wide=True, pairs=w.pairs)
    want = run_contract(oracle, w.specs, data, w.alpha, 0.0, w.idx)
    split(6)
    got = run_contract(gpulib, w.specs, data, w.alpha, 0.0, w.idx)
    assert rel_err(got, want) <= 1e-12


def test_split_ring_permuted_operands(gpulib, oracle, split):
    w = workloads.ring_d2h(24, 72, 5, lo=4, hi=12)
    data = data_for(w.specs, 13)
    want = run_contract(oracle, w.specs, data, 1.0, 0.5, w.idx)
    split(5)
    got, st = run_contract(gpulib, w.specs, data, 1.0, 0.5, w.idx, True)
    assert st["nunfold"] > 0
    assert rel_err(got, want) <= 1e-12


def test_split_error_falls_with_slices(gpulib, oracle, split):
    """More slices, smaller error: s = 3 is visibly inexact (~2^-27), the
    error then drops by orders of magnitude per slice."""
    w = workloads.ladder_d2h(32, 96, 4)
    data = data_for(w.specs, 14, pairs=w.pairs)
    want = run_contract(oracle, w.specs, data, w.alpha, 0.0, w.idx)
    errs = []
    for s in (3, 4, 5, 6):
        split(s)
        errs.append(rel_err(run_contract(gpulib, w.specs, data, w.alpha, 0.0,
                                         w.idx), want))
    assert errs[0] > 1e-10
    assert errs[0] > 100 * errs[1] > 100 * 100 * errs[2]
    assert errs[2] <= 1e-12 and errs[3] <= 1e-13


def test_split_alpha_zero_and_table_case(gpulib, oracle, split):
    from cases import abc_9, abc_12
    for fn, idx in ((abc_9, ("ijcd", "abcd", "ijab")),
                    (abc_12, ("afcb", "bace", "fe"))):
        specs = fn()
        data = data_for(specs, 15)
        split(5)
        for al, be in ((0.0, 0.5), (0.7, 0.0), (0.7, -1.25)):
            got = run_contract(gpulib, specs, data, al, be, idx)
            want = run_contract(oracle, specs, data, al, be, idx)
            assert rel_err(got, want) <= 1e-12, (fn.__name__, al, be)



@pytest.mark.parametrize("n", [7, 8])
@pytest.mark.parametrize("kernel", ["k_gemm_split2", "k_gemm_split"])
def test_split_many_slices(gpulib, oracle, split, monkeypatch, n, kernel):
    """7 and 8 slices on the LDS-staged kernel (two stages) and on the
    per-wave reference kernel (XM_SPLIT_KERNEL=1, fixed per plan), on
    six-decade data: both within 1e-12 of the oracle."""
    monkeypatch.setenv("XM_SPLIT_KERNEL", "1" if kernel == "k_gemm_split"
                       else "0")
    w = workloads.ladder_d2h(32, 128, 2)
    data = data_for(w.specs, 14 + n, wide=True, pairs=w.pairs)
    want = run_contract(oracle, w.specs, data, w.alpha, 0.0, w.idx)
    split(n)
    got, st = run_contract(gpulib, w.specs, data, w.alpha, 0.0, w.idx, True)
    assert st["gemm_kernel"] == kernel and st["nslices"] == n, st
    assert rel_err(got, want) <= 1e-12
