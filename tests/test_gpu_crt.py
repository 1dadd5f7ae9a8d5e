"""The int8 CRT form of FP64 contraction (XM_PRECISION_SPLIT_INT8) on the
GPU, checked against the FP64 CPU oracle: relative Frobenius error <= 1e-12
with the moduli chosen per plan, on ladders (merged factors 2 and 1: one
Horner step per C block), entries spread over six decades, the ring
(operands and C permuted), alpha / beta edge cases, and libxm's
contraction table (cases whose factors are no integer chain run the BF16
form, which the stats report)."""
import numpy as np
import pytest

from libxm_amd import xm, workloads
from libxm_amd.spec import rel_err

from cases import CONTRACT_TESTS, case_id
from helpers import case_seed, magnitude_err
from test_gpu_parity import run_contract

pytestmark = pytest.mark.gpu


@pytest.fixture
def crt(gpulib):
    def use(nmod=0):
        gpulib.set_precision(xm.PRECISION_SPLIT_INT8, None, nmod)
    default = gpulib.lib.xm_get_split_slices()
    yield use
    gpulib.set_precision(xm.PRECISION_NATIVE, default, 0)


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
def test_crt_ladder(w, beta, gpulib, oracle, crt):
    data = data_for(w.specs, 21, pairs=w.pairs)
    want = run_contract(oracle, w.specs, data, w.alpha, beta, w.idx)
    crt()
    got, st = run_contract(gpulib, w.specs, data, w.alpha, beta, w.idx, True)
    assert st["precision"] == xm.PRECISION_SPLIT_INT8
    assert st["gemm_kernel"] == "k_gemm_crt"
    assert 8 <= st["nslices"] <= 13 and 0 < st["err_bound"] <= 1e-13
    e = rel_err(got, want)
    print(w.name, "beta", beta, "moduli", st["nslices"], "err %.2e" % e)
    assert e <= 1e-12


def test_crt_wide_exponent_range(gpulib, oracle, crt):
    """Six decades: rows are scaled by their 2-norm, so the error does not
    depend on the spread (the BF16 form needs 6 slices here)."""
    w = workloads.ladder_d2h(32, 128, 2)
    data = data_for(w.specs, 22, wide=True, pairs=w.pairs)
    want = run_contract(oracle, w.specs, data, w.alpha, 0.0, w.idx)
    crt()
    got = run_contract(gpulib, w.specs, data, w.alpha, 0.0, w.idx)
    assert rel_err(got, want) <= 1e-12


def test_crt_ring_permuted_operands(gpulib, oracle, crt):
    w = workloads.ring_d2h(24, 72, 5, lo=4, hi=12)
    data = data_for(w.specs, 23)
    want = run_contract(oracle, w.specs, data, 1.0, 0.5, w.idx)
    crt()
    got, st = run_contract(gpulib, w.specs, data, 1.0, 0.5, w.idx, True)
    assert st["precision"] == xm.PRECISION_SPLIT_INT8 and st["nunfold"] > 0
    assert rel_err(got, want) <= 1e-12


def test_crt_error_falls_with_moduli(gpulib, oracle, crt):
    """Each modulus adds ~8 bits to the product range, ~4 bits of
    precision per operand: the error falls ~16x per modulus."""
    w = workloads.ladder_d2h(32, 96, 4)
    data = data_for(w.specs, 24, pairs=w.pairs)
    want = run_contract(oracle, w.specs, data, w.alpha, 0.0, w.idx)
    errs = []
    for n in (8, 10, 12):
        crt(n)
        got, st = run_contract(gpulib, w.specs, data, w.alpha, 0.0, w.idx,
                               True)
        assert st["nslices"] == n
        errs.append(rel_err(got, want))
    print("moduli 8, 10, 12:", errs)
    assert errs[0] > 50 * errs[1] > 50 * 50 * errs[2]
    assert errs[2] <= 1e-12


def test_crt_alpha_beta_edges(gpulib, oracle, crt):
    w = workloads.ladder_d2h(32, 96, 6)
    data = data_for(w.specs, 25, pairs=w.pairs)
    crt()
    for al, be in ((0.0, 0.5), (0.0, 0.0), (0.7, 0.0), (-1.5, -1.25)):
        got = run_contract(gpulib, w.specs, data, al, be, w.idx)
        want = run_contract(oracle, w.specs, data, al, be, w.idx)
        assert rel_err(got, want) <= 1e-12, (al, be)
    # beta = 0 never reads C: NaN in C gives the product alone
    nan = [data[0], data[1], {k: np.full_like(v, np.nan)
                              for k, v in data[2].items()}]
    got = run_contract(gpulib, w.specs, nan, 0.5, 0.0, w.idx)
    want = run_contract(oracle, w.specs, data, 0.5, 0.0, w.idx)
    assert rel_err(got, want) <= 1e-12


@pytest.mark.parametrize("case", CONTRACT_TESTS, ids=case_id)
def test_crt_contract_table(case, gpulib, oracle, crt):
    fn, ia, ib, ic = case
    specs = fn(xm.SCALAR_DOUBLE)
    data = data_for(specs, case_seed(case_id(case)))
    want = run_contract(oracle, specs, data, 0.75, -0.5, (ia, ib, ic))
    crt()
    got, st = run_contract(gpulib, specs, data, 0.75, -0.5, (ia, ib, ic),
                           True)
    assert st["precision"] in (xm.PRECISION_SPLIT_INT8,
                               xm.PRECISION_SPLIT_BF16)
    # against the magnitudes: a C of one element, where the products
    # cancel to 1e-3 of their size, is 1e-12 off relative to itself
    assert magnitude_err(got, want, specs, data, 0.75, -0.5,
                         (ia, ib, ic)) <= 1e-12
