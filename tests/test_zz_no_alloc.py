"""xm_contract_noalloc, the contraction that allocates no device memory,
against xm_contract on the same inputs.

Two calls in a row, the second on a smaller contraction (so the shared
scratch array still holds the first call's descriptors), for the ladder
(operands read in place) and the ring (every operand and C block goes
through a permuted panel in the scratch array), each within 1e-13 of
xm_contract (FP64; the no-allocation path runs one tile shape and no K
split, so only the rounding of the K sums may differ) and against the
oracle.  A contraction whose panels exceed the array is refused with -1
and leaves C untouched.
"""
import numpy as np
import pytest

from libxm_amd import workloads, xm
from libxm_amd.spec import Spec, rel_err

from helpers import build_filled

pytestmark = pytest.mark.gpu


def _inputs(w, seed):
    rng = np.random.default_rng(seed)
    data = [s.random_data(rng) for s in w.specs]
    for s, d in zip(w.specs, data):
        workloads.antisymmetrize_diagonal(s, d, getattr(w, "pairs", ()))
    return data


def _run(lib, w, data, noalloc, alpha, beta):
    alloc = lib.allocator()
    x = [build_filled(lib, alloc, s, d) for s, d in zip(w.specs, data)]
    if noalloc:
        rc = lib.contract_noalloc(alpha, x[0], x[1], beta, x[2], *w.idx)
        assert rc == 0
    else:
        lib.contract(alpha, x[0], x[1], beta, x[2], *w.idx)
    out = w.c.read(lib, x[2])
    for z in x:
        lib.tensor_free(z)
    lib.allocator_destroy(alloc)
    return out


CASES = [
    ("ladder", (workloads.ladder_d2h(32, 96, 1), workloads.ladder_small())),
    ("ring", (workloads.ring_d2h(24, 64, 5, lo=4, hi=12),
              workloads.ring_d2h(16, 40, 6, lo=4, hi=10))),
]


@pytest.mark.parametrize("name, ws", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_contract(gpulib, oracle, name, ws):
    for k, w in enumerate(ws):          # the second call is smaller
        data = _inputs(w, 70 + k)
        for alpha, beta in ((0.5, 1.0), (0.75, 0.0)):
            got = _run(gpulib, w, data, True, alpha, beta)
            st = gpulib.stats()
            assert st["flops"] > 0 and st["gemm_launches"] == 1
            want = _run(gpulib, w, data, False, alpha, beta)
            ref = _run(oracle, w, data, False, alpha, beta)
            e1, e2 = rel_err(got, want), rel_err(got, ref)
            print("%s call %d (alpha %g beta %g): vs xm_contract %.2e, "
                  "vs oracle %.2e" % (name, k, alpha, beta, e1, e2))
            assert e1 <= 1e-13 and e2 <= 1e-13


def test_noalloc_refuses_oversize(gpulib):
    # c(i,j) = a(k,i) b(k,j) on 3072 x 3072 blocks: both operands need a
    # permuted (P-contiguous) 72 MiB panel, past the 64 MiB array
    sp = Spec((3072, 3072)).all_canonical()
    alloc = gpulib.allocator()
    a, b, c = (sp.build(gpulib, alloc) for _ in range(3))
    for k, x in enumerate((a, b, c)):
        gpulib.fill_random(x, 80 + k)
    before = sp.read(gpulib, c)
    assert gpulib.contract_noalloc(1.0, a, b, 0.0, c, "ki", "kj", "ij") == -1
    after = sp.read(gpulib, c)
    assert all(np.array_equal(before[i], after[i]) for i in before)
    # the allocating path runs it
    gpulib.contract(1.0, a, b, 0.0, c, "ki", "kj", "ij")
    assert gpulib.stats()["nunfold"] >= 1
    for x in (a, b, c):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)
