"""Host-buffer unfold and fold of blocks (libxm src/tensor.c:436-639) on
tensors whose blocks live in HBM: for canonical and derivative blocks of a
tensor with libxm's 8-permutation symmetry, every mask split of the
indices and a leading dimension larger than the row count, the bytes must
equal the oracle's, and folding an unfolded canonical block must give the
block back."""
import numpy as np
import pytest

from libxm_amd import xm

import cases
from helpers import TYPES, TYPE_NAMES, build_filled

pytestmark = pytest.mark.gpu

MASKS = [((0, 1), (2, 3)), ((2, 0), (3, 1)), ((3,), (1, 0, 2)),
         ((1, 3, 2, 0), ())]


def _unfold_all(lib, spec, data):
    alloc = lib.allocator()
    x = build_filled(lib, alloc, spec, data)
    out = {}
    for idx in spec.grid():
        kind = spec.kind(idx)
        if kind is None:
            continue
        raw = lib.read_block(x, idx)
        dims = spec.block_dims(idx)
        for mi, mj in MASKS:
            m = int(np.prod([dims[k] for k in mi]))
            n = int(np.prod([dims[k] for k in mj]))
            ld = m + 1
            mat = np.zeros(ld * n, dtype=raw.dtype)
            lib.lib.xm_tensor_unfold_block(x, xm.dim(idx), xm.dim(mi),
                                           xm.dim(mj), raw.ctypes.data,
                                           mat.ctypes.data, ld)
            out[(idx, mi, "unfold")] = mat
            if kind == "c":
                back = np.zeros_like(raw)
                lib.lib.xm_tensor_fold_block(x, xm.dim(idx), xm.dim(mi),
                                             xm.dim(mj), mat.ctypes.data,
                                             back.ctypes.data, ld)
                out[(idx, mi, "fold")] = back
                assert np.array_equal(back, raw), (idx, mi)
    lib.tensor_free(x)
    lib.allocator_destroy(alloc)
    return out


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
def test_unfold_fold_match_oracle(gpulib, oracle, t):
    spec = cases.abc_9(t)[1]
    data = spec.random_data(np.random.default_rng(5 + t), -1, 1)
    got = _unfold_all(gpulib, spec, data)
    want = _unfold_all(oracle, spec, data)
    assert got.keys() == want.keys()
    assert any(k[2] == "unfold" and spec.kind(k[0]) != "c" for k in got)
    for k in want:
        assert np.array_equal(got[k], want[k]), k
