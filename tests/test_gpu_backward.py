"""xm_contract_backward on the GPU.

Checked three ways:
  * against the gradient of the dense formula, computed in numpy: G_a =
    alpha * einsum(dc', b) with dc' holding dc's canonical blocks only,
    then every derivative block X of a folded into its source S as
    s_X * transpose(G_a[X], p_X^-1);
  * against central finite differences (h = 1e-6) of
    L = sum over canonical c of dc * c, the forward run by the CPU oracle,
    on 64 seeded canonical entries of a and of b;
  * split precision against the FP64 adjoint (<= 1e-12).
"""
import numpy as np
import pytest

from libxm_amd import xm, workloads
from libxm_amd.spec import rel_err

from cases import abc_9, abc_10, abc_11, abc_12
from helpers import build_filled

pytestmark = pytest.mark.gpu


def canonical_dense(spec, data):
    out = np.zeros(spec.dims)
    for idx in spec.canonical_blocks():
        bd = spec.block_dims(idx)
        off = spec.block_offsets(idx)
        out[tuple(slice(o, o + d) for o, d in zip(off, bd))] = \
            np.asarray(data[idx]).reshape(bd, order="F")
    return out


def fold(spec, G):
    """Gradient w.r.t. the canonical data of `spec` from the gradient G
    w.r.t. every (logical) element of the dense tensor."""
    out = {idx: np.zeros(spec.block_size(idx)) for idx in spec.canonical_blocks()}
    for idx in spec.grid():
        kind = spec.kind(idx)
        if kind is None:
            continue
        bd = spec.block_dims(idx)
        off = spec.block_offsets(idx)
        g = G[tuple(slice(o, o + d) for o, d in zip(off, bd))]
        if kind == "c":
            out[idx] += g.ravel(order="F")
        else:
            _, src, perm, s = kind
            out[src] += (s.real * np.transpose(g, np.argsort(perm))).ravel(order="F")
    return out


def dense_adjoint(specs, data, dcd, alpha, idx):
    sa, sb, sc = specs
    ia, ib, ic = idx
    A, B = sa.dense(data[0]), sb.dense(data[1])
    dC = canonical_dense(sc, dcd)
    GA = alpha * np.einsum("%s,%s->%s" % (ic, ib, ia), dC, B)
    GB = alpha * np.einsum("%s,%s->%s" % (ia, ic, ib), A, dC)
    # only the nonzero blocks of a / b are variables
    return fold(sa, GA), fold(sb, GB)


def gpu_adjoint(lib, specs, data, dcd, alpha, idx):
    sa, sb, sc = specs
    alloc = lib.allocator()
    a = build_filled(lib, alloc, sa, data[0])
    b = build_filled(lib, alloc, sb, data[1])
    dc = build_filled(lib, alloc, sc, dcd)
    da, db = sa.build(lib, alloc), sb.build(lib, alloc)
    lib.contract_backward(alpha, a, b, dc, da, db, *idx)
    out = sa.read(lib, da), sb.read(lib, db), lib.stats()
    for t in (a, b, dc, da, db):
        lib.tensor_free(t)
    lib.allocator_destroy(alloc)
    return out


def loss(lib, specs, data, dcd, alpha, idx):
    """L = sum over canonical c of dc * c, c = alpha * a * b (beta = 0)."""
    sa, sb, sc = specs
    alloc = lib.allocator()
    a = build_filled(lib, alloc, sa, data[0])
    b = build_filled(lib, alloc, sb, data[1])
    c = sc.build(lib, alloc)
    lib.contract(alpha, a, b, 0.0, c, *idx)
    out = sc.read(lib, c)
    for t in (a, b, c):
        lib.tensor_free(t)
    lib.allocator_destroy(alloc)
    return sum(float(np.dot(out[k], dcd[k])) for k in out)


CASES = [
    ("ladder-C2", lambda: workloads.ladder_small().specs, workloads.LADDER_IDX),
    ("abc9-ijab", abc_9, ("ijab", "abcd", "ijcd")),
    ("abc10", abc_10, ("cdab", "ijab", "ijcd")),
    ("abc11", abc_11, ("abcd", "ijcd", "ijab")),
    ("abc12", abc_12, ("afcb", "bace", "fe")),
    ("ring", lambda: workloads.ring_d2h(20, 48, 3, lo=4, hi=10).specs,
     workloads.RING_IDX),
]


@pytest.mark.parametrize("name, make, idx", CASES, ids=[c[0] for c in CASES])
def test_adjoint_vs_dense_gradient(name, make, idx, gpulib):
    specs = make()
    rng = np.random.default_rng(3)
    data = [s.random_data(rng) for s in specs[:2]]
    dcd = specs[2].random_data(rng)
    want_a, want_b = dense_adjoint(specs, data, dcd, 0.75, idx)
    got_a, got_b, st = gpu_adjoint(gpulib, specs, data, dcd, 0.75, idx)
    assert rel_err(got_a, want_a) <= 1e-13
    assert rel_err(got_b, want_b) <= 1e-13
    assert st["flops"] > 0


def test_adjoint_finite_differences(gpulib, oracle):
    """Central differences of L through the CPU oracle's forward on 64
    seeded canonical entries of t and of v (ladder, D2h)."""
    w = workloads.ladder_d2h(24, 64, 6)
    specs = w.specs
    rng = np.random.default_rng(8)
    data = [s.random_data(rng) for s in specs[:2]]
    dcd = specs[2].random_data(rng)
    got_a, got_b, _ = gpu_adjoint(gpulib, specs, data, dcd, w.alpha, w.idx)
    h = 1e-6
    for which, got in ((0, got_a), (1, got_b)):
        spec = specs[which]
        blocks = spec.canonical_blocks()
        for _ in range(32):
            blk = blocks[rng.integers(len(blocks))]
            e = int(rng.integers(spec.block_size(blk)))
            vals = []
            for sgn in (1, -1):
                d = [dict(x) for x in data]
                d[which][blk] = d[which][blk].copy()
                d[which][blk][e] += sgn * h
                vals.append(loss(oracle, specs, d, dcd, w.alpha, w.idx))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert abs(fd - got[blk][e]) <= 1e-6 * max(1.0, abs(fd)), \
                (which, blk, e, fd, got[blk][e])


def test_adjoint_split_precision(gpulib):
    w = workloads.ladder_d2h(32, 96, 7)
    specs = w.specs
    rng = np.random.default_rng(9)
    data = [s.random_data(rng, -1, 1) for s in specs[:2]]
    for s, d in zip(specs[:2], data):
        workloads.antisymmetrize_diagonal(s, d, w.pairs)
    dcd = specs[2].random_data(rng, -1, 1)
    ref_a, ref_b, _ = gpu_adjoint(gpulib, specs, data, dcd, w.alpha, w.idx)
    gpulib.set_precision(xm.PRECISION_SPLIT_BF16, 5)
    try:
        got_a, got_b, st = gpu_adjoint(gpulib, specs, data, dcd, w.alpha, w.idx)
    finally:
        gpulib.set_precision(xm.PRECISION_NATIVE)
    assert st["precision"] == xm.PRECISION_SPLIT_BF16
    assert rel_err(got_a, ref_a) <= 1e-12
    assert rel_err(got_b, ref_b) <= 1e-12


def test_adjoint_one_side_only(gpulib):
    specs = workloads.ladder_small().specs
    rng = np.random.default_rng(4)
    data = [s.random_data(rng) for s in specs[:2]]
    dcd = specs[2].random_data(rng)
    want_a, _ = dense_adjoint(specs, data, dcd, 0.5, workloads.LADDER_IDX)
    lib = gpulib
    alloc = lib.allocator()
    a = build_filled(lib, alloc, specs[0], data[0])
    b = build_filled(lib, alloc, specs[1], data[1])
    dc = build_filled(lib, alloc, specs[2], dcd)
    da = specs[0].build(lib, alloc)
    lib.lib.xm_contract_backward(xm.scalar(0.5), a, b, dc, da, None,
                                 *[s.encode() for s in workloads.LADDER_IDX])
    assert rel_err(specs[0].read(lib, da), want_a) <= 1e-13
    for t in (a, b, dc, da):
        lib.tensor_free(t)
    lib.allocator_destroy(alloc)


def test_adjoint_repeated_calls(gpulib):
    """Repeated calls reuse their auxiliary tensors and plans: each call
    must follow the current data of a, b and dc, and a structure change of
    da (a block made zero, then canonical again) or a new dc tensor, against
    the dense gradient every time."""
    specs = workloads.ladder_small().specs
    idx = workloads.LADDER_IDX
    rng = np.random.default_rng(12)
    lib = gpulib
    alloc = lib.allocator()
    data = [s.random_data(rng) for s in specs[:2]]
    dcd = specs[2].random_data(rng)
    a = build_filled(lib, alloc, specs[0], data[0])
    b = build_filled(lib, alloc, specs[1], data[1])
    dc = build_filled(lib, alloc, specs[2], dcd)
    da, db = specs[0].build(lib, alloc), specs[1].build(lib, alloc)
    for step in range(4):
        if step == 1:           # new values in a, b and dc, same tensors
            data = [s.random_data(rng) for s in specs[:2]]
            dcd = specs[2].random_data(rng)
            specs[0].write(lib, a, data[0])
            specs[1].write(lib, b, data[1])
            specs[2].write(lib, dc, dcd)
        if step == 2:           # da's structure changes (and back)
            blk = specs[0].canonical_blocks()[0]
            lib.set_zero(da, blk)
            lib.set_canonical(da, blk)
        if step == 3:           # a new dc tensor
            lib.tensor_free(dc)
            dcd = specs[2].random_data(rng)
            dc = build_filled(lib, alloc, specs[2], dcd)
        lib.contract_backward(0.75, a, b, dc, da, db, *idx)
        want_a, want_b = dense_adjoint(specs, data, dcd, 0.75, idx)
        assert rel_err(specs[0].read(lib, da), want_a) <= 1e-13, step
        assert rel_err(specs[1].read(lib, db), want_b) <= 1e-13, step
    for t in (a, b, dc, da, db):
        lib.tensor_free(t)
    lib.allocator_destroy(alloc)
