"""Dense classes: C blocks whose terms form one GEMM of stacked blocks.

The ring c(i,j,a,b) = sum t(i,k,a,c) w(k,b,c,j) over tensors that store
every symmetry-allowed block is one GEMM per irrep of (i,a): the X blocks
t(i,k,a,c) sit at row (i,a) and column (k,c) of one panel, the Y blocks
w(k,b,c,j) at (k,c), (j,b) of another, and C is computed in one panel and
folded back.  Each term's K is padded with zeros to a multiple of 4, so
every output sums the same 4-k MFMA groups in the same order as when each
term is a pair of its own: the class form gives the same bits as the
per-C-block form (XM_CLASS_GEMM=0) and matches the oracle.  Also covered:
alpha == 0 / beta == 0 rules, classes forced (XM_CLASS_GEMM=1) where the
per-C-block form reads operands and C in place, the swapped orientation
(C's b indices first), a zero block, which splits a class, and seeded
random structures.
"""
import os

import numpy as np
import pytest

from libxm_amd import workloads, xm
from libxm_amd.spec import Spec, rel_err

from helpers import build_filled, round_scalar

pytestmark = pytest.mark.gpu

TOL = {xm.SCALAR_FLOAT: 1e-5, xm.SCALAR_FLOAT_COMPLEX: 1e-5,
       xm.SCALAR_DOUBLE: 1e-13, xm.SCALAR_DOUBLE_COMPLEX: 1e-13}
NAMES = {xm.SCALAR_FLOAT: "f32", xm.SCALAR_FLOAT_COMPLEX: "c64",
         xm.SCALAR_DOUBLE: "f64", xm.SCALAR_DOUBLE_COMPLEX: "c128"}


@pytest.fixture
def env():
    keys = ("XM_CLASS_GEMM", "XM_GEMM_KSPLIT")
    old = {k: os.environ.get(k) for k in keys}
    # the per-C-block form's K split sums halves: other bits
    os.environ["XM_GEMM_KSPLIT"] = "0"
    yield os.environ
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def _data(specs, seed):
    rng = np.random.default_rng(seed)
    return [s.random_data(rng, -1, 1) for s in specs]


def _contract(lib, specs, data, al, be, idx, cdata=None, stats=False):
    """C after al * a b + be * C (C from cdata when given), and the
    call's stats when asked."""
    alloc = lib.allocator()
    x = [build_filled(lib, alloc, s, d) for s, d in zip(specs, data)]
    if cdata is not None:
        for k, v in cdata.items():
            lib.write_block(x[2], k, v)
    lib.contract(al, x[0], x[1], be, x[2], *idx)
    st = lib.stats() if stats else None
    out = specs[2].read(lib, x[2])
    for z in x:
        lib.tensor_free(z)
    lib.allocator_destroy(alloc)
    return out, st


def _forms(gpulib, specs, data, al, be, idx, mode, cdata=None):
    """Results with classes (mode) and without, and the class stats."""
    os.environ["XM_CLASS_GEMM"] = mode
    got, st = _contract(gpulib, specs, data, al, be, idx, cdata, True)
    os.environ["XM_CLASS_GEMM"] = "0"
    ref, st0 = _contract(gpulib, specs, data, al, be, idx, cdata, True)
    assert st0["classes"] == 0, st0["classes"]
    return got, ref, st


def _same_bits(got, ref):
    for k in ref:
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), k


@pytest.mark.parametrize("t", list(NAMES), ids=list(NAMES.values()))
def test_ring_classes(gpulib, oracle, env, t):
    w = workloads.ring_d2h(24, 72, 5, t, lo=4, hi=12)
    data = _data(w.specs, 3)
    al, be = round_scalar(0.75, t), round_scalar(-0.5, t)
    want, _ = _contract(oracle, w.specs, data, al, be, w.idx)
    # by default the classes that cut the padded work by 8% or more (a
    # plan of classes and single C blocks); forced, one class per irrep of
    # (i, a), every C block in one
    for mode in ("", "1"):
        got, ref, st = _forms(gpulib, w.specs, data, al, be, w.idx, mode)
        if mode:
            assert st["classes"] == 8, st
            assert st["class_cblocks"] == st["nblocks_c"], st
        else:
            assert 0 < st["class_cblocks"] < st["nblocks_c"], st
        _same_bits(got, ref)
        assert rel_err(got, want) <= TOL[t]


def test_ring_class_alpha_beta_zero(gpulib, env):
    """beta == 0 writes the products without reading C (NaN in C
    vanishes); alpha == 0 writes beta * C exactly (signed zeros kept)."""
    t = xm.SCALAR_DOUBLE
    w = workloads.ring_d2h(24, 60, 7, t, lo=4, hi=12)
    data = _data(w.specs, 5)
    got, ref, st = _forms(gpulib, w.specs, data, 1.0, 0.0, w.idx, "")
    assert st["classes"] > 0
    _same_bits(got, ref)
    nan = {k: np.full_like(np.asarray(v), np.nan) for k, v in data[2].items()}
    got2, _, _ = _forms(gpulib, w.specs, data, 1.0, 0.0, w.idx, "", nan)
    _same_bits(got2, ref)
    rng = np.random.default_rng(9)
    cz = {}
    for k, v in data[2].items():
        x = rng.uniform(-1, 1, np.asarray(v).shape)
        x[x < -0.5] = -0.0
        cz[k] = x
    got3, ref3, st3 = _forms(gpulib, w.specs, data, 0.0, -1.5, w.idx, "", cz)
    assert st3["classes"] > 0
    for k in cz:
        want = -1.5 * cz[k]
        g = np.asarray(got3[k])
        assert np.array_equal(g, want), k
        assert np.array_equal(np.signbit(g), np.signbit(want)), k
    _same_bits(got3, ref3)


@pytest.mark.parametrize("t", [xm.SCALAR_DOUBLE, xm.SCALAR_FLOAT_COMPLEX],
                         ids=["f64", "c64"])
def test_forced_class_in_place(gpulib, oracle, env, t):
    """c(i,a) = 0.7 t(i,c,d) v(d,c,a) + 0.3 c: t and C could be read in
    place, so only XM_CLASS_GEMM=1 forms the class (2 x 2 C blocks over
    4 terms); t's blocks and C are then copied into and out of the class
    panels.  c(a,i) as well: C's b index first (the swapped orientation)."""
    st_ = Spec((16, 160, 160), {0: [8], 1: [80], 2: [80]}, t).all_canonical()
    sv = Spec((160, 160, 24), {0: [80], 1: [80], 2: [12]}, t).all_canonical()
    for sc, ic in ((Spec((16, 24), {0: [8], 1: [12]}, t).all_canonical(),
                    "ia"),
                   (Spec((24, 16), {0: [12], 1: [8]}, t).all_canonical(),
                    "ai")):
        specs = (st_, sv, sc)
        data = _data(specs, 11)
        idx = ("icd", "dca", ic)
        al, be = round_scalar(0.7, t), round_scalar(0.3, t)
        os.environ["XM_CLASS_GEMM"] = ""
        _, st_def = _contract(gpulib, specs, data, al, be, idx,
                              stats=True)
        assert st_def["classes"] == 0, st_def
        got, ref, st = _forms(gpulib, specs, data, al, be, idx, "1")
        assert st["classes"] == 1 and st["class_cblocks"] == 4, st
        _same_bits(got, ref)
        want, _ = _contract(oracle, specs, data, al, be, idx)
        assert rel_err(got, want) <= TOL[t], ic


def test_forced_class_swapped_ring(gpulib, oracle, env):
    """The ring with c stored (j,b,i,a): the C blocks' b indices come
    first, so X is w's block and Y t's; C is read in place by the
    per-C-block form, so XM_CLASS_GEMM=1 forms the classes."""
    t = xm.SCALAR_DOUBLE
    w = workloads.ring_d2h(24, 72, 5, t, lo=4, hi=12)
    o, oi, v, vi = w.o_sizes, w.o_irreps, w.v_sizes, w.v_irreps
    sc = workloads.block_tensor([o, v, o, v], [oi, vi, oi, vi], (), t)
    specs = (w.t, w.v, sc)
    idx = ("ikac", "kbcj", "jbia")
    data = _data(specs, 13)
    got, ref, st = _forms(gpulib, specs, data, 1.0, 0.5, idx, "1")
    assert st["classes"] == 8, st
    _same_bits(got, ref)
    want, _ = _contract(oracle, specs, data, 1.0, 0.5, idx)
    assert rel_err(got, want) <= 1e-13


def test_zero_block_splits_class(gpulib, oracle, env):
    """C(i,a) = sum_k A(i,k) B(k,a) with block A(i1, k1) zero: the C
    blocks of row i1 have one term, those of row i0 two, so they form
    two classes (XM_CLASS_GEMM=1) of one row each."""
    t = xm.SCALAR_DOUBLE
    sa = Spec((40, 56), {0: [24], 1: [20]}, t)
    for idx in sa.grid():
        if idx != (1, 1):
            sa.canon(idx)
    sb = Spec((56, 36), {0: [20], 1: [16]}, t).all_canonical()
    sc = Spec((40, 36), {0: [24], 1: [16]}, t).all_canonical()
    specs = (sa, sb, sc)
    data = _data(specs, 17)
    got, ref, st = _forms(gpulib, specs, data, 1.0, 1.0, ("ik", "ka", "ia"),
                          "1")
    assert st["classes"] == 2 and st["class_cblocks"] == 4, st
    _same_bits(got, ref)
    want, _ = _contract(oracle, specs, data, 1.0, 1.0, ("ik", "ka", "ia"))
    assert rel_err(got, want) <= 1e-13


def _random_case(rng, t):
    """A seeded contraction with uneven blocks, permuted storage and some
    zero blocks (which break or split classes)."""
    tmpl = [("ikac", "kbcj", "ijab"), ("ikac", "kbcj", "jbia"),
            ("ick", "kca", "ia"), ("kic", "ack", "ai"), ("ik", "ka", "ia")]
    ia, ib, ic = tmpl[rng.integers(len(tmpl))]
    letters = sorted(set(ia + ib + ic))
    ext = {x: int(rng.integers(12, 40)) for x in letters}
    cuts = {}
    for x in letters:
        n = int(rng.integers(1, 4))
        pts = sorted(set(int(v) for v in rng.integers(1, ext[x], n)))
        cuts[x] = pts

    def spec(idx, zero_frac):
        s = Spec([ext[x] for x in idx], {k: cuts[x] for k, x in
                                         enumerate(idx)}, t)
        for blk in s.grid():
            if rng.random() >= zero_frac:
                s.canon(blk)
        return s

    return (spec(ia, 0.15), spec(ib, 0.15), spec(ic, 0.0)), (ia, ib, ic)


@pytest.mark.parametrize("seed", range(12))
def test_random_structures(gpulib, oracle, env, seed):
    """Seeded random block structures (uneven extents, zero blocks,
    permuted storage of operands and C) in all four types: classes
    forced and not give the same bits and match the oracle."""
    rng = np.random.default_rng(1000 + seed)
    t = list(NAMES)[seed % 4]
    specs, idx = _random_case(rng, t)
    data = _data(specs, seed)
    al = round_scalar(complex(0.7, 0.2), t)
    be = round_scalar(complex(-0.4, 0.1), t)
    got, ref, st = _forms(gpulib, specs, data, al, be, idx, "1")
    # forced, every C block with terms is in a class (one of its own at
    # least: its terms then run as one K-concatenated pair)
    assert st["classes"] > 0, st
    _same_bits(got, ref)
    want, _ = _contract(oracle, specs, data, al, be, idx)
    assert rel_err(got, want) <= TOL[t], (idx, st["classes"])
