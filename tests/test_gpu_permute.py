"""LDS-tiled permutes and the element-operation job cache, against the CPU
oracle.

The permute kernels move a source block into the destination's order and
apply the block's permutation and scalar on the way (libxm
src/tensor.c:507-639 for the unfold, src/xm.c:66-413 for the element
operations).  Jobs whose source-fastest dim lies far from the destination's
first dim run the LDS-tiled kernel; the cases below are built so that they
do (checked through xm_amd_last_permute and the contraction stats), with
blocks of up to 12^4 elements and derivative sources whose permutation
moves every dim.  Results must match the oracle bit for bit (complex
division to 1e-15 relative, 1e-7 for complex64: its quotient formula
differs).

Repeated calls launch cached job lists; the sequences below change the
factor, the structure of either tensor, and the tensors themselves between
calls, and must still match the oracle run through the same sequence.
"""
import numpy as np
import pytest

from libxm_amd import xm
from libxm_amd.spec import Spec, rel_err

from helpers import TYPES, TYPE_NAMES, build_filled, round_scalar

pytestmark = pytest.mark.gpu

SPLIT = [12, 21]        # blocks of 12, 9, 9 per dim
FAR = ("lkji", "klij", "ljki")
NEAR = ("ijkl", "jikl", "kjil")


def b_spec(t, zeros=True):
    """30^4 tensor: block (i,j,k,l) is a derivative of (k,l,i,j) with
    permutation (2,3,0,1) and scalar -0.5 where (i,j,k,l) > (k,l,i,j);
    blocks with i+j+k+l == 5 are zero (with zeros)."""
    sp = Spec((30,) * 4, {d: SPLIT for d in range(4)}, t)
    later = []
    for idx in sp.grid():
        if zeros and sum(idx) == 5:
            continue
        mirror = idx[2:] + idx[:2]
        if idx > mirror:
            later.append((idx, mirror))
        else:
            sp.canon(idx)
    for idx, mirror in later:
        sp.deriv(idx, mirror, (2, 3, 0, 1), -0.5)
    return sp


def a_spec(t):
    return Spec((30,) * 4, {d: SPLIT for d in range(4)}, t).all_canonical()


def data_for(specs, seed, shift=0.0):
    rng = np.random.default_rng(seed)
    out = []
    for s in specs:
        d = s.random_data(rng, -1.0, 1.0)
        out.append({k: v + shift for k, v in d.items()})
    return out


def same(x, y, t, div=False):
    if div and t in (xm.SCALAR_FLOAT_COMPLEX, xm.SCALAR_DOUBLE_COMPLEX):
        return rel_err(x, y) <= (1e-7 if t == xm.SCALAR_FLOAT_COMPLEX
                                 else 1e-15)
    return all(np.array_equal(x[k].view(np.uint8), y[k].view(np.uint8))
               for k in y)


def run_op(lib, op, t, perm, data, zeros=True):
    sa, sb = a_spec(t), b_spec(t, zeros)
    alloc = lib.allocator()
    a = build_filled(lib, alloc, sa, data[0])
    b = build_filled(lib, alloc, sb, data[1])
    if op == "copy":
        lib.copy(a, round_scalar(0.75, t), b, "ijkl", perm)
    elif op == "add":
        lib.add(round_scalar(0.5, t), a, round_scalar(-1.25, t), b,
                "ijkl", perm)
    else:
        getattr(lib, op)(a, b, "ijkl", perm)
    used = lib.last_permute() if lib.has("xm_amd_last_permute") else None
    out = sa.read(lib, a)
    lib.tensor_free(a)
    lib.tensor_free(b)
    lib.allocator_destroy(alloc)
    return out, used


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
@pytest.mark.parametrize("perm", FAR + NEAR)
@pytest.mark.parametrize("op", ["copy", "add", "mul", "div"])
def test_tiled_permute_ops(op, perm, t, gpulib, oracle):
    zeros = op != "div"     # a zero divisor block is fatal in libxm
    data = data_for((a_spec(t), b_spec(t, zeros)), 7 + t,
                    shift=2.0 if op == "div" else 0.0)
    got, used = run_op(gpulib, op, t, perm, data, zeros)
    want, _ = run_op(oracle, op, t, perm, data, zeros)
    assert same(got, want, t, div=op == "div"), (op, perm)
    tiled, gather = used
    assert tiled + gather == len(a_spec(t).canonical_blocks())
    if perm in FAR:
        assert tiled > 0, used


@pytest.mark.parametrize("t", [xm.SCALAR_DOUBLE, xm.SCALAR_FLOAT_COMPLEX],
                         ids=["f64", "c64"])
def test_contraction_with_tiled_unfold(t, gpulib, oracle):
    """c(i,a) = 0.7 t(i,c,d) v(d,c,a) + 0.3 c: v's panel lists (c, d)
    while v stores d fastest, with 80 x 80 (c, d) blocks."""
    st = Spec((16, 160, 160), {0: [8], 1: [80], 2: [80]}, t).all_canonical()
    sv = Spec((160, 160, 24), {0: [80], 1: [80], 2: [12]}, t).all_canonical()
    sc = Spec((16, 24), {0: [8], 1: [12]}, t).all_canonical()
    data = data_for((st, sv, sc), 11)
    outs = []
    for lib in (gpulib, oracle):
        alloc = lib.allocator()
        x = [build_filled(lib, alloc, s, d) for s, d in zip((st, sv, sc), data)]
        lib.contract(round_scalar(0.7, t), x[0], x[1], round_scalar(0.3, t),
                     x[2], "icd", "dca", "ia")
        if lib is gpulib:
            stats = lib.stats()
        outs.append(sc.read(lib, x[2]))
        for y in x:
            lib.tensor_free(y)
        lib.allocator_destroy(alloc)
    assert stats["permute_tiled"] > 0, stats
    tol = 1e-13 if t == xm.SCALAR_DOUBLE else 1e-5
    assert rel_err(outs[0], outs[1]) <= tol


@pytest.mark.parametrize("t", [xm.SCALAR_DOUBLE, xm.SCALAR_DOUBLE_COMPLEX],
                         ids=["f64", "c128"])
def test_cached_jobs_follow_factor_and_structure(t, gpulib, oracle):
    """The same call sequence on both libraries: repeated adds with changed
    factors, a structure change of b (a canonical block set to zero), a
    zero block of a made canonical, then xm_set and xm_copy repeated across
    the same changes."""
    sa = a_spec(t)
    sb = b_spec(t)
    data = data_for((sa, sb), 19)
    z = (0, 1, 2, 2)            # zero in b's recipe (sum 5)
    k = (1, 2, 1, 2)            # canonical in b, no derivative reads it
    outs = []
    for lib in (gpulib, oracle):
        alloc = lib.allocator()
        a = build_filled(lib, alloc, sa, data[0])
        b = build_filled(lib, alloc, sb, data[1])
        seq = []
        for al, be in ((0.5, 1.0), (0.5, -2.0), (1.0, -2.0), (0.0, 0.25)):
            lib.add(round_scalar(al, t), a, round_scalar(be, t), b,
                    "ijkl", "lkji")
            seq.append(sa.read(lib, a))
        lib.set_zero(b, k)
        lib.add(round_scalar(0.5, t), a, round_scalar(1.0, t), b,
                "ijkl", "lkji")
        seq.append(sa.read(lib, a))
        # a with a zero block turned canonical afresh
        a2s = Spec(sa.dims, {d: SPLIT for d in range(4)}, t)
        for idx in sa.grid():
            if idx != z:
                a2s.canon(idx)
        a2 = build_filled(lib, alloc, a2s, {i: data[0][i] for i in
                                            a2s.canonical_blocks()})
        lib.copy(a2, round_scalar(2.0, t), b, "ijkl", "klij")
        seq.append(a2s.read(lib, a2))
        lib.set_canonical(a2, z)
        lib.write_block(a2, z, data[0][z])
        lib.copy(a2, round_scalar(2.0, t), b, "ijkl", "klij")
        seq.append(sa.read(lib, a2))
        for v in (1.5, -3.0):
            lib.set(a2, round_scalar(v, t))
            seq.append(sa.read(lib, a2))
        lib.set_zero(a2, z)
        lib.set(a2, round_scalar(4.0, t))
        seq.append(a2s.read(lib, a2))
        for x in (a, b, a2):
            lib.tensor_free(x)
        lib.allocator_destroy(alloc)
        outs.append(seq)
    for n, (g, w) in enumerate(zip(*outs)):
        assert same(g, w, t), n


@pytest.mark.parametrize("t", [xm.SCALAR_DOUBLE], ids=["f64"])
def test_cached_jobs_new_tensors_same_addresses(t, gpulib, oracle):
    """Tensors freed and rebuilt (the arena may hand out the same
    addresses) with a different structure: the second add must not reuse
    the first one's jobs."""
    outs = []
    for lib in (gpulib, oracle):
        alloc = lib.allocator()
        seq = []
        for zeros in (True, False):
            sa, sb = a_spec(t), b_spec(t, zeros)
            data = data_for((sa, sb), 29)
            a = build_filled(lib, alloc, sa, data[0])
            b = build_filled(lib, alloc, sb, data[1])
            lib.add(round_scalar(0.5, t), a, round_scalar(1.0, t), b,
                    "ijkl", "ljki")
            seq.append(sa.read(lib, a))
            lib.tensor_free(a)
            lib.tensor_free(b)
        lib.allocator_destroy(alloc)
        outs.append(seq)
    for n, (g, w) in enumerate(zip(*outs)):
        assert same(g, w, t), n
