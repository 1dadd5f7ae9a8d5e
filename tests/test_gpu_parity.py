"""Parity of the MI355X library with the CPU oracle, through the C ABI.

Same structures (replayed recipes), same input data, same calls; results
must agree to
  * FP64 / complex128: relative Frobenius error <= 1e-13,
  * FP32 / complex64:  relative Frobenius error <= 1e-5,
  * bit-exactly for xm_set, xm_add, xm_copy, xm_mul, xm_div (element work
    whose rounding sequence is reproduced), and for the pair lists.
"""
import numpy as np
import pytest

from libxm_amd import xm, workloads
from libxm_amd.spec import Spec, rel_err

from cases import (ADD_TESTS, ALPHA_BETA, CONTRACT_TESTS, MULDIV_TESTS,
                   DOT_TESTS, case_id)
from helpers import (TYPES, TYPE_NAMES, build_filled, case_seed, is_single,
                     max_abs_diff, resolve, round_scalar)

pytestmark = pytest.mark.gpu

TOL = {xm.SCALAR_FLOAT: 1e-5, xm.SCALAR_FLOAT_COMPLEX: 1e-5,
       xm.SCALAR_DOUBLE: 1e-13, xm.SCALAR_DOUBLE_COMPLEX: 1e-13}


def run_contract(lib, specs, data, alpha, beta, idx, stats=False):
    alloc = lib.allocator()
    a = build_filled(lib, alloc, specs[0], data[0])
    b = build_filled(lib, alloc, specs[1], data[1])
    c = build_filled(lib, alloc, specs[2], data[2])
    lib.contract(alpha, a, b, beta, c, *idx)
    out = specs[2].read(lib, c)
    st = lib.stats() if stats else None
    for t in (a, b, c):
        lib.tensor_free(t)
    lib.allocator_destroy(alloc)
    return (out, st) if stats else out


def make_data(specs, seed, lo=-0.5, hi=0.5):
    rng = np.random.default_rng(seed)
    return [s.random_data(rng, lo, hi) for s in specs]


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
@pytest.mark.parametrize("case", CONTRACT_TESTS, ids=case_id)
def test_contract_table(case, t, gpulib, oracle):
    """libxm's 46-case contraction table x 4 (alpha, beta) x 4 types."""
    fn, ia, ib, ic = case
    rng = np.random.default_rng(case_seed(case_id(case)) + t)
    for k, (al, be) in enumerate(ALPHA_BETA):
        al = round_scalar(resolve(al, rng, t), t)
        be = round_scalar(resolve(be, rng, t), t)
        specs = fn(t)
        data = make_data(specs, 100 * k + t)
        got = run_contract(gpulib, specs, data, al, be, (ia, ib, ic))
        want = run_contract(oracle, specs, data, al, be, (ia, ib, ic))
        assert rel_err(got, want) <= TOL[t], (al, be)


def test_example_known_answer(gpulib):
    """libxm example.c through the GPU path: c = 2*a*b + 1."""
    lib = gpulib
    alloc = lib.allocator()
    bsa = lib.bs_create((4, 5), {0: [2], 1: [2, 3]})
    bsb = lib.bs_create((5, 3), {0: [2, 3]})
    bsc = lib.bs_create((4, 3), {0: [2]})
    a = lib.tensor(bsa, xm.SCALAR_DOUBLE, alloc)
    b = lib.tensor(bsb, xm.SCALAR_DOUBLE, alloc)
    c = lib.tensor_canonical(bsc, xm.SCALAR_DOUBLE, alloc)
    for bs in (bsa, bsb, bsc):
        lib.bs_free(bs)
    lib.set_canonical(a, (0, 0))
    lib.write_block(a, (0, 0), np.array([1.0, 2, 3, 4]))
    lib.set_derivative(a, (1, 2), (0, 0), (1, 0), -1.0)
    lib.set_canonical(b, (0, 0))
    lib.write_block(b, (0, 0), np.array([6.0, 5, -4, 3, 2, -1]))
    lib.set_derivative(b, (2, 0), (0, 0), (0, 1), -0.5)
    lib.set(c, 1.0)
    lib.contract(2.0, a, b, 1.0, c, "ik", "kj", "ij")
    got = [[lib.get_element(c, (i, j)).real for j in range(3)]
           for i in range(4)]
    assert got == [[43, 11, -1], [65, 9, 1], [17, 3, 1], [39, 1, 3]]
    for x in (a, b, c):
        lib.tensor_free(x)
    lib.allocator_destroy(alloc)


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
@pytest.mark.parametrize("case", ADD_TESTS, ids=lambda c: "%s-%s-%s" % (
    c[0].__name__, c[1], c[2]))
def test_add_table_bit_exact(case, t, gpulib, oracle):
    fn, ia, ib = case
    rng = np.random.default_rng(17 + t)
    for al, be in ALPHA_BETA:
        al = round_scalar(resolve(al, rng, t), t)
        be = round_scalar(resolve(be, rng, t), t)
        sa, sb = fn(t)
        data = make_data((sa, sb), 5)
        outs = []
        for lib in (gpulib, oracle):
            alloc = lib.allocator()
            a = build_filled(lib, alloc, sa, data[0])
            b = build_filled(lib, alloc, sb, data[1])
            lib.add(al, a, be, b, ia, ib)
            outs.append(sa.read(lib, a))
            lib.tensor_free(a)
            lib.tensor_free(b)
            lib.allocator_destroy(alloc)
        for k in outs[1]:
            assert np.array_equal(outs[0][k].view(np.uint8),
                                  outs[1][k].view(np.uint8)), (al, be, k)


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
def test_add_aliased_antisymmetrize(t, gpulib, oracle):
    """xm_add(0.5, t, -0.5, t, "abcd", "bacd") with a == b: the ladder's
    diagonal-block antisymmetrization, bit-exact with libxm's order."""
    w = workloads.ladder_small(t)
    data = make_data([w.t], 3)[0]
    outs = []
    for lib in (gpulib, oracle):
        alloc = lib.allocator()
        x = build_filled(lib, alloc, w.t, data)
        for ia, ib in workloads.antisym_strings(4, w.pairs):
            lib.add(0.5, x, -0.5, x, ia, ib)
        outs.append(w.t.read(lib, x))
        lib.tensor_free(x)
        lib.allocator_destroy(alloc)
    for k in outs[1]:
        assert np.array_equal(outs[0][k].view(np.uint8),
                              outs[1][k].view(np.uint8)), k
    ref = workloads.antisymmetrize_diagonal(w.t, dict(data), w.pairs)
    for k in ref:
        assert np.array_equal(outs[0][k], ref[k].astype(outs[0][k].dtype)), k


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
def test_set_copy_mul_div_dot(t, gpulib, oracle):
    rng = np.random.default_rng(23)
    x = round_scalar(resolve("r", rng, t), t)
    sp = Spec((3, 5, 1, 6, 3, 6, 9, 4), {0: [1], 1: [1, 2], 3: [3], 5: [4],
                                        6: [3, 7]}, t).all_canonical()
    alloc = gpulib.allocator()
    a = sp.build(gpulib, alloc)
    gpulib.set(a, x)
    want = x if np.dtype(xm.NP_DTYPE[t]).kind == "c" else x.real
    for v in sp.read(gpulib, a).values():
        assert np.all(v == np.asarray(want).astype(v.dtype))
    gpulib.tensor_free(a)
    gpulib.allocator_destroy(alloc)
    for fn, ia, ib in MULDIV_TESTS:
        sa, sb = fn(t)
        data = make_data((sa, sb), 31)
        data[1] = {k: v + (2.0 if v.dtype.kind == "f" else 2.0 + 1j)
                   for k, v in data[1].items()}
        for op in ("mul", "div", "copy"):
            outs = []
            for lib in (gpulib, oracle):
                al = lib.allocator()
                aa = build_filled(lib, al, sa, data[0])
                bb = build_filled(lib, al, sb, data[1])
                if op == "copy":
                    lib.copy(aa, 0.75, bb, ia, ib)
                else:
                    getattr(lib, op)(aa, bb, ia, ib)
                outs.append(sa.read(lib, aa))
                lib.tensor_free(aa)
                lib.tensor_free(bb)
                lib.allocator_destroy(al)
            if op == "div" and t in (xm.SCALAR_FLOAT_COMPLEX,
                                     xm.SCALAR_DOUBLE_COMPLEX):
                assert rel_err(outs[0], outs[1]) <= TOL[t] / 100, op
            else:
                for k in outs[1]:
                    assert np.array_equal(outs[0][k], outs[1][k]), (op, k)
    for fn, ia, ib in DOT_TESTS:
        sa, sb = fn(t)
        data = make_data((sa, sb), 37)
        res = []
        for lib in (gpulib, oracle):
            al = lib.allocator()
            aa = build_filled(lib, al, sa, data[0])
            bb = build_filled(lib, al, sb, data[1])
            res.append(lib.dot(aa, bb, ia, ib))
            lib.tensor_free(aa)
            lib.tensor_free(bb)
            lib.allocator_destroy(al)
        scale = max(1.0, abs(res[1]))
        assert abs(res[0] - res[1]) <= (1e-4 if is_single(t) else 1e-12) * scale


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
def test_copy_transposed_scaled(t, gpulib, oracle):
    """xm_copy(a, 0.75, b, "ji", "ij") from a b holding a derivative block
    (scalar -4, transposed source; libxm test.c:458-477), bit-exact."""
    from cases import ab_3
    sa, sb = ab_3(t)
    data = make_data((sa, sb), 43)
    outs = []
    for lib in (gpulib, oracle):
        al = lib.allocator()
        aa = build_filled(lib, al, sa, data[0])
        bb = build_filled(lib, al, sb, data[1])
        lib.copy(aa, 0.75, bb, "ji", "ij")
        outs.append(sa.read(lib, aa))
        lib.tensor_free(aa)
        lib.tensor_free(bb)
        lib.allocator_destroy(al)
    for k in outs[1]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("t", [xm.SCALAR_FLOAT, xm.SCALAR_DOUBLE],
                         ids=["f32", "f64"])
def test_copy_converts_precision(t, gpulib, oracle):
    """xm_copy into the other precision with a derivative source block
    (libxm test.c:1684-1726)."""
    other = xm.SCALAR_DOUBLE if t == xm.SCALAR_FLOAT else xm.SCALAR_FLOAT
    sa = Spec((2, 9, 11, 3), {0: [1], 1: [2, 4], 2: [3, 4, 8]}, t)
    sa.canon((0, 0, 0, 0))
    sa.deriv((1, 1, 3, 0), (0, 0, 0, 0), (0, 1, 2, 3), -3)
    sa.canon((0, 1, 1, 0)).canon((1, 2, 0, 0)).canon((1, 2, 2, 0))
    sb = Spec(sa.dims, {k: p for k, p in enumerate(sa.points)},
              other).all_canonical()
    data = make_data([sa], 41)[0]
    outs = []
    for lib in (gpulib, oracle):
        al = lib.allocator()
        aa = build_filled(lib, al, sa, data)
        bb = sb.build(lib, al)
        lib.copy(bb, 1, aa, "abcd", "abcd")
        outs.append(sb.read(lib, bb))
        lib.tensor_free(aa)
        lib.tensor_free(bb)
        lib.allocator_destroy(al)
    for k in outs[1]:
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("w", [workloads.ladder_small(),
                               workloads.ladder_d2h(32, 96, 1),
                               workloads.ladder_d2h(48, 192, 3)],
                         ids=["ladder-o16-v64-C2", "ladder-o32-v96-D2h",
                              "ladder-o48-v192-D2h"])
def test_ladder_vs_oracle(w, gpulib, oracle):
    specs = w.specs
    data = make_data(specs, 0, -1, 1)
    for s, d in zip(specs, data):
        workloads.antisymmetrize_diagonal(s, d, w.pairs)
    got, st = run_contract(gpulib, specs, data, w.alpha, w.beta, w.idx, True)
    want = run_contract(oracle, specs, data, w.alpha, w.beta, w.idx)
    assert rel_err(got, want) <= 1e-13
    assert st["npairs"] > 0 and st["flops"] > 0
    # the ladder needs no operand permutes: merged pairs are read in place
    assert st["nunfold"] == 0


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
def test_ring_vs_oracle(t, gpulib, oracle):
    """Ring contraction: both operands and C need permutes; odd block
    extents (4-12), so the FP64 DMA-staged kernel cannot take it while the
    FP32 / complex MFMA kernels (dword staging) do."""
    w = workloads.ring_d2h(24, 72, 5, t, lo=4, hi=12)
    specs = w.specs
    data = make_data(specs, 9, -1, 1)
    got, st = run_contract(gpulib, specs, data, 1.0, 0.5, w.idx, True)
    want = run_contract(oracle, specs, data, 1.0, 0.5, w.idx)
    assert rel_err(got, want) <= TOL[t]
    assert st["nunfold"] > 0


def test_alpha_beta_edge_cases(gpulib, oracle):
    """beta == 0 discards C (NaN included); alpha == 0 skips the products;
    a cached plan is reused with new scalars."""
    w = workloads.ladder_small()
    specs = w.specs
    data = make_data(specs, 2, -1, 1)
    data[2] = {k: np.full_like(v, np.nan) for k, v in data[2].items()}
    got = run_contract(gpulib, specs, data, 0.5, 0.0, w.idx)
    want = run_contract(oracle, specs, data, 0.5, 0.0, w.idx)
    assert rel_err(got, want) <= 1e-13
    assert all(np.all(np.isfinite(v)) for v in got.values())
    data = make_data(specs, 4, -1, 1)
    got = run_contract(gpulib, specs, data, 0.0, -2.0, w.idx)
    want = run_contract(oracle, specs, data, 0.0, -2.0, w.idx)
    assert max_abs_diff(got, want) == 0.0
    # plan reuse: same tensors, several (alpha, beta)
    alloc = gpulib.allocator()
    a = build_filled(gpulib, alloc, specs[0], data[0])
    b = build_filled(gpulib, alloc, specs[1], data[1])
    c = build_filled(gpulib, alloc, specs[2], data[2])
    refs = []
    for al, be in [(0.5, 1.0), (-1.0, 0.25), (2.0, 0.0)]:
        gpulib.contract(al, a, b, be, c, *w.idx)
        st = gpulib.stats()
        refs.append((specs[2].read(gpulib, c), st["plan_cached"]))
    assert [r[1] for r in refs] == [0, 1, 1]
    cur = data[2]
    for (al, be), (out, _) in zip([(0.5, 1.0), (-1.0, 0.25), (2.0, 0.0)], refs):
        cur = run_contract(oracle, specs, [data[0], data[1], cur], al, be, w.idx)
        assert rel_err(out, cur) <= 1e-13
    for x in (a, b, c):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def test_device_fill_random_reproducible(gpulib):
    w = workloads.ladder_small()
    alloc = gpulib.allocator()
    x = w.v.build(gpulib, alloc)
    y = w.v.build(gpulib, alloc)
    gpulib.fill_random(x, 7)
    gpulib.fill_random(y, 7)
    dx, dy = w.v.read(gpulib, x), w.v.read(gpulib, y)
    for k in dx:
        assert np.array_equal(dx[k], dy[k])
        assert np.all(np.abs(dx[k]) <= 1.0)
    vals = np.concatenate(list(dx.values()))
    assert abs(vals.mean()) < 0.01 and 0.3 < vals.std() < 0.62
    gpulib.tensor_free(x)
    gpulib.tensor_free(y)
    gpulib.allocator_destroy(alloc)
