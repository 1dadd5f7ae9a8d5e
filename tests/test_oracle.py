"""Pins the CPU oracle (oracle/xm_oracle.c) before it is trusted.

* against libxm itself, compiled from its sources into oracle/_ref (same
  BLAS, same calls): results must agree to the last bit or to 1 ulp-level
  differences, and the ?gemm_ call sequence (sizes and factors) of the
  reference must equal the oracle's merged pair list;
* against a dense brute-force expansion (numpy einsum over the full
  tensors built with X(e) = s*S(permute(e,p))), independent of the block
  algorithm;
* against the closed-form result of libxm's tutorial (example.c).
"""
import numpy as np
import pytest

from libxm_amd import xm
from libxm_amd.spec import Spec, rel_err
from libxm_amd import workloads

from cases import (ADD_TESTS, ALPHA_BETA, CONTRACT_TESTS, DOT_TESTS,
                   MULDIV_TESTS, case_id)
from helpers import (TYPES, TYPE_NAMES, build_filled, case_seed, einsum_ref,
                     is_single, max_abs_diff, resolve, round_scalar)

TOL_DENSE = {xm.SCALAR_FLOAT: 2e-5, xm.SCALAR_FLOAT_COMPLEX: 2e-5,
             xm.SCALAR_DOUBLE: 1e-13, xm.SCALAR_DOUBLE_COMPLEX: 1e-13}


def run_contract(lib, specs, data, alpha, beta, idx):
    alloc = lib.allocator()
    sa, sb, sc = specs
    a = build_filled(lib, alloc, sa, data[0])
    b = build_filled(lib, alloc, sb, data[1])
    c = build_filled(lib, alloc, sc, data[2])
    lib.contract(alpha, a, b, beta, c, *idx)
    out = sc.read(lib, c)
    for t in (a, b, c):
        lib.tensor_free(t)
    lib.allocator_destroy(alloc)
    return out


def make_data(specs, seed):
    rng = np.random.default_rng(seed)
    return [s.random_data(rng, -0.5, 0.5) for s in specs]


def dense_contract(specs, data, alpha, beta, idx):
    sa, sb, sc = specs
    A, B, Cd = (s.dense(d) for s, d in zip(specs, data))
    return sc.blocks_of(einsum_ref(alpha, A, B, beta, Cd, *idx))


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
@pytest.mark.parametrize("case", CONTRACT_TESTS, ids=case_id)
def test_contract_table(case, t, oracle, ref):
    """libxm's 46-case contraction table x 4 (alpha, beta) x 4 types."""
    fn, ia, ib, ic = case
    rng = np.random.default_rng(case_seed(case_id(case)) + t)
    for k, (al, be) in enumerate(ALPHA_BETA):
        al = round_scalar(resolve(al, rng, t), t)
        be = round_scalar(resolve(be, rng, t), t)
        specs = fn(t)
        data = make_data(specs, 1000 * k + t)
        got = run_contract(oracle, specs, data, al, be, (ia, ib, ic))
        want = run_contract(ref, specs, data, al, be, (ia, ib, ic))
        # same BLAS, same calls: agreement to rounding of the BLAS kernel
        scale = max(1.0, max(float(np.max(np.abs(v))) if v.size else 0.0
                             for v in want.values()))
        tol = 1e-5 if is_single(t) else 1e-14
        assert max_abs_diff(got, want) <= tol * scale, (al, be)
        dense = dense_contract(specs, data, al, be, (ia, ib, ic))
        assert rel_err(got, dense) <= TOL_DENSE[t], (al, be)


def _recorded(recorder):
    out = []
    buf = (np.zeros(6)).astype(np.float64)
    import ctypes
    p = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    for i in range(recorder.gemm_recorder_count()):
        recorder.gemm_recorder_get(i, p)
        out.append((int(buf[0]), int(buf[1]), int(buf[2]), int(buf[3]),
                    complex(buf[4], buf[5])))
    return out


def _omp_threads(n):
    import ctypes
    gomp = ctypes.CDLL("libgomp.so.1")
    gomp.omp_set_num_threads(n)


def expected_calls(pairs, alpha, t, b_first):
    """(type, m, n, k, al) of each ?gemm_ call libxm makes for the merged
    pair list (libxm src/contract.c:193-216)."""
    out = []
    for p in pairs:
        al = round_scalar(complex(alpha) * p["alpha"], t)
        if t in (xm.SCALAR_FLOAT, xm.SCALAR_FLOAT_COMPLEX):
            al = round_scalar(np.complex64(complex(alpha)) *
                              np.complex64(p["alpha"]), t)
        m, n = (p["n"], p["m"]) if b_first else (p["m"], p["n"])
        out.append((t, m, n, p["k"], al))
    return out


@pytest.mark.parametrize("case", CONTRACT_TESTS, ids=case_id)
def test_pair_list_matches_reference_gemm_calls(case, oracle, ref, recorder):
    """The oracle's merged pair list (order, sizes, factors) equals the
    sequence of ?gemm_ calls libxm makes, single-threaded."""
    fn, ia, ib, ic = case
    t = xm.SCALAR_DOUBLE
    specs = fn(t)
    data = make_data(specs, 7)
    alloc = oracle.allocator()
    a, b, c = (s.build(oracle, alloc) for s in specs)
    pairs = oracle.get_pairs(a, b, c, ia, ib, ic)
    for x in (a, b, c):
        oracle.tensor_free(x)
    oracle.allocator_destroy(alloc)
    masks = _masks(ic, ib)
    b_first = len(masks) > 0 and masks[0] == 0
    _omp_threads(1)
    try:
        recorder.gemm_recorder_start()
        run_contract(ref, specs, data, 0.75, 0.5, (ia, ib, ic))
        recorder.gemm_recorder_stop()
    finally:
        _omp_threads(4)
    assert _recorded(recorder) == expected_calls(pairs, 0.75, t, b_first)


def _masks(s1, s2):
    return [i for i, ch in enumerate(s1) for ch2 in s2 if ch == ch2]


def test_ladder_small_matches_reference(oracle, ref):
    """The o=16/v=64 C2 ladder, oracle vs libxm and vs dense."""
    w = workloads.ladder_small()
    rng = np.random.default_rng(0)
    data = [s.random_data(rng) for s in w.specs]
    for s, d in zip(w.specs, data):
        workloads.antisymmetrize_diagonal(s, d, w.pairs)
    got = run_contract(oracle, w.specs, data, w.alpha, w.beta, w.idx)
    want = run_contract(ref, w.specs, data, w.alpha, w.beta, w.idx)
    assert max_abs_diff(got, want) <= 1e-13
    dense = dense_contract(w.specs, data, w.alpha, w.beta, w.idx)
    assert rel_err(got, dense) <= 1e-13


def test_ladder_small_work_counts(oracle):
    """184 canonical C blocks, 5888 nonzero pairs before the merge, 3424
    after, 0.772 -> 0.449 GFLOP, all 16x64x64 GEMMs."""
    w = workloads.ladder_small()
    alloc = oracle.allocator()
    a, b, c = (s.build(oracle, alloc) for s in w.specs)
    pairs = oracle.get_pairs(a, b, c, *w.idx)
    st = oracle.stats()
    assert len(oracle.canonical_list(c)) == 184
    assert st["npairs_pre"] == 5888 and st["npairs"] == 3424 == len(pairs)
    assert st["flops"] == 3424 * 2 * 16 * 64 * 64
    assert {(p["m"], p["n"], p["k"]) for p in pairs} == {(16, 64, 64)}
    assert {p["alpha"] for p in pairs} <= {1, 2}
    for x in (a, b, c):
        oracle.tensor_free(x)
    oracle.allocator_destroy(alloc)


def test_example_known_answer(oracle):
    """libxm example.c: c = 2*a*b + 1 with a derivative (transposed,
    negated) block in a and a scaled copy in b (libxm example.c:50-74)."""
    lib = oracle
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
    got = np.array([[lib.get_element(c, (i, j)).real for j in range(3)]
                    for i in range(4)])
    # a = [[1,3,0,0,0],[2,4,0,0,0],[0,0,0,-1,-2],[0,0,0,-3,-4]],
    # b = [[6,-4,2],[5,3,-1],[0,0,0],[-3,2,-1],[-2.5,-1.5,0.5]]
    A = np.array([[1, 3, 0, 0, 0], [2, 4, 0, 0, 0], [0, 0, 0, -1, -2],
                  [0, 0, 0, -3, -4]], float)
    B = np.array([[6, -4, 2], [5, 3, -1], [0, 0, 0], [-3, 2, -1],
                  [-2.5, -1.5, 0.5]])
    assert np.array_equal(got, 2 * A @ B + 1)
    assert np.array_equal(got, [[43, 11, -1], [65, 9, 1], [17, 3, 1],
                                [39, 1, 3]])
    for x in (a, b, c):
        lib.tensor_free(x)
    lib.allocator_destroy(alloc)


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
@pytest.mark.parametrize("case", ADD_TESTS, ids=lambda c: "%s-%s-%s" % (
    c[0].__name__, c[1], c[2]))
def test_add_table(case, t, oracle, ref):
    """libxm's add table x 4 (alpha, beta); oracle bit-identical to libxm
    and equal to the dense formula."""
    fn, ia, ib = case
    rng = np.random.default_rng(11 + t)
    for al, be in ALPHA_BETA:
        al = round_scalar(resolve(al, rng, t), t)
        be = round_scalar(resolve(be, rng, t), t)
        sa, sb = fn(t)
        data = make_data((sa, sb), 3)
        outs = []
        for lib in (oracle, ref):
            alloc = lib.allocator()
            a = build_filled(lib, alloc, sa, data[0])
            b = build_filled(lib, alloc, sb, data[1])
            lib.add(al, a, be, b, ia, ib)
            outs.append(sa.read(lib, a))
            lib.tensor_free(a)
            lib.tensor_free(b)
            lib.allocator_destroy(alloc)
        assert max_abs_diff(outs[0], outs[1]) == 0.0
        A, B = sa.dense(data[0]), sb.dense(data[1])
        want = sa.blocks_of(al * A + be * np.einsum("%s->%s" % (ib, ia), B))
        assert rel_err(outs[0], want) <= (1e-6 if is_single(t) else 1e-15)


@pytest.mark.parametrize("t", TYPES, ids=[TYPE_NAMES[t] for t in TYPES])
def test_set_and_copy(t, oracle, ref):
    """xm_set fills canonical blocks (libxm test.c:1826-1861); xm_copy with
    a permutation and a scale equals libxm (libxm test.c:1548-1580)."""
    rng = np.random.default_rng(5)
    x = round_scalar(resolve("r", rng, t), t)
    sp = Spec((3, 5, 1, 6, 3, 6, 9, 4), {0: [1], 1: [1, 2], 3: [3], 5: [4],
                                        6: [3, 7]}, t).all_canonical()
    for lib in (oracle, ref):
        alloc = lib.allocator()
        a = sp.build(lib, alloc)
        lib.set(a, x)
        want = x if np.dtype(xm.NP_DTYPE[t]).kind == "c" else x.real
        for v in sp.read(lib, a).values():
            assert np.all(v == np.asarray(want).astype(v.dtype))
        lib.tensor_free(a)
        lib.allocator_destroy(alloc)
    sa = Spec((43, 43), {0: [4, 7, 21, 27, 38], 1: [4, 7, 21, 27, 38]},
              t).all_canonical()
    data = sa.random_data(rng)
    s = round_scalar(resolve("r", rng, t), t)
    outs = []
    for lib in (oracle, ref):
        alloc = lib.allocator()
        a = build_filled(lib, alloc, sa, data)
        b = sa.build(lib, alloc)
        lib.copy(b, s, a, "ji", "ij")
        outs.append(sa.read(lib, b))
        lib.tensor_free(a)
        lib.tensor_free(b)
        lib.allocator_destroy(alloc)
    assert max_abs_diff(outs[0], outs[1]) == 0.0
    want = sa.blocks_of(s * sa.dense(data).T)
    assert rel_err(outs[0], want) <= (1e-6 if is_single(t) else 1e-15)


@pytest.mark.parametrize("t", [xm.SCALAR_DOUBLE, xm.SCALAR_DOUBLE_COMPLEX],
                         ids=["f64", "c128"])
def test_mul_div_dot(t, oracle, ref):
    """xm_mul / xm_div / xm_dot tables (libxm test.c:1890-1910)."""
    for fn, ia, ib in MULDIV_TESTS:
        sa, sb = fn(t)
        data = make_data((sa, sb), 9)
        for op in ("mul", "div"):
            outs = []
            for lib in (oracle, ref):
                alloc = lib.allocator()
                a = build_filled(lib, alloc, sa, data[0])
                b = build_filled(lib, alloc, sb, data[1])
                getattr(lib, op)(a, b, ia, ib)
                outs.append(sa.read(lib, a))
                lib.tensor_free(a)
                lib.tensor_free(b)
                lib.allocator_destroy(alloc)
            assert max_abs_diff(outs[0], outs[1]) == 0.0
    for fn, ia, ib in DOT_TESTS:
        sa, sb = fn(t)
        data = make_data((sa, sb), 13)
        res = []
        for lib in (oracle, ref):
            alloc = lib.allocator()
            a = build_filled(lib, alloc, sa, data[0])
            b = build_filled(lib, alloc, sb, data[1])
            res.append(lib.dot(a, b, ia, ib))
            lib.tensor_free(a)
            lib.tensor_free(b)
            lib.allocator_destroy(alloc)
        A, B = sa.dense(data[0]), sb.dense(data[1])
        want = np.sum(A * np.einsum("%s->%s" % (ib, ia), B))
        assert abs(res[0] - want) <= 1e-12 * max(1.0, abs(want))
        assert abs(res[0] - res[1]) <= 1e-12 * max(1.0, abs(want))


def test_dim_and_blockspace(oracle, ref):
    """libxm test_dim / test_blockspace (libxm test.c:1728-1824)."""
    for lib in (oracle, ref):
        d = xm.dim((5, 2, 3, 2, 1, 2, 5, 4))
        idx = xm.dim((0,) * 8)
        n = 0
        while xm.undim(idx) != xm.undim(d):
            assert lib.lib.xm_dim_offset(idx, d) == n
            back = lib.lib.xm_dim_from_offset(n, d)
            assert xm.undim(back) == xm.undim(idx)
            lib.lib.xm_dim_inc(idx, d)
            n += 1
        bs = lib.bs_create((10, 20, 30, 40, 50, 60, 70, 80))
        lib.lib.xm_block_space_autosplit(bs)
        assert lib.bs_nblocks(bs) == (1, 1, 1, 2, 2, 2, 3, 3)
        lib.bs_free(bs)
        bs = lib.bs_create((30, 31, 32, 33, 34))
        lib.lib.xm_block_space_autosplit(bs)
        assert lib.bs_nblocks(bs) == (1, 1, 1, 2, 2)
        assert xm.undim(lib.lib.xm_block_space_get_block_dims(
            bs, xm.dim((0,) * 5))) == (30, 31, 32, 16, 18)
        lib.bs_free(bs)
