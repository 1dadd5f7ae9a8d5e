"""The C-ABI library without a GPU: it loads, exports every symbol the
headers declare, and its host logic (block structure, index masks, pair
enumeration and merge, fatal errors) equals the oracle's.

Structures are built with set_canonical_block_raw and distinct fake data
pointers, which exercises the whole host path without device storage.
"""
import itertools
import os
import re
import subprocess
import sys
import textwrap

import pytest

from libxm_amd import xm, workloads

from cases import CONTRACT_TESTS, case_id
from conftest import LIBXM_SO, ORACLE_SO, ROOT

HEADERS = ["dim.h", "scalar.h", "alloc.h", "blockspace.h", "tensor.h",
           "xm.h", "util.h", "xm_amd.h"]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(xm_\w+)\s*\(", text):
            if not re.search(r"#define\s+\w*\s*$", text[:m.start()].split("\n")[-1]):
                names.add(m.group(1))
    names.discard("xm_fatal") if False else None
    return sorted(names)


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_library_exports_every_declared_function(gpulib):
    decl = declared_functions()
    assert len(decl) > 90
    missing = [n for n in decl if n not in exported(LIBXM_SO)]
    assert missing == []


def test_static_library_built():
    assert os.path.exists(os.path.join(ROOT, "libxm_amd", "libxm.a"))


def _fake():
    return itertools.count(0x10000, 0x1000)


@pytest.mark.parametrize("case", CONTRACT_TESTS, ids=case_id)
def test_pair_lists_match_oracle(case, gpulib, oracle):
    """The merged pair list (libxm order, factors, GEMM sizes) of the
    product's planner equals the oracle's, which equals libxm's own ?gemm_
    call sequence (test_oracle.py)."""
    fn, ia, ib, ic = case
    lists = []
    for lib in (gpulib, oracle):
        alloc = lib.allocator()
        fk = _fake()
        a, b, c = (s.build(lib, alloc, fk) for s in fn())
        lists.append((lib.get_pairs(a, b, c, ia, ib, ic), lib.stats()))
        for t in (a, b, c):
            lib.tensor_free(t, data=False)
        lib.allocator_destroy(alloc)
    assert lists[0][0] == lists[1][0]
    for k in ("flops", "flops_pre", "npairs", "npairs_pre", "nblocks_c"):
        assert lists[0][1][k] == lists[1][1][k], k


@pytest.mark.parametrize("w", [workloads.ladder_small(),
                               workloads.ladder_d2h(32, 96, 1),
                               workloads.ring_d2h(24, 60, 2)],
                         ids=["ladder-C2", "ladder-D2h-small", "ring-small"])
def test_workload_pairs_match_oracle(w, gpulib, oracle):
    lists = []
    for lib in (gpulib, oracle):
        alloc = lib.allocator()
        fk = _fake()
        a, b, c = (s.build(lib, alloc, fk) for s in w.specs)
        lists.append(lib.get_pairs(a, b, c, *w.idx))
        for t in (a, b, c):
            lib.tensor_free(t, data=False)
        lib.allocator_destroy(alloc)
    assert lists[0] == lists[1]
    assert len(lists[0]) > 0


def test_structure_api_matches_oracle(gpulib, oracle):
    """Block types, permutations, scalars, data-pointer sharing of
    derivative blocks and the canonical list order."""
    w = workloads.ladder_small()
    views = []
    for lib in (gpulib, oracle):
        alloc = lib.allocator()
        t = w.v.build(lib, alloc, _fake())
        view = []
        for idx in w.v.grid():
            view.append((lib.block_type(t, idx), lib.block_perm(t, idx),
                         lib.block_scalar(t, idx),
                         lib.block_data_ptr(t, idx)))
        views.append((view, lib.canonical_list(t), lib.nblocks(t),
                      lib.abs_dims(t),
                      lib.lib.xm_tensor_get_largest_block_size(t)))
        lib.tensor_free(t, data=False)
        lib.allocator_destroy(alloc)
    assert views[0] == views[1]


FATAL_SNIPPET = textwrap.dedent("""
    import sys
    sys.path.insert(0, %(root)r)
    from libxm_amd import xm
    from libxm_amd.spec import Spec
    import itertools
    lib = xm.XmLib(%(lib)r)
    alloc = lib.allocator()
    fk = itertools.count(0x10000, 0x1000)
    a = Spec((4, 4), {0: [2], 1: [2]}).all_canonical().build(lib, alloc, fk)
    b = Spec((4, 4), {0: [2], 1: [2]}).all_canonical().build(lib, alloc, fk)
    c = Spec((4, 4), {0: [2], 1: [2]}).all_canonical().build(lib, alloc, fk)
    d = Spec((4, 4), {0: [1], 1: [2]}).all_canonical().build(lib, alloc, fk)
    z = Spec((4, 4), {0: [2], 1: [2]}).build(lib, alloc, fk)
    case = %(case)r
    if case == "idx_len":
        lib.contract(1, a, b, 0, c, "ik", "kj", "ijk")
    elif case == "idx_cover":
        lib.contract(1, a, b, 0, c, "ik", "kl", "ij")
    elif case == "first_index":
        lib.contract(1, a, b, 0, c, "ki", "jk", "kj")
    elif case == "split_ab":
        lib.contract(1, a, d, 0, c, "ik", "kj", "ij")
    elif case == "write_zero":
        import numpy as np
        lib.write_block(z, (0, 0), np.zeros(4))
    elif case == "set_nonzero":
        lib.set_derivative(a, (0, 1), (0, 0), (1, 0), -1)
    elif case == "deriv_source":
        lib.set_derivative(z, (0, 1), (1, 1), (1, 0), -1)
    elif case == "add_cover":
        lib.add(1, a, 1, b, "ij", "jk")
    elif case == "types":
        e = Spec((4, 4), {0: [2], 1: [2]}, xm.SCALAR_FLOAT).all_canonical()
        e = e.build(lib, alloc, fk)
        lib.contract(1, a, b, 0, e, "ik", "kj", "ij")
    print("no error")
""")

FATAL_CASES = {
    "idx_len": "libxm: xm_contract: bad contraction indices",
    "idx_cover": "libxm: xm_contract: bad contraction indices",
    "first_index": "libxm: xm_contract: bad contraction indices",
    "split_ab": "libxm: xm_contract: inconsistent a and b tensor block-spaces",
    "write_zero": "libxm: xm_tensor_write_block: can only write to canonical blocks",
    "set_nonzero": "libxm: xm_tensor_set_derivative_block: block must be zero",
    "deriv_source": "libxm: xm_tensor_set_derivative_block: derivative blocks "
                    "must have canonical source blocks",
    "add_cover": "libxm: xm_add: index spaces do not match",
    "types": "libxm: xm_contract: tensors must have same scalar type",
}


@pytest.mark.parametrize("case", sorted(FATAL_CASES))
def test_fatal_messages_match_libxm(case, gpulib):
    """Errors keep libxm's convention: the message on stderr, exit status 1
    (libxm src/util.c:24-35), with libxm's texts; the oracle agrees."""
    for path in (LIBXM_SO, ORACLE_SO):
        r = subprocess.run([sys.executable, "-c", FATAL_SNIPPET % dict(
            root=ROOT, lib=path, case=case)], capture_output=True, text=True,
            timeout=120)
        assert r.returncode == 1, (path, r.stdout, r.stderr)
        assert r.stderr.strip().splitlines()[-1] == FATAL_CASES[case], path


def test_masks_and_dim_algebra(gpulib, oracle):
    for lib in (gpulib, oracle):
        m1, m2 = xm.Dim(), xm.Dim()
        lib.lib.xm_make_masks(b"ijcd", b"abcd", m1, m2)
        assert (xm.undim(m1), xm.undim(m2)) == ((2, 3), (2, 3))
        lib.lib.xm_make_masks(b"ijab", b"kbcj", m1, m2)
        assert (xm.undim(m1), xm.undim(m2)) == ((1, 3), (3, 1))
        p = lib.lib.xm_dim_permute(xm.dim((5, 6, 7)), xm.dim((2, 0, 1)))
        assert xm.undim(p) == (6, 7, 5)
