"""Full-size parity cases on the GPU.

* Precision: the D2h ladder at o=64, v=512 with entries U[-1,1]*10^U[-6,0],
  run as plain FP64, as split precision (6 and 7 BF16 slices) and in FP32;
  each against the FP64 oracle on a seeded sample of C blocks.
* Batching: the ring c(i,j,a,b) = sum t(i,k,a,c) w(k,b,c,j) at o=96, v=480
  with block sizes U[4,28] (~2.5e5 block pairs in one launch, unfolds on
  both operands), in FP64 and complex64, then xm_add(1, c, 0.5, c2) and
  xm_set(c, 0).
* Headline: the product term (beta = 0) of the o=64, v=384 ladder in FP64
  and split precision against the oracle, on a seeded sample of C blocks
  that includes the one with the largest P x Q extent.
* Differentiation: xm_contract_backward of the o=64, v=384 ladder, by
  central finite differences of L = sum over canonical c of dc * c on 64
  seeded canonical entries of t and of v, split against FP64, and against
  the oracle's contractions of dc' (dc's canonical blocks) with v and with
  t into every block of t / v that a sampled canonical block feeds,
  folded on the host into the sampled blocks.

The oracle recomputes a sample (fullsize.Sample) from the GPU tensors' own
inputs, so the GPU and CPU results see the same data.
"""
import numpy as np
import pytest

from libxm_amd import workloads, xm
from libxm_amd.spec import Spec

from fullsize import (Sample, canonical_view, err, fill, fold_block,
                      fold_sets, pick_cblocks, pick_with_largest,
                      read_blocks, source)

pytestmark = pytest.mark.gpu

BUDGET = 1e11          # FLOPs the oracle recomputes per sample


def _contract_sample(gpu, oracle, w, tensors, seed, mult=2.0):
    t, v, c = tensors
    pairs = gpu.get_pairs(t, v, c, *w.idx)
    chosen, fsum = pick_cblocks(pairs, w.c.canonical_blocks(), BUDGET, seed,
                                mult)
    assert chosen and fsum > 0
    return pairs, chosen


def test_precision_ladder_o64_v512(gpulib, oracle):
    w = workloads.ladder_d2h(64, 512, 0, value_exp=(-6.0, 0.0))
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    for k, x in enumerate((t, v, c)):
        fill(gpulib, w, x, 100 + k, (-6.0, 0.0))
    pairs, chosen = _contract_sample(gpulib, oracle, w, (t, v, c), 1)
    c0 = read_blocks(gpulib, c, chosen)
    ref = Sample(gpulib, oracle, w, w.specs, (t, v, c), c0, chosen, w.alpha,
                 w.beta, w.idx, pairs)
    errs = {}
    work = w.c.build(gpulib, alloc)
    default = gpulib.lib.xm_get_split_slices()
    assert default == 6
    for name, prec, ns in [("fp64", xm.PRECISION_NATIVE, None),
                           ("split6", xm.PRECISION_SPLIT_BF16, 6),
                           ("split7", xm.PRECISION_SPLIT_BF16, 7),
                           ("split_default", xm.PRECISION_SPLIT_BF16,
                            default),
                           ("int8", xm.PRECISION_SPLIT_INT8, None)]:
        gpulib.copy(work, 1.0, c, "abcd", "abcd")
        gpulib.set_precision(prec, ns)
        gpulib.contract(w.alpha, t, v, w.beta, work, *w.idx)
        errs[name] = err(read_blocks(gpulib, work, chosen), ref.out)
    gpulib.set_precision(xm.PRECISION_NATIVE)
    # FP32: the same data converted by xm_copy, contracted in single
    wf = workloads.ladder_d2h(64, 512, 0, type_=xm.SCALAR_FLOAT)
    tf, vf, cf = (s.build(gpulib, alloc) for s in wf.specs)
    for dst, src in ((tf, t), (vf, v), (cf, c)):
        gpulib.copy(dst, 1.0, src, "abcd", "abcd")
    gpulib.contract(w.alpha, tf, vf, w.beta, cf, *w.idx)
    errs["fp32"] = err(read_blocks(gpulib, cf, chosen), ref.out)
    ref.free()
    print("precision ladder o=64 v=512, %d of %d C blocks:" % (
        len(chosen), len(w.c.canonical_blocks())), errs)
    assert errs["fp64"] <= 1e-13
    assert errs["split6"] <= 1e-12
    assert errs["split7"] <= 1e-12
    assert errs["int8"] <= 1e-12      # default settings: moduli per plan
    assert errs["split_default"] <= 1e-12
    assert errs["fp32"] <= 1e-5
    for x in (t, v, c, work, tf, vf, cf):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


@pytest.mark.parametrize("type_", [xm.SCALAR_DOUBLE, xm.SCALAR_FLOAT_COMPLEX],
                         ids=["f64", "c64"])
def test_ring_o96_v480(gpulib, oracle, type_):
    w = workloads.ring_d2h(96, 480, 0, type_)
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    for k, x in enumerate((t, v, c)):
        fill(gpulib, w, x, 200 + k)
    pairs = gpulib.get_pairs(t, v, c, *w.idx)
    assert len(pairs) > 100000
    cplx = type_ == xm.SCALAR_FLOAT_COMPLEX
    _, chosen = _contract_sample(gpulib, oracle, w, (t, v, c), 2,
                                 8.0 if cplx else 2.0)
    c0 = read_blocks(gpulib, c, chosen)
    gpulib.contract(w.alpha, t, v, w.beta, c, *w.idx)
    st = gpulib.stats()
    assert st["npairs"] == len(pairs) and st["gemm_launches"] == 1
    ref = Sample(gpulib, oracle, w, w.specs, (t, v, c), c0, chosen, w.alpha,
                 w.beta, w.idx, pairs)
    e = err(read_blocks(gpulib, c, chosen), ref.out)
    ref.free()
    print("ring o=96 v=480 %s: %d pairs, %d sampled C blocks, rel err %.2e" % (
        "c64" if cplx else "f64", len(pairs), len(chosen), e))
    assert e <= (1e-5 if cplx else 1e-13)

    # xm_add(1, c, 0.5, c2): bit-identical to the oracle's on the sample
    c2 = w.c.build(gpulib, alloc)
    fill(gpulib, w, c2, 300)
    before = read_blocks(gpulib, c, chosen)
    other = read_blocks(gpulib, c2, chosen)
    gpulib.add(1.0, c, 0.5, c2, "ijab", "ijab")
    after = read_blocks(gpulib, c, chosen)
    oa = oracle.allocator()
    oc, oc2 = w.c.build(oracle, oa), w.c.build(oracle, oa)
    for i in chosen:
        oracle.write_block(oc, i, before[i])
        oracle.write_block(oc2, i, other[i])
    oracle.add(1.0, oc, 0.5, oc2, "ijab", "ijab")
    for i in chosen:
        assert np.array_equal(np.asarray(oracle.read_block(oc, i)),
                              np.asarray(after[i]))
    for x in (oc, oc2):
        oracle.tensor_free(x)
    oracle.allocator_destroy(oa)

    # xm_set(c, 0): every canonical entry, checked on the whole tensor
    gpulib.set(c, 0.0)
    assert abs(gpulib.dot(c, c, "ijab", "ijab")) == 0.0
    for i in chosen:
        assert not np.any(np.asarray(gpulib.read_block(c, i)))
    for x in (t, v, c, c2):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def _fd(lib, w, t, v, c, dc, x, entries, h):
    """Central differences of L = dot(dc, c(t, v)) (beta = 0) at `entries`
    (block index, offset) of canonical data of x.  dc must hold c's
    canonical blocks only: xm_dot sums over every block position where both
    operands are nonzero, derivative blocks included."""
    out = []
    for idx, off in entries:
        blk = np.array(lib.read_block(x, idx), dtype=np.float64)
        vals = []
        for s in (h, -h):
            p = blk.copy()
            p[off] += s
            lib.write_block(x, idx, p)
            lib.contract(w.alpha, t, v, 0.0, c, *w.idx)
            vals.append(lib.dot(dc, c, "ijab", "ijab").real)
        lib.write_block(x, idx, blk)
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def test_backward_ladder_o64_v384(gpulib):
    w = workloads.ladder_d2h(64, 384, 0)
    alloc = gpulib.allocator()
    t, v, c, dc = (s.build(gpulib, alloc) for s in (*w.specs, w.c))
    for k, x in enumerate((t, v, c, dc)):
        fill(gpulib, w, x, 400 + k)
    dt, dv = w.t.build(gpulib, alloc), w.v.build(gpulib, alloc)
    gpulib.contract_backward(w.alpha, t, v, dc, dt, dv, *w.idx)
    # L = sum over canonical c of dc * c: dc's canonical blocks alone
    cv = Spec(w.c.dims, {k: p for k, p in enumerate(w.c.points)}, w.c.type)
    for idx in w.c.canonical_blocks():
        cv.canon(idx)
    dcv = cv.build(gpulib, alloc)
    gpulib.copy(dcv, 1.0, dc, "ijab", "ijab")
    rng = np.random.default_rng(5)
    for spec, x, g in ((w.t, t, dt), (w.v, v, dv)):
        cb = spec.canonical_blocks()
        entries = []
        for _ in range(64):
            idx = cb[rng.integers(len(cb))]
            entries.append((idx, int(rng.integers(spec.block_size(idx)))))
        an = np.array([np.asarray(gpulib.read_block(g, i))[o] for i, o in
                       entries], dtype=np.float64)
        # L is linear in each operand, so central differences carry no
        # truncation error; what remains is the rounding of L (~1e-10 at
        # |L| ~ 7e4) over 2h: a ~1e-6 floor at h = 1e-6, ~1e-10 at 1e-2
        for h, tol in ((1e-6, 5e-5), (1e-2, 1e-8)):
            fd = _fd(gpulib, w, t, v, c, dcv, x, entries, h)
            rel = np.linalg.norm(fd - an) / np.linalg.norm(an)
            print("backward ladder o=64 v=384: FD (h=%g) vs adjoint, 64 "
                  "entries of %s, rel %.2e" % (h, "t" if x is t else "v", rel))
            assert rel <= tol
    # split precision adjoint against the FP64 one, on the whole tensors
    st, sv = w.t.build(gpulib, alloc), w.v.build(gpulib, alloc)
    gpulib.set_precision(xm.PRECISION_SPLIT_BF16, 5)
    gpulib.contract_backward(w.alpha, t, v, dc, st, sv, *w.idx)
    gpulib.set_precision(xm.PRECISION_NATIVE)
    for spec, a, b in ((w.t, dt, st), (w.v, dv, sv)):
        e = err(spec.read(gpulib, b), spec.read(gpulib, a))
        print("backward split vs FP64:", e)
        assert e <= 1e-12
    for x in (t, v, c, dc, dcv, dt, dv, st, sv):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def test_ladder_o64_v384_forward(gpulib, oracle):
    w = workloads.ladder_d2h(64, 384, 0)
    alloc = gpulib.allocator()
    t, v, c = (s.build(gpulib, alloc) for s in w.specs)
    for k, x in enumerate((t, v, c)):
        fill(gpulib, w, x, k)           # the benchmark's data (seed 0)
    pairs = gpulib.get_pairs(t, v, c, *w.idx)
    chosen, fsum = pick_with_largest(pairs, w.c.canonical_blocks(), BUDGET,
                                     11)
    zero = {i: np.zeros(w.c.block_size(i)) for i in chosen}
    ref = Sample(gpulib, oracle, w, w.specs, (t, v, c), zero, chosen,
                 w.alpha, 0.0, w.idx, pairs)
    errs = {}
    for name, prec, ns in [("fp64", xm.PRECISION_NATIVE, None),
                           ("split5", xm.PRECISION_SPLIT_BF16, 5),
                           ("int8", xm.PRECISION_SPLIT_INT8, None)]:
        gpulib.set_precision(prec, ns)
        gpulib.contract(w.alpha, t, v, 0.0, c, *w.idx)
        st = gpulib.stats()
        assert st["precision"] == prec
        errs[name] = err(read_blocks(gpulib, c, chosen), ref.out)
    gpulib.set_precision(xm.PRECISION_NATIVE)
    ref.free()
    print("ladder o=64 v=384 product term, %d of %d C blocks (%.3g GFLOP, "
          "largest %r):" % (len(chosen), len(w.c.canonical_blocks()),
                            fsum / 1e9, chosen[0]), errs)
    assert errs["fp64"] <= 1e-13
    assert errs["split5"] <= 1e-12
    assert errs["int8"] <= 1e-12
    for x in (t, v, c):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)


def _oracle_gradient(gpu, cpu, w, dcv, other, other_spec, g_spec, members,
                     idx):
    """The oracle's G = alpha * dc' (x) other into every block of `members`
    (g_spec's block space), with dc' and the operand read from the GPU."""
    cv = canonical_view(w.c)
    gs = Spec(g_spec.dims, {k: p for k, p in enumerate(g_spec.points)},
              g_spec.type)
    for x in members:
        gs.canon(x)
    alloc = cpu.allocator()
    cdc = cv.build(cpu, alloc)
    co = other_spec.build(cpu, alloc)
    cg = gs.build(cpu, alloc)
    ia, ib, ic = idx
    need_dc, need_o = set(), set()
    # g = "ijcd" (dt) or "abcd" (dv); dc' = "ijab"; other = "abcd" / "ijcd"
    for x in members:
        for d in cv.canonical_blocks():
            if (ic == "ijcd" and d[:2] == x[:2]) or \
                    (ic == "abcd" and d[2:] == x[:2]):
                need_dc.add(d)
                oi = d[2:] + x[2:] if ic == "ijcd" else d[:2] + x[2:]
                if other_spec.kind(oi) is not None:
                    need_o.add(source(other_spec, oi))
    for d in need_dc:
        cpu.write_block(cdc, d, gpu.read_block(dcv, d))
    for d in need_o:
        cpu.write_block(co, d, gpu.read_block(other, d))
    cpu.contract(w.alpha, cdc, co, 0.0, cg, ia, ib, ic)
    out = {x: cpu.read_block(cg, x) for x in members}
    for x in (cdc, co, cg):
        cpu.tensor_free(x)
    cpu.allocator_destroy(alloc)
    return out


def test_backward_ladder_o64_v384_vs_oracle(gpulib, oracle):
    w = workloads.ladder_d2h(64, 384, 0)
    alloc = gpulib.allocator()
    t, v, c, dc = (s.build(gpulib, alloc) for s in (*w.specs, w.c))
    for k, x in enumerate((t, v, c, dc)):
        fill(gpulib, w, x, 400 + k)
    dcv = canonical_view(w.c).build(gpulib, alloc)
    gpulib.copy(dcv, 1.0, dc, "ijab", "ijab")
    rng = np.random.default_rng(17)

    def sample(spec, n_off, n_diag):
        cb = spec.canonical_blocks()
        off = [i for i in cb if i[0] < i[1] and i[2] < i[3]]
        diag = [i for i in cb if i[0] == i[1] or i[2] == i[3]]
        return [off[k] for k in rng.choice(len(off), n_off, replace=False)] \
            + [diag[k] for k in rng.choice(len(diag), n_diag, replace=False)]

    # dt: 15 GFLOP per t block on the host; dv: ~3 GFLOP per v block
    plan = ((w.t, t, v, w.v, sample(w.t, 2, 1), ("ijab", "abcd", "ijcd")),
            (w.v, v, t, w.t, sample(w.v, 6, 2), ("ijab", "ijcd", "abcd")))
    refs = []
    for spec, x, other, ospec, srcs, idx in plan:
        fs = fold_sets(spec, srcs)
        members = [m[0] for s in srcs for m in fs[s]]
        g = _oracle_gradient(gpulib, oracle, w, dcv, other, ospec, spec,
                             members, idx)
        refs.append({s: fold_block(spec, s, fs[s], g) for s in srcs})
    for name, prec, ns, tol in (("fp64", xm.PRECISION_NATIVE, None, 1e-13),
                                ("split5", xm.PRECISION_SPLIT_BF16, 5, 1e-12),
                                ("int8", xm.PRECISION_SPLIT_INT8, None,
                                 1e-12)):
        gpulib.set_precision(prec, ns)
        dt, dv = w.t.build(gpulib, alloc), w.v.build(gpulib, alloc)
        gpulib.contract_backward(w.alpha, t, v, dc, dt, dv, *w.idx)
        for (spec, *_), g, ref in zip(plan, (dt, dv), refs):
            e = err(read_blocks(gpulib, g, list(ref)), ref)
            print("backward %s vs oracle contractions, %d canonical %s "
                  "blocks: rel %.2e" % (name, len(ref),
                                        "t" if spec is w.t else "v", e))
            assert e <= tol
        gpulib.tensor_free(dt)
        gpulib.tensor_free(dv)
    gpulib.set_precision(xm.PRECISION_NATIVE)
    for x in (t, v, c, dc, dcv):
        gpulib.tensor_free(x)
    gpulib.allocator_destroy(alloc)
