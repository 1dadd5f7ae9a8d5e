#!/usr/bin/env python3
"""Benchmark of xm_contract on one MI355X.

Workload: the particle-particle ladder c(i,j,a,b) = 0.5 t(i,j,c,d)
v(a,b,c,d) + 1.0 c(i,j,a,b), o = 64 split into seeded blocks from
{12,16,20}, v = 384 from {24,32,40,48}, D2h irreps round-robin with the XOR
selection rule, t/v/c antisymmetric in (i,j)/(a,b) and (c,d) with
derivative blocks, diagonal blocks antisymmetrized, canonical entries
U[-1,1] (seed 0).  A "step" is one xm_contract call on tensors resident in
HBM.

FLOPs F = sum over merged block pairs with nonzero factor of 2*m*n*k (the
GEMM work libxm itself executes).  value = F / (wall time per step) of the
FP64 path (the drop-in); the split-precision (BF16 slices) form is reported
beside it with the same F, its error against FP64 results, and its BF16
matrix-core utilisation.  roofline.achieved = F / (average device time of
the grouped GEMM stage, HIP events on the library's stream): the GEMM
kernel and, for the FP64 tiles run as two K halves, the kernel that sums
the halves into C.

The CPU baseline runs libxm itself (compiled from its sources into
oracle/_ref, sequential MKL per OpenMP thread) on the same workload's
canonical C blocks (about 2 s a run on 16 host cores, median of 5 runs),
timed on the host cores as TFLOP/s of the same F.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK = 78.6          # TFLOP/s, MI355X dense FP64 matrix
BF16_PEAK = 2516.6        # TFLOP/s, MI355X dense BF16 matrix
INT8_PEAK = 5033.2        # TOP/s, MI355X dense int8 matrix
HBM_PEAK = 8000.0         # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def products(nslices):
    """BF16 slice products per FP64 product: pairs (t, u), t + u <= s + 1."""
    return nslices * (nslices + 1) // 2


def fill_tensor(lib, w, x, seed):
    from libxm_amd import workloads
    lib.fill_random(x, seed)
    for ia, ib in workloads.antisym_strings(4, w.pairs):
        lib.add(0.5, x, -0.5, x, ia, ib)


def build_workload(lib, w, seed=0):
    alloc = lib.allocator()
    t, v, c = (s.build(lib, alloc) for s in w.specs)
    for k, x in enumerate((t, v, c)):
        fill_tensor(lib, w, x, seed * 16 + k)
    lib.synchronize()
    return alloc, t, v, c


def time_steps(lib, w, t, v, c, steps, warmup):
    for _ in range(warmup):
        lib.contract(w.alpha, t, v, w.beta, c, *w.idx)
    lib.synchronize()
    gemm, prep = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.contract(w.alpha, t, v, w.beta, c, *w.idx)
        st = lib.stats()
        gemm.append(st["gemm_ms"])
        prep.append(st["prep_ms"])
    lib.synchronize()
    dt = time.perf_counter() - t0
    return (dt / steps * 1e3, sum(gemm) / len(gemm), sum(prep) / len(prep),
            lib.stats())


def sample_blocks(lib, w, t, v, c, frac=1.0):
    """Canonical C blocks of the first (a,b) virtual block pairs, about
    `frac` of the workload's FLOPs (all of them by default), and the a/b
    source blocks they read."""
    cblocks = w.c.canonical_blocks()
    abs_ = []
    for idx in cblocks:
        if idx[2:] not in abs_:
            abs_.append(idx[2:])
    pairs_all = lib.get_pairs(t, v, c, *w.idx)
    f_of = {}
    for p in pairs_all:
        f_of[p["c"]] = f_of.get(p["c"], 0) + 2.0 * p["m"] * p["n"] * p["k"]
    target = frac * sum(f_of.values())
    chosen, fsum = [], 0.0
    for ab in abs_:
        blocks = [x for x in cblocks if x[2:] == ab]
        if frac < 1.0 and chosen and \
                fsum + sum(f_of.get(x, 0) for x in blocks) > target:
            break
        chosen += blocks
        fsum += sum(f_of.get(x, 0) for x in blocks)
    sel = set(chosen)

    def src(spec, idx):
        k = spec.kind(idx)
        return idx if k == "c" else k[1]
    need_a = {src(w.t, p["a"]) for p in pairs_all if p["c"] in sel}
    need_b = {src(w.v, p["b"]) for p in pairs_all if p["c"] in sel}
    return chosen, fsum, need_a, need_b, len(cblocks)


def cpu_baseline(lib, w, alloc_gpu, t, v, c, c0, split_slices=None,
                 runs=5):
    """libxm (or, if it is absent, the oracle) on the workload's canonical C
    blocks (about 2 s a run on 16 host cores: median of `runs`, about 10 s
    of host work); the same blocks are recomputed on the GPU (FP64 and
    split) for parity.  Both start from c0, the C blocks as they were
    before any timed call (the workload's initial c), and the errors are
    reported relative to the whole result and relative to the contraction
    term alone (result - beta * c0)."""
    from libxm_amd import xm
    from libxm_amd.spec import Spec, rel_err
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libxmref.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    kind = "reference" if os.path.exists(ref_so) else "port"
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    cpu = xm.XmLib(ref_so if kind == "reference" else port_so)
    oracle = xm.XmLib(port_so)
    ncores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    chosen, fsum, need_a, need_b, ncb = sample_blocks(lib, w, t, v, c)
    cs = Spec(w.c.dims, {k: p for k, p in enumerate(w.c.points)}, w.c.type)
    for idx in chosen:
        cs.canon(idx)
    alloc = cpu.allocator()
    ct, cv = w.t.build(cpu, alloc), w.v.build(cpu, alloc)
    cc = cs.build(cpu, alloc)
    for i in need_a:
        cpu.write_block(ct, i, lib.read_block(t, i))
    for i in need_b:
        cpu.write_block(cv, i, lib.read_block(v, i))
    times = []
    for _ in range(runs):
        for i in chosen:
            cpu.write_block(cc, i, c0[i])
        t0 = time.perf_counter()
        cpu.contract(w.alpha, ct, cv, w.beta, cc, *w.idx)
        times.append(time.perf_counter() - t0)
    dt = sorted(times)[len(times) // 2]
    cpu_out = {i: cpu.read_block(cc, i) for i in chosen}
    errs = {}
    gc = cs.build(lib, alloc_gpu)
    for prec, ns in [(xm.PRECISION_NATIVE, None)] + (
            [(xm.PRECISION_SPLIT_BF16, split_slices),
             (xm.PRECISION_SPLIT_INT8, None)] if split_slices else []):
        lib.set_precision(prec, ns)
        for i in chosen:
            lib.write_block(gc, i, c0[i])
        lib.contract(w.alpha, t, v, w.beta, gc, *w.idx)
        gpu_out = {i: lib.read_block(gc, i) for i in chosen}
        errs[prec] = (rel_err(gpu_out, cpu_out),
                      rel_err_term(gpu_out, cpu_out, c0, w.beta))
    lib.set_precision(xm.PRECISION_NATIVE)
    lib.tensor_free(gc)
    for x in (ct, cv, cc):
        cpu.tensor_free(x)
    cpu.allocator_destroy(alloc)
    blas = oracle.lib.xm_oracle_blas_name().decode()
    return {"value": fsum / dt / 1e12, "unit": "TFLOP/s", "cores": ncores,
            "kind": kind,
            "sample": "%d of %d canonical C blocks (%d (a,b) virtual "
                      "block pairs), %.3g GFLOP, median of %d runs %.2f s, "
                      "OpenMP %d threads, BLAS %s" % (
                          len(chosen), ncb, len({x[2:] for x in chosen}),
                          fsum / 1e9, runs, dt, ncores, blas),
            "rel_err_fp64_gpu_vs_cpu": errs[xm.PRECISION_NATIVE][0],
            "rel_err_split_gpu_vs_cpu":
                errs.get(xm.PRECISION_SPLIT_BF16, (None,))[0],
            # the same differences relative to the contraction term alone
            "rel_err_fp64_gpu_vs_cpu_product_term":
                errs[xm.PRECISION_NATIVE][1],
            "rel_err_split_gpu_vs_cpu_product_term":
                errs.get(xm.PRECISION_SPLIT_BF16, (None, None))[1],
            "rel_err_int8_gpu_vs_cpu":
                errs.get(xm.PRECISION_SPLIT_INT8, (None,))[0],
            "rel_err_int8_gpu_vs_cpu_product_term":
                errs.get(xm.PRECISION_SPLIT_INT8, (None, None))[1]}


def rel_err_term(got, ref, c0, beta):
    """||got - ref|| / ||ref - beta * c0|| over the blocks of ref."""
    import numpy as np
    num = den = 0.0
    for i, r in ref.items():
        r = np.asarray(r, np.complex128)
        num += float(np.sum(np.abs(np.asarray(got[i], np.complex128) - r) ** 2))
        den += float(np.sum(np.abs(r - beta * np.asarray(c0[i])) ** 2))
    return (num / den) ** 0.5 if den else None


def cpu_ring_sample(lib, w, galloc, t, v, c, c2, frac=0.1):
    """libxm on the host cores (or the oracle, if it is absent) beside the
    GPU for the ring: the first canonical C blocks holding about `frac` of
    the FLOPs, contracted from the same operand blocks, then xm_add and
    xm_set on those C blocks; each also run on the GPU on the same sample
    (contraction within 1e-13, element operations bit for bit)."""
    import numpy as np
    from libxm_amd import xm
    from libxm_amd.spec import Spec, rel_err
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libxmref.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    kind = "reference" if os.path.exists(ref_so) else "port"
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    cpu = xm.XmLib(ref_so if kind == "reference" else port_so)
    ncores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    pairs_all = lib.get_pairs(t, v, c, *w.idx)
    f_of = {}
    for p in pairs_all:
        f_of[p["c"]] = f_of.get(p["c"], 0) + 2.0 * p["m"] * p["n"] * p["k"]
    target = frac * sum(f_of.values())
    chosen, fsum = [], 0.0
    for idx in w.c.canonical_blocks():
        if chosen and fsum + f_of.get(idx, 0) > target:
            break
        chosen.append(idx)
        fsum += f_of.get(idx, 0)
    sel = set(chosen)
    need_a = {p["a"] for p in pairs_all if p["c"] in sel}
    need_b = {p["b"] for p in pairs_all if p["c"] in sel}
    cs = Spec(w.c.dims, {k: p for k, p in enumerate(w.c.points)}, w.c.type)
    for idx in chosen:
        cs.canon(idx)
    alloc = cpu.allocator()
    ct, cv = w.specs[0].build(cpu, alloc), w.specs[1].build(cpu, alloc)
    cc, cc2 = cs.build(cpu, alloc), cs.build(cpu, alloc)
    for i in need_a:
        cpu.write_block(ct, i, lib.read_block(t, i))
    for i in need_b:
        cpu.write_block(cv, i, lib.read_block(v, i))
    c0 = {i: lib.read_block(c, i) for i in chosen}
    d2 = {i: lib.read_block(c2, i) for i in chosen}
    gc, gc2 = cs.build(lib, galloc), cs.build(lib, galloc)
    out = {"kind": kind, "cores": ncores,
           "sample": "%d of %d canonical C blocks, %.3g GFLOP, %.2f GB of C" % (
               len(chosen), len(w.c.canonical_blocks()), fsum / 1e9,
               cs.canonical_bytes() / 1e9)}
    for x, y, d in ((cc, gc, c0), (cc2, gc2, d2)):
        for i in chosen:
            cpu.write_block(x, i, d[i])
            lib.write_block(y, i, d[i])
    t0 = time.perf_counter()
    cpu.contract(w.alpha, ct, cv, w.beta, cc, *w.idx)
    tc = time.perf_counter() - t0
    lib.contract(w.alpha, t, v, w.beta, gc, *w.idx)
    out["contract"] = {"tflops": fsum / tc / 1e12, "seconds": tc,
                       "rel_err_gpu_vs_cpu": rel_err(cs.read(lib, gc),
                                                     cs.read(cpu, cc))}
    cb = cs.canonical_bytes()
    for i in chosen:        # both from the same C again
        cpu.write_block(cc, i, c0[i])
        lib.write_block(gc, i, c0[i])
    t0 = time.perf_counter()
    cpu.add(1.0, cc, 0.5, cc2, "ijab", "ijab")
    ta = time.perf_counter() - t0
    lib.add(1.0, gc, 0.5, gc2, "ijab", "ijab")
    ga, ca = cs.read(lib, gc), cs.read(cpu, cc)
    t0 = time.perf_counter()
    cpu.set(cc, 0.0)
    tsv = time.perf_counter() - t0
    out["xm_add"] = {"gbs": 3 * cb / ta / 1e9, "seconds": ta,
                     "bit_exact": all(np.array_equal(ga[k], ca[k])
                                      for k in ca)}
    out["xm_set"] = {"gbs": cb / tsv / 1e9, "seconds": tsv}
    for x in (ct, cv, cc, cc2):
        cpu.tensor_free(x)
    cpu.allocator_destroy(alloc)
    for x in (gc, gc2):
        lib.tensor_free(x)
    return out


def split_accuracy(lib, w, alloc, t, v, nslices,
                   precision=None):
    """Product term alone (beta = 0): split vs FP64 on the whole workload."""
    from libxm_amd import xm
    from libxm_amd.spec import rel_err
    c1, c2 = w.c.build(lib, alloc), w.c.build(lib, alloc)
    lib.set_precision(xm.PRECISION_NATIVE)
    lib.contract(w.alpha, t, v, 0.0, c1, *w.idx)
    lib.set_precision(xm.PRECISION_SPLIT_BF16 if precision is None
                      else precision, nslices)
    lib.contract(w.alpha, t, v, 0.0, c2, *w.idx)
    lib.set_precision(xm.PRECISION_NATIVE)
    err = rel_err(w.c.read(lib, c2), w.c.read(lib, c1))
    lib.tensor_free(c1)
    lib.tensor_free(c2)
    return err


def copy_calibration(lib, steps):
    """xm_copy(b, 1, a, "ij", "ij") of 2^29 doubles in 64 blocks: every
    launch reads 4 GiB and writes 4 GiB."""
    from libxm_amd import xm
    from libxm_amd.spec import Spec
    sp = Spec((1 << 15, 1 << 14), {0: [k << 9 for k in range(1, 64)]},
              xm.SCALAR_DOUBLE).all_canonical()
    alloc = lib.allocator()
    a, b = sp.build(lib, alloc), sp.build(lib, alloc)
    lib.fill_random(a, 1)
    dt = 0.0
    for _ in range(max(steps, 1)):
        t0 = time.perf_counter()
        lib.copy(b, 1.0, a, "ij", "ij")
        dt = time.perf_counter() - t0
    print(json.dumps({"mode": "copy", "bytes_read": 8 << 29,
                      "bytes_written": 8 << 29, "seconds": dt,
                      "GBps": 2 * (8 << 29) / dt / 1e9}), flush=True)


def canonical_bytes(spec, esize):
    return sum(spec.block_size(i) for i in spec.canonical_blocks()) * esize


def timed(lib, fn, steps):
    """Median wall time (ms) of a synchronous library call."""
    fn()
    lib.synchronize()
    ts = []
    for _ in range(steps):
        t0 = time.perf_counter()
        fn()
        lib.synchronize()
        ts.append((time.perf_counter() - t0) * 1e3)
    return sorted(ts)[len(ts) // 2]


def ring_mode(lib, args):
    """Ring c(i,j,a,b) = sum t(i,k,a,c) w(k,b,c,j), o=96 v=480, block sizes
    U[4,28] (~2.5e5 block pairs, unfolds of both operands), FP64 and
    complex64; then xm_add(1, c, 0.5, c2) and xm_set(c, 0) on the result.
    The FP64 form also times libxm on the host cores on a sample of C
    blocks (contraction, xm_add, xm_set), unless --no-cpu."""
    from libxm_amd import xm, workloads
    out = {"mode": "ring", "unit": "TFLOP/s", "forms": {}}
    for name, tp, esize, mult in (("f64", xm.SCALAR_DOUBLE, 8, 2.0),
                                  ("c64", xm.SCALAR_FLOAT_COMPLEX, 8, 8.0)):
        w = workloads.ring_d2h(96, 480, args.seed, tp)
        alloc = lib.allocator()
        t, v, c = (s.build(lib, alloc) for s in w.specs)
        for k, x in enumerate((t, v, c)):
            lib.fill_random(x, args.seed * 16 + k)
        tp0 = time.perf_counter()
        lib.contract(w.alpha, t, v, w.beta, c, *w.idx)
        lib.synchronize()
        first = (time.perf_counter() - tp0) * 1e3
        # two more calls finish the tile-shape tuning
        ms, gemm_ms, prep_ms, st = time_steps(lib, w, t, v, c, args.steps,
                                              max(args.warmup, 2))
        F = st["flops"]         # 8 real flops per complex multiply-add
        unf = st["panel_bytes"] + (st["cin_bytes"] if w.beta != 0 else 0.0)
        f = {"ms_per_call": ms, "tflops": F / ms / 1e9, "flops": F,
             "gemm_ms": gemm_ms, "gemm_tflops": F / gemm_ms / 1e9,
             "block_pairs": st["npairs"], "c_blocks": st["nblocks_c"],
             "tile": [st["tile_p"], st["tile_q"]],
             "gemm_kernel": st["gemm_kernel"],
             "dense_classes": st["classes"],
             "class_c_blocks": st["class_cblocks"],
             "first_call_ms_incl_plan": first,
             "unfold_panels": st["nunfold"],
             "unfold": {"bytes": unf, "ms": prep_ms,
                        "gbs": unf / prep_ms / 1e6 if prep_ms else None,
                        "frac_of_8tbs": unf / prep_ms / 1e6 / HBM_PEAK
                        if prep_ms else None},
             "fold": {"bytes": st["cout_bytes"], "ms": st["post_ms"],
                      "gbs": st["cout_bytes"] / st["post_ms"] / 1e6
                      if st["post_ms"] else None}}
        c2 = w.c.build(lib, alloc)
        lib.fill_random(c2, args.seed * 16 + 7)
        cb = canonical_bytes(w.c, esize)
        add_ms = timed(lib, lambda: lib.add(1.0, c, 0.5, c2, "ijab", "ijab"),
                       args.steps)
        set_ms = timed(lib, lambda: lib.set(c, 0.0), args.steps)
        f["xm_add"] = {"ms": add_ms, "bytes": 3 * cb,
                       "gbs": 3 * cb / add_ms / 1e6,
                       "frac_of_8tbs": 3 * cb / add_ms / 1e6 / HBM_PEAK}
        f["xm_set"] = {"ms": set_ms, "bytes": cb, "gbs": cb / set_ms / 1e6,
                       "frac_of_8tbs": cb / set_ms / 1e6 / HBM_PEAK}
        if name == "f64" and not args.no_cpu:
            f["cpu_baseline"] = cpu_ring_sample(lib, w, alloc, t, v, c, c2)
        log("ring %s: %.3f ms/call %.2f TFLOP/s (GEMM %.3f ms, %s, %d "
            "classes), unfold %.1f GB %.3f ms, fold %.1f GB %.3f ms, add "
            "%.1f GB/s, set %.1f GB/s" % (
                name, ms, F / ms / 1e9, gemm_ms, st["gemm_kernel"],
                st["classes"], unf / 1e9, prep_ms,
                st["cout_bytes"] / 1e9, st["post_ms"], f["xm_add"]["gbs"],
                f["xm_set"]["gbs"]))
        out["forms"][name] = f
        for x in (t, v, c, c2):
            lib.tensor_free(x)
        lib.allocator_destroy(alloc)
    print(json.dumps(out), flush=True)


def precision_mode(lib, args):
    """Ladder o=64 v=512, entries U[-1,1]*10^U[-6,0]: FP64, split (6 and 7
    BF16 slices; the int8 CRT form with the moduli chosen per plan) and
    FP32, each timed and compared with the FP64 GPU result."""
    from libxm_amd import xm, workloads
    from libxm_amd.spec import rel_err
    w = workloads.ladder_d2h(64, 512, args.seed, value_exp=(-6.0, 0.0))
    alloc = lib.allocator()
    t, v, c = (s.build(lib, alloc) for s in w.specs)
    for k, x in enumerate((t, v, c)):
        lib.fill_random(x, args.seed * 16 + k, -1.0, 1.0, -6.0, 0.0)
        for ia, ib in workloads.antisym_strings(4, w.pairs):
            lib.add(0.5, x, -0.5, x, ia, ib)
    out = {"mode": "precision", "unit": "TFLOP/s (FP64-equivalent)",
           "workload": "ladder o=64 v=512 D2h, U[-1,1]*10^U[-6,0]", "forms": {}}
    c1 = w.c.build(lib, alloc)
    results = {}
    F = None
    for name, prec, ns in (("fp64", xm.PRECISION_NATIVE, None),
                           ("split6", xm.PRECISION_SPLIT_BF16, 6),
                           ("split7", xm.PRECISION_SPLIT_BF16, 7),
                           ("int8", xm.PRECISION_SPLIT_INT8, None)):
        lib.set_precision(prec, ns, 0 if name == "int8" else None)
        ms, gemm_ms, prep_ms, st = time_steps(lib, w, t, v, c, args.steps,
                                              max(args.warmup, 3))
        lib.contract(w.alpha, t, v, 0.0, c1, *w.idx)
        results[name] = w.c.read(lib, c1)
        F = F or st["flops"]
        out["forms"][name] = {"ms_per_call": ms, "gemm_ms": gemm_ms,
                              "slice_ms": prep_ms if prec else None,
                              "tflops": F / ms / 1e9}
        if name == "int8":
            out["forms"][name]["moduli"] = st["nslices"]
            out["forms"][name]["err_bound_model"] = st["err_bound"]
    lib.set_precision(xm.PRECISION_NATIVE)
    wf = workloads.ladder_d2h(64, 512, args.seed, type_=xm.SCALAR_FLOAT)
    tf, vf, cf = (s.build(lib, alloc) for s in wf.specs)
    for dst, src in ((tf, t), (vf, v), (cf, c)):
        lib.copy(dst, 1.0, src, "abcd", "abcd")
    ms, gemm_ms, _, _ = time_steps(lib, wf, tf, vf, cf, args.steps, args.warmup)
    lib.contract(w.alpha, tf, vf, 0.0, cf, *w.idx)
    results["fp32"] = wf.c.read(lib, cf)
    out["forms"]["fp32"] = {"ms_per_call": ms, "gemm_ms": gemm_ms,
                            "tflops": F / ms / 1e9}
    for name in ("split6", "split7", "int8", "fp32"):
        out["forms"][name]["rel_err_vs_fp64_gpu"] = rel_err(
            results[name], results["fp64"])
    out["flops"] = F
    for name, f in out["forms"].items():
        log("precision %s: %.3f ms/call %.2f TFLOP/s%s" % (
            name, f["ms_per_call"], f["tflops"],
            "" if name == "fp64" else ", rel err %.2e" % f["rel_err_vs_fp64_gpu"]))
    print(json.dumps(out), flush=True)


# roofline.traffic is not measured in the timed run: it is the HBM bytes per
# launch from rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this workload,
# scaled by the same counters' readings for a copy of known size
TRAFFIC_SOURCE = ("profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE / "
                  "WRITE_SIZE of this workload, copy-calibrated (read x2)")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary of
    this workload (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        return json.load(open(path)).get(kernel)
    except (OSError, ValueError):
        return None


def stage_traffic(kernels):
    """Sum of measured_traffic over the kernels of one stage, or None if
    the first (the stage's main kernel) has no figure."""
    first = measured_traffic(kernels[0])
    if first is None:
        return None
    return first + sum(measured_traffic(k) or 0 for k in kernels[1:])


def permute_mode(lib, args):
    """xm_copy(b, 1, a, "ijkl", perm) on 4-D FP64 tensors: 128^4 split into
    blocks of 32 and of 16, and 112^4 split into uneven blocks of 4-28 (the
    ring's sizes); GB/s = (bytes read + written) / median wall time of the
    call (warm: job lists cached), with the jobs' permute form."""
    import random
    from libxm_amd import xm
    from libxm_amd.spec import Spec
    rng = random.Random(args.seed + 3)
    cuts, x = [], 0
    while True:
        x += rng.randint(4, 28)
        if x >= 112 - 3:
            break
        cuts.append(x)
    cases = (("128^4 blocks 32", 128, [k * 32 for k in range(1, 4)]),
             ("128^4 blocks 16", 128, [k * 16 for k in range(1, 8)]),
             ("112^4 blocks 4-28", 112, cuts))
    out = {"mode": "permute", "unit": "GB/s", "peak": HBM_PEAK, "cases": []}
    for name, n, pts in cases:
        sp = Spec((n,) * 4, {d: pts for d in range(4)},
                  xm.SCALAR_DOUBLE).all_canonical()
        alloc = lib.allocator()
        a, b = sp.build(lib, alloc), sp.build(lib, alloc)
        lib.fill_random(a, 1)
        nbytes = 2 * 8 * n ** 4
        for perm in ("ijkl", "jikl", "ikjl", "lkji", "klij", "ljki"):
            ms = timed(lib, lambda: lib.copy(b, 1.0, a, "ijkl", perm),
                       max(args.steps, 5))
            tiled, gather = lib.last_permute()
            out["cases"].append({"tensor": name, "perm": perm,
                                 "bytes": nbytes, "ms": ms,
                                 "gbs": nbytes / ms / 1e6,
                                 "frac": nbytes / ms / 1e6 / HBM_PEAK,
                                 "jobs_tiled": tiled,
                                 "jobs_gather": gather})
        lib.tensor_free(a)
        lib.tensor_free(b)
        lib.allocator_destroy(alloc)
        if name == cases[0][0] and not args.no_cpu:
            out["cpu_baseline"] = cpu_permute(sp)
    print(json.dumps(out), flush=True)


def cpu_permute(sp):
    """libxm's xm_copy on the host cores (or the oracle's) for the same
    tensors and permutations, one run each after a warm-up run."""
    from libxm_amd import xm
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libxmref.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    kind = "reference" if os.path.exists(ref_so) else "port"
    cpu = xm.XmLib(ref_so if kind == "reference" else port_so)
    alloc = cpu.allocator()
    a, b = sp.build(cpu, alloc), sp.build(cpu, alloc)
    cpu.set(a, 1.0)
    cpu.copy(b, 1.0, a, "ijkl", "ijkl")
    nbytes = 2 * sp.canonical_bytes()
    res = {"kind": kind, "cores": int(os.environ.get("OMP_NUM_THREADS",
                                                     os.cpu_count() or 1)),
           "tensor": "x".join(str(d) for d in sp.dims), "gbs": {}}
    for perm in ("ijkl", "jikl", "ikjl", "lkji", "klij", "ljki"):
        t0 = time.perf_counter()
        cpu.copy(b, 1.0, a, "ijkl", perm)
        res["gbs"][perm] = nbytes / (time.perf_counter() - t0) / 1e9
    cpu.tensor_free(a)
    cpu.tensor_free(b)
    cpu.allocator_destroy(alloc)
    return res


def backward_mode(lib, args):
    """xm_contract_backward of the o=64 v=384 ladder: dt and dv from dc (c's
    structure, U[-1,1]); FP64, split BF16 (5 slices) and int8 CRT; TFLOP/s
    of the two adjoint contractions' F together (each canonical block's
    images merged into its own GEMM terms), wall time per call."""
    from libxm_amd import xm, workloads
    w = workloads.ladder_d2h(args.o, args.v, args.seed)
    alloc, t, v, c = build_workload(lib, w, args.seed)
    dc = w.c.build(lib, alloc)
    fill_tensor(lib, w, dc, args.seed * 16 + 5)
    dt, dv = w.t.build(lib, alloc), w.v.build(lib, alloc)
    out = {"mode": "backward", "workload": w.name, "forms": {}}
    for name, prec, ns in (("fp64", xm.PRECISION_NATIVE, None),
                           ("split5", xm.PRECISION_SPLIT_BF16, 5),
                           ("int8", xm.PRECISION_SPLIT_INT8, None)):
        lib.set_precision(prec, ns, 0)
        call = lambda: lib.contract_backward(w.alpha, t, v, dc, dt, dv,
                                             *w.idx)
        for _ in range(max(args.warmup, 3)):    # plans and tile tuning
            call()
        ms = timed(lib, call, max(args.steps, 3))
        st = lib.stats()
        out["forms"][name] = {"ms_per_call": ms, "flops": st["flops"],
                              "tflops": st["flops"] / ms / 1e9,
                              "gemm_ms": st["gemm_ms"],
                              "prep_ms": st["prep_ms"],
                              "post_ms": st["post_ms"]}
        log("backward %s: %.3f ms/call %.2f TFLOP/s (GEMM %.3f, permutes "
            "and slices %.3f + %.3f ms)" % (name, ms, st["flops"] / ms / 1e9,
                                          st["gemm_ms"], st["prep_ms"],
                                          st["post_ms"]))
    lib.set_precision(xm.PRECISION_NATIVE)
    print(json.dumps(out), flush=True)


def overflow_mode(lib, args):
    """The o=64 v=384 ladder with t and c in HBM and v (7.4 GB) in pinned
    host memory (arena capped at 1 GiB before v is built): one contraction
    per step, staged (host v blocks copied to HBM on the side stream in
    batches of C blocks) or, with XM_HOST_STAGING=0, read in place."""
    from libxm_amd import xm, workloads
    w = workloads.ladder_d2h(args.o, args.v, args.seed)
    os.environ["XM_ARENA_BYTES"] = str(1 << 30)
    alloc = lib.allocator()
    del os.environ["XM_ARENA_BYTES"]
    t = w.t.build(lib, alloc)
    c = w.c.build(lib, alloc)
    held = []
    while lib.allocator_usage(alloc)[1] == 0:
        held.append(lib.lib.xm_allocator_allocate(alloc, 64 << 20))
    v = w.v.build(lib, alloc)
    for k, x in enumerate((t, v, c)):
        fill_tensor(lib, w, x, args.seed * 16 + k)
    lib.synchronize()
    dev, host = lib.allocator_usage(alloc)
    lib.contract(w.alpha, t, v, w.beta, c, *w.idx)
    lib.synchronize()
    ms = timed(lib, lambda: lib.contract(w.alpha, t, v, w.beta, c, *w.idx),
               max(args.steps, 3))
    st = lib.stats()
    out = {"mode": "overflow", "workload": w.name + ", v in host memory",
           "hbm_bytes": dev, "host_bytes": host,
           "staging": os.environ.get("XM_HOST_STAGING", "1") != "0",
           "ms_per_call": ms, "tflops": st["flops"] / ms / 1e9,
           "gemm_ms": st["gemm_ms"], "batches": st["staged_batches"],
           "staged_bytes": st["staged_bytes"],
           "staged_gbs": st["staged_bytes"] / ms / 1e6}
    print(json.dumps(out), flush=True)
    for p in held:
        lib.lib.xm_allocator_deallocate(alloc, p)


def stream_mode(args):
    """The operand-stream model of the int8 GEMM step
    (libxm_amd/stream_ceiling, built with the library): int8 TOP/s and
    global_load_lds TB/s of its 256 x 128 and 256 x 256 tile steps with the
    operands L2-resident, streamed from HBM, or not fetched."""
    import subprocess
    exe = os.path.join(ROOT, "libxm_amd", "stream_ceiling")
    if not os.path.exists(exe):
        raise RuntimeError("%s is not built (make -C libxm_amd/csrc)" % exe)
    out = subprocess.run([exe], check=True, stdout=subprocess.PIPE,
                         universal_newlines=True).stdout
    for line in out.splitlines():
        rec = json.loads(line)
        rec["mode"] = "stream"
        print(json.dumps(rec), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the first 3 calls of a plan tune its GEMM tile shape
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--o", type=int, default=64)
    ap.add_argument("--v", type=int, default=384)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slices", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-split", action="store_true")
    ap.add_argument("--mode", choices=["contract", "copy", "ring", "precision",
                                       "permute", "overflow", "backward",
                                       "stream"],
                    default="contract",
                    help="copy: xm_copy of a 4 GiB FP64 tensor, a plain "
                         "copy of known size (HBM counter calibration); "
                         "ring: the o=96 v=480 ring in FP64 and complex64 "
                         "with xm_add / xm_set; precision: the o=64 v=512 "
                         "ladder over six decades in FP64, split and FP32; "
                         "permute: xm_copy under six permutations of 4-D "
                         "tensors with blocks of 32, 16 and 4-28; overflow: "
                         "the ladder with v in pinned host memory; backward: "
                         "xm_contract_backward of the ladder; stream: "
                         "the operand-stream model of the int8 GEMM step")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the library runs on a single "
              "MI355X (--gpus 1); sharding is described, not built.",
              file=sys.stderr)
        sys.exit(2)

    if args.mode == "stream":
        return stream_mode(args)
    from libxm_amd import xm, workloads
    lib = xm.load()
    if args.mode == "copy":
        return copy_calibration(lib, args.steps)
    if args.mode == "ring":
        return ring_mode(lib, args)
    if args.mode == "precision":
        return precision_mode(lib, args)
    if args.mode == "permute":
        return permute_mode(lib, args)
    if args.mode == "overflow":
        return overflow_mode(lib, args)
    if args.mode == "backward":
        return backward_mode(lib, args)
    w = workloads.ladder_d2h(args.o, args.v, args.seed)
    log("workload", json.dumps(w.describe()))
    tb = time.perf_counter()
    alloc, t, v, c = build_workload(lib, w, args.seed)
    dev_bytes, host_bytes = lib.allocator_usage(alloc)
    log("inputs resident: %.2f GB HBM (%.1f s)" % (dev_bytes / 1e9,
                                                   time.perf_counter() - tb))

    c0 = None
    if not args.no_cpu:     # the initial C blocks of the CPU sample
        c0 = {i: lib.read_block(c, i) for i in
              sample_blocks(lib, w, t, v, c)[0]}
    lib.set_precision(xm.PRECISION_NATIVE)
    ms, gemm_ms, prep_ms, st = time_steps(lib, w, t, v, c, args.steps,
                                          args.warmup)
    F = st["flops"]
    log("fp64: %.3f ms/step, gemm %.3f ms, %.2f TFLOP/s" % (
        ms, gemm_ms, F / ms / 1e9))
    result = {
        "metric": "xm_contract nonzero-block TFLOP/s: FP64 vs split-BF16 at rel err<=1e-12",
        "value": F / (ms * 1e-3) / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (device-generated U[-1,1], seed %d)" % args.seed,
        "config": dict(w.describe(), **{
            "workload": "ladder o=%d v=%d D2h, fp64" % (args.o, args.v),
            "flops_per_step": F, "flops_pre_merge": st["flops_pre"],
            "block_pairs": st["npairs"], "c_blocks": st["nblocks_c"],
            "hbm_bytes_resident": dev_bytes}),
        "roofline": {"bound": "mfma", "achieved": F / (gemm_ms * 1e-3) / 1e12,
                     "peak": FP64_PEAK, "unit": "TFLOP/s",
                     "frac": F / (gemm_ms * 1e-3) / 1e12 / FP64_PEAK,
                     # the ladder's panels are all P-contiguous with even
                     # extents: the DMA-staged kernels run it (the stacked
                     # 128 x 128 form or 64 x 128 tiles, whose K-split
                     # halves k_ksplit_sum_s / k_ksplit_sum add)
                     "traffic": stage_traffic((st["gemm_kernel"],
                                               "k_ksplit_sum_s"
                                               if st["gemm_kernel"] ==
                                               "k_gemm_f64_dma_s" else
                                               "k_ksplit_sum")
                                              if st["ksplit_tiles"] else
                                              (st["gemm_kernel"],)),
                     "traffic_source": TRAFFIC_SOURCE,
                     "kernel": st["gemm_kernel"], "kernel_ms": gemm_ms,
                     "ksplit_tiles": st["ksplit_tiles"],
                     "tile": [st["tile_p"], st["tile_q"]],
                     "tile_tuning_in_timed_steps": bool(st["tile_tuning"])},
    }
    if not args.no_split:
        lib.set_precision(xm.PRECISION_SPLIT_BF16, args.slices)
        sms, sgemm, sprep, sst = time_steps(lib, w, t, v, c, args.steps,
                                            args.warmup)
        lib.set_precision(xm.PRECISION_NATIVE)
        nprod = products(args.slices)
        err = split_accuracy(lib, w, alloc, t, v, args.slices)
        log("split: %.3f ms/step (slice %.3f, gemm %.3f), %.2f TFLOP/s, "
            "rel err %.2e" % (sms, sprep, sgemm, F / sms / 1e9, err))
        result["split_bf16"] = {
            "value": F / (sms * 1e-3) / 1e12,
            "unit": "TFLOP/s (FP64-equivalent, same F)",
            "ms_per_step": sms, "slice_ms": sprep, "gemm_ms": sgemm,
            "nslices": args.slices, "bf16_products_per_fp64_product": nprod,
            "speedup_vs_fp64": ms / sms,
            "rel_err_vs_fp64_gpu": err, "meets_1e-12": err <= 1e-12,
            "roofline": {"bound": "mfma",
                         "achieved": F * nprod / (sgemm * 1e-3) / 1e12,
                         "peak": BF16_PEAK, "unit": "TFLOP/s (BF16 MFMA)",
                         "frac": F * nprod / (sgemm * 1e-3) / 1e12 / BF16_PEAK,
                         "traffic": measured_traffic("k_gemm_split2"),
                         "traffic_source": TRAFFIC_SOURCE,
                         "kernel": "k_gemm_split2", "kernel_ms": sgemm}}
    if not args.no_split:
        lib.set_precision(xm.PRECISION_SPLIT_INT8, None, 0)
        ims, igemm, iprep, ist = time_steps(lib, w, t, v, c, args.steps,
                                            args.warmup)
        lib.set_precision(xm.PRECISION_NATIVE)
        nmod = ist["nslices"]
        ierr = split_accuracy(lib, w, alloc, t, v, None,
                              xm.PRECISION_SPLIT_INT8)
        log("int8 CRT: %.3f ms/step (residues %.3f, gemm %.3f), %.2f "
            "TFLOP/s, %d moduli, rel err %.2e" % (
                ims, iprep, igemm, F / ims / 1e9, nmod, ierr))
        result["split_int8"] = {
            "value": F / (ims * 1e-3) / 1e12,
            "unit": "TFLOP/s (FP64-equivalent, same F)",
            "ms_per_step": ims, "residue_ms": iprep, "gemm_ms": igemm,
            "moduli": nmod, "int8_products_per_fp64_product": nmod,
            "speedup_vs_fp64": ms / ims, "err_bound_model": ist["err_bound"],
            "rel_err_vs_fp64_gpu": ierr, "meets_1e-12": ierr <= 1e-12,
            "form_ran": ist["precision"], "kernel": ist["gemm_kernel"],
            "roofline": {"bound": "mfma",
                         # the GEMM stage is the int8 GEMM and the
                         # reconstruction of C
                         "achieved": F * nmod / (igemm * 1e-3) / 1e12,
                         "peak": INT8_PEAK, "unit": "TOP/s (int8 MFMA)",
                         "frac": F * nmod / (igemm * 1e-3) / 1e12 / INT8_PEAK,
                         "traffic": stage_traffic(("k_gemm_crt",
                                                   "k_crt_recon")),
                         "traffic_source": TRAFFIC_SOURCE,
                         "kernel": "k_gemm_crt", "kernel_ms": igemm}}
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(
                lib, w, alloc, t, v, c, c0,
                None if args.no_split else args.slices)
        except Exception as e:  # reported, never silently replaced
            result["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
