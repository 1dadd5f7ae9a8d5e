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
GEMM work libxm itself executes).  value = F / (wall time per step).
roofline.achieved = F / (average device time of the grouped GEMM kernel,
HIP events on the library's stream), against 78.6 TFLOP/s FP64 (or the BF16
matrix peak, for the split form).

The CPU baseline runs libxm itself (compiled from its sources into
oracle/_ref, sequential MKL per OpenMP thread) on a bounded sample of the
same workload: the canonical C blocks of the first few (a,b) virtual block
pairs, timed on the host cores, scaled as TFLOP/s of the sample's own F.
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
HBM_PEAK = 8000.0         # GB/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(lib, w, seed=0):
    alloc = lib.allocator()
    t, v, c = (s.build(lib, alloc) for s in w.specs)
    for k, x in enumerate((t, v, c)):
        lib.fill_random(x, seed * 16 + k)
        for ia, ib in __import__("libxm_amd.workloads",
                                 fromlist=["x"]).antisym_strings(4, w.pairs):
            lib.add(0.5, x, -0.5, x, ia, ib)
    lib.synchronize()
    return alloc, t, v, c


def time_steps(lib, w, t, v, c, steps, warmup):
    for _ in range(warmup):
        lib.contract(w.alpha, t, v, w.beta, c, *w.idx)
    lib.synchronize()
    gemm = []
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.contract(w.alpha, t, v, w.beta, c, *w.idx)
        gemm.append(lib.stats()["gemm_ms"])
    lib.synchronize()
    dt = time.perf_counter() - t0
    return dt / steps * 1e3, sum(gemm) / len(gemm), lib.stats()


def cpu_baseline(lib, w, alloc_gpu, t, v, c, budget_s=20.0):
    """libxm (or, if it is absent, the oracle) on a bounded sample."""
    from libxm_amd import xm
    from libxm_amd.spec import Spec, rel_err
    ref_so = os.path.join(ROOT, "oracle", "_ref", "libxmref.so")
    port_so = os.path.join(ROOT, "oracle", "liboracle.so")
    kind = "reference" if os.path.exists(ref_so) else "port"
    os.environ.setdefault("MKL_THREADING_LAYER", "SEQUENTIAL")
    cpu = xm.XmLib(ref_so if kind == "reference" else port_so)
    oracle = xm.XmLib(port_so)
    ncores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # sample: canonical C blocks of the first (a,b) virtual block pairs
    cblocks = w.c.canonical_blocks()
    abs_ = []
    for idx in cblocks:
        if idx[2:] not in abs_:
            abs_.append(idx[2:])
    pairs_all = lib.get_pairs(t, v, c, *w.idx)
    f_of = {}
    for p in pairs_all:
        f_of[p["c"]] = f_of.get(p["c"], 0) + 2.0 * p["m"] * p["n"] * p["k"]
    # ~budget_s at an assumed 0.3 TFLOP/s
    target = 0.3e12 * budget_s
    chosen, fsum = [], 0.0
    for ab in abs_:
        blocks = [x for x in cblocks if x[2:] == ab]
        if chosen and fsum + sum(f_of.get(x, 0) for x in blocks) > target:
            break
        chosen += blocks
        fsum += sum(f_of.get(x, 0) for x in blocks)
    cs = Spec(w.c.dims, {k: p for k, p in enumerate(w.c.points)}, w.c.type)
    for idx in chosen:
        cs.canon(idx)
    need_a = {p["a"] for p in pairs_all if p["c"] in set(chosen)}
    need_b = {p["b"] for p in pairs_all if p["c"] in set(chosen)}

    def src(spec, idx):
        k = spec.kind(idx)
        return idx if k == "c" else k[1]
    need_a = {src(w.t, i) for i in need_a}
    need_b = {src(w.v, i) for i in need_b}
    alloc = cpu.allocator()
    ct, cv = w.t.build(cpu, alloc), w.v.build(cpu, alloc)
    cc = cs.build(cpu, alloc)
    for i in need_a:
        cpu.write_block(ct, i, lib.read_block(t, i))
    for i in need_b:
        cpu.write_block(cv, i, lib.read_block(v, i))
    c0 = {i: lib.read_block(c, i) for i in chosen}
    for i in chosen:
        cpu.write_block(cc, i, c0[i])
    cpu.contract(w.alpha, ct, cv, w.beta, cc, *w.idx)      # warm
    for i in chosen:
        cpu.write_block(cc, i, c0[i])
    t0 = time.perf_counter()
    cpu.contract(w.alpha, ct, cv, w.beta, cc, *w.idx)
    dt = time.perf_counter() - t0
    cpu_out = {i: cpu.read_block(cc, i) for i in chosen}
    # same sample on the GPU: parity of the timed CPU result with ours
    gc = cs.build(lib, alloc_gpu)
    for i in chosen:
        lib.write_block(gc, i, c0[i])
    lib.contract(w.alpha, t, v, w.beta, gc, *w.idx)
    gpu_out = {i: lib.read_block(gc, i) for i in chosen}
    err = rel_err(gpu_out, cpu_out)
    lib.tensor_free(gc)
    for x in (ct, cv, cc):
        cpu.tensor_free(x)
    cpu.allocator_destroy(alloc)
    blas = oracle.lib.xm_oracle_blas_name().decode()
    return {"value": fsum / dt / 1e12, "unit": "TFLOP/s", "cores": ncores,
            "kind": kind,
            "sample": "%d of %d canonical C blocks (first %d (a,b) virtual "
                      "block pairs), %.3g GFLOP, %.2f s; BLAS %s" % (
                          len(chosen), len(cblocks), len(
                              {x[2:] for x in chosen}), fsum / 1e9, dt, blas),
            "rel_err_vs_gpu": err}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--o", type=int, default=64)
    ap.add_argument("--v", type=int, default=384)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--split", choices=["auto", "off", "only"], default="auto")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; this build runs on a single "
              "MI355X (--gpus 1).", file=sys.stderr)
        sys.exit(2)

    from libxm_amd import xm, workloads
    lib = xm.load()
    w = workloads.ladder_d2h(args.o, args.v, args.seed)
    log("workload", json.dumps(w.describe()))
    tb = time.perf_counter()
    alloc, t, v, c = build_workload(lib, w, args.seed)
    dev_bytes, host_bytes = lib.allocator_usage(alloc)
    log("inputs resident: %.2f GB HBM (%.1f s)" % (dev_bytes / 1e9,
                                                   time.perf_counter() - tb))

    lib.set_precision(xm.PRECISION_NATIVE)
    ms, gemm_ms, st = time_steps(lib, w, t, v, c, args.steps, args.warmup)
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
                     "traffic": None, "kernel": "k_gemm_f64",
                     "kernel_ms": gemm_ms},
    }
    if not args.no_cpu:
        try:
            result["cpu_baseline"] = cpu_baseline(lib, w, alloc, t, v, c)
        except Exception as e:  # reported, never silently replaced
            result["cpu_baseline"] = {"error": repr(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
