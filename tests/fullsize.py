"""Full-size workloads on the GPU checked against the oracle on a sample.

A contraction at the workloads' real sizes runs whole on the GPU; the CPU
oracle then recomputes a bounded, seeded sample of canonical C blocks (up
to `budget` FLOPs) from the GPU tensors' own input blocks, into a C
structure in which only the sampled blocks are canonical.  Only the source
blocks those C blocks read are copied to the host.
"""
import numpy as np

from libxm_amd import workloads, xm
from libxm_amd.spec import Spec, rel_err


def fill(lib, w, x, seed, wexp=(0.0, 0.0)):
    """Canonical entries U[-1,1] * 10^U[wexp]; diagonal blocks of an
    antisymmetric pair antisymmetrised (ladder)."""
    lib.fill_random(x, seed, -1.0, 1.0, wexp[0], wexp[1])
    for ia, ib in workloads.antisym_strings(4, getattr(w, "pairs", ())):
        lib.add(0.5, x, -0.5, x, ia, ib)


def source(spec, idx):
    k = spec.kind(idx)
    return idx if k == "c" else k[1]


def pick_cblocks(pairs, cblocks, budget, seed, mult=2.0):
    """Seeded sample of canonical C blocks with nonzero work, at most
    `budget` FLOPs, small and large blocks alike."""
    f_of = {}
    for p in pairs:
        f_of[p["c"]] = f_of.get(p["c"], 0.0) + mult * p["m"] * p["n"] * p["k"]
    rng = np.random.default_rng(seed)
    order = [cblocks[i] for i in rng.permutation(len(cblocks))]
    chosen, fsum = [], 0.0
    for idx in order:
        f = f_of.get(idx, 0.0)
        if f == 0.0 or fsum + f > budget:
            continue
        chosen.append(idx)
        fsum += f
    return chosen, fsum


class Sample:
    """The oracle's recomputation of `chosen` C blocks of
    c = alpha * a * b + beta * c0, inputs read from the GPU tensors."""

    def __init__(self, gpu, cpu, w, specs, tensors, c0, chosen, alpha, beta,
                 idx, pairs):
        sa, sb, sc = specs
        a, b, _ = tensors
        sel = set(chosen)
        need_a = {source(sa, p["a"]) for p in pairs if p["c"] in sel}
        need_b = {source(sb, p["b"]) for p in pairs if p["c"] in sel}
        cs = Spec(sc.dims, {k: pt for k, pt in enumerate(sc.points)}, sc.type)
        for i in chosen:
            cs.canon(i)
        self.cpu, self.alloc = cpu, cpu.allocator()
        self.ca, self.cb = sa.build(cpu, self.alloc), sb.build(cpu, self.alloc)
        self.cc = cs.build(cpu, self.alloc)
        for i in need_a:
            cpu.write_block(self.ca, i, gpu.read_block(a, i))
        for i in need_b:
            cpu.write_block(self.cb, i, gpu.read_block(b, i))
        for i in chosen:
            cpu.write_block(self.cc, i, c0[i])
        cpu.contract(alpha, self.ca, self.cb, beta, self.cc, *idx)
        self.out = {i: cpu.read_block(self.cc, i) for i in chosen}

    def free(self):
        for x in (self.ca, self.cb, self.cc):
            self.cpu.tensor_free(x)
        self.cpu.allocator_destroy(self.alloc)


def read_blocks(lib, t, chosen):
    return {i: lib.read_block(t, i) for i in chosen}


def err(gpu_blocks, ref_blocks):
    return rel_err({k: np.asarray(v, np.complex128) for k, v in gpu_blocks.items()},
                   {k: np.asarray(v, np.complex128) for k, v in ref_blocks.items()})


__all__ = ["fill", "pick_cblocks", "Sample", "read_blocks", "err", "xm"]


def pick_with_largest(pairs, cblocks, budget, seed, mult=2.0):
    """pick_cblocks, with the C block of the largest P x Q extent first."""
    f_of, size = {}, {}
    for p in pairs:
        f_of[p["c"]] = f_of.get(p["c"], 0.0) + mult * p["m"] * p["n"] * p["k"]
        size[p["c"]] = p["m"] * p["n"]
    big = max(size, key=lambda i: (size[i], f_of[i]))
    rest, fsum = pick_cblocks(pairs, [i for i in cblocks if i != big],
                              budget - f_of[big], seed, mult)
    return [big] + rest, fsum + f_of[big]


def canonical_view(spec):
    """Spec with spec's canonical blocks only (derivative blocks zero)."""
    cv = Spec(spec.dims, {k: p for k, p in enumerate(spec.points)}, spec.type)
    for idx in spec.canonical_blocks():
        cv.canon(idx)
    return cv


def fold_sets(spec, sources):
    """For each canonical block S in `sources`, the blocks X whose data is S
    (S itself first) with X's permutation and scalar: X = s * S(perm)."""
    out = {s: [(s, tuple(range(spec.ndim)), 1.0)] for s in sources}
    for idx in spec.grid():
        k = spec.kind(idx)
        if isinstance(k, tuple) and k[1] in out:
            out[k[1]].append((idx, k[2], k[3].real))
    return out


def fold_block(spec, src, members, g_blocks):
    """Gradient w.r.t. canonical block src: sum over its members X of
    s_X * G_X read back through X's permutation (the inverse of
    X(e) = s * S(permute(e, p)))."""
    acc = np.zeros(spec.block_dims(src))
    for x, perm, s in members:
        g = np.asarray(g_blocks[x], np.float64).reshape(spec.block_dims(x),
                                                          order="F")
        acc += s * np.transpose(g, np.argsort(perm))
    return acc.ravel(order="F")
