"""Block structures of the benchmark workloads.

Ladder  c(i,j,a,b) = 0.5 * sum_{c,d} t(i,j,c,d) v(a,b,c,d) + 1.0 c(i,j,a,b)
        xm_contract(0.5, t, v, 1.0, c, "ijcd", "abcd", "ijab")
        (the index strings of libxm test.c:1949).
Ring    c(i,j,a,b) = sum_{k,c} t(i,k,a,c) w(k,b,c,j)
        xm_contract(1.0, t, w, beta, c, "ikac", "kbcj", "ijab").

Occupied (o) and virtual (v) dimensions are split into blocks, each block
carries an irreducible representation (irrep) of an abelian point group
(C2: 2 irreps, D2h: 8); a block is nonzero iff the XOR of its irreps is 0.
Antisymmetric index pairs give derivative blocks exactly as libxm's ladder
test builds them (libxm test.c:835-852): with pairs (0,1) and (2,3),
block (I,J,C,D) with I<=J and C<=D is canonical, and
  (J,I,C,D) = -P01 (I,J,C,D),  (I,J,D,C) = -P23,  (J,I,D,C) = +P01 P23.
"""
import itertools

import numpy as np

from . import xm
from .spec import Spec

LADDER_IDX = ("ijcd", "abcd", "ijab")
RING_IDX = ("ikac", "kbcj", "ijab")


def split_sizes(total, choices, rng):
    """Seeded block sizes drawn from `choices` that sum exactly to total."""
    choices = sorted(set(int(c) for c in choices))
    reach = {0}
    for t in range(1, total + 1):
        if any(t - c in reach for c in choices if c <= t):
            reach.add(t)
    if total not in reach:
        raise ValueError("%d is not a sum of %r" % (total, choices))
    out, rem = [], total
    while rem > 0:
        ok = [c for c in choices if c <= rem and (rem - c) in reach]
        s = int(ok[rng.integers(len(ok))])
        out.append(s)
        rem -= s
    return out


def split_uniform(total, lo, hi, rng):
    """Seeded block sizes uniform in [lo, hi] that sum exactly to total."""
    out, rem = [], total
    while rem > 0:
        ok = [s for s in range(lo, min(hi, rem) + 1)
              if rem - s == 0 or rem - s >= lo]
        s = int(ok[rng.integers(len(ok))])
        out.append(s)
        rem -= s
    return out


def _points(sizes):
    return list(np.cumsum(sizes)[:-1])


def block_tensor(sizes, irreps, pairs=(), type_=xm.SCALAR_DOUBLE):
    """Spec of a tensor whose dimension k is split into sizes[k] with block
    irreps irreps[k]; nonzero iff XOR of irreps is 0; each (p, q) in pairs
    is an antisymmetric index pair (same split on p and q)."""
    nd = len(sizes)
    spec = Spec([sum(s) for s in sizes],
                {k: _points(sizes[k]) for k in range(nd)}, type_)

    def nonzero(idx):
        x = 0
        for k in range(nd):
            x ^= irreps[k][idx[k]]
        return x == 0

    canon = [idx for idx in spec.grid() if nonzero(idx) and
             all(idx[p] <= idx[q] for p, q in pairs)]
    for idx in canon:
        spec.canon(idx)
    for idx in canon:
        swappable = [(p, q) for p, q in pairs if idx[p] < idx[q]]
        for r in range(1, len(swappable) + 1):
            for sub in itertools.combinations(swappable, r):
                perm = list(range(nd))
                tgt = list(idx)
                for p, q in sub:
                    perm[p], perm[q] = perm[q], perm[p]
                    tgt[p], tgt[q] = tgt[q], tgt[p]
                spec.deriv(tuple(tgt), idx, tuple(perm), (-1) ** len(sub))
    return spec


def antisymmetrize_diagonal(spec, data, pairs):
    """Make diagonal blocks antisymmetric within each pair, in place, with
    the arithmetic xm_add(0.5, t, -0.5, t, ...) performs: 0.5*x + (-0.5*y)."""
    for idx, v in data.items():
        bd = spec.block_dims(idx)
        x = np.asarray(v).reshape(bd, order="F")
        for p, q in pairs:
            if idx[p] == idx[q]:
                perm = list(range(len(bd)))
                perm[p], perm[q] = perm[q], perm[p]
                x = 0.5 * x + (-0.5) * np.transpose(x, perm)
        data[idx] = np.ascontiguousarray(x.ravel(order="F"))
    return data


def antisym_strings(ndim, pairs):
    """Index-string pairs for xm_add calls that antisymmetrize diagonal
    blocks: one (idxa, idxb) per antisymmetric pair."""
    base = "abcdefgh"[:ndim]
    out = []
    for p, q in pairs:
        s = list(base)
        s[p], s[q] = s[q], s[p]
        out.append((base, "".join(s)))
    return out


class Ladder:
    """Particle-particle ladder: t(o,o,v,v), v(v,v,v,v), c(o,o,v,v)."""
    idx = LADDER_IDX
    alpha, beta = 0.5, 1.0
    pairs = ((0, 1), (2, 3))

    def __init__(self, o_sizes, o_irreps, v_sizes, v_irreps,
                 type_=xm.SCALAR_DOUBLE, name="ladder", value_exp=(0.0, 0.0)):
        self.o_sizes, self.o_irreps = list(o_sizes), list(o_irreps)
        self.v_sizes, self.v_irreps = list(v_sizes), list(v_irreps)
        self.type = type_
        self.name = name
        self.value_exp = value_exp
        o, oi, v, vi = self.o_sizes, self.o_irreps, self.v_sizes, self.v_irreps
        self.t = block_tensor([o, o, v, v], [oi, oi, vi, vi], self.pairs, type_)
        self.v = block_tensor([v, v, v, v], [vi, vi, vi, vi], self.pairs, type_)
        self.c = block_tensor([o, o, v, v], [oi, oi, vi, vi], self.pairs, type_)

    @property
    def specs(self):
        return self.t, self.v, self.c

    def describe(self):
        return {"workload": self.name, "o_blocks": self.o_sizes,
                "o_irreps": self.o_irreps, "v_blocks": self.v_sizes,
                "v_irreps": self.v_irreps, "index": "%s,%s->%s" % self.idx,
                "alpha": self.alpha, "beta": self.beta}


class Ring:
    """Ring term: t(o,o,v,v) "ikac", w(o,v,v,o) "kbcj", c(o,o,v,v) "ijab";
    every nonzero block canonical (no permutational symmetry)."""
    idx = RING_IDX
    alpha, beta = 1.0, 0.0

    def __init__(self, o_sizes, o_irreps, v_sizes, v_irreps,
                 type_=xm.SCALAR_DOUBLE, name="ring"):
        self.o_sizes, self.o_irreps = list(o_sizes), list(o_irreps)
        self.v_sizes, self.v_irreps = list(v_sizes), list(v_irreps)
        self.type = type_
        self.name = name
        o, oi, v, vi = self.o_sizes, self.o_irreps, self.v_sizes, self.v_irreps
        self.t = block_tensor([o, o, v, v], [oi, oi, vi, vi], (), type_)
        self.v = block_tensor([o, v, v, o], [oi, vi, vi, oi], (), type_)
        self.c = block_tensor([o, o, v, v], [oi, oi, vi, vi], (), type_)

    @property
    def specs(self):
        return self.t, self.v, self.c

    def describe(self):
        return {"workload": self.name, "o_blocks": self.o_sizes,
                "v_blocks": self.v_sizes, "index": "%s,%s->%s" % self.idx,
                "alpha": self.alpha, "beta": self.beta}


def d2h(n):
    """Irreps assigned to n blocks round-robin over the 8 D2h irreps."""
    return [k % 8 for k in range(n)]


def ladder_small(type_=xm.SCALAR_DOUBLE):
    """o=16 in [4,4,4,4] (C2 irreps 0,1,0,1), v=64 in 8 blocks of 8 (C2
    irreps alternating 0,1)."""
    return Ladder([4] * 4, [0, 1, 0, 1], [8] * 8, [k % 2 for k in range(8)],
                  type_, name="ladder o=16 v=64 C2")


def ladder_d2h(o=64, v=384, seed=0, type_=xm.SCALAR_DOUBLE,
               o_choices=(12, 16, 20), v_choices=(24, 32, 40, 48),
               value_exp=(0.0, 0.0)):
    rng = np.random.default_rng(seed)
    os_ = split_sizes(o, o_choices, rng)
    vs_ = split_sizes(v, v_choices, rng)
    return Ladder(os_, d2h(len(os_)), vs_, d2h(len(vs_)), type_,
                  name="ladder o=%d v=%d D2h" % (o, v), value_exp=value_exp)


def ladder_scaled(o, v, seed=0, type_=xm.SCALAR_DOUBLE):
    """A smaller D2h ladder with the same block-size rules, for tests."""
    return ladder_d2h(o, v, seed, type_)


def ring_d2h(o=96, v=480, seed=0, type_=xm.SCALAR_DOUBLE, lo=4, hi=28):
    rng = np.random.default_rng(seed)
    os_ = split_uniform(o, lo, hi, rng)
    vs_ = split_uniform(v, lo, hi, rng)
    return Ring(os_, d2h(len(os_)), vs_, d2h(len(vs_)), type_,
                name="ring o=%d v=%d D2h" % (o, v))
