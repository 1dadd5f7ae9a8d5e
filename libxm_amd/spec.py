"""Replayable block-tensor recipes.

A `Spec` records the calls that build a block tensor through libxm's C API
(block-space splits, then set_canonical / set_derivative calls in order) so
the same structure can be built, call for call, in any library exporting
the API.  It also knows how to expand canonical block data into the full
dense tensor following libxm's derivative-block rule
X(e) = s * S(permute(e, p)) (libxm src/tensor.c:154-178), which the test
suite uses as an implementation-independent check.
"""
import itertools

import numpy as np

from . import xm


class Spec:
    def __init__(self, dims, splits=None, type_=xm.SCALAR_DOUBLE):
        self.dims = tuple(int(d) for d in dims)
        # per dimension: sorted interior split points
        self.points = [sorted(set(int(x) for x in (splits or {}).get(k, [])
                                  if 0 < int(x) < self.dims[k]))
                       for k in range(len(self.dims))]
        self.type = type_
        self.ops = []          # ("c", idx) | ("d", idx, src, perm, s)
        self._kind = {}        # idx -> "c" / ("d", src, perm, s)

    # ---- geometry ------------------------------------------------------
    @property
    def ndim(self):
        return len(self.dims)

    def edges(self, k):
        return [0] + self.points[k] + [self.dims[k]]

    @property
    def nblocks(self):
        return tuple(len(self.points[k]) + 1 for k in range(self.ndim))

    def block_dims(self, idx):
        return tuple(self.edges(k)[idx[k] + 1] - self.edges(k)[idx[k]]
                     for k in range(self.ndim))

    def block_size(self, idx):
        return int(np.prod(self.block_dims(idx), dtype=np.int64))

    def block_offsets(self, idx):
        return tuple(self.edges(k)[idx[k]] for k in range(self.ndim))

    def grid(self):
        """Block indices, first index fastest (libxm's iteration order)."""
        nb = self.nblocks
        for rev in itertools.product(*[range(n) for n in reversed(nb)]):
            yield tuple(reversed(rev))

    # ---- structure -----------------------------------------------------
    def canon(self, idx):
        idx = tuple(idx)
        if idx in self._kind:
            raise ValueError("block %r already set" % (idx,))
        self._kind[idx] = "c"
        self.ops.append(("c", idx))
        return self

    def deriv(self, idx, src, perm, s):
        idx, src, perm = tuple(idx), tuple(src), tuple(perm)
        if idx in self._kind:
            raise ValueError("block %r already set" % (idx,))
        if self._kind.get(src) != "c":
            raise ValueError("source %r is not canonical" % (src,))
        self._kind[idx] = ("d", src, perm, complex(s))
        self.ops.append(("d", idx, src, perm, complex(s)))
        return self

    def all_canonical(self):
        for idx in self.grid():
            if idx not in self._kind:
                self.canon(idx)
        return self

    def kind(self, idx):
        return self._kind.get(tuple(idx))

    def canonical_blocks(self):
        return [idx for idx in self.grid() if self._kind.get(idx) == "c"]

    def copy_structure(self, type_=None):
        s = Spec(self.dims, {k: p for k, p in enumerate(self.points)},
                 self.type if type_ is None else type_)
        s.ops = list(self.ops)
        s._kind = dict(self._kind)
        return s

    def canonical_bytes(self):
        esz = xm.NP_DTYPE[self.type].itemsize
        return sum(self.block_size(i) for i in self.canonical_blocks()) * esz

    # ---- building in a library ------------------------------------------
    def build(self, lib, alloc, fake_ptrs=None):
        """Replay the recipe through lib's C API; returns the tensor.
        With fake_ptrs (an iterator of distinct integers) canonical blocks
        get those data pointers through set_canonical_block_raw instead of
        storage: enough for structure-only work, no data access allowed."""
        bs = lib.bs_create(self.dims, {k: p for k, p in enumerate(self.points)})
        t = lib.tensor(bs, self.type, alloc)
        lib.bs_free(bs)
        for op in self.ops:
            if op[0] == "c":
                if fake_ptrs is None:
                    lib.set_canonical(t, op[1])
                else:
                    lib.lib.xm_tensor_set_canonical_block_raw(
                        t, xm.dim(op[1]), next(fake_ptrs))
            else:
                lib.set_derivative(t, op[1], op[2], op[3], op[4])
        return t

    # ---- data ------------------------------------------------------------
    def random_data(self, rng, lo=-1.0, hi=1.0):
        dt = xm.NP_DTYPE[self.type]
        out = {}
        for idx in self.canonical_blocks():
            n = self.block_size(idx)
            v = rng.uniform(lo, hi, n)
            if dt.kind == "c":
                v = v + 1j * rng.uniform(lo, hi, n)
            out[idx] = v.astype(dt)
        return out

    def write(self, lib, t, data):
        for idx, v in data.items():
            lib.write_block(t, idx, v)

    def read(self, lib, t):
        return {idx: lib.read_block(t, idx) for idx in self.canonical_blocks()}

    def dense(self, data, dtype=None):
        """Full tensor from canonical data (1-D arrays in storage order)."""
        dt = dtype or (np.complex128 if xm.NP_DTYPE[self.type].kind == "c"
                       else np.float64)
        out = np.zeros(self.dims, dtype=dt)
        for idx in self.grid():
            kind = self._kind.get(idx)
            if kind is None:
                continue
            bd = self.block_dims(idx)
            if kind == "c":
                blk = np.asarray(data[idx]).reshape(bd, order="F").astype(dt)
            else:
                _, src, perm, s = kind
                sd = self.block_dims(src)
                S = np.asarray(data[src]).reshape(sd, order="F").astype(dt)
                blk = (s if np.dtype(dt).kind == "c" else s.real) * \
                    np.transpose(S, perm)
            off = self.block_offsets(idx)
            out[tuple(slice(o, o + d) for o, d in zip(off, bd))] = blk
        return out

    def blocks_of(self, dense):
        """Canonical blocks of a dense tensor, 1-D in storage order."""
        out = {}
        for idx in self.canonical_blocks():
            off = self.block_offsets(idx)
            bd = self.block_dims(idx)
            out[idx] = np.asarray(dense[tuple(slice(o, o + d) for o, d in
                                              zip(off, bd))]).ravel(order="F")
        return out


def rel_err(got, ref):
    """Relative Frobenius error over matching canonical blocks."""
    num = den = 0.0
    for idx, r in ref.items():
        g = np.asarray(got[idx], dtype=np.complex128)
        r = np.asarray(r, dtype=np.complex128)
        num += float(np.sum(np.abs(g - r) ** 2))
        den += float(np.sum(np.abs(r) ** 2))
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return (num / den) ** 0.5
