"""Helpers shared by the test modules."""
import zlib

import numpy as np

from libxm_amd import xm
from libxm_amd.spec import rel_err  # noqa: F401

TYPES = [xm.SCALAR_FLOAT, xm.SCALAR_FLOAT_COMPLEX, xm.SCALAR_DOUBLE,
         xm.SCALAR_DOUBLE_COMPLEX]
TYPE_NAMES = {xm.SCALAR_FLOAT: "f32", xm.SCALAR_FLOAT_COMPLEX: "c64",
              xm.SCALAR_DOUBLE: "f64", xm.SCALAR_DOUBLE_COMPLEX: "c128"}


def is_complex(t):
    return t in (xm.SCALAR_FLOAT_COMPLEX, xm.SCALAR_DOUBLE_COMPLEX)


def is_single(t):
    return t in (xm.SCALAR_FLOAT, xm.SCALAR_FLOAT_COMPLEX)


def rand_scalar(rng, t):
    """libxm's random_scalar: U(-0.5, 0.5) per part (libxm test.c:72-78)."""
    x = rng.uniform(-0.5, 0.5)
    if is_complex(t):
        x = complex(x, rng.uniform(-0.5, 0.5))
    return x


def resolve(v, rng, t):
    return rand_scalar(rng, t) if v == "r" else v


def round_scalar(x, t):
    """A host scalar as the element type holds it."""
    if t == xm.SCALAR_FLOAT:
        return complex(np.float32(complex(x).real))
    if t == xm.SCALAR_FLOAT_COMPLEX:
        return complex(np.complex64(x))
    if t == xm.SCALAR_DOUBLE:
        return complex(complex(x).real)
    return complex(x)


def build_filled(lib, alloc, spec, data):
    t = spec.build(lib, alloc)
    spec.write(lib, t, data)
    return t


def einsum_ref(alpha, A, B, beta, Cd, ia, ib, ic):
    """c = alpha * a * b + beta * c on dense arrays (complex128/float64)."""
    return alpha * np.einsum("%s,%s->%s" % (ia, ib, ic), A, B) + beta * Cd


def max_abs_diff(x, y):
    m = 0.0
    for k in x:
        d = np.max(np.abs(np.asarray(x[k], np.complex128) -
                          np.asarray(y[k], np.complex128))) if np.size(x[k]) else 0.0
        m = max(m, float(d))
    return m


def case_seed(name):
    """A seed fixed by a case's name (Python's str hash changes from
    process to process)."""
    return zlib.crc32(name.encode())


def magnitude_err(got, want, specs, data, alpha, beta, idx):
    """Frobenius norm of got - want over that of the magnitudes the result
    is built from, |alpha| sum_k |a| |b| + |beta| |c|: the error measure of
    the split forms' error bounds, which cancellation in the sums does not
    inflate as it does the error relative to the result."""
    ia, ib, ic = idx
    A = np.abs(specs[0].dense(data[0]))
    B = np.abs(specs[1].dense(data[1]))
    C = np.abs(specs[2].dense(data[2]))
    S = abs(alpha) * np.einsum("%s,%s->%s" % (ia, ib, ic), A, B) + \
        abs(beta) * C
    scale = specs[2].blocks_of(S)
    num = den = 0.0
    for k in want:
        num += float(np.sum(np.abs(np.asarray(got[k], dtype=np.complex128) -
                                   np.asarray(want[k])) ** 2))
        den += float(np.sum(np.abs(scale[k]) ** 2))
    if den == 0.0:
        return 0.0 if num == 0.0 else float("inf")
    return (num / den) ** 0.5
