"""ctypes binding of the libxm C ABI (include/xm.h, include/xm_amd.h).

One `XmLib` object wraps one shared library that exports libxm's C API.
The MI355X library (`libxm_amd/libxm.so`) is loaded with `load()`; the
same class can wrap any other library exporting the API, which is how the
test suite drives a checker beside it.

ABI notes (libxm src/scalar.h:49-53, src/dim.h:33-38):
  * xm_scalar_t is `double complex`; under the SysV x86-64 ABI it is passed
    and returned in two SSE registers, exactly like a struct of two doubles,
    which is how it is declared here.
  * xm_dim_t is a 72-byte struct passed by value.
"""
import ctypes as C
import os

import numpy as np

SCALAR_FLOAT = 0
SCALAR_FLOAT_COMPLEX = 1
SCALAR_DOUBLE = 2
SCALAR_DOUBLE_COMPLEX = 3

BLOCK_ZERO = 0
BLOCK_CANONICAL = 1
BLOCK_DERIVATIVE = 2

PRECISION_NATIVE = 0
PRECISION_SPLIT_BF16 = 1
PRECISION_SPLIT_INT8 = 2

NULL_PTR = (1 << 64) - 1
MAX_DIM = 8

NP_DTYPE = {
    SCALAR_FLOAT: np.dtype(np.float32),
    SCALAR_FLOAT_COMPLEX: np.dtype(np.complex64),
    SCALAR_DOUBLE: np.dtype(np.float64),
    SCALAR_DOUBLE_COMPLEX: np.dtype(np.complex128),
}


class Dim(C.Structure):
    _fields_ = [("n", C.c_size_t), ("i", C.c_size_t * MAX_DIM)]


class Scalar(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class ContractPair(C.Structure):
    _fields_ = [("blkidxc", Dim), ("blkidxa", Dim), ("blkidxb", Dim),
                ("alpha", Scalar), ("m", C.c_size_t), ("n", C.c_size_t),
                ("k", C.c_size_t)]


class ContractStats(C.Structure):
    _fields_ = [("flops", C.c_double), ("flops_pre", C.c_double),
                ("npairs", C.c_size_t), ("npairs_pre", C.c_size_t),
                ("nblocks_c", C.c_size_t), ("nunfold", C.c_size_t),
                ("unfold_bytes", C.c_double), ("gemm_ms", C.c_double),
                ("prep_ms", C.c_double), ("post_ms", C.c_double),
                ("total_ms", C.c_double), ("plan_cached", C.c_int),
                ("precision", C.c_int), ("nslices", C.c_int),
                ("gemm_launches", C.c_int), ("panel_bytes", C.c_double),
                ("cin_bytes", C.c_double), ("cout_bytes", C.c_double),
                ("tile_p", C.c_int), ("tile_q", C.c_int),
                ("tile_tuning", C.c_int), ("graph", C.c_int),
                ("permute_tiled", C.c_size_t),
                ("permute_gather", C.c_size_t),
                ("staged_batches", C.c_int), ("staged_bytes", C.c_double),
                ("ksplit_tiles", C.c_size_t), ("err_bound", C.c_double),
                ("gemm_kernel", C.c_char * 32), ("file_bytes", C.c_double),
                ("classes", C.c_size_t), ("class_cblocks", C.c_size_t)]


def dim(seq):
    """Python sequence -> xm_dim_t."""
    seq = tuple(int(x) for x in seq)
    if len(seq) > MAX_DIM:
        raise ValueError("at most %d dimensions" % MAX_DIM)
    d = Dim()
    d.n = len(seq)
    for k, v in enumerate(seq):
        d.i[k] = v
    return d


def undim(d):
    return tuple(int(d.i[k]) for k in range(d.n))


def scalar(x):
    x = complex(x)
    return Scalar(x.real, x.imag)


def unscalar(s):
    return complex(s.re, s.im)


_P = C.c_void_p
_S = C.c_char_p
_SZ = C.c_size_t

# name: (restype, argtypes); functions of include/*.h
_SIGNATURES = {
    # dim.h
    "xm_dim_offset": (_SZ, [C.POINTER(Dim), C.POINTER(Dim)]),
    "xm_dim_from_offset": (Dim, [_SZ, C.POINTER(Dim)]),
    "xm_dim_permute": (Dim, [C.POINTER(Dim), C.POINTER(Dim)]),
    "xm_dim_inc": (None, [C.POINTER(Dim), C.POINTER(Dim)]),
    "xm_dim_inc_mask": (None, [C.POINTER(Dim), C.POINTER(Dim), C.POINTER(Dim)]),
    # util.h
    "xm_make_masks": (None, [_S, _S, C.POINTER(Dim), C.POINTER(Dim)]),
    # blockspace.h
    "xm_block_space_create": (_P, [Dim]),
    "xm_block_space_clone": (_P, [_P]),
    "xm_block_space_permute_clone": (_P, [_P, Dim]),
    "xm_block_space_get_ndims": (_SZ, [_P]),
    "xm_block_space_get_abs_dims": (Dim, [_P]),
    "xm_block_space_get_nblocks": (Dim, [_P]),
    "xm_block_space_split": (None, [_P, _SZ, _SZ]),
    "xm_block_space_autosplit": (None, [_P]),
    "xm_block_space_get_split": (_SZ, [_P, _SZ, _SZ]),
    "xm_block_space_get_block_dims": (Dim, [_P, Dim]),
    "xm_block_space_get_block_size": (_SZ, [_P, Dim]),
    "xm_block_space_get_largest_block_size": (_SZ, [_P]),
    "xm_block_space_decompose_index": (None, [_P, Dim, C.POINTER(Dim), C.POINTER(Dim)]),
    "xm_block_space_eq": (C.c_int, [_P, _P]),
    "xm_block_space_eq1": (C.c_int, [_P, _SZ, _P, _SZ]),
    "xm_block_space_free": (None, [_P]),
    # alloc.h
    "xm_allocator_create": (_P, [_S]),
    "xm_allocator_get_path": (_S, [_P]),
    "xm_allocator_allocate": (C.c_uint64, [_P, _SZ]),
    "xm_allocator_read": (None, [_P, C.c_uint64, _P, _SZ]),
    "xm_allocator_write": (None, [_P, C.c_uint64, _P, _SZ]),
    "xm_allocator_deallocate": (None, [_P, C.c_uint64]),
    "xm_allocator_destroy": (None, [_P]),
    # scalar.h
    "xm_scalar_sizeof": (_SZ, [C.c_int]),
    "xm_scalar_mul": (Scalar, [Scalar, Scalar, C.c_int]),
    "xm_scalar_add": (Scalar, [Scalar, Scalar, C.c_int]),
    # tensor.h
    "xm_tensor_create": (_P, [_P, C.c_int, _P]),
    "xm_tensor_create_canonical": (_P, [_P, C.c_int, _P]),
    "xm_tensor_create_structure": (_P, [_P, C.c_int, _P]),
    "xm_tensor_get_block_space": (_P, [_P]),
    "xm_tensor_get_scalar_type": (C.c_int, [_P]),
    "xm_tensor_get_allocator": (_P, [_P]),
    "xm_tensor_get_abs_dims": (Dim, [_P]),
    "xm_tensor_get_nblocks": (Dim, [_P]),
    "xm_tensor_get_element": (Scalar, [_P, Dim]),
    "xm_tensor_get_block_type": (C.c_int, [_P, Dim]),
    "xm_tensor_get_block_dims": (Dim, [_P, Dim]),
    "xm_tensor_get_block_size": (_SZ, [_P, Dim]),
    "xm_tensor_get_block_bytes": (_SZ, [_P, Dim]),
    "xm_tensor_get_largest_block_size": (_SZ, [_P]),
    "xm_tensor_get_largest_block_bytes": (_SZ, [_P]),
    "xm_tensor_get_block_data_ptr": (C.c_uint64, [_P, Dim]),
    "xm_tensor_get_block_permutation": (Dim, [_P, Dim]),
    "xm_tensor_get_block_scalar": (Scalar, [_P, Dim]),
    "xm_tensor_set_zero_block": (None, [_P, Dim]),
    "xm_tensor_set_canonical_block": (None, [_P, Dim]),
    "xm_tensor_set_canonical_block_raw": (None, [_P, Dim, C.c_uint64]),
    "xm_tensor_set_derivative_block": (None, [_P, Dim, Dim, Dim, Scalar]),
    "xm_tensor_get_canonical_block_list": (None, [_P, C.POINTER(C.POINTER(Dim)), C.POINTER(_SZ)]),
    "xm_tensor_read_block": (None, [_P, Dim, _P]),
    "xm_tensor_write_block": (None, [_P, Dim, _P]),
    "xm_tensor_unfold_block": (None, [_P, Dim, Dim, Dim, _P, _P, _SZ]),
    "xm_tensor_fold_block": (None, [_P, Dim, Dim, Dim, _P, _P, _SZ]),
    "xm_tensor_free_block_data": (None, [_P]),
    "xm_tensor_free": (None, [_P]),
    # xm.h
    "xm_print_banner": (None, []),
    "xm_set": (None, [_P, Scalar]),
    "xm_copy": (None, [_P, Scalar, _P, _S, _S]),
    "xm_add": (None, [Scalar, _P, Scalar, _P, _S, _S]),
    "xm_mul": (None, [_P, _P, _S, _S]),
    "xm_div": (None, [_P, _P, _S, _S]),
    "xm_dot": (Scalar, [_P, _P, _S, _S]),
    "xm_contract": (None, [Scalar, _P, _P, Scalar, _P, _S, _S, _S]),
    # xm_amd.h (optional in other libraries)
    "xm_contract_get_pairs": (_SZ, [_P, _P, _P, _S, _S, _S, C.POINTER(ContractPair), _SZ]),
    "xm_contract_get_stats": (None, [C.POINTER(ContractStats)]),
    "xm_set_contract_precision": (None, [C.c_int]),
    "xm_get_contract_precision": (C.c_int, []),
    "xm_set_split_slices": (None, [C.c_int]),
    "xm_get_split_slices": (C.c_int, []),
    "xm_set_split_moduli": (None, [C.c_int]),
    "xm_get_split_moduli": (C.c_int, []),
    "xm_set_graph_launch": (None, [C.c_int]),
    "xm_get_graph_launch": (C.c_int, []),
    "xm_contract_backward": (None, [Scalar, _P, _P, _P, _P, _P, _S, _S, _S]),
    "xm_contract_noalloc": (C.c_int, [Scalar, _P, _P, Scalar, _P, _S, _S, _S]),
    "xm_amd_fill_random": (None, [_P, C.c_uint64, C.c_double, C.c_double, C.c_double, C.c_double]),
    "xm_amd_allocator_usage": (None, [_P, C.POINTER(_SZ), C.POINTER(_SZ)]),
    "xm_amd_allocator_file_bytes": (_SZ, [_P]),
    "xm_amd_synchronize": (C.c_int, []),
    "xm_amd_last_permute": (None, [C.POINTER(_SZ), C.POINTER(_SZ)]),
    "xm_oracle_blas_name": (_S, []),
}

# Symbols every library implementing libxm's API must export.
CORE_SYMBOLS = [n for n in _SIGNATURES
                if not n.startswith(("xm_amd_", "xm_oracle_"))
                and n not in ("xm_contract_get_pairs", "xm_contract_get_stats",
                              "xm_set_contract_precision",
                              "xm_get_contract_precision",
                              "xm_set_split_slices", "xm_get_split_slices",
                              "xm_set_split_moduli", "xm_get_split_moduli",
                              "xm_set_graph_launch", "xm_get_graph_launch",
                              "xm_contract_backward",
                              "xm_contract_noalloc")]

_libc = C.CDLL(None)
_libc.free.argtypes = [C.c_void_p]


def _b(s):
    return s.encode() if isinstance(s, str) else s


class XmLib:
    """A loaded library exporting libxm's C API."""

    def __init__(self, path, mode=os.RTLD_LOCAL):
        self.path = os.path.abspath(path)
        self.lib = C.CDLL(self.path, mode=mode)
        self.missing = []
        for name, (res, args) in _SIGNATURES.items():
            try:
                fn = getattr(self.lib, name)
            except AttributeError:
                self.missing.append(name)
                continue
            fn.restype = res
            fn.argtypes = args

    def has(self, name):
        return name not in self.missing and hasattr(self.lib, name)

    # ---- block spaces -------------------------------------------------
    def bs_create(self, dims, splits=None):
        """Block space with absolute dims; splits: {dim: [points]}."""
        bs = self.lib.xm_block_space_create(dim(dims))
        if not bs:
            raise MemoryError("xm_block_space_create failed")
        for d, pts in (splits or {}).items():
            for x in pts:
                self.lib.xm_block_space_split(bs, d, x)
        return bs

    def bs_free(self, bs):
        self.lib.xm_block_space_free(bs)

    def bs_nblocks(self, bs):
        return undim(self.lib.xm_block_space_get_nblocks(bs))

    # ---- allocators ----------------------------------------------------
    def allocator(self, path=None):
        a = self.lib.xm_allocator_create(_b(path) if path else None)
        if not a:
            raise MemoryError("xm_allocator_create failed")
        return a

    def allocator_destroy(self, a):
        self.lib.xm_allocator_destroy(a)

    # ---- tensors ---------------------------------------------------------
    def tensor(self, bs, type_, alloc):
        return self.lib.xm_tensor_create(bs, type_, alloc)

    def tensor_canonical(self, bs, type_, alloc):
        return self.lib.xm_tensor_create_canonical(bs, type_, alloc)

    def tensor_structure(self, t, type_=None, alloc=None):
        if type_ is None:
            type_ = self.scalar_type(t)
        return self.lib.xm_tensor_create_structure(t, type_, alloc)

    def tensor_free(self, t, data=True):
        if data:
            self.lib.xm_tensor_free_block_data(t)
        self.lib.xm_tensor_free(t)

    def scalar_type(self, t):
        return self.lib.xm_tensor_get_scalar_type(t)

    def nblocks(self, t):
        return undim(self.lib.xm_tensor_get_nblocks(t))

    def abs_dims(self, t):
        return undim(self.lib.xm_tensor_get_abs_dims(t))

    def block_type(self, t, idx):
        return self.lib.xm_tensor_get_block_type(t, dim(idx))

    def block_dims(self, t, idx):
        return undim(self.lib.xm_tensor_get_block_dims(t, dim(idx)))

    def block_perm(self, t, idx):
        return undim(self.lib.xm_tensor_get_block_permutation(t, dim(idx)))

    def block_scalar(self, t, idx):
        return unscalar(self.lib.xm_tensor_get_block_scalar(t, dim(idx)))

    def block_data_ptr(self, t, idx):
        return int(self.lib.xm_tensor_get_block_data_ptr(t, dim(idx)))

    def set_canonical(self, t, idx):
        self.lib.xm_tensor_set_canonical_block(t, dim(idx))

    def set_derivative(self, t, idx, src, perm, s):
        self.lib.xm_tensor_set_derivative_block(t, dim(idx), dim(src),
                                                dim(perm), scalar(s))

    def set_zero(self, t, idx):
        self.lib.xm_tensor_set_zero_block(t, dim(idx))

    def canonical_list(self, t):
        ptr = C.POINTER(Dim)()
        n = C.c_size_t()
        self.lib.xm_tensor_get_canonical_block_list(t, C.byref(ptr), C.byref(n))
        out = [undim(ptr[k]) for k in range(n.value)]
        if ptr:
            _libc.free(C.cast(ptr, C.c_void_p))
        return out

    def block_array(self, t, idx):
        """Empty host array of the block's size and element type."""
        n = self.lib.xm_tensor_get_block_size(t, dim(idx))
        return np.empty(n, dtype=NP_DTYPE[self.scalar_type(t)])

    def read_block(self, t, idx):
        """Block data (for derivative blocks: the source's raw data) as a
        1-D array in storage order (first index fastest)."""
        out = self.block_array(t, idx)
        self.lib.xm_tensor_read_block(t, dim(idx), out.ctypes.data)
        return out

    def write_block(self, t, idx, data):
        want = NP_DTYPE[self.scalar_type(t)]
        arr = np.ascontiguousarray(np.asarray(data).ravel(order="F"), dtype=want)
        n = self.lib.xm_tensor_get_block_size(t, dim(idx))
        if arr.size != n:
            raise ValueError("block %r holds %d elements, got %d" % (idx, n, arr.size))
        self.lib.xm_tensor_write_block(t, dim(idx), arr.ctypes.data)

    def get_element(self, t, idx):
        return unscalar(self.lib.xm_tensor_get_element(t, dim(idx)))

    # ---- operations ----------------------------------------------------
    def contract(self, alpha, a, b, beta, c, ia, ib, ic):
        self.lib.xm_contract(scalar(alpha), a, b, scalar(beta), c,
                             _b(ia), _b(ib), _b(ic))

    def add(self, alpha, a, beta, b, ia, ib):
        self.lib.xm_add(scalar(alpha), a, scalar(beta), b, _b(ia), _b(ib))

    def set(self, a, x):
        self.lib.xm_set(a, scalar(x))

    def copy(self, a, s, b, ia, ib):
        self.lib.xm_copy(a, scalar(s), b, _b(ia), _b(ib))

    def mul(self, a, b, ia, ib):
        self.lib.xm_mul(a, b, _b(ia), _b(ib))

    def div(self, a, b, ia, ib):
        self.lib.xm_div(a, b, _b(ia), _b(ib))

    def dot(self, a, b, ia, ib):
        return unscalar(self.lib.xm_dot(a, b, _b(ia), _b(ib)))

    def contract_noalloc(self, alpha, a, b, beta, c, ia, ib, ic):
        return self.lib.xm_contract_noalloc(scalar(alpha), a, b, scalar(beta),
                                            c, _b(ia), _b(ib), _b(ic))

    def contract_backward(self, alpha, a, b, dc, da, db, ia, ib, ic):
        self.lib.xm_contract_backward(scalar(alpha), a, b, dc, da, db,
                                      _b(ia), _b(ib), _b(ic))

    def get_pairs(self, a, b, c, ia, ib, ic):
        n = self.lib.xm_contract_get_pairs(a, b, c, _b(ia), _b(ib), _b(ic),
                                           None, 0)
        buf = (ContractPair * max(n, 1))()
        self.lib.xm_contract_get_pairs(a, b, c, _b(ia), _b(ib), _b(ic),
                                       buf, n)
        return [dict(c=undim(p.blkidxc), a=undim(p.blkidxa),
                     b=undim(p.blkidxb), alpha=unscalar(p.alpha),
                     m=int(p.m), n=int(p.n), k=int(p.k)) for p in buf[:n]]

    def stats(self):
        s = ContractStats()
        self.lib.xm_contract_get_stats(C.byref(s))
        out = {name: getattr(s, name) for name, _ in ContractStats._fields_}
        out["gemm_kernel"] = out["gemm_kernel"].decode()
        return out

    def set_precision(self, precision, nslices=None, nmoduli=None):
        """nslices: BF16 slices (SPLIT_BF16); nmoduli: moduli (SPLIT_INT8,
        0 = chosen per plan)."""
        self.lib.xm_set_contract_precision(int(precision))
        if nslices is not None:
            self.lib.xm_set_split_slices(int(nslices))
        if nmoduli is not None:
            self.lib.xm_set_split_moduli(int(nmoduli))

    def set_graph_launch(self, on):
        self.lib.xm_set_graph_launch(1 if on else 0)

    def fill_random(self, t, seed, lo=-1.0, hi=1.0, wlo=0.0, whi=0.0):
        self.lib.xm_amd_fill_random(t, int(seed), lo, hi, wlo, whi)

    def synchronize(self):
        return self.lib.xm_amd_synchronize()

    def allocator_usage(self, alloc):
        d, h = C.c_size_t(), C.c_size_t()
        self.lib.xm_amd_allocator_usage(alloc, C.byref(d), C.byref(h))
        return d.value, h.value

    def allocator_file_bytes(self, alloc):
        return int(self.lib.xm_amd_allocator_file_bytes(alloc))

    def last_permute(self):
        """(tiled, gather) jobs of the last add/copy/mul/div."""
        t, g = C.c_size_t(), C.c_size_t()
        self.lib.xm_amd_last_permute(C.byref(t), C.byref(g))
        return t.value, g.value


HERE = os.path.dirname(os.path.abspath(__file__))
LIBXM_PATH = os.path.join(HERE, "libxm.so")

_loaded = None


def load(path=None):
    """Load the MI355X libxm (libxm_amd/libxm.so).  Raises if it has not
    been built: there is no fallback implementation."""
    global _loaded
    if path is None and _loaded is not None:
        return _loaded
    p = path or LIBXM_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            "libxm.so is not built (%s); run __graft_entry__.build() or "
            "make -C libxm_amd/csrc" % p)
    lib = XmLib(p)
    if path is None:
        _loaded = lib
    return lib
