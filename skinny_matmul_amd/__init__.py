"""skinny_matmul_amd: tall-and-skinny matrix products on AMD MI355X (gfx950).

Python binding of the C ABI in ``include/skinny.h`` (libskinny.so, hand-written
HIP kernels).  The callables mirror the reference operator boundary

    MatmulFunctionType = std::function<bool(blockCount, M, N, K, A, lda,
                                            alpha, B, ldb, beta, C, ldc)>
                                                    (versions.hpp:43-45)

and return ``True`` on success and ``False`` where the reference variant would
return ``false`` (unsupported arguments, the SKIP of test.cu:151-155).  Other
failures raise :class:`SkinnyError` (the reference aborts through GPU_ERROR,
gpu_error.cuh:4-13).  Matrices are torch tensors on the GPU (any shape; only
their storage and the explicit leading dimensions matter) or raw device
addresses given as ``int``.

There is no CPU fallback: if libskinny.so is missing every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Callable, List, Tuple

__all__ = [
    "SkinnyError", "lib", "library_path", "available", "Handle",
    "tsmttsm", "tsmm", "dtsmttsm", "stsmttsm", "dtsmm", "stsmm",
    "ztsmttsm", "ctsmttsm", "ztsmm", "ctsmm", "htsmttsm", "htsmttsm_dd",
    "fill_uniform", "fill_uniform_f", "fill_int",
    "get_enabled_tsmttsm_versions", "get_enabled_tsmm_versions",
    "types", "flops_per_cell",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "lib", "libskinny.so")
_lib = None

OK, UNSUPPORTED, INVALID_ARG, HIP_ERROR, OOM = 0, 1, 2, 3, 4

i64 = C.c_int64
u64 = C.c_uint64
sz = C.c_size_t
vp = C.c_void_p
dbl = C.c_double
flt = C.c_float


class SkinnyError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"skinny status {status}: {msg}")
        self.status = status


class _DComplex(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double)]


class _FComplex(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float)]


class BatchOp(C.Structure):
    """skinny_batch_op (op 0: TSMTTSM, 1: TSMM)."""
    _fields_ = [("op", C.c_int), ("M", C.c_int64), ("N", C.c_int64), ("K", C.c_int64),
                ("A", C.c_void_p), ("lda", C.c_int64), ("alpha", C.c_double),
                ("B", C.c_void_p), ("ldb", C.c_int64), ("beta", C.c_double),
                ("C", C.c_void_p), ("ldc", C.c_int64)]


class SvqbBlock(C.Structure):
    """skinny_svqb_block: X is K x M row-major with leading dimension ldx."""
    _fields_ = [("M", C.c_int64), ("K", C.c_int64), ("X", C.c_void_p), ("ldx", C.c_int64)]


def library_path() -> str:
    return _LIB_PATH


def available() -> bool:
    return os.path.exists(_LIB_PATH)


_SIGS = {
    "skinny_create": (C.c_int, [C.POINTER(vp)]),
    "skinny_destroy": (C.c_int, [vp]),
    "skinny_reserve": (C.c_int, [vp, sz]),
    "skinny_reserve_stream": (C.c_int, [vp, sz, vp]),
    "skinny_tsmttsm_workspace": (sz, [vp, sz, i64, i64, i64, C.c_int]),
    "skinny_default_handle": (vp, []),
    "skinny_status_string": (C.c_char_p, [C.c_int]),
    "skinny_last_error": (C.c_char_p, []),
    "skinny_cu_count": (C.c_int, [vp]),
    "skinny_tuned_blockcount_tsmttsm": (C.c_int, [vp, i64, i64, i64, C.c_int]),
    "skinny_tuned_blockcount_tsmm": (C.c_int, [vp, i64, i64, i64, C.c_int]),
    "skinny_tuned_blockcount_tsmm_ex": (C.c_int, [vp, i64, i64, i64, C.c_int, C.c_int, C.c_int]),
    "skinny_dtsmttsm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64, vp]),
    "skinny_dtsmttsm_noalloc": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64, vp]),
    "skinny_stsmttsm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, flt, vp, i64, flt, vp, i64, vp]),
    "skinny_dtsmm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64, vp]),
    "skinny_stsmm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, flt, vp, i64, flt, vp, i64, vp]),
    "skinny_ztsmttsm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, _DComplex, vp, i64, _DComplex, vp, i64, vp]),
    "skinny_ctsmttsm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, _FComplex, vp, i64, _FComplex, vp, i64, vp]),
    "skinny_ztsmm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, _DComplex, vp, i64, _DComplex, vp, i64, vp]),
    "skinny_ctsmm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, _FComplex, vp, i64, _FComplex, vp, i64, vp]),
    "skinny_dtsmm_inplace": (C.c_int, [vp, vp, vp, dbl, dbl, i64, i64, i64, i64, i64, vp]),
    "skinny_stsmm_inplace": (C.c_int, [vp, vp, vp, flt, flt, i64, i64, i64, i64, i64, vp]),
    "skinny_ztsmm_inplace": (C.c_int, [vp, vp, vp, _DComplex, _DComplex, i64, i64, i64, i64, i64, vp]),
    "skinny_ctsmm_inplace": (C.c_int, [vp, vp, vp, _FComplex, _FComplex, i64, i64, i64, i64, i64, vp]),
    "skinny_dtsmttsm_ozaki": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64, C.c_int, vp]),
    "skinny_probe_i8_layout": (C.c_int, [vp]),
    "skinny_htsmttsm": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64, vp]),
    "skinny_htsmttsm_dd": (C.c_int, [vp, sz, i64, i64, i64, vp, i64, dbl, vp, i64, dbl, vp, i64, vp]),
    "skinny_fill_uniform": (C.c_int, [vp, i64, u64, u64, i64, vp]),
    "skinny_fill_uniform_f": (C.c_int, [vp, i64, u64, u64, i64, vp]),
    "skinny_fill_int": (C.c_int, [vp, i64, u64, u64, i64, C.c_int, C.c_int, vp]),
    "skinny_autotune": (C.c_int, [vp, C.c_int, C.c_int, i64, i64, i64, C.c_int, C.POINTER(C.c_double)]),
    "skinny_tune_export": (sz, [vp, C.c_char_p, sz]),
    "skinny_tune_import": (C.c_int, [vp, C.c_char_p]),
    "skinny_tune_clear": (None, [vp]),
    "skinny_dbatch_create": (C.c_int, [vp, C.c_int, vp, i64, C.POINTER(vp)]),
    "skinny_dbatch_run": (C.c_int, [vp, vp]),
    "skinny_dbatch_tune": (C.c_int, [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int)]),
    "skinny_dbatch_set_class": (C.c_int, [vp, C.c_int, C.c_int]),
    "skinny_dbatch_classes": (C.c_int, [vp, vp]),
    "skinny_dbatch_destroy": (C.c_int, [vp]),
    "skinny_dbatch_items": (i64, [vp]),
    "skinny_dsvqb": (C.c_int, [vp, C.c_int, vp, C.c_int, vp]),
    "skinny_dsvqb_graph": (C.c_int, [vp, C.c_int, vp, C.c_int, vp, C.POINTER(vp)]),
    "skinny_dsvqb_fused": (C.c_int, [vp, C.c_int, vp, C.c_int, vp]),
    "skinny_graph_launch": (C.c_int, [vp, vp]),
    "skinny_graph_destroy": (C.c_int, [vp]),
    "skinny_dorth_error": (C.c_int, [vp, i64, i64, vp, i64, C.POINTER(C.c_double), vp]),
    "skinny_mb_stream": (C.c_int, [C.c_int, sz, C.c_int, C.POINTER(C.c_double), vp]),
    "skinny_mb_peak": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double), vp]),
    "skinny_mb_clock": (C.c_int, [C.c_int, C.POINTER(C.c_double), vp]),
    "skinny_mb_cache": (C.c_int, [C.c_int, sz, C.c_int, C.POINTER(C.c_double), vp]),
    "skinny_mb_barrier": (C.c_int, [C.c_int, C.c_int, C.POINTER(C.c_double), vp]),
}

# every symbol include/skinny.h declares (checked by the CPU test suite)
EXPORTED = tuple(_SIGS)


def lib():
    """Load libskinny.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise SkinnyError(HIP_ERROR, f"{_LIB_PATH} is not built: run "
                              "`python -m skinny_matmul_amd.build`")
        # torch (when installed) carries its own copy of the HIP runtime; load
        # it first so that libskinny binds to that same copy.  Loading the
        # system runtime first and torch afterwards puts two HIP runtimes in
        # one process, and device discovery then fails.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(_LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _last_error() -> str:
    return lib().skinny_last_error().decode(errors="replace")


def _addr(x) -> int:
    if x is None:
        return 0
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        if getattr(x, "device", None) is not None and x.device.type != "cuda":
            raise SkinnyError(INVALID_ARG, "matrices must live on the GPU")
        return int(x.data_ptr())
    raise TypeError(f"cannot take a device address of {type(x)!r}")


def _stream(stream) -> int:
    if stream is None:
        import sys
        torch = sys.modules.get("torch")
        if torch is None or not torch.cuda.is_initialized():
            return 0  # the null stream
        return int(torch.cuda.current_stream().cuda_stream)
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)


def _check(status: int, allow_unsupported: bool = True) -> bool:
    if status == OK:
        return True
    if status == UNSUPPORTED and allow_unsupported:
        return False
    raise SkinnyError(status, _last_error())


class Handle:
    """A library handle owning its device workspace (skinny_create)."""

    def __init__(self):
        h = vp()
        _check(lib().skinny_create(C.byref(h)), False)
        self._h = h.value

    @property
    def value(self) -> int:
        return self._h

    def reserve(self, nbytes: int, stream=None) -> None:
        """Size the workspace of `stream` (default: torch's current stream)
        now, e.g. before capturing calls on it into a graph."""
        _check(lib().skinny_reserve_stream(self._h, nbytes, _stream(stream)), False)

    def close(self) -> None:
        if self._h:
            lib().skinny_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _h(handle) -> int:
    if handle is None:
        return 0
    return handle.value if isinstance(handle, Handle) else int(handle)


# ---------------------------------------------------------------------------
# typed entry points (blockCount, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc)
# ---------------------------------------------------------------------------
def dtsmttsm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
             handle=None, stream=None) -> bool:
    return _check(lib().skinny_dtsmttsm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                        float(alpha), _addr(B), ldb, float(beta),
                                        _addr(Cm), ldc, _stream(stream)))


def dtsmttsm_noalloc(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
                     handle=None, stream=None) -> bool:
    """skinny_dtsmttsm with the partials in the library's fixed per-GPU
    scratch array (no device allocation; overlapping calls on one GPU race)."""
    return _check(lib().skinny_dtsmttsm_noalloc(_h(handle), blockCount, M, N, K, _addr(A),
                                                lda, float(alpha), _addr(B), ldb,
                                                float(beta), _addr(Cm), ldc, _stream(stream)))


NOALLOC_PARTIAL_BYTES = 16 << 20


def stsmttsm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
             handle=None, stream=None) -> bool:
    return _check(lib().skinny_stsmttsm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                        float(alpha), _addr(B), ldb, float(beta),
                                        _addr(Cm), ldc, _stream(stream)))


def dtsmm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
          handle=None, stream=None) -> bool:
    return _check(lib().skinny_dtsmm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                     float(alpha), _addr(B), ldb, float(beta),
                                     _addr(Cm), ldc, _stream(stream)))


def stsmm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
          handle=None, stream=None) -> bool:
    return _check(lib().skinny_stsmm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                     float(alpha), _addr(B), ldb, float(beta),
                                     _addr(Cm), ldc, _stream(stream)))


def dtsmttsm_ozaki(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
                   slices=3, handle=None, stream=None) -> bool:
    """FP64 TSMTTSM through INT8 limbs on the matrix cores (Ozaki split)."""
    return _check(lib().skinny_dtsmttsm_ozaki(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                              float(alpha), _addr(B), ldb, float(beta),
                                              _addr(Cm), ldc, slices, _stream(stream)))


def probe_i8_layout():
    """(rows, cols) of each (lane, element) of the INT8 32x32x32 MFMA result."""
    import numpy as np
    out = np.zeros(2048, dtype=np.int32)
    _check(lib().skinny_probe_i8_layout(out.ctypes.data_as(vp)), False)
    return out[:1024].reshape(64, 16) // 32, out[1024:].reshape(64, 16) // 32


def htsmttsm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
             handle=None, stream=None) -> bool:
    """HR: double in/out, double-double accumulation (types.hpp:49-57)."""
    return _check(lib().skinny_htsmttsm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                        float(alpha), _addr(B), ldb, float(beta),
                                        _addr(Cm), ldc, _stream(stream)))


def htsmttsm_dd(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cdd, ldc,
                handle=None, stream=None) -> bool:
    """HR with the dd result kept: Cdd holds (s, t) pairs (2 * ldc * N doubles)."""
    return _check(lib().skinny_htsmttsm_dd(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                           float(alpha), _addr(B), ldb, float(beta),
                                           _addr(Cdd), ldc, _stream(stream)))


def _z(v) -> _DComplex:
    v = complex(v)
    return _DComplex(v.real, v.imag)


def _c(v) -> _FComplex:
    v = complex(v)
    return _FComplex(v.real, v.imag)


def ztsmttsm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
             handle=None, stream=None) -> bool:
    return _check(lib().skinny_ztsmttsm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                        _z(alpha), _addr(B), ldb, _z(beta),
                                        _addr(Cm), ldc, _stream(stream)))


def ctsmttsm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
             handle=None, stream=None) -> bool:
    return _check(lib().skinny_ctsmttsm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                        _c(alpha), _addr(B), ldb, _c(beta),
                                        _addr(Cm), ldc, _stream(stream)))


def ztsmm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
          handle=None, stream=None) -> bool:
    return _check(lib().skinny_ztsmm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                     _z(alpha), _addr(B), ldb, _z(beta),
                                     _addr(Cm), ldc, _stream(stream)))


def ctsmm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
          handle=None, stream=None) -> bool:
    return _check(lib().skinny_ctsmm(_h(handle), blockCount, M, N, K, _addr(A), lda,
                                     _c(alpha), _addr(B), ldb, _c(beta),
                                     _addr(Cm), ldc, _stream(stream)))


def tsmm_inplace(x, w, alpha, beta, nrows, ncolsin, ncolsout, stridex, stridew,
                 handle=None, stream=None) -> bool:
    """x <- alpha x w + beta x in place, in the GHOST kernels' argument order
    (tsmm/fix_ip_ghost.cuh:4-7): x is nrows x max(ncolsin, ncolsout) row-major
    with row stride stridex, w is ncolsin x ncolsout with column stride
    stridew; the routine is picked from x's dtype."""
    key = _dtype_key(x)
    L = lib()
    args = (_h(handle), _addr(x), _addr(w))
    rest = (nrows, ncolsin, ncolsout, stridex, stridew, _stream(stream))
    if key == "DR":
        return _check(L.skinny_dtsmm_inplace(*args, float(alpha), float(beta), *rest))
    if key == "FR":
        return _check(L.skinny_stsmm_inplace(*args, float(alpha), float(beta), *rest))
    if key == "DC":
        return _check(L.skinny_ztsmm_inplace(*args, _z(alpha), _z(beta), *rest))
    return _check(L.skinny_ctsmm_inplace(*args, _c(alpha), _c(beta), *rest))


def _dtype_key(t) -> str:
    import torch
    return {torch.float64: "DR", torch.float32: "FR",
            torch.complex128: "DC", torch.complex64: "FC"}[t.dtype]


def tsmttsm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
            handle=None, stream=None) -> bool:
    """C = alpha A^T B + beta C, the routine picked from C's dtype."""
    fn = {"DR": dtsmttsm, "FR": stsmttsm, "DC": ztsmttsm, "FC": ctsmttsm}[_dtype_key(Cm)]
    return fn(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc, handle, stream)


def tsmm(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc,
         handle=None, stream=None) -> bool:
    """C = alpha A B + beta C, the routine picked from C's dtype."""
    fn = {"DR": dtsmm, "FR": stsmm, "DC": ztsmm, "FC": ctsmm}[_dtype_key(Cm)]
    return fn(blockCount, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc, handle, stream)


# ---------------------------------------------------------------------------
# registries (versions.hpp:47-177): the variants a test or perf sweep runs
# ---------------------------------------------------------------------------
types = ("FR", "DR", "FC", "DC", "HR")
flops_per_cell = {"FC": 8, "DC": 8, "FR": 2, "DR": 2, "HR": 9}  # types.hpp:17-57

MatmulFunction = Callable[..., bool]


def get_enabled_tsmttsm_versions(type_key: str = "DR") -> List[Tuple[MatmulFunction, str]]:
    table = {"DR": dtsmttsm, "FR": stsmttsm, "DC": ztsmttsm, "FC": ctsmttsm,
             "HR": htsmttsm}
    return [(table[type_key], "SKINNY")]


def get_enabled_tsmm_versions(type_key: str = "DR") -> List[Tuple[MatmulFunction, str]]:
    table = {"DR": dtsmm, "FR": stsmm, "DC": ztsmm, "FC": ctsmm, "HR": dtsmm}
    return [(table[type_key], "SKINNY")]


# ---------------------------------------------------------------------------
# device data generation
# ---------------------------------------------------------------------------
def _fill_target(out, n, itemsize: int) -> int:
    """Device address of a fill target.  A tensor must be contiguous and hold
    at least n elements of the routine's element size (8 bytes for the double
    fills, 4 for the float one) in its own storage: the C routines write n
    consecutive elements from the address and cannot see the buffer's extent
    (an expanded view has more elements than bytes behind it, a strided view
    would have its neighbours overwritten)."""
    if hasattr(out, "element_size") and hasattr(out, "numel"):
        if out.element_size() != itemsize:
            raise SkinnyError(INVALID_ARG, f"fill: the tensor's elements are {out.element_size()} "
                                           f"bytes, the routine writes {itemsize}-byte elements")
        if n > out.numel():
            raise SkinnyError(INVALID_ARG, f"fill: {n} elements into a tensor of {out.numel()}")
        if n > 1 and not out.is_contiguous():
            raise SkinnyError(INVALID_ARG, "fill: the tensor is not contiguous (the routine "
                                           "writes consecutive elements)")
        room = out.untyped_storage().nbytes() - out.storage_offset() * itemsize
        if n * itemsize > room:
            raise SkinnyError(INVALID_ARG, f"fill: {n * itemsize} bytes into {room} bytes "
                                           "of storage")
    return _addr(out)


def fill_uniform(out, n, seed, stream_id, offset=0, stream=None) -> None:
    """n doubles U[-1,1) from the counter RNG (float64 tensors or addresses)."""
    _check(lib().skinny_fill_uniform(_fill_target(out, n, 8), n, seed, stream_id, offset,
                                     _stream(stream)), False)


def fill_uniform_f(out, n, seed, stream_id, offset=0, stream=None) -> None:
    """n floats U[-1,1) from the counter RNG (float32 tensors or addresses)."""
    _check(lib().skinny_fill_uniform_f(_fill_target(out, n, 4), n, seed, stream_id, offset,
                                       _stream(stream)), False)


def fill_int(out, n, seed, stream_id, lo=-2, hi=2, offset=0, stream=None) -> None:
    """n integer-valued doubles in [lo, hi] (float64 tensors or addresses)."""
    _check(lib().skinny_fill_int(_fill_target(out, n, 8), n, seed, stream_id, offset, lo, hi,
                                 _stream(stream)), False)


TUNE_TYPES = {"DR": 0, "FR": 1, "DC": 2, "FC": 3, "HR": 4}


def autotune(op, M, N, K, reps=3, handle=None, dtype="DR") -> float:
    """Tune one shape class of one type ("DR", "FR", "DC", "FC", "HR"; complex
    widths for complex types); returns the best time in ms."""
    kind = {"tsmttsm": 0, "tsmm": 1}.get(op, op)
    out = C.c_double()
    _check(lib().skinny_autotune(_h(handle), kind, TUNE_TYPES.get(dtype, dtype), M, N, K,
                                 reps, C.byref(out)), False)
    return out.value


def tune_export(handle=None) -> str:
    n = lib().skinny_tune_export(_h(handle), None, 0)
    buf = C.create_string_buffer(n)
    lib().skinny_tune_export(_h(handle), buf, n)
    return buf.value.decode()


def tune_import(text: str, handle=None) -> None:
    _check(lib().skinny_tune_import(_h(handle), text.encode()), False)


def tune_clear(handle=None) -> None:
    lib().skinny_tune_clear(_h(handle))


class Batch:
    """A persistent batched launch over many independent FP64 products.

    ops: iterable of (op, M, N, K, A, lda, alpha, B, ldb, beta, C, ldc) with
    op "tsmttsm" or "tsmm" (or 0 / 1)."""

    def __init__(self, ops, item_bytes=0, handle=None):
        ops = list(ops)
        arr = (BatchOp * len(ops))()
        for i, (op, M, N, K, A, lda, alpha, B, ldb, beta, Cm, ldc) in enumerate(ops):
            kind = {"tsmttsm": 0, "tsmm": 1}.get(op, op)
            arr[i] = BatchOp(kind, M, N, K, _addr(A), lda, alpha, _addr(B), ldb, beta,
                             _addr(Cm), ldc)
        b = vp()
        _check(lib().skinny_dbatch_create(_h(handle), len(ops), C.cast(arr, vp), item_bytes,
                                          C.byref(b)), False)
        self._b = b.value
        self._n = len(ops)

    @property
    def items(self) -> int:
        return int(lib().skinny_dbatch_items(self._b))

    def run(self, stream=None) -> None:
        _check(lib().skinny_dbatch_run(self._b, _stream(stream)), False)

    def tune(self, reps=3):
        """Tune the tile class of each op shape, then time plans of the batch
        (item sizes, grids) and keep the fastest; returns (ms, item_bytes
        (0: automatic), grid)."""
        ms, ib, g = C.c_double(), C.c_int64(), C.c_int()
        _check(lib().skinny_dbatch_tune(self._b, reps, C.byref(ms), C.byref(ib), C.byref(g)),
               False)
        return ms.value, ib.value, g.value

    def set_class(self, op: int, cls: int) -> bool:
        """Run op `op` with tile class `cls` (-1: the default rule); False if
        that class cannot compute the op (the plan is then unchanged)."""
        return _check(lib().skinny_dbatch_set_class(self._b, op, cls))

    def classes(self):
        """The tile class of every op in the current plan."""
        arr = (C.c_int * self._n)()
        _check(lib().skinny_dbatch_classes(self._b, C.cast(arr, vp)), False)
        return list(arr)

    def close(self) -> None:
        if self._b:
            lib().skinny_dbatch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _blocks(blocks):
    """[(M, K, X, ldx), ...] -> ctypes array of skinny_svqb_block"""
    arr = (SvqbBlock * len(blocks))()
    for i, (M, K, X, ldx) in enumerate(blocks):
        arr[i] = SvqbBlock(M, K, _addr(X), ldx)
    return arr


def dsvqb(blocks, sweeps=2, handle=None, stream=None) -> None:
    """SVQB sweeps over blocks [(M, K, X, ldx), ...], in place (svqb.cu:72-130)."""
    arr = _blocks(blocks)
    _check(lib().skinny_dsvqb(_h(handle), len(blocks), C.cast(arr, vp), sweeps,
                              _stream(stream)), False)


def dsvqb_fused(blocks, sweeps=2, handle=None, stream=None) -> None:
    """The same SVQB sweeps as one cooperative launch with grid barriers
    between the Gram, eigensolve and update phases."""
    arr = _blocks(blocks)
    _check(lib().skinny_dsvqb_fused(_h(handle), len(blocks), C.cast(arr, vp), sweeps,
                                    _stream(stream)), False)


class Graph:
    """An executable hipGraph of SVQB sweeps (skinny_dsvqb_graph)."""

    def __init__(self, blocks, sweeps=2, stream=None):
        arr = _blocks(blocks)
        g = vp()
        _check(lib().skinny_dsvqb_graph(0, len(blocks), C.cast(arr, vp), sweeps,
                                        _stream(stream), C.byref(g)), False)
        self._g = g.value

    def launch(self, stream=None) -> None:
        _check(lib().skinny_graph_launch(self._g, _stream(stream)), False)

    def close(self) -> None:
        if self._g:
            lib().skinny_graph_destroy(self._g)
            self._g = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def orth_error(M, K, X, ldx, handle=None, stream=None) -> float:
    """||X^T X - I||_F with a double-double Gram (svqb.cu:184-211)."""
    out = C.c_double()
    _check(lib().skinny_dorth_error(_h(handle), M, K, _addr(X), ldx, C.byref(out),
                                    _stream(stream)), False)
    return out.value


STREAM_KINDS = {"read": 0, "copy": 1, "triad": 2, "write": 3, "update": 4, "copy_variants": 5}
PEAK_KINDS = {"fp64_fma": 0, "fp64_mfma": 1, "int8_mfma": 2, "fp64_mfma_4x4x4": 3,
              "fp32_mfma": 4, "fp32_mfma_32x32x2": 5, "fp32_mfma_4x4x1": 6}


def mb_stream(kind: str, bytes_per_array: int = 4 << 30, reps: int = 5, stream=None) -> float:
    """Measured HBM GB/s for one access mix (microbenches/membw.cu)."""
    out = C.c_double()
    _check(lib().skinny_mb_stream(STREAM_KINDS[kind], bytes_per_array, reps, C.byref(out),
                                  _stream(stream)), False)
    return out.value


def mb_peak(kind: str, reps: int = 3, stream=None) -> float:
    """Measured compute peak in TFLOP/s (TOP/s for int8)."""
    out = C.c_double()
    _check(lib().skinny_mb_peak(PEAK_KINDS[kind], reps, C.byref(out), _stream(stream)), False)
    return out.value


def mb_clock(kind: str = "fp64_mfma", stream=None) -> float:
    """Shader clock in MHz while the peak loop of `kind` runs ("idle": none)."""
    out = C.c_double()
    k = -1 if kind == "idle" else PEAK_KINDS[kind]
    _check(lib().skinny_mb_clock(k, C.byref(out), _stream(stream)), False)
    return out.value


CACHE_KINDS = {"latency": 0, "bandwidth": 1, "lds_bandwidth": 2, "lds_latency": 3}


def mb_cache(kind: str, nbytes: int = 1 << 20, reps: int = 3, stream=None) -> float:
    """Memory hierarchy microbenchmark (microbenches/instruction_rates.cu):
    ns per dependent load over an nbytes chain ("latency"), GB/s re-reading
    an nbytes buffer ("bandwidth"), LDS GB/s ("lds_bandwidth") or ns per
    dependent LDS read ("lds_latency")."""
    out = C.c_double()
    _check(lib().skinny_mb_cache(CACHE_KINDS[kind], nbytes, reps, C.byref(out),
                                 _stream(stream)), False)
    return out.value


def mb_barrier(per_cu: int = 1, reps: int = 3, stream=None):
    """Microseconds per grid-wide barrier with per_cu workgroups per CU
    (global_sync/sync.cu); None if that many do not fit on a CU."""
    out = C.c_double()
    ok = _check(lib().skinny_mb_barrier(per_cu, reps, C.byref(out), _stream(stream)))
    return out.value if ok else None


def cu_count(handle=None) -> int:
    return int(lib().skinny_cu_count(_h(handle)))


def tuned_blockcount_tsmttsm(M, N, K, elem_bytes=8, handle=None) -> int:
    return int(lib().skinny_tuned_blockcount_tsmttsm(_h(handle), M, N, K, elem_bytes))


def tuned_blockcount_tsmm(M, N, K, elem_bytes=8, handle=None, beta0=True,
                          inplace=False) -> int:
    """The grid a real TSMM call with blockCount 0 launches (beta == 0 or not,
    C == A or not)."""
    return int(lib().skinny_tuned_blockcount_tsmm_ex(_h(handle), M, N, K, elem_bytes,
                                                     int(bool(beta0)), int(bool(inplace))))
