"""ctypes binding of libdgamd.so (the C ABI declared in include/dgamd.h).

The functions mirror the C entry points one to one; the `*_t` helpers take
torch tensors (device memory and streams are torch's) and raise DGError on
a non-zero status.  Loading fails loudly when the library is not built:
there is no fallback path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdgamd.so")

DG_OK = 0
DG_ERR_INVALID = 1
DG_ERR_UNSUPPORTED = 2
DG_ERR_HIP = 3
DG_ERR_NO_DEVICE = 4
DG_ERR_ALLOC = 5
DG_ERR_SCHEDULE = 6
DG_ERR_WORKSPACE = 7

DG_SCHED_AUTO = 0
DG_SCHED_REFERENCE = 1
DG_SCHED_STANDARD = 2

DG_OZAKI_INT8 = 0
DG_OZAKI_BF16 = 1
DG_OZAKI_INT8_CRT = 2

DG_DIST_UNIFORM_PM1 = 0
DG_DIST_WIDE = 1
DG_DIST_UNIFORM_01 = 2

_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_ptr = ctypes.c_void_p
_chr = ctypes.c_char
_u64 = ctypes.c_uint64
_i64p = ctypes.POINTER(ctypes.c_int64)

DG_PLAN_AUTOTUNE = 0
DG_PLAN_NONPERSISTENT = 1
DG_PLAN_PERSISTENT = 2
DG_FORM_FP64 = 0
DG_FORM_OZAKI_INT8 = 1


class GemmDesc(ctypes.Structure):
    """struct dg_gemm_desc of include/dgamd.h."""
    _fields_ = [("opA", ctypes.c_char), ("opB", ctypes.c_char), ("reserved", ctypes.c_char * 6),
                ("m", ctypes.c_int64), ("n", ctypes.c_int64), ("k", ctypes.c_int64),
                ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64)]


_descp = ctypes.POINTER(GemmDesc)

# name -> (restype, argtypes); every symbol declared in include/dgamd.h.
SIGNATURES = {
    "dg_last_error": (ctypes.c_char_p, []),
    "dg_version": (_int, []),
    "dg_device_ok": (_int, []),
    "dg_dgemm": (_int, [_chr, _chr, _i64, _i64, _i64, _dbl, _ptr, _i64, _ptr, _i64, _dbl, _ptr, _i64, _ptr]),
    "dg_sgemm": (_int, [_chr, _chr, _i64, _i64, _i64, ctypes.c_float, _ptr, _i64, _ptr, _i64, ctypes.c_float, _ptr,
                        _i64, _ptr]),
    "dg_block_splits": (_int, [_i64, _int, _i64p]),
    "dg_virtual_grid": (_int, [_int, ctypes.POINTER(_int), ctypes.POINTER(_int)]),
    "dg_pack_block2d": (_int, [_ptr, _i64, _i64, _i64, _i64p, _int, _i64p, _int, _ptr, _ptr]),
    "dg_unpack_block2d": (_int, [_ptr, _i64p, _int, _i64p, _int, _ptr, _i64, _i64, _i64, _ptr]),
    "dg_copy_block": (_int, [_i64, _i64, _ptr, _i64, _ptr, _i64, _ptr]),
    "dg_copy_block_f32": (_int, [_i64, _i64, _ptr, _i64, _ptr, _i64, _ptr]),
    "dg_copy2d_h2d": (_int, [_ptr, _i64, _ptr, _i64, _i64, _i64, _ptr]),
    "dg_copy2d_d2h": (_int, [_ptr, _i64, _ptr, _i64, _i64, _i64, _ptr]),
    "dg_numroc": (_i64, [_i64, _i64, _int, _int, _int]),
    "dg_block_cyclic_size": (_i64, [_i64, _i64, _i64, _i64, _int, _int]),
    "dg_pack_block_cyclic": (_int, [_ptr, _i64, _i64, _i64, _i64, _i64, _int, _int, _ptr, _ptr]),
    "dg_unpack_block_cyclic": (_int, [_ptr, _i64, _i64, _int, _int, _ptr, _i64, _i64, _i64, _ptr]),
    "dg_cannon25d_workspace": (_i64, [_i64, _i64, _int, _int]),
    "dg_cannon25d": (_int, [_chr, _chr, _i64, _i64, _i64, _dbl, _ptr, _ptr, _dbl, _ptr, _int, _int, _int,
                            _ptr, _i64, _ptr]),
    "dg_cannon_m_index": (_int, [_int, _int, _int, _int, _int, _int, _int]),
    "dg_summa": (_int, [_chr, _chr, _i64, _i64, _i64, _dbl, _ptr, _ptr, _dbl, _ptr, _i64, _i64, _i64, _int,
                        _int, _ptr]),
    "dg_layer_reduce": (_int, [_ptr, _ptr, _i64, _int, _ptr]),
    "dg_generate": (_int, [_u64, _int, _ptr, _i64, _ptr]),
    "dg_generate_f32": (_int, [_u64, _int, _ptr, _i64, _ptr]),
    "dg_grouped_dgemm": (_int, [_descp, _int, _ptr]),
    "dg_grouped_dgemm_noalloc": (_int, [_descp, _int, _ptr]),
    "dg_grouped_plan_create": (_int, [_descp, _int, _int, ctypes.POINTER(ctypes.c_void_p)]),
    "dg_grouped_plan_create_form": (_int, [_descp, _int, _int, _int, _int, ctypes.POINTER(ctypes.c_void_p)]),
    "dg_grouped_plan_problem_form": (_int, [ctypes.c_void_p, _int, ctypes.POINTER(_int)]),
    "dg_grouped_plan_execute": (_int, [_ptr, _ptr]),
    "dg_grouped_plan_graph_launch": (_int, [_ptr, _ptr]),
    "dg_grouped_plan_variant": (_int, [_ptr]),
    "dg_grouped_plan_destroy": (_int, [_ptr]),
    "dg_ozaki_workspace": (_i64, [_i64, _i64, _i64, _int, _int]),
    "dg_ozaki_dgemm": (_int, [_chr, _chr, _i64, _i64, _i64, _dbl, _ptr, _i64, _ptr, _i64, _dbl, _ptr, _i64, _int,
                              _int, _ptr, _i64, ctypes.POINTER(ctypes.c_double), _ptr]),
}


class DGError(RuntimeError):
    def __init__(self, func, status, msg):
        super().__init__("%s failed with status %d: %s" % (func, status, msg))
        self.status = status


_lib = None


def lib():
    """Loads libdgamd.so once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libdgamd.so is not built (%s); run build.py or __graft_entry__.build()" % LIB_PATH)
        try:
            # Load torch's HIP runtime first so that the library binds to the
            # same runtime instance that owns torch's allocations and streams.
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().dg_last_error().decode()


def check(func, status):
    if status != DG_OK:
        raise DGError(func, status, last_error())
    return status


def _op(c):
    return c.encode() if isinstance(c, str) else c


def _offsets(seq):
    arr = (ctypes.c_int64 * len(seq))(*[int(v) for v in seq])
    return arr


# ------------------------------------------------------------------ host-side helpers (no GPU)

def block_splits(n, parts):
    out = (ctypes.c_int64 * (parts + 1))()
    check("dg_block_splits", lib().dg_block_splits(n, parts, out))
    return list(out)


def virtual_grid(size):
    r, c = _int(), _int()
    check("dg_virtual_grid", lib().dg_virtual_grid(size, ctypes.byref(r), ctypes.byref(c)))
    return r.value, c.value


def numroc(n, nb, iproc, isrc, nprocs):
    return lib().dg_numroc(n, nb, iproc, isrc, nprocs)


def block_cyclic_size(rows, cols, mb, nb, P, Q):
    return lib().dg_block_cyclic_size(rows, cols, mb, nb, P, Q)


def cannon_m_index(p, c, schedule, i, j, k, r):
    return lib().dg_cannon_m_index(p, c, schedule, i, j, k, r)


def cannon25d_workspace(M, N, p, c):
    return lib().dg_cannon25d_workspace(M, N, p, c)


# ------------------------------------------------------------------ torch-tensor wrappers (GPU)

def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _dp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def dgemm_t(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    check("dg_dgemm", lib().dg_dgemm(_op(opA), _op(opB), m, n, k, alpha, _dp(A), lda, _dp(B), ldb, beta,
                                      _dp(C), ldc, _stream()))


def sgemm_t(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    check("dg_sgemm", lib().dg_sgemm(_op(opA), _op(opB), m, n, k, alpha, _dp(A), lda, _dp(B), ldb, beta,
                                      _dp(C), ldc, _stream()))


def generate_t(seed, dist, t):
    import torch
    if t.dtype == torch.float32:
        check("dg_generate_f32", lib().dg_generate_f32(seed, dist, _dp(t), t.numel(), _stream()))
    else:
        check("dg_generate", lib().dg_generate(seed, dist, _dp(t), t.numel(), _stream()))


def pack_block2d_t(G, ldg, rows, cols, row_off, col_off, packed):
    ro, co = _offsets(row_off), _offsets(col_off)
    check("dg_pack_block2d", lib().dg_pack_block2d(_dp(G), ldg, rows, cols, ro, len(row_off) - 1, co,
                                                    len(col_off) - 1, _dp(packed), _stream()))


def unpack_block2d_t(packed, row_off, col_off, G, ldg, rows, cols):
    ro, co = _offsets(row_off), _offsets(col_off)
    check("dg_unpack_block2d", lib().dg_unpack_block2d(_dp(packed), ro, len(row_off) - 1, co, len(col_off) - 1,
                                                        _dp(G), ldg, rows, cols, _stream()))


def copy_block_t(rows, cols, src, lds, dst, ldd):
    import torch
    if src.dtype == torch.float32:
        check("dg_copy_block_f32", lib().dg_copy_block_f32(rows, cols, _dp(src), lds, _dp(dst), ldd, _stream()))
    else:
        check("dg_copy_block", lib().dg_copy_block(rows, cols, _dp(src), lds, _dp(dst), ldd, _stream()))


def _host_ptr(a):
    """Address of a host buffer: a numpy array or a CPU torch tensor (pinned or not)."""
    import numpy as np
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


def copy2d_h2d(dev, dev_pitch, host, host_pitch, width, height):
    """Host-to-device copy of `height` columns of `width` bytes (pitches in bytes)."""
    check("dg_copy2d_h2d", lib().dg_copy2d_h2d(_dp(dev), dev_pitch, _host_ptr(host), host_pitch, width, height,
                                                _stream()))


def copy2d_d2h(host, host_pitch, dev, dev_pitch, width, height):
    """Device-to-host copy of `height` columns of `width` bytes (pitches in bytes)."""
    check("dg_copy2d_d2h", lib().dg_copy2d_d2h(_host_ptr(host), host_pitch, _dp(dev), dev_pitch, width, height,
                                                _stream()))


def pack_block_cyclic_t(G, ldg, rows, cols, mb, nb, P, Q, packed):
    check("dg_pack_block_cyclic", lib().dg_pack_block_cyclic(_dp(G), ldg, rows, cols, mb, nb, P, Q, _dp(packed),
                                                              _stream()))


def unpack_block_cyclic_t(packed, mb, nb, P, Q, G, ldg, rows, cols):
    check("dg_unpack_block_cyclic", lib().dg_unpack_block_cyclic(_dp(packed), mb, nb, P, Q, _dp(G), ldg, rows,
                                                                  cols, _stream()))


def cannon25d_t(opA, opB, M, N, K, alpha, Ap, Bp, beta, Cp, p, c, schedule=DG_SCHED_AUTO, workspace=None):
    ws_elems = workspace.numel() if workspace is not None else 0
    check("dg_cannon25d", lib().dg_cannon25d(_op(opA), _op(opB), M, N, K, alpha, _dp(Ap), _dp(Bp), beta, _dp(Cp),
                                              p, c, schedule, _dp(workspace), ws_elems, _stream()))


def summa_t(opA, opB, M, N, K, alpha, Ap, Bp, beta, Cp, mb, nb, kb, P, Q):
    check("dg_summa", lib().dg_summa(_op(opA), _op(opB), M, N, K, alpha, _dp(Ap), _dp(Bp), beta, _dp(Cp), mb, nb,
                                      kb, P, Q, _stream()))


def layer_reduce_t(C, partial, n, nlayers):
    check("dg_layer_reduce", lib().dg_layer_reduce(_dp(C), _dp(partial), n, nlayers, _stream()))


def ozaki_workspace(m, n, k, slices, kind):
    return lib().dg_ozaki_workspace(m, n, k, slices, kind)


def ozaki_dgemm_t(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, slices, kind=DG_OZAKI_INT8,
                  workspace=None):
    """Runs dg_ozaki_dgemm; allocates the workspace with torch when none is
    given.  Returns the a-priori error bound factor."""
    import torch
    need = ozaki_workspace(m, n, k, slices, kind)
    if need < 0:
        raise DGError("dg_ozaki_workspace", DG_ERR_INVALID, "bad arguments")
    if workspace is None:
        workspace = torch.empty(max(need, 1), dtype=torch.uint8, device=C.device)
    bound = ctypes.c_double(0.0)
    check("dg_ozaki_dgemm", lib().dg_ozaki_dgemm(_op(opA), _op(opB), m, n, k, alpha, _dp(A), lda, _dp(B), ldb, beta,
                                                  _dp(C), ldc, slices, kind, _dp(workspace), workspace.numel(),
                                                  ctypes.byref(bound), _stream()))
    return bound.value


def gemm_desc(opA, opB, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    """A GemmDesc over torch tensors (or raw device addresses as ints)."""
    addr = lambda t: t if isinstance(t, int) or t is None else t.data_ptr()
    return GemmDesc(_op(opA), _op(opB), b"\0" * 6, m, n, k, alpha, beta, addr(A), lda, addr(B), ldb, addr(C), ldc)


def _desc_array(descs):
    arr = (GemmDesc * len(descs))(*descs)
    return arr


def grouped_dgemm_t(descs):
    arr = _desc_array(descs)
    check("dg_grouped_dgemm", lib().dg_grouped_dgemm(arr, len(descs), _stream()))


def grouped_dgemm_noalloc_t(descs):
    arr = _desc_array(descs)
    check("dg_grouped_dgemm_noalloc", lib().dg_grouped_dgemm_noalloc(arr, len(descs), _stream()))


class GroupedPlan:
    """dg_grouped_plan over a list of GemmDesc; destroyed with close()."""

    def __init__(self, descs, mode=DG_PLAN_AUTOTUNE, form=DG_FORM_FP64, moduli=0):
        arr = _desc_array(descs)
        h = ctypes.c_void_p()
        if form == DG_FORM_FP64 and moduli == 0:
            check("dg_grouped_plan_create", lib().dg_grouped_plan_create(arr, len(descs), mode, ctypes.byref(h)))
        else:
            check("dg_grouped_plan_create_form",
                  lib().dg_grouped_plan_create_form(arr, len(descs), mode, form, moduli, ctypes.byref(h)))
        self.handle = h

    def problem_form(self, index):
        """(DG_FORM_*, moduli) chosen for problem `index`; (-1, 0) before the first execution."""
        mod = ctypes.c_int(0)
        f = lib().dg_grouped_plan_problem_form(self.handle, index, ctypes.byref(mod))
        return f, mod.value

    def execute(self):
        check("dg_grouped_plan_execute", lib().dg_grouped_plan_execute(self.handle, _stream()))

    def graph_launch(self):
        check("dg_grouped_plan_graph_launch", lib().dg_grouped_plan_graph_launch(self.handle, _stream()))

    def variant(self):
        return lib().dg_grouped_plan_variant(self.handle)

    def close(self):
        if self.handle:
            lib().dg_grouped_plan_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
