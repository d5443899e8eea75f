"""ispm_amd -- Python (ctypes) binding of the MI355X SlELL SpMV / pipelined-PCG library.

Thin mirror of the C-ABI in include/ispm.h; the library itself is libispm0.so, built from
ispm-sparse-lib_amd/csrc by ``make`` in ispm-sparse-lib_amd/.  Every call goes to the HIP
library: there is no CPU fallback, and importing this module fails loudly if the shared
library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "libispm0.so")

F32, F64, BF16 = 0, 1, 2
H2D, D2H, D2D, DEFAULT = 0, 1, 2, 3
PLAN_DEFAULT, PLAN_ESTIMATE, PLAN_EXHAUSTIVE, PLAN_MIN_BYTES = 0, 1, 2, 3
SPMV_NO_TEXTURE_MEMORY, SPMV_PLAN_NO_DEVICE, SPMV_PLAN_SPLIT_DIAG, SPMV_PLAN_SYMMETRIC = 1, 2, 4, 8

_NP = {F32: np.float32, F64: np.float64, BF16: np.uint16}
_TSIZE = {F32: 4, F64: 8, BF16: 2}

ERRORS = {0: "OK", 1: "INVALID", 2: "UNSUPPORTED", 3: "RUNTIME", 4: "NOMEM", 5: "DEVICE",
          6: "NOT_CONVERGED"}
(ERR_INVALID, ERR_UNSUPPORTED, ERR_RUNTIME, ERR_NOMEM, ERR_DEVICE, ERR_NOT_CONVERGED) = range(1, 7)


class IspmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ispm error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


# ------------------------------------------------------------------ structures -------
class SlellDesc(C.Structure):
    _fields_ = [("n_rows", C.c_int), ("n_cols", C.c_int), ("n_nz", C.c_int),
                ("n_slices", C.c_uint), ("slice_height", C.c_uint), ("hblock", C.c_uint),
                ("var_height", C.c_int), ("diags", C.c_int),
                ("slice_ptr", C.c_void_p), ("slice_ptr_len", C.c_size_t),
                ("cols", C.c_void_p), ("cols_len", C.c_size_t),
                ("elms", C.c_void_p), ("elms_len", C.c_size_t), ("elm_type", C.c_int)]


class Launch(C.Structure):
    _fields_ = [("N", C.c_uint), ("S", C.c_uint), ("H", C.c_uint),
                ("use_tex_sliceptr", C.c_int), ("use_tex_x", C.c_int), ("mode", C.c_int)]


class Csr(C.Structure):
    _fields_ = [("n_rows", C.c_int), ("n_cols", C.c_int), ("n_nz", C.c_int),
                ("row_ptr", C.c_void_p), ("cols", C.c_void_p), ("elms", C.c_void_p),
                ("elm_type", C.c_int)]


class SlellHost(C.Structure):
    _fields_ = [("n_rows", C.c_int), ("n_cols", C.c_int), ("n_nz", C.c_int),
                ("n_slices", C.c_uint), ("slice_height", C.c_uint), ("hblock", C.c_uint),
                ("var_height", C.c_int), ("diags", C.c_int),
                ("slice_ptr", C.c_void_p), ("slice_ptr_len", C.c_size_t),
                ("cols", C.c_void_p), ("cols_len", C.c_size_t),
                ("elms", C.c_void_p), ("elms_len", C.c_size_t), ("elm_type", C.c_int)]


class PlanInfo(C.Structure):
    _fields_ = [("S", C.c_uint), ("hblock", C.c_uint), ("V", C.c_uint), ("D", C.c_uint),
                ("launch", Launch), ("tuned_seconds", C.c_double),
                ("stored_bytes", C.c_size_t), ("csr_bytes", C.c_size_t),
                ("device", SlellDesc), ("pattern_planes", C.c_uint), ("pattern_slices", C.c_uint),
                ("window_slices", C.c_uint)]


class CsrOut(C.Structure):
    _fields_ = [("n_rows", C.c_int), ("n_cols", C.c_int), ("n_nz", C.c_int),
                ("row_ptr", C.POINTER(C.c_int)), ("cols", C.POINTER(C.c_int)),
                ("elms", C.c_void_p), ("elm_type", C.c_int)]


PROGRESS_FN = C.CFUNCTYPE(None, C.c_void_p)

# name -> (restype, argtypes); every symbol declared in include/ispm.h
_SIGS = {
    "ispm_last_error": (C.c_char_p, []),
    "ispm_init": (C.c_int, []),
    "ispm_device_string": (C.c_int, [C.c_char_p, C.c_int]),
    "ispm_malloc": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "ispm_free": (C.c_int, [C.c_void_p]),
    "ispm_memcpy": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "ispm_memcpy_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]),
    "ispm_zero": (C.c_int, [C.c_void_p, C.c_size_t]),
    "ispm_host_alloc_mapped": (C.c_int, [C.POINTER(C.c_void_p), C.c_size_t]),
    "ispm_host_free": (C.c_int, [C.c_void_p]),
    "ispm_host_device_pointer": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p]),
    "ispm_synchronize": (C.c_int, []),
    "ispm_get_stream": (C.c_void_p, []),
    "ispm_set_stream": (C.c_int, [C.c_void_p]),
    "ispm_event_create": (C.c_int, [C.POINTER(C.c_void_p)]),
    "ispm_event_record": (C.c_int, [C.c_void_p]),
    "ispm_event_synchronize": (C.c_int, [C.c_void_p]),
    "ispm_event_elapsed_ms": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_float)]),
    "ispm_event_destroy": (C.c_int, [C.c_void_p]),
    "ispm_spmv": (C.c_int, [C.POINTER(Launch), C.POINTER(SlellDesc), C.c_int, C.c_void_p, C.c_void_p]),
    "ispm_spmm": (C.c_int, [C.POINTER(Launch), C.POINTER(SlellDesc), C.c_int, C.c_int,
                            C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t]),
    "ispm_copy_indexed": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "ispm_add_indexed": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "ispm_dot": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]),
    "ispm_ppcg_update_scalars": (C.c_int, [C.c_int] + [C.c_void_p] * 5),
    "ispm_ppcg_update_vectors": (C.c_int, [C.c_int] + [C.c_void_p] * 15 + [C.c_int]),
    "ispm_ppcg_update_vectors_noalloc": (C.c_int, [C.c_int] + [C.c_void_p] * 15 + [C.c_int]),
    "ispm_plan_execute_guarded": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint]),
    "ispm_plan_x_guard": (C.c_uint, [C.c_void_p]),
    "ispm_csr_to_slell": (C.c_int, [C.POINTER(Csr), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.POINTER(SlellHost)]),
    "ispm_csr_to_slell_guess": (C.c_int, [C.POINTER(Csr), C.c_int, C.c_int, C.c_int,
                                          C.POINTER(SlellHost)]),
    "ispm_slell_host_free": (None, [C.POINTER(SlellHost)]),
    "ispm_slell_upload": (C.c_int, [C.POINTER(SlellHost), C.POINTER(SlellDesc)]),
    "ispm_slell_device_free": (None, [C.POINTER(SlellDesc)]),
    "ispm_plan_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(Csr),
                                   C.c_int, C.c_uint]),
    "ispm_plan_create_explicit": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.POINTER(Csr),
                                            C.c_uint, C.c_uint, C.c_int, C.c_int, C.POINTER(Launch)]),
    "ispm_plan_info_get": (C.c_int, [C.c_void_p, C.POINTER(PlanInfo)]),
    "ispm_plan_default_launch": (C.c_int, [C.c_uint, C.c_int, C.POINTER(Launch)]),
    "ispm_plan_estimate": (C.c_int, [C.c_int, C.c_int, C.POINTER(Csr), C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                     C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(Launch)]),
    "ispm_plan_tune": (C.c_int, [C.c_int, C.c_int, C.POINTER(Csr), C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                 C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(Launch),
                                 C.POINTER(C.c_double)]),
    "ispm_plan_execute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ispm_plan_execute_multi": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t,
                                          C.c_void_p, C.c_size_t]),
    "ispm_plan_update_coefs": (C.c_int, [C.c_void_p, C.POINTER(Csr)]),
    "ispm_plan_update_coefs_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int]),
    "ispm_plan_destroy": (None, [C.c_void_p]),
    "ispm_fastainv_sym": (C.c_int, [C.c_int, C.POINTER(Csr), PROGRESS_FN, C.c_void_p,
                                    C.c_longlong, C.c_double, C.POINTER(CsrOut), C.POINTER(CsrOut)]),
    "ispm_csr_out_free": (None, [C.POINTER(CsrOut)]),
    "ispm_csr_multiply": (C.c_int, [C.POINTER(Csr), C.POINTER(Csr), C.c_int, C.POINTER(CsrOut)]),
    "ispm_plan_time_formats": (C.c_int, [C.c_int, C.c_int, C.POINTER(Csr), C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p]),
    "ispm_plan_time_formats_shared": (C.c_int, [C.c_int, C.c_int, C.POINTER(Csr), C.c_int, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                                C.c_int]),
    "ispm_mlp_param_count": (C.c_int, [C.c_int]),
    "ispm_mlp_forward": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "ispm_mlp_gradient": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_int, C.c_float, C.c_void_p, C.c_void_p]),
    "ispm_mlp_train_step": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float,
                                      C.c_float, C.c_void_p]),
    "ispm_mm_read": (C.c_int, [C.c_char_p, C.POINTER(CsrOut)]),
    "ispm_mm_write": (C.c_int, [C.c_char_p, C.POINTER(Csr)]),
    "ispm_csr_fingerprint": (C.c_int, [C.POINTER(Csr), C.POINTER(C.c_ulonglong)]),
    "ispm_plan_cache_lookup": (C.c_int, [C.c_char_p, C.c_ulonglong, C.c_int, C.c_int, C.POINTER(C.c_uint),
                                         C.POINTER(C.c_uint), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                         C.POINTER(Launch), C.POINTER(C.c_double)]),
    "ispm_plan_cache_store": (C.c_int, [C.c_char_p, C.c_ulonglong, C.c_int, C.c_int, C.c_uint, C.c_uint,
                                        C.c_int, C.c_int, C.POINTER(Launch), C.c_double]),
    "ispm_solver_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ispm_solver_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ispm_solver_destroy": (None, [C.c_void_p]),
    "ispm_solver_set_residual_plan": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ispm_solver_solve_normalised": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double,
                                               C.c_void_p]),
    "ispm_ldu_to_csr": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.POINTER(CsrOut), C.c_void_p]),
    "ispm_ldu_values": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ispm_ldu_values_device": (C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    "ispm_droptol_tuner_create": (C.c_int, [C.c_double, C.c_double, C.POINTER(C.c_void_p)]),
    "ispm_droptol_tuner_droptol": (C.c_double, [C.c_void_p]),
    "ispm_droptol_tuner_best": (C.c_double, [C.c_void_p]),
    "ispm_droptol_tuner_note_time": (None, [C.c_void_p, C.c_double]),
    "ispm_droptol_tuner_restart": (None, [C.c_void_p]),
    "ispm_droptol_tuner_iterations": (C.c_int, [C.c_void_p]),
    "ispm_droptol_tuner_destroy": (None, [C.c_void_p]),
    "ispm_block_solver_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "ispm_block_solver_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_void_p]),
    "ispm_block_solver_destroy": (None, [C.c_void_p]),
    "ispm_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "ispm_batch_layout": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.c_void_p]),
    "ispm_batch_matrix_bytes": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong)]),
    "ispm_batch_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_void_p, C.c_void_p]),
    "ispm_batch_destroy": (None, [C.c_void_p]),
}

# Extra entry points appended by later headers (solver); registered when present.
_OPTIONAL_SIGS: dict = {}

_lib = None


def lib():
    """Load libispm0.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built: run `make -C {PKG_DIR}` "
                              "(or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in {**_SIGS, **_OPTIONAL_SIGS}.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def solve_available():
    return hasattr(lib(), "ispm_solver_create")


def declared_symbols():
    return list(_SIGS) + list(_OPTIONAL_SIGS)


def check(code):
    if code != 0:
        raise IspmError(code, lib().ispm_last_error().decode(errors="replace"))
    return code


def init():
    check(lib().ispm_init())


def device_string():
    buf = C.create_string_buffer(64)
    check(lib().ispm_device_string(buf, 64))
    return buf.value.decode()


def synchronize():
    check(lib().ispm_synchronize())


# ------------------------------------------------------------------ device memory ----
class DeviceArray:
    """A zero-filled device allocation of n elements of a numpy dtype."""

    def __init__(self, n, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.n = int(n)
        p = C.c_void_p()
        check(lib().ispm_malloc(C.byref(p), max(1, self.n) * self.dtype.itemsize))
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().ispm_memcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D))
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size <= self.n
        if a.size:
            check(lib().ispm_memcpy(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D))

    def numpy(self, n=None):
        n = self.n if n is None else n
        out = np.empty(n, self.dtype)
        if n:
            check(lib().ispm_memcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr),
                                    out.nbytes, D2H))
        return out

    def zero(self):
        check(lib().ispm_zero(C.c_void_p(self.ptr), self.nbytes))

    def free(self):
        if self.ptr:
            lib().ispm_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    @property
    def vp(self):
        return C.c_void_p(self.ptr)


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().ispm_event_create(C.byref(p)))
        self.p = p

    def record(self):
        check(lib().ispm_event_record(self.p))

    def synchronize(self):
        check(lib().ispm_event_synchronize(self.p))

    def elapsed_ms(self, later):
        ms = C.c_float()
        check(lib().ispm_event_elapsed_ms(self.p, later.p, C.byref(ms)))
        return ms.value

    def __del__(self):
        try:
            lib().ispm_event_destroy(self.p)
        except Exception:
            pass


# ------------------------------------------------------------------ host CSR ---------
class CSR:
    """Host CSR matrix (csr_matrix<T>, formats/formats.h:18-37): int32 row_ptr / cols and
    float64 (or float32) values."""

    def __init__(self, n_rows, n_cols, row_ptr, cols, vals):
        self.n_rows = int(n_rows)
        self.n_cols = int(n_cols)
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int32)
        self.cols = np.ascontiguousarray(cols, dtype=np.int32)
        self.vals = np.ascontiguousarray(vals)
        if self.vals.dtype not in (np.float64, np.float32):
            self.vals = self.vals.astype(np.float64)
        assert self.row_ptr.size == self.n_rows + 1

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    @property
    def elm_type(self):
        return F64 if self.vals.dtype == np.float64 else F32

    def c(self):
        s = Csr(self.n_rows, self.n_cols, self.nnz, self.row_ptr.ctypes.data, self.cols.ctypes.data,
                self.vals.ctypes.data, self.elm_type)
        s._keep = (self.row_ptr, self.cols, self.vals)
        return s

    @classmethod
    def from_scipy(cls, m):
        m = m.tocsr()
        m.sort_indices()
        return cls(m.shape[0], m.shape[1], m.indptr, m.indices, m.data)

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.vals, self.cols, self.row_ptr), shape=(self.n_rows, self.n_cols))


class HostSlell:
    """slell_matrix<E, host> as numpy arrays (copied out of the library)."""

    def __init__(self, h: SlellHost):
        self.n_rows, self.n_cols, self.n_nz = h.n_rows, h.n_cols, h.n_nz
        self.n_slices, self.slice_height, self.hblock = h.n_slices, h.slice_height, h.hblock
        self.var_height, self.diags, self.elm_type = bool(h.var_height), bool(h.diags), h.elm_type
        self.slice_ptr = np.ctypeslib.as_array(C.cast(h.slice_ptr, C.POINTER(C.c_uint)),
                                               (max(1, h.slice_ptr_len),))[:h.slice_ptr_len].copy()
        self.cols = np.ctypeslib.as_array(C.cast(h.cols, C.POINTER(C.c_uint)),
                                          (max(1, h.cols_len),))[:h.cols_len].copy()
        ct = {F32: C.c_float, F64: C.c_double, BF16: C.c_uint16}[h.elm_type]
        self.elms = np.ctypeslib.as_array(C.cast(h.elms, C.POINTER(ct)),
                                          (max(1, h.elms_len),))[:h.elms_len].copy()

    def elms_f64(self):
        if self.elm_type == BF16:
            return (self.elms.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
        return self.elms.astype(np.float64)

    def params(self):
        return dict(S=self.slice_height, hblock=self.hblock, V=self.var_height, D=self.diags)


def csr_to_slell(A: CSR, S, hblock=1, var_height=False, diags=False, min_height=None, out_type=-1):
    if min_height is None:
        min_height = 4 if var_height or S >= 4 else int(S)
    h = SlellHost()
    check(lib().ispm_csr_to_slell(C.byref(A.c()), int(S), int(hblock), int(var_height), int(diags),
                                  int(min_height), int(out_type), C.byref(h)))
    try:
        return HostSlell(h)
    finally:
        lib().ispm_slell_host_free(C.byref(h))


def csr_to_slell_guess(A: CSR, examine_ratio=1, min_height=4, out_type=-1):
    h = SlellHost()
    check(lib().ispm_csr_to_slell_guess(C.byref(A.c()), int(examine_ratio), int(min_height),
                                        int(out_type), C.byref(h)))
    try:
        return HostSlell(h)
    finally:
        lib().ispm_slell_host_free(C.byref(h))


class DeviceSlell:
    """A SlELL matrix uploaded to the device (slell_matrix<E, device>)."""

    def __init__(self, hs: HostSlell):
        h = SlellHost(hs.n_rows, hs.n_cols, hs.n_nz, hs.n_slices, hs.slice_height, hs.hblock,
                      int(hs.var_height), int(hs.diags),
                      hs.slice_ptr.ctypes.data, hs.slice_ptr.size, hs.cols.ctypes.data, hs.cols.size,
                      hs.elms.ctypes.data, hs.elms.size, hs.elm_type)
        self.desc = SlellDesc()
        check(lib().ispm_slell_upload(C.byref(h), C.byref(self.desc)))
        self.host = hs

    def spmv(self, x: DeviceArray, y: DeviceArray, vec_type, N=64, S=None, H=1, mode=0):
        S = self.desc.slice_height if S is None else S
        lp = Launch(N, S, H, 0, 0, mode)
        return lib().ispm_spmv(C.byref(lp), C.byref(self.desc), vec_type, x.vp, y.vp)

    def spmm(self, k, X: DeviceArray, ldx, Y: DeviceArray, ldy, vec_type, N=64, S=None, H=1, mode=0):
        S = self.desc.slice_height if S is None else S
        lp = Launch(N, S, H, 0, 0, mode)
        return lib().ispm_spmm(C.byref(lp), C.byref(self.desc), vec_type, k, X.vp, ldx, Y.vp, ldy)

    def __del__(self):
        try:
            lib().ispm_slell_device_free(C.byref(self.desc))
        except Exception:
            pass


class Plan:
    """spmv_plan<T,E> (plan/plan.h:20-161) through the C-ABI."""

    def __init__(self, A: CSR, vec_type=F64, elm_type=-1, kind=PLAN_DEFAULT, flags=0,
                 explicit=None, launch=None):
        p = C.c_void_p()
        if explicit is not None:
            lp = C.byref(launch) if launch is not None else None
            check(lib().ispm_plan_create_explicit(C.byref(p), vec_type, elm_type, C.byref(A.c()),
                                                  explicit["S"], explicit.get("hblock", 1),
                                                  int(explicit.get("V", 0)), int(explicit.get("D", 0)), lp))
        else:
            check(lib().ispm_plan_create(C.byref(p), vec_type, elm_type, C.byref(A.c()), kind, flags))
        self.p = p
        self.vec_type = vec_type
        self.flags = flags
        self.n_rows_padded = self.info().device.n_rows

    def info(self):
        inf = PlanInfo()
        check(lib().ispm_plan_info_get(self.p, C.byref(inf)))
        return inf

    def _check_vectors(self, need_x, need_y, *arrays):
        """The C-ABI takes bare device pointers: refuse arrays of another element type than
        the plan's vectors, or shorter than the product reads (x) and writes (y)."""
        want = _NP[self.vec_type]
        for a, need in zip(arrays, (need_x, need_y)):
            if a.dtype != want:
                raise ValueError(f"plan vectors are {np.dtype(want).name}, got {a.dtype.name}")
            if a.n < need:
                raise ValueError(f"device array of {a.n} elements, the product needs {need}")

    def execute(self, x: DeviceArray, y: DeviceArray):
        inf = self.info()
        self._check_vectors(inf.device.n_cols, self.n_rows_padded, x, y)
        check(lib().ispm_plan_execute(self.p, x.vp, y.vp))

    def x_guard(self):
        """Zero elements x needs on each side for ispm_plan_execute_guarded's fast kernel."""
        return int(lib().ispm_plan_x_guard(self.p))

    def execute_guarded(self, xg: DeviceArray, guard: int, y: DeviceArray):
        """y = A x where xg holds `guard` zeros, x (n_cols entries, then the plan's padding)
        and `guard` zeros again (ispm_plan_execute_guarded)."""
        inf = self.info()
        self._check_vectors(inf.device.n_cols + 2 * guard, self.n_rows_padded, xg, y)
        ptr = C.c_void_p(xg.ptr + guard * np.dtype(xg.dtype).itemsize)
        check(lib().ispm_plan_execute_guarded(self.p, ptr, y.vp, int(guard)))

    def execute_multi(self, k, X: DeviceArray, ldx, Y: DeviceArray, ldy):
        inf = self.info()
        self._check_vectors((k - 1) * ldx + inf.device.n_cols, (k - 1) * ldy + self.n_rows_padded, X, Y)
        check(lib().ispm_plan_execute_multi(self.p, k, X.vp, ldx, Y.vp, ldy))

    def update_coefs(self, A: CSR):
        check(lib().ispm_plan_update_coefs(self.p, C.byref(A.c())))

    def update_coefs_device(self, vals: DeviceArray):
        t = F64 if vals.dtype == np.float64 else F32
        check(lib().ispm_plan_update_coefs_device(self.p, vals.vp, t))

    def __del__(self):
        try:
            lib().ispm_plan_destroy(self.p)
        except Exception:
            pass


# ------------------------------------------------------------------ sblas ------------
def dot(type_, x: DeviceArray, y: DeviceArray, s: DeviceArray, n):
    check(lib().ispm_dot(type_, x.vp, y.vp, s.vp, n))


def copy_indexed(type_, dst, src, idx, n):
    check(lib().ispm_copy_indexed(type_, dst.vp, src.vp, idx.vp, n))


def add_indexed(type_, dst, src, idx, n):
    check(lib().ispm_add_indexed(type_, dst.vp, src.vp, idx.vp, n))


def ppcg_update_scalars(type_, alpha, beta, gamma, gammaold, delta):
    check(lib().ispm_ppcg_update_scalars(type_, alpha.vp, beta.vp, gamma.vp, gammaold.vp, delta.vp))


def ppcg_update_vectors(type_, resnorm, gamma, delta, alpha, beta, n, m, p, s, q, z, x, r, u, w, n_elts):
    check(lib().ispm_ppcg_update_vectors(type_, resnorm.vp, gamma.vp, delta.vp, alpha.vp, beta.vp,
                                         n.vp, m.vp, p.vp, s.vp, q.vp, z.vp, x.vp, r.vp, u.vp, w.vp,
                                         n_elts))


# ------------------------------------------------------------------ fastainv ---------
def _csr_from_out(o: CsrOut):
    n = o.n_rows
    rp = np.ctypeslib.as_array(o.row_ptr, (n + 1,)).copy()
    ci = np.ctypeslib.as_array(o.cols, (max(1, o.n_nz),))[:o.n_nz].copy()
    ct = C.c_float if o.elm_type == F32 else C.c_double
    v = np.ctypeslib.as_array(C.cast(o.elms, C.POINTER(ct)), (max(1, o.n_nz),))[:o.n_nz].copy()
    return CSR(n, o.n_cols, rp, ci, v)


def fastainv_sym(A: CSR, out_type=F32, max_elts=8192000, droptol=0.125, progress=None):
    """fastainv_sym<double, float|double> (fastainv/fastainv.h:8-30): returns (Ainv1, Ainv2)."""
    a1, a2 = CsrOut(), CsrOut()
    cb = PROGRESS_FN(progress) if progress else PROGRESS_FN()
    Af = A if A.vals.dtype == np.float64 else CSR(A.n_rows, A.n_cols, A.row_ptr, A.cols,
                                                  A.vals.astype(np.float64))
    check(lib().ispm_fastainv_sym(out_type, C.byref(Af.c()), cb, None, int(max_elts),
                                  float(droptol), C.byref(a1), C.byref(a2)))
    try:
        return _csr_from_out(a1), _csr_from_out(a2)
    finally:
        lib().ispm_csr_out_free(C.byref(a1))
        lib().ispm_csr_out_free(C.byref(a2))


def csr_multiply(A: CSR, B: CSR, out_type=F64) -> CSR:
    """C = A * B on the host (FP64 accumulation, sorted columns)."""
    o = CsrOut()
    check(lib().ispm_csr_multiply(C.byref(A.c()), C.byref(B.c()), out_type, C.byref(o)))
    try:
        return _csr_from_out(o)
    finally:
        lib().ispm_csr_out_free(C.byref(o))


# ------------------------------------------------------------------ on-disk ----------
def mm_read(path) -> CSR:
    o = CsrOut()
    check(lib().ispm_mm_read(str(path).encode(), C.byref(o)))
    try:
        return _csr_from_out(o)
    finally:
        lib().ispm_csr_out_free(C.byref(o))


def mm_write(path, A: CSR):
    check(lib().ispm_mm_write(str(path).encode(), C.byref(A.c())))


def fingerprint(A: CSR) -> int:
    h = C.c_ulonglong()
    check(lib().ispm_csr_fingerprint(C.byref(A.c()), C.byref(h)))
    return h.value
