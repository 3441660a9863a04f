"""ctypes binding of libflare_mi.so (include/flare_mi.h).

The library is built in-tree (``python flare_amd/_build.py`` or
``__graft_entry__.build()``). There is no fallback: if the HIP library is
missing or fails to load, every entry point raises.
"""
import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflare_mi.so")

FM_OK = 0
FM_ERR_INVALID = 1
FM_ERR_HIP = 2
FM_ERR_STATE = 3
FM_ERR_CAPACITY = 4

FORM_FP64 = 0
FORM_CHOSEN = 1

CUTOFF_QUADRATIC = 0
CUTOFF_COSINE = 1

PREDICT_MEAN = 1
PREDICT_LOCAL = 2
DENSE_DTC = 1
DENSE_SOR = 2

# Every symbol include/flare_mi.h declares.
EXPORTS = [
    "fm_last_error", "fm_context_create", "fm_context_destroy", "fm_context_synchronize",
    "fm_context_stream", "fm_structure_create", "fm_structure_create_device",
    "fm_structure_destroy", "fm_structure_info", "fm_structure_get_neighbors",
    "fm_structure_compute_b2", "fm_structure_b2_sizes", "fm_structure_get_b2",
    "fm_model_create", "fm_model_destroy", "fm_predict", "fm_last_timings",
    "fm_model_mapping_coefficients", "fm_mapped_create", "fm_mapped_read_file",
    "fm_mapped_info", "fm_mapped_destroy", "fm_mapped_compute", "fm_mapped_compute_noalloc",
    "fm_batch_create", "fm_batch_create_device", "fm_batch_info",
    "fm_structure_get_b2_dervs", "fm_envs_struc", "fm_self_kernel_struc", "fm_predict_dense",
    "fm_structure_compute_b3", "fm_context_set_row_tile",
    "fm_server_create", "fm_server_destroy", "fm_submit", "fm_poll",
    "fm_last_kernel_timings",
]

DESCRIPTOR_B2 = 2
DESCRIPTOR_B3 = 3


class FlareMIError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class B2Settings(C.Structure):
    _fields_ = [("n_species", C.c_int), ("n_max", C.c_int), ("l_max", C.c_int),
                ("cutoff_function", C.c_int), ("radial_r1", C.c_double),
                ("cutoffs", C.POINTER(C.c_double))]


class MappedDesc(C.Structure):
    _fields_ = [("power", C.c_int), ("normalized", C.c_int), ("n_species", C.c_int),
                ("n_max", C.c_int), ("l_max", C.c_int), ("cutoff_function", C.c_int),
                ("cutoffs", C.POINTER(C.c_double)), ("beta", C.POINTER(C.c_double))]


class ModelDesc(C.Structure):
    _fields_ = [("sigma", C.c_double), ("power", C.c_double), ("n_species", C.c_int),
                ("n_descriptors", C.c_int), ("n_sparse_by_type", C.POINTER(C.c_int32)),
                ("sparse_descriptors", C.POINTER(C.c_double)),
                ("sparse_norms", C.POINTER(C.c_double)), ("alpha", C.POINTER(C.c_double)),
                ("L_inv", C.POINTER(C.c_double)), ("R_inv", C.POINTER(C.c_double))]


class ServerDesc(C.Structure):
    _fields_ = [("n_kernels", C.c_int), ("models", C.POINTER(ModelDesc)),
                ("descriptors", C.POINTER(B2Settings)), ("descriptor_kinds", C.POINTER(C.c_int32)),
                ("cutoff", C.c_double), ("form", C.c_int), ("flags", C.c_int),
                ("n_workers", C.c_int), ("hbm_budget", C.c_int64),
                ("max_batch_atoms", C.c_int32), ("max_density", C.c_double)]


_lib = None
_lock = threading.Lock()


def load():
    """Load libflare_mi.so; raises if it is absent (no CPU fallback)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise FlareMIError(FM_ERR_STATE, f"{LIB_PATH} is not built; run flare_amd/_build.py")
        L = C.CDLL(LIB_PATH)
        vp, ip, dp = C.c_void_p, C.c_int, C.c_double
        pvp = C.POINTER(C.c_void_p)
        L.fm_last_error.restype = C.c_char_p
        L.fm_context_create.argtypes = [ip, pvp]
        L.fm_context_destroy.argtypes = [vp]
        L.fm_context_synchronize.argtypes = [vp]
        L.fm_context_set_row_tile.argtypes = [vp, C.c_int64]
        L.fm_context_stream.argtypes = [vp, pvp]
        L.fm_structure_create.argtypes = [vp, vp, ip, vp, vp, dp, pvp]
        L.fm_structure_create_device.argtypes = [vp, vp, ip, vp, vp, dp, pvp]
        L.fm_structure_destroy.argtypes = [vp]
        L.fm_structure_info.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                        C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.fm_structure_get_neighbors.argtypes = [vp] * 7
        L.fm_structure_compute_b2.argtypes = [vp, C.POINTER(B2Settings), ip]
        L.fm_structure_compute_b3.argtypes = [vp, C.POINTER(B2Settings), ip]
        L.fm_structure_b2_sizes.argtypes = [vp, C.POINTER(C.c_int), vp, vp]
        L.fm_structure_get_b2.argtypes = [vp, ip] + [vp] * 7
        L.fm_model_create.argtypes = [vp, C.POINTER(ModelDesc), pvp]
        L.fm_model_destroy.argtypes = [vp]
        L.fm_predict.argtypes = [vp, vp, ip, ip, vp, vp, vp]
        L.fm_last_timings.argtypes = [vp, vp, vp]
        L.fm_last_kernel_timings.argtypes = [vp, ip, vp, vp]
        L.fm_model_mapping_coefficients.argtypes = [vp, vp, C.POINTER(C.c_int64)]
        L.fm_mapped_create.argtypes = [vp, C.POINTER(MappedDesc), pvp]
        L.fm_mapped_read_file.argtypes = [vp, C.c_char_p, pvp]
        L.fm_mapped_info.argtypes = [vp] * 8
        L.fm_mapped_destroy.argtypes = [vp]
        L.fm_mapped_compute.argtypes = [vp, vp, vp, vp, vp]
        L.fm_mapped_compute_noalloc.argtypes = [vp, vp, vp, vp, vp]
        L.fm_batch_create.argtypes = [vp, ip, vp, vp, vp, vp, dp, pvp]
        L.fm_batch_create_device.argtypes = [vp, ip, vp, vp, vp, vp, dp, pvp]
        L.fm_batch_info.argtypes = [vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                    C.POINTER(C.c_double)]
        L.fm_server_create.argtypes = [ip, C.POINTER(ServerDesc), pvp]
        L.fm_server_destroy.argtypes = [vp]
        L.fm_submit.argtypes = [vp, vp, ip, vp, vp, C.POINTER(C.c_int64)]
        L.fm_poll.argtypes = [vp, C.c_int64, ip, C.POINTER(C.c_int), vp, vp]
        for name in EXPORTS:
            getattr(L, name).restype = C.c_char_p if name == "fm_last_error" else C.c_int
        _lib = L
        return L


def check(rc):
    if rc != FM_OK:
        msg = load().fm_last_error().decode()
        if rc == FM_ERR_STATE:
            raise RuntimeError(msg)
        if rc == FM_ERR_INVALID:
            raise ValueError(msg)
        raise FlareMIError(rc, msg)


def ptr(a):
    if a is None:
        return None
    return C.c_void_p(a.ctypes.data)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """A HIP stream and device-buffer pool on one GPU."""

    def __init__(self, device=0):
        self.L = load()
        h = C.c_void_p()
        check(self.L.fm_context_create(int(device), C.byref(h)))
        self.h = h
        self.device = device

    def synchronize(self):
        check(self.L.fm_context_synchronize(self.h))

    def close(self):
        """Releases the context. Structures, models and potentials made on it
        stay usable; the library deletes the context (its streams and buffer
        pool) when the last of them is destroyed."""
        if self.h is not None:
            check(self.L.fm_context_destroy(self.h))
            self.h = None

    def __del__(self):
        # The library defers the deletion until the objects made on the
        # context are gone, whatever order the collector finalises them in;
        # at interpreter exit the process releases the device anyway.
        try:
            import sys
            if not sys.is_finalizing():
                self.close()
        except Exception:
            pass

    def set_row_tile(self, rows):
        """Row tiles for B3 descriptors computed on this context afterwards
        (0: tile only structures whose descriptors would exceed 32 GB)."""
        check(self.L.fm_context_set_row_tile(self.h, int(rows)))

    @property
    def stream(self):
        s = C.c_void_p()
        check(self.L.fm_context_stream(self.h, C.byref(s)))
        return s.value

    STAGES = ["neighbors", "descriptors", "kernel_contraction", "gradient_contraction",
              "variance", "adjoint_gather"]

    KERNELS = ["descriptor", "slice", "kernel_contraction", "gradient_contraction",
               "variance_contraction", "adjoint_dEdc", "bond_forces", "gather"]

    def kernel_timings(self):
        """{kernel: (ms summed over launches, launches)} of the last calls."""
        ms = np.zeros(len(self.KERNELS))
        n = np.zeros(len(self.KERNELS), np.int32)
        check(self.L.fm_last_kernel_timings(self.h, len(self.KERNELS), ptr(ms), ptr(n)))
        return {k: (float(a), int(b)) for k, a, b in zip(self.KERNELS, ms, n)}

    def timings(self):
        """{stage: (ms, launches)} of the stages last run on this context."""
        ms = np.zeros(6)
        n = np.zeros(6, np.int32)
        check(self.L.fm_last_timings(self.h, ptr(ms), ptr(n)))
        return {k: (float(a), int(b)) for k, a, b in zip(self.STAGES, ms, n)}


_contexts = {}


def default_context(device=0):
    with _lock:
        ctx = _contexts.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _contexts[device] = ctx
    return ctx
