"""ctypes binding of liblime_hip.so (the C ABI declared in include/lime_hip.h).

The library is built in-tree by ``python -m lime_amd.build``.  There is no CPU fallback:
if the library or a GPU is missing, every entry point raises.
"""
import ctypes
import os
import threading

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "liblime_hip.so")

LIME_OK = 0
LIME_ERR_ARG = -1
LIME_ERR_UNSUPPORTED = -2
LIME_ERR_HIP = -3

LIME_F_RENORM_TRACE = 1
LIME_F_PREC_FP32 = 0x00
LIME_F_PREC_3XBF16 = 0x10
LIME_F_PREC_FP64 = 0x20
LIME_F_FORCE_TILED = 0x100
LIME_F_ONE_SIDED = 0x200
LIME_NOALLOC_SCRATCH_BYTES = 64 << 20

#: every symbol include/lime_hip.h declares (checked by tests/test_abi.py)
EXPORTS = (
    "lime_lindblad_rk4",
    "lime_lindblad_rk4_noalloc",
    "lime_lindblad_rk4_ragged",
    "lime_ragged_variants",
    "lime_lindblad_max_n",
    "lime_lindblad_workspace_bytes",
    "lime_lindblad_variants",
    "lime_lindblad_apply",
    "lime_adjoint_contract",
    "lime_adjoint_contract_pairs",
    "lime_adjoint_seed",
    "lime_last_error",
    "lime_version",
)


class LimeError(RuntimeError):
    """A negative status returned by liblime_hip.so."""

    def __init__(self, code, msg):
        super().__init__(f"lime_hip error {code}: {msg}")
        self.code = code


class LimeUnsupported(LimeError):
    pass


class LindbladBatch(ctypes.Structure):
    _fields_ = [
        ("batch", ctypes.c_int32),
        ("n", ctypes.c_int32),
        ("group", ctypes.c_int32),
        ("n_dense", ctypes.c_int32),
        ("n_eops", ctypes.c_int32),
        ("eop_group", ctypes.c_int32),
        ("flags", ctypes.c_int32),
        ("variant", ctypes.c_int32),
        ("heff", ctypes.c_void_p),
        ("dmat", ctypes.c_void_p),
        ("lops", ctypes.c_void_p),
        ("eops_t", ctypes.c_void_p),
        ("rho", ctypes.c_void_p),
        ("obs", ctypes.c_void_p),
        ("states", ctypes.c_void_p),
        ("rops", ctypes.c_void_p),
        ("work", ctypes.c_void_p),
        ("work_bytes", ctypes.c_int64),
        ("stage_states", ctypes.c_void_p),
    ]


class RaggedInstance(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int32),
        ("n_dense", ctypes.c_int32),
        ("heff", ctypes.c_int64),
        ("dmat", ctypes.c_int64),
        ("lops", ctypes.c_int64),
        ("rho", ctypes.c_int64),
        ("work", ctypes.c_int64),
    ]


class RaggedBatch(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_int32),
        ("variant", ctypes.c_int32),
        ("max_groups", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("inst", ctypes.c_void_p),
        ("order", ctypes.c_void_p),
        ("ops", ctypes.c_void_p),
        ("rho", ctypes.c_void_p),
        ("work", ctypes.c_void_p),
        ("queue", ctypes.c_void_p),
    ]


class StepParams(ctypes.Structure):
    _fields_ = [
        ("dt", ctypes.c_double),
        ("nsteps", ctypes.c_int32),
        ("nout", ctypes.c_int32),
        ("step_offset", ctypes.c_int32),
        ("rec_capacity", ctypes.c_int32),
    ]


_lock = threading.Lock()
_lib = None


def load(path=LIB_PATH):
    """Load (once) and return the ctypes handle; raises if the library is not built."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"{path} is missing: build it with `python -m lime_amd.build` "
                "(lime_amd has no CPU fallback)")
        lib = ctypes.CDLL(path)
        lib.lime_lindblad_rk4.argtypes = [ctypes.POINTER(LindbladBatch),
                                          ctypes.POINTER(StepParams), ctypes.c_void_p]
        lib.lime_lindblad_rk4.restype = ctypes.c_int
        lib.lime_lindblad_rk4_noalloc.argtypes = [ctypes.POINTER(LindbladBatch),
                                                  ctypes.POINTER(StepParams), ctypes.c_void_p]
        lib.lime_lindblad_rk4_noalloc.restype = ctypes.c_int
        lib.lime_lindblad_rk4_ragged.argtypes = [ctypes.POINTER(RaggedBatch),
                                                 ctypes.POINTER(StepParams), ctypes.c_void_p]
        lib.lime_lindblad_rk4_ragged.restype = ctypes.c_int
        lib.lime_ragged_variants.argtypes = []
        lib.lime_ragged_variants.restype = ctypes.c_int
        lib.lime_lindblad_max_n.argtypes = [ctypes.c_int32, ctypes.c_int32]
        lib.lime_lindblad_max_n.restype = ctypes.c_int
        lib.lime_lindblad_workspace_bytes.argtypes = [ctypes.POINTER(LindbladBatch)]
        lib.lime_lindblad_workspace_bytes.restype = ctypes.c_int64
        lib.lime_lindblad_variants.argtypes = [ctypes.POINTER(LindbladBatch)]
        lib.lime_lindblad_variants.restype = ctypes.c_int
        vp, dp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)
        lib.lime_lindblad_apply.argtypes = [ctypes.POINTER(LindbladBatch), vp, vp, vp, vp, vp,
                                            dp, dp, dp, vp]
        lib.lime_lindblad_apply.restype = ctypes.c_int
        lib.lime_adjoint_contract.argtypes = [ctypes.c_int32, ctypes.c_int32, vp, vp,
                                              ctypes.c_double, vp, vp, vp, vp]
        lib.lime_adjoint_contract.restype = ctypes.c_int
        lib.lime_adjoint_contract_pairs.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                                    ctypes.POINTER(ctypes.c_void_p),
                                                    ctypes.POINTER(ctypes.c_void_p), ctypes.c_double,
                                                    vp, vp, vp, vp]
        lib.lime_adjoint_contract_pairs.restype = ctypes.c_int
        lib.lime_adjoint_seed.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, vp, vp, vp,
                                          ctypes.c_double, vp, vp, vp, vp]
        lib.lime_adjoint_seed.restype = ctypes.c_int
        lib.lime_last_error.argtypes = []
        lib.lime_last_error.restype = ctypes.c_char_p
        lib.lime_version.argtypes = []
        lib.lime_version.restype = ctypes.c_char_p
        _lib = lib
        return lib


def check(code):
    if code == LIME_OK:
        return
    msg = load().lime_last_error().decode(errors="replace")
    if code == LIME_ERR_UNSUPPORTED:
        raise LimeUnsupported(code, msg)
    raise LimeError(code, msg)
