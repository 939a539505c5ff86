"""ctypes binding of libscaletrack.so (the C-ABI in include/scaletrack.h).

The library is the only compute path: if it is missing this module raises at import
time instead of falling back to anything else.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libscaletrack.so")

ST_OK = 0
ST_E_INVALID = -1
ST_E_DEVICE = -2
ST_E_NOMEM = -3
ST_E_STATE = -4
ST_E_UNSUPPORTED = -5
ST_E_RANGE = -6

ST_FP64 = 0
ST_FP32 = 1
ST_WALL = 0
ST_PERIODIC = 1
ST_CHANNEL = 2
ST_METHOD_PLAIN = 0
ST_METHOD_SORTED = 1
ST_FLUSH_OWNER = 0
ST_FLUSH_ATOMIC = 1

STATUS_NAMES = {
    ST_E_INVALID: "ST_E_INVALID",
    ST_E_DEVICE: "ST_E_DEVICE",
    ST_E_NOMEM: "ST_E_NOMEM",
    ST_E_STATE: "ST_E_STATE",
    ST_E_UNSUPPORTED: "ST_E_UNSUPPORTED",
    ST_E_RANGE: "ST_E_RANGE",
}


class st_mesh(C.Structure):
    _fields_ = [("n", C.c_int32 * 3), ("origin", C.c_double * 3), ("ending", C.c_double * 3)]


class st_params(C.Structure):
    _fields_ = [
        ("boundary", C.c_int32),
        ("method", C.c_int32),
        ("flush", C.c_int32),
        ("substeps_per_pass", C.c_int32),
        ("cd", C.c_double),
        ("brick", C.c_int32 * 3),
        ("block", C.c_int32),
        ("resort_fraction", C.c_double),
        ("passes_per_launch", C.c_int32),
    ]


class st_tune_result(C.Structure):
    _fields_ = [
        ("target_parcels_per_brick", C.c_int32),
        ("block", C.c_int32),
        ("blocks_per_cu", C.c_int32),
        ("from_cache", C.c_int32),
        ("ms_per_pass", C.c_double),
        ("runner_up_ms_per_pass", C.c_double),
    ]


class ScaleTrackError(RuntimeError):
    """A nonzero status from the C-ABI, with the library's message."""

    def __init__(self, status, message):
        self.status = status
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol declared in include/scaletrack.h
SIGNATURES = {
    "st_version": (C.c_int, []),
    "st_last_error": (C.c_char_p, []),
    "st_params_default": (None, [C.POINTER(st_params)]),
    "st_device_count": (C.c_int, [C.POINTER(C.c_int32)]),
    "st_ctx_create": (C.c_int, [C.c_int32, _pp]),
    "st_ctx_destroy": (C.c_int, [_vp]),
    "st_ctx_synchronize": (C.c_int, [_vp]),
    "st_cloud_create": (C.c_int, [_vp, C.c_int32, C.c_int64, C.c_double, C.c_double, C.c_double,
                                  C.c_double, C.c_int32, _pp]),
    "st_cloud_destroy": (C.c_int, [_vp]),
    "st_cloud_upload": (C.c_int, [_vp] + [_vp] * 8),
    "st_cloud_download": (C.c_int, [_vp] + [_vp] * 7),
    "st_cloud_count": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "st_cloud_reserve": (C.c_int, [_vp, C.c_int64]),
    "st_cloud_inject": (C.c_int, [_vp, C.c_int64] + [_vp] * 8),
    "st_cloud_compact": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "st_field_create": (C.c_int, [_vp, C.POINTER(st_mesh), C.c_int32, _pp]),
    "st_field_destroy": (C.c_int, [_vp]),
    "st_field_upload_U": (C.c_int, [_vp, _vp, C.c_int32]),
    "st_field_zero_source": (C.c_int, [_vp]),
    "st_field_merge_source": (C.c_int, [_vp]),
    "st_field_download_source": (C.c_int, [_vp, _vp, C.c_int32]),
    "st_evolve": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int32, C.POINTER(st_params),
                            C.POINTER(C.c_double)]),
    "st_evolve_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int32, C.POINTER(st_params),
                                    C.POINTER(C.c_double)]),
    "st_cloud_stray_fraction": (C.c_int, [_vp, C.POINTER(C.c_double)]),
    "st_cloud_sort": (C.c_int, [_vp, _vp, _vp, C.POINTER(st_params)]),
    "st_cloud_resort_count": (C.c_int, [_vp, C.POINTER(C.c_int64)]),
    "st_graph_capture": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_int32, C.POINTER(st_params),
                                   C.c_int32, _pp]),
    "st_graph_launch": (C.c_int, [_vp]),
    "st_graph_destroy": (C.c_int, [_vp]),
    "st_last_kernel_ms": (C.c_int, [_vp, C.POINTER(C.c_double), C.POINTER(C.c_int32)]),
    "st_host_alloc": (C.c_int, [C.c_int64, _pp]),
    "st_host_free": (C.c_int, [_vp]),
    "st_diag_copy_bandwidth": (C.c_int, [_vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "st_diag_kernel_times": (C.c_int, [_vp, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_int32)]),
    "st_batch_create": (C.c_int, [_vp, C.c_int32, _pp, _pp, C.POINTER(st_params), C.c_int32, C.c_int32, _pp]),
    "st_batch_destroy": (C.c_int, [_vp]),
    "st_batch_sort": (C.c_int, [_vp]),
    "st_batch_evolve": (C.c_int, [_vp, C.c_double, C.c_int32]),
    "st_batch_graph_capture": (C.c_int, [_vp, C.c_double, C.c_int32, C.c_int32, _pp]),
    "st_batch_info": (C.c_int, [_vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "st_autotune_batch": (C.c_int, [_vp, C.c_int32, _pp, _pp, C.POINTER(st_params), C.c_double, C.c_int32,
                                    C.c_char_p, C.POINTER(st_tune_result)]),
    "st_adjoint_create": (C.c_int, [_vp, _vp, _vp, C.c_int32, C.c_int32, C.POINTER(st_params), _pp]),
    "st_adjoint_destroy": (C.c_int, [_vp]),
    "st_adjoint_forward": (C.c_int, [_vp, C.c_double]),
    "st_adjoint_source_misfit": (C.c_int, [_vp, _vp, C.POINTER(C.c_double)]),
    "st_adjoint_set_source_cotangent": (C.c_int, [_vp, _vp]),
    "st_adjoint_set_state_cotangent": (C.c_int, [_vp, _vp, _vp]),
    "st_adjoint_backward": (C.c_int, [_vp]),
    "st_adjoint_gradients": (C.c_int, [_vp, _vp, _vp, C.POINTER(C.c_double), _vp, _vp]),
}


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise ImportError(
            f"libscaletrack.so not found at {path}; build it first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = load()


def check(status):
    if status != ST_OK:
        msg = lib.st_last_error()
        raise ScaleTrackError(status, msg.decode() if msg else "")
    return status


def default_params(**overrides):
    p = st_params()
    lib.st_params_default(C.byref(p))
    for k, v in overrides.items():
        if k == "brick":
            for i in range(3):
                p.brick[i] = int(v[i])
        else:
            setattr(p, k, v)
    return p
