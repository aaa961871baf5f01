"""ctypes binding of libcms_hip.so (the C ABI declared in include/cms_hip.h).

The library is built in-tree by ``cms_cuda._build`` (``__graft_entry__.build``).
Loading fails loudly when it is missing: there is no CPU fallback.
"""

import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcms_hip.so")

CMS_OK = 0
CMS_ERR_INVALID = 1
CMS_ERR_HIP = 2
CMS_ERR_COINCIDENT = 3
CMS_ERR_CAPACITY = 4
CMS_ERR_STATE = 5

PREC_FP64 = 0
PREC_FP32_RESCORE = 1

NOALLOC_MAX_POINTS = 131072

NCOUNTS = 15
SC_NRES = 24


class cms_tuning(ctypes.Structure):
    _fields_ = [("trim_lanes", ctypes.c_int), ("probe_count", ctypes.c_int), ("probe_slots", ctypes.c_int),
                ("cell_sweep", ctypes.c_int), ("search_split", ctypes.c_int)]


class cms_settings(ctypes.Structure):
    _fields_ = [("rp", ctypes.c_double), ("density", ctypes.c_double), ("band", ctypes.c_double),
                ("sep", ctypes.c_double), ("weight", ctypes.c_double),
                ("binwidth_dist", ctypes.c_double), ("binwidth_norm", ctypes.c_double)]


_c = ctypes
_vp = _c.c_void_p
_i32 = _c.c_int32
_i64 = _c.c_int64
_PROTOS = {
    "cms_version": (_c.c_int, []),
    "cms_create": (_c.c_int, [_c.c_int, _vp, _c.POINTER(_vp)]),
    "cms_destroy": (_c.c_int, [_vp]),
    "cms_last_error": (_c.c_char_p, [_vp]),
    "cms_synchronize": (_c.c_int, [_vp]),
    "cms_load": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _vp, _vp, _vp, _c.POINTER(cms_settings)]),
    "cms_generate": (_c.c_int, [_vp, _c.c_int, _c.c_int]),
    "cms_status": (_c.c_int, [_vp, _vp]),
    "cms_failed_count": (_c.c_int, [_vp, _c.POINTER(_i32)]),
    "cms_set_tuning": (_c.c_int, [_vp, _vp]),
    "cms_get_tuning": (_c.c_int, [_vp, _vp]),
    "cms_coincident": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_i32), _c.POINTER(_i32)]),
    "cms_counts": (_c.c_int, [_vp, _vp]),
    "cms_contact_ms": (_c.c_int, [_vp, _c.c_int, _vp, _vp]),
    "cms_max_contact_ms": (_c.c_int, [_vp, _vp, _vp]),
    "cms_shape_complementarity": (_c.c_int, [_vp, _c.c_int, _vp]),
    "cms_run": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _vp]),
    "cms_set_graph": (_c.c_int, [_vp, _c.c_int]),
    "cms_set_keep_descriptors": (_c.c_int, [_vp, _c.c_int]),
    "cms_run_stats": (_c.c_int, [_vp, _vp]),
    "cms_dot_count": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.POINTER(_i64)]),
    "cms_copy_dots": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cms_copy_atom_state": (_c.c_int, [_vp, _vp, _vp]),
    "cms_copy_dot_desc": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _vp]),
    "cms_probe_count": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_i64)]),
    "cms_copy_probes": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _vp, _vp]),
    "cms_copy_rows": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _vp, _c.c_int]),
    "cms_trimmed_count": (_c.c_int, [_vp, _c.c_int, _c.c_int, _c.POINTER(_i64)]),
    "cms_copy_trimmed_dots": (_c.c_int, [_vp, _c.c_int, _c.c_int, _vp, _vp, _vp, _vp, _vp, _vp]),
    "cms_toroid_count": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_i64)]),
    "cms_copy_toroid_queue": (_c.c_int, [_vp, _c.c_int, _vp, _vp, _vp, _vp, _vp]),
    "cms_nearest": (_c.c_int, [_vp, _c.c_int, _vp, _i64, _vp, _i64, _vp, _vp]),
    "cms_nearest_noalloc": (_c.c_int, [_vp, _c.c_int, _vp, _i64, _vp, _i64, _vp, _vp]),
    "cms_profile": (_c.c_int, [_vp, _c.c_char_p]),
    "cms_backward_cms": (_c.c_int, [_vp, _vp, _vp, _vp]),
    "cms_frozen_cms": (_c.c_int, [_vp, _vp, _vp]),
    "cms_check_dots": (_c.c_int, [_vp, _c.POINTER(_i64), _c.POINTER(_c.c_double)]),
    "cms_search_stats": (_c.c_int, [_vp, _c.POINTER(_c.c_uint64)]),
    "cms_profile_read": (_c.c_int, [_vp, _c.c_char_p, _c.POINTER(_c.c_double), _c.POINTER(_i64)]),
}
OPTIONAL = {}

_lock = threading.Lock()
_lib = None


def exported_symbols():
    """Names of the C ABI functions this binding expects."""
    return sorted(_PROTOS)


def load():
    """Load libcms_hip.so (once) and attach prototypes; raise if missing."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libcms_hip.so is not built (expected at %s); run __graft_entry__.build() "
                "or python cms-cuda_amd/cms_cuda/_build.py" % LIB_PATH)
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _PROTOS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        for name, (res, args) in OPTIONAL.items():
            fn = getattr(lib, name, None)
            if fn is not None:
                fn.restype = res
                fn.argtypes = args
        _lib = lib
        return lib


class CmsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


def check(code, ctx=None):
    if code == CMS_OK:
        return
    lib = load()
    msg = lib.cms_last_error(ctx)
    msg = msg.decode() if msg else "error %d" % code
    if code == CMS_ERR_INVALID:
        raise ValueError(msg)
    raise CmsError(code, msg)


class Context:
    """Owns one cms_ctx bound to a device and a HIP stream."""

    def __init__(self, device=0, stream=None):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.cms_create(int(device), stream, ctypes.byref(h)), None)
        self.handle = h
        self.device = device
        self.stream = stream
        self.lib = lib

    def close(self):
        if self.handle:
            self.lib.cms_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __call__(self, name, *args):
        code = getattr(self.lib, name)(self.handle, *args)
        check(code, self.handle)
        return code


_pool = threading.local()


def acquire_context(device, stream):
    """A warm context from this thread's pool for (device, stream)."""
    pools = getattr(_pool, "pools", None)
    if pools is None:
        pools = _pool.pools = {}
    free = pools.setdefault((device, stream), [])
    if free:
        return free.pop()
    return Context(device, stream)


def release_context(ctx):
    pools = getattr(_pool, "pools", None)
    if pools is None:
        pools = _pool.pools = {}
    free = pools.setdefault((ctx.device, ctx.stream), [])
    if len(free) < 8:
        free.append(ctx)
    else:
        ctx.close()
