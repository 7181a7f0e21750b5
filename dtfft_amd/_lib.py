"""ctypes binding of libdtfft_hip.so (the C-ABI declared in include/dtfft_hip.h).

The shared library is built in-tree by ``python -m dtfft_amd.build``.  There is no
fallback: if the library is missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libdtfft_hip.so")


class DtfftError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        msg = f"{where}: dtfft_hip error {code} ({error_string(code)})"
        if detail:
            msg += f"\n{detail}"
        super().__init__(msg)


class Config(C.Structure):
    """dtfft_hip_config_t"""
    _fields_ = [
        ("enable_z_slab", C.c_int32),
        ("enable_fused", C.c_int32),
        ("enable_kernel_autotune", C.c_int32),
        ("measure_warmup_iters", C.c_int32),
        ("measure_iters", C.c_int32),
        ("stream", C.c_void_p),
        ("lds_dma", C.c_int32),
        ("twl", C.c_int32),
        ("blocked", C.c_int32),
    ]


_i32, _i64, _i8, _vp, _sz = C.c_int32, C.c_int64, C.c_int8, C.c_void_p, C.c_size_t
_pi32, _pi64, _psz, _pvp = C.POINTER(_i32), C.POINTER(_i64), C.POINTER(_sz), C.POINTER(_vp)

# name -> (restype, argtypes); every symbol of include/dtfft_hip.h
SIGNATURES = {
    "dtfft_hip_get_version": (_i32, []),
    "dtfft_hip_get_error_string": (C.c_char_p, [_i32]),
    "dtfft_hip_get_last_error_message": (C.c_char_p, []),
    "dtfft_hip_set_cache_dir": (_i32, [C.c_char_p]),
    "dtfft_hip_set_compile_only": (_i32, [_i32]),
    "dtfft_hip_kernel_create": (_i32, [_i32, _i32, _pi32, _i64, _pi32, _i32, _i32, _i32, _pvp]),
    "dtfft_hip_kernel_execute": (_i32, [_vp, _vp, _vp, _vp, _i32, _i32]),
    "dtfft_hip_kernel_destroy": (_i32, [_vp]),
    "dtfft_hip_fft_create": (_i32, [_i8, _i8, _i32, _i32, _i32, _i32, _pi32, _pi32, _pi32, _pi32, _i32, _vp,
                                    _pvp]),
    "dtfft_hip_fft_execute": (_i32, [_vp, _vp, _vp, _i8]),
    "dtfft_hip_fft_destroy": (_i32, [_vp]),
    "dtfft_hip_mem_alloc": (_i32, [_sz, _pvp]),
    "dtfft_hip_mem_free": (_i32, [_vp]),
    "dtfft_hip_create_config": (_i32, [C.POINTER(Config)]),
    "dtfft_hip_set_config": (_i32, [C.POINTER(Config)]),
    "dtfft_hip_create_plan_c2c": (_i32, [_i8, _pi32, _i32, _i32, _i32, _pvp]),
    "dtfft_hip_create_plan_r2c": (_i32, [_i8, _pi32, _i32, _i32, _i32, _pvp]),
    "dtfft_hip_create_plan_r2r": (_i32, [_i8, _pi32, _pi32, _i32, _i32, _i32, _pvp]),
    "dtfft_hip_execute": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "dtfft_hip_execute_no_alloc": (_i32, [_vp, _vp, _vp, _i32]),
    "dtfft_hip_transpose": (_i32, [_vp, _vp, _vp, _i32, _vp]),
    "dtfft_hip_destroy": (_i32, [_pvp]),
    "dtfft_hip_execute_cycle_graph": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "dtfft_hip_get_local_sizes": (_i32, [_vp, _pi32, _pi32, _pi32, _pi32, _psz]),
    "dtfft_hip_get_alloc_size": (_i32, [_vp, _psz]),
    "dtfft_hip_get_element_size": (_i32, [_vp, _psz]),
    "dtfft_hip_get_alloc_bytes": (_i32, [_vp, _psz]),
    "dtfft_hip_get_aux_bytes": (_i32, [_vp, _psz]),
    "dtfft_hip_get_z_slab_enabled": (_i32, [_vp, _pi32]),
    "dtfft_hip_plan_mem_alloc": (_i32, [_vp, _sz, _pvp]),
    "dtfft_hip_plan_mem_free": (_i32, [_vp, _vp]),
    "dtfft_hip_get_stream": (_i32, [_vp, _pvp]),
    "dtfft_hip_report": (_i32, [_vp]),
    "dtfft_hip_plan_execute_stage": (_i32, [_vp, _vp, _vp, _i32, _vp, _i32, _pi32]),
    "dtfft_hip_unpack_fft_create": (_i32, [_i32, _pi32, _pi32, _i32, _i32, _i32, _vp, _pvp]),
    "dtfft_hip_unpack_fft_execute": (_i32, [_vp, _vp, _vp, _i8]),
    "dtfft_hip_unpack_fft_destroy": (_i32, [_vp]),
    "dtfft_hip_cycle_create": (_i32, [_i32, _pi32, _i32, _i32, _vp, _pvp]),
    "dtfft_hip_cycle_get_layout": (_i32, [_vp, _pi64, _psz]),
    "dtfft_hip_cycle_execute": (_i32, [_vp, _vp, _vp, _vp, _i32]),
    "dtfft_hip_cycle_execute_graph": (_i32, [_vp, _vp, _vp, _vp, _vp]),
    "dtfft_hip_cycle_destroy": (_i32, [_pvp]),
    "dtfft_hip_memcpy_h2d": (_i32, [_vp, _vp, _sz]),
    "dtfft_hip_memcpy_d2h": (_i32, [_vp, _vp, _sz]),
    "dtfft_hip_memcpy_d2d_async": (_i32, [_vp, _vp, _sz, _vp]),
    "dtfft_hip_device_synchronize": (_i32, []),
    "dtfft_hip_stream_create": (_i32, [_pvp]),
    "dtfft_hip_stream_destroy": (_i32, [_vp]),
    "dtfft_hip_stream_synchronize": (_i32, [_vp]),
    "dtfft_hip_event_create": (_i32, [_pvp]),
    "dtfft_hip_event_destroy": (_i32, [_vp]),
    "dtfft_hip_event_record": (_i32, [_vp, _vp]),
    "dtfft_hip_event_elapsed_ms": (_i32, [_vp, _vp, C.POINTER(C.c_float)]),
    "dtfft_hip_get_device_arch": (_i32, [C.c_char_p, _i32]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m dtfft_amd.build` "
                "(dtfft_amd has no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def error_string(code: int) -> str:
    try:
        return lib().dtfft_hip_get_error_string(int(code)).decode()
    except Exception:  # pragma: no cover - only when the library itself failed to load
        return "?"


def check(code: int, where: str) -> None:
    if code != 0:
        detail = lib().dtfft_hip_get_last_error_message() or b""
        raise DtfftError(code, where, detail.decode(errors="replace"))


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args), name)
