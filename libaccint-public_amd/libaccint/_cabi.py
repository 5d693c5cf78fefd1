"""ctypes binding of the C ABI in include/lai_device.h.

Used by the Python layer for raw device pointers (the torch.autograd path)
and by tests that call the boundary directly.  Loading fails loudly when the
HIP library is not built: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblai_device.so")

LAI_OK = 0
LAI_JK_FP64 = 0
LAI_JK_MIXED = 1
LAI_JK_ANTISYMMETRIC = 2
LAI_JK_PERSISTENT = 4
LAI_NOALLOC_MAX_QUARTETS = 65536

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)



class EriBlock(ctypes.Structure):
    """lai_eri_block: one ERI block of a Fock digestion."""
    _fields_ = [("off", ctypes.c_int64), ("fa", ctypes.c_int32), ("fb", ctypes.c_int32), ("fc", ctypes.c_int32),
                ("fd", ctypes.c_int32), ("na", ctypes.c_int32), ("nb", ctypes.c_int32), ("nc", ctypes.c_int32),
                ("nd", ctypes.c_int32)]


class ClassTune(ctypes.Structure):
    """lai_class_tune: one dispatch-table entry."""
    _fields_ = [("cls", ctypes.c_int), ("strategy", ctypes.c_int), ("chunk", ctypes.c_int),
                ("blocks_per_cu", ctypes.c_int), ("ms", ctypes.c_double), ("ms_default", ctypes.c_double)]


_tp = ctypes.POINTER(ClassTune)

_SIGNATURES = {
    "lai_dev_init": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int),
    "lai_dev_fini": ([_vp], ctypes.c_int),
    "lai_dev_last_error": ([], ctypes.c_char_p),
    "lai_dev_device_count": ([_ip], ctypes.c_int),
    "lai_dev_malloc": ([_vp, ctypes.c_size_t, ctypes.POINTER(_vp)], ctypes.c_int),
    "lai_dev_free": ([_vp, _vp], ctypes.c_int),
    "lai_dev_memcpy_h2d": ([_vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "lai_dev_memcpy_d2h": ([_vp, _vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "lai_dev_zero": ([_vp, ctypes.c_size_t, _vp], ctypes.c_int),
    "lai_dev_stream_sync": ([_vp], ctypes.c_int),
    "lai_dev_stream_wait": ([_vp, _vp], ctypes.c_int),
    "lai_dev_upload_basis": ([_vp, ctypes.c_int, _ip, _dp, _ip, _dp, _dp, ctypes.POINTER(_vp)], ctypes.c_int),
    "lai_dev_basis_free": ([_vp], ctypes.c_int),
    "lai_dev_basis_nbf": ([_vp], ctypes.c_int),
    "lai_dev_upload_batch": ([_vp, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _vp], ctypes.c_int),
    "lai_dev_batch_layout": ([_vp, _vp, _vp, _vp], ctypes.c_int),
    "lai_dev_eri_quartets": ([_vp, _i64, _ip, _dp, ctypes.POINTER(_i64)], ctypes.c_int),
    "lai_dev_eri_quartets_device": ([_vp, _i64, _ip, _vp, ctypes.POINTER(_i64), _vp], ctypes.c_int),
    "lai_dev_eri_quartets_device_noalloc": ([_vp, _i64, _ip, _vp, ctypes.POINTER(_i64), _vp], ctypes.c_int),
    "lai_dev_fock_accumulate": ([_vp, _vp, _i64, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "lai_dev_schwarz": ([_vp, _dp], ctypes.c_int),
    "lai_dev_set_schwarz": ([_vp, _dp], ctypes.c_int),
    "lai_dev_build_worklist": ([_vp, ctypes.c_double, _dp, ctypes.c_int, ctypes.c_double,
                                ctypes.POINTER(_i64), ctypes.POINTER(_i64),
                                ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
    "lai_quartet_mix64": ([ctypes.c_int] * 4, ctypes.c_uint64),
    "lai_dev_worklist_stats": ([_vp, ctypes.POINTER(_i64), _dp, _dp, ctypes.POINTER(_i64)], ctypes.c_int),
    "lai_dev_worklist_class_flops": ([_vp, _dp, _dp], ctypes.c_int),
    "lai_dev_worklist_grad_flops": ([_vp, _dp, _dp], ctypes.c_int),
    "lai_dev_jk": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "lai_dev_jk_profile": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _dp, _dp], ctypes.c_int),
    "lai_dev_jk_graph_capture": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "lai_dev_jk_graph_launch": ([_vp, _vp], ctypes.c_int),
    "lai_dev_set_mixed_prim_cut": ([_vp, ctypes.c_double], ctypes.c_int),
    "lai_dev_class_strategies": ([ctypes.c_int, _ip, _ip], ctypes.c_int),
    "lai_dev_autotune": ([_vp, _vp, ctypes.c_int, ctypes.c_int, _tp, ctypes.c_int, _ip, _vp], ctypes.c_int),
    "lai_dev_get_tuning": ([_vp, _tp, ctypes.c_int, _ip], ctypes.c_int),
    "lai_dev_set_tuning": ([_vp, _tp, ctypes.c_int], ctypes.c_int),
    "lai_dev_set_shell_atoms": ([_vp, _ip, ctypes.c_int], ctypes.c_int),
    "lai_dev_jk_grad": ([_vp, _vp, ctypes.c_double, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "lai_dev_onee": ([_vp, ctypes.c_int, ctypes.c_int, _dp, _dp, _vp, _vp], ctypes.c_int),
    "lai_dev_df_create": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp,
                           ctypes.POINTER(_vp)], ctypes.c_int),
    "lai_dev_df_free": ([_vp], ctypes.c_int),
    "lai_dev_df_dims": ([_vp, _ip, _ip], ctypes.c_int),
    "lai_dev_df_jk": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp], ctypes.c_int),
    "lai_dev_df_jk_profile": ([_vp, _vp, _vp, _vp, ctypes.c_int, _vp, _dp], ctypes.c_int),
    "lai_dev_rys": ([ctypes.c_int, ctypes.c_int, _vp, _vp, _vp, _vp], ctypes.c_int),
    "lai_dev_boys": ([ctypes.c_int, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
}

_lib = None


class DeviceError(RuntimeError):
    pass


def lib():
    """The loaded C-ABI library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise DeviceError("liblai_device.so is not built (run libaccint-public_amd/build.py)")
        l = ctypes.CDLL(_LIB_PATH)
        for name, (args, res) in _SIGNATURES.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = res
        _lib = l
    return _lib


def check(rc):
    if rc != LAI_OK:
        msg = lib().lai_dev_last_error()
        raise DeviceError("lai_dev error %d: %s" % (rc, msg.decode() if msg else ""))
    return rc


def declared_symbols():
    return list(_SIGNATURES)
