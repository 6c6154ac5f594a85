"""ctypes binding of libqtxhip.so (the C ABI declared in include/qtx_hip.h).

torch is imported first so the HIP runtime it ships is the one the library
binds to (both carry the soname libamdhip64.so.7).  There is no fallback: if
the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import torch  # noqa: F401  (loads the HIP runtime before the library)

from .build import LIB

QTX_OK, QTX_EINVAL, QTX_EHIP, QTX_ESOLVER, QTX_EWORKSPACE = 0, -1, -2, -3, -4
QTX_F32, QTX_F64 = 0, 1
QTX_LOCALFLIP, QTX_EXCHANGE = 0, 1
QTX_GRAM_FP64, QTX_GRAM_FP32, QTX_GRAM_BF16X3 = 0, 1, 2
QTX_SOLVE_CHOL_SHIFT, QTX_SOLVE_EIGH_PINV, QTX_SOLVE_EIGH_PINV_DIRECT = 0, 1, 2
QTX_SOLVE_MINNORM, QTX_SOLVE_LSTSQ, QTX_SOLVE_SNR, QTX_SOLVE_BLAS_WS = 0x100, 0x200, 0x400, 0x800
QTX_SOLVE_CHEB = 0x1000
QTX_CONV_FP32, QTX_CONV_BF16X3 = 0, 1
QTX_ACT_GELU_TANH, QTX_ACT_GELU_ERF, QTX_ACT_RELU = 0, 1, 2
TAG_SWEEP, TAG_INIT = 0x51A7, 0x1D17

vp, i32, i64, u32, u64, f64, sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double, C.c_size_t


class QtxHam(C.Structure):
    _fields_ = [
        ("N", i32), ("ndiag", i32), ("diag_c", vp), ("diag_a", vp), ("diag_b", vp), ("diag_kind", vp),
        ("noff", i32), ("off_a", vp), ("off_b", vp), ("off_ra", vp), ("off_rb", vp), ("off_c", vp),
    ]


class QtxRbmJob(C.Structure):
    _fields_ = [
        ("packed", vp), ("M", i32), ("N", i32), ("kind", i32), ("width", i32), ("nbr", vp), ("hopping", i32),
        ("nsteps", i32), ("seed", u64), ("step0", i64), ("ns", i64), ("s", vp), ("logabs", vp), ("n_accept", vp),
        ("n_near", vp), ("reweight", f64), ("tie_tol", f64), ("H", QtxHam), ("eloc", vp), ("nconn", vp), ("O", vp),
    ]


class QtxMinsrTask(C.Structure):
    _fields_ = [
        ("stream", vp), ("o_dtype", i32), ("mode", i32), ("gram", i32), ("status", i32), ("O", vp), ("mean", vp),
        ("r", vp), ("ebar", vp), ("ns", i64), ("np", i64), ("rshift", f64), ("ashift", f64), ("rtol", f64),
        ("atol", f64), ("step", vp), ("ws", vp), ("ws_bytes", sz), ("path", i32), ("fallback", i32),
        ("tol_snr", f64),
    ]


class QtxResConvDesc(C.Structure):
    _fields_ = [("nblocks", i32), ("channels", i32), ("L0", i32), ("L1", i32), ("precision", i32),
                ("activation", i32)]


# name -> (restype, argtypes)
PROTOS = {
    "qtx_last_error": (C.c_char_p, []),
    "qtx_device_sync": (C.c_int, []),
    "qtx_philox_fill": (C.c_int, [vp, u64, u32, i64, i64, i64, i64, vp]),
    "qtx_init_spins": (C.c_int, [vp, u64, i64, i32, i32, vp]),
    "qtx_ham_diag": (C.c_int, [vp, C.POINTER(QtxHam), vp, i64, vp]),
    "qtx_conn_count": (C.c_int, [vp, C.POINTER(QtxHam), vp, i64, vp]),
    "qtx_conn_list": (C.c_int, [vp, C.POINTER(QtxHam), vp, i64, vp, vp, vp, vp, vp]),
    "qtx_rbm_pack": (C.c_int, [vp, C.c_int, vp, i32, i32, vp]),
    "qtx_rbm_logpsi": (C.c_int, [vp, C.c_int, vp, i32, i32, vp, i64, vp, vp]),
    "qtx_rbm_metropolis": (C.c_int, [vp, C.c_int, vp, i32, i32, C.c_int, vp, i32, i32, vp, i32, i64, f64, f64,
                                     vp, vp, vp, vp]),
    "qtx_rbm_eloc": (C.c_int, [vp, C.c_int, vp, i32, i32, C.POINTER(QtxHam), vp, i64, vp, vp]),
    "qtx_rbm_jacobian": (C.c_int, [vp, C.c_int, vp, i32, i32, vp, i64, C.c_int, vp]),
    "qtx_rbm_group_ws_bytes": (sz, [i32]),
    "qtx_rbm_group_logpsi": (C.c_int, [vp, C.POINTER(QtxRbmJob), i32, vp, sz]),
    "qtx_rbm_group_metropolis": (C.c_int, [vp, C.POINTER(QtxRbmJob), i32, vp, sz]),
    "qtx_rbm_group_eloc": (C.c_int, [vp, C.POINTER(QtxRbmJob), i32, vp, sz]),
    "qtx_rbm_group_jacobian": (C.c_int, [vp, C.POINTER(QtxRbmJob), i32, vp, sz]),
    "qtx_resconv_nparams": (i64, [C.POINTER(QtxResConvDesc)]),
    "qtx_resconv_packed_bytes": (sz, [C.POINTER(QtxResConvDesc)]),
    "qtx_resconv_pack": (C.c_int, [vp, C.POINTER(QtxResConvDesc), vp, vp]),
    "qtx_resconv_logpsi": (C.c_int, [vp, C.POINTER(QtxResConvDesc), vp, vp, i64, vp]),
    "qtx_resconv_metropolis": (C.c_int, [vp, C.POINTER(QtxResConvDesc), vp, C.c_int, vp, i32, i32, vp, i32, i64,
                                         f64, f64, vp, vp, vp, vp]),
    "qtx_resconv_eloc": (C.c_int, [vp, C.POINTER(QtxResConvDesc), vp, C.POINTER(QtxHam), vp, vp, i64, vp, vp]),
    "qtx_resconv_jacobian_ws_bytes": (sz, [C.POINTER(QtxResConvDesc), i64]),
    "qtx_resconv_jacobian": (C.c_int, [vp, C.POINTER(QtxResConvDesc), vp, vp, i64, vp, vp, sz]),
    "qtx_resconv_jacobian_noalloc": (C.c_int, [vp, C.POINTER(QtxResConvDesc), vp, vp, i64, vp]),
    "qtx_colmean": (C.c_int, [vp, C.c_int, vp, i64, i64, vp]),
    "qtx_ebar": (C.c_int, [vp, vp, vp, i64, vp, vp]),
    "qtx_obar_matvec": (C.c_int, [vp, C.c_int, vp, vp, vp, i64, i64, C.c_double, vp, C.c_double, vp]),
    "qtx_scale_columns": (C.c_int, [vp, C.c_int, vp, vp, i64, i64, vp]),
    "qtx_symm_images": (C.c_int, [vp, vp, i64, C.c_int32, vp, C.c_int32, C.c_int32, vp]),
    "qtx_symmetrize": (C.c_int, [vp, vp, vp, vp, i64, C.c_int32, vp, vp, vp]),
    "qtx_rows_combine": (C.c_int, [vp, C.c_int, vp, vp, i64, C.c_int32, i64, vp]),
    "qtx_propose": (C.c_int, [vp, C.c_int, vp, C.c_int32, C.c_int32, vp, vp, i64, C.c_int32, vp, vp]),
    "qtx_accept": (C.c_int, [vp, vp, vp, vp, vp, vp, i64, C.c_int32, C.c_double, C.c_double, vp, vp, vp, vp, vp]),
    "qtx_conn_combine": (C.c_int, [vp, vp, vp, vp, vp, vp, vp, vp, i64, vp]),
    "qtx_minsr_ws_bytes": (sz, [i64, i64, C.c_int, C.c_int]),
    "qtx_minsr_solve": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, i64, i64, C.c_int, C.c_int, f64, f64, f64, f64, vp,
                                  vp, sz]),
    "qtx_minsr_blas_ws_bytes": (C.c_int, [vp, i64, i64, C.c_int, C.c_int, vp, sz, C.POINTER(C.c_size_t)]),
    "qtx_minsr_solve_many": (C.c_int, [C.POINTER(QtxMinsrTask), i32]),
    "qtx_minsr_solve_snr": (C.c_int, [vp, C.c_int, vp, vp, vp, vp, i64, i64, C.c_int, C.c_int, f64, f64, f64, vp,
                                      vp, sz]),
    "qtx_pinv_eig_ws_bytes": (sz, [i64]),
    "qtx_pinv_eig": (C.c_int, [vp, vp, i64, vp, f64, f64, f64, vp, vp, sz]),
    "qtx_gram_ws_bytes": (sz, [i64, i64, C.c_int]),
    "qtx_gram": (C.c_int, [vp, C.c_int, vp, vp, vp, i64, i64, C.c_int, vp, vp, sz]),
    "qtx_minsr_last_path": (C.c_int, [C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(f64), C.POINTER(C.c_int)]),
    "qtx_apply_update": (C.c_int, [vp, C.c_int, vp, vp, f64, i64, vp]),
}

_lib = None


class QtxError(RuntimeError):
    pass


def lib():
    """Load the library once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise QtxError(f"{LIB} not found: build it with `python -m quantax_amd.build`")
        h = C.CDLL(LIB)
        for name, (res, args) in PROTOS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def call(name, *args):
    """Call an int-returning entry point and raise on a non-zero status."""
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        msg = lib().qtx_last_error().decode(errors="replace")
        raise QtxError(f"{name} failed ({rc}): {msg}")
    return rc


def ptr(t):
    """Device (or host) address of a torch tensor, None for None."""
    if t is None:
        return None
    return t.data_ptr()


def stream():
    """The current torch HIP stream as a raw handle."""
    return torch.cuda.current_stream().cuda_stream


def device():
    if not torch.cuda.is_available():
        raise QtxError("no HIP device visible: quantax_amd runs only on the GPU")
    return torch.device("cuda", torch.cuda.current_device())
