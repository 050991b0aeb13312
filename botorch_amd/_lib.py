"""ctypes binding of ``libbotorch_amd.so`` (declared in include/botorch_amd.h).

The library is loaded once, from this package directory.  There is no other
implementation: if the library is missing or no GPU is visible, the calls
raise.  Device buffers are passed as raw pointers of torch tensors; every call
is enqueued on torch's current HIP stream.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

from .exceptions import NanError, NotPSDError, UnsupportedError

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libbotorch_amd.so")

BTA_OK = 0
BTA_ERR_INVALID = 1
BTA_ERR_NOT_PSD = 2
BTA_ERR_NAN = 3
BTA_ERR_UNSUPPORTED = 4
BTA_ERR_HIP = 5
BTA_ERR_WORKSPACE = 6

KERNEL_MATERN52 = 0
KERNEL_RBF = 1
PREC_FP64 = 0
PREC_MIXED = 1
PREC_MIXED_F32 = 2

STATUS_JITTER = 1
STATUS_NOT_PSD = 2
STATUS_NAN = 4

GEMM_A_LOWER = 1
GEMM_A_UPPER = 2
GEMM_B_LOWER = 4
GEMM_B_UPPER = 8
GEMM_C_LOWER = 16
GEMM_B_KN = 32
GEMM_F32 = 64
GEMM_A_KM = 128
GEMM_ACC64 = 256

_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_f64 = ctypes.c_double
_sz = ctypes.c_size_t


class GpModel(ctypes.Structure):
    """bta_gp_model"""
    _fields_ = [
        ("kind", _i32), ("d", _i32), ("n", _i64), ("n_pad", _i64),
        ("train_xs", _p), ("center", _p), ("lengthscale", _p),
        ("outputscale", _f64), ("noise", _f64), ("mean_constant", _f64),
        ("y_mean", _f64), ("y_std", _f64), ("linv", _p), ("alpha", _p),
        ("chol_jitter", _f64), ("refine_steps", _i32),
    ]


class QLogNEIState(ctypes.Structure):
    """bta_qlognei_state"""
    _fields_ = [
        ("precision", _i32), ("q", _i32), ("n_b", _i32), ("S", _i32),
        ("xb_s", _p), ("linv", _p), ("ext", _p), ("Lb", _p), ("Zt", _p), ("best_f", _p),
        ("tau_relu", _f64), ("tau_max", _f64), ("max_tries", _i32),
        ("linv_scale", ctypes.c_float), ("jitter", _f64),
        ("Lb_levels", _p), ("lb_levels_mask", _i32),
    ]


class QLogNEIBatchItem(ctypes.Structure):
    """bta_qlognei_batch_item"""
    _fields_ = [
        ("model", ctypes.POINTER(GpModel)), ("state", ctypes.POINTER(QLogNEIState)),
        ("X", _p), ("b", _i64), ("value", _p), ("grad", _p), ("status", _p),
    ]


class QLogNEIBatchPlan(ctypes.Structure):
    """bta_qlognei_batch_plan"""
    _fields_ = [("ws", _p)] + [(f, _i32) for f in (
        "n_items", "q", "kind", "kd", "grad", "covar_blocks", "mc_blocks", "fin_blocks",
        "fwd_tiles", "adj_tiles", "fwd_workers", "adj_workers")] + [(f, _sz) for f in (
            "off_cov", "off_fwd", "off_adj", "off_mc", "off_grad", "off_fin", "off_cov_start",
            "off_mc_start", "off_fin_start", "off_fwd_tiles", "off_adj_tiles")]


# name -> (restype, argtypes); the exact list of extern "C" entry points.
SIGNATURES = {
    "bta_version": (ctypes.c_char_p, []),
    "bta_abi_version": (ctypes.c_int, []),
    "bta_last_error": (ctypes.c_char_p, []),
    "bta_scale_inputs": (ctypes.c_int, [_p, _i64, _i32, _p, _p, _p, _p, _i64, _p]),
    "bta_covar": (ctypes.c_int, [_i32, _p, _i64, _p, _i64, _i32, _f64, _f64, _i32, _p, _i64,
                                 _i64, _i64, _p]),
    "bta_covar_grad": (ctypes.c_int, [_i32, _p, _i64, _p, _i64, _i32, _f64, _p, _i32, _i64, _p,
                                      _p, _p, _p]),
    "bta_gemm": (ctypes.c_int, [_i32, _i64, _i64, _i64, _f64, _p, _i64, _p, _i64, _f64, _p,
                                _i64, _p, _i64, _i64, _i64, _i64, _i64, _i64, _p]),
    "bta_split_f16": (ctypes.c_int, [_i32, _p, _i64, _i64, _i64, _i64, _i64, ctypes.c_float,
                                     _i32, _p, _p, _i64, _p]),
    "bta_gemm_f16x3": (ctypes.c_int, [_i32, _i64, _i64, _i64, _f64, _p, _p, _i64, _p,
                                      ctypes.c_float, _p, _p, _i64, ctypes.c_float, _p, _i64,
                                      _p, _p, ctypes.c_float, _i32, _p]),
    "bta_potrf_workspace_size": (_sz, [_i64]),
    "bta_potrf": (ctypes.c_int, [_p, _i64, _i64, _p, _i64, _i32, _f64,
                                 ctypes.POINTER(_f64), _p, _sz, _p]),
    "bta_potrf_inverse": (ctypes.c_int, [_p, _i64, _i64, _p, _i64, _i32, _f64,
                                         ctypes.POINTER(_f64), _p, _sz, _p]),
    "bta_trtri_lower": (ctypes.c_int, [_p, _i64, _p, _i64, _i64, _p, _sz, _p]),
    "bta_chol_small_batched": (ctypes.c_int, [_p, _i32, _i64, _i64, _i64, _i32, _f64, _p, _p]),
    "bta_gp_cache_workspace_size": (_sz, [_i64]),
    "bta_gp_cache_build": (ctypes.c_int, [ctypes.POINTER(GpModel), _p, _p, _p, _p, _i32,
                                          ctypes.POINTER(_f64), _p, _sz, _p]),
    "bta_gp_posterior_workspace_size": (_sz, [ctypes.POINTER(GpModel), _i64, _i32]),
    "bta_gp_posterior": (ctypes.c_int, [ctypes.POINTER(GpModel), _p, _i64, _i32, _p, _p, _p,
                                        _p, _sz, _p]),
    "bta_gp_posterior_backward_workspace_size": (_sz, [ctypes.POINTER(GpModel), _i64, _i32]),
    "bta_gp_posterior_backward": (ctypes.c_int, [ctypes.POINTER(GpModel), _p, _i64, _i32, _p,
                                                 _p, _p, _p, _sz, _p]),
    "bta_chol_small_backward": (ctypes.c_int, [_p, _p, _i32, _i64, _i64, _p, _p]),
    "bta_qlognei_prepare_workspace_size": (_sz, [ctypes.POINTER(GpModel), _i32, _i32]),
    "bta_qlognei_prepare": (ctypes.c_int, [ctypes.POINTER(GpModel),
                                           ctypes.POINTER(QLogNEIState), _p, _p, _p, _p, _p,
                                           ctypes.POINTER(_f64), ctypes.POINTER(ctypes.c_float),
                                           _p, _sz, _p]),
    "bta_qlognei_workspace_size": (_sz, [ctypes.POINTER(GpModel), ctypes.POINTER(QLogNEIState),
                                         _i64]),
    "bta_qlognei_autotune": (ctypes.c_int, [ctypes.POINTER(GpModel),
                                            ctypes.POINTER(QLogNEIState), _p, _i64, _p, _p, _p,
                                            _p, _sz, ctypes.POINTER(_i32), _p]),
    "bta_qlognei_value_grad": (ctypes.c_int, [ctypes.POINTER(GpModel),
                                              ctypes.POINTER(QLogNEIState), _p, _i64, _p, _p,
                                              _p, _p, _sz, _p]),
    "bta_qlognei_batch_workspace_size": (_sz, [ctypes.POINTER(QLogNEIBatchItem), _i32]),
    "bta_qlognei_batch_prepare": (ctypes.c_int, [ctypes.POINTER(QLogNEIBatchItem), _i32, _p,
                                                 _sz, ctypes.POINTER(QLogNEIBatchPlan), _p]),
    "bta_qlognei_batch_value_grad": (ctypes.c_int, [ctypes.POINTER(QLogNEIBatchPlan), _p]),
    "bta_qlognei_value_grad_noalloc": (ctypes.c_int, [ctypes.POINTER(GpModel),
                                                      ctypes.POINTER(QLogNEIState), _p, _i64,
                                                      _p, _p, _p, _p]),
    "bta_log_areas_forward": (ctypes.c_int, [_i32, _p, _i64, _i64, _i32, _i32, _p, _p, _i64,
                                             _i64, _f64, _f64, _p, _p]),
    "bta_log_areas_backward": (ctypes.c_int, [_i32, _p, _p, _i64, _i64, _i32, _i32, _p, _p,
                                              _i64, _i64, _f64, _f64, _p, _p]),
    "bta_lbfgsb_state_size": (_sz, [_i32, _i32]),
    "bta_lbfgsb_init": (ctypes.c_int, [_p, _i64, _i32, _i32, _p, _p, _p, _p, _f64, _f64, _i32,
                                       _i32, _i32, _p, _p, _p]),
    "bta_lbfgsb_step": (ctypes.c_int, [_p, _i64, _i32, _i32, _p, _p, _p, _p, _p, _f64, _p, _p,
                                       _p, _p]),
    "bta_lbfgsb_results": (ctypes.c_int, [_p, _i64, _i32, _i32, _p, _p, _p, _p]),
    "bta_mll_workspace_size": (_sz, [_i64, _i32]),
    "bta_mll_value_grad": (ctypes.c_int, [_i32, _p, _p, _i64, _i32, ctypes.POINTER(_f64), _f64,
                                          _f64, _f64, _i32, ctypes.POINTER(_f64), _p, _sz, _p]),
}

ABI_VERSION = 5  # BTA_ABI_VERSION of include/botorch_amd.h

_lock = threading.Lock()
_lib = None


def load(path: str = LIB_PATH) -> ctypes.CDLL:
    """Loads the library and declares every signature (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(path):
                raise RuntimeError(
                    f"{path} is missing: build it with `python -m botorch_amd.build`.")
            lib = ctypes.CDLL(path)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            if lib.bta_abi_version() != ABI_VERSION:
                raise RuntimeError(f"{path} has C ABI {lib.bta_abi_version()}, this binding "
                                   f"expects {ABI_VERSION}: rebuild the library.")
            _lib = lib
    return _lib


def lib() -> ctypes.CDLL:
    """The library, after checking that a HIP device is present."""
    L = load()
    if not torch.cuda.is_available():
        raise RuntimeError("botorch_amd needs a HIP GPU (torch.cuda.is_available() is False).")
    return L


def check(rc: int, what: str = "") -> None:
    if rc == BTA_OK:
        return
    msg = load().bta_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == BTA_ERR_NOT_PSD:
        raise NotPSDError(text)
    if rc == BTA_ERR_NAN:
        raise NanError(text)
    if rc == BTA_ERR_UNSUPPORTED:
        raise UnsupportedError(text)
    raise RuntimeError(f"botorch_amd error {rc}: {text}")


def stream() -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t: torch.Tensor | None) -> ctypes.c_void_p | None:
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("botorch_amd kernels need device tensors")
    if not t.is_contiguous():
        raise ValueError("botorch_amd kernels need contiguous tensors")
    return ctypes.c_void_p(t.data_ptr())


def workspace(nbytes: int, device) -> torch.Tensor:
    """A caller-owned scratch buffer from torch's caching allocator."""
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)


def round_up(x: int, m: int) -> int:
    return (x + m - 1) // m * m
