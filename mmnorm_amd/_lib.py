"""ctypes binding of libmmnorm_hip.so (the C ABI declared in include/mmnorm_amd.h).

The library is built in-tree by ``mmnorm_amd.build`` (hipcc, gfx950).  It is
loaded lazily, after ``torch`` so that both share the one HIP runtime that
torch has already mapped (the library's DT_NEEDED libamdhip64.so.7 resolves
to torch's copy by soname).  There is no fallback: if the library is missing
every entry point raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (load torch's HIP runtime first)

LIB_NAME = "libmmnorm_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

MMN_BP_PLAIN = 1
MMN_BP_REC = 2
MMN_BP_PLAIN_FAST = 3

_c_int = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_f = ctypes.c_float
_c_d = ctypes.c_double
_p = ctypes.c_void_p
_sz = ctypes.c_size_t

# name -> (restype, argtypes); kept in the order of include/mmnorm_amd.h
SIGNATURES = {
    "mmn_abi_version": (_c_int, []),
    "mmn_last_error": (ctypes.c_char_p, []),
    "mmn_cuda_image": (_c_int, [_p, _p, _p, _p, _p, _p, _p, _c_f, _c_f, _c_f,
                                _c_int, _c_int, _c_int, _c_int, _c_int, _c_int,
                                _c_int, _c_int, _p]),
    "mmn_cuda_sn_est": (_c_int, [_p, _p, _c_i64, _p, _p, _p, _p, _p, _p, _p,
                                 _c_f, _c_f, _c_f, _c_int, _c_int, _c_int,
                                 _c_int, _c_int, _c_int, _c_int, _p]),
    "mmn_bp_forward_workspace_size": (_sz, [_c_i64, _c_int, _c_int]),
    "mmn_bp_forward": (_c_int, [_p, _p, _p, _p, _p, _p, _c_i64, _p, _p, _p,
                                _c_d, _c_d, _c_int, _c_int, _c_int, _c_int,
                                _c_int, _c_int, _c_int, _c_int, _c_int, _p, _sz,
                                _p]),
    "mmn_bp_profile_workspace_size": (_sz, [_c_i64, _c_int, _c_int, _c_d, _c_d, _p,
                                             _c_int]),
    "mmn_bp_profile": (_c_int, [_p, _p, _p, _p, _p, _p, _c_i64, _p, _p, _p, _c_d, _c_d,
                                _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _p,
                                _p, _sz, _p]),
    "mmn_bp_profile_backward_workspace_size": (_sz, [_c_i64, _c_int, _c_int, _c_d, _c_d,
                                                      _p]),
    "mmn_bp_profile_backward": (_c_int, [_p, _p, _p, _c_d, _p, _p, _p, _p, _c_i64, _p, _p,
                                         _p, _c_d, _c_d, _c_int, _c_int, _c_int, _c_int,
                                         _c_int, _c_int, _c_int, _p, _p, _sz, _p]),
    "mmn_sqrt_rn_check": (_c_int, [ctypes.c_uint64, ctypes.c_uint64, _p, _p]),
    "mmn_bp_interp": (_c_int, [_p, _p, _p, _p, _p, _p, _c_i64, _p, _p, _p, _c_f, _c_f,
                               _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _c_int, _p]),
    "mmn_sn_from_moment": (_c_int, [_p, _p, _p, _p, _c_i64, _p]),
    "mmn_bp_backward_workspace_size": (_sz, [_c_int, _c_int, _c_i64]),
    "mmn_bp_backward": (_c_int, [_p, _p, _p, _c_d, _p, _p, _p, _p, _c_i64, _p,
                                 _p, _p, _c_d, _c_d, _c_int, _c_int, _c_int,
                                 _c_int, _c_int, _c_int, _c_int, _c_int, _p,
                                 _sz, _p]),
    "mmn_rsdf_bfs_workspace_size": (_sz, [_c_int, _c_int, _c_int]),
    "mmn_rsdf_bfs": (_c_int, [_p, _p, _p, _p, _p, _c_int, _c_d, _c_int, _c_int,
                              _c_int, _p, _sz, _p]),
    "mmn_rsdf_bfs_noalloc": (_c_int, [_p, _p, _p, _p, _p, _c_int, _c_d, _c_int, _c_int,
                                      _c_int, _p]),
    "mmn_li_threshold_workspace_size": (_sz, [_c_i64]),
    "mmn_li_threshold": (_c_int, [_p, _p, _p, _p, _p, _c_i64, _c_d, _c_int, _p, _sz,
                                  _p]),
    "mmn_ccl6_workspace_size": (_sz, [_c_int, _c_int, _c_int]),
    "mmn_ccl6": (_c_int, [_p, _p, _p, _c_int, _c_int, _c_int, _p, _sz, _p]),
    "mmn_label_sizes": (_c_int, [_p, _p, _c_i64, _c_int, _p]),
    "mmn_ray_matching": (_c_int, [_p, _p, _p, _p, _c_int, _c_int, _c_f, _c_f, _p, _p, _p,
                                  _p, _p]),
    "mmn_simulate_channel_workspace_size": (_sz, [_c_int, _c_int, _c_int]),
    "mmn_simulate_channel": (_c_int, [_p, _p, _c_f, _c_int, _c_int, _c_f, _c_f, _p, _p, _p,
                                      _p, _p, _c_f, _c_d, _c_d, _c_int, _c_int, _p, _sz,
                                      _p]),
    "mmn_simulate_channels_workspace_size": (_sz, [_c_int, _c_int, _c_int, _c_int]),
    "mmn_simulate_channels": (_c_int, [_p, _p, _c_int, _p, _p, _p, _p, _p, _c_int, _c_int,
                                       _c_int, _c_f, _c_f, _p, _p, _c_f, _c_d, _c_d, _c_int,
                                       _c_int, _p, _sz, _p]),
    "mmn_ray_prefix_size": (_sz, [_c_int, _c_int, _c_int]),
    "mmn_ray_prefix": (_c_int, [_p, _sz, _p, _p, _c_int, _p, _p, _p, _c_int, _c_int, _c_int,
                                _c_f, _c_f, _p, _p, _c_f, _p]),
    "mmn_simulate_channels_prefixed_workspace_size": (_sz, [_c_int, _c_int, _c_int, _c_int,
                                                            _c_int]),
    "mmn_simulate_channels_prefixed": (_c_int, [_p, _p, _sz, _p, _p, _c_int, _p, _p, _p, _p,
                                                _p, _c_int, _c_int, _c_int, _c_f, _c_f, _p,
                                                _p, _c_f, _c_d, _c_d, _c_int, _c_int, _c_f,
                                                _p, _sz, _p]),
    "mmn_channel_loss": (_c_int, [_p, _p, _p, _c_int, _c_int, _c_int, _c_int, _c_int,
                                  _c_int, _p]),
}

_lock = threading.Lock()
_lib = None


class MmnError(RuntimeError):
    """A non-zero status from libmmnorm_hip.so."""


def load() -> ctypes.CDLL:
    """Load the library once and attach the C signatures."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_NAME} is not built ({LIB_PATH} missing); run "
                "`python -m mmnorm_amd.build` (hipcc, gfx950) first")
        lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        if lib.mmn_abi_version() != 1:
            raise RuntimeError("libmmnorm_hip.so ABI version mismatch")
        _lib = lib
        return lib


def call(name: str, *args) -> None:
    """Call a status-returning entry point; raise on a non-zero status."""
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        msg = lib.mmn_last_error().decode(errors="replace")
        if rc < 0:
            raise ValueError(f"{name}: {msg} (status {rc})")
        raise MmnError(f"{name}: {msg} (hipError {rc})")


def require_gpu() -> None:
    if not torch.cuda.is_available():
        raise RuntimeError("mmnorm_amd needs a ROCm GPU (torch.cuda.is_available() is False)")


def ptr(t) -> int | None:
    """Device address of a tensor (None for None)."""
    return None if t is None else t.data_ptr()


def stream_handle(device=None) -> int:
    return torch.cuda.current_stream(device).cuda_stream
