"""Loads libsf_align.so (the HIP library) and declares its C ABI (include/sf_align.h).

There is no CPU fallback: if the library is missing or was built for another device, every
call fails with NativeError.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
ALIGN_LIB = os.path.join(HERE, "libsf_align.so")
WORKLOAD_LIB = os.path.join(HERE, "libsf_workload.so")
DIAG_LIB = os.path.join(HERE, "libsf_diag.so")

SF_OK, SF_EINVAL, SF_ENOMEM, SF_EDEVICE, SF_ENODEV, SF_ERANGE, SF_ETIMEDOUT, SF_ECANCELED = range(8)
ABI_VERSION = 2


class NativeError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"sf status {status}: {msg}")
        self.status = status


class Scoring(C.Structure):
    """sf_scoring: align.rs:660-691 `Affine` (mat, mis, open, ext)."""

    _fields_ = [("mat", C.c_int32), ("mis", C.c_int32), ("open", C.c_int32), ("ext", C.c_int32)]


class Pairs(C.Structure):
    _fields_ = [
        ("pat", C.c_void_p),
        ("pat_off", C.c_void_p),
        ("pat_len", C.c_void_p),
        ("txt", C.c_void_p),
        ("txt_off", C.c_void_p),
        ("txt_len", C.c_void_p),
        ("n", C.c_uint64),
        ("packed", C.c_int32),
        ("pat_bases", C.c_uint64),
        ("txt_bases", C.c_uint64),
    ]


class Dp(C.Structure):
    """sf_dp: align.rs:2111-2129 `Dp`."""

    _fields_ = [
        ("score", C.c_int32),
        ("score_col", C.c_uint32),
        ("clip_best", C.c_int32),
        ("clip_i", C.c_uint32),
        ("clip_j", C.c_uint32),
        ("text_read", C.c_uint32),
        ("last_row", C.c_uint32),
        ("cut_best", C.c_int32),
        ("cut_score", C.c_int32),
    ]


class ExtendWindow(C.Structure):
    _fields_ = [("score", C.c_int32), ("pat_bases", C.c_uint32), ("txt_bases", C.c_uint32), ("txt_read", C.c_uint32)]


class ExtendCut(C.Structure):
    _fields_ = [("whole", C.c_int32), ("cut", C.c_int32), ("cols", C.c_uint32)]


class Timing(C.Structure):
    _fields_ = [
        ("myers_ms", C.c_double),
        ("locate_ms", C.c_double),
        ("affine_ms", C.c_double),
        ("walk_ms", C.c_double),
        ("myers_cells", C.c_uint64),
        ("affine_cells", C.c_uint64),
        ("affine_launches", C.c_uint64),
        ("graph_ms", C.c_double),
        ("graph_replays", C.c_uint64),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_uint64),
        ("slots", C.c_uint32),
        ("max_reads_per_batch", C.c_uint32),
        ("myers_mode", C.c_int32),
        ("affine_mode", C.c_int32),
        ("sc", Scoring),
        ("band", C.c_uint64),
        ("batch_wait_us", C.c_uint32),
    ]


class ReadRequest(C.Structure):
    _fields_ = [
        ("read", C.c_void_p),
        ("read_len", C.c_uint32),
        ("nseg", C.c_uint32),
        ("seg_read_off", C.c_void_p),
        ("seg_read_len", C.c_void_p),
        ("seg_tx_off", C.c_void_p),
        ("seg_tx_len", C.c_void_p),
        ("dist", C.c_void_p),
        ("end", C.c_void_p),
        ("score", C.c_void_p),
        ("score_col", C.c_void_p),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("reads_done", C.c_uint64),
        ("reads_failed", C.c_uint64),
        ("batches", C.c_uint64),
        ("segments", C.c_uint64),
        ("cells", C.c_uint64),
        ("mean_reads_per_batch", C.c_double),
        ("p50_ms", C.c_double),
        ("p99_ms", C.c_double),
        ("max_ms", C.c_double),
        ("span_s", C.c_double),
        ("cap_segments_per_batch", C.c_uint64),
        ("cap_read_bases_per_batch", C.c_uint64),
        ("max_segments_per_batch", C.c_uint64),
    ]


class TbSegment(C.Structure):
    """sf_tb_segment: one read segment of the tie-break and its candidates."""
    _fields_ = [
        ("read_off", C.c_uint64),
        ("read_len", C.c_uint32),
        ("cand0", C.c_uint32),
        ("ncand", C.c_uint32),
        ("mode", C.c_int32),
        ("flags", C.c_uint32),
        ("window", C.c_uint64),
        ("cut", C.c_uint64),
    ]


class TbParams(C.Structure):
    _fields_ = [("sc", Scoring), ("band", C.c_uint64), ("end_bonus", C.c_int32), ("xdrop", C.c_int32)]


class TbTiming(C.Structure):
    _fields_ = [("ext_ms", C.c_double), ("sorted_ms", C.c_double), ("ext_cells", C.c_uint64),
                ("sorted_cells", C.c_uint64)]


SEG_REVERSE, SEG_EXTEND, SEG_DIAGONAL, SEG_TRACK_CUT, SEG_BONUS_WITHIN_CUT = 1, 2, 4, 8, 16


# Every symbol include/sf_align.h declares, with its ctypes signature.
_vp, _u32, _u64, _i32, _int = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32, C.c_int
SIGNATURES = {
    "sf_strerror": ([_int], C.c_char_p),
    "sf_last_error": ([], C.c_char_p),
    "sf_abi_version": ([], _int),
    "sf_init": ([_int], _int),
    "sf_myers": ([C.POINTER(Pairs), _int, _u32, _vp, _vp, _vp], _int),
    "sf_myers_noalloc": ([C.POINTER(Pairs), _int, _u32, _vp, _vp, _vp], _int),
    "sf_myers_banded": ([C.POINTER(Pairs), _vp, _u64, _int, _u32, _int, _vp], _int),
    "sf_affine": ([C.POINTER(Pairs), _int, C.POINTER(Scoring), _vp, _vp, _vp], _int),
    "sf_affine_cigar": ([C.POINTER(Pairs), _int, C.POINTER(Scoring), _vp, _vp, _vp, _vp, _vp, _u64, _vp], _int),
    "sf_batch_create": ([C.POINTER(Pairs), C.POINTER(_vp)], _int),
    "sf_batch_destroy": ([_vp], None),
    "sf_batch_size": ([_vp], _u64),
    "sf_batch_myers": ([_vp, _int, _u32, _int], _int),
    "sf_batch_affine": ([_vp, _int, C.POINTER(Scoring), _vp, _int], _int),
    "sf_batch_use_located": ([_vp, _int], _int),
    "sf_batch_sync": ([_vp], _int),
    "sf_batch_myers_results": ([_vp, _vp, _vp, _vp], _int),
    "sf_batch_affine_results": ([_vp, _vp, _vp, _vp, _vp, _u64, _vp], _int),
    "sf_batch_timing": ([_vp, C.POINTER(Timing)], _int),
    "sf_batch_reset_timing": ([_vp], None),
    "sf_batch_capture_begin": ([_vp], _int),
    "sf_batch_capture_end": ([_vp], _int),
    "sf_batch_replay": ([_vp, _u32], _int),
    "sf_set_persistent": ([_int], _int),
    "sf_batch_autotune": ([_vp, _int, C.POINTER(Scoring), _vp], _int),
    "sf_tune_save": ([C.c_char_p], _int),
    "sf_tune_load": ([C.c_char_p], _int),
    "sf_tune_reset": ([], _int),
    "sf_tune_classes": ([], _u64),
    "sf_server_create": ([_vp, _u64, C.POINTER(ServerConfig), C.POINTER(_vp)], _int),
    "sf_server_submit": ([_vp, C.POINTER(ReadRequest), C.POINTER(_u64)], _int),
    "sf_server_wait": ([_vp, _u64, C.c_double], _int),
    "sf_server_cancel": ([_vp, _u64], _int),
    "sf_server_get_stats": ([_vp, C.POINTER(ServerStats), _int], _int),
    "sf_server_destroy": ([_vp], None),
    "sf_server_replay": ([_vp, C.POINTER(ReadRequest), C.POINTER(C.c_double), _u64, C.POINTER(C.c_double),
                          C.POINTER(_u64)], _int),
    "sf_batch_affine_dp": ([_vp, _int, _vp, _vp, _vp, _int, C.c_int32, _int, _vp], _int),
    "sf_batch_dp_results": ([_vp, _vp], _int),
    "sf_batch_local_band": ([_vp, _vp, _vp, _vp, C.c_int32, C.c_int32], _int),
    "sf_batch_local_results": ([_vp, _vp], _int),
    "sf_affine_dp": ([_vp, _int, _vp, _vp, _vp, _int, C.c_int32, _int, _vp, _vp], _int),
    "sf_affine_extend": ([_vp, _vp, _vp, C.c_int32, C.c_int32, _vp], _int),
    "sf_affine_extend_window": ([_vp, _vp, _vp, _vp, C.c_int32, C.c_int32, _vp], _int),
    "sf_affine_extend_cut_bonus": ([_vp, _vp, _vp, C.c_int32, C.c_int32, _int, _vp, _vp, _int, _int, _vp], _int),
    "sf_local_band_ends": ([_vp, _vp, _vp, _vp, C.c_int32, C.c_int32, _vp], _int),
    "sf_store_create": ([_vp, _u64, C.POINTER(_vp)], _int),
    "sf_store_destroy": ([_vp], None),
    "sf_store_codes_eq": ([_vp, _vp, _vp, _vp, _u64, _vp], _int),
    "sf_tiebreak_segments": ([_vp, _vp, _u64, _vp, _u64, _vp, _vp, _u64, C.POINTER(TbParams), _vp, _vp], _int),
    "sf_tb_batch_create": ([_vp, _vp, _u64, _vp, _u64, _vp, _vp, _u64, C.POINTER(_vp)], _int),
    "sf_tb_batch_run": ([_vp, C.POINTER(TbParams)], _int),
    "sf_tb_batch_results": ([_vp, _vp, _vp], _int),
    "sf_tb_batch_destroy": ([_vp], None),
    "sf_tb_batch_timing": ([_vp, C.POINTER(TbTiming), _int], _int),
    "sf_set_affine_lp16_max": ([C.c_uint64], _int),
    "sf_set_affine_lp8": ([C.c_int], _int),
    "sf_set_affine_pass_packed": ([C.c_int], _int),
    "sf_set_myers_ukk": ([C.c_int], _int),
}

# Every symbol include/sf_diag.h declares (libsf_diag.so: diagnostics for tests and measurements).
DIAG_SIGNATURES = {
    "sf_diag_last_error": ([], C.c_char_p),
    "sf_selftest_lanes": ([_vp], _int),
    "sf_probe_issue": ([_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_float)], _int),
}

_lib = None


def load(path: str = ALIGN_LIB) -> C.CDLL:
    """The HIP library; raises NativeError(SF_ENODEV) when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise NativeError(SF_ENODEV, f"{path} is missing: run `python -m spearfish_amd.build` (no CPU fallback)")
    lib = C.CDLL(path)
    for name, (args, res) in SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    if lib.sf_abi_version() != ABI_VERSION:
        raise NativeError(SF_ENODEV, "ABI version mismatch")
    _lib = lib
    return lib


_diag = None


def load_diag(path: str = DIAG_LIB) -> C.CDLL:
    """The diagnostics library (lane-move self-test, VALU issue probe)."""
    global _diag
    if _diag is not None:
        return _diag
    if not os.path.exists(path):
        raise NativeError(SF_ENODEV, f"{path} is missing: run `python -m spearfish_amd.build`")
    lib = C.CDLL(path)
    for name, (args, res) in DIAG_SIGNATURES.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _diag = lib
    return lib


def check_diag(status: int) -> None:
    if status != SF_OK:
        raise NativeError(status, (load_diag().sf_diag_last_error() or b"").decode())


def check(status: int) -> None:
    if status != SF_OK:
        lib = load()
        raise NativeError(status, (lib.sf_last_error() or b"").decode() or lib.sf_strerror(status).decode())
