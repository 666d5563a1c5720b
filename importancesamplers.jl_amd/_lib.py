"""ctypes binding of libissampler.so (include/issampler.h).

The library is required: importing this module on a machine without the
built library raises, and every entry point that needs a GPU fails loudly
when none is present.  There is no CPU fallback.
"""

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libissampler.so")

IS_OK = 0
IS_EINVAL = -1
IS_ENOMEM = -2
IS_EBUSY = -3
IS_EDEVICE = -4
IS_EUNSUPPORTED = -5
IS_ESTATE = -6
IS_EFAILED = 1

F64P = ctypes.POINTER(ctypes.c_double)
I64P = ctypes.POINTER(ctypes.c_int64)
U64P = ctypes.POINTER(ctypes.c_uint64)
I32P = ctypes.POINTER(ctypes.c_int32)
VP = ctypes.c_void_p


# is_batch_target_fn: (user, samples, coord_stride, dim, count, values, is_float, stream) -> int
BATCH_TARGET_FN = ctypes.CFUNCTYPE(ctypes.c_int, VP, VP, ctypes.c_int64, ctypes.c_int32, ctypes.c_int64, VP,
                                   ctypes.c_int32, VP)


class TransformBlock(ctypes.Structure):
    _fields_ = [("first", ctypes.c_int32), ("count", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("scalar", ctypes.c_int32), ("lower", ctypes.c_double), ("upper", ctypes.c_double)]


class JobDesc(ctypes.Structure):
    _fields_ = [
        ("method", ctypes.c_int32),
        ("family", ctypes.c_int32),
        ("scale_kind", ctypes.c_int32),
        ("scalar", ctypes.c_int32),
        ("dim", ctypes.c_int32),
        ("dof", ctypes.c_double),
        ("rounds", ctypes.c_int32),
        ("round_size", I64P),
        ("location", F64P),
        ("scale", F64P),
        ("bank_size", ctypes.c_int32),
        ("bank_locations", F64P),
        ("bank_scales", F64P),
        ("bank_masses", F64P),
        ("resampling", ctypes.c_int32),
        ("target_kind", ctypes.c_int32),
        ("target_params", F64P),
        ("target_params_len", ctypes.c_int64),
        ("precision", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
        ("stream", ctypes.c_uint64),
        ("mis_scheme", ctypes.c_int32),
        ("mis_group_count", ctypes.c_int32),
        ("mis_group_offsets", I32P),
        ("mis_group_members", I32P),
        ("transform_kinds", I32P),
        ("transform_bounds", F64P),
        ("target_batch", BATCH_TARGET_FN),
        ("target_user", VP),
        ("transform_block_count", ctypes.c_int32),
        ("transform_blocks", ctypes.POINTER(TransformBlock)),
    ]


class ServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_int64), ("max_active", ctypes.c_int32), ("use_graph", ctypes.c_int32)]


class ServerTiming(ctypes.Structure):
    _fields_ = [("submit_ns", ctypes.c_int64), ("admit_ns", ctypes.c_int64), ("done_ns", ctypes.c_int64),
                ("device_bytes", ctypes.c_int64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in ("submitted", "completed", "failed", "active", "queued", "peak_active",
                                              "bytes_in_use", "peak_bytes")]


class Failure(ctypes.Structure):
    _fields_ = [
        ("round", ctypes.c_int32),
        ("phase", ctypes.c_int32),
        ("count", ctypes.c_uint64),
        ("first_logical_index", ctypes.c_int64),
        ("first_block", ctypes.c_int32),
        ("reason_bits", ctypes.c_uint16),
        ("draw_first_logical_index", ctypes.c_int64),
        ("draw_reason_bits", ctypes.c_uint16),
        ("words", ctypes.c_uint64 * 3),
        ("draw_first_block", ctypes.c_int32),
    ]


# name -> (restype, argtypes); the list must match include/issampler.h
SIGNATURES = {
    "is_last_error": (ctypes.c_char_p, []),
    "is_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(VP)]),
    "is_ctx_destroy": (ctypes.c_int, [VP]),
    "is_job_create": (ctypes.c_int, [VP, ctypes.POINTER(JobDesc), ctypes.POINTER(VP)]),
    "is_job_destroy": (ctypes.c_int, [VP]),
    "is_job_device_bytes": (ctypes.c_int64, [ctypes.POINTER(JobDesc)]),
    "is_job_run": (ctypes.c_int, [VP]),
    "is_job_set_normals": (ctypes.c_int, [VP, ctypes.c_int32, F64P, ctypes.c_int64]),
    "is_job_set_radial": (ctypes.c_int, [VP, ctypes.c_int32, F64P, ctypes.c_int64, F64P, ctypes.c_int64]),
    "is_job_set_uniforms": (ctypes.c_int, [VP, ctypes.c_int32, F64P, ctypes.c_int64]),
    "is_job_set_ancestors": (ctypes.c_int, [VP, ctypes.c_int32, I64P, ctypes.c_int64]),
    "is_job_clear_prefilled": (ctypes.c_int, [VP]),
    "is_job_export_normals": (ctypes.c_int, [VP, ctypes.c_int32, F64P, ctypes.c_int64]),
    "is_job_export_radial": (ctypes.c_int, [VP, ctypes.c_int32, F64P, ctypes.c_int64, F64P, ctypes.c_int64]),
    "is_job_export_uniforms": (ctypes.c_int, [VP, ctypes.c_int32, F64P, ctypes.c_int64]),
    "is_job_set_timing": (ctypes.c_int, [VP, ctypes.c_int32]),
    "is_job_set_graph": (ctypes.c_int, [VP, ctypes.c_int32]),
    "is_job_round_kernel_ms": (ctypes.c_int, [VP, F64P]),
    "is_job_total": (ctypes.c_int64, [VP]),
    "is_job_logical_dim": (ctypes.c_int32, [VP]),
    "is_group_create": (ctypes.c_int, [VP, ctypes.POINTER(JobDesc), ctypes.c_int32, ctypes.POINTER(VP)]),
    "is_group_run": (ctypes.c_int, [VP]),
    "is_group_size": (ctypes.c_int32, [VP]),
    "is_group_job": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.POINTER(VP)]),
    "is_group_reset": (ctypes.c_int, [VP]),
    "is_group_round_summary": (ctypes.c_int, [VP, ctypes.c_int32, F64P, F64P]),
    "is_group_destroy": (ctypes.c_int, [VP]),
    "is_job_failure": (ctypes.c_int, [VP, ctypes.POINTER(Failure)]),
    "is_job_get_samples": (ctypes.c_int, [VP, F64P]),
    "is_job_get_logweights": (ctypes.c_int, [VP, F64P]),
    "is_job_get_round_ids": (ctypes.c_int, [VP, I64P]),
    "is_job_get_proposal_ids": (ctypes.c_int, [VP, I64P]),
    "is_job_get_round_summary": (ctypes.c_int, [VP, F64P, F64P, F64P]),
    "is_job_get_proposal": (ctypes.c_int, [VP, F64P, F64P, F64P]),
    "is_job_get_history": (ctypes.c_int, [VP, F64P, F64P, F64P]),
    "is_job_get_locations": (ctypes.c_int, [VP, F64P]),
    "is_job_get_ancestors": (ctypes.c_int, [VP, I64P]),
    "is_job_device_results": (ctypes.c_int, [VP, ctypes.POINTER(VP), I64P, ctypes.POINTER(VP), I32P]),
    "is_job_logweight_offset": (ctypes.c_int, [VP, F64P]),
    "is_job_reset": (ctypes.c_int, [VP]),
    "is_job_lognormalizer": (ctypes.c_int, [VP, F64P]),
    "is_job_moments": (ctypes.c_int, [VP, F64P, F64P, F64P]),
    "is_job_functional_moments": (ctypes.c_int, [VP, VP, ctypes.c_int32, ctypes.c_int32, F64P, F64P]),
    "is_functional_moments_device": (ctypes.c_int, [VP, VP, ctypes.c_int32, VP, ctypes.c_int32, ctypes.c_int64,
                                                    F64P, F64P, VP]),
    "is_functional_moments_device_noalloc": (ctypes.c_int, [VP, VP, ctypes.c_int32, VP, ctypes.c_int32,
                                                            ctypes.c_int64, F64P, F64P, VP]),
    "is_job_resample": (ctypes.c_int, [VP, F64P, ctypes.c_int64, ctypes.c_int32, I64P]),
    "is_job_result": (ctypes.c_int, [VP, ctypes.POINTER(VP)]),
    "is_result_destroy": (ctypes.c_int, [VP]),
    "is_result_count": (ctypes.c_int64, [VP]),
    "is_result_is_weighted": (ctypes.c_int, [VP]),
    "is_result_get_samples": (ctypes.c_int, [VP, F64P]),
    "is_result_get_logweights": (ctypes.c_int, [VP, F64P]),
    "is_result_device_arrays": (ctypes.c_int, [VP, ctypes.POINTER(VP), I64P, ctypes.POINTER(VP), I32P]),
    "is_result_logweight_offset": (ctypes.c_int, [VP, F64P]),
    "is_result_lognormalizer": (ctypes.c_int, [VP, F64P]),
    "is_result_moments": (ctypes.c_int, [VP, F64P, F64P, F64P]),
    "is_result_functional_moments": (ctypes.c_int, [VP, VP, ctypes.c_int32, ctypes.c_int32, F64P, F64P]),
    "is_result_resample": (ctypes.c_int, [VP, F64P, ctypes.c_int64, ctypes.c_int32, I64P, ctypes.POINTER(VP)]),
    "is_cdf_device": (ctypes.c_int, [VP, VP, ctypes.c_int64, VP, VP]),
    "is_select_device": (ctypes.c_int, [VP, VP, VP, ctypes.c_int64, VP, ctypes.c_int64, ctypes.c_int32, VP, VP]),
    "is_cdf": (ctypes.c_int, [VP, F64P, ctypes.c_int64, F64P]),
    "is_select": (ctypes.c_int, [VP, F64P, F64P, ctypes.c_int64, F64P, ctypes.c_int64, ctypes.c_int32, I64P]),
    "is_proposal_logdensity": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_int32,
                                              F64P, F64P, F64P, ctypes.c_int64, ctypes.c_int32, F64P]),
    "is_rng_raw": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64,
                                  ctypes.c_uint64, ctypes.c_int64, U64P]),
    "is_server_create": (ctypes.c_int, [VP, ctypes.POINTER(ServerConfig), ctypes.POINTER(VP)]),
    "is_server_submit": (ctypes.c_int, [VP, ctypes.POINTER(JobDesc), U64P]),
    "is_server_poll": (ctypes.c_int, [VP, ctypes.c_uint64, I32P]),
    "is_server_wait": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.c_int32, I32P]),
    "is_server_take": (ctypes.c_int, [VP, ctypes.c_uint64, ctypes.POINTER(VP), ctypes.POINTER(ServerTiming)]),
    "is_server_get_stats": (ctypes.c_int, [VP, ctypes.POINTER(ServerStats)]),
    "is_server_destroy": (ctypes.c_int, [VP]),
    "is_batch_create": (ctypes.c_int, [VP, ctypes.POINTER(JobDesc), ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(VP)]),
    "is_batch_destroy": (ctypes.c_int, [VP]),
    "is_batch_run": (ctypes.c_int, [VP]),
    "is_batch_set_tuning": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32]),
    "is_batch_get_tuning": (ctypes.c_int, [VP, I32P, I32P]),
    "is_batch_set_order": (ctypes.c_int, [VP, I32P]),
    "is_batch_get_order": (ctypes.c_int, [VP, I32P]),
    "is_batch_trial": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32, F64P]),
    "is_batch_set_timing": (ctypes.c_int, [VP, ctypes.c_int32]),
    "is_batch_kernel_ms": (ctypes.c_int, [VP, F64P]),
    "is_batch_status": (ctypes.c_int, [VP, I32P]),
    "is_batch_job_total": (ctypes.c_int64, [VP, ctypes.c_int32]),
    "is_batch_job_failure": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.POINTER(Failure)]),
    "is_batch_job_samples": (ctypes.c_int, [VP, ctypes.c_int32, F64P]),
    "is_batch_job_logweights": (ctypes.c_int, [VP, ctypes.c_int32, F64P]),
    "is_batch_job_round_summary": (ctypes.c_int, [VP, ctypes.c_int32, F64P, F64P]),
    "is_batch_final_summary": (ctypes.c_int, [VP, F64P, F64P]),
    "is_batch_job_history": (ctypes.c_int, [VP, ctypes.c_int32, F64P, F64P, F64P]),
    "is_batch_job_proposal": (ctypes.c_int, [VP, ctypes.c_int32, F64P, F64P, F64P]),
    "is_batch_export_normals": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32, F64P, ctypes.c_int64]),
    "is_batch_export_radial": (ctypes.c_int, [VP, ctypes.c_int32, ctypes.c_int32, F64P, ctypes.c_int64, F64P,
                                              ctypes.c_int64]),
}

_lib = None


def load():
    """Load libissampler.so; raises OSError when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: build the HIP extension first (__graft_entry__.build())")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def last_error():
    return load().is_last_error().decode("utf-8", "replace")
