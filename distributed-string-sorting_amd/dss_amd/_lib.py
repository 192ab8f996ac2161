"""ctypes binding of libdss_gpu.so (the C ABI declared in include/dss_gpu.h).

The library is built in-tree by distributed-string-sorting_amd/build.py.  There
is no CPU fallback: importing this module without the built library raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdss_gpu.so")

u8p = C.POINTER(C.c_uint8)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
set_p = C.c_void_p
pool_p = C.c_void_p
codec_p = C.c_void_p


class SetInfo(C.Structure):
    _fields_ = [("num_strings", C.c_uint64), ("num_bytes", C.c_uint64),
                ("num_pes", C.c_uint32), ("max_len", C.c_uint32)]


class KernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64),
                ("total_ms", C.c_double), ("alg_bytes", C.c_double)]


class SortParams(C.Structure):
    _fields_ = [("large_min", C.c_uint32), ("wave_class_max", C.c_uint32),
                ("skip_lanes", C.c_uint32), ("skip_strings", C.c_uint32),
                ("merge_fan_in", C.c_uint32), ("pass_mode", C.c_uint32), ("merge_tile", C.c_uint32),
                ("digit_chars", C.c_uint32)]


class CheckResult(C.Structure):
    _fields_ = [("num_strings", C.c_uint64), ("order_errors", C.c_uint64),
                ("lcp_errors", C.c_uint64), ("dist_prefix_size", C.c_uint64),
                ("sorted", C.c_int32), ("complete", C.c_int32)]


class DistCheck(C.Structure):
    _fields_ = [("num_strings", C.c_uint64), ("lower_violations", C.c_uint64),
                ("upper_violations", C.c_uint64), ("exact_sum", C.c_uint64),
                ("results_sum", C.c_uint64)]


class DssError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dss error {code}: {msg}")
        self.code = code


_lib = None

# name -> (restype, argtypes); every function returns int status unless noted
_SIGS = {
    "dss_last_error": (C.c_char_p, []),
    "dss_device_info": (C.c_int, [C.c_char_p, C.c_size_t]),
    "dss_set_create": (C.c_int, [u8p, C.c_uint64, C.c_uint32, C.POINTER(set_p)]),
    "dss_set_create_dton": (C.c_int, [C.c_uint64, C.c_uint32, C.c_double, C.c_uint64, C.c_uint32,
                                      C.POINTER(set_p)]),
    "dss_set_create_skewed": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_double, C.c_double,
                                        C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint32,
                                        C.POINTER(set_p)]),
    "dss_set_create_text": (C.c_int, [u8p, C.c_uint64, C.c_uint64, C.c_uint32, C.POINTER(set_p)]),
    "dss_set_load_file": (C.c_int, [C.c_char_p, C.c_uint64, C.c_uint32, C.POINTER(set_p)]),
    "dss_set_create_pes": (C.c_int, [u8p, u64p, C.c_uint32, C.POINTER(set_p)]),
    "dss_set_create_suffixes": (C.c_int, [u8p, C.c_uint64, C.c_uint32, C.c_uint32, C.POINTER(set_p)]),
    "dss_set_create_markov_suffixes": (C.c_int, [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                                 C.c_uint64, C.c_uint32, C.POINTER(set_p)]),
    "dss_set_destroy": (C.c_int, [set_p]),
    "dss_set_stream": (C.c_int, [set_p, C.c_void_p]),
    "dss_set_info": (C.c_int, [set_p, C.POINTER(SetInfo)]),
    "dss_set_pe_begin": (C.c_int, [set_p, u64p]),
    "dss_set_download": (C.c_int, [set_p, u8p, u64p, u32p]),
    "dss_sort_lcp": (C.c_int, [set_p]),
    "dss_get_sorted": (C.c_int, [set_p, u32p, u32p]),
    "dss_set_local_order": (C.c_int, [set_p, u32p, u32p]),
    "dss_set_sort_params": (C.c_int, [set_p, C.POINTER(SortParams)]),
    "dss_set_reserve": (C.c_int, [set_p, C.c_uint64, C.c_uint32]),
    "dss_get_sort_params": (C.c_int, [set_p, C.POINTER(SortParams)]),
    "dss_set_interval": (C.c_int, [set_p, u64p]),
    "dss_set_dist": (C.c_int, [set_p, u32p]),
    "dss_avg_lcp": (C.c_int, [set_p, u64p]),
    "dss_partition": (C.c_int, [set_p, C.c_int, C.c_uint64, C.c_uint64, C.c_int, u64p]),
    "dss_exchange": (C.c_int, [set_p, C.c_int, u64p, u64p, u64p]),
    "dss_get_exchange": (C.c_int, [set_p, u8p, u32p, u32p, u32p]),
    "dss_merge_lcp": (C.c_int, [set_p]),
    "dss_get_merged": (C.c_int, [set_p, u32p, u32p, u32p]),
    "dss_set_merged": (C.c_int, [set_p, u32p, u32p]),
    "dss_extend_prefix": (C.c_int, [set_p, u64p]),
    "dss_get_extended": (C.c_int, [set_p, u8p, u64p]),
    "dss_prefix_doubling": (C.c_int, [set_p, C.c_uint64, C.c_uint32, u32p]),
    "dss_get_prefix_doubling": (C.c_int, [set_p, u32p, u8p]),
    "dss_hash_prefixes": (C.c_int, [set_p, u32p, C.c_uint64, C.c_uint32, C.c_uint64, u64p]),
    "dss_merge_sort": (C.c_int, [set_p, C.c_int, C.c_int]),
    "dss_prefix_doubling_sort": (C.c_int, [set_p, C.c_int, C.c_uint64]),
    "dss_get_permutation": (C.c_int, [set_p, u32p, u32p]),
    "dss_check": (C.c_int, [set_p, C.POINTER(CheckResult)]),
    "dss_exact_dist": (C.c_int, [set_p, u32p]),
    "dss_check_dist": (C.c_int, [set_p, C.POINTER(DistCheck)]),
    "dss_sync": (C.c_int, [set_p]),
    "dss_set_timing": (C.c_int, [set_p, C.c_int]),
    "dss_kernel_stats": (C.c_int, [set_p, C.POINTER(KernelStat), C.c_uint32, u32p, C.c_int]),
    "dss_pool_create": (C.c_int, [C.c_uint64, C.c_uint64, C.POINTER(pool_p)]),
    "dss_pool_destroy": (C.c_int, [pool_p]),
    "dss_pool_add_job": (C.c_int, [pool_p, u8p, C.c_uint64, u32p]),
    "dss_pool_add_job_dton": (C.c_int, [pool_p, C.c_uint64, C.c_uint32, C.c_double, C.c_uint64, u32p]),
    "dss_pool_add_job_skewed": (C.c_int, [pool_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_double,
                                          C.c_double, C.c_uint32, C.c_uint32, C.c_uint64, u32p]),
    "dss_pool_free_job": (C.c_int, [pool_p, C.c_uint32]),
    "dss_pool_job_info": (C.c_int, [pool_p, C.c_uint32, u64p, u64p]),
    "dss_pool_usage": (C.c_int, [pool_p, u64p, u64p, u64p, u64p]),
    "dss_pool_batch": (C.c_int, [pool_p, u32p, C.c_uint32, C.POINTER(set_p)]),
    "dss_pool_rebatch": (C.c_int, [pool_p, set_p, u32p, C.c_uint32]),
    "dss_pe_dist_sizes": (C.c_int, [set_p, u64p]),
    "dss_get_sorted_pe": (C.c_int, [set_p, C.c_uint32, u32p, u32p]),
    "dss_codec_create": (C.c_int, [C.POINTER(codec_p)]),
    "dss_codec_destroy": (C.c_int, [codec_p]),
    "dss_codec_sync": (C.c_int, [codec_p]),
    "dss_codec_set_timing": (C.c_int, [codec_p, C.c_int]),
    "dss_codec_kernel_stats": (C.c_int, [codec_p, C.POINTER(KernelStat), C.c_uint32, u32p, C.c_int]),
    "dss_varint_write_ranges": (C.c_int, [codec_p, u64p, u64p, C.c_uint32, u8p, C.c_uint64, u64p, u64p]),
    "dss_varint_write_ranges_dev": (C.c_int, [codec_p, C.c_void_p, u64p, C.c_uint32, C.c_void_p, C.c_uint64,
                                              u64p, u64p]),
    "dss_varint_write_ranges_noalloc": (C.c_int, [codec_p, C.c_void_p, u64p, C.c_uint32, C.c_void_p, C.c_uint64,
                                                  u64p, u64p]),
    "dss_varint_read_ranges": (C.c_int, [codec_p, u8p, C.c_uint64, u64p, C.c_uint32, u64p]),
    "dss_varint_read_ranges_dev": (C.c_int, [codec_p, C.c_void_p, C.c_uint64, u64p, C.c_uint32, C.c_void_p]),
    "dss_golomb_b": (C.c_int, [C.c_uint64, C.c_uint64, u64p]),
    "dss_golomb_encode": (C.c_int, [codec_p, C.c_int, u64p, u64p, u64p, C.c_uint32, u64p, C.c_uint64, u64p,
                                    u64p]),
    "dss_golomb_encode_dev": (C.c_int, [codec_p, C.c_int, C.c_void_p, u64p, u64p, C.c_uint32, C.c_void_p,
                                        C.c_uint64, u64p, u64p]),
    "dss_golomb_decode": (C.c_int, [codec_p, C.c_int, u64p, C.c_uint64, C.c_uint32, C.c_uint64, u64p,
                                    C.c_uint64, u64p, u32p]),
    "dss_golomb_decode_dev": (C.c_int, [codec_p, C.c_int, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64,
                                        C.c_void_p, C.c_uint64, u64p, u32p]),
}


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run `python distributed-string-sorting_amd/build.py` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def symbols():
    return list(_SIGS)


def check(rc: int):
    if rc != 0:
        msg = lib().dss_last_error()
        raise DssError(rc, msg.decode() if msg else "")
    return rc
