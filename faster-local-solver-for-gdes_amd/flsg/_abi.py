"""ctypes mirror of include/flsg.h (structs, constants, function signatures)."""
from __future__ import annotations

import ctypes as C

FLSG_OK, FLSG_EARG, FLSG_EASSERT, FLSG_EHIP, FLSG_ENOMEM, FLSG_ECAP, FLSG_ENOCONV = range(7)
FLSG_PPR, FLSG_KATZ, FLSG_HK = 0, 1, 2
FLSG_GD, FLSG_CHEB = 0, 1
FLSG_FP64, FLSG_FP32 = 0, 1

ERROR_NAMES = {
    FLSG_EARG: "invalid argument", FLSG_EASSERT: "assertion", FLSG_EHIP: "HIP runtime error",
    FLSG_ENOMEM: "out of device memory", FLSG_ECAP: "capacity exceeded",
    FLSG_ENOCONV: "no convergence within the iteration cap",
}


class Params(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("method", C.c_int32), ("local", C.c_int32), ("precision", C.c_int32),
        ("alpha", C.c_double), ("eps", C.c_double), ("l_max", C.c_double), ("mu_min", C.c_double),
        ("tau", C.c_double), ("n_terms", C.c_int64), ("max_iters", C.c_int64),
    ]


class Result(C.Structure):
    _fields_ = [
        ("oper", C.c_int64), ("iters", C.c_int64), ("oper_cert", C.c_int64), ("n_corr", C.c_int64),
        ("supp", C.c_int64), ("nodes", C.c_int64), ("cert", C.c_double), ("ms", C.c_double),
        ("ms_push", C.c_double), ("status", C.c_int32), ("touched", C.c_int32),
        ("ms_recompute", C.c_double), ("ms_correct", C.c_double), ("r_l1", C.c_double),
        ("ms_node", C.c_double), ("ms_pairs", C.c_double), ("ms_bins", C.c_double), ("ms_direct", C.c_double),
        ("bin_iters", C.c_int64), ("corr_edges", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Query(C.Structure):
    _fields_ = [("seed", C.c_int64), ("kind", C.c_int32), ("precision", C.c_int32), ("alpha", C.c_double),
                ("eps", C.c_double)]


class HkConfig(C.Structure):
    _fields_ = [("tau", C.c_double), ("eps", C.c_double), ("n_terms", C.c_int64)]


class Trace(C.Structure):
    _fields_ = [
        ("d_nodes", C.c_void_p), ("cap_nodes", C.c_int64), ("h_off", C.c_void_p), ("cap_iters", C.c_int64),
        ("n_iters", C.c_int64), ("overflow", C.c_int32), ("pad", C.c_int32),
    ]


# name -> (restype, argtypes); the set of symbols include/flsg.h declares
SIGNATURES = {
    "flsg_version": (C.c_int, []),
    "flsg_last_error": (C.c_char_p, []),
    "flsg_device_info": (C.c_int, [C.POINTER(C.c_int32), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "flsg_graph_create": (C.c_int, [C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32,
                                    C.POINTER(C.c_void_p)]),
    "flsg_graph_destroy": (C.c_int, [C.c_void_p]),
    "flsg_solve": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int64, C.c_void_p, C.c_void_p,
                             C.POINTER(Result), C.POINTER(Trace), C.c_void_p]),
    "flsg_hk_solve": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int64, C.c_void_p, C.c_void_p,
                                C.POINTER(Result), C.c_void_p, C.POINTER(C.c_int64), C.POINTER(Trace),
                                C.c_void_p]),
    "flsg_hk_estimate_coefficient": (C.c_double, [C.c_double, C.c_int64]),
    "flsg_hk_estimate_n": (C.c_int64, [C.c_double, C.c_double]),
    "flsg_solve_local_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_int64, C.c_void_p, C.c_void_p,
                                           C.POINTER(Result), C.c_void_p]),
    "flsg_solve_local_noalloc_scratch_bytes": (C.c_int64, []),
    "flsg_batch_solve": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_void_p, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(Result), C.c_void_p]),
    "flsg_batch_slot_bytes": (C.c_int64, [C.c_int64]),
    "flsg_batch_set_slots": (C.c_int, [C.c_void_p, C.c_int32]),
    "flsg_batch_set_block": (C.c_int, [C.c_void_p, C.c_int32]),
    "flsg_batch_set_hub_cache": (C.c_int, [C.c_void_p, C.c_int32]),
    "flsg_batch_set_degree_bins": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "flsg_batch_set_binning": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32]),
    "flsg_batch_set_bin_load": (C.c_int, [C.c_void_p, C.c_int32]),
    "flsg_batch_reserve": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p]),
    "flsg_graph_set_out_map": (C.c_int, [C.c_void_p, C.c_void_p]),
    "flsg_batch_launch": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_int64,
                                    C.c_void_p, C.c_void_p]),
    "flsg_stream_poll": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32)]),
    "flsg_session_create": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.POINTER(C.c_void_p)]),
    "flsg_session_open": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(HkConfig), C.c_int64,
                                    C.c_double]),
    "flsg_session_submit": (C.c_int, [C.c_void_p, C.POINTER(Query), C.c_int64, C.POINTER(C.c_int64)]),
    "flsg_session_done": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p]),
    "flsg_session_result": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(Result), C.POINTER(C.c_int64)]),
    "flsg_session_close": (C.c_int, [C.c_void_p]),
    "flsg_session_poll": (C.c_int, [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
    "flsg_session_output": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p]),
    "flsg_session_destroy": (C.c_int, [C.c_void_p]),
    "flsg_batch_refresh_degrees": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    "flsg_batch_arena_dirty": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]),
    "flsg_boundary_bound": (C.c_double, [C.c_float, C.c_int32, C.c_int64, C.c_int64]),
    "flsg_support_bound": (C.c_double, [C.c_float, C.c_float, C.c_int32, C.c_int64, C.c_double, C.c_int64]),
    "flsg_sor_solve": (C.c_int, [C.c_void_p, C.POINTER(Params), C.c_double, C.c_int64, C.c_void_p, C.c_void_p,
                                 C.POINTER(Result), C.c_void_p]),
    "flsg_sor_ignn_column": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64,
                                       C.POINTER(Result), C.c_void_p]),
    "flsg_graph_colouring": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int64),
                                       C.c_void_p]),
    "flsg_graph_insert_edges": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                          C.POINTER(C.c_int64), C.c_void_p, C.POINTER(C.c_int64), C.c_void_p]),
    "flsg_ignn_propagate": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                      C.c_int64, C.POINTER(Result), C.c_void_p]),
    "flsg_ignn_update": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_int64, C.c_void_p, C.c_void_p]),
}


def bind(lib):
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib
