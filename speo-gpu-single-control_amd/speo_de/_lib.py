"""ctypes binding of libspeo_de.so (the C ABI declared in include/speo_de.h).

The library is built in-tree by ``make -C speo-gpu-single-control_amd`` (or
``__graft_entry__.build()``).  There is no fallback: when the library is
missing, every call raises.
"""
from __future__ import annotations

import ctypes
from ctypes import (POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int64, c_longlong,
                    c_uint8, c_uint32, c_uint64, c_void_p)
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "libspeo_de.so"
HEADER = PKG_DIR.parent / "include" / "speo_de.h"

SPEO_OK, SPEO_ERR_ARG, SPEO_ERR_HIP, SPEO_ERR_UNSUPPORTED, SPEO_ERR_NO_DEVICE, SPEO_ERR_ALLOC = (
    0, -1, -2, -3, -4, -5)
OBJ_ELLIPTIC, OBJ_RASTRIGIN, OBJ_ACKLEY, OBJ_BENT_CIGAR, OBJ_DISCUS = 1, 2, 3, 4, 5
OBJ_HAPPYCAT, OBJ_HGBAT, OBJ_SCHWEFEL = 6, 7, 8
OBJ_ROSENBROCK, OBJ_WEIERSTRASS, OBJ_GRIEWANK, OBJ_KATSUURA = 9, 10, 11, 12
OBJ_GRIE_ROSEN, OBJ_ESCAFFER6 = 13, 14
OBJ_HYBRID_1, OBJ_HYBRID_2, OBJ_HYBRID_3, OBJ_HYBRID_4, OBJ_HYBRID_5, OBJ_HYBRID_6 = range(17, 23)
PREC_FP32, PREC_BF16, PREC_FP16, PREC_BF16X3, PREC_FP16X3, PREC_BF16X6 = range(6)
PREC_NAMES = {PREC_FP32: "fp32", PREC_BF16: "bf16", PREC_FP16: "fp16", PREC_BF16X3: "bf16x3",
              PREC_FP16X3: "fp16x3", PREC_BF16X6: "bf16x6"}
MODE_PLAIN, MODE_FUSED, MODE_EPOCH = 0, 1, 2
REGROUP_STATIC_ORDERED, REGROUP_STATIC_RANDOM, REGROUP_DYNAMIC_RANDOM, REGROUP_DYNAMIC_ORDERED = range(4)
CONFIGURE_CONSTANT, CONFIGURE_DYNAMIC = 0, 1


class SpeoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"speo error {code}: {msg}")
        self.code = code


# ---- drop-in structs (CUDA_EA_library/EA_CUDA.h:21-39, config.h:48-59) ----

class GPU_Population(Structure):
    _fields_ = [("elements", c_void_p), ("fitness_value", c_void_p)]


class GPU_IslandInfo(Structure):
    _fields_ = [("island_size", c_int), ("subisland_num", c_int), ("permutated_index", c_void_p)]


class DEInfo(Structure):
    _fields_ = [("CR", c_float), ("F", c_float), ("strategy_ID", c_int), ("pop_size", c_int)]


class ProblemInfo(Structure):
    _fields_ = [("dim", c_int), ("function_ID", c_int), ("run_ID", c_int),
                ("max_base_FEs", c_int), ("seed", c_int), ("running_time", c_int),
                ("computing_time", c_int), ("max_bound", c_float), ("min_bound", c_float)]


class SPEO_RandState(Structure):
    _fields_ = [("seed", c_uint64), ("generation", c_uint32), ("row_base", c_uint32)]


class EngineConfig(Structure):
    _fields_ = [("num_islands", c_int), ("island_size", POINTER(c_int)),
                ("island_problem", POINTER(c_int)), ("num_problems", c_int),
                ("problems", POINTER(c_void_p)), ("island_F", POINTER(c_float)),
                ("island_CR", POINTER(c_float)), ("F", c_float), ("CR", c_float),
                ("min_bound", c_float), ("max_bound", c_float), ("seed", c_uint64),
                ("interval", c_int), ("migration_size", c_int), ("precision", c_int),
                ("mode", c_int), ("use_graph", c_int), ("rows_per_block", c_int),
                ("surrogate_candidates", c_int), ("surrogate_train_interval", c_int),
                ("surrogate_archive", c_int), ("surrogate_steps", c_int),
                ("surrogate_batch", c_int), ("surrogate_lr", c_float),
                ("surrogate_rho_min", c_float), ("epoch_parts", c_int),
                ("subisland_size", c_int), ("de_strategy", c_int), ("configure_EA", c_int),
                ("regroup_option", c_int)]


class RunRecord(Structure):
    """speo_run_record: the 21 columns of the reference's result row."""
    _fields_ = [("function_ID", c_int), ("run_ID", c_int), ("dim", c_int),
                ("best_fitness", c_double), ("time_period", c_double),
                ("communication_percentage", c_double), ("total_FEs", c_int64),
                ("node_num", c_int), ("GPU_num", c_int), ("island_num", c_int),
                ("total_pop_size", c_int), ("subisland_size", c_int), ("interval", c_int),
                ("migration_size", c_int), ("migration_rate", c_double), ("buffer_capacity", c_int),
                ("migration_topology", c_char_p), ("buffer_manage", c_char_p),
                ("regroup_option", c_char_p), ("configure_EA", c_char_p), ("EA_parameters", c_char_p)]


class DriverConfig(Structure):
    """speo_driver_config: the reference's function x run sweep on one GPU."""
    _fields_ = [("data_dir", c_char_p), ("record_file", c_char_p), ("dim", c_int),
                ("function_IDs", c_void_p), ("num_functions", c_int), ("run_IDs", c_void_p),
                ("num_runs", c_int), ("node_ID", c_int), ("node_num", c_int), ("GPU_num", c_int),
                ("island_num", c_int), ("island_size", c_int), ("max_base_FEs", c_int),
                ("interval", c_int), ("migration_size", c_int), ("migration_rate", c_double),
                ("buffer_capacity", c_int), ("subisland_size", c_int), ("configure_EA", c_int),
                ("regroup_option", c_int), ("mode", c_int), ("precision", c_int),
                ("use_graph", c_int), ("min_bound", c_float), ("max_bound", c_float)]


class EngineStats(Structure):
    _fields_ = [("generations", c_int64), ("evaluations", c_int64), ("migrations", c_int64),
                ("graph_launches", c_int64), ("rows_per_block", c_int),
                ("surrogate_predictions", c_int64), ("surrogate_rounds", c_int64),
                ("epoch_parts", c_int)]


_SIGS = {
    "SPEO_NextPow2Dim": (c_int, [c_int]),
    "API_Initilize": (c_int, [GPU_Population, POINTER(SPEO_RandState), GPU_IslandInfo,
                              ProblemInfo]),
    "API_CheckCUDAError": (c_int, []),
    "API_DE_GenerateNewPopulation": (c_int, [GPU_Population, GPU_Population, c_void_p,
                                             GPU_IslandInfo, c_void_p, POINTER(SPEO_RandState),
                                             ProblemInfo]),
    "API_DE_SelectSurvival": (c_int, [GPU_Population, GPU_Population, GPU_IslandInfo,
                                      ProblemInfo]),
    "API_evaluateFitness": (None, [c_void_p] * 10 + [c_int, c_int, c_float, c_int]),
    "speo_last_error": (c_char_p, []),
    "speo_device_check": (c_int, []),
    "speo_problem_create": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, POINTER(c_void_p)]),
    "speo_problem_create_ex": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p,
                                       c_void_p, c_void_p, c_void_p, c_void_p,
                                       POINTER(c_void_p)]),
    "speo_problem_destroy": (c_int, [c_void_p]),
    "speo_problem_dim": (c_int, [c_void_p]),
    "speo_problem_ncomp": (c_int, [c_void_p]),
    "speo_generate_problem_data": (c_int, [c_uint64, c_int, c_void_p, c_void_p]),
    "speo_problem_seed": (c_uint64, [c_uint64, c_uint32]),
    "speo_load_cec_data": (c_int, [c_char_p, c_int, c_int, c_int, c_void_p, c_void_p]),
    "speo_problem_create_cec": (c_int, [c_char_p, c_int, c_int, POINTER(c_void_p)]),
    "speo_evaluate": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_void_p]),
    "speo_evaluate_f64": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "speo_wide_eval": (c_int, [c_int, POINTER(c_longlong)]),
    "speo_evaluate_noalloc": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int,
                                      c_void_p]),
    "speo_init_population": (c_int, [c_void_p, c_int, c_int, c_int, c_float, c_float, c_uint64,
                                     c_uint32, c_void_p]),
    "speo_de_generate": (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p,
                                 c_void_p, c_void_p, c_float, c_float, SPEO_RandState, c_void_p,
                                 c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "speo_de_select": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int,
                               c_void_p, c_void_p]),
    "speo_segment_argmin": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "speo_engine_create": (c_int, [POINTER(EngineConfig), c_void_p, POINTER(c_void_p)]),
    "speo_engine_destroy": (c_int, [c_void_p]),
    "speo_engine_run": (c_int, [c_void_p, c_int]),
    "speo_engine_prepare": (c_int, [c_void_p, c_int]),
    "speo_engine_surrogate_dump": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                           c_void_p, c_void_p, c_void_p, c_void_p]),
    "speo_engine_surrogate_set_labels": (c_int, [c_void_p, c_int, c_void_p]),
    "speo_engine_run_fes_timed": (c_int, [c_void_p, c_int64, POINTER(c_int), POINTER(c_double),
                                          POINTER(c_double)]),
    "speo_island_driver": (c_int, [c_void_p, c_void_p]),
    "speo_engine_subislands": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "speo_island_regroup": (c_int, [c_uint64, c_int, c_int, c_int, c_int, c_int, c_int, c_void_p,
                                    c_void_p, c_void_p, c_void_p]),
    "speo_engine_sync": (c_int, [c_void_p]),
    "speo_engine_best": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "speo_engine_read": (c_int, [c_void_p, c_void_p, c_void_p]),
    "speo_engine_read_trials": (c_int, [c_void_p, c_void_p, c_void_p]),
    "speo_engine_write": (c_int, [c_void_p, c_void_p, c_void_p]),
    "speo_engine_get_stats": (c_int, [c_void_p, POINTER(EngineStats)]),
    "speo_engine_profile": (c_int, [c_void_p, c_int, c_void_p]),
    "speo_engine_surrogate_stats": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "speo_engine_total_rows": (c_int, [c_void_p]),
    "speo_engine_max_dim": (c_int, [c_void_p]),
    "speo_engine_buffers": (c_int, [c_void_p, POINTER(c_void_p), POINTER(c_void_p), c_void_p,
                                    c_void_p, c_void_p]),
    "speo_mlp_create": (c_int, [c_int, c_int, c_uint64, POINTER(c_void_p)]),
    "speo_mlp_destroy": (c_int, [c_void_p]),
    "speo_mlp_num_params": (c_int64, [c_void_p]),
    "speo_mlp_params": (c_int, [c_void_p, c_void_p, c_void_p]),
    "speo_mlp_forward": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "speo_mlp_infer": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "speo_mlp_loss_grad": (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_float, c_float,
                                   c_void_p, c_void_p]),
    "speo_mlp_adam_step": (c_int, [c_void_p, c_float]),
    "speo_mlp_train_step": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_uint64, c_uint32,
                                    c_uint32, c_int, c_float, c_float, c_float, c_void_p]),
    "speo_spearman": (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    "speo_generations_for_fes": (c_int64, [c_int64, c_int64, c_int64]),
    "speo_engine_run_fes": (c_int, [c_void_p, c_int64, POINTER(c_int)]),
    "speo_run_seed": (c_uint64, [c_int, c_int, c_int]),
    "speo_de_parameters": (c_int, [c_double, c_double, c_int, c_char_p, c_int]),
    "speo_result_header": (c_char_p, []),
    "speo_result_record": (c_int, [POINTER(RunRecord), c_char_p, c_int]),
}

_lib = None


def lib():
    """Load libspeo_de.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        # libspeo_de.so and torch both need libamdhip64.so.7; the first one
        # loaded serves the process.  Load torch's first so that the library
        # and torch share torch's HIP runtime (the one that matches the
        # driver on the GPU boxes).
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built: run `make -C {PKG_DIR}` "
                               "(or __graft_entry__.build()); there is no CPU fallback")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error() -> str:
    return lib().speo_last_error().decode()


def check(rc: int) -> None:
    if rc != SPEO_OK:
        raise SpeoError(rc, last_error())


def header_symbols() -> list[str]:
    """Every SPEO_EXPORT function declared in include/speo_de.h."""
    import re
    text = HEADER.read_text()
    names = re.findall(r"^SPEO_EXPORT\s+[\w\s\*]+?\**\s*(\w+)\s*\(", text, re.M)
    return sorted(set(names))
