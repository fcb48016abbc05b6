"""ctypes binding of include/lbm_capi.h (the C-ABI of liblbm_hip.so).

Loading never falls back to anything else: if the HIP library is missing the
import of the solver classes raises, naming the build step.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_hip.so")

# status codes / constants (include/lbm_capi.h)
LBM_OK = 0
LBM_ERR_INVALID_ARGUMENT = 1
LBM_ERR_OUT_OF_RANGE = 2
LBM_ERR_DEVICE = 3
LBM_ERR_OUT_OF_MEMORY = 4
LBM_ERR_PARAMETRIZATION = 5
LBM_ERR_DECOMPOSITION = 6
LBM_ERR_STATE = 7
LBM_ERR_CAPACITY = 8

FLAG_OBSTACLE = 1
FLAG_FLUID = 2
FLAG_VELOCITY_INJECTION = 4
FLAG_GHOST_LAYER = 8

STORAGE_F32 = 0
STORAGE_F64 = 1
STORAGE_F16_SHIFTED = 2
STORAGE_BF16_SHIFTED = 3

FUSED_X_AUTO = 2  # x faces in the step kernel for FP32/FP64 storage, copy launch for FP16/BF16

FORCING_REFERENCE = 0
FORCING_BODY = 1

FIELD_DD = 0
FIELD_VELOCITY = 1
FIELD_DENSITY = 2
FIELD_FLAGS = 3
FIELD_DD_RAW = 4

CASE_CAVITY = 0
CASE_CHANNEL = 1
CASE_QUEUED = 0
CASE_RUNNING = 1
CASE_DONE = 2
CASE_FAILED = 3

EXPORTED_SYMBOLS = (
    "lbm_version", "lbm_last_error", "lbm_device_synchronize", "lbm_solver_config_default",
    "lbm_device_copy_bandwidth",
    "lbm_solver_create", "lbm_solver_destroy", "lbm_solver_error", "lbm_solver_reset",
    "lbm_solver_init_perturbed", "lbm_solver_simulation_step", "lbm_solver_simulation_step_alpha",
    "lbm_solver_simulation_step_beta", "lbm_solver_run", "lbm_solver_wait",
    "lbm_solver_get_step_counter", "lbm_solver_set_step_counter", "lbm_solver_get_config",
    "lbm_solver_set_lid_velocity", "lbm_solver_set_relaxation", "lbm_solver_store", "lbm_solver_set",
    "lbm_solver_velocity_checksum", "lbm_solver_mass_energy", "lbm_solver_mass_energy_noalloc",
    "lbm_solver_moments", "lbm_copy_box",
    "lbm_solver_device_pointers", "lbm_group_create", "lbm_group_destroy", "lbm_group_error",
    "lbm_group_add_comm", "lbm_group_set_exchange_mode", "lbm_group_set_graph", "lbm_group_set_graph_pairs",
    "lbm_group_prepare_steps", "lbm_group_get_launch_counts", "lbm_group_step",
    "lbm_group_wait", "lbm_group_set_timing", "lbm_group_get_timing", "lbm_group_sync",
    "lbm_group_get_timing_detail", "lbm_solver_place_spheres", "lbm_group_set_launch",
    "lbm_group_get_launch", "lbm_group_set_alpha_work_group", "lbm_group_get_alpha_work_group",
    "lbm_group_autotune",
    "lbm_solver_required_bytes", "lbm_solver_create_in", "lbm_solver_colocate", "lbm_group_create_empty", "lbm_group_add_member",
    "lbm_group_remove_member", "lbm_group_member_count", "lbm_group_set_member_stop", "lbm_group_get_stream",
    "lbm_server_config_default", "lbm_case_default", "lbm_server_create", "lbm_server_destroy",
    "lbm_server_error", "lbm_server_case_bytes", "lbm_server_submit", "lbm_server_progress",
    "lbm_server_next_done", "lbm_server_status", "lbm_server_fetch", "lbm_server_stats_get", "lbm_server_time",
    "lbm_group_set_fused_x",
)


class LbmError(RuntimeError):
    """A C-ABI call returned a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"[lbm status {code}] {message}")
        self.code = code


class SolverConfig(C.Structure):
    _fields_ = [
        ("dims", C.c_int * 3),
        ("bc", C.c_int * 6),
        ("storage", C.c_int),
        ("forcing", C.c_int),
        ("inv_tau", C.c_double),
        ("tau", C.c_double),
        ("smagorinsky_cs", C.c_double),
        ("gravitation", C.c_double * 3),
        ("lid_velocity", C.c_double),
        ("store_velocity", C.c_int),
        ("store_density", C.c_int),
        ("work_group_size", C.c_int),
        ("device", C.c_int),
        ("kernel_variant", C.c_int),
        ("lds_per_block", C.c_int),
    ]


class ServerConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int),
        ("memory_budget", C.c_size_t),
        ("max_active", C.c_int),
        ("chunk_ms", C.c_double),
        ("max_chunk_steps", C.c_int),
        ("work_group_size", C.c_int),
        ("kernel_variant", C.c_int),
        ("autotune", C.c_int),
        ("backfill", C.c_int),
        ("max_active_cells", C.c_longlong),
    ]


class Case(C.Structure):
    _fields_ = [
        ("kind", C.c_int),
        ("dims", C.c_int * 3),
        ("steps", C.c_int),
        ("storage", C.c_int),
        ("tau", C.c_double),
        ("smagorinsky_cs", C.c_double),
        ("lid_velocity", C.c_double),
        ("body_force", C.c_double),
        ("obstacle_fill", C.c_double),
        ("r_min", C.c_double),
        ("r_max", C.c_double),
        ("perturbation", C.c_double),
        ("seed", C.c_uint64),
    ]


class CaseStatus(C.Structure):
    _fields_ = [
        ("id", C.c_longlong),
        ("state", C.c_int),
        ("steps_done", C.c_longlong),
        ("cells", C.c_longlong),
        ("submit_s", C.c_double),
        ("admit_s", C.c_double),
        ("done_s", C.c_double),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_longlong),
        ("completed", C.c_longlong),
        ("failed", C.c_longlong),
        ("chunks", C.c_longlong),
        ("steps_cells", C.c_longlong),
        ("bytes_in_use", C.c_size_t),
        ("bytes_peak", C.c_size_t),
        ("queued", C.c_int),
        ("active", C.c_int),
        ("ready", C.c_int),
        ("device_ms", C.c_double),
        ("cells_peak", C.c_longlong),
    ]


_lib = None


def load_library(path: str | None = None):
    """Load liblbm_hip.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or os.environ.get("LBM_HIP_LIBRARY", LIB_PATH)
    if not os.path.exists(path):
        raise ImportError(
            f"{path} not found: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')"
        )
    lib = C.CDLL(path)
    P = C.c_void_p
    I = C.c_int
    IP = C.POINTER(C.c_int)
    D = C.c_double

    def sig(name, res, *args):
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = list(args)

    sig("lbm_version", C.c_char_p)
    sig("lbm_last_error", C.c_char_p)
    sig("lbm_device_synchronize", I, I)
    sig("lbm_solver_config_default", None, C.POINTER(SolverConfig))
    sig("lbm_device_copy_bandwidth", I, I, C.c_size_t, I, I, C.POINTER(D), C.POINTER(D))
    sig("lbm_solver_create", I, C.POINTER(SolverConfig), C.POINTER(P))
    sig("lbm_solver_destroy", I, P)
    sig("lbm_solver_error", C.c_char_p, P)
    sig("lbm_solver_reset", I, P)
    sig("lbm_solver_init_perturbed", I, P, D, C.c_uint64, C.c_uint64, IP, IP)
    sig("lbm_solver_simulation_step", I, P)
    sig("lbm_solver_simulation_step_alpha", I, P)
    sig("lbm_solver_simulation_step_beta", I, P)
    sig("lbm_solver_run", I, P, I)
    sig("lbm_solver_wait", I, P)
    sig("lbm_solver_get_step_counter", I, P, C.POINTER(C.c_longlong))
    sig("lbm_solver_set_step_counter", I, P, C.c_longlong)
    sig("lbm_solver_get_config", I, P, C.POINTER(SolverConfig))
    sig("lbm_solver_set_lid_velocity", I, P, D)
    sig("lbm_solver_set_relaxation", I, P, D, D, D)
    sig("lbm_solver_store", I, P, I, IP, IP, P)
    sig("lbm_solver_set", I, P, I, IP, IP, P, IP)
    sig("lbm_solver_velocity_checksum", I, P, C.POINTER(C.c_float))
    sig("lbm_solver_mass_energy", I, P, C.POINTER(D), C.POINTER(D))
    sig("lbm_solver_mass_energy_noalloc", I, P, C.POINTER(D), C.POINTER(D))
    sig("lbm_solver_moments", I, P, IP, IP, P, P)
    sig("lbm_copy_box", I, P, IP, P, IP, IP, IP)
    sig("lbm_solver_device_pointers", I, P, C.POINTER(P), C.POINTER(P), C.POINTER(C.c_longlong))
    sig("lbm_solver_place_spheres", I, P, C.c_uint64, D, D, D, IP, IP, C.POINTER(D))
    sig("lbm_group_create", I, C.POINTER(P), I, C.POINTER(P))
    sig("lbm_group_destroy", I, P)
    sig("lbm_group_error", C.c_char_p, P)
    sig("lbm_group_add_comm", I, P, I, I, IP, IP, IP, IP)
    sig("lbm_group_set_exchange_mode", I, P, I)
    sig("lbm_group_set_graph", I, P, I)
    sig("lbm_group_set_graph_pairs", I, P, I)
    sig("lbm_group_prepare_steps", I, P, I)
    LL = C.POINTER(C.c_longlong)
    sig("lbm_group_get_launch_counts", I, P, LL, LL, LL)
    sig("lbm_group_set_fused_x", I, P, I)
    sig("lbm_group_step", I, P, I)
    sig("lbm_group_wait", I, P)
    sig("lbm_group_sync", I, P, I)
    sig("lbm_group_get_timing_detail", I, P, C.POINTER(D), C.POINTER(D), C.POINTER(D))
    sig("lbm_group_set_launch", I, P, I, I)
    sig("lbm_group_get_launch", I, P, IP, IP)
    sig("lbm_group_set_alpha_work_group", I, P, I)
    sig("lbm_group_get_alpha_work_group", I, P, IP)
    sig("lbm_group_autotune", I, P, I, IP, IP, C.POINTER(D))
    sig("lbm_solver_required_bytes", I, C.POINTER(SolverConfig), C.POINTER(C.c_size_t))
    sig("lbm_solver_create_in", I, C.POINTER(SolverConfig), P, C.c_size_t, P, C.POINTER(P))
    sig("lbm_solver_colocate", I, C.POINTER(P), I)
    sig("lbm_group_create_empty", I, I, I, I, I, C.POINTER(P))
    sig("lbm_group_add_member", I, P, P, IP)
    sig("lbm_group_remove_member", I, P, I)
    sig("lbm_group_member_count", I, P, IP)
    sig("lbm_group_set_member_stop", I, P, I, C.c_longlong)
    sig("lbm_group_get_stream", I, P, C.POINTER(P))
    sig("lbm_server_config_default", None, C.POINTER(ServerConfig))
    sig("lbm_case_default", None, C.POINTER(Case))
    sig("lbm_server_create", I, C.POINTER(ServerConfig), C.POINTER(P))
    sig("lbm_server_destroy", I, P)
    sig("lbm_server_error", C.c_char_p, P)
    sig("lbm_server_case_bytes", I, C.POINTER(Case), C.POINTER(C.c_size_t))
    sig("lbm_server_submit", I, P, C.POINTER(Case), C.POINTER(C.c_longlong))
    sig("lbm_server_progress", I, P, IP)
    sig("lbm_server_next_done", I, P, C.POINTER(CaseStatus))
    sig("lbm_server_status", I, P, C.c_longlong, C.POINTER(CaseStatus))
    sig("lbm_server_fetch", I, P, C.c_longlong, P, P)
    sig("lbm_server_stats_get", I, P, C.POINTER(ServerStats))
    sig("lbm_server_time", D, P)
    sig("lbm_group_set_timing", I, P, I)
    sig("lbm_group_get_timing", I, P, C.POINTER(D), C.POINTER(C.c_longlong), C.POINTER(D),
        C.POINTER(C.c_longlong))
    _lib = lib
    return lib


def ivec(values):
    """int[3]/int[6] argument from a sequence (None stays NULL)."""
    if values is None:
        return None
    return (C.c_int * len(values))(*[int(v) for v in values])


def check(code: int, message_fn) -> None:
    if code != LBM_OK:
        raise LbmError(code, message_fn())
