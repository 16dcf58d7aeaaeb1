"""ctypes binding of libsimcore_amd.so (declared in include/simcore_amd.h).

The library is loaded from the package directory only.  There is no
fallback: if the shared object is missing, importing this module raises.
"""
import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_int, c_int32, c_int64,
                    c_size_t, c_uint64, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsimcore_amd.so")

SC_OK = 0
SC_ERR_ARG = -1
SC_ERR_HIP = -2
SC_ERR_STATE = -3
SC_ERR_UNPHYSICAL = -4
SC_ERR_NOMEM = -5
SC_ERR_PARSE = -6
SC_PREC_FP64 = 0
SC_PREC_MIXED = 1
SC_NOISE_COUNTER = 0
SC_NOISE_BUFFER = 1

# every symbol include/simcore_amd.h declares
EXPORTS = [
    "sc_abi_version", "sc_create", "sc_load_filaments", "sc_step",
    "sc_get_sites", "sc_get_bonds", "sc_get_bond_forces", "sc_get_tensions",
    "sc_get_cell_keys", "sc_get_candidates", "sc_get_slot_forces",
    "sc_get_dr_state", "sc_get_counters", "sc_set_sites", "sc_set_noise", "sc_device_bytes",
    "sc_thermo_enable", "sc_calculate_pressure", "sc_get_bond_energies",
    "sc_profile_enable", "sc_profile_read", "sc_last_error", "sc_destroy",
    "sc_params_from_yaml", "sc_params_init", "sc_filament_layout",
    "sc_generate_filaments", "sc_copy_probe", "sc_time_pair", "sc_step_async",
    "sc_step_poll",
    "sc_filament_stats", "sc_theta_histogram", "sc_theta_histogram_noalloc",
    "sc_output_header", "sc_spec_frame",
    "sc_posit_frame", "sc_thermo_header", "sc_thermo_frame", "sc_checkpoint_save", "sc_checkpoint_load", "sc_autotune",
    "sc_set_launch_shape", "sc_get_launch_shape",
    "sc_load_crosslinks", "sc_get_crosslinks", "sc_generate_crosslinks",
    "sc_ensemble_create", "sc_ensemble_submit", "sc_ensemble_progress",
    "sc_ensemble_drain", "sc_ensemble_request", "sc_ensemble_result",
    "sc_ensemble_stats_get", "sc_ensemble_clock", "sc_ensemble_last_error",
    "sc_ensemble_destroy",
]

SC_REQ_QUEUED = 0
SC_REQ_RUNNING = 1
SC_REQ_DONE = 2


class ScParams(ctypes.Structure):
    _fields_ = [
        ("n_dim", c_int), ("n_periodic", c_int),
        ("unit_cell", c_double * 9), ("unit_cell_inv", c_double * 9),
        ("system_radius", c_double), ("min_cell_length", c_double),
        ("delta", c_double), ("wca_eps", c_double), ("wca_sig", c_double),
        ("f_cutoff", c_double), ("stoch_flag", c_int), ("precision", c_int),
        ("use_graph", c_int), ("spatial_order", c_int), ("seed", c_uint64),
        ("capacity_factor", c_double),
    ]


class ScFilamentParams(ctypes.Structure):
    _fields_ = [
        ("bond_length", c_double), ("persistence_length", c_double),
        ("gamma_par", c_double), ("gamma_perp", c_double),
        ("diameter", c_double),
    ]


class ScEnsembleParams(ctypes.Structure):
    _fields_ = [("hbm_budget", c_size_t), ("max_active", c_int32),
                ("steps_per_launch", c_int32), ("batch_size", c_int32)]


class ScRequestInfo(ctypes.Structure):
    _fields_ = [("status", c_int32), ("n_filaments", c_int32),
                ("n_sites", c_int64), ("n_bonds", c_int64), ("n_steps", c_int64),
                ("steps_done", c_int64), ("t_submit", c_double), ("t_admit", c_double),
                ("t_done", c_double), ("device_bytes", c_size_t)]


class ScEnsembleStats(ctypes.Structure):
    _fields_ = [("submitted", c_int64), ("completed", c_int64), ("failed", c_int64),
                ("running", c_int64), ("queued", c_int64),
                ("latency_p50", c_double), ("latency_p99", c_double),
                ("latency_mean", c_double), ("latency_max", c_double),
                ("segment_steps", c_double), ("span", c_double),
                ("peak_active", c_int32), ("peak_bytes", c_size_t),
                ("arena_bytes", c_size_t), ("batched_steps", c_int64),
                ("solo_steps", c_int64), ("host_admit_s", c_double),
                ("host_chunk_s", c_double), ("host_retire_s", c_double),
                ("host_capture_s", c_double), ("graph_captures", c_int64)]


class ScCrosslinkParams(ctypes.Structure):
    _fields_ = [("k_spring", c_double), ("rest_length", c_double), ("k_compress", c_double),
                ("k_off", c_double), ("r_bind", c_double), ("seed", c_uint64)]


class ScLaunchShape(ctypes.Structure):
    _fields_ = [("pair_block", c_int32), ("pair_xcd", c_int32),
                ("integrate_waves", c_int32), ("integrate_layout", c_int32),
                ("integrate_streams", c_int32), ("pair_parts", c_int32),
                ("early_long", c_int32)]


class ScTuneReport(ctypes.Structure):
    _fields_ = [("shape", ScLaunchShape), ("n_trials", c_int32),
                ("step_ms_start", c_double), ("step_ms_best", c_double)]


class SimcoreError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"simcore_amd error {code}: {msg}")
        self.code = code


class UnphysicalResult(SimcoreError):
    """The reference throws std::runtime_error("Unphysical result") when a
    filament's tracked displacement exceeds 10000 (species.hpp:114-118)."""


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` "
            "(make -C simcore_amd/csrc)")
    lib = ctypes.CDLL(LIB_PATH)
    H = c_void_p
    dp, ip, lp = POINTER(c_double), POINTER(c_int32), POINTER(c_int64)
    sig = {
        "sc_abi_version": ([], c_int),
        "sc_create": ([POINTER(ScParams), POINTER(H)], c_int),
        "sc_load_filaments": ([H, c_int32, ip, dp, POINTER(ScFilamentParams)], c_int),
        "sc_step": ([H, c_int32], c_int),
        "sc_step_async": ([H, c_int32], c_int),
        "sc_step_poll": ([H, POINTER(c_int)], c_int),
        "sc_get_sites": ([H, dp, dp, dp], c_int),
        "sc_get_bonds": ([H, dp, dp, dp, dp], c_int),
        "sc_get_bond_forces": ([H, dp, dp], c_int),
        "sc_get_tensions": ([H, dp], c_int),
        "sc_get_cell_keys": ([H, ip], c_int),
        "sc_get_candidates": ([H, ip, ip, c_size_t, POINTER(c_size_t)], c_int),
        "sc_get_slot_forces": ([H, dp, c_size_t], c_int),
        "sc_get_dr_state": ([H, dp], c_int),
        "sc_get_counters": ([H, lp, lp, lp, ip], c_int),
        "sc_set_sites": ([H, dp], c_int),
        "sc_set_noise": ([H, c_int, c_uint64, dp, c_size_t], c_int),
        "sc_thermo_enable": ([H, c_int], c_int),
        "sc_calculate_pressure": ([H, c_int32, dp, dp, dp], c_int),
        "sc_get_bond_energies": ([H, dp], c_int),
        "sc_filament_stats": ([H, dp, dp], c_int),
        "sc_checkpoint_save": ([H, c_void_p, c_size_t, POINTER(c_size_t)], c_int),
        "sc_checkpoint_load": ([H, c_void_p, c_size_t], c_int),
        "sc_autotune": ([H, c_int32, POINTER(ScTuneReport)], c_int),
        "sc_set_launch_shape": ([H, POINTER(ScLaunchShape)], c_int),
        "sc_get_launch_shape": ([H, POINTER(ScLaunchShape)], c_int),
        "sc_load_crosslinks": ([H, c_int64, ip, dp, POINTER(ScCrosslinkParams)], c_int),
        "sc_get_crosslinks": ([H, ip, dp, dp, lp], c_int),
        "sc_generate_crosslinks": ([POINTER(ScParams), c_int32, ip, dp, POINTER(ScFilamentParams),
                                    c_int64, c_double, c_uint64, ip, dp], c_int),
        "sc_theta_histogram": ([H, c_int32, c_int32, lp], c_int),
        "sc_theta_histogram_noalloc": ([H, c_int32, c_int32, lp], c_int),
        "sc_device_bytes": ([H, POINTER(c_size_t)], c_int),
        "sc_profile_enable": ([H, c_int], c_int),
        "sc_profile_read": ([H, dp, lp, c_int], c_int),
        "sc_last_error": ([H], c_char_p),
        "sc_copy_probe": ([c_size_t, c_int, dp], c_int),
        "sc_time_pair": ([H, c_int32, dp], c_int),
        "sc_destroy": ([H], c_int),
        "sc_ensemble_create": ([POINTER(ScEnsembleParams), POINTER(H)], c_int),
        "sc_ensemble_submit": ([H, POINTER(ScParams), c_int32, ip, dp,
                                POINTER(ScFilamentParams), c_int64, c_double, lp], c_int),
        "sc_ensemble_progress": ([H, ip, ip], c_int),
        "sc_ensemble_drain": ([H, c_double], c_int),
        "sc_ensemble_request": ([H, c_int64, POINTER(ScRequestInfo)], c_int),
        "sc_ensemble_result": ([H, c_int64, dp], c_int),
        "sc_ensemble_stats_get": ([H, POINTER(ScEnsembleStats)], c_int),
        "sc_ensemble_clock": ([H], c_double),
        "sc_ensemble_last_error": ([H], c_char_p),
        "sc_ensemble_destroy": ([H], c_int),
        "sc_params_from_yaml": ([c_char_p, POINTER(ScParams), dp], c_int),
        "sc_params_init": ([POINTER(ScParams), c_double], c_int),
        "sc_output_header": ([c_char_p, c_int32, c_int32, c_double], c_int),
        "sc_thermo_header": ([c_char_p, c_int32, c_int32, c_double, c_int32], c_int),
        "sc_thermo_frame": ([c_char_p, dp, dp, c_double, c_double], c_int),
        "sc_spec_frame": ([c_char_p, c_int32, ip, dp, POINTER(ScFilamentParams)], c_int),
        "sc_posit_frame": ([c_char_p, POINTER(ScParams), c_int32, ip, dp,
                            POINTER(ScFilamentParams)], c_int),
        "sc_filament_layout": ([c_double, c_double, c_double, c_double, ip, dp], c_int),
        "sc_generate_filaments": ([POINTER(ScParams), c_int32, ip, dp, dp, dp, c_double,
                                   c_uint64, ip, dp, POINTER(ScFilamentParams), lp], c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()
