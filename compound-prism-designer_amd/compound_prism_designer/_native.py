"""ctypes binding of libcpd.so (include/cpd.h).

The library is loaded from this package directory.  There is no fallback:
if libcpd.so is missing or fails to load, importing the package raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

# CPD_LIBRARY selects an alternate build of libcpd.so (same ABI).
LIB_PATH = Path(os.environ.get("CPD_LIBRARY") or Path(__file__).resolve().parent / "libcpd.so")

MAX_PRISMS = 6
COEFFS = 6
MAX_BINS = 1024

OK = 0
ERR_NO_SURFACE_INTERSECTION = 1
ERR_OUT_OF_BOUNDS = 2
ERR_TOTAL_INTERNAL_REFLECTION = 3
ERR_ANGULAR_RESPONSE_TOO_WEAK = 4
ERR_INVALID_ARGUMENT = 16
ERR_NOT_CONVERGED = 17
ERR_NOT_READY = 18
ERR_NO_DEVICE = 32
ERR_DEVICE = 33
ERR_OUT_OF_MEMORY = 34
ERR_CAPACITY = 35

NOALLOC_MAX_DESIGNS = 4096
NOALLOC_MAX_RAYS = 65536
NOALLOC_MAX_ROWS = 1 << 18
NOALLOC_MAX_CELLS = 1 << 21

MODE_FIXED = 0
MODE_ADAPTIVE = 1

# acceptance of cpd_build_designs / cpd_plan_run_params: det_position is
# (distance, angle) from the exit face (designer.py:236-243)
POSITION_FROM_EXIT = -1.0

BEAM_GAUSSIAN = 0
BEAM_FIBER = 1


class Params(ctypes.Structure):
    _fields_ = [
        ("n_glass", ctypes.c_uint32),
        ("ar_coated", ctypes.c_uint32),
        ("glass", (ctypes.c_double * COEFFS) * MAX_PRISMS),
        ("angles", ctypes.c_double * (MAX_PRISMS + 1)),
        ("lengths", ctypes.c_double * MAX_PRISMS),
        ("curvature", ctypes.c_double),
        ("height", ctypes.c_double),
        ("width", ctypes.c_double),
        ("bin_count", ctypes.c_uint32),
        ("flipped", ctypes.c_uint32),
        ("bin_size", ctypes.c_double),
        ("linear_slope", ctypes.c_double),
        ("linear_intercept", ctypes.c_double),
        ("det_length", ctypes.c_double),
        ("max_incident_angle", ctypes.c_double),
        ("det_angle", ctypes.c_double),
        ("det_position", ctypes.c_double * 2),
        ("beam_width", ctypes.c_double),
        ("y_mean", ctypes.c_double),
        ("wavelength_lo", ctypes.c_double),
        ("wavelength_hi", ctypes.c_double),
    ]


class Design(ctypes.Structure):
    _fields_ = [
        ("n_glass", ctypes.c_uint32),
        ("bin_count", ctypes.c_uint32),
        ("ar_coated", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("glass", (ctypes.c_double * COEFFS) * MAX_PRISMS),
        ("plane_point", (ctypes.c_double * 2) * MAX_PRISMS),
        ("plane_normal", (ctypes.c_double * 2) * MAX_PRISMS),
        ("height", ctypes.c_double),
        ("arc_center", ctypes.c_double * 2),
        ("arc_radius", ctypes.c_double),
        ("arc_midpt", ctypes.c_double * 2),
        ("arc_max_dist_sq", ctypes.c_double),
        ("det_position", ctypes.c_double * 2),
        ("det_direction", ctypes.c_double * 2),
        ("det_normal", ctypes.c_double * 2),
        ("det_min_ci", ctypes.c_double),
        ("det_length", ctypes.c_double),
        ("det_slope", ctypes.c_double),
        ("det_intercept", ctypes.c_double),
        ("det_bin_size", ctypes.c_double),
        ("beam_width", ctypes.c_double),
        ("y_mean", ctypes.c_double),
        ("wavelength_lo", ctypes.c_double),
        ("wavelength_hi", ctypes.c_double),
    ]


class Params3(ctypes.Structure):
    """cpd_params3: a 3D spectrometer's parameters (toric lenses, fiber or Gaussian beam)."""

    _fields_ = [
        ("n_glass", ctypes.c_uint32),
        ("ar_coated", ctypes.c_uint32),
        ("glass", (ctypes.c_double * COEFFS) * MAX_PRISMS),
        ("angles", ctypes.c_double * (MAX_PRISMS + 1)),
        ("lengths", ctypes.c_double * MAX_PRISMS),
        ("initial_curvature", ctypes.c_double * 2),
        ("final_curvature", ctypes.c_double * 2),
        ("height", ctypes.c_double),
        ("width", ctypes.c_double),
        ("bin_count", ctypes.c_uint32),
        ("flipped", ctypes.c_uint32),
        ("bin_size", ctypes.c_double),
        ("linear_slope", ctypes.c_double),
        ("linear_intercept", ctypes.c_double),
        ("det_length", ctypes.c_double),
        ("max_incident_angle", ctypes.c_double),
        ("det_angle", ctypes.c_double),
        ("det_position", ctypes.c_double * 2),
        ("beam_kind", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("beam_width", ctypes.c_double),
        ("y_mean", ctypes.c_double),
        ("numerical_aperture", ctypes.c_double),
        ("wavelength_lo", ctypes.c_double),
        ("wavelength_hi", ctypes.c_double),
    ]


class Torus(ctypes.Structure):
    _fields_ = [
        ("center", ctypes.c_double * 2),
        ("toroidal_normal", ctypes.c_double * 2),
        ("toroidal_radius", ctypes.c_double),
        ("poloidal_radius", ctypes.c_double),
        ("bound_center", ctypes.c_double * 2),
        ("bound_normal", ctypes.c_double * 2),
        ("bound_height", ctypes.c_double),
    ]


class Design3(ctypes.Structure):
    """cpd_design3: a constructed 3D spectrometer."""

    _fields_ = [
        ("n_glass", ctypes.c_uint32),
        ("bin_count", ctypes.c_uint32),
        ("ar_coated", ctypes.c_uint32),
        ("beam_kind", ctypes.c_uint32),
        ("glass", (ctypes.c_double * COEFFS) * MAX_PRISMS),
        ("lens", Torus * 2),
        ("plane_point", (ctypes.c_double * 2) * MAX_PRISMS),
        ("plane_normal", (ctypes.c_double * 2) * MAX_PRISMS),
        ("height", ctypes.c_double),
        ("half_width", ctypes.c_double),
        ("det_position", ctypes.c_double * 2),
        ("det_direction", ctypes.c_double * 2),
        ("det_normal", ctypes.c_double * 2),
        ("det_min_ci", ctypes.c_double),
        ("det_length", ctypes.c_double),
        ("det_slope", ctypes.c_double),
        ("det_intercept", ctypes.c_double),
        ("det_bin_size", ctypes.c_double),
        ("beam_width", ctypes.c_double),
        ("y_mean", ctypes.c_double),
        ("fiber_min_cos", ctypes.c_double),
        ("wavelength_lo", ctypes.c_double),
        ("wavelength_hi", ctypes.c_double),
    ]


class Sampling(ctypes.Structure):
    _fields_ = [
        ("n_wavelengths", ctypes.c_uint32),
        ("n_rays", ctypes.c_uint32),
        ("mode", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class RunStats(ctypes.Structure):
    _fields_ = [
        ("trace_ms", ctypes.c_double),
        ("finalize_ms", ctypes.c_double),
        ("ops", ctypes.c_uint64),
        ("items", ctypes.c_uint64),
        ("rays", ctypes.c_uint64),
        ("grid", ctypes.c_uint32),
        ("tile_rays", ctypes.c_uint32),
    ]


OPT_PROFILE = 1
OPT_COUNT_OPS = 2
OPT_TILE_RAYS = 3
OPT_WAVES_PER_CU = 4
OPT_ORDER = 5
OPT_ACCUMULATE = 6
OPT_ROW_TRACE = 7

ACC_EXACT = 0
ACC_F32 = 1
ACC_F32C = 2
ACC_F64 = 3

ORDER_GIVEN = 0
ORDER_COST = 1


class Schedule(ctypes.Structure):
    _fields_ = [("tile_rays", ctypes.c_uint32), ("waves_per_cu", ctypes.c_uint32), ("order", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]

    def as_dict(self) -> dict:
        return {"tile_rays": self.tile_rays, "waves_per_cu": self.waves_per_cu,
                "order": "cost" if self.order == ORDER_COST else "given"}


SERVER_STEPS = 0
SERVER_RESIDENT = 1


class ServerOpts(ctypes.Structure):
    _fields_ = [("memory_budget", ctypes.c_uint64), ("step_items", ctypes.c_uint32), ("max_designs", ctypes.c_uint32),
                ("mode", ctypes.c_uint32), ("max_bins", ctypes.c_uint32), ("max_rays", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("rejected", ctypes.c_uint64),
        ("steps", ctypes.c_uint64), ("items", ctypes.c_uint64), ("designs", ctypes.c_uint64),
        ("queued", ctypes.c_uint32), ("admitted", ctypes.c_uint32), ("busy_ms", ctypes.c_double),
    ]


class Fitness(ctypes.Structure):
    _fields_ = [("size", ctypes.c_double), ("info", ctypes.c_double), ("deviation", ctypes.c_double)]


def _load() -> ctypes.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python compound-prism-designer_amd/build.py` "
            "(or __graft_entry__.build()) first"
        )
    L = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    st = ctypes.c_int
    d, u32, u64, sz, i32 = ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_size_t, ctypes.c_int
    vp = ctypes.c_void_p
    sigs = {
        "cpd_abi_version": ([], i32),
        "cpd_status_name": ([st], ctypes.c_char_p),
        "cpd_last_error": ([], ctypes.c_char_p),
        "cpd_position_detector_array": ([P(Params), d, P(d), P(u32)], st),
        "cpd_build_design": ([P(Params), P(Design)], st),
        "cpd_size_and_deviation": ([P(Design), P(d), P(d)], st),
        "cpd_exit_ray": ([P(Design), d, d, P(d)], st),
        "cpd_ray_trace": ([P(Design), d, d, P(d), P(u32)], st),
        "cpd_fit_sellmeier1": ([P(d), P(d), d, d, u32, P(d), P(d)], st),
        "cpd_fit_sellmeier1_batch": ([vp, P(d), P(d), sz, d, d, u32, P(d), P(d), P(st)], st),
        "cpd_generate_design": ([u32, u64, P(Params), P(Sampling), P(u32)], st),
        "cpd_generate_designs": ([u32, u64, sz, u32, P(Params), P(Design), P(Sampling), P(u32)], st),
        "cpd_hit_fraction": ([P(Design), u32, u32, P(d)], st),
        "cpd_ctx_create": ([i32, P(vp)], st),
        "cpd_ctx_destroy": ([vp], None),
        "cpd_ctx_info": ([vp, ctypes.c_char_p, sz, P(i32), P(i32)], st),
        "cpd_ctx_set_option": ([vp, i32, ctypes.c_int64], st),
        "cpd_ctx_last_stats": ([vp, P(RunStats)], st),
        "cpd_server_create": ([vp, P(ServerOpts), P(vp)], st),
        "cpd_server_submit": ([vp, P(Design), P(Sampling), sz, P(u64)], st),
        "cpd_server_poll": ([vp, u64, P(Fitness), P(st), P(d)], st),
        "cpd_server_wait": ([vp, u64, ctypes.c_int64, P(Fitness), P(st), P(d)], st),
        "cpd_server_stats_get": ([vp, P(ServerStats)], st),
        "cpd_server_destroy": ([vp], None),
        "cpd_fitness_gpu": ([vp, P(Design), P(Sampling), P(Fitness)], st),
        "cpd_fitness_batch": ([vp, P(Design), P(Sampling), sz, P(Fitness), P(st)], st),
        "cpd_fitness_batch_noalloc": ([vp, P(Design), P(Sampling), sz, P(Fitness), P(st)], st),
        "cpd_p_dets_l_wavelength": ([vp, P(Design), P(d), sz, u32, u32, P(d)], st),
        "cpd_propagate_rays": ([vp, P(Design), P(ctypes.c_float), P(ctypes.c_float), sz, P(u32),
                                P(ctypes.c_float)], st),
        "cpd_gpu_fitness": ([vp, P(Design), P(d), sz, u32, u32, u32, P(Fitness)], st),
        "cpd_gpu_fitness_batch": ([vp, P(Design), sz, P(d), sz, u32, u32, u32, P(Fitness), P(st)], st),
        "cpd_autotune": ([vp, P(Design), P(Sampling), sz, P(Schedule), P(d)], st),
        "cpd_ctx_set_schedule": ([vp, P(Schedule)], st),
        "cpd_plan_create": ([vp, P(Design), P(Sampling), sz, P(Schedule), P(vp)], st),
        "cpd_plan_run": ([vp, P(Design), P(Fitness), P(st)], st),
        "cpd_plan_run_params": ([vp, P(Params), d, P(Fitness), P(st)], st),
        "cpd_build_designs": ([vp, P(Params), sz, d, P(Design), P(st)], st),
        "cpd_plan_schedule": ([vp, P(Schedule)], st),
        "cpd_plan_destroy": ([vp], None),
        "cpd3_build_design": ([P(Params3), P(Design3)], st),
        "cpd3_position_detector_array": ([P(Params3), d, P(d), P(u32)], st),
        "cpd3_size_and_deviation": ([P(Design3), P(d), P(d)], st),
        "cpd3_exit_ray": ([P(Design3), d, d, P(d)], st),
        "cpd3_generate_design": ([u64, P(Params3), P(Sampling), P(u32)], st),
        "cpd3_generate_designs": ([u64, sz, u32, P(Params3), P(Design3), P(Sampling)], st),
        "cpd3_fitness_batch": ([vp, P(Design3), P(Sampling), sz, P(Fitness), P(st)], st),
        "cpd3_p_dets_l_wavelength": ([vp, P(Design3), P(d), sz, u32, u32, P(d)], st),
        "cpd3_gpu_fitness_batch": ([vp, P(Design3), sz, P(d), sz, u32, u32, u32, P(Fitness), P(st)], st),
        "cpd3_propagate_rays": ([vp, P(Design3), P(d), P(d), sz, P(u32), P(d)], st),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    if L.cpd_abi_version() != 1:
        raise ImportError("libcpd.so ABI version mismatch")
    return L


lib = _load()


def status_name(status: int) -> str:
    return lib.cpd_status_name(status).decode()


def last_error() -> str:
    return lib.cpd_last_error().decode()
