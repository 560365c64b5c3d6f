"""ctypes wrapper of the CPU oracle (oracle/cpd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker or the timed CPU baseline.  The
product package never imports it.

The oracle restates jdtatz/compound-prism-designer's FP64/FP32 CPU fitness
path (src/fitness.rs and src/spectrometer/*, 2D variant).  What pins each part
(tests/test_oracle.py holds the checks):

  part                               pinned by
  ---------------------------------  ---------------------------------------------
  whole path, adaptive FP64           known-answer test src/fitness.rs:153-258
                                      (size 41.3 +-0.5%, info 0.7407 +-1%,
                                      deviation 0.377159 +-0.1%)
  whole path, FP32 vs FP64            benches/benchmark.rs:147-157 (info rmax 1e-2)
  norminv (erf.rs:62-73)              scipy.special.ndtri at erf.rs:142-161 points,
                                      tolerances of erf.rs
  QMC (qrng.rs)                       closed form frac(0.5 + k/phi), exact decimal
  Welford / plog2p / MI (fitness.rs)  brute force with fractions on tiny inputs;
                                      I = 0 when p(d|l) does not depend on l;
                                      0 <= I <= log2(nbin)
  refraction / Fresnel (ray.rs)       Snell's law, normal-incidence Fresnel
                                      ((n-1)/(n+1))^2, TIR threshold
  planes / arc / detector             hand-built single-face geometries with
                                      closed-form hit points and bins
  construction (geometry.rs)          closed-form trapezoid geometry, arc sagitta
  detector fit (geometry.rs:495-542)  both end rays land on the array ends
  anti-reflection branch (ray.rs:      T = 0.99 per coated face at normal
  336-340)                            incidence; the strict ci > 0.5 threshold
                                      on either side of 60 degrees
  gap bins, intercept (spectrometer.  a flat geometry where a ray in a gap
  rs:291-303)                         misses, the intercept shifts the bins and
                                      a position before it saturates to bin 0
  early-exit stop points              closed forms: one bin with constant t
  (fitness.rs:20-28,108-114,          stops at k + 1 > t / e (the phantom zero
  welford.rs:34-58)                   and skip's post-update mean, which a
                                      textbook skip would move); two split bins
                                      in exact decimals; two equal or all-zero
                                      rows stop the outer loop
  GPU estimator (kernel.rs:89-193,    a row equals the direct mean of t over the
  cuda_fitness.rs:347-423)            rays it folded (brute force per ray); the
                                      32-lane chunk (count 32, mean, m2) merge
                                      and its stop on chunks of known bins; the
                                      KAT's info within the bench's 1e-2
                                      (benches/benchmark.rs:147-157).  Unpinned:
                                      the reference's own FP32 fast-math bits
                                      (FastFloat is not vendored); this
                                      restatement folds in FP64
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libcpd_oracle.so"

MAX_PRISMS = 6
COEFFS = 6


class Params(ctypes.Structure):
    """Mirror of cpd_params (include/cpd.h)."""

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
    """Mirror of cpd_design (include/cpd.h)."""

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


QRNG_ITERATIVE = 0
QRNG_FIXED_POINT = 1

_lib = None


def build() -> Path:
    """Compile libcpd_oracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        build()
    L = ctypes.CDLL(str(LIB_PATH))
    P = ctypes.POINTER
    d, u32, u64, i32, sz = ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
    f32 = ctypes.c_float
    L.ora_build_design.argtypes = [P(Params), P(Design)]
    L.ora_build_design_f32.argtypes = [P(Params), d, P(Design)]
    L.ora_position_detector_array.argtypes = [P(Params), d, P(d), P(u32)]
    L.ora_size_and_deviation.argtypes = [P(Design), P(d), P(d)]
    L.ora_norminv_f64.argtypes = [d]
    L.ora_norminv_f64.restype = d
    L.ora_norminv_many.argtypes = [sz, P(d), P(d)]
    L.ora_norminv_f32.argtypes = [f32]
    L.ora_norminv_f32.restype = f32
    L.ora_qmc_f64.argtypes = [u64, i32, P(d)]
    L.ora_qmc_f32.argtypes = [u64, i32, P(f32)]
    L.ora_propagate_f64.argtypes = [P(Design), d, d, P(u32), P(d)]
    L.ora_propagate_f32.argtypes = [P(Design), f32, f32, P(u32), P(f32)]
    L.ora_propagate_many.argtypes = [P(Design), i32, sz, P(d), P(d), P(u32), P(d)]
    L.ora_propagation_test_f32.argtypes = [P(Design), sz, P(f32), P(f32), P(u32), P(f32)]
    L.ora_ray_trace_f64.argtypes = [P(Design), d, d, P(d)]
    L.ora_ray_trace_f64.restype = sz
    L.ora_p_dets_l_wavelength.argtypes = [P(Design), d, u64, i32, i32, i32, P(d), P(u64)]
    L.ora_fitness.argtypes = [P(Design), u64, u64, i32, i32, i32, P(d), P(u64)]
    L.ora_fitness_counts.argtypes = [P(Design), u64, u64, i32, i32, i32, P(d), P(u64)]
    L.ora_fitness_many.argtypes = [P(Design), P(u32), P(u32), sz, i32, i32, i32, i32, P(d)]
    L.ora_fitness_many.restype = u64
    L.ora_gpu_fitness.argtypes = [P(Design), P(d), sz, u32, u32, u32, P(d), P(i32), P(u64)]
    L.ora_gpu_row.argtypes = [P(Design), d, d, u32, P(d)]
    L.ora_gpu_row.restype = u64
    L.ora_sizeof_params.restype = sz
    L.ora_sizeof_design.restype = sz
    _lib = L
    return L


def _as(obj, cls):
    """Reinterpret a same-layout ctypes struct (e.g. the package's binding of
    cpd_params / cpd_design) as this module's mirror type."""
    if isinstance(obj, cls):
        return obj
    if ctypes.sizeof(obj) != ctypes.sizeof(cls):
        raise TypeError(f"{type(obj).__name__} does not have the layout of {cls.__name__}")
    out = cls()
    ctypes.memmove(ctypes.byref(out), ctypes.byref(obj), ctypes.sizeof(cls))
    return out


def _ptr(a: np.ndarray, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def build_design(params: Params) -> Design:
    params = _as(params, Params)
    d = Design()
    st = lib().ora_build_design(ctypes.byref(params), ctypes.byref(d))
    if st != 0:
        raise ValueError(f"oracle build_design status {st}")
    return d


def build_design_f32(params: Params, acceptance: float) -> Design:
    params = _as(params, Params)
    d = Design()
    st = lib().ora_build_design_f32(ctypes.byref(params), acceptance, ctypes.byref(d))
    if st != 0:
        raise ValueError(f"oracle build_design_f32 status {st}")
    return d


def position_detector_array(params: Params, acceptance: float):
    params = _as(params, Params)
    pos = (ctypes.c_double * 2)()
    flipped = ctypes.c_uint32()
    st = lib().ora_position_detector_array(ctypes.byref(params), acceptance, pos, ctypes.byref(flipped))
    return st, (pos[0], pos[1]), bool(flipped.value)


def size_and_deviation(design: Design):
    design = _as(design, Design)
    s, dv = ctypes.c_double(), ctypes.c_double()
    lib().ora_size_and_deviation(ctypes.byref(design), ctypes.byref(s), ctypes.byref(dv))
    return s.value, dv.value


def norminv(u: float, precision: int = 64) -> float:
    return lib().ora_norminv_f64(u) if precision == 64 else lib().ora_norminv_f32(u)


def norminv_many(u) -> np.ndarray:
    """norminv (erf.rs:62-73) in FP64 of every element of u."""
    u = np.ascontiguousarray(u, np.float64)
    out = np.empty_like(u)
    lib().ora_norminv_many(u.size, _ptr(u, ctypes.c_double), _ptr(out, ctypes.c_double))
    return out


def qmc(n: int, kind: int = QRNG_ITERATIVE, precision: int = 64) -> np.ndarray:
    if precision == 64:
        out = np.empty(n, np.float64)
        lib().ora_qmc_f64(n, kind, _ptr(out, ctypes.c_double))
    else:
        out = np.empty(n, np.float32)
        lib().ora_qmc_f32(n, kind, _ptr(out, ctypes.c_float))
    return out


def propagate(design: Design, wavelength: float, y: float, precision: int = 64):
    """(status, bin, t) of one ray starting at (0, y)."""
    design = _as(design, Design)
    b = ctypes.c_uint32()
    if precision == 64:
        t = ctypes.c_double()
        st = lib().ora_propagate_f64(ctypes.byref(design), wavelength, y, ctypes.byref(b), ctypes.byref(t))
    else:
        t = ctypes.c_float()
        st = lib().ora_propagate_f32(ctypes.byref(design), wavelength, y, ctypes.byref(b), ctypes.byref(t))
    return st, b.value, t.value


def propagate_many(design: Design, wavelength, y, precision: int = 64):
    design = _as(design, Design)
    w = np.ascontiguousarray(np.broadcast_to(np.asarray(wavelength, np.float64), np.shape(y)))
    y = np.ascontiguousarray(y, np.float64)
    n = y.size
    bins = np.empty(n, np.uint32)
    ts = np.empty(n, np.float64)
    lib().ora_propagate_many(ctypes.byref(design), precision, n, _ptr(w, ctypes.c_double),
                             _ptr(y, ctypes.c_double), _ptr(bins, ctypes.c_uint32), _ptr(ts, ctypes.c_double))
    return bins, ts


def propagation_test_f32(design: Design, wavelength_cdf: np.ndarray, ray_cdf: np.ndarray):
    design = _as(design, Design)
    w = np.ascontiguousarray(wavelength_cdf, np.float32)
    r = np.ascontiguousarray(ray_cdf, np.float32)
    n = w.size
    bins = np.empty(n, np.uint32)
    ts = np.empty(n, np.float32)
    lib().ora_propagation_test_f32(ctypes.byref(design), n, _ptr(w, ctypes.c_float), _ptr(r, ctypes.c_float),
                                   _ptr(bins, ctypes.c_uint32), _ptr(ts, ctypes.c_float))
    return bins, ts


def ray_trace(design: Design, wavelength: float, y: float) -> np.ndarray:
    design = _as(design, Design)
    out = np.zeros((MAX_PRISMS + 3, 4), np.float64)
    k = lib().ora_ray_trace_f64(ctypes.byref(design), wavelength, y, _ptr(out, ctypes.c_double))
    return out[:k].copy()


def p_dets_l_wavelength(design: Design, wavelength: float, max_m: int, precision: int = 64,
                        adaptive: bool = False, qrng: int = QRNG_ITERATIVE):
    design = _as(design, Design)
    out = np.empty(design.bin_count, np.float64)
    rays = ctypes.c_uint64()
    lib().ora_p_dets_l_wavelength(ctypes.byref(design), wavelength, max_m, precision, int(adaptive), qrng,
                                  _ptr(out, ctypes.c_double), ctypes.byref(rays))
    return out, rays.value


def fitness(design: Design, max_n: int, max_m: int, precision: int = 64, adaptive: bool = False,
            qrng: int = QRNG_ITERATIVE):
    """(size, info, deviation), rays traced."""
    design = _as(design, Design)
    out = (ctypes.c_double * 3)()
    rays = ctypes.c_uint64()
    lib().ora_fitness(ctypes.byref(design), max_n, max_m, precision, int(adaptive), qrng, out, ctypes.byref(rays))
    return (out[0], out[1], out[2]), rays.value


def fitness_counts(design: Design, max_n: int, max_m: int, precision: int = 64, adaptive: bool = True,
                   qrng: int = QRNG_ITERATIVE):
    """(size, info, deviation), (wavelength rows, rays) of one evaluation."""
    design = _as(design, Design)
    out = (ctypes.c_double * 3)()
    cnt = (ctypes.c_uint64 * 2)()
    lib().ora_fitness_counts(ctypes.byref(design), max_n, max_m, precision, int(adaptive), qrng, out, cnt)
    return (out[0], out[1], out[2]), (cnt[0], cnt[1])


def gpu_fitness(design: Design, seeds, max_n: int = 256, nwarp: int = 2, max_eval: int = 16384):
    """The reference's GPU estimator (kernel.rs:89-193 + cuda_fitness.rs:347-423) in FP64:
    ((size, info, deviation), converged, (rows, rays))."""
    design = _as(design, Design)
    sd = np.ascontiguousarray(seeds, dtype=np.float64).ravel()
    out = (ctypes.c_double * 3)()
    conv = ctypes.c_int()
    cnt = (ctypes.c_uint64 * 2)()
    lib().ora_gpu_fitness(ctypes.byref(design), _ptr(sd, ctypes.c_double), sd.size, max_n, nwarp, max_eval, out,
                          ctypes.byref(conv), cnt)
    return (out[0], out[1], out[2]), bool(conv.value), (cnt[0], cnt[1])


def gpu_row(design: Design, wavelength: float, seed: float, max_eval: int):
    """One wavelength row of the GPU estimator: (p(d|l) [bin_count], rays folded)."""
    design = _as(design, Design)
    out = np.empty(design.bin_count, np.float64)
    rays = lib().ora_gpu_row(ctypes.byref(design), wavelength, seed, max_eval, _ptr(out, ctypes.c_double))
    return out, rays


def fitness_many(designs, max_n, max_m, precision: int = 64, adaptive: bool = False,
                 qrng: int = QRNG_FIXED_POINT, threads: int = 0):
    """Fitness of many designs over a host thread pool; returns (out[n,3], rays)."""
    n = len(designs)
    arr = (Design * n)(*[_as(d, Design) for d in designs])
    mn = np.ascontiguousarray(np.broadcast_to(np.asarray(max_n, np.uint32), (n,)))
    mm = np.ascontiguousarray(np.broadcast_to(np.asarray(max_m, np.uint32), (n,)))
    out = np.zeros((n, 3), np.float64)
    if threads <= 0:
        threads = os.cpu_count() or 1
    rays = lib().ora_fitness_many(arr, _ptr(mn, ctypes.c_uint32), _ptr(mm, ctypes.c_uint32), n, precision,
                                  int(adaptive), qrng, threads, _ptr(out, ctypes.c_double))
    return out, rays
