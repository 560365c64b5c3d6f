"""ctypes wrapper of the CPU oracle's 3D part (the ora3_* functions of
oracle/cpd_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/ and the CPU-baseline leg of
bench.py, as the checker or the timed CPU baseline.  The product package never
imports it.

It restates the reference's 3D variant (Dim3: ToricLens entrance and exit
faces, planar interior faces bounded in height and width, FiberBeam or 3D
GaussianBeam) in FP64 as cpu_fitness runs it.  What pins each part
(tests/test_oracle3.py holds the checks):

  part                                 pinned by
  -----------------------------------  ---------------------------------------------
  approx_cbrt (cbrt.rs:49-68)           the reference's own test (cbrt.rs:97-111):
                                        relative error <= 3e-5 against cbrt on
                                        [1e-5, 1024]; odd symmetry
  sympy_soln (toric.rs:41-141)          with an exact cube root: numpy.roots of
                                        the quartic whose coefficients
                                        toric.rs:175-186 states, where it has two
                                        real roots; NaN where it has four (the
                                        resolvent cubic's casus irreducibilis,
                                        which real arithmetic cannot take) or
                                        none
  ToricSurface::intersection            with an exact cube root in place of
  (toric.rs:143-236)                    approx_cbrt: brute force -- the first
                                        sign change of the implicit torus
                                        equation (toric.rs:26-38) along the ray
                                        among in-bounds points -- and the normal
                                        parallel to the implicit function's
                                        gradient (approx_cbrt itself is pinned
                                        on its own, above)
  ToricLens::from_hyperplane            in the z = 0 plane the torus is the 2D
  (toric.rs:244-291)                    CurvedPlane arc of the same face and
                                        curvature (already pinned by the 2D
                                        oracle): same hit point and normal
  refraction in 3D (ray.rs:309-367)     Snell's law in 3D (unit direction,
                                        n1 sin i = n2 sin r, coplanarity);
                                        Fresnel at normal incidence
  PrismBounds<3> (geometry.rs:69-77)    rays through the width bound are lost
  FiberBeam / 3D GaussianBeam           closed forms of the samples (disc radius,
  (spectrometer.rs:32-46, 113-131)      cone angle, unit directions)
  Qrng<[f64; 2]>, Qrng<[f64; 4]>        the fixed-point form agrees with the
                                        iterative recurrence; exact rational
                                        frac(0.5 + k alpha_i)
  fitness / p_dets_l_wavelength         the 2D fold, which the 2D known-answer
  / the GPU estimator (3D)              test pins, over the 3D trace; I <= log2
                                        of the bins; stop points parity
                                        unpinned (no 3D golden data in the
                                        reference)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import oracle as O

MAX_PRISMS = O.MAX_PRISMS
COEFFS = O.COEFFS
BEAM_GAUSSIAN = 0
BEAM_FIBER = 1


class Params3(ctypes.Structure):
    """Mirror of cpd_params3 (include/cpd.h)."""

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
    """Mirror of cpd_design3 (include/cpd.h)."""

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


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    L = O.lib()
    P = ctypes.POINTER
    d, u32, u64, i32, sz = ctypes.c_double, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
    L.ora3_approx_cbrt.argtypes = [d]
    L.ora3_approx_cbrt.restype = d
    L.ora3_torus_roots.argtypes = [d, d, d, d, d, d, P(d)]
    L.ora3_torus_roots_exact_cbrt.argtypes = [d, d, d, d, d, d, P(d)]
    L.ora3_lens_intersection.argtypes = [P(Design3), i32, P(d), P(d), P(d), P(d), i32]
    L.ora3_refract.argtypes = [P(d), P(d), d, d, i32, P(d)]
    L.ora3_build_design.argtypes = [P(Params3), P(Design3)]
    L.ora3_position_detector_array.argtypes = [P(Params3), d, P(d), P(u32)]
    L.ora3_sample_ray.argtypes = [P(Design3), P(d), P(d)]
    L.ora3_propagate.argtypes = [P(Design3), d, P(d), P(d), P(u32), P(d)]
    L.ora3_exit_ray.argtypes = [P(Design3), d, d, P(d)]
    L.ora3_qmc.argtypes = [i32, i32, u64, P(d)]
    L.ora3_p_dets_l_wavelength.argtypes = [P(Design3), d, u64, i32, i32, P(d)]
    L.ora3_p_dets_l_wavelength.restype = u64
    L.ora3_fitness.argtypes = [P(Design3), u64, u64, i32, i32, P(d), P(u64)]
    L.ora3_fitness_many.argtypes = [P(Design3), sz, u64, u64, i32, i32, i32, P(d)]
    L.ora3_gpu_fitness.argtypes = [P(Design3), P(d), sz, u32, u32, u32, P(d), P(i32), P(u64)]
    L.ora_sizeof_params3.restype = sz
    L.ora_sizeof_design3.restype = sz
    _lib = L
    return L


def _vec(v, n=3):
    return (ctypes.c_double * n)(*[float(x) for x in v])


def approx_cbrt(x: float) -> float:
    return lib().ora3_approx_cbrt(x)


def torus_roots(r_2, R_2, dp_magn_sq, dp_v, dp_n, n_v, exact_cbrt: bool = False) -> np.ndarray:
    """toric.rs:41-141's four candidates (with approx_cbrt, or an exact cube root)."""
    out = (ctypes.c_double * 4)()
    fn = lib().ora3_torus_roots_exact_cbrt if exact_cbrt else lib().ora3_torus_roots
    fn(r_2, R_2, dp_magn_sq, dp_v, dp_n, n_v, out)
    return np.array(out[:])


def build_design(params) -> Design3:
    d = Design3()
    st = lib().ora3_build_design(ctypes.byref(O._as(params, Params3)), ctypes.byref(d))
    if st:
        raise ValueError(f"ora3_build_design: {st}")
    return d


def position_detector_array(params, acceptance: float):
    pos = (ctypes.c_double * 2)()
    fl = ctypes.c_uint32()
    st = lib().ora3_position_detector_array(ctypes.byref(O._as(params, Params3)), acceptance, pos, ctypes.byref(fl))
    return st, (pos[0], pos[1]), bool(fl.value)


def lens_intersection(design, which: int, origin, direction, exact_cbrt: bool = False):
    """(dist, normal[3]) of the ray with lens `which` (0 entrance, 1 exit), or None."""
    dist = ctypes.c_double()
    nrm = (ctypes.c_double * 3)()
    hit = lib().ora3_lens_intersection(ctypes.byref(O._as(design, Design3)), which, _vec(origin), _vec(direction),
                                       ctypes.byref(dist), nrm, int(exact_cbrt))
    return (dist.value, np.array(nrm[:])) if hit else None


def refract(direction, normal, n1: float, n2: float, ar: bool = False):
    """(status, direction[3], ts, tp)"""
    out = (ctypes.c_double * 5)()
    st = lib().ora3_refract(_vec(direction), _vec(normal), n1, n2, int(ar), out)
    return st, np.array(out[:3]), out[3], out[4]


def sample_ray(design, q) -> tuple:
    out = (ctypes.c_double * 6)()
    lib().ora3_sample_ray(ctypes.byref(O._as(design, Design3)), _vec(q, 4), out)
    return np.array(out[:3]), np.array(out[3:])


def propagate(design, wavelength: float, origin, direction):
    """(status, bin, t) of one ray."""
    b = ctypes.c_uint32()
    t = ctypes.c_double()
    st = lib().ora3_propagate(ctypes.byref(O._as(design, Design3)), wavelength, _vec(origin), _vec(direction),
                              ctypes.byref(b), ctypes.byref(t))
    return st, b.value, t.value


def exit_ray(design, y: float, wavelength: float):
    out = (ctypes.c_double * 6)()
    st = lib().ora3_exit_ray(ctypes.byref(O._as(design, Design3)), y, wavelength, out)
    return st, np.array(out[:])


def qmc(dims: int, n: int, kind: int = O.QRNG_FIXED_POINT) -> np.ndarray:
    out = np.zeros((n, 4), np.float64)
    lib().ora3_qmc(dims, kind, n, O._ptr(out, ctypes.c_double))
    return out[:, :dims]


def p_dets_l_wavelength(design, wavelength: float, max_m: int, adaptive: bool = True,
                        qrng: int = O.QRNG_FIXED_POINT):
    """(p(d|l) [bin_count], rays traced)"""
    design = O._as(design, Design3)
    out = np.empty(design.bin_count, np.float64)
    rays = lib().ora3_p_dets_l_wavelength(ctypes.byref(design), wavelength, max_m, int(adaptive), qrng,
                                          O._ptr(out, ctypes.c_double))
    return out, rays


def fitness(design, max_n: int, max_m: int, adaptive: bool = True, qrng: int = O.QRNG_FIXED_POINT):
    """((size, info, deviation), (rows, rays))"""
    out = (ctypes.c_double * 3)()
    cnt = (ctypes.c_uint64 * 2)()
    lib().ora3_fitness(ctypes.byref(O._as(design, Design3)), max_n, max_m, int(adaptive), qrng, out, cnt)
    return (out[0], out[1], out[2]), (cnt[0], cnt[1])


def fitness_many(designs, max_n: int, max_m: int, adaptive: bool = True, qrng: int = O.QRNG_FIXED_POINT,
                 threads: int = 0) -> np.ndarray:
    n = len(designs)
    arr = (Design3 * n)(*[O._as(d, Design3) for d in designs])
    out = np.zeros((n, 3), np.float64)
    lib().ora3_fitness_many(arr, n, max_n, max_m, int(adaptive), qrng, threads or (os.cpu_count() or 1),
                            O._ptr(out, ctypes.c_double))
    return out


def gpu_fitness(design, seeds, max_n: int = 256, nwarp: int = 2, max_eval: int = 16384):
    """((size, info, deviation), converged, (rows, rays)) of the GPU estimator in FP64."""
    sd = np.ascontiguousarray(seeds, dtype=np.float64).ravel()
    out = (ctypes.c_double * 3)()
    conv = ctypes.c_int()
    cnt = (ctypes.c_uint64 * 2)()
    st = lib().ora3_gpu_fitness(ctypes.byref(O._as(design, Design3)), O._ptr(sd, ctypes.c_double), sd.size, max_n,
                                nwarp, max_eval, out, ctypes.byref(conv), cnt)
    if st:
        raise ValueError(f"ora3_gpu_fitness: {st}")
    return (out[0], out[1], out[2]), bool(conv.value), (cnt[0], cnt[1])
