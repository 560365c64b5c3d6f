"""compound_prism_designer -- MI355X fitness evaluator for compound-prism spectrometers.

The Python surface mirrors the reference's pyo3 module (src/pylib.rs,
compound_prism_designer/compound_prism_designer.pyi): Glass, DesignFitness,
Vector2D, UnitVector2D, UniformWavelengthDistribution, GaussianBeam,
CompoundPrism, position_detector_array, DetectorArray, Spectrometer,
RayTraceError and CudaError.  Construction runs on the host in FP64; every
fitness and probability evaluation runs on the GPU through libcpd.so
(include/cpd.h).  There is no CPU fitness path: the methods that the
reference ran on the CPU (cpu_fitness, transmission_probability) evaluate the
same estimator on the GPU.

A CompoundPrism given a curvature pair ((ip, it), (fp, ft)) is the
reference's 3D variant (toric entrance and exit faces); its spectrometers take
a FiberBeam or a GaussianBeam and run on the cpd3_* entry points.

Beyond the reference: fitness_batch() evaluates many designs in one
persistent launch, and Context selects a GPU.
"""
from __future__ import annotations

import ctypes
import math
from typing import Iterable, Optional, Sequence

import numpy as np

from . import _native as N
from ._native import Design, Design3, Fitness, Params, Params3, Sampling

__all__ = [
    "autotune",
    "build_designs",
    "gpu_fitness_batch",
    "Plan",
    "RayTraceError",
    "CudaError",
    "GpuError",
    "Glass",
    "DesignFitness",
    "Vector2D",
    "UnitVector2D",
    "UniformWavelengthDistribution",
    "GaussianBeam",
    "FiberBeam",
    "CompoundPrism",
    "position_detector_array",
    "DetectorArray",
    "Spectrometer",
    "Context",
    "fitness_batch",
    "Server",
    "generate_design",
    "generate_designs",
    "hit_fraction",
    "fitness_batch3",
    "gpu_fitness_batch3",
    "generate_designs3",
]


class RayTraceError(Exception):
    """A ray (or the design's construction) failed: src/pylib.rs:16-24."""


class CudaError(Exception):
    """A GPU error.  The reference's name (src/pylib.rs:26-35) is kept for source compatibility."""


GpuError = CudaError

_libm = ctypes.CDLL("libm.so.6")
_libm.fma.argtypes = [ctypes.c_double] * 3
_libm.fma.restype = ctypes.c_double
_fma = _libm.fma

_RAY_ERRORS = {N.ERR_NO_SURFACE_INTERSECTION, N.ERR_OUT_OF_BOUNDS, N.ERR_TOTAL_INTERNAL_REFLECTION,
               N.ERR_ANGULAR_RESPONSE_TOO_WEAK}


def _check(status: int) -> None:
    if status == N.OK:
        return
    if status in _RAY_ERRORS:
        raise RayTraceError(N.status_name(status))
    msg = N.last_error() or N.status_name(status)
    if status == N.ERR_INVALID_ARGUMENT:
        raise ValueError(msg)
    raise CudaError(f"{N.status_name(status)}: {msg}")


class Glass:
    """Glass n(l) as a degree-5 polynomial in Horner order (src/pylib.rs:37-95)."""

    ORDER = 5

    def __init__(self, name: str, coefficents: Sequence[float]):
        c = np.asarray(coefficents, dtype=np.float64).ravel()
        if c.size != self.ORDER + 1:
            raise ValueError(f"expected {self.ORDER + 1} coefficients")
        self.name = str(name)
        self._c = c

    @property
    def glass(self) -> list:
        return self._c.tolist()

    def __call__(self, wavelength: float) -> float:
        # Horner with fused multiply-adds, as Glass::calc_n (src/spectrometer/glasscat.rs:17-22)
        s = float(self._c[0])
        for c in self._c[1:]:
            s = _fma(s, float(wavelength), float(c))
        return s

    def __getnewargs__(self):
        return (self.name, self._c.copy())

    def __repr__(self):
        return f"Glass({self.name!r}, {self._c.tolist()})"


class DesignFitness:
    """size, info (bits), deviation: src/pylib.rs:103-153."""

    __slots__ = ("size", "info", "deviation")

    def __init__(self, size: float, info: float, deviation: float):
        self.size = float(size)
        self.info = float(info)
        self.deviation = float(deviation)

    def __getnewargs__(self):
        return (self.size, self.info, self.deviation)

    def __iter__(self):
        return iter((self.size, self.info, self.deviation))

    def __repr__(self):
        return f"DesignFitness(size={self.size}, info={self.info}, deviation={self.deviation})"


class Vector2D:
    def __init__(self, x: float, y: float):
        self.x, self.y = float(x), float(y)

    def __iter__(self):
        return iter((self.x, self.y))

    def __getnewargs__(self):
        return (self.x, self.y)

    def __repr__(self):
        return f"Vector2D({self.x}, {self.y})"


class UnitVector2D(Vector2D):
    def __init__(self, x: float, y: float):
        super().__init__(x, y)
        if not math.isclose(math.hypot(x, y), 1.0, rel_tol=1e-3):
            raise ValueError("not a unit vector")

    def __repr__(self):
        return f"UnitVector2D({self.x}, {self.y})"


class UniformWavelengthDistribution:
    """Uniform wavelengths in bounds (um): src/pylib.rs:293-325."""

    def __init__(self, bounds):
        lo, hi = (float(b) for b in bounds)
        if hi < lo:
            raise ValueError("lower bound > upper bound")
        self.bounds = (lo, hi)

    def __getnewargs__(self):
        return (self.bounds,)


class GaussianBeam:
    """Collimated Gaussian beam, 1/e^2 width and mean height: src/pylib.rs:351-378."""

    def __init__(self, width: float, y_mean: float):
        self.width = float(width)
        self.y_mean = float(y_mean)

    def __getnewargs__(self):
        return (self.width, self.y_mean)


class FiberBeam:
    """Multi-mode fiber beam (src/pylib.rs:380-430, spectrometer.rs:70-149): a uniform disc of
    the core radius around center_y, directions uniform over the acceptance cone of the
    numerical aperture.  3D spectrometers only (the reference's 2D fiber beam is unimplemented)."""

    def __init__(self, core_radius: float, numerical_aperture: float, center_y: float):
        self.core_radius = float(core_radius)
        self.numerical_aperture = float(numerical_aperture)
        self.center_y = float(center_y)

    @property
    def y_mean(self):
        return self.center_y

    @property
    def width(self):
        return self.core_radius


class CompoundPrism:
    """CompoundPrism(glasses, angles, lengths, curvature, height, width, ar_coated): src/pylib.rs:666-812."""

    def __init__(self, glasses: Sequence[Glass], angles: Sequence[float], lengths: Sequence[float],
                 curvature, height: float, width: float, ar_coated: bool):
        glasses = list(glasses)
        angles = [float(a) for a in angles]
        lengths = [float(v) for v in lengths]
        if len(angles) < 2:
            raise ValueError("`angles` must have at least 2 elements")
        if len(glasses) > N.MAX_PRISMS:
            raise ValueError(f"Maximum of {N.MAX_PRISMS} prisms are allowed")
        if len(glasses) != len(angles) - 1:
            raise ValueError("len(glasses) + 1 != len(angles)")
        if len(glasses) != len(lengths):
            raise ValueError("len(glasses) != len(lengths)")
        self.glasses = glasses
        self.angles = angles
        self.lengths = lengths
        if isinstance(curvature, (tuple, list)):
            # DimensionedCurvature::Dim3 (src/pylib.rs:651-655): ((signed normalised poloidal,
            # normalised toroidal) of the entrance face, the same of the exit face)
            (ip, it), (fp, ft) = curvature
            self.curvature = ((float(ip), float(it)), (float(fp), float(ft)))
            self.dim = 3
        else:
            self.curvature = float(curvature)
            self.dim = 2
        self.height = float(height)
        self.width = float(width)
        self.ar_coated = bool(ar_coated)

    def _fill(self, p) -> None:
        """Prism fields of a Params (2D) or Params3 (3D)."""
        p.n_glass = len(self.glasses)
        p.ar_coated = int(self.ar_coated)
        for i, g in enumerate(self.glasses):
            for j, c in enumerate(g.glass):
                p.glass[i][j] = c
        for i, a in enumerate(self.angles):
            p.angles[i] = a
        for i, v in enumerate(self.lengths):
            p.lengths[i] = v
        if self.dim == 3:
            (p.initial_curvature[0], p.initial_curvature[1]), (p.final_curvature[0], p.final_curvature[1]) = \
                self.curvature
        else:
            p.curvature = self.curvature
        p.height = self.height
        p.width = self.width

    def _design(self):
        p = Params3() if self.dim == 3 else Params()
        self._fill(p)
        p.bin_count = 1
        p.linear_slope = 1.0
        p.bin_size = 1.0
        if self.dim == 3:
            d = Design3()
            _check(N.lib.cpd3_build_design(ctypes.byref(p), ctypes.byref(d)))
        else:
            d = Design()
            _check(N.lib.cpd_build_design(ctypes.byref(p), ctypes.byref(d)))
        return d

    def exit_ray(self, y: float, wavelength: float):
        """The ray leaving the prism for a start ray at height y along x (its xy part in 3D):
        src/pylib.rs:763-768."""
        if self.dim == 3:
            out = (ctypes.c_double * 6)()
            _check(N.lib.cpd3_exit_ray(ctypes.byref(self._design()), y, wavelength, out))
            return Vector2D(out[0], out[1]), UnitVector2D(out[3], out[4])
        out = (ctypes.c_double * 4)()
        _check(N.lib.cpd_exit_ray(ctypes.byref(self._design()), y, wavelength, out))
        return Vector2D(out[0], out[1]), UnitVector2D(out[2], out[3])

    def final_midpt_and_normal(self):
        d = self._design()
        a = self.angles[-1]
        if self.dim == 3:
            # the arc's midpoint: the face's mid-height point moved by the sagitta against the
            # bound normal (ToricLens drawing, toric.rs:294-321)
            t = d.lens[1]
            mid = (t.bound_center[0] + t.bound_normal[0] * t.bound_height,
                   t.bound_center[1] + t.bound_normal[1] * t.bound_height)
            return Vector2D(*mid), UnitVector2D(math.cos(a), math.sin(a))
        return Vector2D(d.arc_midpt[0], d.arc_midpt[1]), UnitVector2D(math.cos(a), math.sin(a))

    def surfaces(self):
        """Drawing geometry (src/pylib.rs:795-808); drawing is not part of this build."""
        raise NotImplementedError("CompoundPrism.surfaces (drawing) is not part of the MI355X evaluator")

    def polygons(self):
        """matplotlib Paths of the prisms (src/pylib.rs:578-630); drawing is not part of this build."""
        raise NotImplementedError("CompoundPrism.polygons (drawing) is not part of the MI355X evaluator")

    def __getnewargs__(self):
        return (self.glasses, self.angles, self.lengths, self.curvature, self.height, self.width, self.ar_coated)


def _beam_fields(beam) -> tuple:
    if not isinstance(beam, GaussianBeam):
        # FiberBeam in 2D: unimplemented in the reference (spectrometer.rs:96-106)
        raise NotImplementedError("No kernel compiled for this Spectrometer")
    return beam.width, beam.y_mean


def _beam_fields3(p: Params3, beam) -> None:
    if isinstance(beam, FiberBeam):
        p.beam_kind = N.BEAM_FIBER
        p.beam_width, p.y_mean, p.numerical_aperture = beam.core_radius, beam.center_y, beam.numerical_aperture
    elif isinstance(beam, GaussianBeam):
        p.beam_kind = N.BEAM_GAUSSIAN
        p.beam_width, p.y_mean = beam.width, beam.y_mean
    else:
        raise TypeError("beam must be a GaussianBeam or a FiberBeam")


def position_detector_array(length: float, angle: float, compound_prism: CompoundPrism,
                            wavelengths: UniformWavelengthDistribution, beam, acceptance: float):
    """src/pylib.rs:814-838: ((x, y), flipped) of the detector array."""
    pos = (ctypes.c_double * 2)()
    flipped = ctypes.c_uint32()
    if compound_prism.dim == 3:
        p3 = Params3()
        compound_prism._fill(p3)
        _beam_fields3(p3, beam)
        p3.wavelength_lo, p3.wavelength_hi = wavelengths.bounds
        p3.det_length = length
        p3.det_angle = angle
        _check(N.lib.cpd3_position_detector_array(ctypes.byref(p3), acceptance, pos, ctypes.byref(flipped)))
        return (pos[0], pos[1]), bool(flipped.value)
    p = Params()
    compound_prism._fill(p)
    p.beam_width, p.y_mean = _beam_fields(beam)
    p.wavelength_lo, p.wavelength_hi = wavelengths.bounds
    p.det_length = length
    p.det_angle = angle
    _check(N.lib.cpd_position_detector_array(ctypes.byref(p), acceptance, pos, ctypes.byref(flipped)))
    return (pos[0], pos[1]), bool(flipped.value)


class DetectorArray:
    """DetectorArray(bin_count, bin_size, linear_slope, linear_intercept, length,
    max_incident_angle, angle, position, flipped): src/pylib.rs:840-955."""

    def __init__(self, bin_count: int, bin_size: float, linear_slope: float, linear_intercept: float,
                 length: float, max_incident_angle: float, angle: float, position, flipped: bool):
        self.bin_count = int(bin_count)
        self.bin_size = float(bin_size)
        self.linear_slope = float(linear_slope)
        self.linear_intercept = float(linear_intercept)
        self.length = float(length)
        self.max_incident_angle = float(max_incident_angle)
        self.angle = float(angle)
        self.position = (float(position[0]), float(position[1]))
        self.flipped = bool(flipped)

    def __getnewargs__(self):
        return (self.bin_count, self.bin_size, self.linear_slope, self.linear_intercept, self.length,
                self.max_incident_angle, self.angle, self.position, self.flipped)


class Context:
    """A device context (HIP stream and buffers) on one GPU."""

    def __init__(self, device: int = 0):
        h = ctypes.c_void_p()
        _check(N.lib.cpd_ctx_create(device, ctypes.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def info(self) -> dict:
        name = ctypes.create_string_buffer(256)
        cu, clk = ctypes.c_int(), ctypes.c_int()
        _check(N.lib.cpd_ctx_info(self._h, name, 256, ctypes.byref(cu), ctypes.byref(clk)))
        return {"name": name.value.decode(), "compute_units": cu.value, "clock_khz": clk.value}

    def set_option(self, option: int, value: int) -> None:
        _check(N.lib.cpd_ctx_set_option(self._h, option, int(value)))

    def last_stats(self) -> dict:
        st = N.RunStats()
        _check(N.lib.cpd_ctx_last_stats(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in N.RunStats._fields_}

    def set_schedule(self, schedule: Optional["N.Schedule"]) -> None:
        """Schedule of fitness_batch on this context (None: the default)."""
        _check(N.lib.cpd_ctx_set_schedule(self._h, ctypes.byref(schedule) if schedule is not None else None))

    def close(self):
        if self._h:
            N.lib.cpd_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ctx_handle(ctx: Optional[Context]):
    return ctx.handle if ctx is not None else None


class Spectrometer:
    """Spectrometer(compound_prism, detector_array, wavelengths, beam): src/pylib.rs:957-1170."""

    def __init__(self, compound_prism: CompoundPrism, detector_array: DetectorArray,
                 wavelengths: UniformWavelengthDistribution, beam):
        self.compound_prism = compound_prism
        self.detector_array = detector_array
        self.wavelengths = wavelengths
        self.beam = beam

    def __getnewargs__(self):
        return (self.compound_prism, self.detector_array, self.wavelengths, self.beam)

    @property
    def dim(self) -> int:
        return self.compound_prism.dim

    def params(self):
        """Params (2D) or Params3 (3D) of the spectrometer."""
        p = Params3() if self.dim == 3 else Params()
        self.compound_prism._fill(p)
        da = self.detector_array
        p.bin_count = da.bin_count
        p.flipped = int(da.flipped)
        p.bin_size = da.bin_size
        p.linear_slope = da.linear_slope
        p.linear_intercept = da.linear_intercept
        p.det_length = da.length
        p.max_incident_angle = da.max_incident_angle
        p.det_angle = da.angle
        p.det_position[0], p.det_position[1] = da.position
        if self.dim == 3:
            _beam_fields3(p, self.beam)
        else:
            p.beam_width, p.y_mean = _beam_fields(self.beam)
        p.wavelength_lo, p.wavelength_hi = self.wavelengths.bounds
        return p

    def design(self):
        """The constructed record: Design (2D) or Design3 (3D)."""
        if self.dim == 3:
            d = Design3()
            _check(N.lib.cpd3_build_design(ctypes.byref(self.params()), ctypes.byref(d)))
            return d
        d = Design()
        _check(N.lib.cpd_build_design(ctypes.byref(self.params()), ctypes.byref(d)))
        return d

    @property
    def position(self):
        d = self.design()
        return (d.det_position[0], d.det_position[1])

    @property
    def direction(self):
        d = self.design()
        return (d.det_direction[0], d.det_direction[1])

    def size_and_deviation(self):
        s, dv = ctypes.c_double(), ctypes.c_double()
        fn = N.lib.cpd3_size_and_deviation if self.dim == 3 else N.lib.cpd_size_and_deviation
        _check(fn(ctypes.byref(self.design()), ctypes.byref(s), ctypes.byref(dv)))
        return s.value, dv.value

    def _fitness3(self, max_n: int, max_m: int, mode: int, ctx) -> DesignFitness:
        s = Sampling(max_n, max_m, mode, 0)
        f = Fitness()
        per = ctypes.c_int()
        _check(N.lib.cpd3_fitness_batch(_ctx_handle(ctx), ctypes.byref(self.design()), ctypes.byref(s), 1,
                                        ctypes.byref(f), ctypes.byref(per)))
        _check(per.value)
        return DesignFitness(f.size, f.info, f.deviation)

    def cpu_fitness(self, *, max_n: int = 16_384, max_m: int = 16_384, ctx: Optional[Context] = None
                    ) -> DesignFitness:
        """fitness(spec, max_n, max_m) (src/pylib.rs:1008-1026, src/fitness.rs:130-147) with the
        reference's SEM early exits: rows of rays traced in FP64 on the GPU, folded exactly as
        fitness() folds them."""
        if self.dim == 3:
            return self._fitness3(max_n, max_m, N.MODE_ADAPTIVE, ctx)
        s = Sampling(max_n, max_m, N.MODE_ADAPTIVE, 0)
        f = Fitness()
        _check(N.lib.cpd_fitness_gpu(_ctx_handle(ctx), ctypes.byref(self.design()), ctypes.byref(s),
                                     ctypes.byref(f)))
        return DesignFitness(f.size, f.info, f.deviation)

    def gpu_fitness(self, seeds, *, max_n: int = 256, nwarp: int = 2, max_eval: int = 16_384,
                    ctx: Optional[Context] = None) -> Optional[DesignFitness]:
        """cuda_fitness(spec, seeds, max_n, nwarp, max_eval) (src/pylib.rs:1087-1135,
        src/cuda_fitness.rs:347-423): None when the bin errors stay above 5e-3 after every seed."""
        sd = np.ascontiguousarray(seeds, dtype=np.float64).ravel()
        f = Fitness()
        if self.dim == 3:
            if not isinstance(self.beam, FiberBeam):
                raise NotImplementedError("No kernel compiled for this Spectrometer")
            per = ctypes.c_int()
            _check(N.lib.cpd3_gpu_fitness_batch(_ctx_handle(ctx), ctypes.byref(self.design()), 1,
                                                sd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), sd.size,
                                                int(max_n), int(nwarp), int(max_eval), ctypes.byref(f),
                                                ctypes.byref(per)))
            if per.value == N.ERR_NOT_CONVERGED:
                return None
            _check(per.value)
            return DesignFitness(f.size, f.info, f.deviation)
        st = N.lib.cpd_gpu_fitness(_ctx_handle(ctx), ctypes.byref(self.design()),
                                   sd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), sd.size, int(max_n),
                                   int(nwarp), int(max_eval), ctypes.byref(f))
        if st == N.ERR_NOT_CONVERGED:
            return None
        _check(st)
        return DesignFitness(f.size, f.info, f.deviation)

    def fitness(self, *, max_n: int = 64, max_m: int = 16_384, ctx: Optional[Context] = None) -> DesignFitness:
        """Fixed-count fitness: every one of max_n wavelength samples x max_m rays traced (the
        benchmark mode; the reference's estimator with both SEM thresholds at 0)."""
        if self.dim == 3:
            return self._fitness3(max_n, max_m, N.MODE_FIXED, ctx)
        s = Sampling(max_n, max_m, N.MODE_FIXED, 0)
        f = Fitness()
        _check(N.lib.cpd_fitness_gpu(_ctx_handle(ctx), ctypes.byref(self.design()), ctypes.byref(s),
                                     ctypes.byref(f)))
        return DesignFitness(f.size, f.info, f.deviation)

    def transmission_probability(self, wavelengths, *, max_m: int = 16_384, adaptive: bool = True,
                                 ctx: Optional[Context] = None) -> np.ndarray:
        """p(d|l) rows for the given wavelengths: src/pylib.rs:1028-1058 (with the reference's SEM
        early exit per row; adaptive=False traces all max_m rays)."""
        w = np.ascontiguousarray(wavelengths, dtype=np.float64).ravel()
        out = np.empty((w.size, self.detector_array.bin_count), np.float64)
        if w.size and self.dim == 3:
            _check(N.lib.cpd3_p_dets_l_wavelength(
                _ctx_handle(ctx), ctypes.byref(self.design()), w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                w.size, max_m, N.MODE_ADAPTIVE if adaptive else N.MODE_FIXED,
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        elif w.size:
            _check(N.lib.cpd_p_dets_l_wavelength(
                _ctx_handle(ctx), ctypes.byref(self.design()), w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                w.size, max_m, N.MODE_ADAPTIVE if adaptive else N.MODE_FIXED,
                out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return out

    def ray_trace(self, wavelength: float, inital_y: float) -> np.ndarray:
        """Propagation path rows (x, y, ux, uy): src/pylib.rs:1060-1083 (2D spectrometers)."""
        if self.dim == 3:
            raise NotImplementedError("ray_trace draws 2D paths; use CompoundPrism.exit_ray for a 3D prism")
        out = np.zeros((N.MAX_PRISMS + 3, 4), np.float64)
        n = ctypes.c_uint32()
        _check(N.lib.cpd_ray_trace(ctypes.byref(self.design()), wavelength, inital_y,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(n)))
        return out[: n.value].copy()

    def propagate_rays(self, wavelength_cdf, ray_cdf, *, ctx: Optional[Context] = None):
        """Per-ray bin and transmittance (the reference's propagation test kernel).  3D: ray_cdf
        holds 4 values per ray (fiber: rho, theta, deflection, rotation; Gaussian: the first 2),
        traced in FP64."""
        if self.dim == 3:
            w = np.ascontiguousarray(wavelength_cdf, dtype=np.float64).ravel()
            r = np.ascontiguousarray(ray_cdf, dtype=np.float64).reshape(-1, 4)
            if w.size != r.shape[0]:
                raise ValueError("wavelength_cdf and ray_cdf differ in length")
            bins = np.empty(w.size, np.uint32)
            ts = np.empty(w.size, np.float64)
            if w.size:
                dp = ctypes.POINTER(ctypes.c_double)
                _check(N.lib.cpd3_propagate_rays(_ctx_handle(ctx), ctypes.byref(self.design()), w.ctypes.data_as(dp),
                                                 r.ctypes.data_as(dp), w.size,
                                                 bins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                                 ts.ctypes.data_as(dp)))
            return bins, ts
        w = np.ascontiguousarray(wavelength_cdf, dtype=np.float32).ravel()
        r = np.ascontiguousarray(ray_cdf, dtype=np.float32).ravel()
        if w.size != r.size:
            raise ValueError("wavelength_cdf and ray_cdf differ in length")
        bins = np.empty(w.size, np.uint32)
        ts = np.empty(w.size, np.float32)
        if w.size:
            fp = ctypes.POINTER(ctypes.c_float)
            _check(N.lib.cpd_propagate_rays(_ctx_handle(ctx), ctypes.byref(self.design()), w.ctypes.data_as(fp),
                                            r.ctypes.data_as(fp), w.size,
                                            bins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                            ts.ctypes.data_as(fp)))
        return bins, ts


def fitness_batch(designs: Sequence[Design], sampling, *, ctx: Optional[Context] = None):
    """Evaluate many designs in one persistent launch.

    designs: cpd_design records (Spectrometer.design() or generate_design()).
    sampling: one (n_wavelengths, n_rays) pair for all, or a sequence of Sampling.
    Returns (fitness float64 [n, 3], status int32 [n]).
    """
    n = len(designs)
    darr = (Design * n)(*designs)
    if isinstance(sampling, Sampling):
        sarr = (Sampling * n)(*([sampling] * n))
    elif isinstance(sampling, tuple) and len(sampling) == 2 and not isinstance(sampling[0], Sampling):
        s = Sampling(int(sampling[0]), int(sampling[1]), N.MODE_FIXED, 0)
        sarr = (Sampling * n)(*([s] * n))
    else:
        sarr = (Sampling * n)(*sampling)
    out = (Fitness * n)()
    status = (ctypes.c_int * n)()
    st = N.lib.cpd_fitness_batch(_ctx_handle(ctx), darr, sarr, n, out, status)
    if st != N.OK and st != N.ERR_INVALID_ARGUMENT:
        _check(st)
    res = np.array([[f.size, f.info, f.deviation] for f in out], np.float64).reshape(n, 3)
    return res, np.array(status[:], np.int32)


def gpu_fitness_batch(designs: Sequence[Design], seeds, *, max_n: int = 256, nwarp: int = 2,
                      max_eval: int = 16_384, ctx: Optional[Context] = None):
    """The reference's GPU estimator (Spectrometer.gpu_fitness) for many designs in shared launches.
    Returns (fitness float64 [n, 3], status int32 [n]); status ERR_NOT_CONVERGED where gpu_fitness
    returns None."""
    n = len(designs)
    darr = (Design * n)(*designs)
    sd = np.ascontiguousarray(seeds, dtype=np.float64).ravel()
    out = (Fitness * n)()
    status = (ctypes.c_int * n)()
    _check(N.lib.cpd_gpu_fitness_batch(_ctx_handle(ctx), darr, n, sd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                       sd.size, int(max_n), int(nwarp), int(max_eval), out, status))
    res = np.array([[f.size, f.info, f.deviation] for f in out], np.float64).reshape(n, 3)
    return res, np.array(status[:], np.int32)


def autotune(designs: Sequence[Design], sampling, *, ctx: Optional[Context] = None):
    """Time candidate schedules on this batch; returns (Schedule, milliseconds).
    Results never depend on the schedule (exact integer bin sums)."""
    n = len(designs)
    darr = (Design * n)(*designs)
    sarr = _sampling_array(sampling, n)
    best, ms = N.Schedule(), ctypes.c_double()
    _check(N.lib.cpd_autotune(_ctx_handle(ctx), darr, sarr, n, ctypes.byref(best), ctypes.byref(ms)))
    return best, ms.value


class Plan:
    """A captured generation (cpd_plan_* in include/cpd.h): per slot a fixed
    sampling and bin count.  run(designs) evaluates a generation of that shape
    by replaying one HIP graph."""

    def __init__(self, shape: Sequence[Design], sampling, *, schedule: Optional["N.Schedule"] = None,
                 ctx: Optional[Context] = None):
        self.n = len(shape)
        darr = (Design * self.n)(*shape)
        sarr = _sampling_array(sampling, self.n)
        h = ctypes.c_void_p()
        _check(N.lib.cpd_plan_create(_ctx_handle(ctx), darr, sarr, self.n,
                                     ctypes.byref(schedule) if schedule is not None else None, ctypes.byref(h)))
        self._h = h
        self._out = (Fitness * self.n)()
        self._status = (ctypes.c_int * self.n)()

    def schedule(self) -> "N.Schedule":
        sc = N.Schedule()
        _check(N.lib.cpd_plan_schedule(self._h, ctypes.byref(sc)))
        return sc

    def run_raw(self, designs_array) -> int:
        """Evaluate a ctypes array of n Design records into the plan's own
        output arrays; returns the status code (no conversion)."""
        return N.lib.cpd_plan_run(self._h, designs_array, self._out, self._status)

    def run(self, designs: Sequence[Design]):
        """Returns (fitness float64 [n, 3], status int32 [n])."""
        if len(designs) != self.n:
            raise ValueError(f"plan holds {self.n} designs, got {len(designs)}")
        darr = designs if isinstance(designs, ctypes.Array) else (Design * self.n)(*designs)
        _check(self.run_raw(darr))
        return self._results()

    def run_params(self, params: Sequence["N.Params"], acceptance: float = 1.0):
        """One generation from parameter vectors, built into designs on the GPU
        (cpd_plan_run_params): acceptance > 0 positions each detector as
        position_detector_array does, 0 takes params' position as given and
        N.POSITION_FROM_EXIT reads it as (distance, angle) from the exit face.
        Returns (fitness float64 [n, 3], status int32 [n])."""
        if len(params) != self.n:
            raise ValueError(f"plan holds {self.n} designs, got {len(params)}")
        parr = params if isinstance(params, ctypes.Array) else (N.Params * self.n)(*params)
        _check(N.lib.cpd_plan_run_params(self._h, parr, float(acceptance), self._out, self._status))
        return self._results()

    def _results(self):
        res = np.frombuffer(self._out, np.float64).reshape(self.n, 3).copy()
        return res, np.frombuffer(self._status, np.int32).copy()

    def close(self):
        if getattr(self, "_h", None):
            N.lib.cpd_plan_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_designs(params: Sequence["N.Params"], acceptance: float = 1.0, *, ctx: Optional[Context] = None):
    """Design records built on the GPU from parameter vectors
    (cpd_build_designs): acceptance > 0 positions each detector as
    position_detector_array does, 0 takes params' position as given and
    N.POSITION_FROM_EXIT reads it as (distance, angle) from the exit face
    (designer.py:236-243).  Returns
    (list of Design, status int32 [n]); a design is meaningful where its status
    is 0."""
    n = len(params)
    if n == 0:
        return [], np.zeros(0, np.int32)
    parr = params if isinstance(params, ctypes.Array) else (N.Params * n)(*params)
    out = (Design * n)()
    st = (ctypes.c_int * n)()
    _check(N.lib.cpd_build_designs(_ctx_handle(ctx), parr, n, float(acceptance), out, st))
    return list(out), np.array(st[:], np.int32)


def _sampling_array(sampling, n):
    if isinstance(sampling, Sampling):
        return (Sampling * n)(*([sampling] * n))
    if isinstance(sampling, tuple) and len(sampling) == 2 and not isinstance(sampling[0], Sampling):
        s = Sampling(int(sampling[0]), int(sampling[1]), N.MODE_FIXED, 0)
        return (Sampling * n)(*([s] * n))
    return (Sampling * n)(*sampling)


class Server:
    """Continuous-batching fitness server on one GPU (cpd_server_* in include/cpd.h).

    submit() queues a batch of designs and returns a request id; poll()/wait()
    return (fitness [n, 3], status [n], latency_ms) once the request is done.
    Requests that arrive while others run join the next scheduling step, or,
    with resident=True, the device-resident loop, which admits, traces and
    finalizes them on the GPU without a launch per request (designs of at most
    max_bins bins and max_rays rays per wavelength)."""

    def __init__(self, ctx: Optional[Context] = None, *, memory_budget: int = 0, step_items: int = 0,
                 max_designs: int = 0, resident: bool = False, max_bins: int = 0, max_rays: int = 0):
        h = ctypes.c_void_p()
        opts = N.ServerOpts(memory_budget, step_items, max_designs,
                            N.SERVER_RESIDENT if resident else N.SERVER_STEPS, max_bins, max_rays, 0)
        _check(N.lib.cpd_server_create(_ctx_handle(ctx), ctypes.byref(opts), ctypes.byref(h)))
        self._h = h
        self._sizes = {}

    def submit(self, designs: Sequence[Design], sampling) -> int:
        n = len(designs)
        darr = (Design * n)(*designs)
        sarr = _sampling_array(sampling, n)
        rid = ctypes.c_uint64()
        _check(N.lib.cpd_server_submit(self._h, darr, sarr, n, ctypes.byref(rid)))
        self._sizes[rid.value] = n
        return rid.value

    def _collect(self, rid: int, timeout_ms: Optional[int]):
        n = self._sizes[rid]
        out = (Fitness * max(n, 1))()
        status = (ctypes.c_int * max(n, 1))()
        lat = ctypes.c_double()
        if timeout_ms is None:
            st = N.lib.cpd_server_poll(self._h, rid, out, status, ctypes.byref(lat))
        else:
            st = N.lib.cpd_server_wait(self._h, rid, timeout_ms, out, status, ctypes.byref(lat))
        if st == N.ERR_NOT_READY:
            return None
        _check(st)
        del self._sizes[rid]
        res = np.array([[f.size, f.info, f.deviation] for f in out[:n]], np.float64).reshape(n, 3)
        return res, np.array(status[:n], np.int32), lat.value

    def poll(self, rid: int):
        return self._collect(rid, None)

    def wait(self, rid: int, timeout_ms: int = -1):
        return self._collect(rid, timeout_ms)

    def stats(self) -> dict:
        st = N.ServerStats()
        _check(N.lib.cpd_server_stats_get(self._h, ctypes.byref(st)))
        return {f: getattr(st, f) for f, _ in N.ServerStats._fields_}

    def close(self):
        if self._h:
            N.lib.cpd_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate_design(config: int, index: int):
    """Synthetic design `index` of workload configuration `config` (0, 1, 2, 4, or 5..7 with the
    detector and coating settings of the reference's shipped designer configurations; see
    cpd_generate_design): returns (Params, Design, Sampling, attempts)."""
    p, s, att = Params(), Sampling(), ctypes.c_uint32()
    _check(N.lib.cpd_generate_design(config, index, ctypes.byref(p), ctypes.byref(s), ctypes.byref(att)))
    d = Design()
    _check(N.lib.cpd_build_design(ctypes.byref(p), ctypes.byref(d)))
    return p, d, s, att.value


def generate_designs(config: int, first: int, n: int, threads: int = 0):
    """Designs first..first+n-1 of a benchmark configuration, generated on host
    threads: returns (params list, design list, sampling list, attempts array)."""
    P, D, S = (Params * n)(), (Design * n)(), (Sampling * n)()
    att = (ctypes.c_uint32 * n)()
    _check(N.lib.cpd_generate_designs(config, first, n, threads, P, D, S, att))
    return list(P), list(D), list(S), np.array(att[:], np.uint32)


def hit_fraction(design: Design, n_lambda: int = 4, n_rays: int = 256) -> float:
    """Fraction of rays k = 1..n_rays at wavelength samples j = 1..n_lambda that land in a bin
    (FP64 host probe; a workload statistic, see cpd_hit_fraction)."""
    out = ctypes.c_double()
    _check(N.lib.cpd_hit_fraction(ctypes.byref(design), int(n_lambda), int(n_rays), ctypes.byref(out)))
    return out.value


def generate_designs3(first: int, n: int, threads: int = 0):
    """Synthetic 3D designs first..first+n-1 (3 prisms, toric faces, fiber beam, 32 bins; see
    cpd3_generate_design): returns (Params3 list, Design3 list, Sampling list)."""
    P, D, S = (Params3 * n)(), (Design3 * n)(), (Sampling * n)()
    _check(N.lib.cpd3_generate_designs(first, n, threads, P, D, S))
    return list(P), list(D), list(S)


def fitness_batch3(designs: Sequence[Design3], sampling, *, ctx: Optional[Context] = None):
    """fitness() of many 3D designs on the GPU (cpd3_fitness_batch): sampling is one Sampling, an
    (n_wavelengths, n_rays) pair (fixed counts) or a sequence of Sampling; CPD_MODE_ADAPTIVE keeps
    the reference's early exits.  Returns (fitness float64 [n, 3], status int32 [n])."""
    n = len(designs)
    darr = designs if isinstance(designs, ctypes.Array) else (Design3 * n)(*designs)
    sarr = _sampling_array(sampling, n)
    out = (Fitness * n)()
    status = (ctypes.c_int * n)()
    st = N.lib.cpd3_fitness_batch(_ctx_handle(ctx), darr, sarr, n, out, status)
    if st != N.OK and st != N.ERR_INVALID_ARGUMENT:
        _check(st)
    res = np.array([[f.size, f.info, f.deviation] for f in out], np.float64).reshape(n, 3)
    return res, np.array(status[:], np.int32)


def gpu_fitness_batch3(designs: Sequence[Design3], seeds, *, max_n: int = 256, nwarp: int = 2,
                       max_eval: int = 16_384, ctx: Optional[Context] = None):
    """The reference's GPU estimator for fiber-beam 3D designs (cpd3_gpu_fitness_batch).
    Returns (fitness float64 [n, 3], status int32 [n])."""
    n = len(designs)
    darr = designs if isinstance(designs, ctypes.Array) else (Design3 * n)(*designs)
    sd = np.ascontiguousarray(seeds, dtype=np.float64).ravel()
    out = (Fitness * n)()
    status = (ctypes.c_int * n)()
    st = N.lib.cpd3_gpu_fitness_batch(_ctx_handle(ctx), darr, n, sd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      sd.size, int(max_n), int(nwarp), int(max_eval), out, status)
    if st != N.OK and st != N.ERR_INVALID_ARGUMENT:
        _check(st)
    res = np.array([[f.size, f.info, f.deviation] for f in out], np.float64).reshape(n, 3)
    return res, np.array(status[:], np.int32)
