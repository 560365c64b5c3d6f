"""Host-side 3D code (no GPU): construction, detector positioning and exit
rays of the Dim3 spectrometer against the CPU oracle (bit for bit), and the
Python surface for 3D prisms and fiber beams."""
import ctypes

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O
from oracle import oracle3 as O3


def test_struct_layouts_agree():
    assert ctypes.sizeof(N.Params3) == O3.lib().ora_sizeof_params3()
    assert ctypes.sizeof(N.Design3) == O3.lib().ora_sizeof_design3()
    assert [f[0] for f in N.Params3._fields_] == [f[0] for f in O3.Params3._fields_]
    assert [f[0] for f in N.Design3._fields_] == [f[0] for f in O3.Design3._fields_]


def test_construction_and_positioning_match_oracle():
    P, D, S = cpd.generate_designs3(0, 24)
    for p, d in zip(P, D):
        assert bytes(d) == bytes(O3.build_design(p))
        st, pos, flipped = O3.position_detector_array(p, 1.0)
        assert st == 0 and pos == (p.det_position[0], p.det_position[1]) and flipped == bool(p.flipped)
        s, dv = ctypes.c_double(), ctypes.c_double()
        assert N.lib.cpd3_size_and_deviation(ctypes.byref(d), ctypes.byref(s), ctypes.byref(dv)) == 0
        assert s.value > 0 and 0 <= dv.value <= 1
        w = 0.5 * (p.wavelength_lo + p.wavelength_hi)
        out = (ctypes.c_double * 6)()
        st_p = N.lib.cpd3_exit_ray(ctypes.byref(d), p.y_mean, w, out)
        st_o, ref = O3.exit_ray(d, p.y_mean, w)
        assert st_p == st_o
        if st_p == 0:
            assert list(out) == list(ref)
    assert all(s.n_wavelengths == 64 and s.n_rays == 4096 for s in S)


def test_generator_is_deterministic():
    a = cpd.generate_designs3(5, 3)[1]
    b = cpd.generate_designs3(5, 3, threads=1)[1]
    assert [bytes(x) for x in a] == [bytes(x) for x in b]


def test_invalid_parameters():
    P, _, _ = cpd.generate_designs3(0, 1)
    p = P[0]
    d = N.Design3()
    for field, value in (("n_glass", 0), ("height", -1.0), ("width", 0.0), ("bin_count", 0),
                         ("linear_slope", 0.0), ("beam_kind", 7)):
        q = N.Params3()
        ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
        setattr(q, field, value)
        assert N.lib.cpd3_build_design(ctypes.byref(q), ctypes.byref(d)) == N.ERR_INVALID_ARGUMENT, field
    q = N.Params3()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
    q.final_curvature[1] = 0.0
    assert N.lib.cpd3_build_design(ctypes.byref(q), ctypes.byref(d)) == N.ERR_INVALID_ARGUMENT


def spectrometer_from(p, beam=None):
    glasses = [cpd.Glass(f"g{i}", list(p.glass[i])) for i in range(p.n_glass)]
    prism = cpd.CompoundPrism(glasses, list(p.angles[: p.n_glass + 1]), list(p.lengths[: p.n_glass]),
                              ((p.initial_curvature[0], p.initial_curvature[1]),
                               (p.final_curvature[0], p.final_curvature[1])), p.height, p.width, bool(p.ar_coated))
    wl = cpd.UniformWavelengthDistribution((p.wavelength_lo, p.wavelength_hi))
    beam = beam or cpd.FiberBeam(p.beam_width, p.numerical_aperture, p.y_mean)
    (pos, flipped) = cpd.position_detector_array(p.det_length, p.det_angle, prism, wl, beam, 1.0)
    det = cpd.DetectorArray(p.bin_count, p.bin_size, p.linear_slope, p.linear_intercept, p.det_length,
                            p.max_incident_angle, p.det_angle, pos, flipped)
    return cpd.Spectrometer(prism, det, wl, beam)


def test_python_surface_builds_3d_spectrometers():
    P, D, _ = cpd.generate_designs3(0, 3)
    for p, d in zip(P, D):
        spec = spectrometer_from(p)
        assert spec.dim == 3 and spec.compound_prism.dim == 3
        assert bytes(spec.design()) == bytes(d)
        mid, nrm = spec.compound_prism.final_midpt_and_normal()
        assert np.isfinite([mid.x, mid.y, nrm.x, nrm.y]).all()
        with pytest.raises(NotImplementedError):
            spec.ray_trace(0.6, p.y_mean)
    # the reference compiles no GPU kernel for a 3D Gaussian beam
    gauss = spectrometer_from(P[0], cpd.GaussianBeam(0.01, P[0].y_mean))
    assert gauss.design().beam_kind == N.BEAM_GAUSSIAN
    with pytest.raises(NotImplementedError):
        gauss.gpu_fitness([0.5])
    # nor a 2D fiber beam (spectrometer.rs:96-106)
    glasses = [cpd.Glass("g", list(P[0].glass[0]))]
    prism2 = cpd.CompoundPrism(glasses, [0.2, -0.2], [1.0], 0.5, 1.0, 1.0, False)
    with pytest.raises(NotImplementedError):
        cpd.position_detector_array(3.2, 0.0, prism2, cpd.UniformWavelengthDistribution((0.5, 0.8)),
                                    cpd.FiberBeam(0.005, 0.1, 0.5), 1.0)
