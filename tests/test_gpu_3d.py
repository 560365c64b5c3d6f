"""The 3D spectrometer on the GPU (cpd3_* entry points) against the FP64 CPU
oracle: per ray, per row, whole fitness and the GPU estimator.  Beam samples
come from the same host math on both sides and the device trace keeps the
reference's operation order, so every comparison is exact."""
import ctypes
import math

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O
from oracle import oracle3 as O3

pytestmark = pytest.mark.gpu


def designs(n, first=0):
    P, D, S = cpd.generate_designs3(first, n)
    return P, D


def gaussian_variant(p):
    """The same prism with a 3D Gaussian beam, detector positioned anew."""
    q = N.Params3()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
    q.beam_kind = N.BEAM_GAUSSIAN
    q.beam_width = 0.02
    pos = (ctypes.c_double * 2)()
    fl = ctypes.c_uint32()
    if N.lib.cpd3_position_detector_array(ctypes.byref(q), 1.0, pos, ctypes.byref(fl)) != 0:
        return None
    q.det_position[0], q.det_position[1], q.flipped = pos[0], pos[1], fl.value
    d = N.Design3()
    assert N.lib.cpd3_build_design(ctypes.byref(q), ctypes.byref(d)) == 0
    return d


def test_propagate_rays_match_oracle():
    _, D = designs(4)
    rng = np.random.default_rng(0)
    for i, d in enumerate(D):
        print(f"design {i}", flush=True)
        n = 4096
        w = rng.random(n)
        q = rng.random((n, 4))
        bins = np.empty(n, np.uint32)
        ts = np.empty(n, np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        assert N.lib.cpd3_propagate_rays(None, ctypes.byref(d), w.ctypes.data_as(dp), q.ctypes.data_as(dp), n,
                                         bins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                         ts.ctypes.data_as(dp)) == 0
        hits = 0
        for k in range(n):
            lam = d.wavelength_lo + (d.wavelength_hi - d.wavelength_lo) * w[k]
            o, v = O3.sample_ray(d, q[k])
            st, b, t = O3.propagate(d, lam, o, v)
            if st != 0:
                b, t = d.bin_count, 0.0
            assert bins[k] == b and ts[k] == t, (i, k)
            hits += b < d.bin_count
        assert hits > 50


@pytest.mark.parametrize("adaptive", [True, False])
def test_rows_match_oracle(adaptive):
    P, D = designs(3)
    g = gaussian_variant(P[1])
    cases = list(D) + ([g] if g is not None else [])
    for i, d in enumerate(cases):
        print(f"case {i} beam {d.beam_kind} adaptive {adaptive}", flush=True)
        lo, hi = d.wavelength_lo, d.wavelength_hi
        wl = np.array([lo, 0.5 * (lo + hi), hi, lo + 0.3 * (hi - lo), lo + 0.77 * (hi - lo)])
        out = np.empty((wl.size, d.bin_count), np.float64)
        dp = ctypes.POINTER(ctypes.c_double)
        assert N.lib.cpd3_p_dets_l_wavelength(None, ctypes.byref(d), wl.ctypes.data_as(dp), wl.size, 4096,
                                              N.MODE_ADAPTIVE if adaptive else N.MODE_FIXED,
                                              out.ctypes.data_as(dp)) == 0
        for j, w in enumerate(wl):
            ref, _ = O3.p_dets_l_wavelength(d, w, 4096, adaptive=adaptive)
            np.testing.assert_array_equal(out[j], ref)


@pytest.mark.parametrize("mode", [N.MODE_ADAPTIVE, N.MODE_FIXED])
def test_fitness_batch_matches_oracle(mode):
    _, D = designs(8)
    s = N.Sampling(16, 2048, mode, 0)
    res, st = cpd.fitness_batch3(D, s)
    assert (st == 0).all()
    ref = O3.fitness_many(D, 16, 2048, adaptive=mode == N.MODE_ADAPTIVE)
    np.testing.assert_array_equal(res, ref)
    assert np.all(res[:, 1] <= np.log2([d.bin_count for d in D]) + 1e-12)


@pytest.mark.parametrize("mode", [N.MODE_ADAPTIVE, N.MODE_FIXED])
def test_large_launches_match_oracle(mode):
    # launches of at least 16 rows per CU run one wavefront per row (the
    # smaller ones above run four)
    cus = cpd.Context(0).info()["compute_units"]
    n, nw = (1024, 8) if mode == N.MODE_ADAPTIVE else (512, 16)
    assert (4 if mode == N.MODE_ADAPTIVE else nw) * n >= 16 * cus
    _, D = designs(n, first=50)
    res, st = cpd.fitness_batch3(D, N.Sampling(nw, 256, mode, 0))
    assert (st == 0).all()
    ref = O3.fitness_many(D, nw, 256, adaptive=mode == N.MODE_ADAPTIVE)
    bad = np.flatnonzero((res != ref).any(axis=1))
    assert bad.size == 0, (bad[:5], res[bad[:5]], ref[bad[:5]])


def test_batch_results_do_not_depend_on_companions():
    _, D = designs(6)
    s = N.Sampling(12, 1024, N.MODE_ADAPTIVE, 0)
    res, _ = cpd.fitness_batch3(D, s)
    for i in (0, 3, 5):
        solo, _ = cpd.fitness_batch3([D[i]], s)
        np.testing.assert_array_equal(solo[0], res[i])
    # mixed modes and an invalid record in one call
    bad = N.Design3()
    ctypes.memmove(ctypes.byref(bad), ctypes.byref(D[0]), ctypes.sizeof(bad))
    bad.bin_count = 0
    mixed, st = cpd.fitness_batch3([D[1], bad, D[2]], [s, s, N.Sampling(8, 512, N.MODE_FIXED, 0)])
    assert list(st) == [0, N.ERR_INVALID_ARGUMENT, 0] and np.isnan(mixed[1]).all()
    np.testing.assert_array_equal(mixed[0], res[1])
    fixed = O3.fitness_many([D[2]], 8, 512, adaptive=False)
    np.testing.assert_array_equal(mixed[2], fixed[0])


def test_gpu_estimator_matches_oracle():
    _, D = designs(3)
    seeds = [0.37, 0.71]
    res, st = cpd.gpu_fitness_batch3(D, seeds, max_n=32, nwarp=2, max_eval=2048)
    for i, d in enumerate(D):
        print(f"design {i}", flush=True)
        ref, conv, _ = O3.gpu_fitness(d, seeds, 32, 2, 2048)
        assert (st[i] == 0) == conv
        np.testing.assert_array_equal(res[i], ref)


def test_python_surface_3d():
    P, D = designs(2)
    p = P[0]
    glasses = [cpd.Glass(f"g{i}", list(p.glass[i])) for i in range(p.n_glass)]
    prism = cpd.CompoundPrism(glasses, list(p.angles[: p.n_glass + 1]), list(p.lengths[: p.n_glass]),
                              ((p.initial_curvature[0], p.initial_curvature[1]),
                               (p.final_curvature[0], p.final_curvature[1])), p.height, p.width, bool(p.ar_coated))
    wl = cpd.UniformWavelengthDistribution((p.wavelength_lo, p.wavelength_hi))
    beam = cpd.FiberBeam(p.beam_width, p.numerical_aperture, p.y_mean)
    pos, flipped = cpd.position_detector_array(p.det_length, p.det_angle, prism, wl, beam, 1.0)
    det = cpd.DetectorArray(p.bin_count, p.bin_size, p.linear_slope, p.linear_intercept, p.det_length,
                            p.max_incident_angle, p.det_angle, pos, flipped)
    spec = cpd.Spectrometer(prism, det, wl, beam)
    f = spec.cpu_fitness(max_n=16, max_m=2048)
    ref, _ = O3.fitness(spec.design(), 16, 2048, adaptive=True)
    assert (f.size, f.info, f.deviation) == ref
    rows = spec.transmission_probability([0.55, 0.7], max_m=2048)
    assert rows.shape == (2, p.bin_count)
    np.testing.assert_array_equal(rows[1], O3.p_dets_l_wavelength(spec.design(), 0.7, 2048)[0])
    g = spec.gpu_fitness([0.37], max_n=16, max_eval=2048)
    ref_g, conv, _ = O3.gpu_fitness(spec.design(), [0.37], 16, 2, 2048)
    if conv:
        assert (g.size, g.info, g.deviation) == ref_g
    else:
        assert g is None
    b, t = spec.propagate_rays([0.5, 0.5], [[0.3, 0.2, 0.5, 0.1], [0.9, 0.9, 0.9, 0.9]])
    assert b.shape == (2,) and t.dtype == np.float64
