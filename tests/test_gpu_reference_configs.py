"""GPU parity on the detector and coating settings the reference ships.

Every designer configuration of the reference sets them (design_config.toml:23,
35-41; spring.toml:4-19; bch.toml:3-15): anti-reflection coated faces
(src/spectrometer/ray.rs:336-340, ci > 0.5 keeps 0.99 of Ts and Tp), bins
narrower than their pitch with a positive intercept, and a last bin that
overhangs the array (src/spectrometer/spectrometer.rs:291-303).  Configurations
5 (spring / design_config), 6 (bch) and 7 (random gaps, intercepts of up to
three pitches and coatings) draw designs with those settings that land at least
10% of a probe's rays in a bin.

Against the FP64 oracle on the same sample indices:
  fixed count            info within INFO_TOL = 1e-4 bits, size and deviation 1e-5
  p(d|lambda)            within 2e-6 per cell (one ray in another bin moves a
                         cell by ~3e-5)
  adaptive, FP64 rows    equal
  adaptive, FP32 rows    info within INFO_TOL
  GPU estimator          FP64 rows: status, size and deviation equal, info within
                         1e-12 (device log2 in the fold); FP32 rows info within INFO_TOL
  per ray                bins equal, transmittance within 1e-5 relative
The oracle in the reference's own iterative QMC recurrence (qrng.rs:251-254)
is compared within INFO_TOL."""
import ctypes

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O

import flat
from test_gpu_fitness import INFO_TOL, REL_TOL, check_close, same_estimate, to_oracle
from workloads import transmitting

pytestmark = pytest.mark.gpu

CONFIGS = [5, 6, 7]


def designs_of(config, n, first=0):
    _, D, _, _ = cpd.generate_designs(config, first, n)
    return D


def report(label, got, ref):
    err = np.abs(got[:, 1] - ref[:, 1])
    zero = int(np.sum(ref[:, 1] == 0.0))
    print(f"{label}: {len(ref)} designs, info {np.round(np.sort(ref[:, 1]), 3)}, {zero} with info 0, "
          f"max |dI| {err.max():.2e}", flush=True)
    return err


@pytest.mark.parametrize("config", CONFIGS)
def test_fixed_count_matches_oracle(config):
    D = designs_of(config, 16)
    got, st = cpd.fitness_batch(D, (16, 4096))
    assert (st == 0).all()
    ref, _ = O.fitness_many([to_oracle(d) for d in D], 16, 4096, 64, qrng=O.QRNG_FIXED_POINT)
    report(f"config {config} fixed", got, ref)
    for i in range(len(D)):
        check_close(got[i], ref[i], f"config {config} design {i}")
    assert np.sum(ref[:, 1] > 0.05) >= 4


@pytest.mark.parametrize("config", CONFIGS)
def test_bin_probabilities_match_oracle(config):
    dp = ctypes.POINTER(ctypes.c_double)
    for d in designs_of(config, 6, 100):
        w = np.linspace(d.wavelength_lo, d.wavelength_hi, 6)[1:-1].copy()
        ref = np.array([O.p_dets_l_wavelength(to_oracle(d), float(x), 16384, qrng=O.QRNG_FIXED_POINT)[0] for x in w])
        assert ref.sum() > 0.0
        out = np.empty((4, d.bin_count))
        st = N.lib.cpd_p_dets_l_wavelength(None, ctypes.byref(d), w.ctypes.data_as(dp), 4, 16384, N.MODE_FIXED,
                                           out.ctypes.data_as(dp))
        assert st == 0, N.last_error()
        np.testing.assert_allclose(out, ref, atol=2e-6, err_msg=f"config {config}")


def rows_context(trace):
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ROW_TRACE, trace)
    return ctx


@pytest.mark.parametrize("trace", [64, 32])
@pytest.mark.parametrize("config", CONFIGS)
def test_adaptive_fitness_matches_oracle(config, trace):
    D = designs_of(config, 12, 200)
    got, st = cpd.fitness_batch(D, N.Sampling(256, 4096, N.MODE_ADAPTIVE, 0), ctx=rows_context(trace))
    assert (st == 0).all()
    ref, _ = O.fitness_many([to_oracle(d) for d in D], 256, 4096, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    report(f"config {config} adaptive rows {trace}", got, ref)
    for i in range(len(D)):
        if trace == 64:
            assert tuple(got[i]) == tuple(ref[i]), (i, got[i], ref[i])
        else:
            check_close(got[i], ref[i], f"config {config} design {i}")


@pytest.mark.parametrize("trace", [64, 32])
@pytest.mark.parametrize("config", CONFIGS)
def test_gpu_estimator_matches_oracle(config, trace):
    # designer.py:276's call: one seed, max_n 128, max_eval 16384 // 2
    D = designs_of(config, 12, 300)
    seeds = np.array([0.37])
    got, st = cpd.gpu_fitness_batch(D, seeds, max_n=128, nwarp=2, max_eval=8192, ctx=rows_context(trace))
    rows = []
    agree = 0
    for i, d in enumerate(D):
        ref, converged, _ = O.gpu_fitness(to_oracle(d), seeds, 128, 2, 8192)
        rows.append(ref)
        if trace == 64:
            assert st[i] == (0 if converged else N.ERR_NOT_CONVERGED), i
            assert same_estimate(tuple(got[i]), ref), (i, got[i], ref)
        else:
            agree += st[i] == (0 if converged else N.ERR_NOT_CONVERGED)
            assert abs(got[i][1] - ref[1]) <= INFO_TOL, (i, got[i], ref)
    report(f"config {config} gpu_fitness rows {trace}", got, np.array(rows))
    # FP32 rows: a design whose bin statistics sit on the SEM threshold may converge on one side only
    assert trace == 64 or agree >= len(D) - 1


def gpu_rays(d, wcdf, rcdf, ctx=None):
    wcdf = np.ascontiguousarray(wcdf, np.float32)
    rcdf = np.ascontiguousarray(rcdf, np.float32)
    bins = np.empty(wcdf.size, np.uint32)
    ts = np.empty(wcdf.size, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    st = N.lib.cpd_propagate_rays(None if ctx is None else ctx.handle, ctypes.byref(d), wcdf.ctypes.data_as(fp),
                                  rcdf.ctypes.data_as(fp), wcdf.size,
                                  bins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ts.ctypes.data_as(fp))
    assert st == 0, N.last_error()
    return bins, ts


def oracle_rays(d, wcdf, rcdf):
    od = to_oracle(d)
    lam = d.wavelength_lo + (d.wavelength_hi - d.wavelength_lo) * np.asarray(wcdf, np.float32).astype(np.float64)
    y = d.y_mean - d.beam_width * O.norminv_many(np.asarray(rcdf, np.float32).astype(np.float64))
    return O.propagate_many(od, lam, y, precision=64)


def test_per_ray_bins_match_oracle_on_a_million_rays():
    # src/cuda_fitness.rs:533-607's harness (one ray per (wavelength cdf, ray
    # cdf) pair) scaled up: 2^16 random pairs for each of 5 designs of
    # configurations 5, 6, 7 and the KAT-like configuration 0 (hit fraction >=
    # 10%): 1.05e6 rays through the FP32 device trace with its FP64 re-trace of
    # flagged rays.  Every bin is the FP64 oracle's.
    rng = np.random.default_rng(2024)
    total = hits = 0
    groups = [(c, designs_of(c, 5, 400)) for c in CONFIGS] + [(0, transmitting(0, 5)[1])]
    for config, D in groups:
        for d in D:
            w = rng.random(1 << 16, dtype=np.float32)
            r = rng.random(1 << 16, dtype=np.float32)
            gb, gt = gpu_rays(d, w, r)
            ob, ot = oracle_rays(d, w, r)
            bad = np.flatnonzero(gb != ob)
            assert bad.size == 0, (config, bad[:5], gb[bad[:5]], ob[bad[:5]])
            hit = ob < d.bin_count
            np.testing.assert_allclose(gt[hit], ot[hit], rtol=REL_TOL, atol=0)
            total += w.size
            hits += int(hit.sum())
    print(f"per-ray: {total} rays, {hits} hits, all bins equal", flush=True)
    assert total >= 1_000_000 and hits >= 100_000


def test_saturating_bin_zero_and_gaps_on_gpu():
    # the flat design with bins of 0.08 every 0.1 from intercept 0.65 and a
    # beam of width 0.5 over the 2.5 aperture (pos = y + 0.4): positions before the
    # intercept fall in bin 0 by the saturating cast when their remainder is
    # below 0.08; the FP32 device trace must count exactly those
    p = flat.flat_params(N.Params(), ar=True)
    p.bin_size, p.linear_slope, p.linear_intercept = 0.08, 0.1, 0.65
    p.beam_width = 0.5
    d = N.Design()
    assert N.lib.cpd_build_design(ctypes.byref(p), ctypes.byref(d)) == 0
    rng = np.random.default_rng(7)
    w = rng.random(1 << 18, dtype=np.float32)
    r = rng.random(1 << 18, dtype=np.float32)
    gb, gt = gpu_rays(d, w, r)
    ob, ot = oracle_rays(d, w, r)
    assert np.array_equal(gb, ob)
    y = d.y_mean - d.beam_width * O.norminv_many(r.astype(np.float64))
    before = (y + 0.4 < 0.65) & (ob == 0)
    assert before.sum() > 1000  # rays before the intercept counted in bin 0
    gaps = (ob == d.bin_count) & (y > 0.3) & (y < 2.4)
    assert gaps.sum() > 10000   # rays in the gaps between bins
    hit = ob < d.bin_count
    # the FP32 trace multiplies 0.99f by 0.99f: within an ulp of the FP64 product
    np.testing.assert_allclose(gt[hit], flat.T_AR, rtol=1e-7)


def test_flat_designs_stop_where_the_closed_form_says():
    # the early-exit closed forms of tests/test_oracle.py on the GPU: a single
    # bin of constant t stops its row at ray 196 with p = 196 c / 197; equal
    # rows stop fitness() after two wavelengths; the GPU estimator's single-bin
    # row stops after one 32-ray chunk with p = c
    p = flat.single_bin(flat.flat_params(N.Params()))
    d = N.Design()
    assert N.lib.cpd_build_design(ctypes.byref(p), ctypes.byref(d)) == 0
    dp = ctypes.POINTER(ctypes.c_double)
    w = np.array([0.55, 0.7])
    out = np.empty((2, 32))
    assert N.lib.cpd_p_dets_l_wavelength(None, ctypes.byref(d), w.ctypes.data_as(dp), 2, 16384, N.MODE_ADAPTIVE,
                                         out.ctypes.data_as(dp)) == 0
    want = np.zeros(32)
    want[16] = 196 * flat.T_AR / 197
    np.testing.assert_allclose(out, [want, want], rtol=1e-15, atol=0)
    f = N.Fitness()
    s = N.Sampling(64, 16384, N.MODE_ADAPTIVE, 0)
    assert N.lib.cpd_fitness_gpu(None, ctypes.byref(d), ctypes.byref(s), ctypes.byref(f)) == 0
    (size, info, dev), (rows, rays) = O.fitness_counts(to_oracle(d), 64, 16384, 64, adaptive=True,
                                                       qrng=O.QRNG_FIXED_POINT)
    assert (rows, rays) == (2, 392) and info == 0.0
    assert (f.size, f.info, f.deviation) == (size, info, dev)
    # the split design: both estimators against the oracle, whose stop points
    # the closed forms pin
    ps = flat.split_bins(flat.flat_params(N.Params()))
    ds = N.Design()
    assert N.lib.cpd_build_design(ctypes.byref(ps), ctypes.byref(ds)) == 0
    out2 = np.empty((1, 2))
    w1 = np.array([0.6])
    assert N.lib.cpd_p_dets_l_wavelength(None, ctypes.byref(ds), w1.ctypes.data_as(dp), 1, 16384, N.MODE_ADAPTIVE,
                                         out2.ctypes.data_as(dp)) == 0
    ref, rays = O.p_dets_l_wavelength(to_oracle(ds), 0.6, 16384, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    assert rays == 9596 and np.array_equal(out2[0], ref)
    for dd in (d, ds):
        got, st = cpd.gpu_fitness_batch([dd], np.array([0.3]), max_n=8, nwarp=2, max_eval=16384,
                                        ctx=rows_context(64))
        ref, converged, _ = O.gpu_fitness(to_oracle(dd), np.array([0.3]), 8, 2, 16384)
        assert same_estimate(tuple(got[0]), ref) and (st[0] == 0) == converged


@pytest.mark.parametrize("config", [0, 5])
def test_against_the_references_iterative_sequence(config):
    # the reference draws u_k by the FP64 recurrence state = fract(state + 1/phi)
    # (qrng.rs:251-254), within 1e-11 of the exact points this build uses
    D = transmitting(config, 8)[1]
    got, st = cpd.fitness_batch(D, (32, 8192))
    assert (st == 0).all()
    ref, _ = O.fitness_many([to_oracle(d) for d in D], 32, 8192, 64, qrng=O.QRNG_ITERATIVE)
    report(f"config {config} vs the iterative sequence, fixed", got, ref)
    for i in range(len(D)):
        check_close(got[i], ref[i], f"design {i}")
    got_a, st = cpd.fitness_batch(D, N.Sampling(128, 4096, N.MODE_ADAPTIVE, 0), ctx=rows_context(64))
    ref_a, _ = O.fitness_many([to_oracle(d) for d in D], 128, 4096, 64, adaptive=True, qrng=O.QRNG_ITERATIVE)
    report(f"config {config} vs the iterative sequence, adaptive", got_a, ref_a)
    for i in range(len(D)):
        check_close(got_a[i], ref_a[i], f"adaptive design {i}")
