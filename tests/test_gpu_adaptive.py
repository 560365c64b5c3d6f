"""The reference's adaptive estimators on the GPU against the FP64 oracle:
fitness() with the SEM early exits (src/fitness.rs:5-34, 79-123) and the GPU
estimator cuda_fitness (src/cuda_fitness.rs:347-423 with
src/spectrometer/kernel.rs:89-193).

With CPD_OPT_ROW_TRACE = 64 every ray is traced in FP64 in the reference's
operation order from the same beam quantiles and folded in FP64 in the
reference's order, so results are compared for equality (the GPU
estimator's information within 1e-12: its entropy terms use the device's
log2).  With the FP32
trace (the GPU estimator's default) decisions stay the FP64 ones and each
ray's transmittance is rounded to FP32: information is compared within
INFO_TOL = 1e-4 bits, size and deviation for equality."""
import ctypes

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O
import kat

from test_gpu_fitness import INFO_TOL, kat_design, same_estimate, to_oracle
from workloads import transmitting

pytestmark = pytest.mark.gpu


def test_known_design_cpu_fitness_matches_oracle_and_known_answer():
    d = kat_design()
    s = N.Sampling(16384, 16384, N.MODE_ADAPTIVE, 0)
    f = N.Fitness()
    assert N.lib.cpd_fitness_gpu(None, ctypes.byref(d), ctypes.byref(s), ctypes.byref(f)) == 0, N.last_error()
    (size, info, dev), _ = O.fitness(to_oracle(d), 16384, 16384, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    assert (f.size, f.info, f.deviation) == (size, info, dev)
    # src/fitness.rs:250-257
    assert abs(f.info - kat.KAT_INFO) <= kat.KAT_INFO_RMAX * kat.KAT_INFO
    assert abs(f.size - kat.KAT_SIZE) <= kat.KAT_SIZE_RMAX * kat.KAT_SIZE


def test_adaptive_batch_matches_oracle():
    # designs that land >= 10% of a probe's rays in a bin
    _, D, _, _ = transmitting(0, 24)
    samp = N.Sampling(256, 4096, N.MODE_ADAPTIVE, 0)
    got, st = cpd.fitness_batch(D, samp)
    assert (st == 0).all()
    for i, d in enumerate(D):
        ref, _ = O.fitness(to_oracle(d), 256, 4096, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
        assert tuple(got[i]) == ref, (i, got[i], ref)


def test_large_adaptive_launch_matches_oracle():
    # a first round of at least 16 rows per CU runs one wavefront per row (the
    # smaller launches above run four): the same rows, bit for bit
    ctx = cpd.Context(0)
    _, D, _, _ = transmitting(0, 1280, first=300)
    assert 4 * len(D) >= 16 * ctx.info()["compute_units"]
    got, st = cpd.fitness_batch(D, N.Sampling(64, 512, N.MODE_ADAPTIVE, 0), ctx=ctx)
    assert (st == 0).all()
    ref, _ = O.fitness_many([to_oracle(d) for d in D], 64, 512, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    print(f"1280 transmitting designs: {int(np.sum(ref[:, 1] == 0))} with info 0, "
          f"{int(np.sum(ref[:, 1] > 0.1))} above 0.1 bits", flush=True)
    bad = np.flatnonzero((got != ref).any(axis=1))
    assert bad.size == 0, (bad[:5], got[bad[:5]], ref[bad[:5]])
    # the same designs one at a time run four wavefronts per row
    for i in range(0, len(D), 160):
        one, _ = cpd.fitness_batch([D[i]], N.Sampling(64, 512, N.MODE_ADAPTIVE, 0), ctx=ctx)
        assert np.array_equal(one[0], got[i]), i


def test_adaptive_and_fixed_share_a_batch():
    _, D, _, _ = cpd.generate_designs(0, 100, 8)
    samp = [N.Sampling(32, 2048, N.MODE_ADAPTIVE if i % 2 else N.MODE_FIXED, 0) for i in range(8)]
    got, st = cpd.fitness_batch(D, samp)
    assert (st == 0).all()
    for i, d in enumerate(D):
        alone, _ = cpd.fitness_batch([d], [samp[i]])
        assert np.array_equal(got[i], alone[0])


def test_transmission_probability_adaptive_matches_oracle():
    d = kat_design()
    w = np.linspace(0.5, 0.82, 9)
    out = np.empty((9, 32))
    dp = ctypes.POINTER(ctypes.c_double)
    st = N.lib.cpd_p_dets_l_wavelength(None, ctypes.byref(d), w.ctypes.data_as(dp), 9, 16384, N.MODE_ADAPTIVE,
                                       out.ctypes.data_as(dp))
    assert st == 0, N.last_error()
    for j in range(9):
        ref, _ = O.p_dets_l_wavelength(to_oracle(d), float(w[j]), 16384, adaptive=True, qrng=O.QRNG_FIXED_POINT)
        assert np.array_equal(out[j], ref), j


def fp64_rows_context():
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ROW_TRACE, 64)
    return ctx


@pytest.mark.parametrize("trace", [64, 32])
def test_gpu_fitness_matches_oracle_on_known_design(trace):
    # benches/benchmark.rs:147-157's call: seeds [0.798713], max_n 256, nwarp 2, max_eval 16384
    d = kat_design()
    f = N.Fitness()
    seeds = np.array([0.798713])
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ROW_TRACE, trace)
    st = N.lib.cpd_gpu_fitness(ctx.handle, ctypes.byref(d), seeds.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               1, 256, 2, 16384, ctypes.byref(f))
    assert st == 0, N.last_error()
    ref, converged, _ = O.gpu_fitness(to_oracle(d), seeds, 256, 2, 16384)
    assert converged
    if trace == 64:
        assert same_estimate((f.size, f.info, f.deviation), ref), ((f.size, f.info, f.deviation), ref)
    else:
        assert (f.size, f.deviation) == (ref[0], ref[2]) and abs(f.info - ref[1]) <= INFO_TOL


def test_gpu_fitness_not_converged_is_none_like_the_reference():
    d = kat_design()
    seeds = np.array([0.798713, 0.25])
    got, st = cpd.gpu_fitness_batch([d], seeds, max_n=4, nwarp=2, max_eval=16384, ctx=fp64_rows_context())
    ref, converged, _ = O.gpu_fitness(to_oracle(d), seeds, 4, 2, 16384)
    assert not converged and st[0] == N.ERR_NOT_CONVERGED
    assert same_estimate(tuple(got[0]), ref), (got[0], ref)
    spec_like, st32 = cpd.gpu_fitness_batch([d], seeds, max_n=4, nwarp=2, max_eval=16384)
    assert st32[0] == N.ERR_NOT_CONVERGED


@pytest.mark.parametrize("trace", [64, 32])
def test_gpu_fitness_large_batch_matches_single_designs(trace):
    # a batch of at least 16 rows per CU runs one wavefront per row, a single
    # design (128 rows) four: the same results bit for bit, and the FP64 rows
    # the oracle's
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ROW_TRACE, trace)
    _, D, _, _ = transmitting(0, 48, first=400)
    assert 128 * len(D) >= 16 * ctx.info()["compute_units"]
    seeds = np.array([0.37])
    got, st = cpd.gpu_fitness_batch(D, seeds, max_n=128, nwarp=2, max_eval=2048, ctx=ctx)
    for i in range(len(D)):
        one, st1 = cpd.gpu_fitness_batch([D[i]], seeds, max_n=128, nwarp=2, max_eval=2048, ctx=ctx)
        assert st1[0] == st[i] and np.array_equal(one[0], got[i]), (i, one[0], got[i])
        if trace == 64:
            ref, converged, _ = O.gpu_fitness(to_oracle(D[i]), seeds, 128, 2, 2048)
            assert (st[i] == 0) == converged and same_estimate(tuple(got[i]), ref), (i, got[i], ref)


@pytest.mark.parametrize("trace", [64, 32])
def test_gpu_fitness_batch_matches_single_designs(trace):
    _, D, _, _ = transmitting(0, 12, first=200)
    D = list(D) + [kat_design()]
    seeds = np.array([0.1, 0.7])
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ROW_TRACE, trace)
    got, st = cpd.gpu_fitness_batch(D, seeds, max_n=64, nwarp=2, max_eval=4096, ctx=ctx)
    agree = 0
    for i, d in enumerate(D):
        ref, converged, _ = O.gpu_fitness(to_oracle(d), seeds, 64, 2, 4096)
        if trace == 64:
            assert st[i] == (0 if converged else N.ERR_NOT_CONVERGED)
            assert same_estimate(tuple(got[i]), ref), (i, got[i], ref)
        else:
            agree += st[i] == (0 if converged else N.ERR_NOT_CONVERGED)
            assert abs(got[i][1] - ref[1]) <= INFO_TOL, (i, got[i], ref)
    # a design whose bin statistics sit on the SEM threshold may converge on one side only
    assert trace == 64 or agree >= len(D) - 1


def test_fp32_rows_fitness_close_to_oracle():
    _, D, _, _ = transmitting(0, 12)
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ROW_TRACE, 32)
    got, st = cpd.fitness_batch(D, N.Sampling(256, 4096, N.MODE_ADAPTIVE, 0), ctx=ctx)
    assert (st == 0).all()
    for i, d in enumerate(D):
        ref, _ = O.fitness(to_oracle(d), 256, 4096, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
        assert abs(got[i][1] - ref[1]) <= INFO_TOL, (i, got[i], ref)


def wide_design(params, nbin):
    """A generated design re-binned: nbin bins of 3.2 / nbin on the same array."""
    p = N.Params()
    ctypes.memmove(ctypes.byref(p), ctypes.byref(params), ctypes.sizeof(p))
    p.bin_count = nbin
    p.bin_size = p.linear_slope = 3.2 / nbin
    d = N.Design()
    assert N.lib.cpd_build_design(ctypes.byref(p), ctypes.byref(d)) == 0
    return d


@pytest.mark.parametrize("trace", [64, 32])
def test_gpu_estimator_mixed_bin_counts(trace):
    # one launch over designs of 16..128 bins and one of 1024 bins (the largest
    # LDS footprint of the chunk fold), 32 designs x 128 rows = at least 16 rows
    # per CU, so one wavefront per row; each design alone runs four.  FP64 rows
    # are the oracle's bit for bit.
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ROW_TRACE, trace)
    P, D, _, _ = transmitting(2, 31, first=500)
    D = list(D) + [wide_design(transmitting(0, 1, first=50)[0][0], 1024)]
    assert len({d.bin_count for d in D}) >= 4 and 128 * len(D) >= 16 * ctx.info()["compute_units"]
    seeds = np.array([0.61])
    got, st = cpd.gpu_fitness_batch(D, seeds, max_n=128, nwarp=2, max_eval=4096, ctx=ctx)
    for i, d in enumerate(D):
        one, st1 = cpd.gpu_fitness_batch([d], seeds, max_n=128, nwarp=2, max_eval=4096, ctx=ctx)
        assert st1[0] == st[i] and np.array_equal(one[0], got[i]), (i, d.bin_count, one[0], got[i])
        if trace == 64:
            ref, converged, _ = O.gpu_fitness(to_oracle(d), seeds, 128, 2, 4096)
            assert (st[i] == 0) == converged and same_estimate(tuple(got[i]), ref), (i, d.bin_count, got[i], ref)


def test_adaptive_fitness_mixed_bin_counts():
    # fitness() with the SEM exits over a launch that mixes 16..128-bin designs
    # with 512- and 1024-bin ones: large enough for one wavefront per row in its first
    # round; every design against the oracle, and a few alone (four wavefronts)
    ctx = cpd.Context(0)
    P, D, _, _ = transmitting(0, 1000, first=2000)
    P2, D2, _, _ = transmitting(2, 64, first=2000)
    D = list(D) + list(D2) + [wide_design(P[0], 1024), wide_design(P[1], 512)]
    assert len({d.bin_count for d in D}) >= 6
    samp = N.Sampling(64, 2048, N.MODE_ADAPTIVE, 0)
    got, st = cpd.fitness_batch(D, samp, ctx=ctx)
    assert (st == 0).all()
    ref, _ = O.fitness_many([to_oracle(d) for d in D], 64, 2048, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    bad = np.flatnonzero((got != ref).any(axis=1))
    assert bad.size == 0, (bad[:5], got[bad[:5]], ref[bad[:5]])
    for i in (0, 1, len(D) - 2, len(D) - 1):
        one, _ = cpd.fitness_batch([D[i]], samp, ctx=ctx)
        assert np.array_equal(one[0], got[i]), (i, D[i].bin_count)
