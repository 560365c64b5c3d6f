"""GPU parity: the HIP fitness path against the FP64 oracle on the same inputs.

Tolerances (from the fixed-count estimator's definition):
  info        |GPU - oracle FP64| <= 1e-4 bits
  size, dev   relative <= 1e-5 (both are FP64 on either side)
Sample indices are the same integers k on both sides (R1 sequence in 64-bit
fixed point, oracle QRNG_FIXED_POINT)."""
import ctypes

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O
import kat
from workloads import transmitting

pytestmark = pytest.mark.gpu

INFO_TOL = 1e-4
REL_TOL = 1e-5


def to_oracle(d: N.Design) -> O.Design:
    od = O.Design()
    ctypes.memmove(ctypes.byref(od), ctypes.byref(d), ctypes.sizeof(od))
    return od


def kat_design() -> N.Design:
    p = kat.fill_params(N.Params())
    pos = (ctypes.c_double * 2)()
    fl = ctypes.c_uint32()
    assert N.lib.cpd_position_detector_array(ctypes.byref(p), 1.0, pos, ctypes.byref(fl)) == 0
    p.det_position[0], p.det_position[1] = pos[0], pos[1]
    p.flipped = fl.value
    d = N.Design()
    assert N.lib.cpd_build_design(ctypes.byref(p), ctypes.byref(d)) == 0
    return d


def gpu_fitness(d, n_lambda, n_rays):
    s = N.Sampling(n_lambda, n_rays, N.MODE_FIXED, 0)
    f = N.Fitness()
    st = N.lib.cpd_fitness_gpu(None, ctypes.byref(d), ctypes.byref(s), ctypes.byref(f))
    assert st == 0, N.last_error()
    return np.array([f.size, f.info, f.deviation])


# The GPU estimator folds its rows on the device: its per-bin summaries and
# stop decisions are the host fold's bit for bit, its entropy terms use the
# device's log2, so information agrees with the FP64 oracle to this bound.
GPU_FOLD_INFO_TOL = 1e-12


def same_estimate(got, ref):
    """cpd_gpu_fitness with FP64 rows against the oracle's estimator: size and deviation equal,
    information within GPU_FOLD_INFO_TOL."""
    return got[0] == ref[0] and got[2] == ref[2] and abs(got[1] - ref[1]) <= GPU_FOLD_INFO_TOL


def check_close(gpu, ref, label):
    assert abs(gpu[1] - ref[1]) <= INFO_TOL, f"{label}: info {gpu[1]} vs {ref[1]}"
    for i in (0, 2):
        assert abs(gpu[i] - ref[i]) <= REL_TOL * max(abs(ref[i]), 1e-300), f"{label}: {gpu} vs {ref}"


def test_kat_fixed_count_matches_oracle():
    d = kat_design()
    gpu = gpu_fitness(d, 64, 16384)
    ref, rays = O.fitness(to_oracle(d), 64, 16384, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
    assert rays == 64 * 16384
    print("KAT 64x16384 gpu", gpu, "oracle", ref, "dI", gpu[1] - ref[1])
    check_close(gpu, np.array(ref), "KAT")


def test_kat_known_answer_on_gpu():
    # src/fitness.rs:255-257 with the sample counts the reference's benchmark uses
    d = kat_design()
    gpu = gpu_fitness(d, 1024, 16384)
    assert abs(gpu[0] - kat.KAT_SIZE) <= kat.KAT_SIZE_RMAX * kat.KAT_SIZE
    assert abs(gpu[1] - kat.KAT_INFO) <= kat.KAT_INFO_RMAX * kat.KAT_INFO
    assert abs(gpu[2] - kat.KAT_DEVIATION) <= kat.KAT_DEVIATION_RMAX * kat.KAT_DEVIATION


# (4, 6, 4, 65536): 6 prisms and 256 bins, so a 65536-ray item folds its u32
# LDS columns into the u64 row 16 times
@pytest.mark.parametrize("config,n,nl,nr", [(0, 24, 16, 4096), (2, 24, 8, 2048), (4, 6, 4, 65536)])
def test_generated_batch_matches_oracle(config, n, nl, nr):
    # designs that land >= 10% of a probe's rays in a bin
    _, designs, _, hit = transmitting(config, n)
    samp = [N.Sampling(nl, nr, N.MODE_FIXED, 0) for _ in range(n)]
    res, status = cpd.fitness_batch(designs, samp)
    assert (status == 0).all()
    ref, _ = O.fitness_many([to_oracle(d) for d in designs], nl, nr, precision=64, qrng=O.QRNG_FIXED_POINT)
    err = np.abs(res[:, 1] - ref[:, 1])
    print(f"config {config}: max |dI| {err.max():.3e}, hit fractions {np.round(hit, 2)}, "
          f"infos {np.round(ref[:, 1], 4)}, {int(np.sum(ref[:, 1] == 0))} with info 0")
    for i in range(n):
        check_close(res[i], ref[i], f"design {i}")
    assert np.sum(ref[:, 1] > 0.05) >= n // 4


def test_batch_result_independent_of_companions():
    # integer accumulation: a design's result is bit-identical alone and in a batch
    designs = [cpd.generate_design(2, i)[1] for i in range(12)]
    samp = [N.Sampling(8 + i, 1024 + 64 * i, N.MODE_FIXED, 0) for i in range(12)]
    batch, _ = cpd.fitness_batch(designs, samp)
    sub, _ = cpd.fitness_batch(designs[5:9][::-1], samp[5:9][::-1])
    for q, i in enumerate(range(8, 4, -1)):
        alone = gpu_fitness(designs[i], samp[i].n_wavelengths, samp[i].n_rays)
        assert np.array_equal(alone, batch[i]), (i, alone, batch[i])
        assert np.array_equal(sub[q], batch[i])


def test_short_rows_claimed_several_at_a_time():
    # 1024 designs x 64 wavelengths of 1024 rays: each fetch of the persistent
    # grid's work counter claims several consecutive work items (rows this
    # short would otherwise serialise on the counter); the results are the
    # one-launch-per-design results bit for bit
    _, D, _, _ = cpd.generate_designs(1, 0, 1024)
    batch, st = cpd.fitness_batch(D, (64, 1024))
    assert (st == 0).all()
    for i in range(0, 1024, 97):
        assert np.array_equal(gpu_fitness(D[i], 64, 1024), batch[i]), i


def test_propagation_matches_oracle_per_ray():
    # src/cuda_fitness.rs:533-607: the reference harness's 1024 (wavelength cdf,
    # ray cdf) pairs (Xoshiro256** seed_from_u64(0), Standard f32) on the known design
    from xoshiro import harness_pairs

    d = kat_design()
    wcdf, rcdf = harness_pairs(1024)
    bins = np.empty(1024, np.uint32)
    ts = np.empty(1024, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    st = N.lib.cpd_propagate_rays(None, ctypes.byref(d), wcdf.ctypes.data_as(fp), rcdf.ctypes.data_as(fp), 1024,
                                  bins.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), ts.ctypes.data_as(fp))
    assert st == 0
    lam = 0.5 + (0.82 - 0.5) * wcdf.astype(np.float64)
    y = 0.95 - 0.2 * np.array([O.norminv(float(u)) for u in rcdf])
    ob, ot = O.propagate_many(to_oracle(d), lam, y, precision=64)
    assert np.array_equal(bins, ob)
    hit = ob < 32
    assert hit.sum() > 100
    np.testing.assert_allclose(ts[hit], ot[hit], rtol=1e-5)


def test_transmission_probability_matches_oracle():
    d = kat_design()
    w = np.linspace(0.5, 0.82, 7)
    out = np.empty((7, 32))
    dp = ctypes.POINTER(ctypes.c_double)
    st = N.lib.cpd_p_dets_l_wavelength(None, ctypes.byref(d), w.ctypes.data_as(dp), 7, 4096, N.MODE_FIXED,
                                       out.ctypes.data_as(dp))
    assert st == 0
    for j in range(7):
        ref, _ = O.p_dets_l_wavelength(to_oracle(d), float(w[j]), 4096, qrng=O.QRNG_FIXED_POINT)
        np.testing.assert_allclose(out[j], ref, atol=2e-6)


@pytest.mark.parametrize("config", [1, 2, 4])
def test_bin_probabilities_of_generated_designs_match_oracle(config):
    # Per-(wavelength, bin) probabilities of the generated designs that reach
    # the detector, 4 wavelengths x 16384 rays each.  Rays whose FP32 outcome
    # lies within the error margin of a bin edge (a few per mille) are re-traced
    # in FP64, so every ray must land in the oracle's bin: one ray in another bin
    # moves two cells by t / 16384 (about 3e-5 for t = 0.5), far above the 2e-6
    # that FP32 transmittances leave.
    _, D, _, _ = transmitting(config, 12)
    checked = 0
    dp = ctypes.POINTER(ctypes.c_double)
    for d in D:
        w = np.linspace(d.wavelength_lo, d.wavelength_hi, 6)[1:-1].copy()
        ref = np.array([O.p_dets_l_wavelength(to_oracle(d), float(x), 16384, qrng=O.QRNG_FIXED_POINT)[0] for x in w])
        if ref.sum() == 0.0:
            continue
        out = np.empty((4, d.bin_count))
        st = N.lib.cpd_p_dets_l_wavelength(None, ctypes.byref(d), w.ctypes.data_as(dp), 4, 16384, N.MODE_FIXED,
                                           out.ctypes.data_as(dp))
        assert st == 0, N.last_error()
        np.testing.assert_allclose(out, ref, atol=2e-6, err_msg=f"config {config}")
        checked += 1
    assert checked >= 10


@pytest.mark.parametrize("acc", [N.ACC_EXACT, N.ACC_F32, N.ACC_F32C, N.ACC_F64])
def test_accumulation_policies_match_oracle(acc):
    # 6 prisms and 256 bins, 4 wavelengths x 65536 rays: every policy estimates
    # the same information; at this size all are within INFO_TOL of the FP64 oracle
    _, D, _, _ = cpd.generate_designs(4, 0, 6)
    samp = N.Sampling(4, 65536, N.MODE_FIXED, 0)
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ACCUMULATE, acc)
    got, st = cpd.fitness_batch(D, samp, ctx=ctx)
    assert (st == 0).all()
    for i, d in enumerate(D):
        ref, _ = O.fitness(to_oracle(d), 4, 65536, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
        check_close(got[i], np.array(ref), f"policy {acc} design {i}")


@pytest.fixture(scope="module")
def config4_rows():
    # two transmitting designs of configuration 4 (6 prisms, 256 bins) at the
    # configuration's own row length, 2^20 rays, over 4 wavelengths
    _, D, _, _ = transmitting(4, 2)
    ref, _ = O.fitness_many([to_oracle(d) for d in D], 4, 1 << 20, 64, qrng=O.QRNG_FIXED_POINT)
    return D, ref


@pytest.mark.parametrize("acc,tile", [(N.ACC_EXACT, 0), (N.ACC_EXACT, 1 << 20), (N.ACC_F32, 0), (N.ACC_F32C, 0),
                                      (N.ACC_F64, 0)])
def test_config4_rows_of_a_million_rays(config4_rows, acc, tile):
    # tile 2^20 makes a whole wavelength one work item (the largest the trace
    # takes, folding its u32 columns 265 times); FP32C always takes one item per
    # wavelength.  Every policy is within INFO_TOL of the FP64 oracle.
    D, ref = config4_rows
    ctx = cpd.Context(0)
    ctx.set_option(N.OPT_ACCUMULATE, acc)
    if tile:
        ctx.set_option(N.OPT_TILE_RAYS, tile)
    got, st = cpd.fitness_batch(D, N.Sampling(4, 1 << 20, N.MODE_FIXED, 0), ctx=ctx)
    assert (st == 0).all()
    print(f"policy {acc} tile {tile}: dI {got[:, 1] - ref[:, 1]}, info {ref[:, 1]}", flush=True)
    for i in range(len(D)):
        check_close(got[i], ref[i], f"policy {acc} design {i}")


@pytest.mark.parametrize("n_rays", [100, 2500, 4097, 10000, 20000])
def test_ray_counts_off_the_block_size_match_oracle(n_rays):
    # The trace covers whole banded blocks of 4096 positions; in a design's
    # last block the positions of rays past n_rays are idle.  Every ray must be
    # traced exactly once: one ray missed or counted twice moves a cell by
    # t / n_rays (>= 1e-5 here), far above the 2e-6 FP32 transmittances leave.
    _, D, _, _ = transmitting(0, 6)
    dp = ctypes.POINTER(ctypes.c_double)
    for d in D:
        w = np.linspace(d.wavelength_lo, d.wavelength_hi, 5)[1:-1].copy()
        ref = np.array([O.p_dets_l_wavelength(to_oracle(d), float(x), n_rays, qrng=O.QRNG_FIXED_POINT)[0]
                        for x in w])
        out = np.empty((3, d.bin_count))
        st = N.lib.cpd_p_dets_l_wavelength(None, ctypes.byref(d), w.ctypes.data_as(dp), 3, n_rays, N.MODE_FIXED,
                                           out.ctypes.data_as(dp))
        assert st == 0, N.last_error()
        np.testing.assert_allclose(out, ref, atol=2e-6, err_msg=f"{n_rays} rays")


@pytest.mark.parametrize("tile", [1024, 1088, 4160, 12352])
def test_tiles_that_split_the_last_block(tile):
    # work items of 1024 positions split a design's last, partly idle block
    # across items; tiles that are no multiple of the 4096-position block start
    # inside blocks and span two or more, so their band tests bound ranks over
    # partial blocks and their chunks are traced in position order.  The sums
    # (exact integers) match one item per wavelength and the oracle
    _, D, _, _ = transmitting(0, 8)
    S = [N.Sampling(8, 5000 + 320 * i, N.MODE_FIXED, 0) for i in range(len(D))]
    ctx = cpd.Context(0)
    whole, st = cpd.fitness_batch(D, S, ctx=ctx)
    assert (st == 0).all()
    ctx.set_schedule(N.Schedule(tile, 16, N.ORDER_GIVEN, 0))
    split, st = cpd.fitness_batch(D, S, ctx=ctx)
    ctx.set_schedule(None)
    assert (st == 0).all()
    assert np.array_equal(whole, split)
    for i, d in enumerate(D):
        r, _ = O.fitness(to_oracle(d), 8, S[i].n_rays, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
        assert abs(split[i][1] - r[1]) <= INFO_TOL, (i, split[i], r)


def test_many_entries_find_their_items():
    # a launch of 5000 designs of 1-3 items each: work items find their design
    # by a 64-lane search over the item prefix in three rounds (two cover at
    # most 4096 entries); every design matches its own result in launches of
    # 1000 designs
    _, D, _, _ = cpd.generate_designs(0, 0, 5000)
    S = [N.Sampling(1 + i % 3, 256, N.MODE_FIXED, 0) for i in range(len(D))]
    whole, st = cpd.fitness_batch(D, S)
    assert (st == 0).all()
    for q in range(0, len(D), 1000):
        part, st = cpd.fitness_batch(D[q:q + 1000], S[q:q + 1000])
        assert (st == 0).all()
        assert np.array_equal(part, whole[q:q + 1000]), q
