"""Schedules, the autotuner and captured generations (cpd_plan_*).

Bin sums are exact integers, so every schedule and every replay of a captured
generation must give results bit-identical to an eager cpd_fitness_batch of
the same designs."""
import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N

pytestmark = pytest.mark.gpu


def uneven(n, first=0, seed=0):
    """n designs of the uneven distribution (1-6 prisms, 16-128 bins) with
    sampling reduced so the test is quick."""
    _, D, S, _ = cpd.generate_designs(2, first, n)
    rng = np.random.default_rng(seed)
    samp = [N.Sampling(int(rng.integers(4, 20)), int(64 * rng.integers(8, 200)), N.MODE_FIXED, 0) for _ in range(n)]
    return list(D), samp


def test_schedules_give_identical_results():
    D, S = uneven(48)
    ctx = cpd.Context(0)
    ref, st = cpd.fitness_batch(D, S, ctx=ctx)
    assert (st == 0).all()
    for tile, waves, order in [(1024, 4, N.ORDER_GIVEN), (4096, 8, N.ORDER_COST), (0, 24, N.ORDER_COST),
                               (64, 1, N.ORDER_GIVEN)]:
        ctx.set_schedule(N.Schedule(tile, waves, order, 0))
        got, st = cpd.fitness_batch(D, S, ctx=ctx)
        assert (st == 0).all()
        assert np.array_equal(got, ref), (tile, waves, order)
    ctx.set_schedule(None)


def test_item_size_does_not_change_results_with_folds():
    # 256 bins: a whole-wavelength item of 65536 rays folds its columns 16 times;
    # 1024- and 4096-ray items fold never or once
    _, D, _, _ = cpd.generate_designs(4, 0, 6)
    S = N.Sampling(4, 65536, N.MODE_FIXED, 0)
    ctx = cpd.Context(0)
    ref, st = cpd.fitness_batch(D, S, ctx=ctx)
    assert (st == 0).all()
    for tile in (1024, 4096, 16384):
        ctx.set_schedule(N.Schedule(tile, 16, N.ORDER_GIVEN, 0))
        got, _ = cpd.fitness_batch(D, S, ctx=ctx)
        assert np.array_equal(got, ref), tile
    ctx.set_schedule(None)


def test_autotune_returns_a_candidate_schedule():
    D, S = uneven(32, seed=1)
    best, ms = cpd.autotune(D, S)
    assert ms > 0
    assert best.tile_rays in (0, 16384, 8192, 4096)
    assert best.waves_per_cu in (4, 8, 16, 20, 24)
    assert best.order in (N.ORDER_GIVEN, N.ORDER_COST)


def test_plan_replays_match_eager_batches():
    D, S = uneven(40, seed=2)
    # a second generation of the same shape: per slot a design with the same bin count
    _, pool, _, _ = cpd.generate_designs(2, 1000, 400)
    by_bins = {}
    for g in pool:
        by_bins.setdefault(g.bin_count, []).append(g)
    D2 = [by_bins[g.bin_count].pop() for g in D]
    with cpd.Plan(D, S) as plan:
        sched = plan.schedule()
        assert sched.waves_per_cu >= 1
        for gen in (D, D2, D, D2):
            want, _ = cpd.fitness_batch(gen, S)  # eager work between replays
            got, st = plan.run(gen)
            assert (st == 0).all()
            assert np.array_equal(got, want)


def test_plan_rejects_a_design_of_another_shape():
    D, S = uneven(12, seed=3)
    with cpd.Plan(D, S, schedule=N.Schedule(0, 8, N.ORDER_COST, 0)) as plan:
        gen = list(D)
        _, other, _, _ = cpd.generate_designs(2, 5000, 64)
        gen[5] = next(g for g in other if g.bin_count != D[5].bin_count)
        got, st = plan.run(gen)
        assert st[5] == N.ERR_INVALID_ARGUMENT and np.isnan(got[5]).all()
        want, _ = cpd.fitness_batch(D, S)
        keep = np.arange(12) != 5
        assert (st[keep] == 0).all()
        assert np.array_equal(got[keep], want[keep])
