"""The generation-level optimiser client on the GPU: every generation is one
cpd_plan_run_params call (parameter vectors built and evaluated on the
device), and its objectives agree with building the same designs
(cpd_build_designs) and evaluating them eagerly (cpd_fitness_batch)."""
import math

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import optimize as O

pytestmark = pytest.mark.gpu


def problem():
    return O.SpectrometerProblem(O.synthetic_catalog(12, seed=5), 3, n_wavelengths=16, n_rays=4096)


def test_generation_objectives_match_eager_path():
    pb = problem()
    try:
        X = np.random.default_rng(11).random((96, pb.n_var))
        F, G, fit, st = pb.evaluate(X)
        params = pb.params(X)
        designs, bst = cpd.build_designs(params, pb.acceptance)
        np.testing.assert_array_equal(st, bst)
        ok = np.flatnonzero(bst == 0)
        assert ok.size >= 1
        eager, est = cpd.fitness_batch([designs[i] for i in ok], (16, 4096))
        assert (est == 0).all()
        np.testing.assert_array_equal(fit[ok], eager)
        max_info = math.log2(pb.bin_count)
        for i in range(len(X)):
            if st[i] != 0:
                assert F[i].tolist() == [pb.max_height * 400, max_info, 1.0] and G[i] == 1
            elif max_info - fit[i, 1] > 0:
                assert F[i].tolist() == [fit[i, 0], max_info - fit[i, 1], fit[i, 2]]
                assert G[i] == (0.0 if pb.max_height / 2 < fit[i, 0] < pb.max_height * 40 else 1.0)
    finally:
        pb.close()


def test_evolve_keeps_its_best():
    pb = problem()
    try:
        X, F, G, hist = O.evolve(pb, pop_size=48, generations=4, seed=3)
        assert len(hist) == 5 and X.shape == (48, pb.n_var) and np.isfinite(F).all()
        feas = [h["feasible"] for h in hist]
        assert all(b >= a for a, b in zip(feas, feas[1:]))  # elitist: feasible survivors are kept
        best = [h["best"] for h in hist if h["best"] is not None]
        for a, b in zip(best, best[1:]):
            assert all(y <= x for x, y in zip(a, b))  # per-objective minimum over the feasible set
    finally:
        pb.close()
