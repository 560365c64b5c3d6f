"""The generation-level optimiser client on the GPU.

Every estimator of SpectrometerProblem ("plan": a captured generation;
"server": requests to the continuous-batching server; "gpu_fitness": the
designer's estimator, designer.py:276) is checked against the oracle on the
same parameter vectors: the oracle builds each design (CompoundPrism::new,
detector_array_positioning or the designer's exit-face placement,
designer.py:236-243, in FP64) and evaluates it (fitness.rs:130-147 fixed
count, or the GPU estimator's FP64 restatement).  Populations come from a few
generations of `evolve`, so most of their designs transmit."""
import math

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from compound_prism_designer import optimize as O
from oracle import oracle as OR

pytestmark = pytest.mark.gpu

INFO_TOL = 1e-4
REL = 1e-9


def problem(**kw):
    return O.SpectrometerProblem(O.synthetic_catalog(12, seed=5), 3, n_wavelengths=16, n_rays=4096, **kw)


_POP = {}


def evolved():
    """A population after 12 generations (auto-positioned detectors)."""
    if "X" not in _POP:
        pb = problem()
        try:
            X, F, G, _ = O.evolve(pb, pop_size=128, generations=12, seed=3)
        finally:
            pb.close()
        _POP["X"] = X
    return _POP["X"]


def exit_relative(pb_auto, X):
    """The same individuals with their detectors placed from the exit face at
    the auto-positioned location: distance and angle of (position - face
    midpoint) against the face normal, where that angle is in [-pi/2, pi/2]
    and the positioned array is not flipped (the exit-face placement never
    flips it)."""
    designs, st = cpd.build_designs(pb_auto.params(X), pb_auto.acceptance)
    H = pb_auto.max_height
    P = pb_auto.nglass
    rows = []
    for x, d, s in zip(X, designs, st):
        # direction = normal rotated by -90 degrees unless flipped
        if s != 0 or d.det_direction[0] != d.det_normal[1]:
            continue
        a = -math.pi / 2 + math.pi * x[P + P]  # last face angle
        vx, vy = d.det_position[0] - d.arc_midpt[0], d.det_position[1] - d.arc_midpt[1]
        dv = math.hypot(vx, vy)
        da = math.atan2(vy, vx) - a
        da = (da + math.pi) % (2 * math.pi) - math.pi
        if abs(da) >= math.pi / 2 or dv >= 40 * H:
            continue
        rows.append(np.concatenate([x, [dv / (40 * H), (da + math.pi / 2) / math.pi]]))
    return np.array(rows)


def oracle_design(pb, p):
    """The oracle's record for cpd_params p under the problem's placement rule,
    or None where construction fails."""
    q = OR._as(p, OR.Params)
    if pb.acceptance is not None:
        st, pos, fl = OR.position_detector_array(q, pb.acceptance)
        if st != 0:
            return None
        q.det_position[0], q.det_position[1], q.flipped = pos[0], pos[1], int(fl)
    else:
        try:
            prism = OR.build_design(q)
        except ValueError:
            return None
        # designer.py:236-243 in numpy
        dv, da = q.det_position[0], q.det_position[1]
        a = q.angles[q.n_glass]
        R = np.array(((np.cos(da), -np.sin(da)), (np.sin(da), np.cos(da))))
        v = R @ (np.cos(a), np.sin(a))
        q.det_position[0] = prism.arc_midpt[0] + v[0] * dv
        q.det_position[1] = prism.arc_midpt[1] + v[1] * dv
        q.flipped = 0
    try:
        return OR.build_design(q)
    except ValueError:
        return None


def population(acceptance):
    X = evolved()
    if acceptance is None:
        pa = problem()
        try:
            X = exit_relative(pa, X)
        finally:
            pa.close()
    return X


@pytest.mark.parametrize("estimator", ["plan", "server"])
@pytest.mark.parametrize("acceptance", [1.0, None])
def test_client_objectives_match_oracle_fitness(estimator, acceptance):
    X = population(acceptance)
    pb = problem(estimator=estimator, acceptance=acceptance)
    try:
        F, G, fit, st = pb.evaluate(X)
        params = pb.params(X)
        refs = [oracle_design(pb, p) for p in params]
        np.testing.assert_array_equal(st == 0, [r is not None for r in refs])
        ok = np.flatnonzero(st == 0)
        # the device-built records are the oracle's (cos and sin to the last bit)
        dev, _ = cpd.build_designs(params, pb.build_acceptance)
        for i in ok:
            a = np.frombuffer(bytes(dev[i])[16:], np.float64)
            b = np.frombuffer(bytes(refs[i])[16:], np.float64)
            np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-300)
        ref, _ = OR.fitness_many([refs[i] for i in ok], 16, 4096)
        np.testing.assert_allclose(fit[ok][:, [0, 2]], ref[:, [0, 2]], rtol=REL)
        err = np.abs(fit[ok, 1] - ref[:, 1])
        transmitting = int((ref[:, 1] > 0.05).sum())
        print(f"{estimator}, acceptance {acceptance}: {len(X)} individuals, {ok.size} built, "
              f"{transmitting} with info > 0.05 bits, max |dI| {err.max():.2e}")
        assert err.max() <= INFO_TOL
        assert ok.size >= len(X) // 2 and transmitting >= (8 if acceptance else 4)
        # objectives and constraint from the fitness, as designer.py:267-300
        max_info = math.log2(pb.bin_count)
        for i in ok:
            if max_info - fit[i, 1] > 0:
                assert F[i].tolist() == [fit[i, 0], max_info - fit[i, 1], fit[i, 2]]
        if acceptance is None:
            # placed where the positioning put them (to rounding): the same fitness
            pa = problem()
            try:
                auto, ast = pa.fitness(X[:, :-2])
            finally:
                pa.close()
            assert (ast[ok] == 0).all()
            np.testing.assert_allclose(fit[ok][:, [0, 2]], auto[ok][:, [0, 2]], rtol=1e-9)
            assert np.abs(fit[ok, 1] - auto[ok, 1]).max() <= INFO_TOL
    finally:
        pb.close()


def test_plan_and_server_estimators_agree_bit_for_bit():
    X = evolved()
    a, b = problem(estimator="plan"), problem(estimator="server")
    try:
        fa, sa = a.fitness(X)
        fb, sb = b.fitness(X)
        np.testing.assert_array_equal(sa, sb)
        ok = sa == 0
        np.testing.assert_array_equal(fa[ok], fb[ok])
    finally:
        a.close()
        b.close()


def test_designer_gpu_fitness_estimator_matches_oracle():
    # designer.py:276: gpu_fitness(seeds=rand(1), max_n=128, max_eval=8192)
    X = evolved()
    pb = problem(estimator="gpu_fitness", seed=21)
    try:
        fit, st = pb.fitness(X)
        seeds = np.random.default_rng(21).random(1)
        params = pb.params(X)
        built = 0
        agree = 0
        for i, p in enumerate(params):
            d = oracle_design(pb, p)
            if d is None:
                assert st[i] not in (0, N.ERR_NOT_CONVERGED)
                continue
            built += 1
            ref, converged, _ = OR.gpu_fitness(d, seeds, 128, 2, 8192)
            agree += (st[i] == 0) == converged
            # records built on the device (its cos and sin) against the oracle's
            np.testing.assert_allclose(fit[i, [0, 2]], [ref[0], ref[2]], rtol=REL)
            assert abs(fit[i, 1] - ref[1]) <= INFO_TOL, (i, fit[i], ref)
        assert built >= len(X) // 2
        # a design whose bin statistics sit on the SEM threshold may converge on one side only
        assert agree >= built - 1
        F, G, _, st2 = pb.evaluate(X)
        # a design that does not converge gets the RayTraceError objectives
        for i in np.flatnonzero(st2 == N.ERR_NOT_CONVERGED):
            assert F[i].tolist() == [pb.max_height * 400, math.log2(pb.bin_count), 1.0]
    finally:
        pb.close()


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
