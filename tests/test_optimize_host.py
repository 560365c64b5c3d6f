"""Host logic of the generation-level optimiser client (no GPU): the mapping
of normalised variables to cpd_params (designer.py:115-265), the
multi-objective survival, and the Sellmeier catalogue fit."""
import math

import numpy as np

from compound_prism_designer import optimize as O


def test_variable_mapping():
    cat = O.synthetic_catalog(5, seed=3)
    pb = O.SpectrometerProblem(cat, 2, max_height=20.0)
    assert pb.n_var == 3 * 2 + 5 and pb.n_obj == 3
    x = np.array([[0.0, 0.999, 0.0, 0.5, 1.0, 0.0, 0.5, 0.25, 0.5, 0.5, 0.75]])
    p = pb.params(x)[0]
    assert p.n_glass == 2
    assert list(p.glass[0]) == cat[0].glass and list(p.glass[1]) == cat[4].glass
    assert [p.angles[i] for i in range(3)] == [-math.pi / 2, 0.0, math.pi / 2]
    assert p.lengths[0] == 0.0 and math.isclose(p.lengths[1], math.atanh(0.5) * 20.0 / 4)
    assert math.isclose(p.curvature, 1e-5 + (1 - 1e-5) * 0.25)
    assert math.isclose(p.height, 20.0 * (1e-4 + (1 - 1e-4) * 0.5)) and math.isclose(p.y_mean, p.height * 0.5)
    assert math.isclose(p.det_angle, math.pi / 2)
    assert p.bin_count == 32 and p.wavelength_lo == 0.5 and p.wavelength_hi == 0.82


def test_nondominated_ranks_and_survival():
    F = np.array([[1.0, 2.0], [2.0, 1.0], [3.0, 3.0], [1.5, 1.5], [0.5, 4.0], [3.0, 3.5]])
    np.testing.assert_array_equal(O.nondominated_ranks(F), [0, 0, 1, 0, 0, 2])
    # brute force: i is in front 0 iff no j dominates it
    rng = np.random.default_rng(1)
    R = rng.random((40, 3))
    ranks = O.nondominated_ranks(R)
    for i in range(40):
        dominated = any((R[j] <= R[i]).all() and (R[j] < R[i]).any() for j in range(40))
        assert (ranks[i] == 0) == (not dominated)
    # infeasible rows go last, whatever their objectives
    G = np.zeros(6)
    G[4] = 1
    keep = O.survive(F, G, 5)
    assert 4 not in keep and set(keep) >= {0, 1, 3}


def test_sellmeier_catalog_is_smooth_and_dispersive():
    cat = O.synthetic_catalog(4, seed=2)
    w = np.linspace(0.5, 0.82, 9)
    for g in cat:
        n = np.array([g(x) for x in w])
        assert (n > 1.0).all() and (np.diff(n) < 0).all()  # normal dispersion over the range


def test_explicit_position_variables():
    # designer.py:131-132, 236-243: without automatic positioning two more
    # variables give the detector's distance ([0, 40 max_height]) and angle
    # ([-pi/2, pi/2]) from the exit face
    import pytest
    from compound_prism_designer import _native as N

    cat = O.synthetic_catalog(5, seed=3)
    pb = O.SpectrometerProblem(cat, 2, max_height=20.0, acceptance=None)
    assert pb.n_var == 3 * 2 + 5 + 2 and pb.build_acceptance == N.POSITION_FROM_EXIT
    x = np.full((1, pb.n_var), 0.5)
    x[0, -2:] = [0.25, 1.0]
    p = pb.params(x)[0]
    assert p.det_position[0] == 0.25 * 800.0 and p.det_position[1] == math.pi / 2 and p.flipped == 0
    with pytest.raises(ValueError):
        O.SpectrometerProblem(cat, 2, acceptance=0.0)
    with pytest.raises(ValueError):
        O.SpectrometerProblem(cat, 2, estimator="pymoo")
