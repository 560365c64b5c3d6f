"""Design construction on the GPU (cpd_build_designs, cpd_plan_run_params)
against the host construction (cpd_position_detector_array +
cpd_build_design, the FP64 restatement of CompoundPrism::new,
detector_array_positioning and LinearDetectorArray::new).

Both sides run the same construction source; only the device's cos and sin
may differ from the host's in the last bit, so records agree to 1e-12
relative (most bit for bit) and every status agrees."""
import ctypes

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N

pytestmark = pytest.mark.gpu

REL = 1e-12


def doubles(d: N.Design) -> np.ndarray:
    return np.frombuffer(bytes(d)[16:], np.float64)


def ints(d: N.Design):
    return (d.n_glass, d.bin_count, d.ar_coated)


def host_status(p: N.Params, acceptance: float) -> int:
    pos = (ctypes.c_double * 2)()
    fl = ctypes.c_uint32()
    st = N.lib.cpd_position_detector_array(ctypes.byref(p), acceptance, pos, ctypes.byref(fl))
    if st != 0:
        return st
    q = N.Params()
    ctypes.memmove(ctypes.byref(q), ctypes.byref(p), ctypes.sizeof(q))
    q.det_position[0], q.det_position[1], q.flipped = pos[0], pos[1], fl.value
    return N.lib.cpd_build_design(ctypes.byref(q), ctypes.byref(N.Design()))


@pytest.mark.parametrize("config", [0, 2, 4])
@pytest.mark.parametrize("acceptance", [1.0, 0.0])
def test_device_records_match_host(config, acceptance):
    # generated parameters carry the host's position for acceptance 1, so both
    # the positioned (1.0) and the given-position (0.0) paths rebuild D
    P, D, _, _ = cpd.generate_designs(config, 0, 256)
    dev, st = cpd.build_designs(P, acceptance)
    assert (st == 0).all()
    identical = 0
    for a, b in zip(dev, D):
        assert ints(a) == ints(b)
        da, db = doubles(a), doubles(b)
        np.testing.assert_allclose(da, db, rtol=REL, atol=1e-300)
        identical += bytes(a) == bytes(b)
    assert identical >= len(D) // 2


def test_statuses_match_host_on_raw_draws():
    # angles redrawn at random: most draws fail (total internal reflection,
    # no lens or detector intersection, too little transmission)
    rng = np.random.default_rng(7)
    P, _, _, _ = cpd.generate_designs(2, 0, 512)
    for p in P:
        for i in range(p.n_glass + 1):
            p.angles[i] = rng.uniform(-np.pi / 2, np.pi / 2)
        p.det_angle = rng.uniform(-np.pi, np.pi)
    _, st = cpd.build_designs(P, 1.0)
    host = np.array([host_status(p, 1.0) for p in P])
    assert (host != 0).sum() > 100 and (host == 0).sum() > 5
    np.testing.assert_array_equal(st, host)


def test_invalid_parameters():
    P, _, _, _ = cpd.generate_designs(0, 0, 7)
    P[0].n_glass = 0
    P[1].n_glass = 7
    P[2].bin_count = 0
    P[3].linear_slope = 0.0
    P[4].curvature = 0.0
    P[5].height = -1.0
    dev, st = cpd.build_designs(P, 1.0)
    assert list(st[:6]) == [N.ERR_INVALID_ARGUMENT] * 6
    assert st[6] == 0
    with pytest.raises(Exception):
        cpd.build_designs(P, 1.5)


def test_plan_generation_from_parameters():
    P, D, S, _ = cpd.generate_designs(2, 0, 48)
    plan = cpd.Plan(D, S)
    try:
        dev, st = cpd.build_designs(P, 1.0)
        assert (st == 0).all()
        ref, rst = plan.run(dev)
        assert (rst == 0).all()
        # the same device construction inside the captured graph: bit for bit
        for _ in range(3):
            res, pst = plan.run_params(P, 1.0)
            assert (pst == 0).all()
            np.testing.assert_array_equal(res, ref)
        # against the host-built records: within the fitness tolerance
        host, _ = plan.run(D)
        assert np.max(np.abs(res[:, 1] - host[:, 1])) <= 1e-4
        np.testing.assert_allclose(res[:, [0, 2]], host[:, [0, 2]], rtol=1e-9)
        # a slot with another bin count, and one that fails to position
        P[3].bin_count = D[3].bin_count * 2
        P[5].angles[0] = 1.5
        P[5].angles[1] = -1.5
        res2, pst2 = plan.run_params(P, 1.0)
        assert pst2[3] == N.ERR_INVALID_ARGUMENT and np.isnan(res2[3]).all()
        assert pst2[5] != 0 and np.isnan(res2[5]).all()
        ok = [i for i in range(len(P)) if i not in (3, 5)]
        np.testing.assert_array_equal(res2[ok], ref[ok])
    finally:
        plan.close()


def oracle_record(p: N.Params, acceptance: float) -> N.Design:
    # the oracle's construction (CompoundPrism::new, detector_array_positioning,
    # LinearDetectorArray::new restated in FP64): the detector positioned by the
    # oracle first when acceptance > 0
    from oracle import oracle as O

    q = O._as(p, O.Params)
    if acceptance > 0:
        st, pos, fl = O.position_detector_array(q, acceptance)
        assert st == 0
        q.det_position[0], q.det_position[1], q.flipped = pos[0], pos[1], int(fl)
    d = N.Design()
    ctypes.memmove(ctypes.byref(d), ctypes.byref(O.build_design(q)), ctypes.sizeof(d))
    return d


@pytest.mark.parametrize("config", [0, 2, 4, 5, 6, 7])
def test_device_records_match_oracle(config):
    P, _, _, _ = cpd.generate_designs(config, 0, 128)
    dev, st = cpd.build_designs(P, 1.0)
    assert (st == 0).all()
    for p, a in zip(P, dev):
        b = oracle_record(p, 1.0)
        assert ints(a) == ints(b)
        np.testing.assert_allclose(doubles(a), doubles(b), rtol=REL, atol=1e-300)


def test_device_sellmeier_fits_match_golden_and_host():
    # glasscat.py:114-127: fitted indices within 1e-6 of scipy's curve_fit at the
    # golden check wavelengths, and a residual no larger than scipy's; the
    # device's Givens least squares and the host's Householder one reach the
    # same optimum (indices within 1e-10 of each other)
    import json
    from pathlib import Path

    from compound_prism_designer.optimize import fit_sellmeier_catalog

    cases = json.loads((Path(__file__).resolve().parent / "golden" / "sellmeier.json").read_text())
    ctx = cpd.Context(0)
    try:
        for lo, hi in sorted({(cs["lo"], cs["hi"]) for cs in cases}):
            sel = [cs for cs in cases if (cs["lo"], cs["hi"]) == (lo, hi)]
            glasses, chisq = fit_sellmeier_catalog(ctx, [cs["b"] for cs in sel], [cs["c"] for cs in sel], (lo, hi))
            for cs, g, x2 in zip(sel, glasses, chisq):
                n_dev = np.array([g(w) for w in cs["check_w"]])
                np.testing.assert_allclose(n_dev, cs["check_n_fit"], rtol=0, atol=1e-6)
                assert x2 <= cs["chisq"] * (1 + 1e-6) + 1e-18
                b = (ctypes.c_double * 3)(*cs["b"])
                c = (ctypes.c_double * 3)(*cs["c"])
                host = (ctypes.c_double * 6)()
                assert N.lib.cpd_fit_sellmeier1(b, c, lo, hi, 100, host, None) == 0
                h = cpd.Glass("host", list(host))
                np.testing.assert_allclose(n_dev, [h(w) for w in cs["check_w"]], rtol=0, atol=1e-10)
        # a catalogue of 4096 glasses in one call; a pole inside the range fails by name
        rng = np.random.default_rng(3)
        b, c = rng.uniform(0.3, 1.5, (4096, 3)), rng.uniform(0.001, 0.2, (4096, 3))
        glasses, chisq = fit_sellmeier_catalog(ctx, b, c)
        assert len(glasses) == 4096 and np.isfinite(chisq).all()
        for i in (0, 1000, 4095):
            bb, cc = (ctypes.c_double * 3)(*b[i]), (ctypes.c_double * 3)(*c[i])
            host = (ctypes.c_double * 6)()
            assert N.lib.cpd_fit_sellmeier1(bb, cc, 0.5, 0.82, 100, host, None) == 0
            ws = np.linspace(0.5, 0.82, 33)
            np.testing.assert_allclose([glasses[i](w) for w in ws], [cpd.Glass("h", list(host))(w) for w in ws],
                                       rtol=0, atol=1e-10)
        b[7, 1], c[7, 1] = 1.0, 0.6 ** 2
        with pytest.raises(ValueError, match="G007"):
            fit_sellmeier_catalog(ctx, b[:16], c[:16])
    finally:
        ctx.close()
