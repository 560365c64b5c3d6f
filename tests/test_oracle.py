"""Pins the CPU oracle (oracle/cpd_oracle.cpp) against the reference's own known
answers and against what the mathematics fixes, so that a dropped term, a
wrong sign or a transposed operand fails here."""
import ctypes
import json
import math
from decimal import Decimal, getcontext
from fractions import Fraction
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O
import kat

GOLDEN = Path(__file__).resolve().parent / "golden"


def kat_design(precision=64):
    p = kat.fill_params(O.Params())
    st, pos, fl = O.position_detector_array(p, 1.0)
    assert st == 0
    p.det_position[0], p.det_position[1] = pos
    p.flipped = int(fl)
    return p, O.build_design(p)


# ---------------------------------------------------------------------------
# reference known answers
# ---------------------------------------------------------------------------

def test_known_design_fitness_f64_adaptive():
    # src/fitness.rs:250-257: fitness(&spec, 16_384, 16_384) in f64
    _, d = kat_design()
    (size, info, dev), rays = O.fitness(d, 16384, 16384, 64, adaptive=True)
    assert abs(size - kat.KAT_SIZE) <= kat.KAT_SIZE_RMAX * kat.KAT_SIZE
    assert abs(info - kat.KAT_INFO) <= kat.KAT_INFO_RMAX * kat.KAT_INFO
    assert abs(dev - kat.KAT_DEVIATION) <= kat.KAT_DEVIATION_RMAX * kat.KAT_DEVIATION
    assert rays < 16384 * 16384  # the SEM early exits fired


def test_known_design_f32_matches_f64():
    # benches/benchmark.rs:147-157: the f32 evaluation agrees on info to rmax 1e-2;
    # the f32 path builds the design in f32 as the bench does
    p = kat.fill_params(O.Params())
    d32 = O.build_design_f32(p, 1.0)
    (_, info32, _), _ = O.fitness(d32, 16384, 16384, 32, adaptive=True)
    _, d = kat_design()
    (_, info64, _), _ = O.fitness(d, 16384, 16384, 64, adaptive=True)
    assert abs(info32 - info64) <= 1e-2 * abs(info64)


def test_norminv_against_ndtri():
    # src/spectrometer/erf.rs:142-161 with scipy.special.ndtri as the true quantile
    for case in json.loads((GOLDEN / "norminv.json").read_text()):
        z = O.norminv(case["u"])
        assert abs(z - case["z"]) <= case["rmax"] * max(abs(z), abs(case["z"])), case


def test_norminv_symmetry_f32():
    for u in [2.0 ** -12, 0.0078125, 0.25, 0.4921875]:  # 1 - u exact in f32
        assert O.norminv(u, 32) == pytest.approx(-O.norminv(1 - u, 32), rel=1e-5)


# ---------------------------------------------------------------------------
# QMC: u_k = frac(0.5 + k / phi)
# ---------------------------------------------------------------------------

def test_qmc_closed_form():
    getcontext().prec = 50
    alpha = (Decimal(5).sqrt() - 1) / 2
    n = 16384
    it = O.qmc(n, O.QRNG_ITERATIVE, 64)
    fx = O.qmc(n, O.QRNG_FIXED_POINT, 64)
    for k in [1, 2, 3, 100, 1000, 4096, 16384]:
        exact = float((Decimal("0.5") + k * alpha) % 1)
        assert abs(fx[k - 1] - exact) <= 2 ** -52
        assert abs(it[k - 1] - exact) <= 1e-11
    # first sample of the reference sequence: 0.5 + 0.618... - 1
    assert it[0] == pytest.approx(0.1180339887498949, abs=1e-15)


# ---------------------------------------------------------------------------
# estimator logic against brute force on the oracle's own per-ray outcomes
# ---------------------------------------------------------------------------

def _rays(design, wavelength, m):
    u = O.qmc(m, O.QRNG_FIXED_POINT, 64)
    y = np.array([design.y_mean - design.beam_width * O.norminv(float(v)) for v in u])
    return O.propagate_many(design, wavelength, y, 64)


def test_fixed_count_p_dets_is_mean_over_rays():
    # fitness.rs:5-34 with no early exit: p(d|l) = sum_k t_k [bin_k = d] / M, except
    # that the bin hit by the last ray carries one extra zero sample (count M + 1)
    _, d = kat_design()
    m = 2000
    for w in (0.55, 0.66, 0.8):
        bins, ts = _rays(d, w, m)
        want = np.zeros(32)
        for b, t in zip(bins, ts):
            if b < 32:
                want[b] += t
        want /= m
        if bins[-1] < 32:
            want[bins[-1]] *= m / (m + 1)
        got, rays = O.p_dets_l_wavelength(d, w, m, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
        assert rays == m
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15)


def test_mutual_information_from_rows():
    # fitness.rs:79-123: I = mean_l sum_d p log2 p - sum_d pbar log2 pbar, checked
    # with exact rationals of the rows (to 1e-12 of the float rows)
    _, d = kat_design()
    n, m = 6, 1500
    u = O.qmc(n, O.QRNG_FIXED_POINT, 64)
    rows = []
    for j in range(n):
        w = d.wavelength_lo + (d.wavelength_hi - d.wavelength_lo) * u[j]
        rows.append(O.p_dets_l_wavelength(d, w, m, 64, qrng=O.QRNG_FIXED_POINT)[0])
    rows = np.array(rows)
    plog = lambda p: p * math.log2(p) if p > 0 else 0.0
    h_cond = sum(sum(plog(p) for p in r) for r in rows) / n
    pbar = [sum(Fraction(float(rows[j, b])) for j in range(n)) / n for b in range(32)]
    h_det = sum(plog(float(p)) for p in pbar)
    (_, info, _), _ = O.fitness(d, n, m, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
    assert info == pytest.approx(h_cond - h_det, abs=1e-12)


def constant_glass_design(P=3, nbin=32):
    """A design whose glasses do not disperse: p(d|l) cannot depend on l."""
    p = kat.fill_params(O.Params())
    p.n_glass = P
    for i in range(P):
        for j in range(6):
            p.glass[i][j] = 0.0
        p.glass[i][5] = 1.5 + 0.1 * i
    for i in range(P + 1):
        p.angles[i] = math.radians([-20, 25, -30, 10, 0, 0, 0][i])
    p.bin_count = nbin
    p.det_angle = 0.0
    p.det_position[0], p.det_position[1] = 30.0, -1.0
    return p


def test_information_is_zero_without_dispersion():
    p = constant_glass_design()
    d = O.build_design(p)
    (_, info, _), _ = O.fitness(d, 16, 3000, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
    assert abs(info) <= 1e-12


def test_information_bounds_on_known_design():
    _, d = kat_design()
    (_, info, _), _ = O.fitness(d, 32, 4096, 64, adaptive=False, qrng=O.QRNG_FIXED_POINT)
    assert 0.0 < info <= math.log2(32)


# ---------------------------------------------------------------------------
# optics: Snell, Fresnel, total internal reflection, detector bins
# ---------------------------------------------------------------------------

def _unit(v):
    return v / np.linalg.norm(v)


def test_snell_law_along_random_paths():
    # n1 (d_in x n) = n2 (d_out x n) at every surface crossed (ray.rs:309-367)
    _, d = kat_design()
    rng = np.random.default_rng(3)
    checked = 0
    for _ in range(200):
        w = rng.uniform(0.5, 0.82)
        y = rng.uniform(0.2, 2.3)
        path = O.ray_trace(d, w, y)
        ns = [1.0] + [sum(c * w ** (5 - k) for k, c in enumerate(d.glass[i])) for i in range(3)] + [1.0]
        for s in range(1, min(len(path), 5)):
            din, dout = path[s - 1, 2:], path[s, 2:]
            if s <= 3:
                nrm = np.array(d.plane_normal[s - 1])
            else:
                nrm = _unit(np.array(d.arc_center) - path[s, :2])
            cross = lambda a, b: a[0] * b[1] - a[1] * b[0]
            assert ns[s - 1] * cross(din, nrm) == pytest.approx(ns[s] * cross(dout, nrm), abs=1e-12)
            assert np.linalg.norm(dout) == pytest.approx(1.0, abs=1e-12)
            checked += 1
    assert checked > 300


def single_face_params(exit_angle_deg, n=1.5, k=1e-7):
    """One glass, entrance face at 0 deg, flat-ish lens; detector at x = 30
    facing -x, running upwards from y = -0.4 (pos = y + 0.4)."""
    p = kat.fill_params(O.Params())
    p.n_glass = 1
    for j in range(6):
        p.glass[0][j] = 0.0
    p.glass[0][5] = n
    p.angles[0] = 0.0
    p.angles[1] = math.radians(exit_angle_deg)
    p.lengths[0] = 1.0
    p.curvature = k
    p.det_angle = 0.0
    p.det_position[0], p.det_position[1] = 30.0, -0.4
    p.flipped = 0
    p.max_incident_angle = math.radians(60)
    return p


def test_normal_incidence_fresnel_and_bin():
    # both surfaces at normal incidence: T = (1 - ((n-1)/(n+1))^2)^2 for s and p
    n = 1.5
    d = O.build_design(single_face_params(0.0, n))
    st, b, t = O.propagate(d, 0.6, d.height / 2)
    assert st == 0
    r = ((n - 1) / (n + 1)) ** 2
    assert t == pytest.approx((1 - r) ** 2, rel=1e-12)
    # the ray stays at y = h/2 = 1.25: position 1.65 on the array -> bin 16
    assert b == 16


def test_total_internal_reflection_threshold():
    # inside glass n = 1.5 the ray meets the exit face at its angle; TIR iff n sin > 1
    n = 1.5
    crit = math.degrees(math.asin(1 / n))
    d_tir = O.build_design(single_face_params(crit + 2.0, n))
    st, _, _ = O.propagate(d_tir, 0.6, d_tir.height / 2)
    assert st == 3  # TotalInternalReflection
    d_ok = O.build_design(single_face_params(crit - 8.0, n))
    path = O.ray_trace(d_ok, 0.6, d_ok.height / 2)
    assert len(path) >= 3
    # Snell at the exit: sin(out) = n sin(in) relative to the face normal
    nrm = -np.array([math.cos(math.radians(crit - 8.0)), math.sin(math.radians(crit - 8.0))])
    cross = lambda a, b: a[0] * b[1] - a[1] * b[0]
    assert abs(cross(path[2, 2:], nrm)) == pytest.approx(n * abs(cross(path[1, 2:], nrm)), rel=1e-6)


def test_detector_bins_along_the_array():
    # with flat optics the beam height maps straight onto the array: pos = y + 0.4
    d = O.build_design(single_face_params(0.0))
    for y, want in [(0.05, 4), (0.61, 10), (1.25, 16), (2.25, 26)]:
        st, b, _ = O.propagate(d, 0.6, y)
        assert st == 0 and b == want, (y, b)
    # outside the array's length: a miss (NoSurfaceIntersection)
    d.det_position[1] = -5.0
    st, b, _ = O.propagate(d, 0.6, 2.4)
    assert st == 1 and b == 32


# ---------------------------------------------------------------------------
# construction
# ---------------------------------------------------------------------------

def test_trapezoid_faces_are_separated_by_lengths():
    # faces i-1 and i are `lengths[i-1]` apart at their closest end, and face 0
    # touches x = 0 (geometry.rs:430-489)
    rng = np.random.default_rng(5)
    for _ in range(50):
        p = kat.fill_params(O.Params())
        P = int(rng.integers(1, 7))
        p.n_glass = P
        for i in range(P + 1):
            p.angles[i] = rng.uniform(-1.2, 1.2)
        for i in range(P):
            p.lengths[i] = rng.uniform(0.1, 2.0)
            for j in range(6):
                p.glass[i][j] = p.glass[0][j]
        p.height = rng.uniform(0.5, 2.0)
        d = O.build_design(p)
        h = p.height

        def x_at(i, y):
            if i < P:
                px, py = d.plane_point[i]
                nx, ny = d.plane_normal[i]
            else:
                # the exit lens' chord: its end points lie on the last face line
                nx, ny = -math.cos(p.angles[P]), -math.sin(p.angles[P])
                px, py = None, h / 2
                return None
            return px - (y - py) * ny / nx

        assert min(x_at(0, 0.0), x_at(0, h)) == pytest.approx(0.0, abs=1e-12)
        for i in range(1, P):
            gap = min(x_at(i, 0.0) - x_at(i - 1, 0.0), x_at(i, h) - x_at(i - 1, h))
            assert gap == pytest.approx(p.lengths[i - 1], abs=1e-9)


def test_exit_lens_geometry():
    # the arc passes through both chord ends, its bound centre lies on it, and
    # R = chord / (2 |k|) (geometry.rs:280-315)
    p, d = kat_design()
    c = np.array(d.arc_center)
    m = np.array(d.arc_midpt)
    R = d.arc_radius
    theta = p.angles[3]
    nrm = np.array([-math.cos(theta), -math.sin(theta)])
    h = p.height
    assert np.linalg.norm(m - c) == pytest.approx(R, rel=1e-12)
    chord = abs(h / nrm[0])
    assert R == pytest.approx(chord / (2 * abs(p.curvature)), rel=1e-12)
    # chord ends: on the face line at y = 0 and y = h; |end - m|^2 = max_dist_sq
    along = np.array([-nrm[1], nrm[0]])
    mid_chord = c - nrm * math.sqrt(R * R - chord * chord / 4) * np.sign(p.curvature)
    for s in (-0.5, 0.5):
        e = mid_chord + along * chord * s
        assert np.linalg.norm(e - c) == pytest.approx(R, rel=1e-12)
        assert np.sum((e - m) ** 2) == pytest.approx(d.arc_max_dist_sq, rel=1e-9)
    assert min(abs(mid_chord[1] + along[1] * chord * s) for s in (-0.5, 0.5)) == pytest.approx(0, abs=1e-9)


def test_detector_positioning_spans_the_extreme_wavelengths():
    # geometry.rs:495-542: with acceptance 1 the median rays of the lowest and
    # highest wavelength meet the two ends of the array
    p, d = kat_design()
    ends = []
    for w in (p.wavelength_lo, p.wavelength_hi):
        path = O.ray_trace(d, w, p.y_mean)
        o, u = path[4, :2], path[4, 2:]  # the ray leaving the lens
        P0, nrm = np.array(d.det_position), np.array(d.det_normal)
        s_hit = -np.dot(o - P0, nrm) / np.dot(u, nrm)
        pos = np.dot(o + s_hit * u - P0, np.array(d.det_direction))
        ends.append(pos)
    assert sorted(ends) == pytest.approx([0.0, p.det_length], abs=1e-9)


def test_layout_matches_header():
    assert O.lib().ora_sizeof_params() == ctypes.sizeof(O.Params)
    assert O.lib().ora_sizeof_design() == ctypes.sizeof(O.Design)


# ---------------------------------------------------------------------------
# the reference's GPU estimator (kernel.rs:89-193, cuda_fitness.rs:347-423)
# ---------------------------------------------------------------------------

def test_gpu_estimator_known_design_within_bench_tolerance():
    # benches/benchmark.rs:147-157: cuda_fitness(spec, [0.798713], 256, 2, 16384) converges
    # and agrees with fitness(spec, 16384, 16384) on info to rmax 1e-2
    _, d = kat_design()
    (size, info, dev), converged, (rows, rays) = O.gpu_fitness(d, [0.798713], 256, 2, 16384)
    (size_c, info_c, dev_c), _ = O.fitness(d, 16384, 16384, 64, adaptive=True)
    assert converged and rows == 256
    assert abs(info - info_c) <= 1e-2 * abs(info_c)
    assert (size, dev) == (size_c, dev_c)
    assert rays < 256 * 16384  # the per-row early exit fired
    # too few rows: the bin means over the rows stay imprecise and the reference returns None
    _, converged, (rows, _) = O.gpu_fitness(d, [0.798713], 4, 2, 16384)
    assert not converged and rows == 4


def test_gpu_row_is_the_mean_of_t_over_its_rays():
    # a row folds 32-ray chunks into per-bin summaries with the parallel Welford
    # combination and pads with zeros; its means must equal the plain mean of t
    # per bin over the rays folded (k = 32 .. 32 + rays - 1, u_k = frac(seed + k/phi))
    _, d = kat_design()
    alpha = 0x9E3779B97F4A7C16
    for lam, seed, max_eval in [(0.61, 0.3, 16384), (0.52, 0.798713, 2000), (0.8, 0.05, 64)]:
        row, rays = O.gpu_row(d, lam, seed, max_eval)
        assert rays % 32 == 0 and 32 <= rays <= max_eval - max_eval % 32
        seed_fx = int((seed - math.floor(seed)) * 2.0 ** 64)
        ks = np.arange(32, 32 + rays, dtype=np.uint64)
        u = [(((seed_fx + int(k) * alpha) % 2 ** 64) >> 11) * 2.0 ** -53 for k in ks]
        y = np.array([d.y_mean - d.beam_width * O.norminv(x) for x in u])
        bins, ts = O.propagate_many(d, np.full(rays, lam), y, precision=64)
        want = np.zeros(d.bin_count)
        hit = bins < d.bin_count
        np.add.at(want, bins[hit], ts[hit])
        want /= rays
        np.testing.assert_allclose(row, want, rtol=1e-12, atol=1e-15)


# ---------------------------------------------------------------------------
# anti-reflection coating (ray.rs:336-340) and gap / intercept bins
# (spectrometer.rs:291-303 with utils.rs:177-192 and the saturating u32 cast)
# ---------------------------------------------------------------------------

import flat  # noqa: E402


def test_anti_reflection_coating_at_normal_incidence():
    # both faces at ci = 1 > 0.5: Ts = Tp = 0.99 per face, t = 0.99^2, for
    # every ray whatever the glass; uncoated, the Fresnel closed form
    for n in (1.3, 1.5, 1.9):
        d = O.build_design(flat.flat_params(O.Params(), n=n, ar=True))
        for y in (0.3, 1.25, 2.2):
            st, b, t = O.propagate(d, 0.6, y)
            assert st == 0 and t == flat.T_AR, (n, y, t)
        d0 = O.build_design(flat.flat_params(O.Params(), n=n, ar=False))
        st, _, t0 = O.propagate(d0, 0.6, 1.25)
        assert st == 0 and t0 == pytest.approx((1 - ((n - 1) / (n + 1)) ** 2) ** 2, rel=1e-12)


def _fresnel_t(n1, n2, ci):
    cr = math.sqrt(1 - (n1 / n2) ** 2 * (1 - ci * ci))
    rs = (n1 * ci - n2 * cr) / (n1 * ci + n2 * cr)
    rp = (n1 * cr - n2 * ci) / (n1 * cr + n2 * ci)
    return 1 - rs * rs, 1 - rp * rp


def test_anti_reflection_threshold_is_strict_at_60_degrees():
    # the exit face tilted by theta: a ray along +x meets it at ci = cos(theta).
    # ci > 0.5 (theta < 60 deg) keeps the coating's 0.99; beyond, Fresnel applies
    # (n = 1.05, so no total internal reflection below 72 deg).  The detector is
    # re-aimed to face the exit ray 20 units past the lens.
    n = 1.05
    for theta, coated in [(59.9, True), (59.99, True), (60.01, False), (60.1, False), (65.0, False)]:
        p = flat.flat_params(O.Params(), n=n, ar=True, exit_angle_deg=theta)
        path = O.ray_trace(O.build_design(p), 0.6, 1.25)
        assert len(path) >= 3
        (ox, oy), (dx, dy) = path[2, :2], path[2, 2:]
        ang = math.atan2(dy, dx)
        p.det_angle = ang
        # the array's normal Rot(ang) (-1, 0) faces the ray; it starts 1.6 below it
        p.det_position[0] = ox + 20 * dx + 1.6 * math.sin(ang)
        p.det_position[1] = oy + 20 * dy - 1.6 * math.cos(ang)
        st, b, t = O.propagate(O.build_design(p), 0.6, 1.25)
        assert st == 0, (theta, st)
        # the lens is flat to 1e-7, so its normal at the hit is the face normal
        if coated:
            want = 0.99 * 0.99
        else:
            ts, tp = _fresnel_t(n, 1.0, math.cos(math.radians(theta)))
            want = 0.99 * (ts + tp) / 2
        assert t == pytest.approx(want, rel=1e-6), (theta, t, want)


def test_gap_bins_intercept_and_saturating_bin_zero():
    # flat optics: pos = y + 0.4.  Bins of 0.08 every 0.1 from intercept 0.65:
    # bin i covers [0.65 + 0.1 i, 0.73 + 0.1 i).  Before the intercept the
    # quotient is negative and `as u32` saturates it to bin 0, which counts the
    # ray when the euclidean remainder is below 0.08 (pos 0.47: -0.18 = 0.1 (-2)
    # + 0.02; pos 0.54: -0.11 = 0.1 (-2) + 0.09, a gap)
    p = flat.flat_params(O.Params(), ar=False)
    p.bin_size, p.linear_slope, p.linear_intercept = 0.08, 0.1, 0.65
    d = O.build_design(p)
    cases = [
        (0.07, 0),     # pos 0.47, remainder 0.02 of quotient -2: bin 0
        (0.14, None),  # pos 0.54, remainder 0.09: gap
        (0.005, 0),    # pos 0.405, -0.245 = 0.1 (-3) + 0.055: bin 0
        (0.04, None),  # pos 0.44, -0.21 = 0.1 (-3) + 0.09: gap
        (0.3, 0),      # pos 0.70: bin 0 proper
        (0.34, None),  # pos 0.74: gap between bins 0 and 1
        (0.5, 2),      # pos 0.90
        (0.535, None), # pos 0.935: gap
        (2.42, 21),    # pos 2.82
        (2.44, None),  # pos 2.84: gap past bin 21
    ]
    for y, want in cases:
        st, b, _ = O.propagate(d, 0.6, y)
        if want is None:
            assert st == 1 and b == p.bin_count, (y, st, b)
        else:
            assert st == 0 and b == want, (y, st, b)
    # 0.08 bins every 0.1 cover 80% of the array past the intercept
    ys = np.linspace(0.3, 2.45, 4001)
    bins, _ = O.propagate_many(d, 0.6, ys, 64)
    assert abs(np.mean(bins < p.bin_count) - 0.8) < 0.01


def test_last_bin_may_overhang_the_array():
    # bch.toml: 64 bins of 0.42 over 26.6: bin 63 starts at 26.46 and is cut by
    # the array end (pos > length misses before binning)
    p = flat.flat_params(O.Params(), ar=True)
    p.bin_count, p.bin_size, p.linear_slope, p.linear_intercept = 64, 0.42, 0.42, 0.0
    p.det_length = 26.6
    p.det_position[1] = -26.55 + 1.25  # pos = y + 25.3
    d = O.build_design(p)
    for y, want in [(1.2, 63), (1.28, 63), (1.31, None), (1.1, 62)]:
        st, b, _ = O.propagate(d, 0.6, y)
        assert (b if st == 0 else None) == want, (y, st, b)


# ---------------------------------------------------------------------------
# early exits: the stop points of fitness.rs:5-34 / 79-123 (with the Welford
# skip and SEM test of welford.rs:26-58) and of the GPU estimator's row fold
# (kernel.rs:89-193), derived in closed form for the flat designs
# ---------------------------------------------------------------------------

E = 5e-3


def test_inner_stop_single_bin_constant_t():
    # every ray lands in one bin with t = c.  After the first ray that bin's
    # Welford holds a phantom zero (skip(1) from count 0) and c; after k rays it
    # holds {0, c x k}: count k + 1, mean k c / (k + 1), m2 = k c^2 / (k + 1).
    # The SEM test m2 < e^2 count (count - 1) = e^2 (k + 1) k holds iff
    # k + 1 > c / e: with c = 0.9801, c / e = 196.02, so the loop breaks at
    # ray k = 196 with p = 196 c / 197.
    c = flat.T_AR
    k = next(k for k in range(1, 10 ** 6) if k + 1 > c / E)
    assert k == 196
    d = O.build_design(flat.single_bin(flat.flat_params(O.Params())))
    for w in (0.51, 0.8):
        row, rays = O.p_dets_l_wavelength(d, w, 16384, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
        assert rays == k
        want = np.zeros(32)
        want[16] = k * c / (k + 1)
        np.testing.assert_allclose(row, want, rtol=1e-15, atol=0)
    # without the phantom zero (count k, m2 0) the test would pass at k = 2;
    # with <= instead of < nothing changes here, so the next test covers it


def _inner_stop_two_bins(max_m):
    """Stop ray and per-bin means of fitness.rs:5-34 for the split design,
    derived per bin from its hit indices n_1 < n_2 < ... (all with value c).
    Just after hit j the bin holds j values c among n_j + 1 samples (the
    skips pad zeros and the first skip adds the phantom zero), so its mean is
    j c / (n_j + 1).  Its m2 grows at hit j by the skip term
    mu'^2 C (n_j - C) / n_j, with mu' = (j - 1) c / n_j the mean *after*
    padding to n_j and C = n_{j-1} + 1 the count before (welford.rs:37-39),
    plus the sample term (c - mu') (c - mean_j)."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    c, e2 = Decimal(flat.T_AR), Decimal(E) * Decimal(E)
    state = {0: (0, Decimal(0), 0), 1: (0, Decimal(0), 0)}  # j, m2, n_{j-1} + 1
    for k in range(1, max_m + 1):
        b = 1 if flat.r1_below_half(k) else 0
        j, m2, cnt = state[b]
        j += 1
        mu_pad = (j - 1) * c / k
        mean = j * c / (k + 1)
        m2 = m2 + mu_pad * mu_pad * cnt * (k - cnt) / k + (c - mu_pad) * (c - mean)
        state[b] = (j, m2, k + 1)
        if m2 < e2 * (k + 1) * k:
            return k, b, {bb: state[bb][0] for bb in (0, 1)}
    return max_m, None, {bb: state[bb][0] for bb in (0, 1)}


def test_inner_stop_two_split_bins():
    # each bin sees a pattern of hits and zero-padding; the loop breaks at the
    # first ray whose own bin passes the SEM test (fitness.rs:25-27), and the
    # final skip(count) leaves the other bin's mean at hits c / count
    k, b, hits = _inner_stop_two_bins(16384)
    assert b is not None and 1000 < k < 16384
    d = O.build_design(flat.split_bins(flat.flat_params(O.Params())))
    row, rays = O.p_dets_l_wavelength(d, 0.6, 16384, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    assert rays == k
    c = flat.T_AR
    want = [hits[bb] * c / (k + 1 if bb == b else k) for bb in (0, 1)]
    np.testing.assert_allclose(row, want, rtol=1e-13)


def test_inner_stop_distinguishes_the_skip_mean():
    # the same derivation with the pre-padding mean in the skip term (the
    # textbook update) stops elsewhere, so the test above pins the reference's
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    c, e2 = Decimal(flat.T_AR), Decimal(E) * Decimal(E)
    state = {0: (0, Decimal(0), 0), 1: (0, Decimal(0), 0)}
    stop = None
    for k in range(1, 16385):
        bb = 1 if flat.r1_below_half(k) else 0
        j, m2, cnt = state[bb]
        mu_old = (j * c / cnt) if cnt else Decimal(0)
        j += 1
        mu_pad = (j - 1) * c / k
        mean = j * c / (k + 1)
        m2 = m2 + mu_old * mu_old * cnt * (k - cnt) / k + (c - mu_pad) * (c - mean)
        state[bb] = (j, m2, k + 1)
        if m2 < e2 * (k + 1) * k:
            stop = k
            break
    assert stop != _inner_stop_two_bins(16384)[0]


def test_outer_stop_after_two_equal_rows():
    # fitness.rs:108-114: after one row every bin has m2 = 0 < e^2 * 1 * 0 = 0
    # is false; after two equal rows m2 = 0 < e^2 * 2 * 1 for every bin.  The
    # glass does not disperse, so rows are equal and the loop stops at two;
    # information is exactly 0.
    d = O.build_design(flat.single_bin(flat.flat_params(O.Params())))
    (size, info, dev), (rows, rays) = O.fitness_counts(d, 64, 16384, 64, adaptive=True,
                                                       qrng=O.QRNG_FIXED_POINT)
    assert (rows, rays) == (2, 2 * 196) and info == 0.0
    # a design that transmits nothing: two all-zero rows of max_m rays each
    p = flat.single_bin(flat.flat_params(O.Params()))
    p.det_position[1] = -50.0
    d0 = O.build_design(p)
    (_, info0, _), (rows0, rays0) = O.fitness_counts(d0, 64, 4096, 64, adaptive=True, qrng=O.QRNG_FIXED_POINT)
    assert (rows0, rays0) == (2, 2 * 4096) and info0 == 0.0


def _gpu_row_closed_form(seed, max_eval, split):
    """The GPU estimator's row (kernel.rs:113-192) for the flat designs.  Chunk i
    (i = 1, 2, ...) holds rays k = 32 i .. 32 i + 31 with u_k = frac(seed + k alpha).
    A bin touched by chunk i merges the chunk's (32, mean, m2 over all 32
    lanes) with Add; merged statistics are exact, so after the chunks that
    touched it the bin holds N = 32 x (those chunks) samples with h values c:
    mean h c / N, m2 = h c^2 (1 - h / N).  The row stops after the first chunk
    with a hit in which every touched bin has m2 < e^2 N (N - 1); the final
    skip(count) gives p = h c / count."""
    from decimal import Decimal, getcontext
    from fractions import Fraction
    getcontext().prec = 60
    alpha = (Decimal(5).sqrt() - 1) / 2
    c, e2 = Fraction(flat.T_AR), Fraction(E) * Fraction(E)
    sd = Decimal(seed)
    n_bins = 2 if split else 1
    h, N = [0] * n_bins, [0] * n_bins
    chunks = (max_eval - max_eval % 32) // 32
    for i in range(1, chunks + 1):
        lanes = [(1 if (sd + k * alpha) % 1 < Decimal("0.5") else 0) if split else 0 for k in range(32 * i, 32 * i + 32)]
        touched = sorted(set(lanes))
        for b in touched:
            h[b] += lanes.count(b)
            N[b] += 32
        if all(h[b] * c * c * (1 - Fraction(h[b], N[b])) < e2 * N[b] * (N[b] - 1) for b in touched):
            return 32 * i, [hb * c / (32 * i) for hb in h]
    return 32 * chunks, [hb * c / (32 * chunks) for hb in h]


@pytest.mark.parametrize("split", [False, True])
def test_gpu_estimator_row_stop(split):
    p = flat.flat_params(O.Params())
    p = flat.split_bins(p) if split else flat.single_bin(p)
    d = O.build_design(p)
    for seed in (0.3, 0.798713):
        rays, want = _gpu_row_closed_form(seed, 16384, split)
        row, got_rays = O.gpu_row(d, 0.6, seed, 16384)
        assert got_rays == rays, (seed, got_rays, rays)
        if split:
            np.testing.assert_allclose(row, [float(x) for x in want], rtol=1e-14)
        else:
            # one chunk of 32 hits: (32, c, 0) passes at once and p = c
            assert rays == 32 and row[16] == flat.T_AR and row.sum() == flat.T_AR
