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
