"""The 3D part of the CPU oracle, pinned against the reference's own test of
approx_cbrt and against the mathematics (numpy's quartic roots, brute-force
ray-torus intersection, the 2D arc the torus reduces to in its z = 0 plane,
Snell's law).  No GPU."""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import compound_prism_designer as cpd
from compound_prism_designer import _native as N
from oracle import oracle as O
from oracle import oracle3 as O3


def quartic_coeffs(r_2, R_2, q, dv, dn, nv):
    # toric.rs:175-186 (the coefficients the closed form solves)
    return [1.0, 4 * dv, -2 * (r_2 + R_2 - q - 2 * dv ** 2 - 2 * R_2 * nv ** 2),
            -4 * (r_2 + R_2 - q) * dv + 8 * R_2 * dn * nv,
            (r_2 - R_2) ** 2 - 2 * (R_2 + r_2) * q + q ** 2 + 4 * R_2 * dn ** 2]


def test_approx_cbrt_within_the_reference_bound():
    # cbrt.rs:97-111: relative error below 3e-5 against cbrt on [1e-5, 1024]
    v = np.geomspace(1e-5, 1024.0, 20001)
    got = np.array([O3.approx_cbrt(x) for x in v])
    assert np.max(np.abs(got / np.cbrt(v) - 1)) < 3e-5
    assert all(O3.approx_cbrt(-x) == -O3.approx_cbrt(x) for x in v[::997])
    assert O3.approx_cbrt(0.0) == 0.0


def test_closed_form_algebra_against_numpy_roots():
    rng = np.random.default_rng(0)
    errs, kinds = [], {}
    for _ in range(4000):
        R, r = rng.uniform(0.5, 5), rng.uniform(0.3, 4)
        c, n = rng.normal(size=3), rng.normal(size=3)
        n /= np.linalg.norm(n)
        o, d = rng.normal(size=3) * 3, rng.normal(size=3)
        d /= np.linalg.norm(d)
        dp = o - c
        args = (r * r, R * R, dp @ dp, dp @ d, dp @ n, n @ d)
        got = O3.torus_roots(*args, exact_cbrt=True)
        roots = np.roots(quartic_coeffs(*args))
        real = np.sort(roots[np.abs(roots.imag) < 1e-7].real)
        fin = np.sort(got[~np.isnan(got)])
        kinds[(len(real), len(fin))] = kinds.get((len(real), len(fin)), 0) + 1
        if len(real) == 2 and len(fin) == 2:
            errs.append(np.max(np.abs(fin - real) / np.maximum(1.0, np.abs(real))))
    # two real roots: the closed form finds them; four (the resolvent's casus
    # irreducibilis) or none: every candidate is NaN in real arithmetic
    assert set(kinds) <= {(0, 0), (2, 2), (4, 0)}, kinds
    assert kinds.get((2, 2), 0) > 1000 and kinds.get((4, 0), 0) > 100
    errs = np.array(errs)
    assert np.median(errs) < 1e-12 and np.percentile(errs, 99) < 1e-8


def implicit(p, c, n, R, r):
    # toric.rs:26-38
    u = p - c
    q = u @ u
    return (q + R * R - r * r) ** 2 - 4 * R * R * (q - (u @ n) ** 2)


def designs3(n, first=0):
    _, D, _ = cpd.generate_designs3(first, n)
    return D


def torus_of(d, which):
    t = d.lens[which]
    c = np.array([t.center[0], t.center[1], 0.0])
    n = np.array([t.toroidal_normal[0], t.toroidal_normal[1], 0.0])
    bc = np.array([t.bound_center[0], t.bound_center[1], 0.0])
    bn = np.array([t.bound_normal[0], t.bound_normal[1], 0.0])
    return c, n, t.toroidal_radius, t.poloidal_radius, bc, bn, t.bound_height


def brute_first_hit(o, v, c, n, R, r, bc, bn, bh, t_max):
    ts = np.linspace(0.0, t_max, 40001)
    u = o[None, :] + ts[:, None] * v[None, :] - c[None, :]
    q = np.einsum("ij,ij->i", u, u)
    f = (q + R * R - r * r) ** 2 - 4 * R * R * (q - (u @ n) ** 2)
    for i in np.flatnonzero(np.sign(f[:-1]) != np.sign(f[1:])):
        a, b = ts[i], ts[i + 1]
        for _ in range(60):
            m = 0.5 * (a + b)
            if np.sign(implicit(o + m * v, c, n, R, r)) == np.sign(implicit(o + a * v, c, n, R, r)):
                a = m
            else:
                b = m
        t = 0.5 * (a + b)
        h = bn @ (o + t * v - bc)
        if 0.0 <= h <= bh:
            return t
    return None


def test_lens_intersection_against_brute_force():
    rng = np.random.default_rng(1)
    checked = 0
    for d in designs3(6):
        for which in (0, 1):
            c, n, R, r, bc, bn, bh = torus_of(d, which)
            for _ in range(12):
                # rays aimed at the lens from a point before it, leaving the z = 0 plane
                target = bc + bn * rng.uniform(0, bh) + np.array([0, 0, 1.0]) * rng.uniform(-0.3, 0.3)
                o = target - np.array([rng.uniform(0.5, 2.0), rng.uniform(-0.3, 0.3), rng.uniform(-0.3, 0.3)])
                v = target - o
                v /= np.linalg.norm(v)
                got = O3.lens_intersection(d, which, o, v, exact_cbrt=True)
                want = brute_first_hit(o, v, c, n, R, r, bc, bn, bh, 4.0 + 4 * (R + r))
                if got is None:
                    # NaN candidates: the line crosses the torus four times
                    line = np.roots(quartic_coeffs(r * r, R * R, (o - c) @ (o - c), (o - c) @ v, (o - c) @ n, n @ v))
                    assert want is None or np.sum(np.abs(line.imag) < 1e-7) == 4
                    continue
                assert want is not None
                dist, normal = got
                assert abs(dist - want) <= 1e-8 * max(1.0, want)
                p = o + dist * v
                u = p - c
                grad = 4 * (u @ u + R * R - r * r) * u - 8 * R * R * (u - n * (u @ n))
                assert abs(np.linalg.norm(normal) - 1) < 1e-9
                assert np.linalg.norm(np.cross(normal, grad)) <= 1e-8 * np.linalg.norm(grad)
                checked += 1
    assert checked > 30


def arc_hit_2d(center, radius, midpt, max_dist_sq, o, v):
    # the 2D CurvedPlane, as the 2D oracle's design records give it
    delta = o - center
    ud = v @ delta
    disc = ud * ud - delta @ delta + radius * radius
    if disc <= 0:
        return None
    for t in (-ud - math.sqrt(disc), -ud + math.sqrt(disc)):
        p = o + t * v
        if t >= 0 and (p - midpt) @ (p - midpt) <= max_dist_sq:
            return t, (center - p) / radius
    return None


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_exit_torus_is_the_2d_arc_in_its_plane(sign):
    rng = np.random.default_rng(2)
    P3, D3, _ = cpd.generate_designs3(0, 5)
    checked = 0
    for p3 in P3:
        p3.final_curvature[0] = sign * abs(p3.final_curvature[0])
        d3 = O3.build_design(p3)
        # the same faces and exit curvature in the 2D oracle
        p2 = O.Params()
        for f in ("n_glass", "ar_coated", "height", "width", "bin_count", "bin_size", "linear_slope",
                  "det_length", "max_incident_angle", "det_angle"):
            setattr(p2, f, getattr(p3, f))
        ctypes.memmove(p2.glass, p3.glass, ctypes.sizeof(p2.glass))
        ctypes.memmove(p2.angles, p3.angles, ctypes.sizeof(p2.angles))
        ctypes.memmove(p2.lengths, p3.lengths, ctypes.sizeof(p2.lengths))
        p2.curvature = p3.final_curvature[0]
        d2 = O.build_design(p2)
        center = np.array(d2.arc_center[:])
        midpt = np.array(d2.arc_midpt[:])
        bc = np.array(d3.lens[1].bound_center[:])
        bn = np.array(d3.lens[1].bound_normal[:])
        for _ in range(20):
            target = bc + bn * rng.uniform(0, d3.lens[1].bound_height) + \
                np.array([-bn[1], bn[0]]) * rng.uniform(-0.4, 0.4) * p3.height
            o2 = target - np.array([rng.uniform(0.5, 1.5), rng.uniform(-0.2, 0.2)])
            v2 = (target - o2) / np.linalg.norm(target - o2)
            want = arc_hit_2d(center, d2.arc_radius, midpt, d2.arc_max_dist_sq, o2, v2)
            got = O3.lens_intersection(d3, 1, [*o2, 0.0], [*v2, 0.0], exact_cbrt=True)
            if want is None or got is None:
                continue
            assert abs(got[0] - want[0]) <= 1e-9 * max(1.0, want[0])
            np.testing.assert_allclose(got[1][:2], want[1], atol=1e-9)
            assert got[1][2] == 0.0
            checked += 1
    assert checked >= 20


def test_refraction_obeys_snell_in_3d():
    rng = np.random.default_rng(3)
    for _ in range(500):
        nrm = rng.normal(size=3)
        nrm /= np.linalg.norm(nrm)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if d @ nrm > 0:
            d = -d  # incidence against the normal: ci = -d.n >= 0
        n1, n2 = rng.uniform(1.0, 1.9, size=2)
        st, v, ts, tp = O3.refract(d, nrm, n1, n2)
        ci = -(d @ nrm)
        sin_i = math.sqrt(max(0.0, 1 - ci * ci))
        if n1 * sin_i > n2:
            assert st == N.ERR_TOTAL_INTERNAL_REFLECTION
            continue
        assert st == 0
        assert abs(np.linalg.norm(v) - 1) < 1e-12
        sin_r = np.linalg.norm(np.cross(v, nrm))
        assert abs(n1 * sin_i - n2 * sin_r) < 1e-12
        assert abs(v @ np.cross(d, nrm)) < 1e-12  # in the plane of incidence
        assert 0 < ts <= 1 and 0 < tp <= 1
    st, v, ts, tp = O3.refract([1, 0, 0], [-1, 0, 0], 1.0, 1.5)
    assert abs(ts - (1 - (0.5 / 2.5) ** 2)) < 1e-15 and abs(tp - ts) < 1e-15


def test_width_bound_loses_rays_outside():
    P3, _, _ = cpd.generate_designs3(0, 3)
    for p3 in P3:
        d3 = O3.build_design(p3)
        w = 0.5 * (p3.wavelength_lo + p3.wavelength_hi)
        st0, b0, _ = O3.propagate(d3, w, [0.0, p3.y_mean, 0.0], [1.0, 0.0, 0.0])
        st1, _ = O3.exit_ray(d3, p3.y_mean, w)
        assert st1 == 0  # the median ray leaves the prism
        z = 1.5 * d3.half_width
        # a ray running parallel to the axis beyond the half width cannot meet an interior face
        st2, _, _ = O3.propagate(d3, w, [0.0, p3.y_mean, z], [1.0, 0.0, 0.0])
        assert st2 != 0


def test_beam_samples():
    _, D3, _ = cpd.generate_designs3(0, 1)
    d = D3[0]
    q = O3.qmc(4, 4096)
    rays = [O3.sample_ray(d, qq) for qq in q]
    org = np.array([r[0] for r in rays])
    dirs = np.array([r[1] for r in rays])
    rad2 = (org[:, 1] - d.y_mean) ** 2 + org[:, 2] ** 2
    assert np.all(org[:, 0] == 0) and np.all(rad2 <= d.beam_width ** 2 * (1 + 1e-12))
    # uniform disc: half the rays within radius / sqrt(2)
    assert abs(np.mean(rad2 <= d.beam_width ** 2 / 2) - 0.5) < 0.01
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-12)
    # uniform over the cone: cos >= min_cos with mean (1 + min_cos) / 2
    assert np.all(dirs[:, 0] >= d.fiber_min_cos - 1e-15)
    assert abs(np.mean(dirs[:, 0]) - (1 + d.fiber_min_cos) / 2) < 1e-3 * (1 - d.fiber_min_cos)
    # the 3D Gaussian beam: unit x direction, radial spread of the 1/e^2 width
    dg = O3.Design3()
    ctypes.memmove(ctypes.byref(dg), ctypes.byref(O._as(d, O3.Design3)), ctypes.sizeof(dg))
    dg.beam_kind = O3.BEAM_GAUSSIAN
    dg.beam_width = 0.1
    q2 = O3.qmc(2, 8192)
    g = [O3.sample_ray(dg, [*qq, 0, 0]) for qq in q2]
    og = np.array([x[0] for x in g])
    assert all(np.array_equal(x[1], [1.0, 0.0, 0.0]) for x in g)
    r2 = (og[:, 1] - dg.y_mean) ** 2 + og[:, 2] ** 2
    assert abs(np.mean(r2) / dg.beam_width ** 2 - 1) < 0.02


@pytest.mark.parametrize("dims", [2, 4])
def test_qmc_points(dims):
    fx = O3.qmc(dims, 20000, O.QRNG_FIXED_POINT)
    it = O3.qmc(dims, 20000, O.QRNG_ITERATIVE)
    assert np.max(np.abs(fx - it)) < 1e-10
    # exact: u_k,i = floor(frac(1/2 + k alpha_i) 2^53) / 2^53 with the f64 alphas
    phi = 1.167303978261418684256045899854842180720560371525489039140082449275651903429527 if dims == 4 \
        else 1.3247179572447460259609088544780973
    alphas, p = [], phi
    for i in range(dims):
        alphas.append(Fraction(1.0 / p))
        p = p * phi
    for k in (1, 2, 3, 17, 4096, 19999):
        for i in range(dims):
            u = (Fraction(1, 2) + k * alphas[i]) % 1
            assert fx[k - 1, i] == math.floor(u * 2 ** 53) / 2 ** 53


def test_fitness_bounds_and_modes():
    D = designs3(3)
    for d in D:
        (size, info, dev), (rows, rays) = O3.fitness(d, 8, 512, adaptive=False)
        assert rows == 8 and rays == 8 * 512
        assert 0.0 <= info <= math.log2(d.bin_count) + 1e-12
        assert size > 0 and 0 <= dev <= 1
        row, m = O3.p_dets_l_wavelength(d, 0.6, 512, adaptive=False)
        assert m == 512 and np.all(row >= 0) and row.sum() <= 1 + 1e-12


def test_reference_entrance_sign_sends_rays_back():
    # designer.py:209 passes a negative entrance curvature; the torus normal
    # then points along the ray and refraction turns it back, so no design
    # with it can position its detector
    P3, _, _ = cpd.generate_designs3(0, 4)
    for p3 in P3:
        p3.initial_curvature[0] = -abs(p3.initial_curvature[0])
        st, _, _ = O3.position_detector_array(p3, 1.0)
        assert st != 0
