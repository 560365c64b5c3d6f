// trace3.h -- the 3D spectrometer in FP64, for the host (construction,
// positioning, beam tables) and the device (row kernels).
//
// The geometry is the reference's Dim3 variant (src/pylib.rs:636-640): a
// ToricLens entrance face, planar interior faces bounded in height and width,
// a ToricLens exit face (src/spectrometer/ray.rs:121-127), then the linear
// detector array.  Every surface lies in the xy plane (z = 0); rays from the
// fiber beam leave it.  The arithmetic follows the reference's operation order
// with fused multiply-adds exactly at its mul_add sites, dot products as
// ((x x' + y y') + z z') (Vector<3> for SimdVector<T, 4>,
// src/spectrometer/vector.rs:518-560), so a ray traced here and by the FP64
// CPU path lands in the same bin with the same transmittance.
//
// Beam samples: the trigonometric and square-root part of a ray's inverse
// CDF depends only on its quasi-random point, so it is evaluated once per ray
// index on the host (BeamSample3) and read by every row; the device applies
// the beam's scale and offset in the reference's order.
#pragma once

#include <math.h>
#include <stdint.h>
#include <string.h>

#include "cpd.h"
#include "trace_core.h"

namespace cpd {

struct P3 {
    double x, y, z;
};

CPD_HD P3 p3_sub(P3 a, P3 b) { return P3{a.x - b.x, a.y - b.y, a.z - b.z}; }
CPD_HD double p3_dot(P3 a, P3 b) {
#pragma clang fp contract(off)
    return (a.x * b.x + a.y * b.y) + a.z * b.z;
}
// Vector::mul_add: a b + c, fused per component
CPD_HD P3 p3_fma(P3 a, double b, P3 c) { return P3{fma(a.x, b, c.x), fma(a.y, b, c.y), fma(a.z, b, c.z)}; }

// approx_cbrt (src/spectrometer/cbrt.rs:49-68, Moroz et al., Algorithm 6):
// a first guess from the bits of x rounded to FP32, then two polynomial
// corrections at full precision.
CPD_HD double moroz_cbrt(double x) {
#pragma clang fp contract(off)
    const double k1 = (double)1.752319676f, k2 = (double)1.2509524245f, k3 = (double)0.5093818292f;
    const double third = (double)0.333333333333f;
    const bool pos = !signbit(x);
    const double ax = fabs(x);
    const float xf = (float)ax;
    uint32_t bits;
    memcpy(&bits, &xf, 4);
    bits = 0x548c2b4bu - bits / 3u;
    float yf;
    memcpy(&yf, &bits, 4);
    double y = (double)yf;
    double c = ax * y * y * y;
    y = y * (k1 - c * (k2 - k3 * c));
    const double d = ax * y * y;
    c = 1.0 - d * y;
    const double r = d * (1.0 + third * c);
    return pos ? r : -r;
}

// The four real-line candidates t0 <= t1 <= t2 <= t3 of the ray-torus
// quartic t^4 + 4 (dp.v) t^3 + c t^2 + d t + e in the closed form of
// src/spectrometer/toric.rs:41-141 (Ferrari's method through the resolvent
// cubic, as sympy's cse prints it), with dp = o - centre, v the direction,
// n the torus axis, R and r the toroidal and poloidal radii.
CPD_HD void torus_quartic(double r_2, double R_2, double q, double dv, double dn, double nv, double t[4]) {
#pragma clang fp contract(off)
    const double a0 = R_2 * dv, a1 = q * dv, a2 = dv * r_2;
    const double a3 = dn * nv, a4 = R_2 * a3, a5 = nv * nv;
    const double a6 = 2.0 * R_2;
    const double a7 = R_2 - q + r_2 - a5 * a6;
    const double a8 = 8.0 * dv * a7 - 8.0 * a0 + 8.0 * a1 - 8.0 * a2 + 16.0 * a4;
    const double a9 = 2.0 * r_2;
    const double a10 = dv * dv, a11 = R_2 * a5;
    const double a12 = 2.0 * q - 2.0 * a10 + 4.0 * a11 - a6 - a9;
    const double a13 = a12 * a12 * a12;
    const double a14 = 4.0 * dv;
    const double s15 = -4.0 * a0 + 4.0 * a1 + a14 * a7 - 4.0 * a2 + 8.0 * a4;
    const double a15 = s15 * s15;
    const double a16 = R_2 * R_2, a17 = q * q, a18 = r_2 * r_2;
    const double a19 = q * a6, a20 = r_2 * a6, a21 = q * a9;
    const double a22 = 4.0 * R_2 * (dn * dn);
    const double a23 =
        a14 * (-a0 + a1 + a14 * (R_2 / 8.0 - q / 8.0 + r_2 / 8.0 - a10 / 16.0 - a11 / 4.0) - a2 + a3 * a6);
    const double a24 = a16 + a17 + a18 - a19 - a20 - a21 + a22 - a23;
    const double a25 = a12 * a24;
    const double a26 = -a13 / 108.0 - a15 / 8.0 + a25 / 3.0;
    const double a27 = 2.0 * moroz_cbrt(a26);
    const double a28 = 8.0 * R_2 / 3.0;
    const double a29 = 4.0 * R_2 / 3.0 - 4.0 * q / 3.0 + 4.0 * r_2 / 3.0 + 4.0 * a10 / 3.0 - a28 * a5;
    const double a30 = sqrt(-a27 + a29);
    const double a31 = a8 / a30;
    const double a32 = -8.0 * q / 3.0 + 8.0 * r_2 / 3.0 + 8.0 * a10 / 3.0 - 16.0 * a11 / 3.0 + a28;
    const double a33 = a27 + a32;
    const double a34 = sqrt(a31 + a33) / 2.0;
    const double a35 = -dv;
    const double a36 = a30 / 2.0;
    const double a37 = a35 - a36;
    const double a38 = (a12 * a12) / 12.0;
    const bool depressed = a24 + a38 == 0.0;
    const double a40 = -a16 - a17 - a18 + a19 + a20 + a21 - a22 + a23 - a38;
    const double a41 = moroz_cbrt(a13 / 216.0 + a15 / 16.0 - a25 / 6.0 + sqrt((a26 * a26) / 4.0 + (a40 * a40 * a40) / 27.0));
    const double a42 = 2.0 * a41;
    const double a43 = 2.0 * a40 / (3.0 * a41);
    const double a44 = sqrt(a29 + a42 - a43);
    const double a45 = a8 / a44;
    const double a46 = a32 - a42 + a43;
    const double a47 = sqrt(a45 + a46) / 2.0;
    const double a48 = a44 / 2.0;
    const double a49 = a35 - a48;
    const double a50 = sqrt(-a31 + a33) / 2.0;
    const double a51 = a35 + a36;
    const double a52 = sqrt(-a45 + a46) / 2.0;
    const double a53 = a35 + a48;
    t[3] = depressed ? a34 + a37 : a47 + a49;
    t[2] = depressed ? a50 + a51 : a52 + a53;
    t[1] = depressed ? -a34 + a37 : -a47 + a49;
    t[0] = depressed ? -a50 + a51 : -a52 + a53;
}

// ToricSurface::intersection (toric.rs:143-236) with SandwichBounds
// (geometry.rs:104-109): the first candidate that is not negative and lies
// between the chord plane and the plane at the sagitta.  The normal runs from
// the hit point to the nearest point of the tube's centre circle, over r.
CPD_HD bool torus_hit(const cpd_torus &T, P3 o, P3 v, double &dist, P3 &normal) {
#pragma clang fp contract(off)
    const P3 c{T.center[0], T.center[1], 0.0};
    const P3 n{T.toroidal_normal[0], T.toroidal_normal[1], 0.0};
    const P3 dp = p3_sub(o, c);
    double t[4];
    torus_quartic(T.poloidal_radius * T.poloidal_radius, T.toroidal_radius * T.toroidal_radius, p3_dot(dp, dp),
                  p3_dot(dp, v), p3_dot(dp, n), p3_dot(n, v), t);
    const P3 bc{T.bound_center[0], T.bound_center[1], 0.0};
    const P3 bn{T.bound_normal[0], T.bound_normal[1], 0.0};
    bool found = false;
    double tk = 0.0;
    P3 p{0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (found || t[k] < 0.0) continue;
        const P3 q = p3_fma(v, t[k], o);
        const double h = p3_dot(bn, p3_sub(q, bc));
        if (0.0 <= h && h <= T.bound_height) {
            found = true;
            tk = t[k];
            p = q;
        }
    }
    if (!found) return false;
    // oproj(p - c, n) = u - n (u.n), normalised, scaled to R about c
    const P3 u = p3_sub(p, c);
    const double un = p3_dot(u, n);
    const P3 w{u.x - n.x * un, u.y - n.y * un, u.z - n.z * un};
    const double wl = sqrt(p3_dot(w, w));
    const P3 tc = p3_fma(P3{w.x / wl, w.y / wl, w.z / wl}, T.toroidal_radius, c);
    normal = P3{(tc.x - p.x) / T.poloidal_radius, (tc.y - p.y) / T.poloidal_radius, (tc.z - p.z) / T.poloidal_radius};
    dist = tk;
    return true;
}

struct Ray3 {
    P3 o, v;
    double ts, tp;
};

// Ray::refract (src/spectrometer/ray.rs:309-367) in 3D.
CPD_HD int refract3(Ray3 &ray, P3 nrm, double n1, double n2, bool ar) {
#pragma clang fp contract(off)
    const double ci = -p3_dot(ray.v, nrm);
    const double r = n1 / n2;
    const double cr_sq = 1.0 - r * r * (1.0 - ci * ci);
    if (cr_sq < 0.0) return CPD_ERR_TOTAL_INTERNAL_REFLECTION;
    const double cr = sqrt(cr_sq);
    const double k = r * ci - cr;
    ray.v = P3{ray.v.x * r + nrm.x * k, ray.v.y * r + nrm.y * k, ray.v.z * r + nrm.z * k};
    if (ar && ci > 0.5) {
        ray.ts = ray.ts * 0.99;
        ray.tp = ray.tp * 0.99;
    } else {
        const double rs = (n1 * ci - n2 * cr) / (n1 * ci + n2 * cr);
        const double rp = (n1 * cr - n2 * ci) / (n1 * cr + n2 * ci);
        ray.ts = ray.ts * (1.0 - rs * rs);
        ray.tp = ray.tp * (1.0 - rp * rp);
    }
    return CPD_OK;
}

CPD_HD int lens3(Ray3 &ray, const cpd_torus &T, double n1, double n2, bool ar) {
    double d;
    P3 nrm;
    if (!torus_hit(T, ray.o, ray.v, d, nrm)) return CPD_ERR_NO_SURFACE_INTERSECTION;
    ray.o = p3_fma(ray.v, d, ray.o);
    return refract3(ray, nrm, n1, n2, ar);
}

// HyperPlane::intersection (geometry.rs:124-145) with PrismBounds<3>
// (geometry.rs:69-77): 0 <= y <= height, |z| <= half width.
CPD_HD int plane3(Ray3 &ray, const double *pt, const double *nm, double height, double half_width, double n1,
                  double n2, bool ar) {
#pragma clang fp contract(off)
    const P3 n{nm[0], nm[1], 0.0};
    const double ci = p3_dot(ray.v, n);
    if (ci == 0.0) return CPD_ERR_NO_SURFACE_INTERSECTION;
    const double d = p3_dot(p3_sub(P3{pt[0], pt[1], 0.0}, ray.o), n) / ci;
    if (!(d >= 0.0)) return CPD_ERR_NO_SURFACE_INTERSECTION;
    const P3 p = p3_fma(ray.v, d, ray.o);
    if (!(0.0 <= p.y && p.y <= height && -half_width <= p.z && p.z <= half_width)) return CPD_ERR_NO_SURFACE_INTERSECTION;
    ray.o = p;
    return refract3(ray, n, n1, n2, ar);
}

// The interior faces and the exit lens, for a ray already refracted into
// glass 0 (the rest of CompoundPrism::propagate, ray.rs:485-499), given the
// glasses' indices n[i] at the wavelength (glass_n64 of each glass, as calc_n
// computes them).  The faces are unrolled so that n stays in registers.
CPD_HD int prism3_inner(Ray3 &ray, const cpd_design3 &g, const double *n) {
    const bool ar = g.ar_coated != 0;
    double n1 = n[0];
#pragma unroll
    for (int i = 1; i < CPD_MAX_PRISMS; ++i) {
        if (i >= (int)g.n_glass) break;
        const int e = plane3(ray, g.plane_point[i], g.plane_normal[i], g.height, g.half_width, n1, n[i], ar);
        if (e != CPD_OK) return e;
        n1 = n[i];
    }
    return lens3(ray, g.lens[1], n1, 1.0, ar);
}

// CompoundPrism::propagate (ray.rs:485-499).
CPD_HD int prism3(Ray3 &ray, const cpd_design3 &g, const double *n) {
    const int e = lens3(ray, g.lens[0], 1.0, n[0], g.ar_coated != 0);
    if (e != CPD_OK) return e;
    return prism3_inner(ray, g, n);
}

// Where a beam ray meets the entrance lens, and the lens normal there: both
// independent of the wavelength, so computed once per ray and reused by every
// wavelength row.  nx is NaN when the ray misses the lens.
struct EntryHit3 {
    double px, py, pz, nx, ny, nz;
};

CPD_HD EntryHit3 entrance_hit3(const cpd_design3 &g, const Ray3 &ray) {
    double d;
    P3 nrm;
    if (!torus_hit(g.lens[0], ray.o, ray.v, d, nrm)) return EntryHit3{0.0, 0.0, 0.0, NAN, NAN, NAN};
    const P3 p = p3_fma(ray.v, d, ray.o);  // the translation lens3 applies
    return EntryHit3{p.x, p.y, p.z, nrm.x, nrm.y, nrm.z};
}

// LinearDetectorArray intersection and bin (spectrometer.rs:265-303), the
// array in z = 0.
CPD_HD int detector3(const Ray3 &ray, const cpd_design3 &g, uint32_t &bin, double &t) {
#pragma clang fp contract(off)
    const P3 dn{g.det_normal[0], g.det_normal[1], 0.0};
    const P3 dpos{g.det_position[0], g.det_position[1], 0.0};
    const double ci = -p3_dot(ray.v, dn);
    if (ci <= g.det_min_ci) return CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK;
    const double d = p3_dot(p3_sub(ray.o, dpos), dn) / ci;
    const P3 p = p3_fma(ray.v, d, ray.o);
    const double pos = p3_dot(p3_sub(p, dpos), P3{g.det_direction[0], g.det_direction[1], 0.0});
    if (pos < 0.0 || g.det_length < pos) return CPD_ERR_NO_SURFACE_INTERSECTION;
    const double v = pos - g.det_intercept;
    double q = trunc(v / g.det_slope);
    double r = v - q * g.det_slope;
    if (r < 0.0) {
        q = g.det_slope > 0.0 ? q - 1.0 : q + 1.0;
        r = r + fabs(g.det_slope);
    }
    const uint32_t b = !(q > 0.0) ? 0u : (q >= 4294967295.0 ? 4294967295u : (uint32_t)q);
    if (!(b < g.bin_count && r < g.det_bin_size)) return CPD_ERR_NO_SURFACE_INTERSECTION;
    bin = b;
    t = (ray.ts + ray.tp) * 0.5;
    return CPD_OK;
}

// Refractive indices of the design's glasses at a wavelength (1 past the
// last glass).
CPD_HD void glass_indices3(const cpd_design3 &g, double lambda, double *n) {
#pragma unroll
    for (int i = 0; i < CPD_MAX_PRISMS; ++i) n[i] = i < (int)g.n_glass ? glass_n64(g.glass[i], lambda) : 1.0;
}

// UniformDistribution::inverse_cdf (distribution.rs:32-34)
CPD_HD double lambda_of3(const cpd_design3 &g, double u) {
#pragma clang fp contract(off)
    return g.wavelength_lo + (g.wavelength_hi - g.wavelength_lo) * u;
}

// ---------------------------------------------------------------------------
// Beam samples
// ---------------------------------------------------------------------------
// The per-ray-index part of a beam ray:
//   fiber    (spectrometer.rs:113-131, distribution.rs:58-67):
//            a = sqrt(p_rho), b = cos(2 pi p_theta), c = sin(2 pi p_theta),
//            d = p_defl, e = cos(2 pi p_rot), f = sin(2 pi p_rot)
//   Gaussian (spectrometer.rs:32-46):
//            a = norminv(p_r), b = sin(2 pi p_theta), c = cos(2 pi p_theta)
struct BeamSample3 {
    double a, b, c, d, e, f;
};

static constexpr double kTau = 6.283185307179586;

CPD_HD BeamSample3 beam_sample3(uint32_t kind, const double *q) {
#pragma clang fp contract(off)
    BeamSample3 s{0, 0, 0, 0, 0, 0};
    if (kind == CPD_BEAM_FIBER) {
        const double th = q[1] * kTau, rot = q[3] * kTau;
        s.a = sqrt(q[0]);
        s.b = cos(th);
        s.c = sin(th);
        s.d = q[2];
        s.e = cos(rot);
        s.f = sin(rot);
    } else {
        const double th = q[1] * kTau;
        s.a = norminv64(q[0]);
        s.b = sin(th);
        s.c = cos(th);
    }
    return s;
}

// The beam ray of a sample: origin (0, y, z) and direction.
CPD_HD Ray3 beam_ray3(const cpd_design3 &g, const BeamSample3 &s) {
#pragma clang fp contract(off)
    if (g.beam_kind == CPD_BEAM_FIBER) {
        const double r = g.beam_width * s.a;
        const double y = r * s.b, z = r * s.c;
        const double ux = g.fiber_min_cos + s.d * (1.0 - g.fiber_min_cos);
        const double sx = sqrt(1.0 - ux * ux);
        return Ray3{P3{0.0, y + g.y_mean, z}, P3{ux, sx * s.e, sx * s.f}, 1.0, 1.0};
    }
    const double r = g.beam_width * s.a;
    return Ray3{P3{0.0, g.y_mean + r * s.b, r * s.c}, P3{1.0, 0.0, 0.0}, 1.0, 1.0};
}

// One ray: its bin (bin_count on a miss) and transmittance (0 on a miss).
CPD_HD uint32_t trace3(const cpd_design3 &g, const double *n, Ray3 ray, double &t) {
    t = 0.0;
    if (prism3(ray, g, n) != CPD_OK) return g.bin_count;
    uint32_t bin = g.bin_count;
    double tt = 0.0;
    if (detector3(ray, g, bin, tt) != CPD_OK) return g.bin_count;
    t = tt;
    return bin;
}

// The same from the ray's entrance hit (entrance_hit3 of the same ray).
CPD_HD uint32_t trace3_entered(const cpd_design3 &g, const double *n, Ray3 ray, const EntryHit3 &h, double &t) {
    t = 0.0;
    if (isnan(h.nx)) return g.bin_count;
    ray.o = P3{h.px, h.py, h.pz};
    if (refract3(ray, P3{h.nx, h.ny, h.nz}, 1.0, n[0], g.ar_coated != 0) != CPD_OK) return g.bin_count;
    if (prism3_inner(ray, g, n) != CPD_OK) return g.bin_count;
    uint32_t bin = g.bin_count;
    double tt = 0.0;
    if (detector3(ray, g, bin, tt) != CPD_OK) return g.bin_count;
    t = tt;
    return bin;
}

// R1 points in dims 2 (Gaussian) or 4 (fiber): frac(seed + k alpha_i) in
// 64-bit fixed point, alpha_i the reference's f64 constants 1 / phi_d^i
// (qrng.rs:160-177) scaled by 2^64 (exact).
static constexpr double kPhi2 = 1.3247179572447460259609088544780973;
static constexpr double kPhi4 = 1.167303978261418684256045899854842180720560371525489039140082449275651903429527;
CPD_HD void qmc_alpha_fx(uint32_t kind, uint64_t afx[4]) {
#pragma clang fp contract(off)
    double a[4];
    if (kind == CPD_BEAM_FIBER) {
        a[0] = 1.0 / kPhi4;
        a[1] = 1.0 / (kPhi4 * kPhi4);
        a[2] = 1.0 / (kPhi4 * kPhi4 * kPhi4);
        a[3] = 1.0 / (kPhi4 * kPhi4 * kPhi4 * kPhi4);
    } else {
        a[0] = 1.0 / kPhi2;
        a[1] = 1.0 / (kPhi2 * kPhi2);
        a[2] = a[3] = 0.0;
    }
    for (int i = 0; i < 4; ++i) afx[i] = (uint64_t)(a[i] * 0x1.0p+64);
}
CPD_HD void qmc_point3(const uint64_t afx[4], uint64_t seed_fx, uint64_t k, double q[4]) {
    for (int i = 0; i < 4; ++i) q[i] = (double)((seed_fx + k * afx[i]) >> 11) * 0x1.0p-53;
}

// Spectrometer::size_and_deviation (spectrometer.rs:429-437)
CPD_HD void size_and_deviation3(const cpd_design3 &g, double &size, double &deviation) {
#pragma clang fp contract(off)
    const double half = g.det_length * 0.5;
    const double mx = fma(g.det_direction[0], half, g.det_position[0]);
    const double my = fma(g.det_direction[1], half, g.det_position[1]);
    const double dy = my - g.y_mean;
    size = sqrt(mx * mx + dy * dy);
    deviation = fabs(dy / size);
}

}  // namespace cpd
