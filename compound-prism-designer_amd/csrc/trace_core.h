// trace_core.h -- per-ray trace of the 2D compound-prism spectrometer.
//
// Two traces live here:
//  * trace32<P>: the FP32 hot path.  Per-wavelength invariants (refractive
//    ratios, the entrance-face refraction of the collimated beam) arrive
//    precomputed in FP64 and rounded once; the exit lens and the detector are
//    solved relative to the lens' chord midpoint, which avoids the
//    cancellation of |o - c|^2 - R^2 at lens radii far larger than the prism.
//    Every comparison whose outcome decides hit/miss or the bin carries an
//    error margin; a ray that lands inside a margin is flagged for the FP64
//    trace instead of being counted.
//  * trace64: the FP64 trace in the reference's operation order
//    (src/spectrometer/ray.rs:309-382, 485-499; geometry.rs:124-145,
//    161-216; spectrometer.rs:265-303), with fused multiply-adds exactly at the
//    reference's mul_add sites.  Flagged rays are re-traced with it, so their
//    bin decision is the FP64 one.
//
// The code is compiled for the device (HIP) and, for detector positioning and
// construction, for the host.
#pragma once

#include <math.h>
#include <stdint.h>

#include "cpd.h"

#if defined(__HIP__)
#define CPD_HD __host__ __device__ inline
#else
#define CPD_HD inline
#endif

namespace cpd {

// ---------------------------------------------------------------------------
// R1 low-discrepancy sequence u_k = frac(0.5 + k / phi) in 64-bit fixed point
// (the reference iterates state = fract(state + 1/phi) from 0.5,
// src/spectrometer/qrng.rs:203-255; the fixed-point form gives every k the
// same u on host and device).
// ---------------------------------------------------------------------------
static constexpr uint64_t kAlphaFx = 0x9E3779B97F4A7C16ull;  // round(2^64 / phi)
static constexpr uint64_t kHalfFx = 0x8000000000000000ull;

CPD_HD double qmc_u64(uint64_t k) { return (double)((kHalfFx + k * kAlphaFx) >> 11) * 0x1.0p-53; }

// Standard normal quantile, Shaw-Luu-Brickman rational in v = -ln(2 min(u, 1-u))
// (arXiv:0901.0638), as src/spectrometer/erf.rs:27-73 uses it.
CPD_HD double norminv64(double x) {
#pragma clang fp contract(off)
    const double P[8] = {4.3304513840364031401e-8, 0.000036313870818023761224, 0.0034424140686962222423,
                         0.085838533424158257377,  0.73176759583280610539,     2.3884158540184385711,
                         3.0333178251950406994,    1.2533141359896652729};
    const double Q[8] = {1.3141263119543315917e-6, 0.00030617264753008793976, 0.014494272424798068406,
                         0.2168237095066675527,    1.2356513216582148689,     2.9373357991677046357,
                         2.9202373175993672857,    1.0};
    double u = x > 0.5 ? 1.0 - x : x;
    double v = -log(2.0 * u);
    double p = P[0], q = Q[0];
    for (int i = 1; i < 8; ++i) {
        p = fma(p, v, P[i]);
        q = fma(q, v, Q[i]);
    }
    return copysign(v * p / q, x - 0.5);
}

// ---------------------------------------------------------------------------
// FP64 trace in the reference's operation order.
// ---------------------------------------------------------------------------
enum TraceErr : int {
    T_OK = CPD_OK,
    T_NO_SURFACE = CPD_ERR_NO_SURFACE_INTERSECTION,
    T_TIR = CPD_ERR_TOTAL_INTERNAL_REFLECTION,
    T_WEAK = CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK,
};

struct Ray64 {
    double ox, oy, dx, dy, ts, tp;
};

CPD_HD double glass_n64(const double *c, double w) {
    double s = c[0];
    for (int i = 1; i < CPD_GLASS_COEFFS; ++i) s = fma(s, w, c[i]);
    return s;
}

// Snell refraction + Fresnel transmittance at a surface with unit normal n.
CPD_HD int refract64(Ray64 &ray, double nx, double ny, double n1, double n2, bool ar) {
#pragma clang fp contract(off)
    double ci = -(ray.dx * nx + ray.dy * ny);
    double r = n1 / n2;
    double cr_sq = 1.0 - r * r * (1.0 - ci * ci);
    if (cr_sq < 0.0) return T_TIR;
    double cr = sqrt(cr_sq);
    double k = r * ci - cr;
    double vx = ray.dx * r + nx * k;
    double vy = ray.dy * r + ny * k;
    if (ar && ci > 0.5) {
        ray.ts = ray.ts * 0.99;
        ray.tp = ray.tp * 0.99;
    } else {
        double rs = (n1 * ci - n2 * cr) / (n1 * ci + n2 * cr);
        double rp = (n1 * cr - n2 * ci) / (n1 * cr + n2 * ci);
        ray.ts = ray.ts * (1.0 - rs * rs);
        ray.tp = ray.tp * (1.0 - rp * rp);
    }
    ray.dx = vx;
    ray.dy = vy;
    return T_OK;
}

// Bounded plane: intersect, translate, refract.
CPD_HD int plane64(Ray64 &ray, const double *p, const double *n, double height, double n1, double n2, bool ar) {
#pragma clang fp contract(off)
    double ci = ray.dx * n[0] + ray.dy * n[1];
    if (ci == 0.0) return T_NO_SURFACE;
    double d = ((p[0] - ray.ox) * n[0] + (p[1] - ray.oy) * n[1]) / ci;
    if (!(d >= 0.0)) return T_NO_SURFACE;
    double hx = fma(ray.dx, d, ray.ox);
    double hy = fma(ray.dy, d, ray.oy);
    if (!(0.0 <= hy && hy <= height)) return T_NO_SURFACE;
    ray.ox = hx;
    ray.oy = hy;
    return refract64(ray, n[0], n[1], n1, n2, ar);
}

// Circular-arc exit lens: nearest valid root first.
CPD_HD int arc64(Ray64 &ray, const cpd_design &g, double n1, bool ar) {
#pragma clang fp contract(off)
    double ex = ray.ox - g.arc_center[0], ey = ray.oy - g.arc_center[1];
    double ud = ray.dx * ex + ray.dy * ey;
    double disc = ud * ud - (ex * ex + ey * ey) + g.arc_radius * g.arc_radius;
    if (!(disc > 0.0)) return T_NO_SURFACE;
    double sd = sqrt(disc);
    double d1 = -ud - sd, d2 = -ud + sd;
    double p1x = fma(ray.dx, d1, ray.ox), p1y = fma(ray.dy, d1, ray.oy);
    double p2x = fma(ray.dx, d2, ray.ox), p2y = fma(ray.dy, d2, ray.oy);
    double a1x = p1x - g.arc_midpt[0], a1y = p1y - g.arc_midpt[1];
    double a2x = p2x - g.arc_midpt[0], a2y = p2y - g.arc_midpt[1];
    double px, py;
    if (!signbit(d1) && a1x * a1x + a1y * a1y <= g.arc_max_dist_sq) {
        px = p1x;
        py = p1y;
    } else if (!signbit(d2) && a2x * a2x + a2y * a2y <= g.arc_max_dist_sq) {
        px = p2x;
        py = p2y;
    } else {
        return T_NO_SURFACE;
    }
    double nx = (g.arc_center[0] - px) / g.arc_radius;
    double ny = (g.arc_center[1] - py) / g.arc_radius;
    ray.ox = px;
    ray.oy = py;
    return refract64(ray, nx, ny, n1, 1.0, ar);
}

// Through the compound prism only (entrance face, interior faces, lens).
CPD_HD int prism64(Ray64 &ray, const cpd_design &g, double wavelength) {
    const bool ar = g.ar_coated != 0;
    double n1 = glass_n64(g.glass[0], wavelength);
    int e = plane64(ray, g.plane_point[0], g.plane_normal[0], g.height, 1.0, n1, ar);
    if (e != T_OK) return e;
    for (uint32_t i = 1; i < g.n_glass; ++i) {
        double n2 = glass_n64(g.glass[i], wavelength);
        e = plane64(ray, g.plane_point[i], g.plane_normal[i], g.height, n1, n2, ar);
        if (e != T_OK) return e;
        n1 = n2;
    }
    return arc64(ray, g, n1, ar);
}

// Detector array: bin index and mean transmittance.
CPD_HD int detector64(const Ray64 &ray, const cpd_design &g, uint32_t &bin, double &t) {
#pragma clang fp contract(off)
    double ci = -(ray.dx * g.det_normal[0] + ray.dy * g.det_normal[1]);
    if (ci <= g.det_min_ci) return T_WEAK;
    double d = ((ray.ox - g.det_position[0]) * g.det_normal[0] + (ray.oy - g.det_position[1]) * g.det_normal[1]) / ci;
    double px = fma(ray.dx, d, ray.ox), py = fma(ray.dy, d, ray.oy);
    double pos = (px - g.det_position[0]) * g.det_direction[0] + (py - g.det_position[1]) * g.det_direction[1];
    if (pos < 0.0 || g.det_length < pos) return T_NO_SURFACE;
    // euclid_div_rem (src/spectrometer/utils.rs:177-192)
    double v = pos - g.det_intercept;
    double q = trunc(v / g.det_slope);
    double r = v - q * g.det_slope;
    if (r < 0.0) {
        q = g.det_slope > 0.0 ? q - 1.0 : q + 1.0;
        r = r + fabs(g.det_slope);
    }
    // saturating `as u32`
    uint32_t b = !(q > 0.0) ? 0u : (q >= 4294967295.0 ? 4294967295u : (uint32_t)q);
    if (!(b < g.bin_count && r < g.det_bin_size)) return T_NO_SURFACE;
    bin = b;
    t = (ray.ts + ray.tp) * 0.5;
    return T_OK;
}

// Spectrometer::size_and_deviation (src/spectrometer/spectrometer.rs:429-437):
// distance from the beam's mean start to the detector's midpoint, and
// |sin| of the deviation angle.
CPD_HD void size_and_deviation64(const cpd_design &g, double &size, double &deviation) {
#pragma clang fp contract(off)
    double half = g.det_length * 0.5;
    double mx = fma(g.det_direction[0], half, g.det_position[0]);
    double my = fma(g.det_direction[1], half, g.det_position[1]);
    double dy = my - g.y_mean;
    size = sqrt(mx * mx + dy * dy);
    deviation = fabs(dy / size);
}

// One ray from (0, y) with direction (1, 0): returns the bin (bin_count on a
// miss) and its transmittance.
CPD_HD uint32_t trace64(const cpd_design &g, double wavelength, double y, double &t) {
    Ray64 ray{0.0, y, 1.0, 0.0, 1.0, 1.0};
    t = 0.0;
    if (prism64(ray, g, wavelength) != T_OK) return g.bin_count;
    uint32_t bin = g.bin_count;
    double tt = 0.0;
    if (detector64(ray, g, bin, tt) != T_OK) return g.bin_count;
    t = tt;
    return bin;
}

// ---------------------------------------------------------------------------
// FP32 trace record and per-wavelength constants
// ---------------------------------------------------------------------------
struct alignas(16) TraceDesign {
    uint32_t n_glass, nbin, ar_coated, n_lambda;
    uint32_t n_rays, reserved0, reserved1, reserved2;
    // entrance face: hit at x = (c0 - y n0y) / n0x for beam height y in [0, h]
    float c0, n0y, inv_n0x, h;
    // faces 1..P-1 (slot 0 unused)
    float px[CPD_MAX_PRISMS], py[CPD_MAX_PRISMS], nx[CPD_MAX_PRISMS], ny[CPD_MAX_PRISMS];
    // exit lens relative to its chord midpoint m: centre offset c - m,
    // kappa = |c - m|^2 - R^2, 1/R, bound radius^2
    float mx, my, cmx, cmy;
    float kappa, inv_r, max_dist_sq, lens_scale;
    // detector, relative to m: m - P0, normal, direction
    float mpx, mpy, dnx, dny;
    float ddx, ddy, min_ci, length;
    float slope, inv_slope, intercept, bin_size;
    float beam_w, y_mean, reserved5, reserved6;
};

struct LambdaConsts {
    float r[CPD_MAX_PRISMS + 1];   // n_{i-1}/n_i at faces 1..P-1, n_{P-1} at the lens
    float r2[CPD_MAX_PRISMS + 1];  // r^2
    float dx0, dy0, ts0, tp0;      // beam after the entrance face
    float rs0;                     // its s-polarised Fresnel amplitude
    int ok;                        // 0 when the entrance face reflects totally
    double lambda;
};

// Build the FP32 record from the FP64 design (host or device).
CPD_HD void make_trace_design(const cpd_design &g, uint32_t n_lambda, uint32_t n_rays, TraceDesign &t) {
#pragma clang fp contract(off)
    t.n_glass = g.n_glass;
    t.nbin = g.bin_count;
    t.ar_coated = g.ar_coated;
    t.n_lambda = n_lambda;
    t.n_rays = n_rays;
    t.reserved0 = t.reserved1 = t.reserved2 = 0;
    const double n0x = g.plane_normal[0][0], n0y = g.plane_normal[0][1];
    t.c0 = (float)(g.plane_point[0][0] * n0x + g.plane_point[0][1] * n0y);
    t.n0y = (float)n0y;
    t.inv_n0x = (float)(1.0 / n0x);
    t.h = (float)g.height;
    for (int i = 0; i < CPD_MAX_PRISMS; ++i) {
        t.px[i] = (float)g.plane_point[i][0];
        t.py[i] = (float)g.plane_point[i][1];
        t.nx[i] = (float)g.plane_normal[i][0];
        t.ny[i] = (float)g.plane_normal[i][1];
    }
    const double cmx = g.arc_center[0] - g.arc_midpt[0], cmy = g.arc_center[1] - g.arc_midpt[1];
    t.mx = (float)g.arc_midpt[0];
    t.my = (float)g.arc_midpt[1];
    t.cmx = (float)cmx;
    t.cmy = (float)cmy;
    t.kappa = (float)(cmx * cmx + cmy * cmy - g.arc_radius * g.arc_radius);
    t.inv_r = (float)(1.0 / g.arc_radius);
    t.max_dist_sq = (float)g.arc_max_dist_sq;
    t.lens_scale = (float)(g.arc_radius + sqrt(g.arc_max_dist_sq));
    t.mpx = (float)(g.arc_midpt[0] - g.det_position[0]);
    t.mpy = (float)(g.arc_midpt[1] - g.det_position[1]);
    t.dnx = (float)g.det_normal[0];
    t.dny = (float)g.det_normal[1];
    t.ddx = (float)g.det_direction[0];
    t.ddy = (float)g.det_direction[1];
    t.min_ci = (float)g.det_min_ci;
    t.length = (float)g.det_length;
    t.slope = (float)g.det_slope;
    t.inv_slope = (float)(1.0 / g.det_slope);
    t.intercept = (float)g.det_intercept;
    t.bin_size = (float)g.det_bin_size;
    t.beam_w = (float)g.beam_width;
    t.y_mean = (float)g.y_mean;
    t.reserved5 = t.reserved6 = 0.f;
}

// Wavelength sample j (1-based) of the R1 sequence and its per-face
// constants, all in FP64 (src/spectrometer/distribution.rs:32-34,
// glasscat.rs:17-22, and the entrance refraction of ray.rs:309-367 for the
// beam direction (1, 0), which every ray shares).
CPD_HD double lambda_of(const cpd_design &g, double u) {
#pragma clang fp contract(off)
    return g.wavelength_lo + (g.wavelength_hi - g.wavelength_lo) * u;
}

CPD_HD void make_lambda_consts(const cpd_design &g, double lambda, LambdaConsts &lc) {
#pragma clang fp contract(off)
    const int P = (int)g.n_glass;
    double n[CPD_MAX_PRISMS];
    for (int i = 0; i < P; ++i) n[i] = glass_n64(g.glass[i], lambda);
    lc.lambda = lambda;
    Ray64 ray{0.0, 0.0, 1.0, 0.0, 1.0, 1.0};
    lc.ok = refract64(ray, g.plane_normal[0][0], g.plane_normal[0][1], 1.0, n[0], g.ar_coated != 0) == T_OK;
    lc.dx0 = (float)ray.dx;
    lc.dy0 = (float)ray.dy;
    lc.ts0 = (float)ray.ts;
    lc.tp0 = (float)ray.tp;
    {
        // s amplitude of the entrance refraction of the direction (1, 0)
        double n0 = n[0], ci = -g.plane_normal[0][0];
        double cr2 = 1.0 - (1.0 / n0) * (1.0 / n0) * (1.0 - ci * ci);
        double cr = cr2 > 0.0 ? sqrt(cr2) : 0.0;
        lc.rs0 = (float)((ci - n0 * cr) / (ci + n0 * cr));
    }
    for (int i = 0; i <= CPD_MAX_PRISMS; ++i) {
        double r = 1.0;
        if (i >= 1 && i < P) r = n[i - 1] / n[i];
        if (i == P) r = n[P - 1];
        lc.r[i] = (float)r;
        lc.r2[i] = (float)(r * r);
    }
}

// ---------------------------------------------------------------------------
// FP32 trace with error margins
// ---------------------------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__)
CPD_HD float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
CPD_HD float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
#else
CPD_HD float fast_rcp(float x) { return 1.0f / x; }
CPD_HD float fast_sqrt(float x) { return sqrtf(x); }
#endif

// Relative error budget of an FP32 position or distance per unit of path
// length travelled, and of dimensionless quantities (cosines, cos^2).
static constexpr float kGuardRel = 8.0e-6f;
static constexpr float kGuardCos = 2.0e-5f;

// Per-stage operation counts of the reference formulas with per-design and
// per-wavelength invariants hoisted (add/mul/div/sqrt = 1, fma = 2): beam
// height 2; planar face 42 = intersection 13 + refraction 29; lens 58 =
// intersection, bounds and normal 26 + refraction 32; detector 24; hit 1.
// A ray is charged for the stages it reaches.
static constexpr uint32_t kOpsBeam = 2, kOpsFaceHit = 13, kOpsFaceRefract = 29, kOpsLensHit = 26,
                          kOpsLensRefract = 32, kOpsDetector = 24, kOpsHit = 1;

struct Trace32Result {
    uint32_t bin;  // nbin on a miss
    float t;
    bool guard;    // the outcome is within the FP32 error margin of a decision
    uint32_t ops;  // algorithmic operations of the stages the ray reached
};

CPD_HD bool near_f(float a, float b, float eps) { return fabsf(a - b) <= eps; }

// Margin bookkeeping: a refraction with angular amplification
// A = r cos(ti) / cos(tt) multiplies the direction error it receives by A.
// With rs = (A - 1) / (A + 1) (the s-polarised Fresnel amplitude the trace
// computes anyway), A = (1 + rs) / (1 - rs); the running product of
// max(A, 1) is kept as num / den so that margin tests need no division:
// |x - b| <= kGuardRel * scale * num / den  <=>  |x - b| * den <= kGuardRel * scale * num.
struct Margin {
    float num = 1.f, den = 1.f;
    CPD_HD void refraction(float rs) {
        float a = fmaxf(rs, 0.f);
        num *= 1.f + a;
        den *= 1.f - a;
    }
    CPD_HD bool near(float x, float b, float scale) const { return fabsf(x - b) * den <= kGuardRel * scale * num; }
};

template <int P>
CPD_HD Trace32Result trace32(const TraceDesign &d, const LambdaConsts &lc, float y) {
    Trace32Result res{d.nbin, 0.f, false, kOpsBeam + kOpsFaceHit};
    Margin m;
    // entrance face: the beam starts at (0, y) travelling +x and meets the
    // face at height y, so only the height bound decides
    float scale = fabsf(y) + 1.f;
    res.guard = m.near(y, 0.f, scale) || m.near(y, d.h, scale);
    if (!(y >= 0.f && y <= d.h)) return res;
    float ox = (d.c0 - y * d.n0y) * d.inv_n0x;
    float oy = y;
    res.guard = res.guard || m.near(ox, 0.f, scale);
    if (!(ox >= 0.f)) return res;
    float path = fabsf(ox) + scale;
    res.ops += kOpsFaceRefract;
    float dx = lc.dx0, dy = lc.dy0, ts = lc.ts0, tp = lc.tp0;
    m.refraction(lc.rs0);
    const bool ar = d.ar_coated != 0;
#pragma unroll
    for (int i = 1; i < P; ++i) {
        res.ops += kOpsFaceHit;
        const float nx = d.nx[i], ny = d.ny[i];
        float ci_p = dx * nx + dy * ny;
        float num = (d.px[i] - ox) * nx + (d.py[i] - oy) * ny;
        float dist = num * fast_rcp(ci_p);
        path += fabsf(dist);
        res.guard = res.guard || m.near(dist, 0.f, path);
        if (!(dist >= 0.f)) return res;
        float hx = fmaf(dx, dist, ox), hy = fmaf(dy, dist, oy);
        res.guard = res.guard || m.near(hy, 0.f, path) || m.near(hy, d.h, path);
        if (!(hy >= 0.f && hy <= d.h)) return res;
        ox = hx;
        oy = hy;
        res.ops += kOpsFaceRefract;
        float ci = -ci_p;
        float r = lc.r[i];
        float cr_sq = 1.f - lc.r2[i] * (1.f - ci * ci);
        res.guard = res.guard || fabsf(cr_sq) <= kGuardCos;
        if (!(cr_sq >= 0.f)) return res;
        float cr = fast_sqrt(cr_sq);
        float rci = r * ci;
        float k = rci - cr;
        float vx = dx * r + nx * k;
        float vy = dy * r + ny * k;
        float a = k * fast_rcp(rci + cr);
        if (ar && ci > 0.5f) {
            ts *= 0.99f;
            tp *= 0.99f;
        } else {
            float rcr = r * cr;
            float b = (rcr - ci) * fast_rcp(rcr + ci);
            ts *= 1.f - a * a;
            tp *= 1.f - b * b;
        }
        if (ar) res.guard = res.guard || near_f(ci, 0.5f, kGuardCos);
        m.refraction(a);
        dx = vx;
        dy = vy;
    }
    // exit lens, relative to its chord midpoint m: e = o - m, c = m + cm.
    // |o + s dir - c|^2 = R^2  <=>  s^2 + 2 b s + c2 = 0 with
    // b = dir.(e - cm), c2 = |e|^2 - 2 e.cm + (|cm|^2 - R^2); the roots are
    // formed without cancellation (their product is c2).
    float qx, qy;
    res.ops += kOpsLensHit;
    {
        float ex = ox - d.mx, ey = oy - d.my;
        float b = dx * (ex - d.cmx) + dy * (ey - d.cmy);
        float c2 = (ex * ex + ey * ey) - 2.f * (ex * d.cmx + ey * d.cmy) + d.kappa;
        float disc = b * b - c2;
        res.guard = res.guard || m.near(disc, 0.f, 4.f * (fabsf(b) + d.lens_scale) * path);
        if (!(disc > 0.f)) return res;
        float sd = fast_sqrt(disc);
        float d1, d2;
        if (b > 0.f) {
            d1 = -b - sd;
            d2 = c2 * fast_rcp(d1);
        } else {
            d2 = -b + sd;
            d1 = c2 * fast_rcp(d2);
        }
        float q1x = fmaf(dx, d1, ex), q1y = fmaf(dy, d1, ey);
        float q1 = q1x * q1x + q1y * q1y;
        const float bscale = 4.f * (sqrtf(d.max_dist_sq) + 1.f) * path;
        res.guard = res.guard || m.near(d1, 0.f, path) || m.near(q1, d.max_dist_sq, bscale);
        if (!signbit(d1) && q1 <= d.max_dist_sq) {
            qx = q1x;
            qy = q1y;
            path += fabsf(d1);
        } else {
            float q2x = fmaf(dx, d2, ex), q2y = fmaf(dy, d2, ey);
            float q2 = q2x * q2x + q2y * q2y;
            res.guard = res.guard || m.near(d2, 0.f, path) || m.near(q2, d.max_dist_sq, bscale);
            if (!(!signbit(d2) && q2 <= d.max_dist_sq)) return res;
            qx = q2x;
            qy = q2y;
            path += fabsf(d2);
        }
        res.ops += kOpsLensRefract;
        // normal (c - p) / R = (cm - q) / R
        float nx = (d.cmx - qx) * d.inv_r, ny = (d.cmy - qy) * d.inv_r;
        float ci = -(dx * nx + dy * ny);
        float r = lc.r[P];
        float cr_sq = 1.f - lc.r2[P] * (1.f - ci * ci);
        res.guard = res.guard || fabsf(cr_sq) <= kGuardCos;
        if (!(cr_sq >= 0.f)) return res;
        float cr = fast_sqrt(cr_sq);
        float rci = r * ci;
        float k = rci - cr;
        float vx = dx * r + nx * k;
        float vy = dy * r + ny * k;
        float a = k * fast_rcp(rci + cr);
        if (ar && ci > 0.5f) {
            ts *= 0.99f;
            tp *= 0.99f;
        } else {
            float rcr = r * cr;
            float bb = (rcr - ci) * fast_rcp(rcr + ci);
            ts *= 1.f - a * a;
            tp *= 1.f - bb * bb;
        }
        if (ar) res.guard = res.guard || near_f(ci, 0.5f, kGuardCos);
        m.refraction(a);
        dx = vx;
        dy = vy;
    }
    // detector array: w = o - P0 = (m - P0) + q
    res.ops += kOpsDetector;
    {
        float ci = -(dx * d.dnx + dy * d.dny);
        res.guard = res.guard || m.near(ci, d.min_ci, kGuardCos / kGuardRel);
        if (!(ci > d.min_ci)) return res;
        float wx = d.mpx + qx, wy = d.mpy + qy;
        float dist = (wx * d.dnx + wy * d.dny) * fast_rcp(ci);
        float ux = fmaf(dx, dist, wx), uy = fmaf(dy, dist, wy);
        float pos = ux * d.ddx + uy * d.ddy;
        float scale_d = path + fabsf(dist) + fabsf(wx) + fabsf(wy);
        res.guard = res.guard || m.near(pos, 0.f, scale_d) || m.near(pos, d.length, scale_d);
        if (!(pos >= 0.f && pos <= d.length)) return res;
        float v = pos - d.intercept;
        float qb = floorf(v * d.inv_slope);
        float rem = fmaf(-qb, d.slope, v);
        res.guard = res.guard || m.near(rem, 0.f, scale_d) || m.near(rem, d.slope, scale_d) ||
                    m.near(rem, d.bin_size, scale_d);
        if (!(qb >= 0.f && qb < (float)d.nbin && rem < d.bin_size)) return res;
        res.bin = (uint32_t)qb;
        res.t = (ts + tp) * 0.5f;
        res.ops += kOpsHit;
    }
    return res;
}

}  // namespace cpd
