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
// Per-wavefront LDS histogram of one work item: `cols` columns per bin (lane
// l adds into column l % cols), at most kHistWords words for bin counts up to
// 2048, so the histogram stays near 8 KB and several wavefronts fit a CU.
static constexpr uint32_t kHistWords = 2048;
static constexpr uint32_t kMaxTileRays = 16384;
CPD_HD constexpr uint32_t hist_cols(uint32_t nbin) {
    uint32_t c = 64;
    while (c > 1 && c * nbin > kHistWords) c >>= 1;
    return c;
}
// Largest tile whose per-column u32 sums cannot overflow.  A column gathers
// the FP32-traced rays of its 64/cols lanes (tile/cols at most) plus at most
// one re-traced ray per lane per queue drain (ceil(tile/64) drains), each
// worth at most 2^22: 2*tile + 64 < 1024*cols keeps every sum below 2^32.
CPD_HD constexpr uint32_t max_tile_rays(uint32_t cols) {
    return ((512u * cols - 33u) & ~63u) < kMaxTileRays ? ((512u * cols - 33u) & ~63u) : kMaxTileRays;
}

// Rays per work item of a design: the whole wavelength sample rounded up to
// a wavefront multiple, capped by the overflow bound and by tile_cap (a
// multiple of 64, 0 for none).
CPD_HD uint32_t tile_for(uint32_t nbin, uint32_t n_rays, uint32_t tile_cap) {
    uint32_t tile = (n_rays + 63) / 64 * 64;
    const uint32_t bound = max_tile_rays(hist_cols(nbin));
    if (tile > bound) tile = bound;
    if (tile_cap && tile > tile_cap) tile = tile_cap;
    return tile < 64 ? 64 : tile;
}

struct alignas(16) TraceDesign {
    uint32_t n_glass, nbin, ar_coated, n_lambda;
    uint32_t n_rays, hist_cols, tile_rays, n_tiles;  // LDS columns; rays per work item; items per wavelength
    // entrance face: a beam ray at height y meets it at x = a0 + b0 y; the
    // prism spans heights [0, h], tested as |y - h/2| <= h/2
    float a0, b0, h2, nbin_f;
    // faces 1..P-1 (slot 0 unused): normal and point.normal
    float nx[CPD_MAX_PRISMS], ny[CPD_MAX_PRISMS], pn[CPD_MAX_PRISMS];
    // exit lens relative to its chord midpoint m: centre offset cm = c - m,
    // kappa = |cm|^2 - R^2, 1/R, bound radius^2, error scales
    float mx, my, cmx, cmy;
    float kappa, inv_r, max_dist_sq, lens_scale;
    float bound_scale, reserved3, reserved4, reserved5;
    // detector, relative to m: m - P0, normal, direction, |pos - L/2| <= L/2
    float mpx, mpy, dnx, dny;
    float ddx, ddy, min_ci, half_len;
    float slope, inv_slope, intercept, bin_size;
    float beam_w, y_mean, reserved6, reserved7;
};

struct LambdaConsts {
    float r[CPD_MAX_PRISMS + 1];     // n_{i-1}/n_i at faces 1..P-1, n_{P-1} at the lens
    float r2[CPD_MAX_PRISMS + 1];    // r^2
    float omr2[CPD_MAX_PRISMS + 1];  // 1 - r^2
    float dx0, dy0, ts0, tp0;        // beam after the entrance face
    float amp0;                      // angular amplification of the entrance refraction, >= 1
    int ok;                          // 0 when the entrance face reflects totally
    double lambda;
};

// Build the FP32 record from the FP64 design (host or device).
// tile_cap (a multiple of 64, 0 for none) bounds the rays of one work item.
CPD_HD void make_trace_design(const cpd_design &g, uint32_t n_lambda, uint32_t n_rays, TraceDesign &t,
                              uint32_t tile_cap = 0) {
#pragma clang fp contract(off)
    t.n_glass = g.n_glass;
    t.nbin = g.bin_count;
    t.ar_coated = g.ar_coated;
    t.n_lambda = n_lambda;
    t.n_rays = n_rays;
    t.hist_cols = hist_cols(g.bin_count);
    t.tile_rays = tile_for(g.bin_count, n_rays, tile_cap);
    t.n_tiles = (n_rays + t.tile_rays - 1) / t.tile_rays;
    const double n0x = g.plane_normal[0][0], n0y = g.plane_normal[0][1];
    const double c0 = g.plane_point[0][0] * n0x + g.plane_point[0][1] * n0y;
    t.a0 = (float)(c0 / n0x);
    t.b0 = (float)(-n0y / n0x);
    t.h2 = (float)(g.height * 0.5);
    t.nbin_f = (float)g.bin_count;
    for (int i = 0; i < CPD_MAX_PRISMS; ++i) {
        t.nx[i] = (float)g.plane_normal[i][0];
        t.ny[i] = (float)g.plane_normal[i][1];
        t.pn[i] = (float)(g.plane_point[i][0] * g.plane_normal[i][0] + g.plane_point[i][1] * g.plane_normal[i][1]);
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
    t.bound_scale = (float)(4.0 * (sqrt(g.arc_max_dist_sq) + 1.0));
    t.reserved3 = t.reserved4 = t.reserved5 = 0.f;
    t.mpx = (float)(g.arc_midpt[0] - g.det_position[0]);
    t.mpy = (float)(g.arc_midpt[1] - g.det_position[1]);
    t.dnx = (float)g.det_normal[0];
    t.dny = (float)g.det_normal[1];
    t.ddx = (float)g.det_direction[0];
    t.ddy = (float)g.det_direction[1];
    t.min_ci = (float)g.det_min_ci;
    t.half_len = (float)(g.det_length * 0.5);
    t.slope = (float)g.det_slope;
    t.inv_slope = (float)(1.0 / g.det_slope);
    t.intercept = (float)g.det_intercept;
    t.bin_size = (float)g.det_bin_size;
    t.beam_w = (float)g.beam_width;
    t.y_mean = (float)g.y_mean;
    t.reserved6 = t.reserved7 = 0.f;
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
        double ci = -g.plane_normal[0][0], r = 1.0 / n[0];
        double cr2 = 1.0 - r * r * (1.0 - ci * ci);
        double cr = cr2 > 0.0 ? sqrt(cr2) : 0.0;
        double a = cr > 0.0 ? r * ci / cr : 1e30;
        lc.amp0 = (float)(a > 1.0 ? a : 1.0);
    }
    for (int i = 0; i <= CPD_MAX_PRISMS; ++i) {
        double r = 1.0;
        if (i >= 1 && i < P) r = n[i - 1] / n[i];
        if (i == P) r = n[P - 1];
        lc.r[i] = (float)r;
        lc.r2[i] = (float)(r * r);
        lc.omr2[i] = (float)(1.0 - r * r);
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

// Error budget of an FP32 length per unit of path travelled, and of an FP32
// cosine, before angular amplification.
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
    uint32_t ops;  // algorithmic operations of the stages the ray reached (COUNT)
};

// Snell refraction of (dx, dy) at unit normal (nx, ny) with ci = -d.n and
// ratio r = n1/n2 (r2 = r^2, omr2 = 1 - r^2), plus the Fresnel factors
// 1 - rs^2 and 1 - rp^2 with rs = (r ci - cr) / (r ci + cr) and
// rp = (r cr - ci) / (r cr + ci) (the reference's amplitudes divided through
// by n2) from one reciprocal.  The angular amplification of the refraction,
// A = r ci / cr, is folded into the margin as amp_num / amp_den.
struct Refract32 {
    float vx, vy, fs, fp;
    float cr_sq;
    float amp_num, amp_den;
};

CPD_HD Refract32 refract32(float dx, float dy, float nx, float ny, float ci, float r, float r2, float omr2) {
    Refract32 o;
    o.cr_sq = fmaf(r2, ci * ci, omr2);
    const float cr = fast_sqrt(fmaxf(o.cr_sq, 0.f));
    const float rci = r * ci;
    const float k = rci - cr;
    o.vx = fmaf(nx, k, dx * r);
    o.vy = fmaf(ny, k, dy * r);
    const float a = rci + cr;
    const float b = fmaf(r, cr, ci);
    const float inv = fast_rcp(a * b);
    const float as = k * (b * inv);
    const float ap = fmaf(r, cr, -ci) * (a * inv);
    o.fs = fmaf(-as, as, 1.f);
    o.fp = fmaf(-ap, ap, 1.f);
    o.amp_num = fmaxf(rci, cr);
    o.amp_den = cr;
    return o;
}

// The constants one ray of a P-glass design at one wavelength reads: the
// hot subset of TraceDesign and LambdaConsts, small enough to live in scalar
// registers for the whole ray loop.  Faces 1..P-1 are stored at [0, P-2].
template <int P>
struct alignas(4) HotRecord {
    static constexpr int F = P > 1 ? P - 1 : 1;
    uint32_t nbin, ar_coated;
    float a0, b0, h2, y_mean, beam_w, nbin_f;
    float dx0, dy0, ts0, tp0, amp0;
    float fnx[F], fny[F], fpn[F], fr[F], fr2[F], fomr2[F];
    float mx, my, cmx, cmy, kappa, inv_r, max_dist_sq, lens_scale, bound_scale;
    float lr, lr2, lomr2;
    float mpx, mpy, dnx, dny, ddx, ddy, min_ci, half_len;
    float slope, inv_slope, intercept, bin_size;
};

template <int P>
CPD_HD HotRecord<P> make_hot(const TraceDesign &d, const LambdaConsts &lc) {
    HotRecord<P> h;
    h.nbin = d.nbin;
    h.ar_coated = d.ar_coated;
    h.a0 = d.a0;
    h.b0 = d.b0;
    h.h2 = d.h2;
    h.y_mean = d.y_mean;
    h.beam_w = d.beam_w;
    h.nbin_f = d.nbin_f;
    h.dx0 = lc.dx0;
    h.dy0 = lc.dy0;
    h.ts0 = lc.ts0;
    h.tp0 = lc.tp0;
    h.amp0 = lc.amp0;
    for (int i = 0; i < HotRecord<P>::F; ++i) {
        const int f = i + 1 < CPD_MAX_PRISMS ? i + 1 : CPD_MAX_PRISMS - 1;
        h.fnx[i] = d.nx[f];
        h.fny[i] = d.ny[f];
        h.fpn[i] = d.pn[f];
        h.fr[i] = lc.r[f];
        h.fr2[i] = lc.r2[f];
        h.fomr2[i] = lc.omr2[f];
    }
    h.mx = d.mx;
    h.my = d.my;
    h.cmx = d.cmx;
    h.cmy = d.cmy;
    h.kappa = d.kappa;
    h.inv_r = d.inv_r;
    h.max_dist_sq = d.max_dist_sq;
    h.lens_scale = d.lens_scale;
    h.bound_scale = d.bound_scale;
    h.lr = lc.r[P];
    h.lr2 = lc.r2[P];
    h.lomr2 = lc.omr2[P];
    h.mpx = d.mpx;
    h.mpy = d.mpy;
    h.dnx = d.dnx;
    h.dny = d.dny;
    h.ddx = d.ddx;
    h.ddy = d.ddy;
    h.min_ci = d.min_ci;
    h.half_len = d.half_len;
    h.slope = d.slope;
    h.inv_slope = d.inv_slope;
    h.intercept = d.intercept;
    h.bin_size = d.bin_size;
    return h;
}

template <int P, bool COUNT>
CPD_HD Trace32Result trace32(const HotRecord<P> &d, float y) {
    Trace32Result res{d.nbin, 0.f, false, 0};
    const bool ar = d.ar_coated != 0;
    // entrance face: the beam ray (0, y) travelling +x meets it at height y
    const float ay = y - d.h2;
    bool alive = fabsf(ay) <= d.h2;
    bool guard = fabsf(fabsf(ay) - d.h2) <= kGuardRel * (fabsf(y) + 1.f);
    float ox = fmaf(d.b0, y, d.a0), oy = y;
    alive = alive && ox >= 0.f;
    float path = fabsf(ox) + fabsf(y) + 1.f;
    float num = d.amp0, den = 1.f;
    float dx = d.dx0, dy = d.dy0, ts = d.ts0, tp = d.tp0;
    if (COUNT) res.ops = kOpsBeam + kOpsFaceHit + (alive ? kOpsFaceRefract : 0u);
#pragma unroll
    for (int i = 1; i < P; ++i) {
        if (COUNT) res.ops += alive ? kOpsFaceHit : 0u;
        const float nx = d.fnx[i - 1], ny = d.fny[i - 1];
        const float ci_p = fmaf(dx, nx, dy * ny);
        const float dist = (d.fpn[i - 1] - fmaf(ox, nx, oy * ny)) * fast_rcp(ci_p);
        const float hx = fmaf(dx, dist, ox), hy = fmaf(dy, dist, oy);
        path += fabsf(dist);
        const float a = hy - d.h2;
        guard = guard || (alive && fabsf(fabsf(a) - d.h2) * den <= kGuardRel * path * num);
        alive = alive && dist >= 0.f && fabsf(a) <= d.h2;
        if (COUNT) res.ops += alive ? kOpsFaceRefract : 0u;
        ox = hx;
        oy = hy;
        const float ci = -ci_p;
        const Refract32 f = refract32(dx, dy, nx, ny, ci, d.fr[i - 1], d.fr2[i - 1], d.fomr2[i - 1]);
        guard = guard || (alive && fabsf(f.cr_sq) <= kGuardCos);
        alive = alive && f.cr_sq >= 0.f;
        if (ar) {
            guard = guard || (alive && fabsf(ci - 0.5f) * den <= kGuardCos * num);
            const bool coat = ci > 0.5f;
            ts *= coat ? 0.99f : f.fs;
            tp *= coat ? 0.99f : f.fp;
        } else {
            ts *= f.fs;
            tp *= f.fp;
        }
        num *= f.amp_num;
        den *= f.amp_den;
        dx = f.vx;
        dy = f.vy;
    }
    // exit lens, relative to its chord midpoint m: e = o - m, c = m + cm.
    // |o + s dir - c|^2 = R^2  <=>  s^2 + 2 b s + c2 = 0 with
    // b = dir.(e - cm) and c2 = |e|^2 - 2 e.cm + (|cm|^2 - R^2); the root of
    // larger magnitude is -(b + sign(b) sqrt(disc)) and the other one is
    // c2 over it, so neither root cancels.
    if (COUNT) res.ops += alive ? kOpsLensHit : 0u;
    float qx, qy;
    {
        const float ex = ox - d.mx, ey = oy - d.my;
        const float b = fmaf(dx, ex - d.cmx, dy * (ey - d.cmy));
        const float e2 = fmaf(ex, ex, ey * ey);
        const float ec = fmaf(ex, d.cmx, ey * d.cmy);
        const float c2 = fmaf(-2.f, ec, e2 + d.kappa);
        const float disc = fmaf(b, b, -c2);
        const float lim = kGuardRel * path * num;
        guard = guard || (alive && fabsf(disc) * den <= lim * 4.f * (fabsf(b) + d.lens_scale));
        alive = alive && disc > 0.f;
        const float sd = fast_sqrt(fmaxf(disc, 0.f));
        const float big = -(b + copysignf(sd, b));
        const float small = c2 * fast_rcp(big);
        const bool bpos = !signbit(b);
        const float d1 = bpos ? big : small, d2 = bpos ? small : big;
        const float q1x = fmaf(dx, d1, ex), q1y = fmaf(dy, d1, ey);
        const float q2x = fmaf(dx, d2, ex), q2y = fmaf(dy, d2, ey);
        const float s1 = fmaf(q1x, q1x, q1y * q1y), s2 = fmaf(q2x, q2x, q2y * q2y);
        const bool ok1 = d1 >= 0.f && s1 <= d.max_dist_sq;
        const bool ok2 = d2 >= 0.f && s2 <= d.max_dist_sq;
        const float blim = lim * d.bound_scale;
        guard = guard || (alive && (fabsf(s1 - d.max_dist_sq) * den <= blim ||
                                    (!ok1 && fabsf(s2 - d.max_dist_sq) * den <= blim)));
        alive = alive && (ok1 || ok2);
        if (COUNT) res.ops += alive ? kOpsLensRefract : 0u;
        qx = ok1 ? q1x : q2x;
        qy = ok1 ? q1y : q2y;
        path += fabsf(ok1 ? d1 : d2);
        // normal (c - p) / R = (cm - q) / R
        const float nx = (d.cmx - qx) * d.inv_r, ny = (d.cmy - qy) * d.inv_r;
        const float ci = -fmaf(dx, nx, dy * ny);
        const Refract32 f = refract32(dx, dy, nx, ny, ci, d.lr, d.lr2, d.lomr2);
        guard = guard || (alive && fabsf(f.cr_sq) <= kGuardCos);
        alive = alive && f.cr_sq >= 0.f;
        if (ar) {
            guard = guard || (alive && fabsf(ci - 0.5f) * den <= kGuardCos * num);
            const bool coat = ci > 0.5f;
            ts *= coat ? 0.99f : f.fs;
            tp *= coat ? 0.99f : f.fp;
        } else {
            ts *= f.fs;
            tp *= f.fp;
        }
        num *= f.amp_num;
        den *= f.amp_den;
        dx = f.vx;
        dy = f.vy;
    }
    // detector array: w = o - P0 = (m - P0) + q
    if (COUNT) res.ops += alive ? kOpsDetector : 0u;
    {
        const float ci = -fmaf(dx, d.dnx, dy * d.dny);
        guard = guard || (alive && fabsf(ci - d.min_ci) * den <= kGuardCos * num);
        alive = alive && ci > d.min_ci;
        const float wx = d.mpx + qx, wy = d.mpy + qy;
        const float dist = fmaf(wx, d.dnx, wy * d.dny) * fast_rcp(ci);
        const float pos = fmaf(fmaf(dx, dist, wx), d.ddx, fmaf(dy, dist, wy) * d.ddy);
        const float lim = kGuardRel * (path + fabsf(dist) + fabsf(wx) + fabsf(wy)) * num;
        const float ap = pos - d.half_len;
        guard = guard || (alive && fabsf(fabsf(ap) - d.half_len) * den <= lim);
        alive = alive && fabsf(ap) <= d.half_len;
        const float v = pos - d.intercept;
        const float qb = floorf(v * d.inv_slope);
        const float rem = fmaf(-qb, d.slope, v);
        const float edge = fminf(fminf(fabsf(rem), fabsf(d.slope - rem)), fabsf(rem - d.bin_size));
        guard = guard || (alive && edge * den <= lim);
        alive = alive && qb >= 0.f && qb < d.nbin_f && rem < d.bin_size;
        if (alive) {
            res.bin = (uint32_t)qb;
            res.t = (ts + tp) * 0.5f;
        }
        if (COUNT) res.ops += alive ? kOpsHit : 0u;
    }
    res.guard = guard;
    return res;
}

}  // namespace cpd
