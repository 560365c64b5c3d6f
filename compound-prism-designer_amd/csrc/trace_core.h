// trace_core.h -- per-ray trace of the 2D compound-prism spectrometer.
//
// Two traces live here:
//  * trace32: the FP32 hot path.  Per-wavelength invariants arrive
//    precomputed in FP64 and rounded once: the planar faces, which every ray
//    of the collimated beam crosses with the same direction, collapse to an
//    interval of accepted beam heights and an affine exit point.  The exit
//    lens and the detector are solved relative to the lens' chord midpoint,
//    which avoids the cancellation of |o - c|^2 - R^2 at lens radii far
//    larger than the prism.  Every comparison whose outcome decides hit/miss
//    or the bin carries an error margin; a ray that lands inside a margin is
//    flagged for the FP64 trace instead of being counted.  Wave-uniform
//    exits follow the planar-chain test and the detector's acceptance test,
//    which a ray passes or fails before its hit point and direction at the
//    lens are formed.
//  * trace64: the FP64 trace in the reference's operation order
//    (src/spectrometer/ray.rs:309-382, 485-499; geometry.rs:124-145,
//    161-216; spectrometer.rs:265-303), with fused multiply-adds exactly at the
//    reference's mul_add sites; trace64_chain is the same trace with the
//    planar faces' directions and transmittances taken from the wavelength's
//    Chain64 record.  Flagged rays are re-traced with it, so their bin
//    decision is the FP64 one.
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
// The same sequence from any seed (a fraction in 64-bit fixed point).
CPD_HD double qmc_seeded(uint64_t seed_fx, uint64_t k) { return (double)((seed_fx + k * kAlphaFx) >> 11) * 0x1.0p-53; }
// A seed in [0, 1) (any real: its fractional part) as a 64-bit fraction.
CPD_HD uint64_t seed_to_fx(double seed) {
    double f = seed - floor(seed);
    if (!(f >= 0.0 && f < 1.0)) f = 0.0;
    return (uint64_t)(f * 0x1.0p+64);  // f <= 1 - 2^-53, so the product is below 2^64
}

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

// Beam height of a ray from its quantile z: y = y_mean - width z
// (spectrometer.rs:26-30), a separate multiply and subtract.
CPD_HD double beam_y64(const cpd_design &g, double z) {
#pragma clang fp contract(off)
    return g.y_mean - g.beam_width * z;
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
// 1024: 4 KB, so that a CU's LDS holds the 24 wavefronts (six per SIMD)
// the trace kernel runs with.
static constexpr uint32_t kHistWords = 1024;
// Rays per work item: at most kMaxTileRays; kDefaultTileRays unless a
// schedule caps it.
static constexpr uint32_t kMaxTileRays = 1u << 20;
static constexpr uint32_t kDefaultTileRays = 65536;
// Rays 1..kBandRays, kBandRays + 1..2 kBandRays, ... form the blocks within
// which the banded beam tables are sorted by height.
static constexpr uint32_t kBandRays = 4096;
CPD_HD constexpr uint32_t hist_cols(uint32_t nbin) {
    uint32_t c = 64;
    while (c > 1 && c * nbin > kHistWords) c >>= 1;
    return c;
}
// Rays between two folds of the u32 columns into the wavefront's u64 row.
// Over a stretch of S rays a column gathers at most S / cols FP32-traced rays
// of its 64 / cols lanes plus one re-traced ray per lane per queue drain; the
// queue carries < 128 entries in, so there are at most S / 64 + 2 drains.
// Each ray adds at most 2^22, so (2 S + 128) / cols < 1024 keeps every
// column below 2^32.
CPD_HD constexpr uint32_t fold_rays(uint32_t cols) { return (512u * cols - 65u) & ~63u; }

// Positions per work item of a design: the whole wavelength sample's
// positions (n_pos) rounded up to a wavefront multiple, capped by tile_cap (a
// multiple of 64; 0 for the default cap).
CPD_HD uint32_t tile_for(uint32_t nbin, uint32_t n_pos, uint32_t tile_cap) {
    (void)nbin;
    uint32_t tile = (n_pos + 63) / 64 * 64;
    const uint32_t cap = tile_cap ? tile_cap : kDefaultTileRays;
    if (tile > cap) tile = cap;
    return tile < 64 ? 64 : tile;
}

struct alignas(16) TraceDesign {
    uint32_t n_glass, nbin, ar_coated, n_lambda;
    uint32_t n_rays, hist_cols, tile_rays, n_tiles;  // LDS columns; positions per work item; items per wavelength
    // exit lens relative to its chord midpoint m: centre offset cm = c - m,
    // 1/R, bound radius^2, error scales
    float cmx, cmy, inv_r, max_dist_sq;
    float lens_scale, bound_scale, nbin_f;
    uint32_t skip;  // nonzero: the design failed construction; its work items trace nothing
    // detector, relative to m: m - P0, normal, direction, |pos - L/2| <= L/2
    float mpx, mpy, dnx, dny;
    float ddx, ddy, min_ci, half_len;
    float slope, inv_slope, intercept, bin_size;
    float beam_w, y_mean, reserved1, reserved2;
};

// Per-wavelength constants of one design.  Every ray of the collimated beam
// enters travelling +x, and refraction at a plane depends only on the
// direction, so through the planar faces all rays share one direction and
// one pair of Fresnel factors; they are computed once, in FP64 and in the
// reference's operation order, and are the reference's values exactly.  The
// hit point on each face is affine in the beam height y, so the planar chain
// reduces to an interval of accepted heights and an affine map to the point
// where a ray leaves the last face.
struct LambdaConsts {
    float ylo, yhi;            // accepted beam heights (empty when ylo > yhi)
    float e0x, e0y, e1x, e1y;  // last planar hit point minus the lens chord midpoint: e0 + e1 y
    float lb0, lb1;            // lens: b(y) = dir.(e - cm) = lb0 + lb1 y
    float lc0, lc1, lc2;       // lens: |e|^2 - 2 e.cm + |cm|^2 - R^2 = lc0 + lc1 y + lc2 y^2
    float dx, dy, rdx, rdy;    // direction into the lens, and times the lens ratio r
    // arc bounds: on the lens circle |q|^2 - max_dist_sq = 2 q.cm + R^2 -
    // |cm|^2 - max_dist_sq, affine in the hit point q = e + s dir:
    // lg0 + lg1 y + lcmd2 s
    float lg0, lg1, lcmd2;
    float ts, tp;              // Fresnel transmittances through the planar faces
    float amp;                 // angular amplification through the planar faces, >= 1
    float path;                // length scale of the FP32 error budget up to the lens
    float lr, lomr2;           // lens ratio r = n_{P-1}, 1 - r^2
    float lr2r;                // r^2 / R^2 (R the lens radius)
    float nd0, nd1, ddn, rddn; // detector normal dn: (cm - e).dn / R = nd0 + nd1 y, dir.dn / R, r dir.dn
    // The array, for the lens hit q = e + tq dir and w = q + m - P0 (P0 the
    // array's start, dd its direction): w.dn = wdn0 + wdn1 y + wdn2 tq;
    // w.dd - intercept = wdd0 + wdd1 y + wdd2 tq; the lens normal's
    // n.dd = ndd0 + ndd1 y + ndd2 tq; r dir.dd
    float wdn0, wdn1, wdn2, wdd0, wdd1, wdd2, ndd0, ndd1, ndd2, rdd;
    float hc;                  // half the array's length minus the intercept
    float dlen0, dlen1;        // the detector stage's length scale is dlen0 + dlen1 |tq| + |dist|
    int ok;                    // 0 when every ray is lost in the planar faces (TIR, parallel face)
    int dead;                  // 1 when a band test over every beam height proves every ray lost
    double lambda;
};

// The planar faces of one wavelength in FP64, as the per-ray trace meets
// them: the direction arriving at face i and its cosine with the face normal
// (identical for every ray of the collimated beam, since refraction reads
// only the direction), and the direction and transmittances with which rays
// leave the last face towards the lens.  trace64_chain re-traces single rays
// with it.
struct Chain64 {
    double dx[CPD_MAX_PRISMS], dy[CPD_MAX_PRISMS], ci[CPD_MAX_PRISMS];
    double ldx, ldy, ts, tp, n_last;
};

// The lens and detector-acceptance constants of one wavelength in FP64, for
// the band test below: the same quantities as LambdaConsts (relative to the
// lens chord midpoint), before rounding.
struct LensBand64 {
    double lb0, lb1, lc0, lc1, lc2;  // b(y) = lb0 + lb1 y; c(y) = lc0 + lc1 y + lc2 y^2
    double e0x, e0y, e1x, e1y;       // last planar hit point minus the chord midpoint: e0 + e1 y
    double dx, dy;                   // direction into the lens
    double lr2r, lomr2, r;           // cr^2 = lomr2 + lr2r disc; lens ratio r
    double radius, inv_r, max_dist_sq;
    double nd0, nd1, ddn, rddn, min_ci;
    double cmx, cmy;                 // lens centre minus the chord midpoint
    double mpx, mpy;                 // chord midpoint minus the array's start
    double dnx, dny, ddx, ddy, length;  // array normal, direction, length
    double ylo, yhi, path;           // accepted beam heights; length scale of the planar chain
};

// The three per-wavelength records of one (design, wavelength sample), as
// make_lambda_consts builds them; a batch computes them once per wavelength,
// one lane each, before the rays (every work item of that wavelength copies
// them into its LDS).
struct alignas(16) LambdaRecord {
    LambdaConsts lc;
    Chain64 chain;
    LensBand64 band;
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
    // work items cover the banded positions 1..P, the rays rounded up to whole
    // blocks of kBandRays (positions past the rays in the last block idle)
    const uint32_t positions = (uint32_t)(((uint64_t)n_rays + kBandRays - 1) / kBandRays * kBandRays);
    t.tile_rays = tile_for(g.bin_count, positions, tile_cap);
    t.n_tiles = (positions + t.tile_rays - 1) / t.tile_rays;
    t.cmx = (float)(g.arc_center[0] - g.arc_midpt[0]);
    t.cmy = (float)(g.arc_center[1] - g.arc_midpt[1]);
    t.inv_r = (float)(1.0 / g.arc_radius);
    t.max_dist_sq = (float)g.arc_max_dist_sq;
    t.lens_scale = (float)(g.arc_radius + sqrt(g.arc_max_dist_sq));
    t.bound_scale = (float)(2.0 * (sqrt((g.arc_center[0] - g.arc_midpt[0]) * (g.arc_center[0] - g.arc_midpt[0]) +
                                        (g.arc_center[1] - g.arc_midpt[1]) * (g.arc_center[1] - g.arc_midpt[1])) +
                                   1.0));
    t.nbin_f = (float)g.bin_count;
    t.skip = 0;
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
    t.reserved1 = t.reserved2 = 0.f;
}

// Wavelength sample j (1-based) of the R1 sequence (src/spectrometer/
// distribution.rs:32-34).
CPD_HD double lambda_of(const cpd_design &g, double u) {
#pragma clang fp contract(off)
    return g.wavelength_lo + (g.wavelength_hi - g.wavelength_lo) * u;
}

// Fold the linear condition c + m y >= 0 on the beam height into [lo, hi].
CPD_HD void clip_heights(double c, double m, double &lo, double &hi) {
    if (m > 0.0) {
        const double t = -c / m;
        if (t > lo) lo = t;
    } else if (m < 0.0) {
        const double t = -c / m;
        if (t < hi) hi = t;
    } else if (!(c >= 0.0)) {
        lo = INFINITY;
        hi = -INFINITY;
    }
}

// Angular amplification max(r ci, cr) / cr of a refraction (>= 1), the
// factor by which it can grow an error of the incoming direction.
CPD_HD double refraction_gain(double dx, double dy, double nx, double ny, double n1, double n2) {
    const double ci = -(dx * nx + dy * ny), r = n1 / n2;
    const double cr2 = 1.0 - r * r * (1.0 - ci * ci);
    if (!(cr2 > 0.0)) return 1.0;
    const double cr = sqrt(cr2), a = r * ci / cr;
    return a > 1.0 ? a : 1.0;
}

// The planar chain for one wavelength (src/spectrometer/ray.rs:485-499 with
// geometry.rs:124-145 and ray.rs:309-367 per face), in FP64.  The shared
// direction and Fresnel factors use refract64, the reference's operation
// order.  Face i's hit point is o_i(y) = A_i + B_i y and its distance along
// the shared direction d is (p.n - o_i.n) / (d.n), affine in y as well; the
// conditions d >= 0 and 0 <= hit y <= h become bounds on y.
CPD_HD void make_lambda_consts(const cpd_design &g, double lambda, LambdaConsts &lc, Chain64 *chain = nullptr,
                               LensBand64 *band = nullptr) {
#pragma clang fp contract(off)
    const int P = (int)g.n_glass;
    const bool ar = g.ar_coated != 0;
    double n[CPD_MAX_PRISMS];
    for (int i = 0; i < P; ++i) n[i] = glass_n64(g.glass[i], lambda);
    lc.lambda = lambda;
    const double h = g.height;
    // entrance face: the ray from (0, y) along +x meets it at
    // x = ((p - o).n) / n.x = (p.n - y n.y) / n.x, at height y
    const double n0x = g.plane_normal[0][0], n0y = g.plane_normal[0][1];
    const double c0 = g.plane_point[0][0] * n0x + g.plane_point[0][1] * n0y;
    bool ok = n0x != 0.0;
    double ax = ok ? c0 / n0x : 0.0, bx = ok ? -n0y / n0x : 0.0, ay = 0.0, by = 1.0;
    double lo = 0.0, hi = h;
    clip_heights(ax, bx, lo, hi);
    double path = 1.0 + h + fabs(ax) + fabs(bx) * h;
    double amp = refraction_gain(1.0, 0.0, n0x, n0y, 1.0, n[0]);
    Ray64 ray{0.0, 0.0, 1.0, 0.0, 1.0, 1.0};
    if (chain) {
        chain->dx[0] = ray.dx;
        chain->dy[0] = ray.dy;
        chain->ci[0] = ray.dx * n0x + ray.dy * n0y;
    }
    ok = ok && refract64(ray, n0x, n0y, 1.0, n[0], ar) == T_OK;
    for (int i = 1; i < P && ok; ++i) {
        const double nx = g.plane_normal[i][0], ny = g.plane_normal[i][1];
        const double ci = ray.dx * nx + ray.dy * ny;
        if (ci == 0.0) {
            ok = false;
            break;
        }
        if (chain) {
            chain->dx[i] = ray.dx;
            chain->dy[i] = ray.dy;
            chain->ci[i] = ci;
        }
        const double pn = g.plane_point[i][0] * nx + g.plane_point[i][1] * ny;
        const double alpha = (pn - (ax * nx + ay * ny)) / ci;
        const double beta = -(bx * nx + by * ny) / ci;
        clip_heights(alpha, beta, lo, hi);
        ax = ax + ray.dx * alpha;
        ay = ay + ray.dy * alpha;
        bx = bx + ray.dx * beta;
        by = by + ray.dy * beta;
        clip_heights(ay, by, lo, hi);
        clip_heights(h - ay, -by, lo, hi);
        path += fabs(alpha) + fabs(beta) * h;
        amp *= refraction_gain(ray.dx, ray.dy, nx, ny, n[i - 1], n[i]);
        ok = refract64(ray, nx, ny, n[i - 1], n[i], ar) == T_OK;
    }
    lc.ok = ok ? 1 : 0;
    lc.dead = 0;
    if (chain) {
        chain->ldx = ray.dx;
        chain->ldy = ray.dy;
        chain->ts = ray.ts;
        chain->tp = ray.tp;
        chain->n_last = n[P - 1];
    }
    lc.ylo = (float)lo;
    lc.yhi = (float)hi;
    // relative to the lens chord midpoint m; c = m + cm
    const double e0x = ax - g.arc_midpt[0], e0y = ay - g.arc_midpt[1];
    const double cmx = g.arc_center[0] - g.arc_midpt[0], cmy = g.arc_center[1] - g.arc_midpt[1];
    const double kappa = cmx * cmx + cmy * cmy - g.arc_radius * g.arc_radius;
    lc.e0x = (float)e0x;
    lc.e0y = (float)e0y;
    lc.e1x = (float)bx;
    lc.e1y = (float)by;
    lc.lb0 = (float)(ray.dx * (e0x - cmx) + ray.dy * (e0y - cmy));
    lc.lb1 = (float)(ray.dx * bx + ray.dy * by);
    lc.lc0 = (float)(e0x * e0x + e0y * e0y - 2.0 * (e0x * cmx + e0y * cmy) + kappa);
    lc.lc1 = (float)(2.0 * (e0x * bx + e0y * by) - 2.0 * (bx * cmx + by * cmy));
    lc.lc2 = (float)(bx * bx + by * by);
    lc.lg0 = (float)(2.0 * (e0x * cmx + e0y * cmy) + (g.arc_radius * g.arc_radius - (cmx * cmx + cmy * cmy)) -
                     g.arc_max_dist_sq);
    lc.lg1 = (float)(2.0 * (bx * cmx + by * cmy));
    lc.lcmd2 = (float)(2.0 * (ray.dx * cmx + ray.dy * cmy));
    const double r = n[P - 1];
    lc.dx = (float)ray.dx;
    lc.dy = (float)ray.dy;
    lc.rdx = (float)(ray.dx * r);
    lc.rdy = (float)(ray.dy * r);
    lc.ts = (float)ray.ts;
    lc.tp = (float)ray.tp;
    lc.amp = (float)amp;
    lc.path = (float)(path + fabs(e0x) + fabs(e0y) + (fabs(bx) + fabs(by)) * h);
    lc.lr = (float)r;
    lc.lr2r = (float)(r * r / (g.arc_radius * g.arc_radius));
    lc.lomr2 = (float)(1.0 - r * r);
    const double dnx = g.det_normal[0], dny = g.det_normal[1];
    lc.nd0 = (float)(((cmx - e0x) * dnx + (cmy - e0y) * dny) / g.arc_radius);
    lc.nd1 = (float)(-(bx * dnx + by * dny) / g.arc_radius);
    lc.ddn = (float)((ray.dx * dnx + ray.dy * dny) / g.arc_radius);
    lc.rddn = (float)(r * (ray.dx * dnx + ray.dy * dny));
    {
        const double ddx = g.det_direction[0], ddy = g.det_direction[1];
        const double wx0 = g.arc_midpt[0] - g.det_position[0] + e0x, wy0 = g.arc_midpt[1] - g.det_position[1] + e0y;
        const double dir_dd = ray.dx * ddx + ray.dy * ddy, e1_dd = bx * ddx + by * ddy;
        lc.wdn0 = (float)(wx0 * dnx + wy0 * dny);
        lc.wdn1 = (float)(bx * dnx + by * dny);
        lc.wdn2 = (float)(ray.dx * dnx + ray.dy * dny);
        lc.wdd0 = (float)(wx0 * ddx + wy0 * ddy - g.det_intercept);
        lc.wdd1 = (float)e1_dd;
        lc.wdd2 = (float)dir_dd;
        lc.ndd0 = (float)(((cmx - e0x) * ddx + (cmy - e0y) * ddy) / g.arc_radius);
        lc.ndd1 = (float)(-e1_dd / g.arc_radius);
        lc.ndd2 = (float)(-dir_dd / g.arc_radius);
        lc.rdd = (float)(r * dir_dd);
        lc.hc = (float)(g.det_length * 0.5 - g.det_intercept);
        // |w|_1 <= |w(y)|_1 + |dir|_1 |tq| with w(y) = w(0) + e1 y, whose
        // 1-norm is convex in y: its larger value at the FP32 ends of the
        // accepted heights bounds it for every ray that can reach the array.
        // With the path to the lens (which holds |e1|_1 h >= |w(0) - w(y)|_1)
        // and the intercept it bounds the magnitudes the detector stage's
        // FP32 values are formed from.
        const double pl = path + fabs(e0x) + fabs(e0y) + (fabs(bx) + fabs(by)) * h;
        const double ya = (double)lc.ylo, yb = (double)lc.yhi;
        const double wa = fabs(wx0 + bx * ya) + fabs(wy0 + by * ya), wb = fabs(wx0 + bx * yb) + fabs(wy0 + by * yb);
        lc.dlen0 = (float)((pl + (wa > wb ? wa : wb) + fabs(g.det_intercept)) * 1.000001);
        lc.dlen1 = (float)((1.0 + fabs(ray.dx) + fabs(ray.dy)) * 1.000001);
    }
    if (band) {
        band->lb0 = ray.dx * (e0x - cmx) + ray.dy * (e0y - cmy);
        band->lb1 = ray.dx * bx + ray.dy * by;
        band->lc0 = e0x * e0x + e0y * e0y - 2.0 * (e0x * cmx + e0y * cmy) + kappa;
        band->lc1 = 2.0 * (e0x * bx + e0y * by) - 2.0 * (bx * cmx + by * cmy);
        band->lc2 = bx * bx + by * by;
        band->e0x = e0x;
        band->e0y = e0y;
        band->e1x = bx;
        band->e1y = by;
        band->dx = ray.dx;
        band->dy = ray.dy;
        band->r = r;
        band->lr2r = r * r / (g.arc_radius * g.arc_radius);
        band->lomr2 = 1.0 - r * r;
        band->radius = g.arc_radius;
        band->inv_r = 1.0 / g.arc_radius;
        band->max_dist_sq = g.arc_max_dist_sq;
        band->nd0 = (cmx - e0x) * dnx + (cmy - e0y) * dny;
        band->nd1 = -(bx * dnx + by * dny);
        band->ddn = ray.dx * dnx + ray.dy * dny;
        band->rddn = r * band->ddn;
        band->min_ci = g.det_min_ci;
        band->cmx = cmx;
        band->cmy = cmy;
        band->mpx = g.arc_midpt[0] - g.det_position[0];
        band->mpy = g.arc_midpt[1] - g.det_position[1];
        band->dnx = dnx;
        band->dny = dny;
        band->ddx = g.det_direction[0];
        band->ddy = g.det_direction[1];
        band->length = g.det_length;
        band->ylo = lo;
        band->yhi = hi;
        band->path = path + fabs(e0x) + fabs(e0y) + (fabs(bx) + fabs(by)) * h;
    }
}

// ---------------------------------------------------------------------------
// Band test: can any ray with beam height in [ya, yb] reach the detector?
// ---------------------------------------------------------------------------
// Interval arithmetic in FP64 over the lens intersection, the refraction and
// the detector's acceptance cosine as functions of the beam height (the
// planar faces are affine in it).  It returns false only when, for every
// height of the band, every root of the lens the trace could take (the near
// one, and the far one unless the near one is valid for the whole band)
// either does not exist, lies behind the ray or outside the arc's bounds,
// refracts by total internal reflection, leaves at an acceptance cosine at
// most the detector's minimum, or meets the array's line off its ends.  Each
// test keeps a margin of 1e-9 of the
// magnitudes it was formed from (the discriminant's the FP64 trace's own
// cancellation of (R + |o - m|)^2 terms), far above the rounding of either
// computation, so the FP64 trace in the reference's operation order misses
// with every ray of such a band, whatever the FP32 trace would have decided,
// and the band can be skipped.
struct Iv {
    double lo, hi;
};
CPD_HD Iv iv_add(Iv a, Iv b) { return Iv{a.lo + b.lo, a.hi + b.hi}; }
CPD_HD Iv iv_sub(Iv a, Iv b) { return Iv{a.lo - b.hi, a.hi - b.lo}; }
CPD_HD Iv iv_scale(Iv a, double k) { return k >= 0.0 ? Iv{a.lo * k, a.hi * k} : Iv{a.hi * k, a.lo * k}; }
CPD_HD Iv iv_affine(double c0, double c1, Iv y) {  // c0 + c1 y
    const double a = c0 + c1 * y.lo, b = c0 + c1 * y.hi;
    return Iv{fmin(a, b), fmax(a, b)};
}
CPD_HD Iv iv_mul(Iv a, Iv b) {
    const double p1 = a.lo * b.lo, p2 = a.lo * b.hi, p3 = a.hi * b.lo, p4 = a.hi * b.hi;
    return Iv{fmin(fmin(p1, p2), fmin(p3, p4)), fmax(fmax(p1, p2), fmax(p3, p4))};
}
CPD_HD Iv iv_sqr(Iv a) {
    if (a.lo >= 0.0) return Iv{a.lo * a.lo, a.hi * a.hi};
    if (a.hi <= 0.0) return Iv{a.hi * a.hi, a.lo * a.lo};
    return Iv{0.0, fmax(a.lo * a.lo, a.hi * a.hi)};
}
CPD_HD Iv iv_sqrt0(Iv a) {  // sqrt over the non-negative part
    return Iv{sqrt(fmax(a.lo, 0.0)), sqrt(fmax(a.hi, 0.0))};
}
CPD_HD double iv_mag(Iv a) { return fmax(fabs(a.lo), fabs(a.hi)); }

// Whether a ray of the band can pass the acceptance test through lens root
// `far` (0: -b - sqrt(disc), 1: -b + sqrt(disc)); *valid is set when the root
// is certainly in front of the ray and inside the arc for every height of the
// band (so the trace never looks at the other root), *certain when, beyond
// that, every ray of the band certainly refracts, passes the acceptance test
// and meets the array's line between its ends, *never when the root is
// behind the ray or outside the arc for every height of the band.
CPD_HD bool band_root_may_hit(const LensBand64 &L, Iv y, Iv b, Iv sd, Iv disc, bool far, bool *valid,
                              bool *certain, bool *never) {
    *valid = false;
    *certain = false;
    *never = false;
    // root distance d = -b -/+ sqrt(disc)
    const Iv nb{-b.hi, -b.lo};
    const Iv d = far ? iv_add(nb, sd) : iv_sub(nb, sd);
    const double dm = 1e-9 * (iv_mag(d) + iv_mag(b) + iv_mag(sd));
    if (d.hi < -dm) return *never = true, false;  // behind the ray
    // bounds: |e + d dir|^2 <= max_dist_sq, e = e0 + e1 y
    const Iv qx = iv_add(iv_affine(L.e0x, L.e1x, y), iv_scale(d, L.dx));
    const Iv qy = iv_add(iv_affine(L.e0y, L.e1y, y), iv_scale(d, L.dy));
    const Iv qx2 = iv_sqr(qx), qy2 = iv_sqr(qy);
    const Iv s = iv_sub(iv_add(qx2, qy2), Iv{L.max_dist_sq, L.max_dist_sq});
    const double sm = 1e-9 * (L.max_dist_sq + qx2.hi + qy2.hi);
    if (s.lo > sm) return *never = true, false;  // outside the arc
    *valid = d.lo > dm && s.hi < -sm;
    // ci = -/+ sqrt(disc) / R, cr = sqrt(1 - r^2 (1 - ci^2)), k = r ci - cr
    const Iv cr2 = iv_add(Iv{L.lomr2, L.lomr2}, iv_scale(disc, L.lr2r));
    if (cr2.hi < -1e-9) return false;  // total internal reflection
    const Iv cr = iv_sqrt0(cr2);
    const Iv sdr = iv_scale(sd, L.inv_r);
    const Iv ci = far ? sdr : Iv{-sdr.hi, -sdr.lo};
    const Iv k = iv_sub(iv_scale(ci, L.r), cr);
    // acceptance cosine cd = -(r dir.dn + k n.dn), n.dn = (nd0 + nd1 y - d ddn) / R
    const Iv ndn = iv_scale(iv_sub(iv_affine(L.nd0, L.nd1, y), iv_scale(d, L.ddn)), L.inv_r);
    const Iv kn = iv_mul(k, ndn);
    const Iv cd{-(L.rddn + kn.hi), -(L.rddn + kn.lo)};
    const double cm = 1e-9 * (1.0 + fabs(L.rddn) + iv_mag(k) * (1.0 + iv_mag(ndn)));
    if (cd.hi <= L.min_ci - cm) return false;  // below the acceptance cosine
    // array ends (spectrometer.rs:291-295): rays accepted (cd > min_ci) meet the
    // array's line at w + v dist, w = q - P0, v = r dir + k n, n = (cm - q) / R,
    // dist = w.dn / cd; a position off [0, length] misses
    const double cd_lo = fmax(cd.lo, L.min_ci - cm);
    if (!(cd_lo > 1e-3)) return true;
    const Iv nx = iv_scale(iv_sub(Iv{L.cmx, L.cmx}, qx), L.inv_r);
    const Iv ny = iv_scale(iv_sub(Iv{L.cmy, L.cmy}, qy), L.inv_r);
    const Iv vx = iv_add(Iv{L.r * L.dx, L.r * L.dx}, iv_mul(k, nx));
    const Iv vy = iv_add(Iv{L.r * L.dy, L.r * L.dy}, iv_mul(k, ny));
    const Iv wx = iv_add(Iv{L.mpx, L.mpx}, qx), wy = iv_add(Iv{L.mpy, L.mpy}, qy);
    const Iv wn = iv_add(iv_scale(wx, L.dnx), iv_scale(wy, L.dny));
    const double r1 = wn.lo / cd_lo, r2 = wn.lo / cd.hi, r3 = wn.hi / cd_lo, r4 = wn.hi / cd.hi;
    const Iv dist{fmin(fmin(r1, r2), fmin(r3, r4)), fmax(fmax(r1, r2), fmax(r3, r4))};
    const Iv px = iv_add(wx, iv_mul(vx, dist)), py = iv_add(wy, iv_mul(vy, dist));
    const Iv pos = iv_add(iv_scale(px, L.ddx), iv_scale(py, L.ddy));
    const double pm = 1e-9 * (iv_mag(px) + iv_mag(py) + L.length);
    *certain = *valid && cr2.lo > 1e-9 && cd.lo > L.min_ci + cm && pos.lo > pm && pos.hi < L.length - pm;
    return !(pos.hi < -pm || pos.lo > L.length + pm);
}

// The band's class: kBandLost when band_may_hit would be false;
// kBandCertain when every ray of it certainly passes the planar faces, takes
// the near lens root and meets the array's line between its ends (so every
// narrower band of heights inside it may hit as well); kBandMixed otherwise.
enum { kBandLost = 0, kBandMixed = 1, kBandCertain = 2 };
CPD_HD int band_class(const LensBand64 &L, double ya, double yb) {
    const Iv y{fmin(ya, yb), fmax(ya, yb)};
    // every height outside the accepted interval: lost at the planar faces
    const double ym = 1e-9 * (fmax(fabs(ya), fabs(yb)) + L.path);
    if (y.hi < L.ylo - ym || y.lo > L.yhi + ym) return kBandLost;
    const bool inside = y.lo > L.ylo + ym && y.hi < L.yhi - ym;
    const Iv b = iv_affine(L.lb0, L.lb1, y);
    // c(y) = lc0 + y (lc1 + lc2 y)
    const Iv c = iv_add(Iv{L.lc0, L.lc0}, iv_mul(y, iv_affine(L.lc1, L.lc2, y)));
    Iv disc = iv_sub(iv_sqr(b), c);
    // The FP64 trace forms disc = (d.(o - c))^2 - |o - c|^2 + R^2 from terms
    // of size (R + |o - m|)^2, so its own disc is uncertain by a few ulp of
    // that; every decision downstream (roots, bounds, refraction, acceptance)
    // sees the widened interval.
    const Iv ex = iv_affine(L.e0x, L.e1x, y), ey = iv_affine(L.e0y, L.e1y, y);
    const double em = iv_mag(ex) + iv_mag(ey);
    const double tol = 1e-14 * (L.radius + em) * (L.radius + em) + 1e-9 * (iv_mag(b) * iv_mag(b) + iv_mag(c));
    if (disc.hi <= -tol) return kBandLost;
    disc = Iv{disc.lo - tol, disc.hi + tol};
    const Iv sd = iv_sqrt0(disc);
    // the trace takes the near root wherever it is valid (nearest valid root
    // first, geometry.rs:203-209); where it is valid for the whole band the far
    // root is never looked at
    bool near_valid, near_certain, near_never, far_valid, far_certain, far_never;
    if (band_root_may_hit(L, y, b, sd, disc, false, &near_valid, &near_certain, &near_never))
        return inside && disc.lo > 0.0 && near_certain ? kBandCertain : kBandMixed;
    if (near_valid) return kBandLost;
    // the far root, which the trace takes where the near one is not valid
    if (band_root_may_hit(L, y, b, sd, disc, true, &far_valid, &far_certain, &far_never))
        return inside && disc.lo > 0.0 && near_never && far_certain ? kBandCertain : kBandMixed;
    return kBandLost;
}

CPD_HD bool band_may_hit(const LensBand64 &L, double ya, double yb) { return band_class(L, ya, yb) != kBandLost; }

// trace64 of one ray with the planar faces' directions and transmittances
// taken from the wavelength's chain record (built by make_lambda_consts for a
// wavelength whose record has ok set): each face costs the ray only its hit
// point, computed with the operations of plane64 on the same values, and the
// lens and detector stages are trace64's, so the bin and transmittance are
// trace64's bit for bit.
CPD_HD uint32_t trace64_chain(const cpd_design &g, const Chain64 &ch, double y, double &t) {
#pragma clang fp contract(off)
    t = 0.0;
    double ox = 0.0, oy = y;
    for (uint32_t i = 0; i < g.n_glass; ++i) {
        const double *p = g.plane_point[i], *n = g.plane_normal[i];
        const double d = ((p[0] - ox) * n[0] + (p[1] - oy) * n[1]) / ch.ci[i];
        if (!(d >= 0.0)) return g.bin_count;
        const double hx = fma(ch.dx[i], d, ox), hy = fma(ch.dy[i], d, oy);
        if (!(0.0 <= hy && hy <= g.height)) return g.bin_count;
        ox = hx;
        oy = hy;
    }
    Ray64 ray{ox, oy, ch.ldx, ch.ldy, ch.ts, ch.tp};
    if (arc64(ray, g, ch.n_last, g.ar_coated != 0) != T_OK) return g.bin_count;
    uint32_t bin = g.bin_count;
    double tt = 0.0;
    if (detector64(ray, g, bin, tt) != T_OK) return g.bin_count;
    t = tt;
    return bin;
}

// One ray in FP64 from the wavelength's affine planar chain and lens
// constants (LensBand64: the quantities trace32 evaluates, without its FP32
// rounding) and the planar faces' transmittances (Chain64).  It returns false
// when a decision -- accepted height, lens discriminant, root sign, arc
// bounds, reflection, coating, acceptance, the array's ends or a bin edge --
// lies within a relative margin of kAffineRel of the quantities it is formed
// from, far above the FP64 rounding of these few operations and of the
// reference's own; the caller then re-traces the ray with trace64_chain, the
// reference's operation order.  Otherwise bin (bin_count on a miss) is
// trace64's and t agrees with trace64's to ~1e-12 relative (the reference's
// 1 - rs^2 loses digits near grazing exit, the identity used here does not).
static constexpr double kAffineRel = 1e-9;
CPD_HD bool trace64_affine(const LensBand64 &L, const Chain64 &ch, const cpd_design &g, double y, uint32_t &bin,
                           double &t) {
    bin = g.bin_count;
    t = 0.0;
    // planar chain: the faces' conditions are affine in y
    const double ym = kAffineRel * (fabs(y) + L.path);
    if (fabs(y - L.ylo) <= ym || fabs(y - L.yhi) <= ym) return false;
    if (!(y >= L.ylo && y <= L.yhi)) return true;
    // exit lens, relative to its chord midpoint m (as trace32)
    const double ex = fma(L.e1x, y, L.e0x), ey = fma(L.e1y, y, L.e0y);
    const double b = fma(L.lb1, y, L.lb0);
    const double c = fma(fma(L.lc2, y, L.lc1), y, L.lc0);
    const double disc = fma(b, b, -c);
    if (fabs(disc) <= kAffineRel * (b * b + fabs(c) + L.max_dist_sq)) return false;
    if (!(disc > 0.0)) return true;
    const double sd = sqrt(disc);
    const double big = -(b + copysign(sd, b));
    const double small = c / big;
    const double d1 = b >= 0.0 ? big : small, d2 = b >= 0.0 ? small : big;  // ascending
    const double dm = kAffineRel * (fabs(b) + sd + L.radius + L.path);
    double qx = 0.0, qy = 0.0;
    bool found = false;
    for (int i = 0; i < 2 && !found; ++i) {
        // nearest valid root first (geometry.rs:203-209)
        const double d = i == 0 ? d1 : d2;
        qx = fma(L.dx, d, ex);
        qy = fma(L.dy, d, ey);
        const double q2 = qx * qx + qy * qy;
        if (fabs(d) <= dm || fabs(q2 - L.max_dist_sq) <= kAffineRel * (q2 + L.max_dist_sq + L.path * L.path))
            return false;
        found = d >= 0.0 && q2 <= L.max_dist_sq;
    }
    if (!found) return true;
    // refraction out of the lens: n = (c - q) / R, n1 = r, n2 = 1
    const double nx = (L.cmx - qx) * L.inv_r, ny = (L.cmy - qy) * L.inv_r;
    const double ci = -(L.dx * nx + L.dy * ny);
    const double cr_sq = 1.0 - L.r * L.r * (1.0 - ci * ci);
    if (fabs(cr_sq) <= kAffineRel) return false;
    if (!(cr_sq >= 0.0)) return true;
    const double cr = sqrt(cr_sq);
    const double k = L.r * ci - cr;
    const double vx = L.dx * L.r + nx * k, vy = L.dy * L.r + ny * k;
    const bool ar = g.ar_coated != 0;
    if (ar && fabs(ci - 0.5) <= kAffineRel) return false;
    // detector array: its acceptance cosine, the distance along v and the
    // position from its start; errors grow with the lens's angular gain
    const double gain = (fabs(L.r * ci) > cr ? fabs(L.r * ci) : cr) / cr;
    const double cd = -(vx * L.dnx + vy * L.dny);
    if (fabs(cd - L.min_ci) <= kAffineRel * gain) return false;
    if (!(cd > L.min_ci)) return true;
    const double wx = qx + L.mpx, wy = qy + L.mpy;
    const double dist = (wx * L.dnx + wy * L.dny) / cd;
    const double pos = (wx + vx * dist) * L.ddx + (wy + vy * dist) * L.ddy;
    const double pm = kAffineRel * (fabs(wx) + fabs(wy) + fabs(dist) + L.path + L.radius) * gain / cd;
    if (fabs(pos) <= pm || fabs(pos - L.length) <= pm) return false;
    if (pos < 0.0 || L.length < pos) return true;
    // bin: euclid_div_rem (src/spectrometer/utils.rs:177-192), saturating u32
    const double v = pos - g.det_intercept;
    double q = trunc(v / g.det_slope);
    double rem = v - q * g.det_slope;
    if (rem < 0.0) {
        q = g.det_slope > 0.0 ? q - 1.0 : q + 1.0;
        rem = rem + fabs(g.det_slope);
    }
    const double em = pm + kAffineRel * (fabs(v) + fabs(g.det_slope));
    if (rem <= em || fabs(rem - fabs(g.det_slope)) <= em || fabs(rem - g.det_bin_size) <= em) return false;
    const uint32_t bq = !(q > 0.0) ? 0u : (q >= 4294967295.0 ? 4294967295u : (uint32_t)q);
    if (!(bq < g.bin_count && rem < g.det_bin_size)) return true;
    bin = bq;
    if (ar && ci > 0.5) {
        t = (ch.ts * 0.99 + ch.tp * 0.99) * 0.5;
    } else {
        // 1 - rs^2 = 4 r ci cr / (r ci + cr)^2, 1 - rp^2 = 4 r ci cr / (r cr + ci)^2
        const double fa = L.r * ci + cr, fb = L.r * cr + ci, rc4 = 4.0 * L.r * ci * cr;
        t = (ch.ts * (rc4 / (fa * fa)) + ch.tp * (rc4 / (fb * fb))) * 0.5;
    }
    return true;
}

// ---------------------------------------------------------------------------
// FP32 trace with error margins
// ---------------------------------------------------------------------------
#if defined(__HIP_DEVICE_COMPILE__)
// max of two non-NaN values in one instruction (fmaxf would first quiet
// signalling NaNs, one more instruction per operand)
CPD_HD float max2f(float a, float b) {
    float r;
    asm("v_max_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
CPD_HD float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
CPD_HD float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
// Whether any lane of the wavefront holds v (one lane on the host).
CPD_HD bool wave_any(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0; }
#elif defined(CPD_EMULATE_DEVICE_ROUNDING)
// A host model of the device's approximate reciprocal and square root (each
// within 1 ulp of the exact value): the correctly rounded result moved by the
// ulp a harness draws (tests/guard).
float cpd_emu_rcp(float x);
float cpd_emu_sqrt(float x);
inline float max2f(float a, float b) { return fmaxf(a, b); }
inline float fast_rcp(float x) { return cpd_emu_rcp(x); }
inline float fast_sqrt(float x) { return cpd_emu_sqrt(x); }
inline bool wave_any(bool v) { return v; }
#else
CPD_HD float max2f(float a, float b) { return fmaxf(a, b); }
CPD_HD float fast_rcp(float x) { return 1.0f / x; }
CPD_HD float fast_sqrt(float x) { return sqrtf(x); }
CPD_HD bool wave_any(bool v) { return v; }
#endif

// A float whose value does not matter (clang: an arbitrary but defined value,
// so no register is set for it).
#if defined(__clang__)
#define kAnyFloat __builtin_nondeterministic_value(0.f)
#define kAnyU32 __builtin_nondeterministic_value(0u)
#else
#define kAnyFloat 0.f
#define kAnyU32 0u
#endif

// Error budget of an FP32 length per unit of path travelled, and of an FP32
// cosine, before angular amplification.
static constexpr float kGuardRel = 8.0e-6f;
static constexpr float kGuardCos = 2.0e-5f;

// Algorithmic operations per stage a ray reaches (add/mul/div/sqrt = 1,
// fma = 2), after hoisting what the beam's rays share: beam height 2; the
// planar chain 4 (the affine hit point on the last face); lens 54 =
// intersection, bounds and normal 22 + refraction 32; detector 24; hit 1.
// The same stages without the planar-chain hoist cost 42 per planar face
// instead of 4 in total (per ray 85 + 42 P, the reference formulas with
// only per-design and per-wavelength invariants hoisted).
static constexpr uint32_t kOpsBeam = 2, kOpsChain = 4, kOpsLensHit = 22, kOpsLensRefract = 32, kOpsDetector = 24,
                          kOpsHit = 1;

struct Trace32Result {
    uint32_t bin;  // the ray's bin, with hit
    float t;       // its transmittance, with hit
    bool guard;    // the outcome is within the FP32 error margin of a decision
    bool hit;      // the ray lands in a bin
    uint32_t ops;  // algorithmic operations of the stages the ray reached (COUNT)
};

// The constants one ray reads: the hot subset of TraceDesign and
// LambdaConsts, small enough to live in scalar registers for the whole ray
// loop.
struct alignas(4) HotRecord {
    uint32_t nbin, ar_coated;
    float y_mean, beam_w, ylo, yhi, ymarg;
    float ymid, yhalf;  // the accepted heights' midpoint and half-width
    float disc_lim;  // margin of the lens discriminant for every accepted height
    float blim;      // margin of the lens's bounds test
    float coat_lim;  // margin of the coating's cosine test
    float cd_lim;    // margin of the acceptance cosine, times cos(theta_t), for every ray
    float lb0, lb1, lc0, lc1, lc2;
    float g0, g1, cmd2;  // arc bounds: |q|^2 - max_dist_sq = g0 + g1 y + cmd2 s at the root s
    float inv_r;
    float lr2r, lomr2, lr, nd0, nd1, ddn, rddn, min_ci;
    float wdn0, wdn1, wdn2, wdd0, wdd1, wdd2, ndd0, ndd1, ndd2, rdd;
    float num0k;         // kGuardRel A
    float dlen0, dlen1;  // the detector stage's length scale terms, times num0k
    float hc, half_len, slope, inv_slope, bin_size;
    float ts4, tp4;      // 4 tscale Ts, 4 tscale Tp through the planar faces
    float tcoat;         // t of a ray through the coating: 0.99 tscale (Ts + Tp) + toff
    float toff;
};

// tscale multiplies the planar faces' transmittances and toff is added, so
// that trace32's t = tscale (Ts + Tp) + toff (0.5 and 0: the mean the
// reference takes; tscale a power of two, so any other scale gives the same
// bits scaled).
CPD_HD HotRecord make_hot(const TraceDesign &d, const LambdaConsts &lc, float tscale = 0.5f, float toff = 0.f) {
    HotRecord h;
    h.nbin = d.nbin;
    h.ar_coated = d.ar_coated;
    h.y_mean = d.y_mean;
    h.beam_w = d.beam_w;
    h.ylo = lc.ylo;
    h.yhi = lc.yhi;
    // One height margin for the wavelength.  A ray within the per-height
    // margin kGuardRel (|y| + 1) of a bound y_b has |y| <= (|y_b| +
    // kGuardRel) / (1 - kGuardRel), so kGuardRel (max |y_b| + 1) /
    // (1 - kGuardRel), rounded up, covers every such ray.
    h.ymarg = kGuardRel * (fmaxf(fabsf(lc.ylo), fabsf(lc.yhi)) + 1.f) * (1.f + 2.f * kGuardRel) * 1.000001f;
    // ylo <= y <= yhi as |y - ymid| <= yhalf: the rounding of ymid, yhalf
    // and the two differences moves the bounds by a few ulps of max |y_b|,
    // far inside ymarg
    h.ymid = (float)(((double)lc.ylo + (double)lc.yhi) * 0.5);
    h.yhalf = (float)(((double)lc.yhi - (double)lc.ylo) * 0.5);
    // b(y) = lb0 + lb1 y is affine, so its largest magnitude over the accepted
    // heights bounds |b| for every ray the discriminant decides; the margin
    // 4 kGuardRel path A (|b| + lens_scale) at that bound covers each ray's.
    const float bmax = fmaxf(fabsf(fmaf(lc.lb1, lc.ylo, lc.lb0)), fabsf(fmaf(lc.lb1, lc.yhi, lc.lb0))) * 1.000001f;
    h.disc_lim = kGuardRel * lc.path * lc.amp * 4.f * (bmax + d.lens_scale) * 1.000001f;
    // the arc bounds' value 2 q.cm + const moves by 2 |cm| |dq| for a hit
    // point error dq of kGuardRel path A (bound_scale = 2 (|cm| + 1)), plus
    // 8 ulps of the magnitudes its FP32 constants carry
    h.blim = (kGuardRel * lc.amp * lc.path * d.bound_scale +
              4.8e-7f * (fabsf(lc.lg0) + fabsf(lc.lg1) * fmaxf(fabsf(lc.ylo), fabsf(lc.yhi)))) * 1.000001f;
    h.coat_lim = kGuardCos * lc.amp;
    // the acceptance cosine's margin kGuardCos A max(r ci, cr) / cr: with
    // ci, cr <= 1, A max(r, 1) bounds the numerator for every ray (FP32 ci
    // and cr can exceed 1 by ~1e-6, hence the extra 1e-4)
    h.cd_lim = kGuardCos * lc.amp * fmaxf(lc.lr, 1.f) * 1.0001f;
    h.lb0 = lc.lb0;
    h.lb1 = lc.lb1;
    h.lc0 = lc.lc0;
    h.lc1 = lc.lc1;
    h.lc2 = lc.lc2;
    h.g0 = lc.lg0;
    h.g1 = lc.lg1;
    h.cmd2 = lc.lcmd2;
    h.inv_r = d.inv_r;
    h.lr2r = lc.lr2r;
    h.lomr2 = lc.lomr2;
    h.lr = lc.lr;
    h.nd0 = lc.nd0;
    h.nd1 = lc.nd1;
    h.ddn = lc.ddn;
    h.rddn = lc.rddn;
    h.min_ci = d.min_ci;
    h.wdn0 = lc.wdn0;
    h.wdn1 = lc.wdn1;
    h.wdn2 = lc.wdn2;
    h.wdd0 = lc.wdd0;
    h.wdd1 = lc.wdd1;
    h.wdd2 = lc.wdd2;
    h.ndd0 = lc.ndd0;
    h.ndd1 = lc.ndd1;
    h.ndd2 = lc.ndd2;
    h.rdd = lc.rdd;
    h.num0k = kGuardRel * lc.amp;
    h.dlen0 = lc.dlen0 * h.num0k * 1.000001f;
    h.dlen1 = lc.dlen1 * h.num0k * 1.000001f;
    h.hc = lc.hc;
    h.half_len = d.half_len;
    h.slope = d.slope;
    h.inv_slope = d.inv_slope;
    h.bin_size = d.bin_size;
    h.ts4 = lc.ts * (4.f * tscale);
    h.tp4 = lc.tp * (4.f * tscale);
    h.tcoat = fmaf(lc.ts * tscale + lc.tp * tscale, 0.99f, toff);
    h.toff = toff;
    return h;
}

// Saturating conversion of an integral float to u32 (negative and NaN give
// 0, as the reference's `as u32` casts do).
#if defined(__HIP_DEVICE_COMPILE__)
CPD_HD uint32_t sat_u32(float x) {
    uint32_t r;
    asm("v_cvt_u32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}
#else
CPD_HD uint32_t sat_u32(float x) { return x > 0.f ? (x < 4294967296.f ? (uint32_t)x : 0xffffffffu) : 0u; }
#endif

// One ray at beam height y.  Decisions are kept as per-lane booleans and
// combined with bitwise operators, so the whole trace is straight-line code.
// The ray lands in bin `bin` with transmittance t when `hit`; guard marks a
// ray whose outcome is within the FP32 error margin of a decision.
template <bool COUNT>
CPD_HD Trace32Result trace32(const HotRecord &d, float y) {
    // bin and t are read only for a hit, which only the end of the trace
    // sets: the miss paths leave them without a value to materialise
    Trace32Result res{__builtin_bit_cast(uint32_t, y), y, false, false, 0};
    const bool ar = d.ar_coated != 0;
    // planar chain: the ray survives the planar faces iff ylo <= y <= yhi
    const float yo = fabsf(y - d.ymid) - d.yhalf;  // <= 0 inside, |yo| from the nearer bound
    bool alive = yo <= 0.f;
    bool guard = fabsf(yo) <= d.ymarg;
    if (COUNT) res.ops = kOpsBeam + kOpsChain + (alive ? kOpsLensHit : 0u);
    // wavefronts of banded rays often lie wholly outside the accepted heights
    if (!wave_any(alive | guard)) return res;
    // exit lens, relative to its chord midpoint m: e = o - m, c = m + cm.
    // |e + s dir - cm|^2 = R^2  <=>  s^2 + 2 b s + c = 0; the root of larger
    // magnitude is -(b + sign(b) sqrt(disc)) and the other one is c over it,
    // so neither root cancels.
    const float b = fmaf(d.lb1, y, d.lb0);
    const float c = fmaf(fmaf(d.lc2, y, d.lc1), y, d.lc0);
    const float disc = fmaf(b, b, -c);
    guard |= alive & (fabsf(disc) <= d.disc_lim);
    alive &= disc > 0.f;
    // a negative discriminant (a lane already lost) gives NaN, which only
    // reaches values of lost lanes: every later decision is masked by alive
    const float sd = fast_sqrt(disc);
    const float big = -(b + copysignf(sd, b));
    const float small = c * fast_rcp(big);
    const bool bpos = !signbit(b);
    const float d1 = bpos ? big : small, d2 = bpos ? small : big;
    // |q|^2 - max_dist_sq at q = e + s dir on the lens circle (affine in s)
    const float gy = fmaf(d.g1, y, d.g0);
    const float s1 = fmaf(d.cmd2, d1, gy);
    const float s2 = fmaf(d.cmd2, d2, gy);
    const bool ok1 = (d1 >= 0.f) & (s1 <= 0.f);
    const bool ok2 = (d2 >= 0.f) & (s2 <= 0.f);
    guard |= alive & ((fabsf(s1) <= d.blim) | (!ok1 & (fabsf(s2) <= d.blim)));
    alive &= ok1 | ok2;
    if (COUNT) res.ops += alive ? kOpsLensRefract : 0u;
    const float tq = ok1 ? d1 : d2;  // the hit point is q = e + tq dir
    // Incidence cosine -dir.n, n = (c - p) / R = (cm - q) / R: it equals
    // (b + tq) / R, which is -sqrt(disc) / R at the root -b - sqrt(disc) and
    // +sqrt(disc) / R at the other; its error, that of sqrt(disc) over R, is
    // the one the normal carries through q.
    const float sdr = sd * d.inv_r;
    const float ci = ok1 ? -sdr : sdr;
    // 1 - r^2 (1 - ci^2) with ci^2 = disc / R^2
    const float cr_sq = fmaf(d.lr2r, disc, d.lomr2);
    const float cr = fast_sqrt(cr_sq);  // NaN only for lanes lost to reflection
    const float rci = d.lr * ci;
    const float k = rci - cr;
    guard |= alive & (fabsf(cr_sq) <= kGuardCos);
    alive &= cr_sq >= 0.f;
    if (ar) guard |= alive & (fabsf(ci - 0.5f) <= d.coat_lim);
    const float den = cr;
    // detector array: w = o - P0 = (m - P0) + q
    if (COUNT) res.ops += alive ? kOpsDetector : 0u;
    {
        // acceptance cosine -v.dn of v = r dir + k n: -(r dir.dn + k n.dn),
        // n.dn = ((cm - e).dn - tq dir.dn) / R = nd0 + nd1 y - tq ddn (the
        // constants carry the 1 / R)
        const float ndn = fmaf(-tq, d.ddn, fmaf(d.nd1, y, d.nd0));
        const float cd = -fmaf(k, ndn, d.rddn);
        guard |= alive & (fabsf(cd - d.min_ci) * den <= d.cd_lim);
        alive &= cd > d.min_ci;
        // most wavefronts lose every ray here (the acceptance cosine): the
        // rest of the stage runs on the lanes with a ray left or near a
        // decision (a divergent branch, which a wavefront with none skips)
        res.t = cr;  // (any value: a miss's t is not read)
        if (alive | guard) {
            // the ray meets the array's line at w + dist v, dist = -w.dn / v.dn,
            // at u = (w + dist v).dd - intercept along it; every product with
            // dn or dd is affine in y and tq (per-wavelength constants)
            const float wdn = fmaf(d.wdn2, tq, fmaf(d.wdn1, y, d.wdn0));
            const float wdd = fmaf(d.wdd2, tq, fmaf(d.wdd1, y, d.wdd0));
            const float ndd = fmaf(d.ndd2, tq, fmaf(d.ndd1, y, d.ndd0));
            const float vdd = fmaf(k, ndd, d.rdd);
            const float dist = wdn * fast_rcp(cd);
            const float u = fmaf(dist, vdd, wdd);
            // position errors: kGuardRel A per unit of length (the path to the
            // lens, |tq| and |w| through the planar faces' bound, |dist|), times
            // the lens's angular gain max(r ci, cr) / cr (the 1 / cr multiplies
            // the margins instead)
            const float plim = fmaf(fabsf(dist), d.num0k, fmaf(d.dlen1, fabsf(tq), d.dlen0)) * max2f(rci, cr);
            const float apos = u - d.hc;  // position minus half the length
            guard |= alive & (fabsf(fabsf(apos) - d.half_len) * den <= plim);
            alive &= fabsf(apos) <= d.half_len;
            const float qb = floorf(u * d.inv_slope);
            const float rem = fmaf(-qb, d.slope, u);
            const float edge = fminf(fminf(fabsf(rem), fabsf(d.slope - rem)), fabsf(rem - d.bin_size));
            guard |= alive & (edge * den <= plim);
            // the quotient is cast to u32 saturating (bin_index, spectrometer.rs:
            // 291-303): a position before the intercept has a negative quotient
            // and counts in bin 0 when its remainder is below the bin size
            const uint32_t bin = sat_u32(qb);
            alive &= (bin < d.nbin) & (rem < d.bin_size);
            // Fresnel: 1 - rs^2 = 4 r ci cr / (r ci + cr)^2 and 1 - rp^2 =
            // 4 r ci cr / (r cr + ci)^2 (rs = (r ci - cr) / (r ci + cr), rp =
            // (r cr - ci) / (r cr + ci), the reference's amplitudes divided
            // through by n2), one reciprocal
            const float fa = rci + cr, fb = fmaf(d.lr, cr, ci);
            const float inv = fast_rcp(fa * fb);
            const float ia = fb * inv, ib = fa * inv;  // 1 / fa, 1 / fb
            const float sum = fmaf(d.ts4 * ia, ia, (d.tp4 * ib) * ib);
            const bool coat = ar & (ci > 0.5f);
            res.bin = bin;
            res.t = coat ? d.tcoat : fmaf(rci * cr, sum, d.toff);
            res.hit = alive;
            if (COUNT) res.ops += alive ? kOpsHit : 0u;
        }
    }
    res.guard = guard;
    return res;
}

}  // namespace cpd
