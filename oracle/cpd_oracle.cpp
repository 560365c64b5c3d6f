// cpd_oracle.cpp -- CPU restatement of the reference compound-prism fitness path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
// __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, and only as
// the checker or as the timed CPU baseline.  The product library never links,
// loads or calls it.
//
// It restates, file by file, the FP64/FP32 CPU path of
// jdtatz/compound-prism-designer for the 2D spectrometer
// (Spectrometer<F, SimdVector<F,2>, UniformDistribution, GaussianBeam,
// FocusingPlanerCompoundPrism>):
//   src/fitness.rs                       fitness, mutual_information, p_dets_l_wavelength
//   src/spectrometer/welford.rs          Welford
//   src/spectrometer/qrng.rs             R1 additive recurrence
//   src/spectrometer/erf.rs              norminv (Shaw-Luu-Brickman)
//   src/spectrometer/distribution.rs     UniformDistribution
//   src/spectrometer/glasscat.rs         Glass::calc_n
//   src/spectrometer/geometry.rs         planes, arc, trapezoids, detector fit
//   src/spectrometer/ray.rs              Ray, refract, CompoundPrism
//   src/spectrometer/spectrometer.rs     GaussianBeam, LinearDetectorArray, Spectrometer
//   src/spectrometer/vector.rs           2D vector ops
//   src/spectrometer/utils.rs            euclid_div_rem, plog2p
//   src/spectrometer/kernel.rs:195-224   propagation_test_kernel (CPU emulation)
//   src/spectrometer/kernel.rs:89-193    the GPU estimator's per-warp fold (FP64)
//   src/cuda_fitness.rs:347-423          cuda_fitness's loop over seeds (FP64)
//
// Floating-point fidelity: compiled with -ffp-contract=off, so every mul and
// add rounds separately, and std::fma is used exactly where the reference
// calls mul_add (Horner in glasscat.rs:20 and erf.rs:14, Vector::mul_add in
// vector.rs:497-503).  Dot products are x*x' + y*y' (vector.rs:290-292).
//
// Pinning (what checks each part) is listed in the header of oracle/oracle.py.

#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "../include/cpd.h"

namespace ora {

// ---------------------------------------------------------------------------
// src/spectrometer/vector.rs: Vector<2> for SimdVector<T, 2>
// ---------------------------------------------------------------------------
template <class F>
struct V2 {
    F x, y;
};
template <class F> static inline V2<F> vadd(V2<F> a, V2<F> b) { return {a.x + b.x, a.y + b.y}; }
template <class F> static inline V2<F> vsub(V2<F> a, V2<F> b) { return {a.x - b.x, a.y - b.y}; }
template <class F> static inline V2<F> vneg(V2<F> a) { return {-a.x, -a.y}; }
template <class F> static inline V2<F> vmul(V2<F> a, F s) { return {a.x * s, a.y * s}; }
template <class F> static inline V2<F> vdiv(V2<F> a, F s) { return {a.x / s, a.y / s}; }
// vector.rs:290-292 dot = hadamard_product(..).reduce_sum()
template <class F> static inline F dot(V2<F> a, V2<F> b) { return a.x * b.x + a.y * b.y; }
template <class F> static inline F norm_sq(V2<F> a) { return dot(a, a); }
// vector.rs:497-503 Simd::mul_add (fused)
template <class F> static inline V2<F> mul_add(V2<F> a, F b, V2<F> c) {
    return {std::fma(a.x, b, c.x), std::fma(a.y, b, c.y)};
}
// vector.rs:336-343 angled_xy
template <class F> static inline V2<F> angled_xy(F theta) { return {std::cos(theta), std::sin(theta)}; }
// vector.rs:355-380
template <class F> static inline V2<F> rot_90_xy(V2<F> v) { return {-v.y, v.x}; }
template <class F> static inline V2<F> rot_180_xy(V2<F> v) { return {-v.x, -v.y}; }
template <class F> static inline V2<F> rot_90_ccw_xy(V2<F> v) { return {v.y, -v.x}; }
// vector.rs:386-409
template <class F> static inline F sec_xy(V2<F> v, F hyp) { return hyp / v.x; }
template <class F> static inline F sin_xy(V2<F> v, F hyp) { return v.y / hyp; }
template <class F> static inline F tan_xy(V2<F> v) { return v.y / v.x; }

static inline bool is_sign_positive(double v) { return !std::signbit(v); }
static inline bool is_sign_positive(float v) { return !std::signbit(v); }

// ---------------------------------------------------------------------------
// src/spectrometer/utils.rs
// ---------------------------------------------------------------------------
// utils.rs:177-192 euclid_div_rem
template <class F>
static inline void euclid_div_rem(F self, F rhs, F &q_out, F &r_out) {
    F q = std::trunc(self / rhs);
    F r = self - q * rhs;
    if (r < F(0)) {
        q_out = rhs > F(0) ? q - F(1) : q + F(1);
        r_out = r + std::fabs(rhs);
    } else {
        q_out = q;
        r_out = r;
    }
}
// utils.rs:196-202 plog2p
template <class F> static inline F plog2p(F p) { return p > F(0) ? p * std::log2(p) : F(0); }
// Rust `f as u32`: saturating, NaN -> 0 (utils.rs:74 `v as $to`)
template <class F> static inline uint32_t sat_u32(F v) {
    if (!(v > F(0))) return 0u;  // negative, zero and NaN
    if (v >= F(4294967295.0)) return 4294967295u;
    return (uint32_t)v;
}
// Rust `fract`: self - self.trunc()
template <class F> static inline F fract(F v) { return v - std::trunc(v); }

// ---------------------------------------------------------------------------
// src/spectrometer/welford.rs
// ---------------------------------------------------------------------------
template <class F>
struct Welford {
    F count = F(0), mean = F(0), m2 = F(0);
    // welford.rs:26-32
    void next_sample(F x) {
        count += F(1);
        F delta = x - mean;
        mean += delta / count;
        F delta2 = x - mean;
        m2 += delta * delta2;
    }
    // welford.rs:34-42 (m2 uses the updated mean)
    void skip(F new_count) {
        if (count < new_count) {
            F zero_count = new_count - count;
            F c = new_count;
            mean = (count * mean) / c;
            m2 += mean * mean * count * zero_count / c;
            count = c;
        }
    }
    // welford.rs:55-58
    bool sem_le_error_threshold(F error_squared) const { return m2 < error_squared * (count * (count - F(1))); }
    // welford.rs:76-85 Add
    static Welford add(const Welford &a, const Welford &b) {
        Welford r;
        r.count = a.count + b.count;
        F delta = b.mean - a.mean;
        r.mean = (a.count * a.mean + b.count * b.mean) / r.count;
        r.m2 = a.m2 + b.m2 + delta * delta * a.count * b.count / r.count;
        return r;
    }
};

// ---------------------------------------------------------------------------
// src/spectrometer/qrng.rs:156-158, 203-255: R1 sequence, alpha = 1/phi
// ---------------------------------------------------------------------------
static const double PHI_1 = 1.61803398874989484820458683436563;
static const double ALPHA_1 = 1.0 / PHI_1;

// Iterative form of Qrng<[F;1]>::next (qrng.rs:217-219, 251-254).
template <class F>
struct QrngIter {
    F state;
    F alpha;
    explicit QrngIter(F seed) : state(seed), alpha((F)ALPHA_1) {}
    F next() {
        state = fract(state + alpha);
        return state;
    }
};

// Exact form u_k = frac(0.5 + k * alpha) in 64-bit fixed point.  This is the
// sample-index definition the GPU uses; it agrees with the FP64 iterative
// sequence to ~1e-12.
static const uint64_t ALPHA_FX = 0x9E3779B97F4A7C16ull;  // round(2^64 / phi)
static inline uint64_t qmc_fx(uint64_t seed_fx, uint64_t k) { return seed_fx + k * ALPHA_FX; }
static inline double fx_to_f64(uint64_t s) { return (double)(s >> 11) * 0x1.0p-53; }
static inline float fx_to_f32(uint64_t s) { return (float)(s >> 40) * 0x1.0p-24f; }

template <class F> struct FxConv;
template <> struct FxConv<double> { static double conv(uint64_t s) { return fx_to_f64(s); } };
template <> struct FxConv<float> { static float conv(uint64_t s) { return fx_to_f32(s); } };

enum QrngKind { QRNG_ITERATIVE = 0, QRNG_FIXED_POINT = 1 };

template <class F>
struct Qrng {
    int kind;
    QrngIter<F> it;
    uint64_t k = 0;
    Qrng(int kind_) : kind(kind_), it(F(0.5)) {}
    F next() {
        if (kind == QRNG_ITERATIVE) return it.next();
        ++k;
        return FxConv<F>::conv(qmc_fx(0x8000000000000000ull, k));
    }
};

// ---------------------------------------------------------------------------
// src/spectrometer/erf.rs:8-73 norminv (Shaw, Luu & Brickman, arXiv:0901.0638)
// ---------------------------------------------------------------------------
static const double SHAW_P[8] = {
    4.3304513840364031401e-8, 0.000036313870818023761224, 0.0034424140686962222423,
    0.085838533424158257377,  0.73176759583280610539,     2.3884158540184385711,
    3.0333178251950406994,    1.2533141359896652729,
};
static const double SHAW_Q[8] = {
    1.3141263119543315917e-6, 0.00030617264753008793976, 0.014494272424798068406,
    0.2168237095066675527,    1.2356513216582148689,     2.9373357991677046357,
    2.9202373175993672857,    1.0,
};
// erf.rs:8-16 polynomial: coeff.reduce(|s, c| s.mul_add(z, c))
template <class F>
static inline F polynomial8(F z, const double *coeff) {
    F s = (F)coeff[0];
    for (int i = 1; i < 8; ++i) s = std::fma(s, z, (F)coeff[i]);
    return s;
}
template <class F>
static inline F norminv(F x) {
    F u = x > (F)0.5 ? F(1) - x : x;
    F v = -std::log(F(2) * u);
    F p = polynomial8(v, SHAW_P);
    F q = polynomial8(v, SHAW_Q);
    return std::copysign(v * p / q, x - (F)0.5);
}

// ---------------------------------------------------------------------------
// Spectrometer model (geometry.rs, ray.rs, spectrometer.rs)
// ---------------------------------------------------------------------------
template <class F>
struct Plane {  // BoundedHyperSurface<HyperPlane, PrismBounds> (geometry.rs:335-336)
    V2<F> point, normal;
    F height;
};
template <class F>
struct Arc {  // CurvedPlane (geometry.rs:337-342)
    V2<F> center;
    F radius;
    V2<F> midpt;
    F max_dist_sq;
};
template <class F>
struct Detector {  // LinearDetectorArray (spectrometer.rs:158-180)
    uint32_t bin_count;
    F bin_size, slope, intercept, min_ci, length;
    V2<F> normal, position, direction;
};
template <class F>
struct Spec {
    int n_glass;
    bool ar_coated;
    F glass[CPD_MAX_PRISMS][CPD_GLASS_COEFFS];
    Plane<F> planes[CPD_MAX_PRISMS];
    Arc<F> arc;
    Detector<F> det;
    F beam_width, y_mean;
    F wl_lo, wl_hi;
};

template <class F>
struct Ray {  // ray.rs:245-255
    V2<F> origin, direction;
    F ts, tp;
};

enum Err {
    OK = CPD_OK,
    NO_SURFACE_INTERSECTION = CPD_ERR_NO_SURFACE_INTERSECTION,
    TOTAL_INTERNAL_REFLECTION = CPD_ERR_TOTAL_INTERNAL_REFLECTION,
    ANGULAR_RESPONSE_TOO_WEAK = CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK,
};

// glasscat.rs:17-22 calc_n: Horner with mul_add
template <class F>
static inline F calc_n(const F *c, F w) {
    F s = c[0];
    for (int i = 1; i < CPD_GLASS_COEFFS; ++i) s = std::fma(s, w, c[i]);
    return s;
}

// geometry.rs:124-145 HyperPlane::intersection + PrismBounds<2>::in_bounds (62-67)
template <class F>
static inline bool plane_intersection(const Plane<F> &pl, const Ray<F> &ray, F &dist) {
    F ci = dot(ray.direction, pl.normal);
    if (ci == F(0)) return false;
    F d = dot(vsub(pl.point, ray.origin), pl.normal) / ci;
    if (d >= F(0)) {
        V2<F> p = mul_add(ray.direction, d, ray.origin);
        if (F(0) <= p.y && p.y <= pl.height) {
            dist = d;
            return true;
        }
    }
    return false;
}

// geometry.rs:161-180 spherical_ray_intersection + 190-216 SphericalLikeSurface
// + RadialBounds::in_bounds (85-91)
template <class F>
static inline bool arc_intersection(const Arc<F> &arc, const Ray<F> &ray, F &dist, V2<F> &normal) {
    V2<F> delta = vsub(ray.origin, arc.center);
    F ud = dot(ray.direction, delta);
    F radius_squared = arc.radius * arc.radius;
    F disc = ud * ud - norm_sq(delta) + radius_squared;
    if (!(disc > F(0))) return false;  // tangent and none both return None
    F sd = std::sqrt(disc);
    F d1 = -ud - sd;
    F d2 = -ud + sd;
    V2<F> p1 = mul_add(ray.direction, d1, ray.origin);
    V2<F> p2 = mul_add(ray.direction, d2, ray.origin);
    F d;
    V2<F> p;
    if (is_sign_positive(d1) && norm_sq(vsub(p1, arc.midpt)) <= arc.max_dist_sq) {
        d = d1;
        p = p1;
    } else if (is_sign_positive(d2) && norm_sq(vsub(p2, arc.midpt)) <= arc.max_dist_sq) {
        d = d2;
        p = p2;
    } else {
        return false;
    }
    normal = vdiv(vsub(arc.center, p), arc.radius);
    dist = d;
    return true;
}

// ray.rs:309-367 Ray::refract
template <class F>
static inline Err refract(Ray<F> &ray, V2<F> normal, F n1, F n2, bool ar_coated) {
    F ci = -dot(ray.direction, normal);
    F r = n1 / n2;
    F cr_sq = F(1) - r * r * (F(1) - ci * ci);
    if (cr_sq < F(0)) return TOTAL_INTERNAL_REFLECTION;
    F cr = std::sqrt(cr_sq);
    V2<F> v = vadd(vmul(ray.direction, r), vmul(normal, r * ci - cr));
    F ts, tp;
    if (ar_coated && ci > (F)0.5) {
        ts = ray.ts * (F)0.99;
        tp = ray.tp * (F)0.99;
    } else {
        F rs = (n1 * ci - n2 * cr) / (n1 * ci + n2 * cr);
        F rp = (n1 * cr - n2 * ci) / (n1 * cr + n2 * ci);
        rs = rs * rs;
        rp = rp * rp;
        ts = ray.ts * (F(1) - rs);
        tp = ray.tp * (F(1) - rp);
    }
    ray.direction = v;
    ray.ts = ts;
    ray.tp = tp;
    return OK;
}

// ray.rs:281-294 translate + 370-382 surface_propagate (planes)
template <class F>
static inline Err surface_propagate_plane(Ray<F> &ray, const Plane<F> &pl, F n1, F n2, bool ar) {
    F d;
    if (!plane_intersection(pl, ray, d)) return NO_SURFACE_INTERSECTION;
    ray.origin = mul_add(ray.direction, d, ray.origin);
    return refract(ray, pl.normal, n1, n2, ar);
}
template <class F>
static inline Err surface_propagate_arc(Ray<F> &ray, const Arc<F> &arc, F n1, F n2, bool ar) {
    F d;
    V2<F> normal;
    if (!arc_intersection(arc, ray, d, normal)) return NO_SURFACE_INTERSECTION;
    ray.origin = mul_add(ray.direction, d, ray.origin);
    return refract(ray, normal, n1, n2, ar);
}

// ray.rs:485-499 GenericCompoundPrism::propagate (initial_surface 411-418,
// final_surface 426-433)
template <class F>
static inline Err prism_propagate(const Spec<F> &s, Ray<F> &ray, F wavelength) {
    F n1 = calc_n(s.glass[0], wavelength);
    Err e = surface_propagate_plane(ray, s.planes[0], F(1), n1, s.ar_coated);
    if (e != OK) return e;
    for (int i = 1; i < s.n_glass; ++i) {
        F n2 = calc_n(s.glass[i], wavelength);
        e = surface_propagate_plane(ray, s.planes[i], n1, n2, s.ar_coated);
        if (e != OK) return e;
        n1 = n2;
    }
    return surface_propagate_arc(ray, s.arc, n1, F(1), s.ar_coated);
}

// ray.rs:296-299
template <class F> static inline F average_transmittance(const Ray<F> &r) { return (r.ts + r.tp) * (F)0.5; }

// spectrometer.rs:265-277 LinearDetectorArray::intersection, 291-303 bin_index,
// ray.rs:390-402 intersect_detector_array
template <class F>
static inline Err intersect_detector(const Detector<F> &det, const Ray<F> &ray, V2<F> &p, uint32_t &bin, F &t) {
    F ci = -dot(ray.direction, det.normal);
    if (ci <= det.min_ci) return ANGULAR_RESPONSE_TOO_WEAK;
    F d = dot(vsub(ray.origin, det.position), det.normal) / ci;
    p = mul_add(ray.direction, d, ray.origin);
    F pos = dot(vsub(p, det.position), det.direction);
    if (pos < F(0) || det.length < pos) return NO_SURFACE_INTERSECTION;
    F q, rem;
    euclid_div_rem(pos - det.intercept, det.slope, q, rem);
    uint32_t b = sat_u32(q);
    if (!(b < det.bin_count && rem < det.bin_size)) return NO_SURFACE_INTERSECTION;
    bin = b;
    t = average_transmittance(ray);
    return OK;
}

// ray.rs:272-279 Ray::new_from_start
template <class F> static inline Ray<F> ray_from_start(F y) { return Ray<F>{{F(0), y}, {F(1), F(0)}, F(1), F(1)}; }

// spectrometer.rs:421-427 Spectrometer::propagate
template <class F>
static inline Err spec_propagate(const Spec<F> &s, Ray<F> ray, F wavelength, uint32_t &bin, F &t) {
    Err e = prism_propagate(s, ray, wavelength);
    if (e != OK) return e;
    V2<F> p;
    return intersect_detector(s.det, ray, p, bin, t);
}

// spectrometer.rs:26-30 GaussianBeam 2D inverse_cdf
template <class F> static inline F beam_y(const Spec<F> &s, F u) { return s.y_mean - s.beam_width * norminv(u); }
// distribution.rs:32-34 UniformDistribution::inverse_cdf (separate mul and add)
template <class F> static inline F sample_wavelength(const Spec<F> &s, F u) { return s.wl_lo + (s.wl_hi - s.wl_lo) * u; }

// spectrometer.rs:305-308 mid_pt + 429-437 size_and_deviation
template <class F>
static inline void size_and_deviation(const Spec<F> &s, F &size, F &deviation) {
    V2<F> mid = mul_add(s.det.direction, s.det.length * (F)0.5, s.det.position);
    F dx = mid.x;
    F dy = mid.y - s.y_mean;
    V2<F> dv = {dx, dy};
    size = std::sqrt(norm_sq(dv));
    deviation = std::fabs(sin_xy(dv, size));
}

// ---------------------------------------------------------------------------
// Construction (geometry.rs:430-489, 225-243, 280-315; ray.rs:137-183;
// spectrometer.rs:183-219, 322-364; geometry.rs:495-542)
// ---------------------------------------------------------------------------
template <class F>
struct HyperPlane {
    V2<F> point, normal;
};

// geometry.rs:430-489 create_joined_trapezoids; planes[0..P) are the entrance
// and interior faces, *last is the exit face's hyperplane.  Lengths follow
// utils.rs:5-12 array_prepend: plane i (i = 1..P) uses lengths[i - 1].
template <class F>
static void create_joined_trapezoids(F height, const F *angles, const F *lengths, int P, HyperPlane<F> *planes,
                                     HyperPlane<F> *last) {
    F h2 = height * (F)0.5;
    V2<F> normal = rot_180_xy(angled_xy(angles[0]));
    planes[0] = HyperPlane<F>{{std::fabs(tan_xy(normal)) * h2, h2}, normal};
    bool prev_sign = is_sign_positive(normal.y);
    F d1 = planes[0].point.x;
    F mx = planes[0].point.x;
    for (int i = 1; i <= P; ++i) {
        V2<F> nrm = rot_180_xy(angled_xy(angles[i]));
        bool sign = is_sign_positive(nrm.y);
        F d2 = std::fabs(tan_xy(nrm)) * h2;
        F sep_dist = lengths[i - 1] + (prev_sign != sign ? d1 + d2 : std::fabs(d1 - d2));
        prev_sign = sign;
        d1 = d2;
        mx += sep_dist;
        HyperPlane<F> hp{{mx, h2}, nrm};
        if (i < P)
            planes[i] = hp;
        else
            *last = hp;
    }
}

// geometry.rs:280-315 CurvedPlane::from_hyperplane
template <class F>
static Arc<F> curved_plane(HyperPlane<F> hp, F k, F height) {
    F chord = std::fabs(sec_xy(hp.normal, height));
    F radius = chord * (F)0.5 / std::fabs(k);
    F apothem = std::sqrt(radius * radius - chord * chord * (F)0.25);
    F sagitta = radius - apothem;
    Arc<F> a;
    a.center = mul_add(hp.normal, std::copysign(apothem, k), hp.point);
    a.radius = radius;
    a.midpt = mul_add(hp.normal, -std::copysign(sagitta, k), hp.point);
    a.max_dist_sq = sagitta * sagitta + chord * chord * (F)0.25;
    return a;
}

// ray.rs:137-183 CompoundPrism::new (2D: Plane entrance, Plane interior,
// CurvedPlane exit)
template <class F>
static void build_prism(const cpd_params &p, Spec<F> &s) {
    int P = (int)p.n_glass;
    s.n_glass = P;
    s.ar_coated = p.ar_coated != 0;
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) s.glass[i][j] = (F)p.glass[i][j];
    F angles[CPD_MAX_PRISMS + 1], lengths[CPD_MAX_PRISMS];
    for (int i = 0; i <= P; ++i) angles[i] = (F)p.angles[i];
    for (int i = 0; i < P; ++i) lengths[i] = (F)p.lengths[i];
    HyperPlane<F> hps[CPD_MAX_PRISMS]{}, last{};
    F height = (F)p.height;
    create_joined_trapezoids(height, angles, lengths, P, hps, &last);
    for (int i = 0; i < P; ++i) s.planes[i] = Plane<F>{hps[i].point, hps[i].normal, height};
    s.arc = curved_plane(last, (F)p.curvature, height);
    s.beam_width = (F)p.beam_width;
    s.y_mean = (F)p.y_mean;
    s.wl_lo = (F)p.wavelength_lo;
    s.wl_hi = (F)p.wavelength_hi;
}

// spectrometer.rs:183-219 LinearDetectorArray::new; min_ci = cos(max_incident_angle)
// as PyDetectorArray does (pylib.rs:906-920).
template <class F>
static void build_detector(const cpd_params &p, Spec<F> &s) {
    Detector<F> &d = s.det;
    d.bin_count = p.bin_count;
    d.bin_size = (F)p.bin_size;
    d.slope = (F)p.linear_slope;
    d.intercept = (F)p.linear_intercept;
    d.min_ci = std::cos((F)p.max_incident_angle);
    d.length = (F)p.det_length;
    d.normal = rot_180_xy(angled_xy((F)p.det_angle));
    d.direction = p.flipped ? rot_90_xy(d.normal) : rot_90_ccw_xy(d.normal);
    d.position = V2<F>{(F)p.det_position[0], (F)p.det_position[1]};
}

// geometry.rs:495-542 fit_ray_difference_surface
template <class F>
static bool fit_ray_difference_surface(const Ray<F> &lower, const Ray<F> &upper, F d_length, F d_angle,
                                       V2<F> &pos, bool &flipped) {
    V2<F> spec_dir = rot_90_xy(angled_xy(d_angle));
    V2<F> spec = vmul(spec_dir, d_length);
    F a = upper.direction.x, b = -lower.direction.x, c = upper.direction.y, d = -lower.direction.y;
    F det = a * d - b * c;
    if (det == F(0)) return false;
    F i10 = -c / det, i11 = a / det;
    V2<F> temp = vadd(vsub(spec, upper.origin), lower.origin);
    F d2 = i10 * temp.x + i11 * temp.y;
    if (d2 > F(0)) {
        pos = mul_add(lower.direction, d2, lower.origin);
        flipped = false;
        return true;
    }
    temp = vadd(vsub(vneg(spec), upper.origin), lower.origin);
    d2 = i10 * temp.x + i11 * temp.y;
    if (d2 < F(0)) return false;
    pos = mul_add(lower.direction, d2, lower.origin);
    flipped = true;
    return true;
}

// spectrometer.rs:322-364 detector_array_positioning
template <class F>
static Err position_detector(const Spec<F> &s, F length, F angle, F acceptance, V2<F> &pos, bool &flipped) {
    Ray<F> ray = ray_from_start(s.y_mean);
    F uh = std::fmax(F(1) - acceptance, F(0)) * (F)0.5;
    F wmin = sample_wavelength(s, uh);
    F wmax = sample_wavelength(s, F(1) - uh);
    Ray<F> lower = ray, upper = ray;
    Err e = prism_propagate(s, lower, wmin);
    if (e != OK) return e;
    e = prism_propagate(s, upper, wmax);
    if (e != OK) return e;
    if (average_transmittance(lower) <= (F)1e-3 || average_transmittance(upper) <= (F)1e-3)
        return ANGULAR_RESPONSE_TOO_WEAK;
    if (!fit_ray_difference_surface(lower, upper, length, angle, pos, flipped)) return NO_SURFACE_INTERSECTION;
    return OK;
}

template <class F>
static void spec_from_design(const cpd_design &d, Spec<F> &s) {
    s.n_glass = (int)d.n_glass;
    s.ar_coated = d.ar_coated != 0;
    for (int i = 0; i < CPD_MAX_PRISMS; ++i)
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) s.glass[i][j] = (F)d.glass[i][j];
    for (int i = 0; i < CPD_MAX_PRISMS; ++i) {
        s.planes[i].point = V2<F>{(F)d.plane_point[i][0], (F)d.plane_point[i][1]};
        s.planes[i].normal = V2<F>{(F)d.plane_normal[i][0], (F)d.plane_normal[i][1]};
        s.planes[i].height = (F)d.height;
    }
    s.arc.center = V2<F>{(F)d.arc_center[0], (F)d.arc_center[1]};
    s.arc.radius = (F)d.arc_radius;
    s.arc.midpt = V2<F>{(F)d.arc_midpt[0], (F)d.arc_midpt[1]};
    s.arc.max_dist_sq = (F)d.arc_max_dist_sq;
    s.det.bin_count = d.bin_count;
    s.det.bin_size = (F)d.det_bin_size;
    s.det.slope = (F)d.det_slope;
    s.det.intercept = (F)d.det_intercept;
    s.det.min_ci = (F)d.det_min_ci;
    s.det.length = (F)d.det_length;
    s.det.normal = V2<F>{(F)d.det_normal[0], (F)d.det_normal[1]};
    s.det.position = V2<F>{(F)d.det_position[0], (F)d.det_position[1]};
    s.det.direction = V2<F>{(F)d.det_direction[0], (F)d.det_direction[1]};
    s.beam_width = (F)d.beam_width;
    s.y_mean = (F)d.y_mean;
    s.wl_lo = (F)d.wavelength_lo;
    s.wl_hi = (F)d.wavelength_hi;
}

static void design_from_spec(const Spec<double> &s, cpd_design &d) {
    std::memset(&d, 0, sizeof(d));
    d.n_glass = (uint32_t)s.n_glass;
    d.bin_count = s.det.bin_count;
    d.ar_coated = s.ar_coated ? 1u : 0u;
    for (int i = 0; i < s.n_glass; ++i) {
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) d.glass[i][j] = s.glass[i][j];
        d.plane_point[i][0] = s.planes[i].point.x;
        d.plane_point[i][1] = s.planes[i].point.y;
        d.plane_normal[i][0] = s.planes[i].normal.x;
        d.plane_normal[i][1] = s.planes[i].normal.y;
    }
    d.height = s.planes[0].height;
    d.arc_center[0] = s.arc.center.x;
    d.arc_center[1] = s.arc.center.y;
    d.arc_radius = s.arc.radius;
    d.arc_midpt[0] = s.arc.midpt.x;
    d.arc_midpt[1] = s.arc.midpt.y;
    d.arc_max_dist_sq = s.arc.max_dist_sq;
    d.det_position[0] = s.det.position.x;
    d.det_position[1] = s.det.position.y;
    d.det_direction[0] = s.det.direction.x;
    d.det_direction[1] = s.det.direction.y;
    d.det_normal[0] = s.det.normal.x;
    d.det_normal[1] = s.det.normal.y;
    d.det_min_ci = s.det.min_ci;
    d.det_length = s.det.length;
    d.det_slope = s.det.slope;
    d.det_intercept = s.det.intercept;
    d.det_bin_size = s.det.bin_size;
    d.beam_width = s.beam_width;
    d.y_mean = s.y_mean;
    d.wavelength_lo = s.wl_lo;
    d.wavelength_hi = s.wl_hi;
}

// ---------------------------------------------------------------------------
// src/fitness.rs
// ---------------------------------------------------------------------------
struct Counters {
    uint64_t rays = 0;
    uint64_t wavelengths = 0;
};

// fitness.rs:5-34 vector_quasi_monte_carlo_integration, specialised to
// p_dets_l_wavelength (48-69).  max_err = 0 disables the early exit.
template <class F>
static void p_dets_l_wavelength(const Spec<F> &s, F wavelength, uint64_t max_m, F max_err, int qrng_kind,
                                std::vector<Welford<F>> &stats, Counters *ctr) {
    F max_err_squared = max_err * max_err;
    F count = F(0);
    const uint32_t nbin = s.det.bin_count;
    stats.assign(nbin, Welford<F>());
    Qrng<F> qrng(qrng_kind);
    for (uint64_t i = 0; i < max_m; ++i) {
        F u = qrng.next();
        count += F(1);
        if (ctr) ctr->rays++;
        uint32_t idx;
        F t;
        if (spec_propagate(s, ray_from_start(beam_y(s, u)), wavelength, idx, t) == OK) {
            stats[idx].skip(count);
            stats[idx].next_sample(t);
            if (stats[idx].sem_le_error_threshold(max_err_squared)) break;
        }
    }
    for (auto &w : stats) w.skip(count);
}

// fitness.rs:79-123 mutual_information
template <class F>
static F mutual_information(const Spec<F> &s, uint64_t max_n, uint64_t max_m, bool adaptive, int qrng_kind,
                            Counters *ctr) {
    const uint32_t nbin = s.det.bin_count;
    std::vector<Welford<F>> p_dets(nbin), row;
    Welford<F> h_det_l_w;
    Qrng<F> qrng(qrng_kind);
    const F inner_err = adaptive ? (F)5e-3 : F(0);
    const F outer_err_sqr = adaptive ? (F)(5e-3 * 5e-3) : F(0);
    for (uint64_t j = 0; j < max_n; ++j) {
        F u = qrng.next();
        F w = sample_wavelength(s, u);
        if (ctr) ctr->wavelengths++;
        p_dets_l_wavelength(s, w, max_m, inner_err, qrng_kind, row, ctr);
        F h_det_l_ws = F(0);
        for (uint32_t b = 0; b < nbin; ++b) {
            F p = row[b].mean;
            p_dets[b].next_sample(p);
            h_det_l_ws += plog2p(p);
        }
        h_det_l_w.next_sample(h_det_l_ws);
        bool all = true;
        for (uint32_t b = 0; b < nbin; ++b) all = all && p_dets[b].sem_le_error_threshold(outer_err_sqr);
        if (all) break;
    }
    F h_det = F(0);
    for (uint32_t b = 0; b < nbin; ++b) h_det = h_det + plog2p(p_dets[b].mean);
    return h_det_l_w.mean - h_det;
}

// ---------------------------------------------------------------------------
// src/spectrometer/kernel.rs:89-193 (one warp = one wavelength row) and
// src/cuda_fitness.rs:347-423 (the host loop over seeds), in FP64.  The
// reference folds in F = f32 with fast-math; this restatement folds in FP64,
// with each 32-lane warp sum taken in the butterfly order 16, 8, 4, 2, 1.
// ---------------------------------------------------------------------------
static inline uint64_t seed_fx_of(double seed) {
    double f = seed - std::floor(seed);
    if (!(f >= 0.0 && f < 1.0)) f = 0.0;
    return (uint64_t)(f * 0x1.0p+64);
}

static inline double butterfly_sum32(const double *v) {
    double a[32], b[32];
    for (int l = 0; l < 32; ++l) a[l] = v[l];
    for (int off = 16; off > 0; off >>= 1) {
        for (int l = 0; l < 32; ++l) b[l] = a[l] + a[l ^ off];
        for (int l = 0; l < 32; ++l) a[l] = b[l];
    }
    return a[0];
}

// kernel.rs:89-193 for one wavelength: returns the rays folded (count).
static uint64_t gpu_row(const Spec<double> &s, double wavelength, uint64_t seed_fx, uint32_t max_eval,
                        std::vector<Welford<double>> &stats) {
    const uint32_t nbin = s.det.bin_count;
    stats.assign(nbin, Welford<double>());
    const double e2 = 5e-3 * 5e-3;
    const uint32_t max_evals = max_eval - max_eval % 32;
    double count = 0.0;
    for (uint32_t i = 0; 32 * i < max_evals; ++i) {
        count += 32.0;
        uint32_t bin[32];
        double t[32];
        bool pending[32];
        bool any = false;
        for (uint32_t l = 0; l < 32; ++l) {
            const uint64_t k = l + 32 * (uint64_t)(i + 1);
            const double u = fx_to_f64(seed_fx + k * ALPHA_FX);  // frac(seed + k alpha)
            bin[l] = nbin;
            t[l] = 0.0;
            uint32_t b;
            double tt;
            if (spec_propagate(s, ray_from_start(beam_y(s, u)), wavelength, b, tt) == OK) {
                bin[l] = b;
                t[l] = tt;
            }
            pending[l] = bin[l] < nbin;
            any = any || pending[l];
        }
        bool finished = any;
        for (;;) {
            uint32_t bmin = 0xffffffffu;
            for (uint32_t l = 0; l < 32; ++l)
                if (pending[l] && bin[l] < bmin) bmin = bin[l];
            if (bmin == 0xffffffffu) break;
            double tb[32], dev[32];
            for (uint32_t l = 0; l < 32; ++l) tb[l] = bin[l] == bmin ? t[l] : 0.0;
            const double mean = butterfly_sum32(tb) / 32.0;
            for (uint32_t l = 0; l < 32; ++l) dev[l] = (tb[l] - mean) * (tb[l] - mean);
            Welford<double> w;
            w.count = 32.0;
            w.mean = mean;
            w.m2 = butterfly_sum32(dev);
            stats[bmin] = Welford<double>::add(stats[bmin], w);
            finished = finished && stats[bmin].sem_le_error_threshold(e2);
            for (uint32_t l = 0; l < 32; ++l)
                if (bin[l] == bmin) pending[l] = false;
        }
        if (finished) break;
    }
    for (auto &w : stats) w.skip(count);
    return (uint64_t)count;
}

}  // namespace ora

using namespace ora;

// ---------------------------------------------------------------------------
// C ABI of the oracle (ctypes, tests only)
// ---------------------------------------------------------------------------
extern "C" {

int ora_build_design(const cpd_params *p, cpd_design *out) {
    if (!p || !out || p->n_glass < 1 || p->n_glass > CPD_MAX_PRISMS) return CPD_ERR_INVALID_ARGUMENT;
    Spec<double> s{};
    build_prism(*p, s);
    build_detector(*p, s);
    design_from_spec(s, *out);
    return CPD_OK;
}

int ora_position_detector_array(const cpd_params *p, double acceptance, double pos_out[2], uint32_t *flipped_out) {
    if (!p || p->n_glass < 1 || p->n_glass > CPD_MAX_PRISMS) return CPD_ERR_INVALID_ARGUMENT;
    Spec<double> s{};
    build_prism(*p, s);
    V2<double> pos;
    bool flipped = false;
    Err e = position_detector(s, p->det_length, p->det_angle, acceptance, pos, flipped);
    if (e != OK) return e;
    pos_out[0] = pos.x;
    pos_out[1] = pos.y;
    *flipped_out = flipped ? 1u : 0u;
    return CPD_OK;
}

// Construction and positioning entirely in FP32, as the reference's FP32
// tests and bench build their design (cuda_fitness.rs:473-531,
// benches/benchmark.rs:87-145).  The result is widened into cpd_design.
int ora_build_design_f32(const cpd_params *p, double acceptance, cpd_design *out) {
    if (!p || !out || p->n_glass < 1 || p->n_glass > CPD_MAX_PRISMS) return CPD_ERR_INVALID_ARGUMENT;
    Spec<float> s{};
    build_prism(*p, s);
    cpd_params q = *p;
    if (acceptance > 0) {
        V2<float> pos;
        bool flipped = false;
        Err e = position_detector(s, (float)p->det_length, (float)p->det_angle, (float)acceptance, pos, flipped);
        if (e != OK) return e;
        q.det_position[0] = pos.x;
        q.det_position[1] = pos.y;
        q.flipped = flipped;
    }
    build_detector(q, s);
    Spec<double> w{};
    w.n_glass = s.n_glass;
    w.ar_coated = s.ar_coated;
    for (int i = 0; i < s.n_glass; ++i) {
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) w.glass[i][j] = s.glass[i][j];
        w.planes[i] = Plane<double>{{s.planes[i].point.x, s.planes[i].point.y},
                                    {s.planes[i].normal.x, s.planes[i].normal.y},
                                    s.planes[i].height};
    }
    w.arc = Arc<double>{{s.arc.center.x, s.arc.center.y}, s.arc.radius, {s.arc.midpt.x, s.arc.midpt.y}, s.arc.max_dist_sq};
    w.det.bin_count = s.det.bin_count;
    w.det.bin_size = s.det.bin_size;
    w.det.slope = s.det.slope;
    w.det.intercept = s.det.intercept;
    w.det.min_ci = s.det.min_ci;
    w.det.length = s.det.length;
    w.det.normal = {s.det.normal.x, s.det.normal.y};
    w.det.position = {s.det.position.x, s.det.position.y};
    w.det.direction = {s.det.direction.x, s.det.direction.y};
    w.beam_width = s.beam_width;
    w.y_mean = s.y_mean;
    w.wl_lo = s.wl_lo;
    w.wl_hi = s.wl_hi;
    design_from_spec(w, *out);
    return CPD_OK;
}

int ora_size_and_deviation(const cpd_design *d, double *size, double *deviation) {
    Spec<double> s{};
    spec_from_design(*d, s);
    size_and_deviation(s, *size, *deviation);
    return CPD_OK;
}

double ora_norminv_f64(double u) { return norminv<double>(u); }
void ora_norminv_many(size_t n, const double *u, double *out) {
    for (size_t i = 0; i < n; ++i) out[i] = norminv<double>(u[i]);
}
float ora_norminv_f32(float u) { return norminv<float>(u); }

// u_k of the R1 sequence: iterative (the reference) or fixed point (exact).
void ora_qmc_f64(uint64_t n, int kind, double *out) {
    Qrng<double> q(kind);
    for (uint64_t i = 0; i < n; ++i) out[i] = q.next();
}
void ora_qmc_f32(uint64_t n, int kind, float *out) {
    Qrng<float> q(kind);
    for (uint64_t i = 0; i < n; ++i) out[i] = q.next();
}

// One ray from (0, y) at the given wavelength; returns the RayTraceError code.
int ora_propagate_f64(const cpd_design *d, double wavelength, double y, uint32_t *bin, double *t) {
    Spec<double> s{};
    spec_from_design(*d, s);
    uint32_t b = d->bin_count;
    double tt = 0;
    Err e = spec_propagate(s, ray_from_start(y), wavelength, b, tt);
    *bin = e == OK ? b : d->bin_count;
    *t = e == OK ? tt : 0.0;
    return e;
}
int ora_propagate_f32(const cpd_design *d, float wavelength, float y, uint32_t *bin, float *t) {
    Spec<float> s{};
    spec_from_design(*d, s);
    uint32_t b = d->bin_count;
    float tt = 0;
    Err e = spec_propagate(s, ray_from_start(y), wavelength, b, tt);
    *bin = e == OK ? b : d->bin_count;
    *t = e == OK ? tt : 0.0f;
    return e;
}

// Batched single rays (wavelength, y) in FP64 or FP32 (precision 64 / 32).
void ora_propagate_many(const cpd_design *d, int precision, size_t n, const double *wavelength, const double *y,
                        uint32_t *bin, double *t) {
    if (precision == 32) {
        Spec<float> s{};
        spec_from_design(*d, s);
        for (size_t i = 0; i < n; ++i) {
            uint32_t b = d->bin_count;
            float tt = 0;
            Err e = spec_propagate(s, ray_from_start((float)y[i]), (float)wavelength[i], b, tt);
            bin[i] = e == OK ? b : d->bin_count;
            t[i] = e == OK ? tt : 0.0;
        }
    } else {
        Spec<double> s{};
        spec_from_design(*d, s);
        for (size_t i = 0; i < n; ++i) {
            uint32_t b = d->bin_count;
            double tt = 0;
            Err e = spec_propagate(s, ray_from_start(y[i]), wavelength[i], b, tt);
            bin[i] = e == OK ? b : d->bin_count;
            t[i] = e == OK ? tt : 0.0;
        }
    }
}

// kernel.rs:195-224 propagation_test_kernel in FP32 on the CPU (the
// reference's MockGpu, cuda_fitness.rs:549-604).
void ora_propagation_test_f32(const cpd_design *d, size_t n, const float *wavelength_cdf, const float *ray_cdf,
                              uint32_t *bin, float *t) {
    Spec<float> s{};
    spec_from_design(*d, s);
    for (size_t i = 0; i < n; ++i) {
        float w = sample_wavelength(s, wavelength_cdf[i]);
        uint32_t b = d->bin_count;
        float tt = 0;
        Err e = spec_propagate(s, ray_from_start(beam_y(s, ray_cdf[i])), w, b, tt);
        bin[i] = e == OK ? b : d->bin_count;
        t[i] = e == OK ? tt : 0.0f;
    }
}

// spectrometer.rs:439-464 propagation_path (pyo3 ray_trace, pylib.rs:1060-1083):
// rows of [origin.x, origin.y, dir.x, dir.y]; returns the row count (<= P + 3).
size_t ora_ray_trace_f64(const cpd_design *d, double wavelength, double y, double *out) {
    Spec<double> s{};
    spec_from_design(*d, s);
    Ray<double> ray = ray_from_start(y);
    size_t k = 0;
    auto emit = [&](V2<double> o, V2<double> dir) {
        out[4 * k + 0] = o.x;
        out[4 * k + 1] = o.y;
        out[4 * k + 2] = dir.x;
        out[4 * k + 3] = dir.y;
        ++k;
    };
    emit(ray.origin, ray.direction);
    // ray.rs:501-541 propagate_trace: stops at the first error.
    double n1 = calc_n(s.glass[0], wavelength);
    Ray<double> r0 = ray;
    bool ok = surface_propagate_plane(r0, s.planes[0], 1.0, n1, s.ar_coated) == OK;
    if (ok) {
        ray = r0;
        emit(ray.origin, ray.direction);
    }
    for (int i = 1; ok && i < s.n_glass; ++i) {
        double n2 = calc_n(s.glass[i], wavelength);
        Ray<double> r2 = ray;
        ok = surface_propagate_plane(r2, s.planes[i], n1, n2, s.ar_coated) == OK;
        if (ok) {
            ray = r2;
            emit(ray.origin, ray.direction);
        }
        n1 = n2;
    }
    if (ok) {
        Ray<double> r2 = ray;
        ok = surface_propagate_arc(r2, s.arc, n1, 1.0, s.ar_coated) == OK;
        if (ok) {
            ray = r2;
            emit(ray.origin, ray.direction);
        }
    }
    // The detector step uses the last successfully propagated ray
    // (spectrometer.rs:453-459), even after an error.
    V2<double> p;
    uint32_t b;
    double t;
    if (intersect_detector(s.det, ray, p, b, t) == OK) emit(p, ray.direction);
    return k;
}

// fitness.rs:48-69 p_dets_l_wavelength.  precision 64/32; adaptive 0/1;
// qrng 0 = iterative (reference), 1 = fixed point.
int ora_p_dets_l_wavelength(const cpd_design *d, double wavelength, uint64_t max_m, int precision, int adaptive,
                            int qrng_kind, double *out, uint64_t *rays) {
    Counters c;
    if (precision == 32) {
        Spec<float> s{};
        spec_from_design(*d, s);
        std::vector<Welford<float>> row;
        p_dets_l_wavelength<float>(s, (float)wavelength, max_m, adaptive ? 5e-3f : 0.0f, qrng_kind, row, &c);
        for (uint32_t b = 0; b < d->bin_count; ++b) out[b] = row[b].mean;
    } else {
        Spec<double> s{};
        spec_from_design(*d, s);
        std::vector<Welford<double>> row;
        p_dets_l_wavelength<double>(s, wavelength, max_m, adaptive ? 5e-3 : 0.0, qrng_kind, row, &c);
        for (uint32_t b = 0; b < d->bin_count; ++b) out[b] = row[b].mean;
    }
    if (rays) *rays = c.rays;
    return CPD_OK;
}

// fitness.rs:130-147 fitness -> out[3] = {size, info, deviation}.
int ora_fitness(const cpd_design *d, uint64_t max_n, uint64_t max_m, int precision, int adaptive, int qrng_kind,
                double *out, uint64_t *rays) {
    Counters c;
    if (precision == 32) {
        Spec<float> s{};
        spec_from_design(*d, s);
        float size, dev;
        size_and_deviation(s, size, dev);
        float info = mutual_information<float>(s, max_n, max_m, adaptive != 0, qrng_kind, &c);
        out[0] = size;
        out[1] = info;
        out[2] = dev;
    } else {
        Spec<double> s{};
        spec_from_design(*d, s);
        double size, dev;
        size_and_deviation(s, size, dev);
        double info = mutual_information<double>(s, max_n, max_m, adaptive != 0, qrng_kind, &c);
        out[0] = size;
        out[1] = info;
        out[2] = dev;
    }
    if (rays) *rays = c.rays;
    return CPD_OK;
}

// fitness with both counters: out[3], counts[2] = {wavelength rows, rays}.
int ora_fitness_counts(const cpd_design *d, uint64_t max_n, uint64_t max_m, int precision, int adaptive,
                       int qrng_kind, double *out, uint64_t *counts) {
    Counters c;
    if (precision == 32) {
        Spec<float> s{};
        spec_from_design(*d, s);
        float size, dev;
        size_and_deviation(s, size, dev);
        out[1] = mutual_information<float>(s, max_n, max_m, adaptive != 0, qrng_kind, &c);
        out[0] = size;
        out[2] = dev;
    } else {
        Spec<double> s{};
        spec_from_design(*d, s);
        double size, dev;
        size_and_deviation(s, size, dev);
        out[1] = mutual_information<double>(s, max_n, max_m, adaptive != 0, qrng_kind, &c);
        out[0] = size;
        out[2] = dev;
    }
    counts[0] = c.wavelengths;
    counts[1] = c.rays;
    return CPD_OK;
}

// n designs over `threads` host threads, one design per thread at a time
// (designer.py:190-192 ThreadPool + GIL release at pylib.rs:1019).
// out is [n][3]; returns the total number of rays traced.
uint64_t ora_fitness_many(const cpd_design *designs, const uint32_t *max_n, const uint32_t *max_m, size_t n,
                          int precision, int adaptive, int qrng_kind, int threads, double *out) {
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> total{0};
    auto worker = [&]() {
        for (;;) {
            size_t i = next.fetch_add(1);
            if (i >= n) break;
            uint64_t r = 0;
            ora_fitness(&designs[i], max_n[i], max_m[i], precision, adaptive, qrng_kind, &out[3 * i], &r);
            total.fetch_add(r);
        }
    };
    if (threads <= 1) {
        worker();
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t) pool.emplace_back(worker);
        for (auto &th : pool) th.join();
    }
    return total.load();
}

}  // extern "C"

extern "C" {
// One row of the GPU estimator (kernel.rs:89-193): p(d|wavelength) into
// out[bin_count]; returns the rays folded.
uint64_t ora_gpu_row(const cpd_design *d, double wavelength, double seed, uint32_t max_eval, double *out) {
    Spec<double> s{};
    spec_from_design(*d, s);
    std::vector<Welford<double>> row;
    const uint64_t rays = gpu_row(s, wavelength, seed_fx_of(seed), max_eval, row);
    for (uint32_t b = 0; b < s.det.bin_count; ++b) out[b] = row[b].mean;
    return rays;
}

// cuda_fitness (cuda_fitness.rs:347-423) over the given seeds: out[3];
// *converged = 0 where the reference returns None (out then holds the
// estimate after the last seed); counts[2] = {rows, rays}.
int ora_gpu_fitness(const cpd_design *d, const double *seeds, size_t n_seeds, uint32_t max_n, uint32_t nwarp,
                    uint32_t max_eval, double *out, int *converged, uint64_t *counts) {
    if (!d || !out || nwarp == 0) return CPD_ERR_INVALID_ARGUMENT;
    Spec<double> s{};
    spec_from_design(*d, s);
    const uint32_t nbin = s.det.bin_count;
    const uint32_t n_rows = (max_n / nwarp) * nwarp;
    const double e2 = 5e-3 * 5e-3;
    std::vector<Welford<double>> p_dets(nbin), row;
    Welford<double> h_det_l_w;
    uint64_t rows = 0, rays = 0;
    int conv = 0;
    for (size_t si = 0; si < n_seeds && !conv; ++si) {
        const uint64_t seed_fx = seed_fx_of(seeds[si]);
        for (uint32_t id = 0; id < n_rows; ++id) {
            const double u = fx_to_f64(seed_fx + (uint64_t)id * ALPHA_FX);
            rays += gpu_row(s, sample_wavelength(s, u), seed_fx, max_eval, row);
            double h = 0.0;
            for (uint32_t b = 0; b < nbin; ++b) {
                p_dets[b].next_sample(row[b].mean);
                h += plog2p(row[b].mean);
            }
            h_det_l_w.next_sample(h);
            ++rows;
        }
        bool all = rows > 0;
        for (uint32_t b = 0; b < nbin; ++b) all = all && p_dets[b].sem_le_error_threshold(e2);
        conv = all ? 1 : 0;
    }
    double h_det = 0.0;
    for (uint32_t b = 0; b < nbin; ++b) h_det = h_det + plog2p(p_dets[b].mean);
    double size, dev;
    size_and_deviation(s, size, dev);
    out[0] = size;
    out[1] = h_det_l_w.mean - h_det;
    out[2] = dev;
    if (converged) *converged = conv;
    if (counts) {
        counts[0] = rows;
        counts[1] = rays;
    }
    return CPD_OK;
}

// Layout checks for the ctypes mirrors of cpd.h.
size_t ora_sizeof_params(void) { return sizeof(cpd_params); }
size_t ora_sizeof_design(void) { return sizeof(cpd_design); }
}

// ===========================================================================
// 3D variant (Dim3, src/pylib.rs:636-640): ToricLens entrance and exit faces,
// planar interior faces with height and width bounds, FiberBeam or 3D
// GaussianBeam, the same LinearDetectorArray; FP64 as cpu_fitness runs it.
//   src/spectrometer/toric.rs:41-141     sympy_soln (quartic closed form)
//   src/spectrometer/toric.rs:143-236    ToricSurface::intersection
//   src/spectrometer/toric.rs:240-291    ToricLens::from_hyperplane
//   src/spectrometer/cbrt.rs:49-68       approx_cbrt (Moroz et al., Algorithm 6)
//   src/spectrometer/geometry.rs:69-77   PrismBounds<3>; 94-110 SandwichBounds
//   src/spectrometer/spectrometer.rs:32-46, 70-149  3D beams
//   src/spectrometer/distribution.rs:58-67          UniformDiscDistribution
//   src/spectrometer/qrng.rs:160-255     Qrng<[F; 2]>, Qrng<[F; 4]>
//   src/spectrometer/vector.rs:518-560   Vector<3> for SimdVector<T, 4>: dot is
//                                        ((x x' + y y') + z z'), mul_add fused
// ===========================================================================
namespace ora3 {

using ora::Welford;
using ora::calc_n;
using ora::euclid_div_rem;
using ora::plog2p;
using ora::sat_u32;
using ora::fract;

struct V3 {
    double x, y, z;
};
static inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 mul(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
static inline V3 div(V3 a, double s) { return {a.x / s, a.y / s, a.z / s}; }
static inline double dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline V3 mul_add(V3 a, double b, V3 c) { return {std::fma(a.x, b, c.x), std::fma(a.y, b, c.y), std::fma(a.z, b, c.z)}; }
static inline V3 xy(double x, double y) { return {x, y, 0.0}; }

// cbrt.rs:8-11 bit_magic on the f32 rounding of x; cbrt.rs:49-68 approx_cbrt
static inline double approx_cbrt(double x) {
    const double k1 = (double)1.752319676f, k2 = (double)1.2509524245f, k3 = (double)0.5093818292f;
    const double one_third = (double)0.333333333333f;
    const bool s = !std::signbit(x);
    x = std::fabs(x);
    const float xf = (float)x;
    uint32_t bits;
    std::memcpy(&bits, &xf, 4);
    const uint32_t mb = 0x548c2b4bu - bits / 3u;
    float yf;
    std::memcpy(&yf, &mb, 4);
    double y = (double)yf;
    double c = x * y * y * y;
    y = y * (k1 - c * (k2 - k3 * c));
    const double d = x * y * y;
    c = 1.0 - d * y;
    const double res = d * (1.0 + one_third * c);
    return s ? res : -res;
}

// toric.rs:41-141 the four roots of t^4 + b t^3 + c t^2 + d t + e, ascending;
// CBRT is approx_cbrt as the reference calls it (std::cbrt for the check of
// the algebra alone)
static inline double exact_cbrt(double x) { return std::cbrt(x); }
template <double (*CBRT)(double) = approx_cbrt>
static inline void sympy_soln(double r_2, double R_2, double dp_magn_sq, double dp_v, double dp_n, double n_v,
                              double out[4]) {
    const double x0 = R_2 * dp_v;
    const double x1 = dp_magn_sq * dp_v;
    const double x2 = dp_v * r_2;
    const double x3 = dp_n * n_v;
    const double x4 = R_2 * x3;
    const double x5 = n_v * n_v;
    const double x6 = 2.0 * R_2;
    const double x7 = R_2 - dp_magn_sq + r_2 - x5 * x6;
    const double x8 = 8.0 * dp_v * x7 - 8.0 * x0 + 8.0 * x1 - 8.0 * x2 + 16.0 * x4;
    const double x9 = 2.0 * r_2;
    const double x10 = dp_v * dp_v;
    const double x11 = R_2 * x5;
    const double x12 = 2.0 * dp_magn_sq - 2.0 * x10 + 4.0 * x11 - x6 - x9;
    const double x13 = x12 * x12 * x12;
    const double x14 = 4.0 * dp_v;
    const double x15i = -4.0 * x0 + 4.0 * x1 + x14 * x7 - 4.0 * x2 + 8.0 * x4;
    const double x15 = x15i * x15i;
    const double x16 = R_2 * R_2;
    const double x17 = dp_magn_sq * dp_magn_sq;
    const double x18 = r_2 * r_2;
    const double x19 = dp_magn_sq * x6;
    const double x20 = r_2 * x6;
    const double x21 = dp_magn_sq * x9;
    const double x22 = 4.0 * R_2 * (dp_n * dp_n);
    const double x23 =
        x14 * (-x0 + x1 + x14 * (R_2 / 8.0 - dp_magn_sq / 8.0 + r_2 / 8.0 - x10 / 16.0 - x11 / 4.0) - x2 + x3 * x6);
    const double x24 = x16 + x17 + x18 - x19 - x20 - x21 + x22 - x23;
    const double x25 = x12 * x24;
    const double x26 = -x13 / 108.0 - x15 / 8.0 + x25 / 3.0;
    const double x27 = 2.0 * CBRT(x26);
    const double x28 = 8.0 * R_2 / 3.0;
    const double x29 = 4.0 * R_2 / 3.0 - 4.0 * dp_magn_sq / 3.0 + 4.0 * r_2 / 3.0 + 4.0 * x10 / 3.0 - x28 * x5;
    const double x30 = std::sqrt(-x27 + x29);
    const double x31 = x8 / x30;
    const double x32 = -8.0 * dp_magn_sq / 3.0 + 8.0 * r_2 / 3.0 + 8.0 * x10 / 3.0 - 16.0 * x11 / 3.0 + x28;
    const double x33 = x27 + x32;
    const double x34 = std::sqrt(x31 + x33) / 2.0;
    const double x35 = -dp_v;
    const double x36 = x30 / 2.0;
    const double x37 = x35 - x36;
    const double x38 = (x12 * x12) / 12.0;
    const bool x39 = x24 + x38 == 0.0;
    const double x40 = -x16 - x17 - x18 + x19 + x20 + x21 - x22 + x23 - x38;
    const double x41 = CBRT(x13 / 216.0 + x15 / 16.0 - x25 / 6.0 +
                                   std::sqrt((x26 * x26) / 4.0 + (x40 * x40 * x40) / 27.0));
    const double x42 = 2.0 * x41;
    const double x43 = 2.0 * x40 / (3.0 * x41);
    const double x44 = std::sqrt(x29 + x42 - x43);
    const double x45 = x8 / x44;
    const double x46 = x32 - x42 + x43;
    const double x47 = std::sqrt(x45 + x46) / 2.0;
    const double x48 = x44 / 2.0;
    const double x49 = x35 - x48;
    const double x50 = std::sqrt(-x31 + x33) / 2.0;
    const double x51 = x35 + x36;
    const double x52 = std::sqrt(-x45 + x46) / 2.0;
    const double x53 = x35 + x48;
    out[3] = x39 ? x34 + x37 : x47 + x49;
    out[2] = x39 ? x50 + x51 : x52 + x53;
    out[1] = x39 ? -x34 + x37 : -x47 + x49;
    out[0] = x39 ? -x50 + x51 : -x52 + x53;
}

struct Torus {  // ToricSurface + SandwichBounds
    V3 center, tn;
    double R, r;
    V3 bc, bn;
    double bh;
};

// geometry.rs:104-109 SandwichBounds::in_bounds
static inline bool sandwich_in(const Torus &T, V3 p) {
    const double dp = dot(T.bn, sub(p, T.bc));
    return 0.0 <= dp && dp <= T.bh;
}

// toric.rs:143-236 ToricSurface::intersection: the first root in r0..r3 that
// is non-negative with its point in bounds; the normal points from the hit to
// the tube's centre circle.
template <double (*CBRT)(double) = approx_cbrt>
static inline bool torus_intersection(const Torus &T, V3 o, V3 d, double &dist, V3 &normal) {
    const V3 dp = sub(o, T.center);
    const double dp_magn_sq = dot(dp, dp);
    const double dp_v = dot(dp, d);
    const double dp_n = dot(dp, T.tn);
    const double n_v = dot(T.tn, d);
    const double R_2 = T.R * T.R;
    const double r_2 = T.r * T.r;
    double roots[4];
    sympy_soln<CBRT>(r_2, R_2, dp_magn_sq, dp_v, dp_n, n_v, roots);
    int found = -1;
    V3 p{};
    for (int i = 0; i < 4 && found < 0; ++i) {
        if (roots[i] < 0.0) continue;
        const V3 q = mul_add(d, roots[i], o);
        if (sandwich_in(T, q)) {
            found = i;
            p = q;
        }
    }
    if (found < 0) return false;
    // vector.rs:686-692 oproj(a, b) = a - b (a.b); normalize = v / |v|
    const V3 u = sub(p, T.center);
    const V3 op = sub(u, mul(T.tn, dot(u, T.tn)));
    const V3 unit = div(op, std::sqrt(dot(op, op)));
    const V3 t_center = mul_add(unit, T.R, T.center);
    normal = div(sub(t_center, p), T.r);
    dist = roots[found];
    return true;
}

struct Plane3 {
    V3 point, normal;
};

struct Ray3 {
    V3 o, d;
    double ts, tp;
};

enum Err { OK = CPD_OK, NSI = CPD_ERR_NO_SURFACE_INTERSECTION, TIR = CPD_ERR_TOTAL_INTERNAL_REFLECTION,
           WEAK = CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK };

struct Spec3 {
    int n_glass;
    bool ar;
    double glass[CPD_MAX_PRISMS][CPD_GLASS_COEFFS];
    Torus lens[2];
    Plane3 planes[CPD_MAX_PRISMS];  // 1..P-1
    double height, half_width;
    uint32_t bin_count;
    double bin_size, slope, intercept, min_ci, length;
    V3 det_normal, det_position, det_direction;
    int beam_kind;
    double beam_width, y_mean, min_cos;
    double wl_lo, wl_hi;
};

// ray.rs:309-367 refract (3D)
static inline Err refract(Ray3 &ray, V3 normal, double n1, double n2, bool ar) {
    const double ci = -dot(ray.d, normal);
    const double r = n1 / n2;
    const double cr_sq = 1.0 - r * r * (1.0 - ci * ci);
    if (cr_sq < 0.0) return TIR;
    const double cr = std::sqrt(cr_sq);
    const V3 v = add(mul(ray.d, r), mul(normal, r * ci - cr));
    double ts, tp;
    if (ar && ci > 0.5) {
        ts = ray.ts * 0.99;
        tp = ray.tp * 0.99;
    } else {
        double rs = (n1 * ci - n2 * cr) / (n1 * ci + n2 * cr);
        double rp = (n1 * cr - n2 * ci) / (n1 * cr + n2 * ci);
        rs = rs * rs;
        rp = rp * rp;
        ts = ray.ts * (1.0 - rs);
        tp = ray.tp * (1.0 - rp);
    }
    ray.d = v;
    ray.ts = ts;
    ray.tp = tp;
    return OK;
}

static inline Err through_torus(Ray3 &ray, const Torus &T, double n1, double n2, bool ar) {
    double dist;
    V3 normal;
    if (!torus_intersection(T, ray.o, ray.d, dist, normal)) return NSI;
    ray.o = mul_add(ray.d, dist, ray.o);
    return refract(ray, normal, n1, n2, ar);
}

// geometry.rs:124-145 HyperPlane::intersection with PrismBounds<3>
static inline Err through_plane(Ray3 &ray, const Plane3 &pl, double height, double half_width, double n1, double n2,
                                bool ar) {
    const double ci = dot(ray.d, pl.normal);
    if (ci == 0.0) return NSI;
    const double d = dot(sub(pl.point, ray.o), pl.normal) / ci;
    if (!(d >= 0.0)) return NSI;
    const V3 p = mul_add(ray.d, d, ray.o);
    if (!(0.0 <= p.y && p.y <= height && -half_width <= p.z && p.z <= half_width)) return NSI;
    ray.o = mul_add(ray.d, d, ray.o);
    return refract(ray, pl.normal, n1, n2, ar);
}

// ray.rs:485-499 propagate
static inline Err prism_propagate(const Spec3 &s, Ray3 &ray, double w) {
    double n1 = calc_n(s.glass[0], w);
    Err e = through_torus(ray, s.lens[0], 1.0, n1, s.ar);
    if (e != OK) return e;
    for (int i = 1; i < s.n_glass; ++i) {
        const double n2 = calc_n(s.glass[i], w);
        e = through_plane(ray, s.planes[i], s.height, s.half_width, n1, n2, s.ar);
        if (e != OK) return e;
        n1 = n2;
    }
    return through_torus(ray, s.lens[1], n1, 1.0, s.ar);
}

// spectrometer.rs:265-303 detector intersection and bin (3D)
static inline Err detector(const Spec3 &s, const Ray3 &ray, uint32_t &bin, double &t) {
    const double ci = -dot(ray.d, s.det_normal);
    if (ci <= s.min_ci) return WEAK;
    const double d = dot(sub(ray.o, s.det_position), s.det_normal) / ci;
    const V3 p = mul_add(ray.d, d, ray.o);
    const double pos = dot(sub(p, s.det_position), s.det_direction);
    if (pos < 0.0 || s.length < pos) return NSI;
    double q, rem;
    euclid_div_rem(pos - s.intercept, s.slope, q, rem);
    const uint32_t b = sat_u32(q);
    if (!(b < s.bin_count && rem < s.bin_size)) return NSI;
    bin = b;
    t = (ray.ts + ray.tp) * 0.5;
    return OK;
}

static inline Err spec_propagate(const Spec3 &s, Ray3 ray, double w, uint32_t &bin, double &t) {
    const Err e = prism_propagate(s, ray, w);
    if (e != OK) return e;
    return detector(s, ray, bin, t);
}

static const double TAU = 6.283185307179586;

// spectrometer.rs:113-131 FiberBeam inverse_cdf with distribution.rs:58-67
static inline Ray3 fiber_ray(const Spec3 &s, const double q[4]) {
    const double th = q[1] * TAU;
    const double sn = std::sin(th), cs = std::cos(th);
    const double rr = s.beam_width * std::sqrt(q[0]);
    const double y = rr * cs, z = rr * sn;
    const double ux = s.min_cos + q[2] * (1.0 - s.min_cos);
    const double s_ux = std::sqrt(1.0 - ux * ux);
    const double rot = q[3] * TAU;
    const double s2 = std::sin(rot), c2 = std::cos(rot);
    return Ray3{{0.0, y + s.y_mean, z}, {ux, s_ux * c2, s_ux * s2}, 1.0, 1.0};
}
// spectrometer.rs:32-46 GaussianBeam (3D)
static inline Ray3 gaussian_ray(const Spec3 &s, const double q[2]) {
    const double r = s.beam_width * ora::norminv(q[0]);
    const double th = q[1] * TAU;
    const double sn = std::sin(th), cs = std::cos(th);
    return Ray3{{0.0, s.y_mean + r * sn, r * cs}, {1.0, 0.0, 0.0}, 1.0, 1.0};
}
static inline Ray3 sample_ray(const Spec3 &s, const double q[4]) {
    return s.beam_kind == CPD_BEAM_FIBER ? fiber_ray(s, q) : gaussian_ray(s, q);
}
static inline double sample_wavelength(const Spec3 &s, double u) { return s.wl_lo + (s.wl_hi - s.wl_lo) * u; }

// qrng.rs:160-177: PHI_2, PHI_4 and their alphas as f64 computes them
static const double PHI_2 = 1.3247179572447460259609088544780973;
static const double PHI_4 = 1.167303978261418684256045899854842180720560371525489039140082449275651903429527;
static inline void alphas(int dims, double a[4]) {
    if (dims == 2) {
        a[0] = 1.0 / PHI_2;
        a[1] = 1.0 / (PHI_2 * PHI_2);
        a[2] = a[3] = 0.0;
    } else {
        a[0] = 1.0 / PHI_4;
        a[1] = 1.0 / (PHI_4 * PHI_4);
        a[2] = 1.0 / (PHI_4 * PHI_4 * PHI_4);
        a[3] = 1.0 / (PHI_4 * PHI_4 * PHI_4 * PHI_4);
    }
}

// Qrng<[f64; N]> from 0.5: iterative (qrng.rs:241-255 iadd_mod_1) or the
// exact fixed-point form frac(0.5 + k alpha_i) with the f64 alphas.
struct QrngN {
    int dims, kind;
    double state[4], a[4];
    uint64_t afx[4];
    uint64_t k = 0;
    QrngN(int dims_, int kind_) : dims(dims_), kind(kind_) {
        alphas(dims, a);
        for (int i = 0; i < 4; ++i) {
            state[i] = 0.5;
            afx[i] = (uint64_t)(a[i] * 0x1.0p+64);
        }
    }
    void next(double q[4]) {
        ++k;
        for (int i = 0; i < 4; ++i) {
            if (i >= dims) {
                q[i] = 0.0;
                continue;
            }
            if (kind == ora::QRNG_ITERATIVE) {
                state[i] = fract(state[i] + a[i]);
                q[i] = state[i];
            } else {
                q[i] = ora::fx_to_f64(0x8000000000000000ull + k * afx[i]);
            }
        }
    }
};
static inline int beam_dims(const Spec3 &s) { return s.beam_kind == CPD_BEAM_FIBER ? 4 : 2; }

// fitness.rs:5-34 / 48-69 for a 3D spectrometer
static void p_dets_l_wavelength(const Spec3 &s, double w, uint64_t max_m, double max_err, int kind,
                                std::vector<Welford<double>> &stats, uint64_t *rays) {
    const double e2 = max_err * max_err;
    double count = 0.0;
    stats.assign(s.bin_count, Welford<double>());
    QrngN qrng(beam_dims(s), kind);
    double q[4];
    for (uint64_t i = 0; i < max_m; ++i) {
        qrng.next(q);
        count += 1.0;
        if (rays) ++*rays;
        uint32_t idx;
        double t;
        if (spec_propagate(s, sample_ray(s, q), w, idx, t) == OK) {
            stats[idx].skip(count);
            stats[idx].next_sample(t);
            if (stats[idx].sem_le_error_threshold(e2)) break;
        }
    }
    for (auto &w2 : stats) w2.skip(count);
}

// fitness.rs:79-123 mutual_information (wavelengths from Qrng<[f64; 1]>)
static double mutual_information(const Spec3 &s, uint64_t max_n, uint64_t max_m, bool adaptive, int kind,
                                 uint64_t *rays, uint64_t *rows) {
    const uint32_t nbin = s.bin_count;
    std::vector<Welford<double>> p_dets(nbin), row;
    Welford<double> h;
    ora::Qrng<double> qrng(kind);
    const double inner = adaptive ? 5e-3 : 0.0, outer = adaptive ? 5e-3 * 5e-3 : 0.0;
    for (uint64_t j = 0; j < max_n; ++j) {
        const double w = sample_wavelength(s, qrng.next());
        if (rows) ++*rows;
        p_dets_l_wavelength(s, w, max_m, inner, kind, row, rays);
        double hs = 0.0;
        for (uint32_t b = 0; b < nbin; ++b) {
            p_dets[b].next_sample(row[b].mean);
            hs += plog2p(row[b].mean);
        }
        h.next_sample(hs);
        bool all = true;
        for (uint32_t b = 0; b < nbin; ++b) all = all && p_dets[b].sem_le_error_threshold(outer);
        if (all) break;
    }
    double h_det = 0.0;
    for (uint32_t b = 0; b < nbin; ++b) h_det = h_det + plog2p(p_dets[b].mean);
    return h.mean - h_det;
}

// spectrometer.rs:305-308, 429-437 (xy components; the detector lies in z = 0)
static inline void size_and_deviation(const Spec3 &s, double &size, double &dev) {
    const V3 mid = mul_add(s.det_direction, s.length * 0.5, s.det_position);
    const double dx = mid.x, dy = mid.y - s.y_mean;
    size = std::sqrt((dx * dx + dy * dy) + 0.0 * 0.0);
    dev = std::fabs(dy / size);
}

// kernel.rs:89-193 for a 3D design, FP64: the warp fold of the 2D oracle.
static uint64_t gpu_row(const Spec3 &s, double w, uint64_t seed_fx, uint32_t max_eval,
                        std::vector<Welford<double>> &stats) {
    const uint32_t nbin = s.bin_count;
    stats.assign(nbin, Welford<double>());
    const double e2 = 5e-3 * 5e-3;
    const uint32_t max_evals = max_eval - max_eval % 32;
    double a[4];
    uint64_t afx[4];
    alphas(4, a);
    for (int i = 0; i < 4; ++i) afx[i] = (uint64_t)(a[i] * 0x1.0p+64);
    double count = 0.0;
    for (uint32_t i = 0; 32 * i < max_evals; ++i) {
        count += 32.0;
        uint32_t bin[32];
        double t[32];
        bool pending[32];
        bool any = false;
        for (uint32_t l = 0; l < 32; ++l) {
            const uint64_t k = l + 32 * (uint64_t)(i + 1);
            double q[4];
            for (int d = 0; d < 4; ++d) q[d] = ora::fx_to_f64(seed_fx + k * afx[d]);
            bin[l] = nbin;
            t[l] = 0.0;
            uint32_t b;
            double tt;
            if (spec_propagate(s, fiber_ray(s, q), w, b, tt) == OK) {
                bin[l] = b;
                t[l] = tt;
            }
            pending[l] = bin[l] < nbin;
            any = any || pending[l];
        }
        bool finished = any;
        for (;;) {
            uint32_t bmin = 0xffffffffu;
            for (uint32_t l = 0; l < 32; ++l)
                if (pending[l] && bin[l] < bmin) bmin = bin[l];
            if (bmin == 0xffffffffu) break;
            double tb[32], dv[32];
            for (uint32_t l = 0; l < 32; ++l) tb[l] = bin[l] == bmin ? t[l] : 0.0;
            const double mean = ora::butterfly_sum32(tb) / 32.0;
            for (uint32_t l = 0; l < 32; ++l) dv[l] = (tb[l] - mean) * (tb[l] - mean);
            Welford<double> wf;
            wf.count = 32.0;
            wf.mean = mean;
            wf.m2 = ora::butterfly_sum32(dv);
            stats[bmin] = Welford<double>::add(stats[bmin], wf);
            finished = finished && stats[bmin].sem_le_error_threshold(e2);
            for (uint32_t l = 0; l < 32; ++l)
                if (bin[l] == bmin) pending[l] = false;
        }
        if (finished) break;
    }
    for (auto &wf : stats) wf.skip(count);
    return (uint64_t)count;
}

// ---------------------------------------------------------------------------
// Construction: create_joined_trapezoids (geometry.rs:430-489) in xy, the
// ToricLens faces (toric.rs:244-291), interior Planes with
// PrismBounds{height, width / 2} (ray.rs:151-176), detector (2D formulas).
// ---------------------------------------------------------------------------
static Torus toric_lens(ora::HyperPlane<double> hp, double kp, double kt, double height, double width) {
    const double chord = std::fabs(height / hp.normal.x);
    const double r = chord * 0.5 / std::fabs(kp);
    const double R = width * 0.5 / kt;
    const double apothem = std::sqrt(r * r - chord * chord * 0.25);
    const double sagitta = r - apothem;
    Torus T;
    const double off = std::copysign(apothem + R, kp);
    T.center = xy(std::fma(hp.normal.x, off, hp.point.x), std::fma(hp.normal.y, off, hp.point.y));
    T.tn = xy(-hp.normal.y, hp.normal.x);  // rot_90_xy
    T.R = R;
    T.r = r;
    T.bc = xy(hp.point.x, hp.point.y);
    T.bn = !std::signbit(kp) ? xy(-hp.normal.x, -hp.normal.y) : xy(hp.normal.x, hp.normal.y);
    T.bh = sagitta;
    return T;
}

static void build3(const cpd_params3 &p, Spec3 &s) {
    const int P = (int)p.n_glass;
    s.n_glass = P;
    s.ar = p.ar_coated != 0;
    for (int i = 0; i < P; ++i)
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) s.glass[i][j] = p.glass[i][j];
    ora::HyperPlane<double> hps[CPD_MAX_PRISMS]{}, last{};
    ora::create_joined_trapezoids<double>(p.height, p.angles, p.lengths, P, hps, &last);
    s.lens[0] = toric_lens(hps[0], p.initial_curvature[0], p.initial_curvature[1], p.height, p.width);
    s.lens[1] = toric_lens(last, p.final_curvature[0], p.final_curvature[1], p.height, p.width);
    for (int i = 1; i < P; ++i) s.planes[i] = Plane3{xy(hps[i].point.x, hps[i].point.y), xy(hps[i].normal.x, hps[i].normal.y)};
    s.height = p.height;
    s.half_width = p.width * 0.5;
    s.beam_kind = (int)p.beam_kind;
    s.beam_width = p.beam_width;
    s.y_mean = p.y_mean;
    s.min_cos = p.beam_kind == CPD_BEAM_FIBER ? std::sqrt(1.0 - p.numerical_aperture * p.numerical_aperture) : 1.0;
    s.wl_lo = p.wavelength_lo;
    s.wl_hi = p.wavelength_hi;
    // detector (spectrometer.rs:183-219)
    s.bin_count = p.bin_count;
    s.bin_size = p.bin_size;
    s.slope = p.linear_slope;
    s.intercept = p.linear_intercept;
    s.min_ci = std::cos(p.max_incident_angle);
    s.length = p.det_length;
    const ora::V2<double> n = ora::rot_180_xy(ora::angled_xy(p.det_angle));
    const ora::V2<double> dd = p.flipped ? ora::rot_90_xy(n) : ora::rot_90_ccw_xy(n);
    s.det_normal = xy(n.x, n.y);
    s.det_direction = xy(dd.x, dd.y);
    s.det_position = xy(p.det_position[0], p.det_position[1]);
}

static void spec_from3(const cpd_design3 &d, Spec3 &s) {
    s.n_glass = (int)d.n_glass;
    s.ar = d.ar_coated != 0;
    std::memcpy(s.glass, d.glass, sizeof(s.glass));
    for (int k = 0; k < 2; ++k) {
        const cpd_torus &t = d.lens[k];
        s.lens[k] = Torus{xy(t.center[0], t.center[1]), xy(t.toroidal_normal[0], t.toroidal_normal[1]),
                          t.toroidal_radius, t.poloidal_radius, xy(t.bound_center[0], t.bound_center[1]),
                          xy(t.bound_normal[0], t.bound_normal[1]), t.bound_height};
    }
    for (int i = 0; i < CPD_MAX_PRISMS; ++i)
        s.planes[i] = Plane3{xy(d.plane_point[i][0], d.plane_point[i][1]), xy(d.plane_normal[i][0], d.plane_normal[i][1])};
    s.height = d.height;
    s.half_width = d.half_width;
    s.bin_count = d.bin_count;
    s.bin_size = d.det_bin_size;
    s.slope = d.det_slope;
    s.intercept = d.det_intercept;
    s.min_ci = d.det_min_ci;
    s.length = d.det_length;
    s.det_normal = xy(d.det_normal[0], d.det_normal[1]);
    s.det_position = xy(d.det_position[0], d.det_position[1]);
    s.det_direction = xy(d.det_direction[0], d.det_direction[1]);
    s.beam_kind = (int)d.beam_kind;
    s.beam_width = d.beam_width;
    s.y_mean = d.y_mean;
    s.min_cos = d.fiber_min_cos;
    s.wl_lo = d.wavelength_lo;
    s.wl_hi = d.wavelength_hi;
}

static void design3_from(const Spec3 &s, cpd_design3 &d) {
    std::memset(&d, 0, sizeof(d));
    d.n_glass = (uint32_t)s.n_glass;
    d.bin_count = s.bin_count;
    d.ar_coated = s.ar ? 1u : 0u;
    d.beam_kind = (uint32_t)s.beam_kind;
    for (int i = 0; i < s.n_glass; ++i)
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) d.glass[i][j] = s.glass[i][j];
    for (int k = 0; k < 2; ++k) {
        const Torus &T = s.lens[k];
        cpd_torus &t = d.lens[k];
        t.center[0] = T.center.x;
        t.center[1] = T.center.y;
        t.toroidal_normal[0] = T.tn.x;
        t.toroidal_normal[1] = T.tn.y;
        t.toroidal_radius = T.R;
        t.poloidal_radius = T.r;
        t.bound_center[0] = T.bc.x;
        t.bound_center[1] = T.bc.y;
        t.bound_normal[0] = T.bn.x;
        t.bound_normal[1] = T.bn.y;
        t.bound_height = T.bh;
    }
    for (int i = 1; i < s.n_glass; ++i) {
        d.plane_point[i][0] = s.planes[i].point.x;
        d.plane_point[i][1] = s.planes[i].point.y;
        d.plane_normal[i][0] = s.planes[i].normal.x;
        d.plane_normal[i][1] = s.planes[i].normal.y;
    }
    d.height = s.height;
    d.half_width = s.half_width;
    d.det_position[0] = s.det_position.x;
    d.det_position[1] = s.det_position.y;
    d.det_direction[0] = s.det_direction.x;
    d.det_direction[1] = s.det_direction.y;
    d.det_normal[0] = s.det_normal.x;
    d.det_normal[1] = s.det_normal.y;
    d.det_min_ci = s.min_ci;
    d.det_length = s.length;
    d.det_slope = s.slope;
    d.det_intercept = s.intercept;
    d.det_bin_size = s.bin_size;
    d.beam_width = s.beam_width;
    d.y_mean = s.y_mean;
    d.fiber_min_cos = s.min_cos;
    d.wavelength_lo = s.wl_lo;
    d.wavelength_hi = s.wl_hi;
}

// spectrometer.rs:322-364 with the 3D start ray (0, median_y, 0) along x
static Err position3(const Spec3 &s, double length, double angle, double acceptance, double pos[2], bool &flipped) {
    const Ray3 start{{0.0, s.y_mean, 0.0}, {1.0, 0.0, 0.0}, 1.0, 1.0};
    const double uh = std::fmax(1.0 - acceptance, 0.0) * 0.5;
    Ray3 lo = start, hi = start;
    Err e = prism_propagate(s, lo, sample_wavelength(s, uh));
    if (e != OK) return e;
    e = prism_propagate(s, hi, sample_wavelength(s, 1.0 - uh));
    if (e != OK) return e;
    if ((lo.ts + lo.tp) * 0.5 <= 1e-3 || (hi.ts + hi.tp) * 0.5 <= 1e-3) return WEAK;
    ora::Ray<double> l2{{lo.o.x, lo.o.y}, {lo.d.x, lo.d.y}, lo.ts, lo.tp};
    ora::Ray<double> h2{{hi.o.x, hi.o.y}, {hi.d.x, hi.d.y}, hi.ts, hi.tp};
    ora::V2<double> p;
    if (!ora::fit_ray_difference_surface(l2, h2, length, angle, p, flipped)) return NSI;
    pos[0] = p.x;
    pos[1] = p.y;
    return OK;
}

}  // namespace ora3

extern "C" {
using namespace ora3;

double ora3_approx_cbrt(double x) { return approx_cbrt(x); }

void ora3_torus_roots(double r_2, double R_2, double dp_magn_sq, double dp_v, double dp_n, double n_v, double out[4]) {
    sympy_soln(r_2, R_2, dp_magn_sq, dp_v, dp_n, n_v, out);
}

// The same closed form with an exact cube root (tests of the algebra alone).
void ora3_torus_roots_exact_cbrt(double r_2, double R_2, double dp_magn_sq, double dp_v, double dp_n, double n_v,
                                 double out[4]) {
    sympy_soln<exact_cbrt>(r_2, R_2, dp_magn_sq, dp_v, dp_n, n_v, out);
}

// Intersection of a ray (origin o[3], direction d[3]) with a lens of the
// design (which: 0 entrance, 1 exit): 1 and (dist, normal[3]) on a hit.
// exact != 0: with an exact cube root (tests of the selection, bounds and
// normal alone)
int ora3_lens_intersection(const cpd_design3 *des, int which, const double o[3], const double d[3], double *dist,
                           double normal[3], int exact) {
    Spec3 s{};
    spec_from3(*des, s);
    V3 n;
    const V3 oo{o[0], o[1], o[2]}, dd{d[0], d[1], d[2]};
    const bool hit = exact ? torus_intersection<exact_cbrt>(s.lens[which], oo, dd, *dist, n)
                           : torus_intersection(s.lens[which], oo, dd, *dist, n);
    if (!hit) return 0;
    normal[0] = n.x;
    normal[1] = n.y;
    normal[2] = n.z;
    return 1;
}

// refract (ray.rs:309-367) of direction d[3] at normal n[3]: out[5] = v, ts, tp
int ora3_refract(const double d[3], const double n[3], double n1, double n2, int ar, double out[5]) {
    Ray3 r{{0, 0, 0}, {d[0], d[1], d[2]}, 1.0, 1.0};
    const ora3::Err e = refract(r, V3{n[0], n[1], n[2]}, n1, n2, ar != 0);
    if (e != ora3::OK) return e;
    out[0] = r.d.x;
    out[1] = r.d.y;
    out[2] = r.d.z;
    out[3] = r.ts;
    out[4] = r.tp;
    return ora3::OK;
}

int ora3_build_design(const cpd_params3 *p, cpd_design3 *out) {
    if (!p || !out || p->n_glass < 1 || p->n_glass > CPD_MAX_PRISMS) return CPD_ERR_INVALID_ARGUMENT;
    Spec3 s{};
    build3(*p, s);
    design3_from(s, *out);
    return CPD_OK;
}

int ora3_position_detector_array(const cpd_params3 *p, double acceptance, double pos[2], uint32_t *flipped) {
    Spec3 s{};
    build3(*p, s);
    bool f = false;
    const ora3::Err e = position3(s, p->det_length, p->det_angle, acceptance, pos, f);
    if (e == ora3::OK && flipped) *flipped = f ? 1u : 0u;
    return e;
}

// The beam ray of quasi-random point q[4] (origin, direction)
void ora3_sample_ray(const cpd_design3 *d, const double q[4], double out[6]) {
    Spec3 s{};
    spec_from3(*d, s);
    const Ray3 r = sample_ray(s, q);
    out[0] = r.o.x;
    out[1] = r.o.y;
    out[2] = r.o.z;
    out[3] = r.d.x;
    out[4] = r.d.y;
    out[5] = r.d.z;
}

// Full trace of ray (o, d) at the wavelength: status, bin, t
int ora3_propagate(const cpd_design3 *d, double wavelength, const double o[3], const double dir[3], uint32_t *bin,
                   double *t) {
    Spec3 s{};
    spec_from3(*d, s);
    return spec_propagate(s, Ray3{{o[0], o[1], o[2]}, {dir[0], dir[1], dir[2]}, 1.0, 1.0}, wavelength, *bin, *t);
}

// The exit ray of the start ray (0, y, 0) along x: out[6]
int ora3_exit_ray(const cpd_design3 *d, double y, double wavelength, double out[6]) {
    Spec3 s{};
    spec_from3(*d, s);
    Ray3 r{{0.0, y, 0.0}, {1.0, 0.0, 0.0}, 1.0, 1.0};
    const ora3::Err e = prism_propagate(s, r, wavelength);
    if (e != ora3::OK) return e;
    out[0] = r.o.x;
    out[1] = r.o.y;
    out[2] = r.o.z;
    out[3] = r.d.x;
    out[4] = r.d.y;
    out[5] = r.d.z;
    return ora3::OK;
}

// The first n points of the sequence of `dims` dimensions (kind 0 iterative,
// 1 fixed point): out[n][4]
void ora3_qmc(int dims, int kind, uint64_t n, double *out) {
    QrngN q(dims, kind);
    for (uint64_t i = 0; i < n; ++i) q.next(out + 4 * i);
}

// p(d|wavelength) of one row: out[bin_count]; returns rays traced
uint64_t ora3_p_dets_l_wavelength(const cpd_design3 *d, double wavelength, uint64_t max_m, int adaptive, int kind,
                                  double *out) {
    Spec3 s{};
    spec_from3(*d, s);
    std::vector<Welford<double>> row;
    uint64_t rays = 0;
    p_dets_l_wavelength(s, wavelength, max_m, adaptive ? 5e-3 : 0.0, kind, row, &rays);
    for (uint32_t b = 0; b < s.bin_count; ++b) out[b] = row[b].mean;
    return rays;
}

// fitness (fitness.rs:130-147): out[3]; counts[2] = {rows, rays} (may be NULL)
int ora3_fitness(const cpd_design3 *d, uint64_t max_n, uint64_t max_m, int adaptive, int kind, double *out,
                 uint64_t *counts) {
    Spec3 s{};
    spec_from3(*d, s);
    uint64_t rays = 0, rows = 0;
    double size, dev;
    size_and_deviation(s, size, dev);
    out[0] = size;
    out[1] = mutual_information(s, max_n, max_m, adaptive != 0, kind, &rays, &rows);
    out[2] = dev;
    if (counts) {
        counts[0] = rows;
        counts[1] = rays;
    }
    return CPD_OK;
}

// fitness of n designs on `threads` host threads: out[n][3]
void ora3_fitness_many(const cpd_design3 *designs, size_t n, uint64_t max_n, uint64_t max_m, int adaptive, int kind,
                       int threads, double *out) {
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < (threads > 0 ? threads : 1); ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;)
                ora3_fitness(&designs[i], max_n, max_m, adaptive, kind, &out[3 * i], nullptr);
        });
    for (auto &th : pool) th.join();
}

// cuda_fitness for a fiber-beam 3D design (FP64): as ora_gpu_fitness
int ora3_gpu_fitness(const cpd_design3 *d, const double *seeds, size_t n_seeds, uint32_t max_n, uint32_t nwarp,
                     uint32_t max_eval, double *out, int *converged, uint64_t *counts) {
    if (!d || !out || nwarp == 0 || d->beam_kind != CPD_BEAM_FIBER) return CPD_ERR_INVALID_ARGUMENT;
    Spec3 s{};
    spec_from3(*d, s);
    const uint32_t nbin = s.bin_count;
    const uint32_t n_rows = (max_n / nwarp) * nwarp;
    const double e2 = 5e-3 * 5e-3;
    std::vector<Welford<double>> p_dets(nbin), row;
    Welford<double> h;
    uint64_t rows = 0, rays = 0;
    int conv = 0;
    for (size_t si = 0; si < n_seeds && !conv; ++si) {
        const uint64_t seed_fx = ora::seed_fx_of(seeds[si]);
        for (uint32_t id = 0; id < n_rows; ++id) {
            const double u = ora::fx_to_f64(seed_fx + (uint64_t)id * ora::ALPHA_FX);
            rays += gpu_row(s, sample_wavelength(s, u), seed_fx, max_eval, row);
            double hs = 0.0;
            for (uint32_t b = 0; b < nbin; ++b) {
                p_dets[b].next_sample(row[b].mean);
                hs += plog2p(row[b].mean);
            }
            h.next_sample(hs);
            ++rows;
        }
        bool all = rows > 0;
        for (uint32_t b = 0; b < nbin; ++b) all = all && p_dets[b].sem_le_error_threshold(e2);
        conv = all ? 1 : 0;
    }
    double h_det = 0.0;
    for (uint32_t b = 0; b < nbin; ++b) h_det = h_det + plog2p(p_dets[b].mean);
    double size, dev;
    size_and_deviation(s, size, dev);
    out[0] = size;
    out[1] = h.mean - h_det;
    out[2] = dev;
    if (converged) *converged = conv;
    if (counts) {
        counts[0] = rows;
        counts[1] = rays;
    }
    return CPD_OK;
}

size_t ora_sizeof_params3(void) { return sizeof(cpd_params3); }
size_t ora_sizeof_design3(void) { return sizeof(cpd_design3); }
}
