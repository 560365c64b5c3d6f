// design3.cpp -- host-side construction of 3D spectrometers: the Dim3
// CompoundPrism (toric entrance and exit faces), detector positioning with
// the 3D median ray, and the synthetic 3D workload.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "construct.h"
#include "cpd.h"
#include "cpd_internal.h"
#include "trace3.h"

namespace {

using cpd::Vec2;

// ToricLens::from_hyperplane (src/spectrometer/toric.rs:244-291): poloidal
// radius r = chord / (2 |kp|) with chord = |h / n.x|, toroidal radius
// R = width / (2 kt), the torus centre on the face normal at
// sign(kp) (apothem + R) from the face's mid-height point, the axis along the
// face (normal rotated +90 degrees), and bounds between the chord plane and
// the plane at the sagitta, facing away from the bulge.
cpd_torus toric_lens(Vec2 pt, Vec2 nm, double kp, double kt, double height, double width) {
#pragma clang fp contract(off)
    cpd_torus t{};
    const double chord = fabs(height / nm.x);
    const double r = chord * 0.5 / fabs(kp);
    const double R = width * 0.5 / kt;
    const double apothem = sqrt(r * r - chord * chord * 0.25);
    const double sagitta = r - apothem;
    const double off = copysign(apothem + R, kp);
    t.center[0] = fma(nm.x, off, pt.x);
    t.center[1] = fma(nm.y, off, pt.y);
    t.toroidal_normal[0] = -nm.y;
    t.toroidal_normal[1] = nm.x;
    t.toroidal_radius = R;
    t.poloidal_radius = r;
    t.bound_center[0] = pt.x;
    t.bound_center[1] = pt.y;
    const bool up = !signbit(kp);
    t.bound_normal[0] = up ? -nm.x : nm.x;
    t.bound_normal[1] = up ? -nm.y : nm.y;
    t.bound_height = sagitta;
    return t;
}

bool finite_nonzero(double v) { return v != 0.0 && isfinite(v); }

cpd_status check3(const cpd_params3 *p, bool detector) {
    if (!p) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "params is NULL");
    if (p->n_glass < 1 || p->n_glass > CPD_MAX_PRISMS)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "n_glass must be in 1..6, got %u", p->n_glass);
    if (!(p->height > 0.0)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "height must be > 0");
    if (!(p->width > 0.0)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "width must be > 0");
    for (int k = 0; k < 2; ++k)
        if (!finite_nonzero(p->initial_curvature[k]) || !finite_nonzero(p->final_curvature[k]))
            return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "curvatures must be finite and non-zero");
    if (!(p->wavelength_hi >= p->wavelength_lo))
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "lower wavelength bound > upper bound");
    if (p->beam_kind != CPD_BEAM_GAUSSIAN && p->beam_kind != CPD_BEAM_FIBER)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "beam_kind must be CPD_BEAM_GAUSSIAN or CPD_BEAM_FIBER");
    if (p->beam_kind == CPD_BEAM_FIBER && !(fabs(p->numerical_aperture) <= 1.0))
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "numerical aperture must be in [-1, 1]");
    if (detector) {
        if (p->bin_count < 1 || p->bin_count > CPD_MAX_BINS)
            return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "bin_count must be in 1..%d", CPD_MAX_BINS);
        if (!(p->linear_slope > 0.0)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "linear_slope must be > 0");
    }
    return CPD_OK;
}

// The prism and beam fields (what positioning traces through).
void build_prism3(const cpd_params3 &p, cpd_design3 &d) {
#pragma clang fp contract(off)
    const uint32_t P = p.n_glass;
    d.n_glass = P;
    d.ar_coated = p.ar_coated ? 1u : 0u;
    d.beam_kind = p.beam_kind;
    for (uint32_t i = 0; i < P; ++i)
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) d.glass[i][j] = p.glass[i][j];
    double pt[CPD_MAX_PRISMS][2] = {}, nm[CPD_MAX_PRISMS][2] = {};
    Vec2 ept{0, 0}, enm{0, 0};
    cpd::trapezoid_faces(p.height, p.angles, p.lengths, P, pt, nm, ept, enm);
    d.lens[0] = toric_lens(Vec2{pt[0][0], pt[0][1]}, Vec2{nm[0][0], nm[0][1]}, p.initial_curvature[0],
                           p.initial_curvature[1], p.height, p.width);
    d.lens[1] = toric_lens(ept, enm, p.final_curvature[0], p.final_curvature[1], p.height, p.width);
    for (uint32_t i = 1; i < P; ++i) {
        d.plane_point[i][0] = pt[i][0];
        d.plane_point[i][1] = pt[i][1];
        d.plane_normal[i][0] = nm[i][0];
        d.plane_normal[i][1] = nm[i][1];
    }
    d.height = p.height;
    d.half_width = p.width * 0.5;
    d.beam_width = p.beam_width;
    d.y_mean = p.y_mean;
    // FiberBeam::new: min_cos = sqrt(1 - NA^2) (spectrometer.rs:85-93)
    d.fiber_min_cos = p.beam_kind == CPD_BEAM_FIBER ? sqrt(1.0 - p.numerical_aperture * p.numerical_aperture) : 1.0;
    d.wavelength_lo = p.wavelength_lo;
    d.wavelength_hi = p.wavelength_hi;
}

void build_detector3(const cpd_params3 &p, const double pos[2], uint32_t flipped, cpd_design3 &d) {
    d.bin_count = p.bin_count;
    d.det_bin_size = p.bin_size;
    d.det_slope = p.linear_slope;
    d.det_intercept = p.linear_intercept;
    d.det_min_ci = cos(p.max_incident_angle);
    d.det_length = p.det_length;
    const Vec2 n = cpd::face_normal(p.det_angle);
    d.det_normal[0] = n.x;
    d.det_normal[1] = n.y;
    d.det_direction[0] = flipped ? -n.y : n.y;
    d.det_direction[1] = flipped ? n.x : -n.x;
    d.det_position[0] = pos[0];
    d.det_position[1] = pos[1];
}

// The median ray (0, median_y, 0) along x through the prism at a wavelength.
int exit_ray3(const cpd_design3 &d, double y, double lambda, cpd::Ray3 &ray) {
    double n[CPD_MAX_PRISMS];
    cpd::glass_indices3(d, lambda, n);
    ray = cpd::Ray3{cpd::P3{0.0, y, 0.0}, cpd::P3{1.0, 0.0, 0.0}, 1.0, 1.0};
    return cpd::prism3(ray, d, n);
}

cpd_status position3(const cpd_params3 &p, double acceptance, double pos[2], uint32_t &flipped_out) {
#pragma clang fp contract(off)
    cpd_design3 d{};
    build_prism3(p, d);
    const double half_out = fmax(1.0 - acceptance, 0.0) * 0.5;
    cpd::Ray3 lo, hi;
    int e = exit_ray3(d, d.y_mean, cpd::lambda_of3(d, half_out), lo);
    if (e != CPD_OK) return (cpd_status)e;
    e = exit_ray3(d, d.y_mean, cpd::lambda_of3(d, 1.0 - half_out), hi);
    if (e != CPD_OK) return (cpd_status)e;
    if ((lo.ts + lo.tp) * 0.5 <= 1e-3 || (hi.ts + hi.tp) * 0.5 <= 1e-3) return CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK;
    // fit_ray_difference_surface reads the rays' x and y
    const cpd::Ray64 l2{lo.o.x, lo.o.y, lo.v.x, lo.v.y, lo.ts, lo.tp};
    const cpd::Ray64 h2{hi.o.x, hi.o.y, hi.v.x, hi.v.y, hi.ts, hi.tp};
    bool flipped = false;
    if (!cpd::fit_detector(l2, h2, p.det_length, p.det_angle, pos, flipped)) return CPD_ERR_NO_SURFACE_INTERSECTION;
    flipped_out = flipped ? 1u : 0u;
    return CPD_OK;
}

struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    double uniform(double a, double b) { return a + (b - a) * uniform(); }
};

constexpr double kPi = 3.14159265358979323846;

// Fraction of the first n beam rays (R1 from 0.5) that reach a bin, over the
// wavelengths at the centre and the two ends of the range.
double probe_hits3(const cpd_design3 &d, uint32_t n) {
    uint64_t afx[4];
    cpd::qmc_alpha_fx(d.beam_kind, afx);
    uint32_t hits = 0;
    for (double u : {0.5, 0.05, 0.95}) {
        double idx[CPD_MAX_PRISMS];
        cpd::glass_indices3(d, cpd::lambda_of3(d, u), idx);
        for (uint32_t k = 1; k <= n; ++k) {
            double q[4], t;
            cpd::qmc_point3(afx, cpd::kHalfFx, k, q);
            hits += cpd::trace3(d, idx, cpd::beam_ray3(d, cpd::beam_sample3(d.beam_kind, q)), t) < d.bin_count;
        }
    }
    return hits / (3.0 * n);
}

}  // namespace

extern "C" {

cpd_status cpd3_build_design(const cpd_params3 *params, cpd_design3 *out) {
    cpd_status st = check3(params, true);
    if (st != CPD_OK) return st;
    if (!out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    cpd_design3 d{};
    build_prism3(*params, d);
    build_detector3(*params, params->det_position, params->flipped ? 1u : 0u, d);
    *out = d;
    return CPD_OK;
}

cpd_status cpd3_position_detector_array(const cpd_params3 *params, double acceptance, double position_out[2],
                                        uint32_t *flipped_out) {
    cpd_status st = check3(params, false);
    if (st != CPD_OK) return st;
    if (!position_out || !flipped_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    return position3(*params, acceptance, position_out, *flipped_out);
}

cpd_status cpd3_size_and_deviation(const cpd_design3 *design, double *size, double *deviation) {
    if (!design || !size || !deviation) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    cpd::size_and_deviation3(*design, *size, *deviation);
    return CPD_OK;
}

cpd_status cpd3_exit_ray(const cpd_design3 *design, double y, double wavelength, double out[6]) {
    if (!design || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (design->n_glass < 1 || design->n_glass > CPD_MAX_PRISMS)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "n_glass out of range");
    cpd::Ray3 ray;
    const int e = exit_ray3(*design, y, wavelength, ray);
    if (e != CPD_OK) return (cpd_status)e;
    const double v[6] = {ray.o.x, ray.o.y, ray.o.z, ray.v.x, ray.v.y, ray.v.z};
    memcpy(out, v, sizeof v);
    return CPD_OK;
}

// Synthetic 3D workload (no reference counterpart; the reference's optimiser
// draws its 3D designs in designer.py:160-265): 3 prisms of random
// Sellmeier glasses, 32 bins, a fiber of 50 um core and NA 0.1 (lengths in
// cm), toric faces with poloidal curvatures U(0.05, 1) and toroidal
// curvatures U(0.2, 1), the detector positioned for acceptance 1.  Draws
// whose construction or positioning fails, or whose probe of 64 beam rays at
// three wavelengths lands fewer than 10% in bins, are redrawn.  The entrance
// curvature is positive (the face concave towards the beam): with a negative
// one the torus normal at the entrance points along the ray, and the
// reference's refraction sends every ray backwards.  The sampling is 64
// wavelengths x 4096 rays.
cpd_status cpd3_generate_design(uint64_t index, cpd_params3 *params_out, cpd_sampling *sampling_out,
                                uint32_t *attempts_out) {
    if (!params_out || !sampling_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    SplitMix rng{0x3D70C0FFEEull ^ (index * 0xD1B54A32D192ED03ull)};
    rng.next();
    cpd_params3 p;
    memset(&p, 0, sizeof p);
    const uint32_t P = 3;
    p.n_glass = P;
    p.width = 2.0;
    p.bin_count = 32;
    p.bin_size = 3.2 / 32;
    p.linear_slope = 3.2 / 32;
    p.det_length = 3.2;
    p.max_incident_angle = kPi / 4;
    p.beam_kind = CPD_BEAM_FIBER;
    p.beam_width = 0.005;
    p.numerical_aperture = 0.1;
    p.wavelength_lo = 0.5;
    p.wavelength_hi = 0.82;
    for (uint32_t attempt = 1; attempt <= 1000000; ++attempt) {
        for (uint32_t i = 0; i <= P; ++i) p.angles[i] = rng.uniform(-kPi / 2, kPi / 2);
        p.height = rng.uniform(0.5, 2.0);
        for (uint32_t i = 0; i < P; ++i) p.lengths[i] = rng.uniform(0.5, 2.0);
        bool glass_ok = true;
        for (uint32_t i = 0; i < P; ++i) {
            double B[3], C[3];
            for (int j = 0; j < 3; ++j) B[j] = rng.uniform(0.3, 1.5);
            for (int j = 0; j < 3; ++j) C[j] = rng.uniform(0.001, 0.2);
            if (cpd_fit_sellmeier1(B, C, p.wavelength_lo, p.wavelength_hi, 100, p.glass[i], nullptr) != CPD_OK)
                glass_ok = false;
        }
        p.initial_curvature[0] = rng.uniform(0.05, 1.0);
        p.initial_curvature[1] = rng.uniform(0.2, 1.0);
        p.final_curvature[0] = rng.uniform(0.05, 1.0);
        p.final_curvature[1] = rng.uniform(0.2, 1.0);
        p.y_mean = rng.uniform(0.2, 0.8) * p.height;
        p.det_angle = rng.uniform(-kPi, kPi);
        if (!glass_ok) continue;
        double pos[2];
        uint32_t flipped = 0;
        if (cpd3_position_detector_array(&p, 1.0, pos, &flipped) != CPD_OK) continue;
        p.det_position[0] = pos[0];
        p.det_position[1] = pos[1];
        p.flipped = flipped;
        cpd_design3 d;
        if (cpd3_build_design(&p, &d) != CPD_OK) continue;
        if (probe_hits3(d, 64) < 0.1) continue;
        *params_out = p;
        sampling_out->n_wavelengths = 64;
        sampling_out->n_rays = 4096;
        sampling_out->mode = CPD_MODE_FIXED;
        sampling_out->reserved = 0;
        if (attempts_out) *attempts_out = attempt;
        cpd::clear_error();
        return CPD_OK;
    }
    return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "no valid 3D design after 1e6 draws");
}

cpd_status cpd3_generate_designs(uint64_t first_index, size_t n, uint32_t threads, cpd_params3 *params_out,
                                 cpd_design3 *designs_out, cpd_sampling *sampling_out) {
    if (n && (!params_out || !designs_out || !sampling_out)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    const unsigned hw = std::thread::hardware_concurrency();
    const unsigned T = (unsigned)std::max<size_t>(1, std::min<size_t>(n, threads ? threads : (hw ? hw : 1)));
    std::atomic<size_t> next{0};
    std::atomic<int> err{CPD_OK};
    auto work = [&]() {
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n || err.load() != CPD_OK) break;
            cpd_status st = cpd3_generate_design(first_index + i, &params_out[i], &sampling_out[i], nullptr);
            if (st == CPD_OK) st = cpd3_build_design(&params_out[i], &designs_out[i]);
            if (st != CPD_OK) {
                int expected = CPD_OK;
                err.compare_exchange_strong(expected, st);
                break;
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < T; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    return (cpd_status)err.load();
}

}  // extern "C"
