// design.cpp -- host-side (FP64) construction of 2D compound-prism
// spectrometers, detector positioning, Sellmeier glass fits and the synthetic
// benchmark workloads.  None of this touches the GPU.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "cpd.h"
#include "cpd_internal.h"
#include "trace_core.h"

namespace {

thread_local std::string g_last_error;

struct Vec2 {
    double x, y;
};

// Unit vector at angle theta, rotated by 180 degrees: (-cos, -sin)
// (src/spectrometer/vector.rs:336-371, angled_xy().rot_180_xy()).
Vec2 face_normal(double theta) {
    double c = cos(theta), s = sin(theta);
    return Vec2{-c, -s};
}

bool sign_positive(double v) { return !signbit(v); }

cpd_status check_params(const cpd_params *p) {
    if (!p) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "params is NULL");
    if (p->n_glass < 1 || p->n_glass > CPD_MAX_PRISMS)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "n_glass must be in 1..6, got %u", p->n_glass);
    if (!(p->height > 0.0)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "height must be > 0");
    if (!(p->curvature != 0.0) || !isfinite(p->curvature))
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "curvature must be finite and non-zero");
    if (!(p->wavelength_hi >= p->wavelength_lo))
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "lower wavelength bound > upper bound");
    return CPD_OK;
}

// The compound prism part of a design: face planes from the angles and
// separation lengths (create_joined_trapezoids, geometry.rs:430-489), the
// exit lens from the last plane (CurvedPlane::from_hyperplane,
// geometry.rs:280-315).
void build_prism(const cpd_params &p, cpd_design &d) {
#pragma clang fp contract(off)
    const uint32_t P = p.n_glass;
    d.n_glass = P;
    d.ar_coated = p.ar_coated ? 1u : 0u;
    for (uint32_t i = 0; i < P; ++i)
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) d.glass[i][j] = p.glass[i][j];
    d.height = p.height;

    const double h2 = p.height * 0.5;
    // entrance face: mid-height point at x = |tan(theta0)| h/2
    Vec2 n = face_normal(p.angles[0]);
    double x0 = fabs(n.y / n.x) * h2;
    d.plane_point[0][0] = x0;
    d.plane_point[0][1] = h2;
    d.plane_normal[0][0] = n.x;
    d.plane_normal[0][1] = n.y;
    bool prev_up = sign_positive(n.y);
    double prev_half = x0;  // half of the face's x-extent over the height
    double x = x0;
    Vec2 exit_point{0, 0}, exit_normal{0, 0};
    // faces 1..P: face i sits lengths[i-1] beyond the previous one, measured
    // between the facing edges of the two trapezoids
    for (uint32_t i = 1; i <= P; ++i) {
        Vec2 ni = face_normal(p.angles[i]);
        bool up = sign_positive(ni.y);
        double half = fabs(ni.y / ni.x) * h2;
        double gap = (up != prev_up) ? prev_half + half : fabs(prev_half - half);
        x += p.lengths[i - 1] + gap;
        prev_up = up;
        prev_half = half;
        if (i < P) {
            d.plane_point[i][0] = x;
            d.plane_point[i][1] = h2;
            d.plane_normal[i][0] = ni.x;
            d.plane_normal[i][1] = ni.y;
        } else {
            exit_point = Vec2{x, h2};
            exit_normal = ni;
        }
    }
    // exit lens: chord = |h / n.x|, R = chord / (2 |k|), centre on the normal
    // through the chord midpoint at the apothem, bounds centred on the arc's
    // midpoint with radius^2 = sagitta^2 + chord^2 / 4
    const double k = p.curvature;
    double chord = fabs(p.height / exit_normal.x);
    double radius = chord * 0.5 / fabs(k);
    double apothem = sqrt(radius * radius - chord * chord * 0.25);
    double sagitta = radius - apothem;
    double sa = copysign(apothem, k), ss = -copysign(sagitta, k);
    d.arc_center[0] = fma(exit_normal.x, sa, exit_point.x);
    d.arc_center[1] = fma(exit_normal.y, sa, exit_point.y);
    d.arc_radius = radius;
    d.arc_midpt[0] = fma(exit_normal.x, ss, exit_point.x);
    d.arc_midpt[1] = fma(exit_normal.y, ss, exit_point.y);
    d.arc_max_dist_sq = sagitta * sagitta + chord * chord * 0.25;

    d.beam_width = p.beam_width;
    d.y_mean = p.y_mean;
    d.wavelength_lo = p.wavelength_lo;
    d.wavelength_hi = p.wavelength_hi;
}

// LinearDetectorArray::new (spectrometer.rs:183-219) with
// min_ci = cos(max_incident_angle) (pylib.rs:906-920).
void build_detector(const cpd_params &p, cpd_design &d) {
    d.bin_count = p.bin_count;
    d.det_bin_size = p.bin_size;
    d.det_slope = p.linear_slope;
    d.det_intercept = p.linear_intercept;
    d.det_min_ci = cos(p.max_incident_angle);
    d.det_length = p.det_length;
    Vec2 n = face_normal(p.det_angle);
    d.det_normal[0] = n.x;
    d.det_normal[1] = n.y;
    // direction: normal rotated +90 degrees when flipped, -90 otherwise
    if (p.flipped) {
        d.det_direction[0] = -n.y;
        d.det_direction[1] = n.x;
    } else {
        d.det_direction[0] = n.y;
        d.det_direction[1] = -n.x;
    }
    d.det_position[0] = p.det_position[0];
    d.det_position[1] = p.det_position[1];
}

// fit_ray_difference_surface (geometry.rs:495-542): the array of length L at
// angle a spans the two exit rays of the extreme wavelengths.
bool fit_detector(const cpd::Ray64 &lo, const cpd::Ray64 &hi, double length, double angle, double pos[2],
                  bool &flipped) {
#pragma clang fp contract(off)
    // array direction (-sin a, cos a) scaled by its length
    double sx = -sin(angle) * length, sy = cos(angle) * length;
    double a = hi.dx, b = -lo.dx, c = hi.dy, dd = -lo.dy;
    double det = a * dd - b * c;
    if (det == 0.0) return false;
    double m10 = -c / det, m11 = a / det;
    // distance along the lower ray to the array end that the upper ray meets
    double tx = sx - hi.ox + lo.ox, ty = sy - hi.oy + lo.oy;
    double s = m10 * tx + m11 * ty;
    if (s > 0.0) {
        pos[0] = fma(lo.dx, s, lo.ox);
        pos[1] = fma(lo.dy, s, lo.oy);
        flipped = false;
        return true;
    }
    tx = -sx - hi.ox + lo.ox;
    ty = -sy - hi.oy + lo.oy;
    s = m10 * tx + m11 * ty;
    if (s < 0.0) return false;
    pos[0] = fma(lo.dx, s, lo.ox);
    pos[1] = fma(lo.dy, s, lo.oy);
    flipped = true;
    return true;
}

// ---------------------------------------------------------------------------
// Sellmeier1 -> degree-5 polynomial least squares (Householder QR)
// ---------------------------------------------------------------------------
double sellmeier1(const double b[3], const double c[3], double w) {
    double w2 = w * w;
    return sqrt(1.0 + b[0] * w2 / (w2 - c[0]) + b[1] * w2 / (w2 - c[1]) + b[2] * w2 / (w2 - c[2]));
}

// Least squares min |A x - y| for A (m x 6, row-major), by Householder QR
// with column scaling.  Returns false when A is rank deficient.
bool lstsq6(int m, double *A, double *y, double x[6]) {
    const int n = 6;
    double scale[6];
    for (int j = 0; j < n; ++j) {
        double s = 0;
        for (int i = 0; i < m; ++i) s = fmax(s, fabs(A[i * n + j]));
        if (s == 0) return false;
        scale[j] = s;
        for (int i = 0; i < m; ++i) A[i * n + j] /= s;
    }
    for (int k = 0; k < n; ++k) {
        double norm = 0;
        for (int i = k; i < m; ++i) norm = hypot(norm, A[i * n + k]);
        if (norm == 0) return false;
        double alpha = A[k * n + k] > 0 ? -norm : norm;
        double v0 = A[k * n + k] - alpha;
        // v = (v0, A[k+1..m, k]); H = I - 2 v v^T / (v^T v)
        double vtv = v0 * v0;
        for (int i = k + 1; i < m; ++i) vtv += A[i * n + k] * A[i * n + k];
        for (int j = k + 1; j < n; ++j) {
            double s = v0 * A[k * n + j];
            for (int i = k + 1; i < m; ++i) s += A[i * n + k] * A[i * n + j];
            s = 2.0 * s / vtv;
            A[k * n + j] -= s * v0;
            for (int i = k + 1; i < m; ++i) A[i * n + j] -= s * A[i * n + k];
        }
        double s = v0 * y[k];
        for (int i = k + 1; i < m; ++i) s += A[i * n + k] * y[i];
        s = 2.0 * s / vtv;
        y[k] -= s * v0;
        for (int i = k + 1; i < m; ++i) y[i] -= s * A[i * n + k];
        A[k * n + k] = alpha;
    }
    for (int k = n - 1; k >= 0; --k) {
        double s = y[k];
        for (int j = k + 1; j < n; ++j) s -= A[k * n + j] * x[j];
        x[k] = s / A[k * n + k];
    }
    for (int j = 0; j < n; ++j) x[j] /= scale[j];
    return true;
}

// ---------------------------------------------------------------------------
// SplitMix64 (Steele, Lea & Flood 2014) workload generator
// ---------------------------------------------------------------------------
struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0, 1)
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform(); }
    uint32_t below(uint32_t n) { return (uint32_t)(uniform() * n); }
};

const double kPi = 3.14159265358979323846;

}  // namespace

namespace cpd {

cpd_status fail(cpd_status st, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return st;
}

void clear_error() { g_last_error.clear(); }

}  // namespace cpd

extern "C" {

int cpd_abi_version(void) { return CPD_ABI_VERSION; }

const char *cpd_status_name(cpd_status st) {
    switch (st) {
        case CPD_OK: return "Ok";
        case CPD_ERR_NO_SURFACE_INTERSECTION: return "NoSurfaceIntersection";
        case CPD_ERR_OUT_OF_BOUNDS: return "OutOfBounds";
        case CPD_ERR_TOTAL_INTERNAL_REFLECTION: return "TotalInternalReflection";
        case CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK: return "SpectrometerAngularResponseTooWeak";
        case CPD_ERR_INVALID_ARGUMENT: return "InvalidArgument";
        case CPD_ERR_NOT_CONVERGED: return "NotConverged";
        case CPD_ERR_NOT_READY: return "NotReady";
        case CPD_ERR_NO_DEVICE: return "NoDevice";
        case CPD_ERR_DEVICE: return "DeviceError";
        case CPD_ERR_OUT_OF_MEMORY: return "OutOfMemory";
        case CPD_ERR_CAPACITY: return "Capacity";
    }
    return "Unknown";
}

const char *cpd_last_error(void) { return g_last_error.c_str(); }

cpd_status cpd_position_detector_array(const cpd_params *params, double acceptance, double position_out[2],
                                       uint32_t *flipped_out) {
#pragma clang fp contract(off)
    cpd_status st = check_params(params);
    if (st != CPD_OK) return st;
    if (!position_out || !flipped_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    cpd_design d{};
    build_prism(*params, d);
    // the two extreme accepted wavelengths through the prism from the beam's mean height
    double half_out = fmax(1.0 - acceptance, 0.0) * 0.5;
    double w_lo = cpd::lambda_of(d, half_out);
    double w_hi = cpd::lambda_of(d, 1.0 - half_out);
    cpd::Ray64 lo{0.0, d.y_mean, 1.0, 0.0, 1.0, 1.0}, hi = lo;
    int e = cpd::prism64(lo, d, w_lo);
    if (e != CPD_OK) return (cpd_status)e;
    e = cpd::prism64(hi, d, w_hi);
    if (e != CPD_OK) return (cpd_status)e;
    if ((lo.ts + lo.tp) * 0.5 <= 1e-3 || (hi.ts + hi.tp) * 0.5 <= 1e-3) return CPD_ERR_ANGULAR_RESPONSE_TOO_WEAK;
    bool flipped = false;
    if (!fit_detector(lo, hi, params->det_length, params->det_angle, position_out, flipped))
        return CPD_ERR_NO_SURFACE_INTERSECTION;
    *flipped_out = flipped ? 1u : 0u;
    return CPD_OK;
}

cpd_status cpd_build_design(const cpd_params *params, cpd_design *out) {
    cpd_status st = check_params(params);
    if (st != CPD_OK) return st;
    if (!out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    if (params->bin_count < 1 || params->bin_count > CPD_MAX_BINS)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "bin_count must be in 1..%d", CPD_MAX_BINS);
    if (!(params->linear_slope > 0.0)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "linear_slope must be > 0");
    cpd_design d{};
    build_prism(*params, d);
    build_detector(*params, d);
    *out = d;
    return CPD_OK;
}

cpd_status cpd_size_and_deviation(const cpd_design *g, double *size, double *deviation) {
#pragma clang fp contract(off)
    if (!g || !size || !deviation) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    cpd::size_and_deviation64(*g, *size, *deviation);
    return CPD_OK;
}

cpd_status cpd_exit_ray(const cpd_design *g, double y, double wavelength, double out[4]) {
    if (!g || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (g->n_glass < 1 || g->n_glass > CPD_MAX_PRISMS) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "n_glass out of range");
    cpd::Ray64 ray{0.0, y, 1.0, 0.0, 1.0, 1.0};
    int e = cpd::prism64(ray, *g, wavelength);
    if (e != CPD_OK) return (cpd_status)e;
    out[0] = ray.ox;
    out[1] = ray.oy;
    out[2] = ray.dx;
    out[3] = ray.dy;
    return CPD_OK;
}

cpd_status cpd_ray_trace(const cpd_design *g, double wavelength, double y, double *out, uint32_t *n_rows) {
    if (!g || !out || !n_rows) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (g->n_glass < 1 || g->n_glass > CPD_MAX_PRISMS) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "n_glass out of range");
    uint32_t k = 0;
    auto row = [&](double ox, double oy, double dx, double dy) {
        out[4 * k + 0] = ox;
        out[4 * k + 1] = oy;
        out[4 * k + 2] = dx;
        out[4 * k + 3] = dy;
        ++k;
    };
    const bool ar = g->ar_coated != 0;
    cpd::Ray64 ray{0.0, y, 1.0, 0.0, 1.0, 1.0};
    row(ray.ox, ray.oy, ray.dx, ray.dy);
    // surfaces one at a time; the path stops at the first failure
    // (src/spectrometer/ray.rs:501-541)
    double n1 = cpd::glass_n64(g->glass[0], wavelength);
    bool ok;
    {
        cpd::Ray64 r = ray;
        ok = cpd::plane64(r, g->plane_point[0], g->plane_normal[0], g->height, 1.0, n1, ar) == CPD_OK;
        if (ok) {
            ray = r;
            row(ray.ox, ray.oy, ray.dx, ray.dy);
        }
    }
    for (uint32_t i = 1; ok && i < g->n_glass; ++i) {
        double n2 = cpd::glass_n64(g->glass[i], wavelength);
        cpd::Ray64 r = ray;
        ok = cpd::plane64(r, g->plane_point[i], g->plane_normal[i], g->height, n1, n2, ar) == CPD_OK;
        if (ok) {
            ray = r;
            row(ray.ox, ray.oy, ray.dx, ray.dy);
        }
        n1 = n2;
    }
    if (ok) {
        cpd::Ray64 r = ray;
        ok = cpd::arc64(r, *g, n1, ar) == CPD_OK;
        if (ok) {
            ray = r;
            row(ray.ox, ray.oy, ray.dx, ray.dy);
        }
    }
    // the detector step starts from the last ray that made it, even after a
    // failure (src/spectrometer/spectrometer.rs:453-459)
    uint32_t bin;
    double t;
    if (cpd::detector64(ray, *g, bin, t) == CPD_OK) {
        double ci = -(ray.dx * g->det_normal[0] + ray.dy * g->det_normal[1]);
        double d = ((ray.ox - g->det_position[0]) * g->det_normal[0] + (ray.oy - g->det_position[1]) * g->det_normal[1]) / ci;
        row(fma(ray.dx, d, ray.ox), fma(ray.dy, d, ray.oy), ray.dx, ray.dy);
    }
    *n_rows = k;
    return CPD_OK;
}

cpd_status cpd_fit_sellmeier1(const double b[3], const double c[3], double lo, double hi, uint32_t npts,
                              double coeffs_out[CPD_GLASS_COEFFS], double *chisq_out) {
    if (!b || !c || !coeffs_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(lo < hi) || npts < CPD_GLASS_COEFFS || npts > 100000)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "need lo < hi and 6 <= npts <= 100000");
    std::string err;
    double *A = new double[(size_t)npts * 6];
    double *y = new double[npts];
    double *y0 = new double[npts];
    double *w = new double[npts];
    // numpy.linspace(lo, hi, npts)
    double step = (hi - lo) / (double)(npts - 1);
    for (uint32_t i = 0; i < npts; ++i) w[i] = (i == npts - 1) ? hi : lo + (double)i * step;
    for (uint32_t i = 0; i < npts; ++i) {
        double v = sellmeier1(b, c, w[i]);
        y[i] = y0[i] = v;
        double pw = 1.0;
        for (int j = 5; j >= 0; --j) {
            A[i * 6 + j] = pw;
            pw *= w[i];
        }
    }
    bool ok = true;
    for (uint32_t i = 0; i < npts; ++i) ok = ok && isfinite(y[i]);
    double x[6] = {0};
    if (ok) ok = lstsq6((int)npts, A, y, x);
    if (ok) {
        double chisq = 0;
        for (uint32_t i = 0; i < npts; ++i) {
            double r = y0[i] - cpd::glass_n64(x, w[i]);
            chisq += r * r;
        }
        for (int j = 0; j < 6; ++j) coeffs_out[j] = x[j];
        if (chisq_out) *chisq_out = chisq;
    }
    delete[] A;
    delete[] y;
    delete[] y0;
    delete[] w;
    if (!ok) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "Sellmeier fit failed (pole in range or singular)");
    return CPD_OK;
}

// Draw order per design: sampling counts once; then, per attempt, face angles,
// height, separation lengths, glasses (B1..B3, C1..C3 each), lens curvature,
// beam mean height, detector angle.  Attempts whose construction or detector
// positioning fails are redrawn.
cpd_status cpd_generate_design(uint32_t config, uint64_t index, cpd_params *params_out, cpd_sampling *sampling_out,
                               uint32_t *attempts_out) {
    if (!params_out || !sampling_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    if (config != 0 && config != 1 && config != 2 && config != 4)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "config must be 0, 1, 2 or 4");
    SplitMix64 rng{((0xC0FFEE00ull + (config == 1 ? 0 : config)) << 32) ^ (index * 0xD1B54A32D192ED03ull)};
    rng.next();
    uint32_t P = 3, nbin = 32, n_lambda = 64, n_rays = 16384;
    if (config == 2) {
        P = 1 + rng.below(6);
        const uint32_t bins[4] = {16, 32, 64, 128};
        nbin = bins[rng.below(4)];
        n_lambda = 32 + rng.below(225);
        n_rays = 64 * (64 + rng.below(961));
    } else if (config == 4) {
        P = 6;
        nbin = 256;
        n_lambda = 1024;
        n_rays = 1u << 20;
    }
    cpd_params p;
    memset(&p, 0, sizeof p);
    p.n_glass = P;
    p.width = 2.0;
    p.bin_count = nbin;
    p.bin_size = 3.2 / nbin;
    p.linear_slope = 3.2 / nbin;
    p.linear_intercept = 0.0;
    p.det_length = 3.2;
    p.max_incident_angle = kPi / 4;
    p.beam_width = 0.2;
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
        p.curvature = 1.0 - rng.uniform();
        p.y_mean = rng.uniform() * p.height;
        p.det_angle = rng.uniform(-kPi, kPi);
        if (!glass_ok) continue;
        double pos[2];
        uint32_t flipped = 0;
        if (cpd_position_detector_array(&p, 1.0, pos, &flipped) != CPD_OK) continue;
        p.det_position[0] = pos[0];
        p.det_position[1] = pos[1];
        p.flipped = flipped;
        cpd_design d;
        if (cpd_build_design(&p, &d) != CPD_OK) continue;
        *params_out = p;
        sampling_out->n_wavelengths = n_lambda;
        sampling_out->n_rays = n_rays;
        sampling_out->mode = CPD_MODE_FIXED;
        sampling_out->reserved = 0;
        if (attempts_out) *attempts_out = attempt;
        cpd::clear_error();
        return CPD_OK;
    }
    return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "no valid design after 1e6 draws");
}

cpd_status cpd_generate_designs(uint32_t config, uint64_t first_index, size_t n, uint32_t threads,
                                cpd_params *params_out, cpd_design *designs_out, cpd_sampling *sampling_out,
                                uint32_t *attempts_out) {
    if (n && (!params_out || !designs_out || !sampling_out)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    unsigned hw = std::thread::hardware_concurrency();
    const unsigned T = (unsigned)std::max<size_t>(1, std::min<size_t>(n, threads ? threads : (hw ? hw : 1)));
    std::atomic<size_t> next{0};
    std::atomic<int> err{CPD_OK};
    auto work = [&]() {
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n || err.load() != CPD_OK) break;
            uint32_t att = 0;
            cpd_status st = cpd_generate_design(config, first_index + i, &params_out[i], &sampling_out[i], &att);
            if (st == CPD_OK) st = cpd_build_design(&params_out[i], &designs_out[i]);
            if (st != CPD_OK) {
                int expected = CPD_OK;
                err.compare_exchange_strong(expected, st);
                break;
            }
            if (attempts_out) attempts_out[i] = att;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < T; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    return (cpd_status)err.load();
}

}  // extern "C"
