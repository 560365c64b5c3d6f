// design.cpp -- the host entry points of design construction (the FP64
// construction itself is construct.h, shared with the device), Sellmeier
// glass fits and the synthetic benchmark workloads.  None of this touches the
// GPU.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "cpd.h"
#include "construct.h"
#include "cpd_internal.h"
#include "trace_core.h"

namespace {

thread_local std::string g_last_error;

cpd_status check_params(const cpd_params *p) {
    if (!p) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "params is NULL");
    switch (cpd::check_prism_params(*p)) {
        case cpd::PARAM_N_GLASS:
            return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "n_glass must be in 1..6, got %u", p->n_glass);
        case cpd::PARAM_HEIGHT: return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "height must be > 0");
        case cpd::PARAM_CURVATURE: return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "curvature must be finite and non-zero");
        case cpd::PARAM_WAVELENGTHS: return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "lower wavelength bound > upper bound");
        default: return CPD_OK;
    }
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

// Transmission probe of the generated workloads: the fraction of rays k =
// 1..n_rays of the beam at wavelength samples j = 1..n_lambda (the R1
// sequences of fitness.rs:19,94) that land in a bin, traced in FP64.
constexpr uint32_t kProbeLambdas = 4, kProbeRays = 256;
constexpr double kMinProbeHits = 0.1;

double probe_hit_fraction(const cpd_design &g, uint32_t n_lambda, uint32_t n_rays) {
    uint64_t hits = 0;
    for (uint32_t j = 1; j <= n_lambda; ++j) {
        const double w = cpd::lambda_of(g, cpd::qmc_u64(j));
        for (uint32_t k = 1; k <= n_rays; ++k) {
            double t;
            const double y = cpd::beam_y64(g, cpd::norminv64(cpd::qmc_u64(k)));
            hits += cpd::trace64(g, w, y, t) < g.bin_count;
        }
    }
    return (double)hits / ((double)n_lambda * (double)n_rays);
}

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
    cpd_status st = check_params(params);
    if (st != CPD_OK) return st;
    if (!position_out || !flipped_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    return cpd::position_detector(*params, acceptance, position_out, *flipped_out);
}

cpd_status cpd_build_design(const cpd_params *params, cpd_design *out) {
    cpd_status st = check_params(params);
    if (st != CPD_OK) return st;
    if (!out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    if (params->bin_count < 1 || params->bin_count > CPD_MAX_BINS)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "bin_count must be in 1..%d", CPD_MAX_BINS);
    if (!(params->linear_slope > 0.0)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "linear_slope must be > 0");
    return cpd::construct_design(*params, 0.0, *out);
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
// beam mean height, detector angle, and for config 7 the detector's bin size,
// intercept and anti-reflection coating.  Attempts whose construction or
// detector positioning fails are redrawn; for configs 5..7 so are attempts
// whose transmission probe lands fewer than kMinProbeHits of its rays in a bin.
cpd_status cpd_generate_design(uint32_t config, uint64_t index, cpd_params *params_out, cpd_sampling *sampling_out,
                               uint32_t *attempts_out) {
    if (!params_out || !sampling_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    if (config == 3 || config > 7) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "config must be 0, 1, 2, 4, 5, 6 or 7");
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
    } else if (config == 5) {
        P = 4;  // spring.toml nglass; design_config.toml max-count
    } else if (config == 6) {
        nbin = 64;  // bch.toml
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
    // the designer's variable ranges (designer.py:122-130, 216-227) in mm
    double max_height = 0.0;
    if (config == 5) {
        // spring.toml:4-19 and design_config.toml:17-41: 32 bins of 0.8 at
        // slope 1 from intercept 0.1 over 32 mm, anti-reflection coated
        max_height = 25.0;
        p.width = 7.0;
        p.ar_coated = 1;
        p.bin_size = 0.8;
        p.linear_slope = 1.0;
        p.linear_intercept = 0.1;
        p.det_length = 32.0;
        p.beam_width = 3.2;
    } else if (config == 6) {
        // bch.toml:3-15: 64 bins of 0.42 over a 26.6 mm array (the last bin
        // overhangs its end), wavelengths 0.48..1.0 um, anti-reflection coated
        max_height = 20.0;
        p.width = 6.6;
        p.ar_coated = 1;
        p.bin_size = 0.42;
        p.linear_slope = 0.42;
        p.det_length = 26.6;
        p.beam_width = 2.0;
        p.wavelength_lo = 0.48;
        p.wavelength_hi = 1.0;
    }
    for (uint32_t attempt = 1; attempt <= 1000000; ++attempt) {
        for (uint32_t i = 0; i <= P; ++i) p.angles[i] = rng.uniform(-kPi / 2, kPi / 2);
        if (max_height > 0.0) {
            p.height = rng.uniform(1e-4 * max_height, max_height);
            for (uint32_t i = 0; i < P; ++i) p.lengths[i] = atanh(rng.uniform()) * max_height / 4.0;
        } else {
            p.height = rng.uniform(0.5, 2.0);
            for (uint32_t i = 0; i < P; ++i) p.lengths[i] = rng.uniform(0.5, 2.0);
        }
        bool glass_ok = true;
        for (uint32_t i = 0; i < P; ++i) {
            double B[3], C[3];
            for (int j = 0; j < 3; ++j) B[j] = rng.uniform(0.3, 1.5);
            for (int j = 0; j < 3; ++j) C[j] = rng.uniform(0.001, 0.2);
            if (cpd_fit_sellmeier1(B, C, p.wavelength_lo, p.wavelength_hi, 100, p.glass[i], nullptr) != CPD_OK)
                glass_ok = false;
        }
        p.curvature = max_height > 0.0 ? rng.uniform(1e-5, 1.0) : 1.0 - rng.uniform();
        p.y_mean = rng.uniform() * p.height;
        p.det_angle = rng.uniform(-kPi, kPi);
        if (config == 7) {
            // gaps between bins, an intercept of up to three bin pitches (so
            // positions before it fall in bin 0 by the saturating cast when
            // their remainder is below the bin size), a coin for the coating
            p.bin_size = p.linear_slope * rng.uniform(0.5, 1.0);
            p.linear_intercept = p.linear_slope * rng.uniform(0.0, 3.0);
            p.ar_coated = rng.uniform() < 0.5 ? 1u : 0u;
        }
        if (!glass_ok) continue;
        double pos[2];
        uint32_t flipped = 0;
        if (cpd_position_detector_array(&p, 1.0, pos, &flipped) != CPD_OK) continue;
        p.det_position[0] = pos[0];
        p.det_position[1] = pos[1];
        p.flipped = flipped;
        cpd_design d;
        if (cpd_build_design(&p, &d) != CPD_OK) continue;
        if (config >= 5 && probe_hit_fraction(d, kProbeLambdas, kProbeRays) < kMinProbeHits) continue;
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

cpd_status cpd_hit_fraction(const cpd_design *design, uint32_t n_lambda, uint32_t n_rays, double *out) {
    if (!design || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (design->n_glass < 1 || design->n_glass > CPD_MAX_PRISMS || design->bin_count < 1 || n_lambda < 1 ||
        n_rays < 1)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "need 1..6 glasses, bins, wavelengths and rays");
    *out = probe_hit_fraction(*design, n_lambda, n_rays);
    return CPD_OK;
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
