// rows3.hip -- the 3D spectrometer (toric lenses, fiber or Gaussian beam) on
// the GPU: the reference's estimators over the Dim3 variant.
//
//  * cpd3_fitness_batch: fitness() (src/fitness.rs:79-147) -- rows of
//    wavelength samples j = 1, 2, ... (R1 from 0.5), each row's rays
//    k = 1, 2, ... from Qrng<[f64; 4]> (fiber) or Qrng<[f64; 2]> (Gaussian),
//    folded ray by ray with the SEM <= 5e-3 early exit (sem_row); rows are
//    launched in growing blocks and folded on the host in order.
//    CPD_MODE_FIXED drops both early exits.
//  * cpd3_p_dets_l_wavelength: the rows alone (transmission_probability).
//  * cpd3_gpu_fitness_batch: cuda_fitness for fiber designs -- the 32-ray
//    chunk fold of kernel.rs:89-193 (chunk_row) over the seeded 4-D sequence.
//  * cpd3_propagate_rays: one ray per (wavelength cdf, ray cdf) pair.
//
// Every ray is traced in FP64 in the reference's operation order (trace3.h).
// The trigonometric part of each ray's beam sample comes from a table built
// on the host per (beam kind, seed, rays), so rows match the FP64 CPU path
// bit for bit.  Where a ray meets the entrance lens does not depend on the
// wavelength: one launch computes every (design, ray) entrance hit before the
// rows, and each row refracts from the table.  A wavefront owns a row; a
// row's glass indices are computed once.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <vector>

#include "cpd.h"
#include "cpd_internal.h"
#include "outer_fold.h"
#include "rows_fold.h"
#include "runtime.h"
#include "trace3.h"

namespace cpd {

namespace {

#define R3_TRY(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? CPD_ERR_OUT_OF_MEMORY : CPD_ERR_DEVICE, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                                 \
    } while (0)

struct RowArgs3 {
    RowSlots slots;
    const cpd_design3 *designs;
    const BeamSample3 *fiber;  // sample of ray k at k - 1, per beam kind
    const BeamSample3 *gauss;
    const EntryHit3 *entry;    // [design][k - 1] entrance hits, or null: trace the entrance per row
    uint32_t entry_rays;       // rays per design in entry
};

// Rays of a 3D row, all in FP64 (nothing to re-trace), from the design's
// entrance hits when the launch has them.
struct Tracer3 {
    static constexpr bool kRetrace = false;
    static constexpr size_t kExtraLds = 0;
    const cpd_design3 *g;
    const BeamSample3 *samples;
    const EntryHit3 *entry;
    double n[CPD_MAX_PRISMS];
    uint32_t nbin;
    bool live = true;

    __device__ Tracer3(const RowArgs3 &a, uint32_t r) : g(&a.designs[a.slots.row_design[r]]) {
        samples = g->beam_kind == CPD_BEAM_FIBER ? a.fiber : a.gauss;
        entry = a.entry ? a.entry + (size_t)a.slots.row_design[r] * a.entry_rays : nullptr;
        glass_indices3(*g, a.slots.row_lambda[r], n);
        nbin = g->bin_count;
    }
    __device__ uint32_t trace(uint32_t k, bool active, double &t, bool &) {
        t = 0.0;
        if (!active) return nbin;
        const Ray3 ray = beam_ray3(*g, samples[k - 1]);
        return entry ? trace3_entered(*g, n, ray, entry[k - 1], t) : trace3(*g, n, ray, t);
    }
    __device__ uint32_t retrace(uint32_t, double &t) {
        t = 0.0;
        return nbin;
    }
};

template <uint32_t W>
__global__ __launch_bounds__(64 * W) void sem_rows3_kernel(RowArgs3 a) {
    extern __shared__ double lds[];
    const uint32_t r = blockIdx.x;
    clear_row(lds, a.slots.nbin_max, a.designs[a.slots.row_design[r]].bin_count);
    Tracer3 tr(a, r);
    __syncthreads();
    sem_row<W>(a.slots, r, tr, lds);
}

// Three wavefronts per SIMD: the compiler fits the FP64 trace and the chunk
// fold in 166 VGPRs without spilling (182 unconstrained, two wavefronts).
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void chunk_rows3_kernel(RowArgs3 a) {
    extern __shared__ double lds[];
    const uint32_t r = blockIdx.x;
    clear_row(lds, a.slots.nbin_max, a.designs[a.slots.row_design[r]].bin_count);
    Tracer3 tr(a, r);
    __syncthreads();
    chunk_row<1>(a.slots, r, tr, lds);
}

// Entrance hits of rays k = 1..n_rays of every design (at [design][k - 1]).
__global__ __launch_bounds__(256) void entrance3_kernel(const cpd_design3 *designs, uint32_t n_designs,
                                                        const BeamSample3 *fiber, const BeamSample3 *gauss,
                                                        uint32_t n_rays, EntryHit3 *out) {
    const uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    if (i >= (uint64_t)n_designs * n_rays) return;
    const cpd_design3 &g = designs[i / n_rays];
    const BeamSample3 *s = g.beam_kind == CPD_BEAM_FIBER ? fiber : gauss;
    out[i] = entrance_hit3(g, beam_ray3(g, s[i % n_rays]));
}

// propagation_test_kernel (kernel.rs:195-224) for 3D designs: ray i at
// wavelength lambda[i] with beam sample s[i].
__global__ __launch_bounds__(256) void propagate3_kernel(const cpd_design3 *g, const double *lambda,
                                                         const BeamSample3 *s, uint32_t n, uint32_t *bin, double *t) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    double nn[CPD_MAX_PRISMS];
    glass_indices3(*g, lambda[i], nn);
    double tt;
    bin[i] = trace3(*g, nn, beam_ray3(*g, s[i]), tt);
    t[i] = tt;
}

cpd_status check_design3(const cpd_design3 &g) {
    if (g.n_glass < 1 || g.n_glass > CPD_MAX_PRISMS) return fail(CPD_ERR_INVALID_ARGUMENT, "n_glass out of range");
    if (g.bin_count < 1 || g.bin_count > CPD_MAX_BINS) return fail(CPD_ERR_INVALID_ARGUMENT, "bin_count out of range");
    if (g.beam_kind != CPD_BEAM_GAUSSIAN && g.beam_kind != CPD_BEAM_FIBER)
        return fail(CPD_ERR_INVALID_ARGUMENT, "unknown beam kind");
    if (!(g.det_slope > 0.0)) return fail(CPD_ERR_INVALID_ARGUMENT, "detector slope must be > 0");
    return CPD_OK;
}

cpd_status check_sampling3(const cpd_sampling &s) {
    if (s.n_wavelengths < 1 || s.n_rays < 1) return fail(CPD_ERR_INVALID_ARGUMENT, "empty sampling");
    if (s.n_rays > (1u << 24)) return fail(CPD_ERR_INVALID_ARGUMENT, "n_rays above 2^24");
    if (s.mode != CPD_MODE_FIXED && s.mode != CPD_MODE_ADAPTIVE)
        return fail(CPD_ERR_INVALID_ARGUMENT, "unknown sampling mode");
    return CPD_OK;
}

}  // namespace

// Beam samples of rays k = 1..n (at k - 1) for one beam kind and seed,
// built on the host (as the FP64 CPU path evaluates them) and cached per
// kind while the seed is unchanged and n does not grow.
cpd_status Context::ensure_samples3(uint32_t kind, uint64_t seed_fx, uint32_t n, const void **table) {
    DeviceBuffer &buf = kind == CPD_BEAM_FIBER ? samples3 : samples3_gauss;
    uint64_t &key_seed = kind == CPD_BEAM_FIBER ? samples3_key : samples3_gauss_key;
    uint32_t &have = kind == CPD_BEAM_FIBER ? samples3_n : samples3_gauss_n;
    if (key_seed == seed_fx && have >= n) {
        *table = buf.ptr;
        return CPD_OK;
    }
    cpd_status st;
    if ((st = buf.reserve((size_t)n * sizeof(BeamSample3), false, stream)) != CPD_OK) return st;
    if ((st = host_stage.reserve((size_t)n * sizeof(BeamSample3))) != CPD_OK) return st;
    BeamSample3 *h = (BeamSample3 *)host_stage.ptr;
    uint64_t afx[4];
    qmc_alpha_fx(kind, afx);
    parallel_for((n + 1023) / 1024, [&](size_t blk) {
        const uint32_t k0 = (uint32_t)blk * 1024u + 1u, k1 = std::min<uint32_t>(n, k0 + 1023u);
        for (uint32_t k = k0; k <= k1; ++k) {
            double q[4];
            qmc_point3(afx, seed_fx, k, q);
            h[k - 1] = beam_sample3(kind, q);
        }
    });
    R3_TRY(hipMemcpyAsync(buf.ptr, h, (size_t)n * sizeof(BeamSample3), hipMemcpyHostToDevice, stream));
    R3_TRY(hipStreamSynchronize(stream));
    key_seed = seed_fx;
    have = n;
    *table = buf.ptr;
    return CPD_OK;
}

cpd_status Context::stage_designs3(const cpd_design3 *designs, const std::vector<uint32_t> &valid,
                                   const cpd_design3 **designs_dev, uint32_t *nbin_max) {
    const size_t n = valid.size();
    cpd_status st;
    if ((st = design3_stage.reserve(n * sizeof(cpd_design3), false, stream)) != CPD_OK) return st;
    if ((st = host_stage.reserve(n * sizeof(cpd_design3))) != CPD_OK) return st;
    cpd_design3 *h = (cpd_design3 *)host_stage.ptr;
    uint32_t nb = 1;
    for (size_t q = 0; q < n; ++q) {
        h[q] = designs[valid[q]];
        nb = std::max(nb, h[q].bin_count);
    }
    R3_TRY(hipMemcpyAsync(design3_stage.ptr, h, n * sizeof(cpd_design3), hipMemcpyHostToDevice, stream));
    R3_TRY(hipStreamSynchronize(stream));
    *designs_dev = design3_stage.as<cpd_design3>();
    *nbin_max = nb;
    return CPD_OK;
}

// The entrance hits of rays 1..n_rays of the staged designs, or null when the
// table would pass kEntryBudget (the rows then trace their entrances).
static constexpr size_t kEntryBudget = size_t(2) << 30;

cpd_status Context::entrance_table3(const cpd_design3 *designs_dev, uint32_t n_designs, const void *fiber_tab,
                                    const void *gauss_tab, uint32_t n_rays, const void **table) {
    *table = nullptr;
    const size_t bytes = (size_t)n_designs * n_rays * sizeof(EntryHit3);
    if (bytes == 0 || bytes > kEntryBudget) return CPD_OK;
    cpd_status st;
    if ((st = entry3.reserve(bytes, false, stream)) != CPD_OK) return st;
    const uint64_t total = (uint64_t)n_designs * n_rays;
    hipLaunchKernelGGL(entrance3_kernel, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, stream, designs_dev,
                       n_designs, (const BeamSample3 *)fiber_tab, (const BeamSample3 *)gauss_tab, n_rays,
                       entry3.as<EntryHit3>());
    R3_TRY(hipGetLastError());
    *table = entry3.ptr;
    return CPD_OK;
}

cpd_status Context::run_rows3(bool chunked, const cpd_design3 *designs_dev, const std::vector<uint32_t> &row_design,
                              const std::vector<double> &row_lambda, const std::vector<uint64_t> &row_out,
                              const std::vector<uint32_t> &row_cap, uint64_t out_n, uint32_t max_rays,
                              uint32_t nbin_max, double max_err_sq, const void *fiber_tab, const void *gauss_tab,
                              const void *entry, uint32_t entry_rays, const double *&rows_host,
                              const uint64_t *&rays_host) {
    const uint32_t n_rows = (uint32_t)row_design.size();
    if (n_rows == 0) return CPD_OK;
    RowArgs3 a{};
    cpd_status st;
    if ((st = stage_rows(row_design, row_lambda, row_out, row_cap, out_n, max_rays, nbin_max, max_err_sq, a.slots)) !=
        CPD_OK)
        return st;
    a.designs = designs_dev;
    a.fiber = (const BeamSample3 *)fiber_tab;
    a.gauss = (const BeamSample3 *)gauss_tab;
    a.entry = (const EntryHit3 *)entry;
    a.entry_rays = entry_rays;
    const size_t lds = chunked ? chunk_lds_bytes(nbin_max, 1) : sem_lds_bytes(nbin_max);
    if (profile) R3_TRY(hipEventRecord(ev[0], stream));
    if (chunked)
        hipLaunchKernelGGL(chunk_rows3_kernel, dim3(n_rows), dim3(64), lds, stream, a);
    else if (sem_waves(n_rows, cu_count) == 1)
        hipLaunchKernelGGL(sem_rows3_kernel<1>, dim3(n_rows), dim3(64), lds, stream, a);
    else
        hipLaunchKernelGGL(sem_rows3_kernel<kSemWaves>, dim3(n_rows), dim3(64 * kSemWaves), lds, stream, a);
    return finish_rows(a.slots, out_n, rows_host, rays_host);
}

// fitness() for the 3D designs in `valid`, all of one sampling mode.
cpd_status Context::run_fitness3(const cpd_design3 *designs, const cpd_sampling *sampling,
                                 const std::vector<uint32_t> &valid, cpd_fitness *out) {
    const size_t n = valid.size();
    if (n == 0) return CPD_OK;
    R3_TRY(hipSetDevice(device));
    const bool adaptive = sampling[valid[0]].mode == CPD_MODE_ADAPTIVE;
    uint32_t max_m = 0;
    bool fiber = false, gauss = false;
    for (uint32_t i : valid) {
        max_m = std::max(max_m, sampling[i].n_rays);
        (designs[i].beam_kind == CPD_BEAM_FIBER ? fiber : gauss) = true;
    }
    cpd_status st;
    const void *ftab = nullptr, *gtab = nullptr;
    if (fiber && (st = ensure_samples3(CPD_BEAM_FIBER, kHalfFx, max_m, &ftab)) != CPD_OK) return st;
    if (gauss && (st = ensure_samples3(CPD_BEAM_GAUSSIAN, kHalfFx, max_m, &gtab)) != CPD_OK) return st;
    const cpd_design3 *ddev;
    uint32_t nbin_max;
    if ((st = stage_designs3(designs, valid, &ddev, &nbin_max)) != CPD_OK) return st;
    const void *entry = nullptr;
    if ((st = entrance_table3(ddev, (uint32_t)n, ftab, gtab, max_m, &entry)) != CPD_OK) return st;
    row_kernel_ms = 0.0;
    double fold_ms = 0.0;
    const double e2 = adaptive ? kMaxErr * kMaxErr : 0.0;
    std::vector<OuterFold> fold(n);
    std::vector<uint64_t> next_j(n, 1), block(n, adaptive ? 4 : ~0ull);
    for (size_t q = 0; q < n; ++q) {
        fold[q].reset(designs[valid[q]].bin_count);
        fold[q].e2 = e2;
    }
    std::vector<uint32_t> row_design, row_cap;
    std::vector<double> row_lambda;
    std::vector<uint64_t> row_out;
    const double *rows = nullptr;
    const uint64_t *rays = nullptr;
    uint64_t total_rays = 0;
    for (;;) {
        row_design.clear();
        row_lambda.clear();
        row_out.clear();
        row_cap.clear();
        uint64_t out_n = 0;
        for (size_t q = 0; q < n; ++q) {
            if (fold[q].done) continue;
            const cpd_design3 &g = designs[valid[q]];
            const cpd_sampling &s = sampling[valid[q]];
            const uint64_t last = std::min<uint64_t>(s.n_wavelengths, next_j[q] + std::min<uint64_t>(block[q], s.n_wavelengths) - 1);
            for (uint64_t j = next_j[q]; j <= last; ++j) {
                row_design.push_back((uint32_t)q);
                row_lambda.push_back(lambda_of3(g, qmc_u64(j)));
                row_out.push_back(out_n);
                row_cap.push_back(s.n_rays);
                out_n += g.bin_count;
            }
        }
        if (row_design.empty()) break;
        if ((st = run_rows3(false, ddev, row_design, row_lambda, row_out, row_cap, out_n, max_m, nbin_max, e2, ftab,
                            gtab, entry, max_m, rows, rays)) != CPD_OK)
            return st;
        const auto f0 = std::chrono::steady_clock::now();
        for (size_t r = 0; r < row_design.size();) {
            const uint32_t q = row_design[r];
            size_t e = r;
            while (e < row_design.size() && row_design[e] == q) ++e;
            const cpd_sampling &s = sampling[valid[q]];
            for (size_t x = r; x < e && !fold[q].done; ++x) {
                total_rays += rays[x];
                fold[q].rays += rays[x];
                if (fold[q].fold(rows + row_out[x]) || fold[q].rows >= s.n_wavelengths) fold[q].done = true;
            }
            next_j[q] += e - r;
            // 3D rows are trace-bound: rows past the stop cost throughput, so the
            // blocks grow more slowly than the 2D rows'
            if (adaptive) block[q] = std::min<uint64_t>(block[q] * 4, 1024);
            r = e;
        }
        fold_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - f0).count();
    }
    for (size_t q = 0; q < n; ++q) {
        double size, dev;
        size_and_deviation3(designs[valid[q]], size, dev);
        out[valid[q]] = cpd_fitness{size, fold[q].info(), dev};
    }
    last = cpd_run_stats{};
    last.rays = total_rays;
    last.trace_ms = row_kernel_ms;
    last.finalize_ms = fold_ms;
    return CPD_OK;
}

cpd_status Context::rows3(const cpd_design3 &g, const double *wavelengths, size_t n_w, uint32_t max_m, uint32_t mode,
                          double *out) {
    R3_TRY(hipSetDevice(device));
    cpd_status st;
    const void *tab = nullptr;
    if ((st = ensure_samples3(g.beam_kind, kHalfFx, max_m, &tab)) != CPD_OK) return st;
    const cpd_design3 *ddev;
    uint32_t nbin_max;
    std::vector<uint32_t> one{0};
    if ((st = stage_designs3(&g, one, &ddev, &nbin_max)) != CPD_OK) return st;
    const void *entry = nullptr;
    if ((st = entrance_table3(ddev, 1, tab, tab, max_m, &entry)) != CPD_OK) return st;
    std::vector<uint32_t> row_design(n_w, 0), row_cap(n_w, max_m);
    std::vector<double> row_lambda(wavelengths, wavelengths + n_w);
    std::vector<uint64_t> row_out(n_w);
    for (size_t j = 0; j < n_w; ++j) row_out[j] = j * g.bin_count;
    const double *rows = nullptr;
    const uint64_t *rays = nullptr;
    const double e2 = mode == CPD_MODE_ADAPTIVE ? kMaxErr * kMaxErr : 0.0;
    if ((st = run_rows3(false, ddev, row_design, row_lambda, row_out, row_cap, n_w * g.bin_count, max_m, nbin_max, e2,
                        tab, tab, entry, max_m, rows, rays)) != CPD_OK)
        return st;
    memcpy(out, rows, n_w * g.bin_count * sizeof(double));
    return CPD_OK;
}

// cuda_fitness (cuda_fitness.rs:347-423) for the fiber designs in `valid`.
cpd_status Context::run_gpu_fitness3(const cpd_design3 *designs, const std::vector<uint32_t> &valid,
                                     const double *seeds, size_t n_seeds, uint32_t max_n, uint32_t nwarp,
                                     uint32_t max_eval, cpd_fitness *out, cpd_status *per_design) {
    const size_t n = valid.size();
    if (n == 0) return CPD_OK;
    R3_TRY(hipSetDevice(device));
    const uint32_t n_rows_per = (max_n / nwarp) * nwarp;
    const uint32_t max_evals = max_eval - max_eval % 32;
    const uint32_t table_n = max_evals + 32;
    cpd_status st;
    const cpd_design3 *ddev;
    uint32_t nbin_max;
    if ((st = stage_designs3(designs, valid, &ddev, &nbin_max)) != CPD_OK) return st;
    row_kernel_ms = 0.0;
    double fold_ms = 0.0;
    std::vector<OuterFold> fold(n);
    for (size_t q = 0; q < n; ++q) fold[q].reset(designs[valid[q]].bin_count);
    std::vector<uint32_t> row_design;
    std::vector<double> row_lambda;
    std::vector<uint64_t> row_out;
    const double *rows = nullptr;
    const uint64_t *rays = nullptr;
    uint64_t total_rays = 0;
    for (size_t si = 0; si < n_seeds; ++si) {
        const uint64_t seed_fx = seed_to_fx(seeds[si]);
        row_design.clear();
        row_lambda.clear();
        row_out.clear();
        uint64_t out_n = 0;
        for (size_t q = 0; q < n; ++q) {
            if (fold[q].done) continue;
            const cpd_design3 &g = designs[valid[q]];
            for (uint32_t id = 0; id < n_rows_per; ++id) {
                row_design.push_back((uint32_t)q);
                row_lambda.push_back(lambda_of3(g, qmc_seeded(seed_fx, id)));
                row_out.push_back(out_n);
                out_n += g.bin_count;
            }
        }
        if (row_design.empty()) break;
        const void *tab = nullptr, *entry = nullptr;
        if ((st = ensure_samples3(CPD_BEAM_FIBER, seed_fx, table_n, &tab)) != CPD_OK) return st;
        if ((st = entrance_table3(ddev, (uint32_t)n, tab, tab, table_n, &entry)) != CPD_OK) return st;
        if ((st = run_rows3(true, ddev, row_design, row_lambda, row_out, std::vector<uint32_t>(), out_n, max_evals,
                            nbin_max, kMaxErr * kMaxErr, tab, tab, entry, table_n, rows, rays)) != CPD_OK)
            return st;
        const auto f0 = std::chrono::steady_clock::now();
        for (size_t r = 0; r < row_design.size(); ++r) total_rays += rays[r];
        std::vector<size_t> starts;
        for (size_t r = 0; r < row_design.size(); r += n_rows_per) starts.push_back(r);
        parallel_for(starts.size(), [&](size_t i) {
            const size_t r0 = starts[i];
            OuterFold &f = fold[row_design[r0]];
            for (size_t r = r0; r < r0 + n_rows_per; ++r) f.fold(rows + row_out[r]);
            if (f.rows && f.converged()) f.done = true;
        });
        fold_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - f0).count();
    }
    for (size_t q = 0; q < n; ++q) {
        double size, dev;
        size_and_deviation3(designs[valid[q]], size, dev);
        out[valid[q]] = cpd_fitness{size, fold[q].info(), dev};
        if (per_design) per_design[valid[q]] = fold[q].done ? CPD_OK : CPD_ERR_NOT_CONVERGED;
    }
    last = cpd_run_stats{};
    last.rays = total_rays;
    last.trace_ms = row_kernel_ms;
    last.finalize_ms = fold_ms;
    return CPD_OK;
}

cpd_status Context::propagate3(const cpd_design3 &g, const double *wcdf, const double *rcdf, size_t n, uint32_t *bin,
                               double *t) {
    R3_TRY(hipSetDevice(device));
    // staged: design | wavelengths | samples | bins | t
    const size_t off_w = 1024, off_s = off_w + (n * 8 + 255) / 256 * 256;
    const size_t off_b = off_s + (n * sizeof(BeamSample3) + 255) / 256 * 256;
    const size_t off_t = off_b + (n * 4 + 255) / 256 * 256;
    const size_t total = off_t + n * 8;
    cpd_status st;
    if ((st = prop_buf.reserve(total, false, stream)) != CPD_OK) return st;
    if ((st = host_stage.reserve(off_b)) != CPD_OK) return st;
    char *h = (char *)host_stage.ptr;
    memcpy(h, &g, sizeof g);
    double *lam = (double *)(h + off_w);
    BeamSample3 *smp = (BeamSample3 *)(h + off_s);
    for (size_t i = 0; i < n; ++i) {
        lam[i] = lambda_of3(g, wcdf[i]);
        smp[i] = beam_sample3(g.beam_kind, rcdf + 4 * i);
    }
    char *d = (char *)prop_buf.ptr;
    R3_TRY(hipMemcpyAsync(d, h, off_b, hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(propagate3_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, stream,
                       (const cpd_design3 *)d, (const double *)(d + off_w), (const BeamSample3 *)(d + off_s),
                       (uint32_t)n, (uint32_t *)(d + off_b), (double *)(d + off_t));
    R3_TRY(hipGetLastError());
    R3_TRY(hipMemcpyAsync(bin, d + off_b, n * 4, hipMemcpyDeviceToHost, stream));
    R3_TRY(hipMemcpyAsync(t, d + off_t, n * 8, hipMemcpyDeviceToHost, stream));
    R3_TRY(hipStreamSynchronize(stream));
    return CPD_OK;
}

}  // namespace cpd

using cpd::Context;

extern "C" {

cpd_status cpd3_fitness_batch(cpd_ctx *ctx, const cpd_design3 *designs, const cpd_sampling *sampling, size_t n,
                              cpd_fitness *out, cpd_status *per_design) {
    if (n == 0) return CPD_OK;
    if (!designs || !sampling || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > (1u << 20)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "batch above 2^20 designs");
    std::vector<uint32_t> fixed, adaptive;
    for (size_t i = 0; i < n; ++i) {
        cpd_status s = cpd::check_design3(designs[i]);
        if (s == CPD_OK) s = cpd::check_sampling3(sampling[i]);
        if (per_design) per_design[i] = s;
        if (s == CPD_OK)
            (sampling[i].mode == CPD_MODE_ADAPTIVE ? adaptive : fixed).push_back((uint32_t)i);
        else
            out[i].size = out[i].info = out[i].deviation = NAN;
    }
    Context *c;
    cpd_status st = cpd::resolve_context(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    for (const std::vector<uint32_t> *group : {&fixed, &adaptive}) {
        st = c->run_fitness3(designs, sampling, *group, out);
        if (st != CPD_OK) {
            if (per_design)
                for (uint32_t i : *group) per_design[i] = st;
            return st;
        }
    }
    return CPD_OK;
}

cpd_status cpd3_p_dets_l_wavelength(cpd_ctx *ctx, const cpd_design3 *design, const double *wavelengths,
                                    size_t n_wavelengths, uint32_t max_m, uint32_t mode, double *out) {
    if (!design || (n_wavelengths && (!wavelengths || !out))) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n_wavelengths == 0) return CPD_OK;
    if (n_wavelengths > (1u << 24)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "too many wavelengths");
    cpd_status st = cpd::check_design3(*design);
    if (st != CPD_OK) return st;
    if ((st = cpd::check_sampling3(cpd_sampling{(uint32_t)n_wavelengths, max_m, mode, 0})) != CPD_OK) return st;
    Context *c;
    if ((st = cpd::resolve_context(ctx, &c)) != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    return c->rows3(*design, wavelengths, n_wavelengths, max_m, mode, out);
}

cpd_status cpd3_gpu_fitness_batch(cpd_ctx *ctx, const cpd_design3 *designs, size_t n, const double *seeds,
                                  size_t n_seeds, uint32_t max_n, uint32_t nwarp, uint32_t max_eval,
                                  cpd_fitness *out, cpd_status *per_design) {
    if (n == 0) return CPD_OK;
    if (!designs || !out || (n_seeds && !seeds)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nwarp == 0) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "nwarp must be >= 1");
    if (max_eval > (1u << 24)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "max_eval above 2^24");
    if (n > (1u << 20)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "batch above 2^20 designs");
    if ((uint64_t)max_n * n > (1ull << 28)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "max_n x designs above 2^28 rows");
    std::vector<uint32_t> valid;
    for (size_t i = 0; i < n; ++i) {
        cpd_status s = cpd::check_design3(designs[i]);
        if (s == CPD_OK && designs[i].beam_kind != CPD_BEAM_FIBER)
            s = cpd::fail(CPD_ERR_INVALID_ARGUMENT, "No kernel compiled for this Spectrometer (Gaussian beam in 3D)");
        if (per_design) per_design[i] = s;
        if (s == CPD_OK)
            valid.push_back((uint32_t)i);
        else
            out[i].size = out[i].info = out[i].deviation = NAN;
    }
    if (valid.empty() && n == 1) return per_design ? per_design[0] : CPD_ERR_INVALID_ARGUMENT;
    Context *c;
    cpd_status st = cpd::resolve_context(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    st = c->run_gpu_fitness3(designs, valid, seeds, n_seeds, max_n, nwarp, max_eval, out, per_design);
    if (st != CPD_OK && per_design)
        for (uint32_t i : valid) per_design[i] = st;
    return st;
}

cpd_status cpd3_propagate_rays(cpd_ctx *ctx, const cpd_design3 *design, const double *wavelength_cdf,
                               const double *ray_cdf, size_t n, uint32_t *bin_out, double *t_out) {
    if (n == 0) return CPD_OK;
    if (!design || !wavelength_cdf || !ray_cdf || !bin_out || !t_out)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > (1u << 26)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "too many rays");
    cpd_status st = cpd::check_design3(*design);
    if (st != CPD_OK) return st;
    Context *c;
    if ((st = cpd::resolve_context(ctx, &c)) != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    return c->propagate3(*design, wavelength_cdf, ray_cdf, n, bin_out, t_out);
}

}  // extern "C"
