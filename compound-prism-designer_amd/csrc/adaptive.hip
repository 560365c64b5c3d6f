// adaptive.hip -- the reference's adaptive estimators on the GPU.
//
// Two estimators stop tracing a wavelength's rays once its bin statistics
// are precise enough:
//  * CPD_MODE_ADAPTIVE, src/fitness.rs:5-34 + 79-123 (what `fitness()` and
//    pyo3 `cpu_fitness` compute): rays k = 1, 2, ... of one wavelength are
//    folded one at a time into per-bin Welford summaries, padded with zeros
//    (`skip`), and the row stops as soon as the bin just updated has a
//    standard error of the mean below 5e-3.  Rows j = 1, 2, ... stop once
//    every bin's mean over the rows has one.
//  * cpd_gpu_fitness, src/spectrometer/kernel.rs:89-193 +
//    src/cuda_fitness.rs:347-423 (pyo3 `gpu_fitness`): per seed, wavelength
//    row id uses u = frac(seed + id alpha) and lane l of chunk i the ray
//    u = frac(seed + (l + 32 (i + 1)) alpha); each 32-ray chunk is folded
//    into the summaries of the bins it hit (one Welford per distinct bin,
//    lowest bin first), and the row stops when every bin the chunk hit has a
//    small enough error.  Seeds are added until every bin's mean over the
//    rows has one; otherwise the result is "not converged".
//
// Rays are traced with trace64, the FP64 trace in the reference's operation
// order, from beam quantiles built on the host, or (CPD_OPT_ROW_TRACE 32, the
// default of the GPU estimator) with the FP32 trace and its FP64 re-trace of
// flagged rays.  The fold is FP64 in the reference's operation order, so with
// the FP64 trace a row is the FP64 oracle's row bit for bit.  A block owns
// one row (four wavefronts tracing when a launch has fewer than 16 rows per
// CU, else one); the sequential fold runs in one wavefront on wave-uniform
// values (each lane computes the same numbers) over summaries kept in LDS.
// The rows of one design are independent, so the host launches blocks of
// rows speculatively and scans them in order, discarding rows past the stop.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <mutex>
#include <thread>
#include <vector>

#include "cpd.h"
#include "cpd_internal.h"
#include "kernels.h"
#include "outer_fold.h"
#include "rows_fold.h"
#include "runtime.h"
#include "welford.h"

namespace cpd {

namespace {

#define AD_TRY(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? CPD_ERR_OUT_OF_MEMORY : CPD_ERR_DEVICE, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                                 \
    } while (0)

// Arguments of a launch of 2D rows.
struct RowArgs {
    RowSlots slots;
    const cpd_design *designs;
    const double *z64;  // beam quantiles, index k - 1
    const float *z32;   // the same in FP32
};

// Per-row constants: the planar chain of this wavelength (ok: some ray can
// leave it), its FP64 record where chain is given (lane 0 writes it), and
// the FP32 trace's hot record.
template <bool FP32>
__device__ __forceinline__ HotRecord row_hot(const cpd_design &g, double lambda, bool &ok,
                                             Chain64 *chain = nullptr) {
    HotRecord hot{};
    LambdaConsts lc;
    make_lambda_consts(g, lambda, lc, chain);
    ok = lc.ok != 0;
    if (FP32) {
        TraceDesign td;
        make_trace_design(g, 1, 64, td);
        hot = make_hot(td, lc);
    }
    return hot;
}

// Rays of a 2D row: FP64 in the reference's operation order, or the FP32
// trace that flags rays near a decision for the exact FP64 re-trace (the
// decisions are the FP64 ones; t is rounded to FP32).  k is the ray's
// 1-based index.  The FP64 traces start from the wavelength's planar-face
// record in LDS (trace64_chain: the same arithmetic as trace64 with the
// ray-independent directions and transmittances computed once per row).
template <bool FP32>
struct Tracer2 {
    static constexpr bool kRetrace = FP32;
    static constexpr size_t kExtraLds = sizeof(Chain64);
    const cpd_design *g;
    const double *z64;
    const float *z32;
    const Chain64 *chain;
    HotRecord hot;
    double lambda;
    uint32_t nbin;
    bool live;

    // Lane 0 writes the chain record when chain_lds is given; the caller
    // synchronises before tracing.
    __device__ Tracer2(const RowArgs &a, uint32_t r, Chain64 *chain_lds)
        : g(&a.designs[a.slots.row_design[r]]), z64(a.z64), z32(a.z32), chain(chain_lds) {
        lambda = a.slots.row_lambda[r];
        nbin = g->bin_count;
        hot = row_hot<FP32>(*g, lambda, live, threadIdx.x == 0 ? chain_lds : nullptr);
    }
    __device__ uint32_t trace(uint32_t k, bool active, double &t, bool &redo) {
        t = 0.0;
        if (!FP32) return active ? trace64_chain(*g, *chain, beam_y64(*g, z64[k - 1]), t) : nbin;
        const float z = active ? z32[k - 1] : 0.f;
        const Trace32Result tr = trace32<false>(hot, hot.y_mean - hot.beam_w * z);
        redo = active && tr.guard;
        if (!active) return nbin;
        if (!tr.hit) return nbin;
        t = tr.t;
        return tr.bin;
    }
    __device__ uint32_t retrace(uint32_t k, double &t) { return trace64_chain(*g, *chain, beam_y64(*g, z64[k - 1]), t); }
};

template <bool FP32, uint32_t W>
__global__ __launch_bounds__(64 * W) void sem_rows_kernel(RowArgs a) {
    extern __shared__ double lds[];
    const uint32_t r = blockIdx.x;
    clear_row(lds, a.slots.nbin_max, a.designs[a.slots.row_design[r]].bin_count);
    Tracer2<FP32> tr(a, r, (Chain64 *)((char *)lds + sem_lds_bytes(a.slots.nbin_max)));
    __syncthreads();
    sem_row<W>(a.slots, r, tr, lds);
}

template <bool FP32, uint32_t W>
__global__ __launch_bounds__(64 * W) void chunk_rows_kernel(RowArgs a) {
    extern __shared__ double lds[];
    const uint32_t r = blockIdx.x;
    clear_row(lds, a.slots.nbin_max, a.designs[a.slots.row_design[r]].bin_count);
    Tracer2<FP32> tr(a, r, (Chain64 *)((char *)lds + chunk_lds_bytes(a.slots.nbin_max, W)));
    __syncthreads();
    chunk_row<W>(a.slots, r, tr, lds);
}

}  // namespace

// The GPU estimator's outer fold (cuda_fitness.rs:366-380) on the device:
// one wavefront per design of the launch, whose rows are contiguous and in id
// order.  Each row updates every bin's Welford summary of p(d) (lanes own
// bins; the host's operations in the host's order, so bit for bit the host's
// values) and the summary of the row's entropy sum H = sum_d plog2p(p(d|l))
// (each row's sum in bin order, the summary in row order).  The design is done when every bin's mean has an SEM
// below 5e-3; info = mean H - sum_d plog2p(mean p(d)) is refreshed after
// every launch.  log2 is the device's, so info agrees with a host fold to
// ~1e-15 while the summaries and the stop decisions are the host's exactly.
struct OuterArgs {
    const double *rows;          // p(d|l) of the launch's rows at row_out offsets
    const uint64_t *row_out;
    const uint64_t *row_rays;
    const uint32_t *row_design;  // design index q of each row
    const cpd_design *designs;   // indexed by q
    W64 *pw;                     // per-bin summaries, design q's at pw_off[q]
    const uint64_t *pw_off;
    W64 *hw;                     // [q] summary of the rows' entropy sums
    uint32_t *done;              // [q]
    double *info;                // [q]
    unsigned long long *rays;    // rays folded, all designs
    uint32_t rows_per;
    double e2;
};

__device__ __forceinline__ double dev_plog2p(double p) { return p > 0.0 ? p * log2(p) : 0.0; }

__global__ __launch_bounds__(64) void outer_fold_kernel(OuterArgs a) {
    __shared__ double hrow[64];
    extern __shared__ double terms[];  // [nbin]
    const uint32_t lane = threadIdx.x;
    const uint32_t r0 = blockIdx.x * a.rows_per, r_end = r0 + a.rows_per;
    const uint32_t q = a.row_design[r0];
    const uint32_t nbin = a.designs[q].bin_count;
    W64 *pw = a.pw + a.pw_off[q];
    W64 h = a.hw[q];
    unsigned long long rays = 0;
    // 64 rows at a time: each lane sums one row's entropy terms in bin order,
    // then every lane folds its bins through the rows in order, and lane 0
    // folds the rows' sums in order
    for (uint32_t c = r0; c < r_end; c += 64) {
        const uint32_t m = min(64u, r_end - c);
        if (lane < m) {
            const double *row = a.rows + a.row_out[c + lane];
            double hr = 0.0;
            for (uint32_t b = 0; b < nbin; ++b) hr += dev_plog2p(row[b]);
            hrow[lane] = hr;
            rays += a.row_rays[c + lane];
        }
        for (uint32_t b = lane; b < nbin; b += 64) {
            W64 w = pw[b];
            for (uint32_t i = 0; i < m; ++i) w_next(w, a.rows[a.row_out[c + i] + b]);
            pw[b] = w;
        }
        __syncthreads();
        if (lane == 0)
            for (uint32_t i = 0; i < m; ++i) w_next(h, hrow[i]);
        __syncthreads();
    }
    for (int off = 32; off > 0; off >>= 1) rays += __shfl_xor(rays, off);
    bool ok = true;
    for (uint32_t b = lane; b < nbin; b += 64) {
        ok = ok && w_sem_le(pw[b], a.e2);
        terms[b] = dev_plog2p(pw[b].mean);
    }
    const bool conv = __builtin_amdgcn_ballot_w64(!ok) == 0;
    __syncthreads();
    if (lane == 0) {
        double h_det = 0.0;
        for (uint32_t b = 0; b < nbin; ++b) h_det = h_det + terms[b];
        a.hw[q] = h;
        a.done[q] = conv ? 1u : 0u;
        a.info[q] = h.mean - h_det;
        atomicAdd(a.rays, rays);
    }
}

// Reciprocals RN(1/k), k = 0..n - 1 (k = 0 holds 0), computed by IEEE
// division on the host; grow-only.
cpd_status Context::ensure_recip(uint32_t n) {
    if (recip_n >= n) return CPD_OK;
    cpd_status st;
    if ((st = recip.reserve((size_t)n * sizeof(double), false, stream)) != CPD_OK) return st;
    if ((st = host_stage.reserve((size_t)n * sizeof(double))) != CPD_OK) return st;
    double *h = (double *)host_stage.ptr;
    h[0] = 0.0;
    for (uint32_t k = 1; k < n; ++k) h[k] = 1.0 / (double)k;
    AD_TRY(hipMemcpyAsync(recip.ptr, h, (size_t)n * sizeof(double), hipMemcpyHostToDevice, stream));
    AD_TRY(hipStreamSynchronize(stream));
    recip_n = n;
    return CPD_OK;
}

// Stages the rows' descriptors (one pinned copy) and reserves their outputs.
cpd_status Context::stage_rows(const std::vector<uint32_t> &row_design, const std::vector<double> &row_lambda,
                               const std::vector<uint64_t> &row_out, const std::vector<uint32_t> &row_cap,
                               uint64_t out_n, uint32_t max_rays, uint32_t nbin_max, double max_err_sq,
                               RowSlots &a) {
    const uint32_t n_rows = (uint32_t)row_design.size();
    cpd_status st;
    if ((st = ensure_recip(max_rays + 40)) != CPD_OK) return st;
    const size_t off_lam = (n_rows * sizeof(uint32_t) + 255) / 256 * 256;
    const size_t off_out = off_lam + (n_rows * sizeof(double) + 255) / 256 * 256;
    const size_t off_cap = off_out + (n_rows * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t bytes = off_cap + n_rows * sizeof(uint32_t);
    if ((st = row_stage.reserve(bytes, false, stream)) != CPD_OK) return st;
    if ((st = row_buf.reserve(out_n * sizeof(double) + n_rows * sizeof(uint64_t) + 256, false, stream)) != CPD_OK)
        return st;
    // pinned staging both ways: the rows' per-bin means are tens of MB
    const size_t off_rays = (out_n * sizeof(double) + 255) / 256 * 256;
    if ((st = row_args_pin.reserve(bytes)) != CPD_OK) return st;
    if ((st = rows_pin.reserve(off_rays + n_rows * sizeof(uint64_t))) != CPD_OK) return st;
    char *h = (char *)row_args_pin.ptr;
    memcpy(h, row_design.data(), n_rows * sizeof(uint32_t));
    memcpy(h + off_lam, row_lambda.data(), n_rows * sizeof(double));
    memcpy(h + off_out, row_out.data(), n_rows * sizeof(uint64_t));
    if (!row_cap.empty()) memcpy(h + off_cap, row_cap.data(), n_rows * sizeof(uint32_t));
    char *d = row_stage.as<char>();
    AD_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
    a = RowSlots{};
    a.row_design = (const uint32_t *)d;
    a.row_lambda = (const double *)(d + off_lam);
    a.row_out = (const uint64_t *)(d + off_out);
    a.row_cap = (const uint32_t *)(d + off_cap);
    a.out = row_buf.as<double>();
    a.row_rays = (uint64_t *)(row_buf.as<char>() + off_rays);
    a.recip = recip.as<double>();
    a.max_err_sq = max_err_sq;
    a.n_rows = n_rows;
    a.max_rays = max_rays;
    a.nbin_max = nbin_max;
    return CPD_OK;
}

// Copies the rows back after their launch and waits.  rows_host points at
// the per-bin means at row_out offsets, rays_host at the rows' ray counts,
// both in the context's pinned buffers until its next launch of rows.
cpd_status Context::finish_rows(const RowSlots &a, uint64_t out_n, const double *&rows_host,
                                const uint64_t *&rays_host) {
    AD_TRY(hipGetLastError());
    if (profile) AD_TRY(hipEventRecord(ev[1], stream));
    const size_t off_rays = (out_n * sizeof(double) + 255) / 256 * 256;
    char *hr = (char *)rows_pin.ptr;
    AD_TRY(hipMemcpyAsync(hr, a.out, out_n * sizeof(double), hipMemcpyDeviceToHost, stream));
    AD_TRY(hipMemcpyAsync(hr + off_rays, a.row_rays, a.n_rows * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    rows_host = (const double *)hr;
    rays_host = (const uint64_t *)(hr + off_rays);
    AD_TRY(hipStreamSynchronize(stream));
    if (profile) {
        float ms = 0.f;
        AD_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        row_kernel_ms += ms;
    }
    return CPD_OK;
}

// Launch one block of 2D rows and copy the rows back.
cpd_status Context::run_rows(bool chunked, bool fp32, const cpd_design *designs_dev,
                             const std::vector<uint32_t> &row_design, const std::vector<double> &row_lambda,
                             const std::vector<uint64_t> &row_out, const std::vector<uint32_t> &row_cap,
                             uint64_t out_n, uint32_t max_rays, uint32_t nbin_max, const double *z64_dev,
                             const float *z32_dev, const double *&rows_host, const uint64_t *&rays_host,
                             RowSlots *dev_slots) {
    const uint32_t n_rows = (uint32_t)row_design.size();
    if (n_rows == 0) return CPD_OK;
    RowArgs a{};
    cpd_status st;
    if ((st = stage_rows(row_design, row_lambda, row_out, row_cap, out_n, max_rays, nbin_max, kMaxErr * kMaxErr,
                         a.slots)) != CPD_OK)
        return st;
    a.designs = designs_dev;
    a.z64 = z64_dev;
    a.z32 = z32_dev;
    const uint32_t waves = sem_waves(n_rows, cu_count);
    const size_t lds = (chunked ? chunk_lds_bytes(nbin_max, waves) : sem_lds_bytes(nbin_max)) +
                       Tracer2<true>::kExtraLds;
    if (profile) AD_TRY(hipEventRecord(ev[0], stream));
    void (*kernel)(RowArgs);
    if (chunked)
        kernel = waves == 1 ? (fp32 ? chunk_rows_kernel<true, 1> : chunk_rows_kernel<false, 1>)
                            : (fp32 ? chunk_rows_kernel<true, kSemWaves> : chunk_rows_kernel<false, kSemWaves>);
    else
        kernel = waves == 1 ? (fp32 ? sem_rows_kernel<true, 1> : sem_rows_kernel<false, 1>)
                            : (fp32 ? sem_rows_kernel<true, kSemWaves> : sem_rows_kernel<false, kSemWaves>);
    hipLaunchKernelGGL(kernel, dim3(n_rows), dim3(64 * waves), lds, stream, a);
    if (dev_slots) {
        // the rows stay on the device for a fold there
        AD_TRY(hipGetLastError());
        if (profile) AD_TRY(hipEventRecord(ev[1], stream));
        *dev_slots = a.slots;
        return CPD_OK;
    }
    return finish_rows(a.slots, out_n, rows_host, rays_host);
}

cpd_status Context::stage_designs(const cpd_design *designs, const std::vector<uint32_t> &valid,
                                  const cpd_design **designs_dev, uint32_t *nbin_max) {
    const size_t n = valid.size();
    cpd_status st;
    if ((st = design_stage.reserve(n * sizeof(cpd_design), false, stream)) != CPD_OK) return st;
    std::vector<cpd_design> h(n);
    uint32_t nb = 1;
    for (size_t q = 0; q < n; ++q) {
        h[q] = designs[valid[q]];
        nb = std::max(nb, h[q].bin_count);
    }
    AD_TRY(hipMemcpyAsync(design_stage.ptr, h.data(), n * sizeof(cpd_design), hipMemcpyHostToDevice, stream));
    AD_TRY(hipStreamSynchronize(stream));
    *designs_dev = design_stage.as<cpd_design>();
    *nbin_max = nb;
    return CPD_OK;
}

// fitness() with the reference's early exits for the designs in `valid`.
cpd_status Context::run_adaptive(const cpd_design *designs, const cpd_sampling *sampling,
                                 const std::vector<uint32_t> &valid, cpd_fitness *out) {
    const size_t n = valid.size();
    if (n == 0) return CPD_OK;
    AD_TRY(hipSetDevice(device));
    uint32_t max_m = 0;
    for (uint32_t i : valid) max_m = std::max(max_m, sampling[i].n_rays);
    cpd_status st;
    if ((st = ensure_beam_table(max_m)) != CPD_OK) return st;
    const cpd_design *ddev;
    uint32_t nbin_max;
    if ((st = stage_designs(designs, valid, &ddev, &nbin_max)) != CPD_OK) return st;
    row_kernel_ms = 0.0;
    double fold_ms = 0.0;
    std::vector<OuterFold> fold(n);
    std::vector<uint64_t> next_j(n, 1), block(n, 4);
    for (size_t q = 0; q < n; ++q) fold[q].reset(designs[valid[q]].bin_count);
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
            const cpd_design &g = designs[valid[q]];
            const cpd_sampling &s = sampling[valid[q]];
            const uint64_t last = std::min<uint64_t>(s.n_wavelengths, next_j[q] + block[q] - 1);
            for (uint64_t j = next_j[q]; j <= last; ++j) {
                row_design.push_back((uint32_t)q);
                row_lambda.push_back(lambda_of(g, qmc_u64(j)));
                row_out.push_back(out_n);
                row_cap.push_back(s.n_rays);
                out_n += g.bin_count;
            }
        }
        if (row_design.empty()) break;
        if ((st = run_rows(false, sem_fp32(), ddev, row_design, row_lambda, row_out, row_cap, out_n, max_m,
                           nbin_max, beam.z64.as<double>(), beam.z32.as<float>(), rows, rays)) != CPD_OK)
            return st;
        // scan each design's rows in order
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
            // a row's time is mostly its serial fold, and a launch rarely fills the
            // GPU: fewer, larger rounds beat the rows traced past the stop
            block[q] = std::min<uint64_t>(block[q] * 16, 1024);
            r = e;
        }
        fold_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - f0).count();
    }
    for (size_t q = 0; q < n; ++q) {
        const cpd_design &g = designs[valid[q]];
        double size, dev;
        size_and_deviation64(g, size, dev);
        out[valid[q]] = cpd_fitness{size, fold[q].info(), dev};
    }
    last = cpd_run_stats{};
    last.rays = total_rays;
    last.trace_ms = row_kernel_ms;  // with CPD_OPT_PROFILE
    last.finalize_ms = fold_ms;     // the host fold over rows
    return CPD_OK;
}

// p(d|lambda) rows at the given wavelengths with the early exit of
// fitness.rs:5-34.
cpd_status Context::rows_adaptive(const cpd_design &g, const double *wavelengths, size_t n_w, uint32_t max_m,
                                  double *out) {
    AD_TRY(hipSetDevice(device));
    cpd_status st;
    if ((st = ensure_beam_table(max_m)) != CPD_OK) return st;
    const cpd_design *ddev;
    uint32_t nbin_max;
    std::vector<uint32_t> one{0};
    if ((st = stage_designs(&g, one, &ddev, &nbin_max)) != CPD_OK) return st;
    std::vector<uint32_t> row_design(n_w, 0), row_cap(n_w, max_m);
    std::vector<double> row_lambda(wavelengths, wavelengths + n_w);
    std::vector<uint64_t> row_out(n_w);
    const double *rows = nullptr;
    const uint64_t *rays = nullptr;
    for (size_t j = 0; j < n_w; ++j) row_out[j] = j * g.bin_count;
    if ((st = run_rows(false, sem_fp32(), ddev, row_design, row_lambda, row_out, row_cap, n_w * g.bin_count,
                       max_m, nbin_max, beam.z64.as<double>(), beam.z32.as<float>(), rows, rays)) != CPD_OK)
        return st;
    memcpy(out, rows, n_w * g.bin_count * sizeof(double));
    return CPD_OK;
}

// cuda_fitness (cuda_fitness.rs:347-423) for the designs in `valid`.
cpd_status Context::run_gpu_fitness(const cpd_design *designs, const std::vector<uint32_t> &valid,
                                    const double *seeds, size_t n_seeds, uint32_t max_n, uint32_t nwarp,
                                    uint32_t max_eval, cpd_fitness *out, cpd_status *per_design) {
    const size_t n = valid.size();
    if (n == 0) return CPD_OK;
    AD_TRY(hipSetDevice(device));
    const uint32_t n_rows_per = (max_n / nwarp) * nwarp;  // grid max_n / nwarp blocks of nwarp warps
    const uint32_t max_evals = max_eval - max_eval % 32;
    const uint32_t table_n = max_evals + 32;
    cpd_status st;
    if ((st = seed_z32.reserve((size_t)table_n * sizeof(float), false, stream)) != CPD_OK) return st;
    if ((st = seed_z64.reserve((size_t)table_n * sizeof(double), false, stream)) != CPD_OK) return st;
    const cpd_design *ddev;
    uint32_t nbin_max;
    if ((st = stage_designs(designs, valid, &ddev, &nbin_max)) != CPD_OK) return st;
    row_kernel_ms = 0.0;
    // the outer fold's state on the device: per-bin summaries, the entropy
    // summary, done flags, information, the ray count and the offsets
    std::vector<uint64_t> pw_off(n);
    uint64_t n_pw = 0;
    for (size_t q = 0; q < n; ++q) {
        pw_off[q] = n_pw;
        n_pw += designs[valid[q]].bin_count;
    }
    const size_t o_hw = n_pw * sizeof(W64), o_done = o_hw + n * sizeof(W64);
    const size_t o_info = (o_done + n * sizeof(uint32_t) + 7) / 8 * 8, o_rays = o_info + n * sizeof(double);
    const size_t o_off = o_rays + 8, fold_bytes = o_off + n * sizeof(uint64_t);
    if ((st = fold_buf.reserve(fold_bytes, false, stream)) != CPD_OK) return st;
    if ((st = host_out.reserve(std::max(n * sizeof(uint64_t), o_off - o_done))) != CPD_OK) return st;
    char *fb = fold_buf.as<char>();
    // zero summaries, flags and count (a fill kernel, not a memset)
    AD_TRY(launch_zero_u32(stream, (unsigned int *)fb, (uint32_t)(o_off / 4)));
    memcpy(host_out.ptr, pw_off.data(), n * sizeof(uint64_t));
    AD_TRY(hipMemcpyAsync(fb + o_off, host_out.ptr, n * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    OuterArgs oa{};
    oa.designs = ddev;
    oa.pw = (W64 *)fb;
    oa.pw_off = (const uint64_t *)(fb + o_off);
    oa.hw = (W64 *)(fb + o_hw);
    oa.done = (uint32_t *)(fb + o_done);
    oa.info = (double *)(fb + o_info);
    oa.rays = (unsigned long long *)(fb + o_rays);
    oa.rows_per = n_rows_per;
    oa.e2 = kMaxErr * kMaxErr;
    std::vector<uint32_t> done(n, 0u);
    bool launched = false;
    std::vector<uint32_t> row_design;
    std::vector<double> row_lambda;
    std::vector<uint64_t> row_out;
    const double *rows = nullptr;
    const uint64_t *rays = nullptr;
    for (size_t si = 0; si < n_seeds; ++si) {
        const uint64_t seed_fx = seed_to_fx(seeds[si]);
        row_design.clear();
        row_lambda.clear();
        row_out.clear();
        uint64_t out_n = 0;
        for (size_t q = 0; q < n; ++q) {
            if (done[q]) continue;
            const cpd_design &g = designs[valid[q]];
            for (uint32_t id = 0; id < n_rows_per; ++id) {
                row_design.push_back((uint32_t)q);
                row_lambda.push_back(lambda_of(g, qmc_seeded(seed_fx, id)));
                row_out.push_back(out_n);
                out_n += g.bin_count;
            }
        }
        if (row_design.empty()) break;
        if ((st = upload_beam_table(stream, seed_fx, table_n, seed_z32.as<float>(), seed_z64.as<double>())) !=
            CPD_OK)
            return st;
        RowSlots slots;
        if ((st = run_rows(true, chunk_fp32(), ddev, row_design, row_lambda, row_out, std::vector<uint32_t>(), out_n,
                           max_evals, nbin_max, seed_z64.as<double>(), seed_z32.as<float>(), rows, rays, &slots)) !=
            CPD_OK)
            return st;
        oa.rows = slots.out;
        oa.row_out = slots.row_out;
        oa.row_rays = slots.row_rays;
        oa.row_design = slots.row_design;
        const uint32_t blocks = (uint32_t)(row_design.size() / n_rows_per);
        hipLaunchKernelGGL(outer_fold_kernel, dim3(blocks), dim3(64), nbin_max * sizeof(double), stream, oa);
        AD_TRY(hipGetLastError());
        // done flags, information and the ray count in one copy
        AD_TRY(hipMemcpyAsync(host_out.ptr, fb + o_done, o_off - o_done, hipMemcpyDeviceToHost, stream));
        AD_TRY(hipStreamSynchronize(stream));
        if (profile) {
            float ms = 0.f;
            AD_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
            row_kernel_ms += ms;
        }
        memcpy(done.data(), host_out.ptr, n * sizeof(uint32_t));
        launched = true;
    }
    std::vector<double> info(n, 0.0);  // no seed: the empty fold's 0 - 0
    uint64_t total_rays = 0;
    if (launched) {
        memcpy(info.data(), (const char *)host_out.ptr + (o_info - o_done), n * sizeof(double));
        memcpy(&total_rays, (const char *)host_out.ptr + (o_rays - o_done), 8);
    }
    for (size_t q = 0; q < n; ++q) {
        const cpd_design &g = designs[valid[q]];
        double size, dev;
        size_and_deviation64(g, size, dev);
        out[valid[q]] = cpd_fitness{size, info[q], dev};
        if (per_design) per_design[valid[q]] = done[q] ? CPD_OK : CPD_ERR_NOT_CONVERGED;
    }
    last = cpd_run_stats{};
    last.rays = total_rays;
    last.trace_ms = row_kernel_ms;  // with CPD_OPT_PROFILE
    return CPD_OK;
}

}  // namespace cpd
