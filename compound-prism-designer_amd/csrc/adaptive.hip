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
// the FP64 trace a row is the FP64 oracle's row bit for bit.  A wavefront owns one row; the sequential fold runs on wave-uniform
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

struct RowArgs {
    const cpd_design *designs;
    const uint32_t *row_design;  // [n_rows] design of each row
    const double *row_lambda;    // [n_rows] its wavelength
    const uint64_t *row_out;     // [n_rows] offset of its bin_count outputs
    const uint32_t *row_cap;     // [n_rows] rays cap of each row (sem rows)
    double *out;                 // p(d|lambda) rows
    uint64_t *row_rays;          // [n_rows] rays folded into the row
    const double *z64;           // beam quantiles, index k - 1
    const float *z32;            // the same in FP32
    uint32_t n_rows;
    uint32_t max_rays;           // rays cap (chunked: a multiple of 32)
    uint32_t nbin_max;
};

__device__ __forceinline__ double readlane_f64(double v, uint32_t l) {
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)u, l);
    const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(u >> 32), l);
    return __longlong_as_double(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}

// Sum over the 32 lanes of each half-wave, butterfly order 16, 8, 4, 2, 1
// (every lane of a half ends with the same total).
__device__ __forceinline__ double half_sum(double v) {
#pragma clang fp contract(off)
    for (int off = 16; off > 0; off >>= 1) v = v + __shfl_xor(v, off);
    return v;
}

// half_sum's total evaluated by one lane from the 32 values v(0..31), in the
// order every lane of half_sum adds in (each butterfly step adds the same two
// partial sums on both partners): s16(i) = v(i) + v(i + 16), s8(i) = s16(i)
// + s16(i + 8), s4(i) = s8(i) + s8(i + 4), then s2 and s1 alike, so the
// result is half_sum's bit for bit.  Each s4 is formed from its own eight
// values behind a compiler barrier, which keeps eight of the 32 live at once.
template <class V>
__device__ __forceinline__ double butterfly_sum(V v) {
#pragma clang fp contract(off)
    double s4[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        asm volatile("" ::: "memory");
        s4[i] = ((v(i) + v(i + 16)) + (v(i + 8) + v(i + 24))) + ((v(i + 4) + v(i + 20)) + (v(i + 12) + v(i + 28)));
    }
    return (s4[0] + s4[2]) + (s4[1] + s4[3]);
}

__device__ __forceinline__ W64 load_w(const double *lds, uint32_t nbin_max, uint32_t b) {
    return W64{lds[b], lds[nbin_max + b], lds[2 * nbin_max + b]};
}
__device__ __forceinline__ void store_w(double *lds, uint32_t nbin_max, uint32_t b, const W64 &w) {
    lds[b] = w.count;
    lds[nbin_max + b] = w.mean;
    lds[2 * nbin_max + b] = w.m2;
}

// One ray of a row: FP64 in the reference's operation order, or the FP32
// trace with its FP64 re-trace of flagged rays (the decisions are the FP64
// ones; t is rounded to FP32).  k is the ray's 1-based index.
template <bool FP32>
__device__ __forceinline__ uint32_t row_ray(const RowArgs &a, const cpd_design &g, double lambda,
                                            const HotRecord &hot, bool active, uint32_t k, double &t) {
    t = 0.0;
    if (!FP32) return active ? trace64(g, lambda, beam_y64(g, a.z64[k - 1]), t) : g.bin_count;
    const float z = active ? a.z32[k - 1] : 0.f;
    const Trace32Result r = trace32<false>(hot, hot.y_mean - hot.beam_w * z);
    uint32_t bin = r.bin;
    t = r.t;
    const bool redo = active && r.guard;
    if (wave_any(redo) && redo) bin = trace64(g, lambda, beam_y64(g, a.z64[k - 1]), t);
    if (!active) {
        bin = g.bin_count;
        t = 0.0;
    }
    return bin;
}

// Per-row constants of the FP32 trace: the planar chain of this wavelength
// (and its FP64 record, where chain is given).
template <bool FP32>
__device__ __forceinline__ HotRecord row_hot(const cpd_design &g, double lambda, bool &ok,
                                             Chain64 *chain = nullptr) {
    HotRecord hot{};
    ok = true;
    if (FP32) {
        TraceDesign td;
        make_trace_design(g, 1, 64, td);
        LambdaConsts lc;
        make_lambda_consts(g, lambda, lc, chain);
        ok = lc.ok != 0;
        hot = make_hot(td, lc);
    }
    return hot;
}

// fitness.rs:5-34, specialised to p_dets_l_wavelength (48-69).
template <bool FP32>
__global__ __launch_bounds__(64) void sem_rows_kernel(RowArgs a) {
#pragma clang fp contract(off)
    extern __shared__ double lds[];
    const uint32_t r = blockIdx.x, lane = threadIdx.x;
    const cpd_design &g = a.designs[a.row_design[r]];
    const uint32_t nbin = g.bin_count;
    for (uint32_t b = lane; b < nbin; b += 64) store_w(lds, a.nbin_max, b, W64{0.0, 0.0, 0.0});
    __syncthreads();
    const double lambda = a.row_lambda[r];
    const uint32_t cap = a.row_cap[r];
    const double max_err_sq = kMaxErr * kMaxErr;
    bool chain_ok;
    const HotRecord hot = row_hot<FP32>(g, lambda, chain_ok);
    double count = 0.0;
    bool stop = false;
    for (uint32_t base = 1; base <= cap && !stop; base += 64) {
        const uint32_t k = base + lane;
        double t;
        const uint32_t bin = chain_ok ? row_ray<FP32>(a, g, lambda, hot, k <= cap, k, t) : nbin;
        if (!chain_ok) t = 0.0;
        const uint32_t n_here = min(64u, cap - base + 1);
        // visit only the lanes whose ray hit a bin, in ray order; count still
        // advances by one per ray, as in the reference's loop
        uint64_t hits = __ballot(lane < n_here && bin < nbin);
        uint32_t consumed = 0;
        while (hits) {
            const uint32_t l = (uint32_t)__builtin_ctzll(hits);
            hits &= hits - 1;
            count += (double)(l + 1 - consumed);
            consumed = l + 1;
            const uint32_t b = __builtin_amdgcn_readlane(bin, l);
            W64 w = load_w(lds, a.nbin_max, b);
            w_skip(w, count);
            w_next(w, readlane_f64(t, l));
            __builtin_amdgcn_wave_barrier();
            if (lane == 0) store_w(lds, a.nbin_max, b, w);
            __builtin_amdgcn_wave_barrier();
            if (w_sem_le(w, max_err_sq)) {
                stop = true;
                break;
            }
        }
        if (!stop) count += (double)(n_here - consumed);
    }
    __syncthreads();
    for (uint32_t b = lane; b < nbin; b += 64) {
        W64 w = load_w(lds, a.nbin_max, b);
        w_skip(w, count);
        a.out[a.row_out[r] + b] = w.mean;
    }
    if (lane == 0) a.row_rays[r] = (uint64_t)count;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    return (1ull << __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u))) - 1ull;
}

// Passes of 64 rays (two chunks) traced ahead of the fold.
constexpr uint32_t kAhead = 4;

// LDS of chunk_rows_kernel after the 3 nbin_max Welford doubles: a half's
// (transmittance, bin) pairs, the traced passes' transmittances, bins and
// re-trace queue, and the wavelength's FP64 chain record.
__host__ __device__ constexpr size_t chunk_lds_bytes(uint32_t nbin_max) {
    return (size_t)3 * nbin_max * sizeof(double) + 32 * 16 + kAhead * 64 * (sizeof(double) + 2 * sizeof(uint32_t)) +
           sizeof(Chain64);
}

// kernel.rs:89-193: 32-ray chunks folded in order, each chunk's
// (32, mean, m2) over its 32 lanes into every bin it hit, with the SEM test
// after each chunk.  A wavefront traces kAhead passes of two chunks (lanes
// 0-31 and 32-63) before folding them: the FP32 trace's flagged rays of all
// the passes are queued and re-traced together in FP64 through the
// wavelength's chain record, and rays traced beyond the chunk where the row
// stops are not folded.  In the fold, the two lowest bins of a chunk are
// folded one at a time with half-wave butterflies; when more remain, every
// lane that hit one of them evaluates the same butterfly sums for its own
// bin from LDS, so a chunk spread over many bins costs one pass instead of
// one per bin.
template <bool FP32>
__global__ __launch_bounds__(64) void chunk_rows_kernel(RowArgs a) {
#pragma clang fp contract(off)
    extern __shared__ double lds[];
    double2 *tpair = (double2 *)(lds + 3 * a.nbin_max);  // bin ~0u: nothing left to fold
    double *tv = (double *)(tpair + 32);                 // [kAhead * 64] pass p, lane at p * 64 + lane
    uint32_t *bv = (uint32_t *)(tv + kAhead * 64);
    uint32_t *queue = bv + kAhead * 64;
    Chain64 *chain = (Chain64 *)(queue + kAhead * 64);
    const uint32_t r = blockIdx.x, lane = threadIdx.x;
    const cpd_design &g = a.designs[a.row_design[r]];
    const uint32_t nbin = g.bin_count;
    for (uint32_t b = lane; b < nbin; b += 64) store_w(lds, a.nbin_max, b, W64{0.0, 0.0, 0.0});
    const double lambda = a.row_lambda[r];
    const double max_err_sq = kMaxErr * kMaxErr;
    const uint32_t iters = a.max_rays / 32;
    const uint32_t half = lane >> 5, l = lane & 31;
    bool chain_ok;
    const HotRecord hot = row_hot<FP32>(g, lambda, chain_ok, lane == 0 ? chain : nullptr);
    __syncthreads();
    double count = 0.0;
    bool stop = false;
    for (uint32_t i0 = 0; i0 < iters && !stop; i0 += 2 * kAhead) {
        const uint32_t passes = min(kAhead, (iters - i0 + 1) / 2);
        // trace: ray k = l + 32 (c + 1) of chunk c = i0 + 2 p + half
        uint32_t queued = 0;
        for (uint32_t p = 0; p < passes; ++p) {
            const uint32_t c = i0 + 2 * p + half;
            const bool active = chain_ok && c < iters;
            const uint32_t k = l + 32 * (c + 1);
            double t = 0.0;
            uint32_t bin = nbin;
            bool redo = false;
            if (FP32) {
                const float z = active ? a.z32[k - 1] : 0.f;
                const Trace32Result tr = trace32<false>(hot, hot.y_mean - hot.beam_w * z);
                redo = active && tr.guard;
                if (active && !tr.guard) {
                    bin = tr.bin;
                    t = tr.bin < nbin ? (double)tr.t : 0.0;
                }
            } else if (active) {
                bin = trace64(g, lambda, beam_y64(g, a.z64[k - 1]), t);
            }
            const uint32_t slot = p * 64 + lane;
            tv[slot] = t;
            bv[slot] = bin;
            if (FP32) {
                const uint64_t mask = __builtin_amdgcn_ballot_w64(redo);
                if (redo) queue[queued + __popcll(mask & lanemask_lt())] = slot;
                queued += __popcll(mask);
            }
        }
        // FP64 re-trace of the queued rays, 64 at a time
        if (FP32) {
            for (uint32_t q0 = 0; q0 < queued; q0 += 64) {
                __builtin_amdgcn_wave_barrier();
                if (q0 + lane < queued) {
                    const uint32_t slot = queue[q0 + lane];
                    const uint32_t c = i0 + 2 * (slot >> 6) + ((slot >> 5) & 1);
                    const uint32_t k = (slot & 31) + 32 * (c + 1);
                    double t;
                    const uint32_t bin = trace64_chain(g, *chain, beam_y64(g, a.z64[k - 1]), t);
                    tv[slot] = bin < nbin ? t : 0.0;
                    bv[slot] = bin;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
        // fold the passes' chunks in order
        for (uint32_t p = 0; p < passes && !stop; ++p) {
            const uint32_t i = i0 + 2 * p;
            const double t = tv[p * 64 + lane];
            const uint32_t bin = bv[p * 64 + lane];
            for (uint32_t h = 0; h < 2 && i + h < iters; ++h) {
                count += 32.0;
                const bool mine = half == h;
                bool pending = mine && bin < nbin;
                bool finished = __ballot(pending) != 0;
                for (int serial = 0; serial < 2 && __ballot(pending) != 0; ++serial) {
                    const uint32_t bmin = wave_min_u32(pending ? bin : 0xffffffffu);
                    const double tb = (mine && bin == bmin) ? t : 0.0;
                    const double wmean = readlane_f64(half_sum(tb), 32 * h) / 32.0;
                    const double dev = mine ? tb - wmean : 0.0;
                    const double wm2 = readlane_f64(half_sum(dev * dev), 32 * h);
                    const W64 w = w_add(load_w(lds, a.nbin_max, bmin), W64{32.0, wmean, wm2});
                    __builtin_amdgcn_wave_barrier();
                    if (lane == 0) store_w(lds, a.nbin_max, bmin, w);
                    __builtin_amdgcn_wave_barrier();
                    finished = finished && w_sem_le(w, max_err_sq);
                    if (bin == bmin) pending = false;
                }
                if (__ballot(pending) != 0) {
                    // the remaining bins at once; lanes that share a bin
                    // compute and store the same record
                    __builtin_amdgcn_wave_barrier();
                    if (mine) tpair[l] = make_double2(t, __longlong_as_double(pending ? bin : ~0u));
                    __builtin_amdgcn_wave_barrier();
                    bool ok = true;
                    if (pending) {
                        auto v = [&](int j) {
                            const double2 pr = tpair[j];
                            return (uint32_t)__double_as_longlong(pr.y) == bin ? pr.x : 0.0;
                        };
                        const double wmean = butterfly_sum(v) / 32.0;
                        const double wm2 = butterfly_sum([&](int j) {
                            const double dev = v(j) - wmean;
                            return dev * dev;
                        });
                        const W64 w = w_add(load_w(lds, a.nbin_max, bin), W64{32.0, wmean, wm2});
                        store_w(lds, a.nbin_max, bin, w);
                        ok = w_sem_le(w, max_err_sq);
                    }
                    __builtin_amdgcn_wave_barrier();
                    finished = finished && __ballot(pending && !ok) == 0;
                    pending = false;
                }
                if (finished) {
                    stop = true;
                    break;
                }
            }
        }
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    for (uint32_t b = lane; b < nbin; b += 64) {
        W64 w = load_w(lds, a.nbin_max, b);
        w_skip(w, count);
        a.out[a.row_out[r] + b] = w.mean;
    }
    if (lane == 0) a.row_rays[r] = (uint64_t)count;
}

// Per-design state of the outer fold over rows (fitness.rs:89-114,
// cuda_fitness.rs:366-380).
struct OuterFold {
    std::vector<W64> p_dets;
    W64 h{0.0, 0.0, 0.0};
    uint64_t rows = 0, rays = 0;
    bool done = false;

    void reset(uint32_t nbin) {
        p_dets.assign(nbin, W64{0.0, 0.0, 0.0});
        h = W64{0.0, 0.0, 0.0};
        rows = rays = 0;
        done = false;
    }
    // Folds one row; returns whether every bin's mean is precise enough.
    bool fold(const double *row) {
        double h_row = 0.0;
        for (size_t b = 0; b < p_dets.size(); ++b) {
            w_next(p_dets[b], row[b]);
            h_row += plog2p(row[b]);
        }
        w_next(h, h_row);
        ++rows;
        return converged();
    }
    bool converged() const {
        const double e2 = kMaxErr * kMaxErr;
        for (const W64 &w : p_dets)
            if (!w_sem_le(w, e2)) return false;
        return true;
    }
    // H(D) - H(D|Lambda) (fitness.rs:115-122)
    double info() const {
        double h_det = 0.0;
        for (const W64 &w : p_dets) h_det = h_det + plog2p(w.mean);
        return h.mean - h_det;
    }
};

// Runs body(i) for i in [0, n) on up to 16 host threads.
template <class F>
void parallel_for(size_t n, F body) {
    const size_t hw = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
    const size_t threads = std::min(hw, n / 64 + 1);
    if (threads <= 1) {
        for (size_t i = 0; i < n; ++i) body(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (size_t t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) body(i);
        });
    for (auto &th : pool) th.join();
}

}  // namespace

// Launch one block of rows and copy the rows back.  rows_host points at
// n_rows x bin_count values at row_out offsets, rays_host at the rows' ray
// counts, both in the context's pinned buffers until its next run_rows.
cpd_status Context::run_rows(bool chunked, bool fp32, const cpd_design *designs_dev,
                             const std::vector<uint32_t> &row_design, const std::vector<double> &row_lambda,
                             const std::vector<uint64_t> &row_out, const std::vector<uint32_t> &row_cap,
                             uint64_t out_n, uint32_t max_rays, uint32_t nbin_max, const double *z64_dev,
                             const float *z32_dev, const double *&rows_host, const uint64_t *&rays_host) {
    const uint32_t n_rows = (uint32_t)row_design.size();
    if (n_rows == 0) return CPD_OK;
    const size_t off_lam = (n_rows * sizeof(uint32_t) + 255) / 256 * 256;
    const size_t off_out = off_lam + (n_rows * sizeof(double) + 255) / 256 * 256;
    const size_t off_cap = off_out + (n_rows * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t bytes = off_cap + n_rows * sizeof(uint32_t);
    cpd_status st;
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
    RowArgs a{};
    a.designs = designs_dev;
    a.row_design = (const uint32_t *)d;
    a.row_lambda = (const double *)(d + off_lam);
    a.row_out = (const uint64_t *)(d + off_out);
    a.row_cap = (const uint32_t *)(d + off_cap);
    a.out = row_buf.as<double>();
    a.row_rays = (uint64_t *)(row_buf.as<char>() + (out_n * sizeof(double) + 255) / 256 * 256);
    a.z64 = z64_dev;
    a.z32 = z32_dev;
    a.n_rows = n_rows;
    a.max_rays = max_rays;
    a.nbin_max = nbin_max;
    const size_t lds = chunked ? chunk_lds_bytes(nbin_max) : (size_t)3 * nbin_max * sizeof(double);
    if (profile) AD_TRY(hipEventRecord(ev[0], stream));
    if (chunked && fp32)
        hipLaunchKernelGGL(chunk_rows_kernel<true>, dim3(n_rows), dim3(64), lds, stream, a);
    else if (chunked)
        hipLaunchKernelGGL(chunk_rows_kernel<false>, dim3(n_rows), dim3(64), lds, stream, a);
    else if (fp32)
        hipLaunchKernelGGL(sem_rows_kernel<true>, dim3(n_rows), dim3(64), lds, stream, a);
    else
        hipLaunchKernelGGL(sem_rows_kernel<false>, dim3(n_rows), dim3(64), lds, stream, a);
    AD_TRY(hipGetLastError());
    if (profile) AD_TRY(hipEventRecord(ev[1], stream));
    char *hr = (char *)rows_pin.ptr;
    AD_TRY(hipMemcpyAsync(hr, a.out, out_n * sizeof(double), hipMemcpyDeviceToHost, stream));
    AD_TRY(hipMemcpyAsync(hr + off_rays, a.row_rays, n_rows * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
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
            block[q] = std::min<uint64_t>(block[q] * 4, 1024);
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
        if ((st = run_rows(true, chunk_fp32(), ddev, row_design, row_lambda, row_out, std::vector<uint32_t>(), out_n,
                           max_evals, nbin_max, seed_z64.as<double>(), seed_z32.as<float>(), rows, rays)) != CPD_OK)
            return st;
        const auto f0 = std::chrono::steady_clock::now();
        for (size_t r = 0; r < row_design.size(); ++r) total_rays += rays[r];
        // each design's rows are contiguous and in id order; designs fold in parallel
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
        const cpd_design &g = designs[valid[q]];
        double size, dev;
        size_and_deviation64(g, size, dev);
        out[valid[q]] = cpd_fitness{size, fold[q].info(), dev};
        if (per_design) per_design[valid[q]] = fold[q].done ? CPD_OK : CPD_ERR_NOT_CONVERGED;
    }
    last = cpd_run_stats{};
    last.rays = total_rays;
    last.trace_ms = row_kernel_ms;  // with CPD_OPT_PROFILE
    last.finalize_ms = fold_ms;     // the host fold over rows
    return CPD_OK;
}

}  // namespace cpd
