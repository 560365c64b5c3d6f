// kernels.hip -- gfx950 kernels of the compound-prism fitness evaluator.
//
// trace_kernel: one 64-lane wavefront per work item (design, wavelength
// sample, tile of rays).  Lane l traces rays k = k0 + l, k0 + 64 + l, ... of
// the tile in FP32 (trace32) and adds round(t * 2^22) to its own column of a
// [bin][column] LDS histogram, so no two lanes ever contend for a bin.  Rays
// whose FP32 outcome lies within the error margin of a hit/miss or bin
// decision go to a per-wave LDS queue and are re-traced 64 at a time in FP64
// (trace64, the reference's operation order).  At the end of an item the
// columns are summed (skewed so the 64 lanes hit 32 distinct banks) into
// 64-bit integer sums, which are added to the design's per-(wavelength, bin)
// accumulators with integer atomics -- exact, so the result is the same
// whatever the tile order or the batch a design shares the GPU with.
//
// finalize_kernel: one wavefront per design turns the integer sums into
// p(d|l) = sum / (2^22 M) (the bin of the last ray divides by M + 1, as the
// reference's Welford padding does, src/fitness.rs:20-32), and reduces
// H(D|L) = mean_l sum_d p log2 p and H(D) = sum_d pbar log2 pbar in FP64 with
// Neumaier-compensated sums and a fixed reduction order
// (src/fitness.rs:79-123).  It clears the accumulators it read, so the next
// launch starts from zero without a memset.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace cpd {

namespace {

__device__ __forceinline__ uint32_t quant32(float t) { return __float2uint_rn(t * (float)kQuantScale); }
__device__ __forceinline__ uint32_t quant64(double t) { return __double2uint_rn(t * kQuantScale); }

__device__ __forceinline__ uint64_t lanemask_lt() {
    const uint32_t lane = threadIdx.x;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Copy of a wave-uniform record into scalar registers.
template <class T>
__device__ __forceinline__ T to_sgprs(const T &v) {
    static_assert(sizeof(T) % 4 == 0, "record of 32-bit words");
    T out;
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&v);
    uint32_t *dst = reinterpret_cast<uint32_t *>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
    return out;
}

__device__ __forceinline__ uint32_t find_design(const uint64_t *begin, uint32_t n, uint64_t item) {
    uint32_t lo = 0, hi = n;  // begin[lo] <= item < begin[hi]
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (begin[mid] <= item)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Neumaier-compensated FP64 sum.
struct CSum {
    double s = 0.0, c = 0.0;
    __device__ void add(double x) {
#pragma clang fp contract(off)
        double t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ double value() const { return s + c; }
};

__device__ __forceinline__ double plog2p(double p) { return p > 0.0 ? p * log2(p) : 0.0; }

struct TileCtx {
    const cpd_design *g;
    const TraceDesign *td;
    LambdaConsts lc;
    uint32_t k_begin, k_end, n_rays, nbin, cols;
    uint32_t *hist;
    unsigned long long *row;  // the wavefront's u64 sums per bin
    uint32_t *queue;
    uint32_t *last_bin;
    const float *z32;
    const double *z64;
};

// FP64 re-trace of `count` (<= 64) queued rays starting at ring slot `head`.
__device__ void drain_queue(const TileCtx &c, uint32_t head, uint32_t count) {
    const uint32_t lane = threadIdx.x;
    if (lane < count) {
        const uint32_t k = c.queue[(head + lane) & (kQueueCap - 1)];
        const double y = beam_y64(*c.g, c.z64[k - 1]);
        double t;
        const uint32_t b = trace64(*c.g, c.lc.lambda, y, t);
        if (b < c.nbin) atomicAdd(&c.hist[b * c.cols + (lane & (c.cols - 1))], quant64(t));
        if (k == c.n_rays) *c.last_bin = b;
    }
}

// Adds each bin's u32 columns (skewed, conflict-free order) to the u64 row and
// clears them.  Lane b handles bins b, b + 64, ...
__device__ void fold_columns(uint32_t *hist, unsigned long long *row, uint32_t nbin, uint32_t cols) {
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbin; b += 64) {
        uint32_t *cell = hist + b * cols;
        uint64_t sum = 0;
        for (uint32_t q = 0; q < cols; ++q) {
            const uint32_t cc = (q + b) & (cols - 1);
            sum += cell[cc];
            cell[cc] = 0;
        }
        row[b] += sum;
    }
    __syncthreads();
}

template <bool COUNT>
__device__ void tile_loop(const TileCtx &c, unsigned long long &ops) {
    const uint32_t lane = threadIdx.x;
    const uint32_t col = lane & (c.cols - 1);
    // register copies of the per-item records: the loop stores through
    // last_bin, so loads through c.td would be repeated every iteration
    const HotRecord hot = to_sgprs(make_hot(*c.td, c.lc));
    const float *__restrict__ z32 = c.z32;
    uint32_t *__restrict__ hist = c.hist;
    const uint32_t k_end = c.k_end, n_rays = c.n_rays, nbin = c.nbin, cols = c.cols;
    uint32_t head = 0, tail = 0;  // ring of queued ray indices (wave-uniform)
    const uint32_t fold = fold_rays(cols);
    uint32_t since_fold = 0;
    for (uint32_t base = c.k_begin; base < k_end; base += 64) {
        const uint32_t k = base + lane;
        const bool active = k < k_end;
        const float z = active ? z32[k - 1] : 0.f;
        const float y = hot.y_mean - hot.beam_w * z;
        const Trace32Result r = trace32<COUNT>(hot, y);
        if (COUNT && active) ops += r.ops;
        const bool flagged = active && r.guard;
        if (active && !r.guard) {
            if (r.bin < nbin) atomicAdd(&hist[r.bin * cols + col], quant32(r.t));
            if (k == n_rays) *c.last_bin = r.bin;
        }
        const uint64_t mask = __ballot(flagged);
        if (mask) {
            if (flagged) c.queue[(tail + __popcll(mask & lanemask_lt())) & (kQueueCap - 1)] = k;
            tail += __popcll(mask);
            if (tail - head >= 64) {
                __builtin_amdgcn_wave_barrier();
                drain_queue(c, head, 64);
                head += 64;
            }
        }
        since_fold += 64;
        if (since_fold >= fold) {
            fold_columns(hist, c.row, nbin, cols);
            since_fold = 0;
        }
    }
    if (tail != head) {
        __builtin_amdgcn_wave_barrier();
        drain_queue(c, head, tail - head);
    }
}

template <bool COUNT>
__global__ __launch_bounds__(64) void trace_kernel(BatchArgs a) {
    extern __shared__ uint32_t lds[];
    uint32_t *hist = lds;
    uint32_t *queue = lds + a.hist_words;
    unsigned long long *row = (unsigned long long *)(lds + trace_row_offset_words(a.hist_words));
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < a.hist_words; i += 64) hist[i] = 0;
    for (uint32_t i = lane; i < a.row_words; i += 64) row[i] = 0;
    __syncthreads();
    unsigned long long ops = 0;
    uint64_t item = blockIdx.x;
    for (;;) {
        if (a.work_counter) {
            uint32_t v = 0;
            if (lane == 0) v = atomicAdd(a.work_counter, 1u);
            item = (uint64_t)__builtin_amdgcn_readfirstlane(v);
        }
        if (item >= a.n_items) break;
        const uint32_t e = find_design(a.item_begin, a.n_designs, item);
        const uint32_t d = a.slot ? a.slot[e] : e;
        const cpd_design *g = a.designs + d;
        const TraceDesign *td = a.traces + d;
        const uint32_t M = td->n_rays, nbin = td->nbin;
        const uint32_t n_tiles = td->n_tiles, tile_rays = td->tile_rays, cols = td->hist_cols;
        const uint64_t local = item - a.item_begin[e] + (a.item_offset ? a.item_offset[e] : 0);
        const uint32_t j = (uint32_t)(local / n_tiles), tile = (uint32_t)(local % n_tiles);

        TileCtx c;
        c.g = g;
        c.td = td;
        const double lambda = a.wavelengths ? a.wavelengths[j] : lambda_of(*g, qmc_u64(j + 1));
        make_lambda_consts(*g, lambda, c.lc);
        c.k_begin = tile * tile_rays + 1;
        c.k_end = min(c.k_begin + tile_rays, M + 1);
        c.n_rays = M;
        c.nbin = nbin;
        c.cols = cols;
        c.hist = hist;
        c.row = row;
        c.queue = queue;
        c.last_bin = a.last_bin + a.lb_begin[d] + j;
        c.z32 = a.z32;
        c.z64 = a.z64;
        if (!c.lc.ok) {
            // the entrance face reflects every ray of this wavelength
            if (c.k_end == M + 1 && lane == 0) *c.last_bin = nbin;
        } else {
            tile_loop<COUNT>(c, ops);
            // last fold and flush: row + columns into the global sums
            __syncthreads();
            unsigned long long *acc = a.acc + a.acc_begin[d] + (uint64_t)j * nbin;
            for (uint32_t b = lane; b < nbin; b += 64) {
                uint32_t *cell = hist + b * cols;
                uint64_t sum = row[b];
                row[b] = 0;
                for (uint32_t q = 0; q < cols; ++q) {
                    const uint32_t cc = (q + b) & (cols - 1);
                    sum += cell[cc];
                    cell[cc] = 0;
                }
                if (sum) atomicAdd(acc + b, (unsigned long long)sum);
            }
            __syncthreads();
        }
        if (!a.work_counter) break;
    }
    if (COUNT) {
        for (int off = 32; off > 0; off >>= 1) ops += __shfl_xor(ops, off);
        if (lane == 0 && ops) atomicAdd(a.ops, ops);
    }
}

__global__ __launch_bounds__(64) void finalize_kernel(FinalizeArgs a) {
    const uint32_t e = blockIdx.x;
    const uint32_t d = a.slot ? a.slot[e] : e;
    const uint32_t lane = threadIdx.x;
    const cpd_design &g = a.designs[d];
    const TraceDesign &td = a.traces[d];
    const uint32_t N = td.n_lambda, M = td.n_rays, nbin = td.nbin;
    unsigned long long *acc = a.acc + a.acc_begin[d];
    uint32_t *lb = a.last_bin + a.lb_begin[d];
    double *rows = a.rows ? a.rows + a.row_begin[d] : nullptr;
    const double inv_scale = 1.0 / kQuantScale;
    CSum h_cond, h_det;
    for (uint32_t b = lane; b < nbin; b += 64) {
        CSum pb;
        for (uint32_t j = 0; j < N; ++j) {
            const uint64_t s = acc[(uint64_t)j * nbin + b];
            const double div = lb[j] == b ? (double)M + 1.0 : (double)M;
            const double p = (double)s * inv_scale / div;
            h_cond.add(plog2p(p));
            pb.add(p);
            if (rows) rows[(uint64_t)j * nbin + b] = p;
            acc[(uint64_t)j * nbin + b] = 0;
        }
        h_det.add(plog2p(pb.value() / (double)N));
    }
    // fixed-order butterfly over the wavefront on (sum, compensation) pairs
    double hs = h_cond.s, hc = h_cond.c, ds = h_det.s, dc = h_det.c;
    for (int off = 32; off > 0; off >>= 1) {
        CSum x;
        x.s = hs;
        x.c = hc;
        x.add(__shfl_xor(hs, off));
        x.c += __shfl_xor(hc, off);
        hs = x.s;
        hc = x.c;
        CSum y;
        y.s = ds;
        y.c = dc;
        y.add(__shfl_xor(ds, off));
        y.c += __shfl_xor(dc, off);
        ds = y.s;
        dc = y.c;
    }
    __syncthreads();
    for (uint32_t j = lane; j < N; j += 64) lb[j] = nbin;
    if (lane == 0 && a.out) {
        double size, deviation;
        size_and_deviation64(g, size, deviation);
        cpd_fitness f;
        f.size = size;
        f.info = (hs + hc) / (double)N - (ds + dc);
        f.deviation = deviation;
        a.out[a.out_stride_slot ? d : e] = f;
    }
}

__global__ void zero_u32_kernel(unsigned int *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void propagate_kernel(const cpd_design *g, const TraceDesign *td, const float *wcdf, const float *rcdf,
                                 uint32_t n, uint32_t *bin, float *t) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    LambdaConsts lc;
    const double lambda = lambda_of(*g, (double)wcdf[i]);
    make_lambda_consts(*g, lambda, lc);
    const double z = norminv64((double)rcdf[i]);
    Trace32Result r{td->nbin, 0.f, false, 0};
    if (lc.ok) {
        const float y = td->y_mean - td->beam_w * (float)z;
        r = trace32<false>(make_hot(*td, lc), y);
    }
    if (!lc.ok || r.guard) {
        double t64;
        const uint32_t b = trace64(*g, lambda, beam_y64(*g, z), t64);
        r.bin = b;
        r.t = (float)t64;
    }
    bin[i] = r.bin;
    t[i] = r.bin < td->nbin ? r.t : 0.f;
}

// FP32 trace records of n designs from their FP64 records and per-slot
// sampling (one thread per design).
__global__ void build_traces_kernel(const cpd_design *designs, const cpd_sampling *sampling, uint32_t n,
                                    uint32_t tile_cap, TraceDesign *traces) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    make_trace_design(designs[i], sampling[i].n_wavelengths, sampling[i].n_rays, traces[i], tile_cap);
}

}  // namespace

hipError_t launch_build_traces(hipStream_t s, const cpd_design *designs, const cpd_sampling *sampling, uint32_t n,
                               uint32_t tile_cap, TraceDesign *traces) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(build_traces_kernel, dim3((n + 63) / 64), dim3(64), 0, s, designs, sampling, n, tile_cap,
                       traces);
    return hipGetLastError();
}

size_t trace_lds_bytes(uint32_t hist_words, uint32_t row_words) {
    return (size_t)trace_row_offset_words(hist_words) * sizeof(uint32_t) + (size_t)row_words * sizeof(uint64_t);
}

hipError_t launch_trace(hipStream_t s, const BatchArgs &a, uint32_t grid, bool count_ops, hipEvent_t start,
                        hipEvent_t stop) {
    if (grid == 0 || a.n_items == 0) return hipSuccess;
    const size_t lds = trace_lds_bytes(a.hist_words, a.row_words);
    if (!start && !stop) {
        // plain launch: the form a stream capture records as a kernel node
        if (count_ops)
            hipLaunchKernelGGL(trace_kernel<true>, dim3(grid), dim3(64), lds, s, a);
        else
            hipLaunchKernelGGL(trace_kernel<false>, dim3(grid), dim3(64), lds, s, a);
        return hipGetLastError();
    }
    if (count_ops)
        hipExtLaunchKernelGGL(trace_kernel<true>, dim3(grid), dim3(64), lds, s, start, stop, 0, a);
    else
        hipExtLaunchKernelGGL(trace_kernel<false>, dim3(grid), dim3(64), lds, s, start, stop, 0, a);
    return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t s, const FinalizeArgs &a) {
    if (a.n_designs == 0) return hipSuccess;
    hipLaunchKernelGGL(finalize_kernel, dim3(a.n_designs), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_zero_u32(hipStream_t s, unsigned int *p, uint32_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(zero_u32_kernel, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
    return hipGetLastError();
}

hipError_t launch_propagate(hipStream_t s, const cpd_design *g, const TraceDesign *td, const float *wcdf,
                            const float *rcdf, uint32_t n, uint32_t *bin, float *t) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(propagate_kernel, dim3((n + 255) / 256), dim3(256), 0, s, g, td, wcdf, rcdf, n, bin, t);
    return hipGetLastError();
}

}  // namespace cpd
