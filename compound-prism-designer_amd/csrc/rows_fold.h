// rows_fold.h -- the two adaptive row estimators as device templates over a
// ray tracer, shared by the 2D rows (adaptive.hip) and the 3D rows
// (rows3.hip).
//
// A block owns one row (a design at one wavelength) and keeps the row's
// per-bin Welford summaries (count, mean, m2) in LDS.
//  * sem_row (W wavefronts tracing, the first folding): p_dets_l_wavelength with the early exit
//    of src/fitness.rs:5-34 -- rays k = 1, 2, ... traced 64 per wavefront and
//    folded one at a time, in order, into the summary of the bin each hit, padded
//    with zeros (skip); the row stops as soon as the bin just updated has a
//    standard error of the mean below 5e-3.
//  * chunk_row (W wavefronts tracing, the first folding): the reference's GPU
//    estimator, src/spectrometer/kernel.rs:89-193
//    -- 32-ray chunks folded in order, each chunk's (32, mean, m2) over its
//    32 lanes into every bin it hit, lowest bin first, with the SEM test after
//    each chunk.
//
// A tracer Tr supplies, for the row it was built for:
//   uint32_t nbin;                 bins of the design
//   bool live;                     false when no ray can reach the detector
//   uint32_t trace(k, active, t, redo)   ray k's bin (nbin on a miss or when
//                                  inactive) and transmittance (0 then); with
//                                  kRetrace it may flag the ray instead (redo)
//   uint32_t retrace(k, t)         the exact re-trace of a flagged ray
//   static constexpr bool kRetrace;
//   static constexpr size_t kExtraLds;   LDS bytes the tracer keeps after
//                                  sem_lds_bytes or chunk_lds_bytes (a chain
//                                  record, say)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "trace_core.h"
#include "welford.h"

namespace cpd {

// Rows of one launch: which design and wavelength each block traces, where
// its bin_count means go, and the rays it may fold.
struct RowSlots {
    const uint32_t *row_design;  // [n_rows] index into the launch's designs
    const double *row_lambda;    // [n_rows] its wavelength
    const uint64_t *row_out;     // [n_rows] offset of its bin_count outputs
    const uint32_t *row_cap;     // [n_rows] rays cap of each row (sem rows)
    double *out;                 // p(d|lambda) rows
    uint64_t *row_rays;          // [n_rows] rays folded into the row
    const double *recip;         // [k] = RN(1/k), k <= max_rays + 32: the folds' integer divisors
    double max_err_sq;           // SEM^2 threshold of the early exits (0: none)
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

// Minimum over the 32 lanes of each half-wave.
__device__ __forceinline__ uint32_t half_min_u32(uint32_t v) {
    for (int off = 16; off > 0; off >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, off));
    return v;
}

__device__ __forceinline__ uint64_t lanemask_below() {
    return (1ull << __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u))) - 1ull;
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

// Clears the row's summaries (before the tracer is built; the caller
// synchronises after).
__device__ __forceinline__ void clear_row(double *lds, uint32_t nbin_max, uint32_t nbin) {
    for (uint32_t b = threadIdx.x; b < nbin; b += blockDim.x) store_w(lds, nbin_max, b, W64{0.0, 0.0, 0.0});
}

// Pads every summary to the rays folded and writes the row's means.
__device__ __forceinline__ void finish_row(const RowSlots &a, uint32_t r, const double *lds, uint32_t nbin,
                                           double count) {
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbin; b += blockDim.x) {
        W64 w = load_w(lds, a.nbin_max, b);
        w_skip(w, count);
        a.out[a.row_out[r] + b] = w.mean;
    }
    if (threadIdx.x == 0) a.row_rays[r] = (uint64_t)count;
}

// One hit of the per-ray fold: skip to ray k (the count so far), then the
// sample; rk and rk1 are the reciprocals of k and k + 1.
__device__ __forceinline__ void fold_hit_r(W64 &w, double k, double t, double rk, double rk1, bool track_m2) {
    if (track_m2) {
        w_skip_r(w, k, rk);
        w_next_r(w, t, rk1);
    } else {
        w_skip_mean_r(w, k, rk);
        w_next_mean_r(w, t, rk1);
    }
}
// The same with the reciprocals from the table.
__device__ __forceinline__ void fold_hit(W64 &w, double k, double t, const double *recip, bool track_m2) {
    fold_hit_r(w, k, t, recip[(uint32_t)k], recip[(uint32_t)k + 1], track_m2);
}

// Most wavefronts of a sem_row block (its template argument W): each traces
// one pass of 64 rays, and the first folds the passes in ray order.
constexpr uint32_t kSemWaves = 4;

// Wavefronts per row of a sem_row launch of n_rows rows: one when the rows
// alone give every SIMD four wavefronts (the other wavefronts of a block
// would wait through its fold), kSemWaves when fewer rows leave the GPU short
// of wavefronts.
inline uint32_t sem_waves(size_t n_rows, int cu_count) {
    return n_rows >= (size_t)16 * (size_t)(cu_count > 1 ? cu_count : 1) ? 1u : kSemWaves;
}

// LDS of sem_row after the 3 nbin_max Welford doubles: the traced passes'
// transmittances and bins, and the fold's stop flag and ray count; the
// tracer's own bytes follow.
__host__ __device__ constexpr size_t sem_lds_bytes(uint32_t nbin_max) {
    return (size_t)3 * nbin_max * sizeof(double) + kSemWaves * 64 * (sizeof(double) + sizeof(uint32_t)) +
           2 * sizeof(double);
}

// fitness.rs:5-34, specialised to p_dets_l_wavelength (48-69).  The block's
// wavefronts trace as many passes at once; the first wavefront folds them.
// Rays k are folded in order; hits of different bins touch independent
// summaries, so each pass folds every bin's hits as its own chain at once
// (the lowest lane of each bin leads it, in ray order).  The
// early exit stops at the first hit whose bin's SEM passes, which is the
// earliest such hit over the chains: a pass where a chain passes is folded
// again one hit at a time from the summaries as they were, and stops there.
// Passes traced beyond the stop are not folded.  W = 1 keeps the pass in
// registers, with no LDS ring and no barrier.
template <uint32_t W, class Tr>
__device__ void sem_row(const RowSlots &a, uint32_t r, Tr &tr, double *lds) {
#pragma clang fp contract(off)
    static_assert(W >= 1 && W <= kSemWaves, "the LDS ring holds kSemWaves passes");
    double *ring_t = lds + 3 * a.nbin_max;  // [kSemWaves * 64]: pass p, lane at p * 64 + lane
    uint32_t *ring_b = (uint32_t *)(ring_t + kSemWaves * 64);
    double *ctl = (double *)(ring_b + kSemWaves * 64);  // stop flag, rays folded
    constexpr uint32_t n_waves = W;
    const uint32_t lane = threadIdx.x & 63u, wave = W > 1 ? threadIdx.x >> 6 : 0u;
    const uint32_t nbin = tr.nbin;
    const uint32_t cap = a.row_cap[r];
    const double max_err_sq = a.max_err_sq;
    // without an early exit nothing reads m2 (rows give their means): the
    // fold updates counts and means alone, in the same operations
    const bool track_m2 = max_err_sq > 0.0;
    double count = 0.0;
    bool stop = false;
    for (uint32_t group = 1; group <= cap && !stop; group += 64 * n_waves) {
        // this wavefront's pass (one wavefront: kept in registers)
        double t_own = 0.0;
        uint32_t bin_own = nbin;
        if (group + 64 * wave <= cap) {
            const uint32_t k = group + 64 * wave + lane;
            bool redo = false;
            bin_own = tr.live ? tr.trace(k, k <= cap, t_own, redo) : nbin;
            if (Tr::kRetrace && wave_any(redo) && redo) bin_own = tr.retrace(k, t_own);
            if (!tr.live) t_own = 0.0;
        }
        if (n_waves > 1) {
            ring_t[wave * 64 + lane] = t_own;
            ring_b[wave * 64 + lane] = bin_own;
            __syncthreads();
        }
        if (wave == 0) {
            for (uint32_t p = 0; p < n_waves && !stop; ++p) {
                const uint32_t base = group + 64 * p;
                if (base > cap) break;
                const double t = n_waves > 1 ? ring_t[p * 64 + lane] : t_own;
                const uint32_t bin = n_waves > 1 ? ring_b[p * 64 + lane] : bin_own;
                const uint32_t n_here = min(64u, cap - base + 1);
                const bool hit = lane < n_here && bin < nbin;
                const uint64_t hits = __ballot(hit);
                if (hits) {
                    // this lane's bin-mates (leaders only) and whether it leads
                    uint64_t mates = 0;
                    for (uint64_t left = hits; left;) {
                        const uint32_t b = __builtin_amdgcn_readlane(bin, (uint32_t)__builtin_ctzll(left));
                        const uint64_t m = __ballot(hit && bin == b);
                        left &= ~m;
                        if (hit && bin == b && lane == (uint32_t)__builtin_ctzll(m)) mates = m;
                    }
                    const bool lead = mates != 0;
                    W64 w{0.0, 0.0, 0.0};
                    if (lead) w = load_w(lds, a.nbin_max, bin);
                    bool passed = false;
                    for (uint64_t rem = mates; wave_any(rem != 0);) {
                        const uint32_t src = rem ? (uint32_t)__builtin_ctzll(rem) : lane;
                        const double ts = __shfl(t, (int)src);
                        if (rem) {
                            rem &= rem - 1;
                            fold_hit(w, (double)(base + src), ts, a.recip, track_m2);
                            passed = passed || (track_m2 && w_sem_le(w, max_err_sq));
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    if (!wave_any(passed)) {
                        if (lead) store_w(lds, a.nbin_max, bin, w);
                        __builtin_amdgcn_wave_barrier();
                    } else {
                        // again, hit by hit from the summaries in LDS, up to the stop
                        uint64_t left = hits;
                        while (left) {
                            const uint32_t l = (uint32_t)__builtin_ctzll(left);
                            left &= left - 1;
                            const uint32_t b = __builtin_amdgcn_readlane(bin, l);
                            W64 v = load_w(lds, a.nbin_max, b);
                            fold_hit(v, (double)(base + l), readlane_f64(t, l), a.recip, track_m2);
                            __builtin_amdgcn_wave_barrier();
                            if (lane == 0) store_w(lds, a.nbin_max, b, v);
                            __builtin_amdgcn_wave_barrier();
                            if (w_sem_le(v, max_err_sq)) {
                                count = (double)(base + l);
                                stop = true;
                                break;
                            }
                        }
                    }
                }
                if (!stop) count += (double)n_here;
            }
            if (n_waves > 1 && lane == 0) {
                ctl[0] = stop ? 1.0 : 0.0;
                ctl[1] = count;
            }
        }
        if (n_waves > 1) {
            __syncthreads();
            stop = ctl[0] != 0.0;
            count = ctl[1];
        }
    }
    finish_row(a, r, lds, nbin, count);
}

// Passes of 64 rays (two chunks) traced ahead of the fold.
constexpr uint32_t kAhead = 4;

// LDS of chunk_row with W wavefronts after the 3 nbin_max Welford doubles:
// the pass's (transmittance, bin) pairs, the traced passes' transmittances,
// bins and re-trace queues, and the fold's stop flag and chunk count; the
// tracer's own bytes follow.
__host__ __device__ constexpr size_t chunk_lds_bytes(uint32_t nbin_max, uint32_t W) {
    return (size_t)3 * nbin_max * sizeof(double) + 64 * 16 +
           (size_t)W * kAhead * 64 * (sizeof(double) + 2 * sizeof(uint32_t)) + 2 * sizeof(double);
}

// kernel.rs:89-193: each of the block's W wavefronts traces kAhead passes
// of two chunks (lanes 0-31 and 32-63), and the first folds all of them in
// order; a wavefront's flagged rays are queued and re-traced together, and
// rays traced beyond the chunk where the row stops are not folded.  In the fold, each chunk's (32, mean,
// m2) per bin comes from half-wave butterflies, both halves at once: the two
// lowest bins of each half one at a time, then every lane that hit one of the
// remaining bins evaluates the same butterfly sums for its own bin from LDS.
// The merges into the row's summaries follow chunk by chunk.
template <uint32_t W, class Tr>
__device__ void chunk_row(const RowSlots &a, uint32_t r, Tr &tr, double *lds) {
#pragma clang fp contract(off)
    static_assert(W >= 1 && W <= kSemWaves, "chunk_lds_bytes sizes the passes of W wavefronts");
    double2 *tpair = (double2 *)(lds + 3 * a.nbin_max);  // [64]; bin ~0u: nothing left to fold
    double *tv = (double *)(tpair + 64);                 // [W * kAhead * 64] pass p, lane at p * 64 + lane
    uint32_t *bv = (uint32_t *)(tv + W * kAhead * 64);
    uint32_t *queue = bv + W * kAhead * 64;              // wavefront w's at w * kAhead * 64
    double *ctl = (double *)(queue + W * kAhead * 64);   // stop flag, chunks folded
    const uint32_t lane = threadIdx.x & 63u, wave = W > 1 ? threadIdx.x >> 6 : 0u;
    const uint32_t nbin = tr.nbin;
    const double max_err_sq = a.max_err_sq;
    const uint32_t iters = a.max_rays / 32;
    const uint32_t half = lane >> 5, l = lane & 31;
    double count = 0.0;
    bool stop = false;
    for (uint32_t i0 = 0; i0 < iters && !stop; i0 += 2 * kAhead * W) {
        // this wavefront's passes: pass p of the group is chunks i0 + 2 p and
        // i0 + 2 p + 1, and the wavefront traces p = wave kAhead ... + kAhead - 1
        const uint32_t w0 = i0 + 2 * kAhead * wave;
        const uint32_t passes = w0 < iters ? min(kAhead, (iters - w0 + 1) / 2) : 0u;
        uint32_t *wq = queue + wave * kAhead * 64;
        // trace: ray k = l + 32 (c + 1) of chunk c = w0 + 2 p + half
        uint32_t queued = 0;
        for (uint32_t p = 0; p < passes; ++p) {
            const uint32_t c = w0 + 2 * p + half;
            const bool active = tr.live && c < iters;
            const uint32_t k = l + 32 * (c + 1);
            double t = 0.0;
            bool redo = false;
            uint32_t bin = tr.live ? tr.trace(k, active, t, redo) : nbin;
            if (redo || !(bin < nbin)) {
                bin = redo ? bin : nbin;
                t = 0.0;
            }
            const uint32_t slot = (wave * kAhead + p) * 64 + lane;
            tv[slot] = t;
            bv[slot] = bin;
            if (Tr::kRetrace) {
                const uint64_t mask = __builtin_amdgcn_ballot_w64(redo);
                if (redo) wq[queued + __popcll(mask & lanemask_below())] = slot;
                queued += __popcll(mask);
            }
        }
        // exact re-trace of the queued rays, 64 at a time
        if (Tr::kRetrace) {
            for (uint32_t q0 = 0; q0 < queued; q0 += 64) {
                __builtin_amdgcn_wave_barrier();
                if (q0 + lane < queued) {
                    const uint32_t slot = wq[q0 + lane];
                    const uint32_t c = i0 + 2 * (slot >> 6) + ((slot >> 5) & 1);
                    const uint32_t k = (slot & 31) + 32 * (c + 1);
                    double t;
                    const uint32_t bin = tr.retrace(k, t);
                    tv[slot] = bin < nbin ? t : 0.0;
                    bv[slot] = bin;
                }
            }
        }
        if (W > 1)
            __syncthreads();
        else
            __builtin_amdgcn_wave_barrier();
        // fold the passes' chunks in order (the first wavefront).  A chunk's
        // statistics per bin depend on its own 32 lanes only, so both halves
        // form theirs at once; the merges into the row's summaries then run
        // chunk by chunk, every bin of a chunk at once.
        const uint32_t group_passes = wave == 0 ? min(kAhead * W, (iters - i0 + 1) / 2) : 0u;
        for (uint32_t p = 0; p < group_passes && !stop; ++p) {
            const uint32_t i = i0 + 2 * p;
            const double t = tv[p * 64 + lane];
            const uint32_t bin = bv[p * 64 + lane];
            // (32, cmean, cm2): this lane's bin over its chunk
            bool pending = i + half < iters && bin < nbin;
            double cmean = 0.0, cm2 = 0.0;
            for (int serial = 0; serial < 2 && __ballot(pending) != 0; ++serial) {
                const uint32_t bmin = half_min_u32(pending ? bin : 0xffffffffu);
                const double tb = bin == bmin ? t : 0.0;
                const double wmean = half_sum(tb) / 32.0;
                const double dev = tb - wmean;
                const double wm2 = half_sum(dev * dev);
                if (pending && bin == bmin) {
                    cmean = wmean;
                    cm2 = wm2;
                    pending = false;
                }
            }
            if (__ballot(pending) != 0) {
                // the remaining bins from LDS, each lane for its own bin
                __builtin_amdgcn_wave_barrier();
                tpair[lane] = make_double2(t, __longlong_as_double(pending ? bin : ~0u));
                __builtin_amdgcn_wave_barrier();
                if (pending) {
                    const double2 *mine = tpair + 32 * half;
                    auto v = [&](int j) {
                        const double2 pr = mine[j];
                        return (uint32_t)__double_as_longlong(pr.y) == bin ? pr.x : 0.0;
                    };
                    cmean = butterfly_sum(v) / 32.0;
                    cm2 = butterfly_sum([&](int j) {
                        const double dev = v(j) - cmean;
                        return dev * dev;
                    });
                }
                __builtin_amdgcn_wave_barrier();
            }
            for (uint32_t h = 0; h < 2 && i + h < iters; ++h) {
                count += 32.0;
                const bool merge = half == h && bin < nbin;
                bool ok = true;
                if (merge) {
                    // lanes that share a bin compute and store the same record
                    const W64 held = load_w(lds, a.nbin_max, bin);
                    const W64 w = w_add_r(held, W64{32.0, cmean, cm2}, a.recip[(uint32_t)held.count + 32]);
                    store_w(lds, a.nbin_max, bin, w);
                    ok = w_sem_le(w, max_err_sq);
                }
                __builtin_amdgcn_wave_barrier();
                if (__ballot(merge) != 0 && __ballot(merge && !ok) == 0) {
                    stop = true;
                    break;
                }
            }
        }
        if (W > 1) {
            if (wave == 0 && lane == 0) {
                ctl[0] = stop ? 1.0 : 0.0;
                ctl[1] = count;
            }
            __syncthreads();
            stop = ctl[0] != 0.0;
            count = ctl[1];
        } else {
            __builtin_amdgcn_wave_barrier();
        }
    }
    finish_row(a, r, lds, nbin, count);
}

}  // namespace cpd
