// kernels.hip -- gfx950 kernels of the compound-prism fitness evaluator.
//
// trace_kernel: one 64-lane wavefront per work item (design, wavelength
// sample, tile of rays), taken from an atomic counter by a persistent grid.
// Per item the wavefront builds the wavelength's constants in FP64 (the
// planar faces collapse to an interval of beam heights and an affine exit
// point; their FP64 directions and transmittances go to an LDS record), holds
// them in registers (scalar, or vector where an operand would otherwise be
// copied per use), and lane l traces positions k0 + l, k0 + 64 + l, ... of
// the tile in FP32 (trace32) from the banded beam table, adding each hit to
// its own column of a [bin][column] LDS histogram.  Rays whose FP32 outcome
// lies within the error margin of a hit/miss or bin decision go to a
// per-wave LDS queue and are re-traced 64 at a time in FP64 (trace64_chain,
// the reference's operation order).  Items of designs that failed
// construction (a plan's stand-ins) trace nothing.  With the default accumulation a ray adds
// round(t * 2^22) to a u32 column; the columns are folded into a u64 row
// before they can overflow (skewed, so the lanes hit distinct banks), and
// the row is added to the design's per-(wavelength, bin) sums with integer
// atomics -- exact, so the result is the same whatever the tile order or the
// batch a design shares the GPU with.  The FP32 and FP64 accumulation
// policies of the precision study replace the cells and the atomics.
//
// finalize_kernel: one wavefront per design turns the sums into
// p(d|l) = sum / M (the bin of the last ray divides by M + 1, as the
// reference's Welford padding does, src/fitness.rs:20-32), and reduces
// H(D|L) = mean_l sum_d p log2 p and H(D) = sum_d pbar log2 pbar with
// Neumaier-compensated sums and a fixed reduction order
// (src/fitness.rs:79-123).  It clears the sums it read, so the next launch
// starts from zero without a memset.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "construct.h"
#include "kernels.h"

namespace cpd {

namespace {

__device__ __forceinline__ uint32_t quant32(float t) { return __float2uint_rn(t * (float)kQuantScale); }
__device__ __forceinline__ uint32_t quant64(double t) { return __double2uint_rn(t * kQuantScale); }

// Set bits of m below this lane (mbcnt: no lane-mask register kept live).
__device__ __forceinline__ uint32_t count_below(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
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

// A wave-uniform value held in a vector register.  A vector instruction can
// read only one scalar register, so an operand that meets another uniform
// operand in one instruction (fma(e1, y, e0), for instance) would otherwise
// be copied into a vector register at every use; the empty asm makes the
// compiler treat the value as varying per lane, which keeps the single copy
// made here in a vector register for the whole loop.  It also takes the
// value out of the scalar register file, which the ray loop otherwise fills
// until it spills.
template <class T>
__device__ __forceinline__ T in_vgpr(T v) {
    static_assert(sizeof(T) == 4, "32-bit value");
    asm("; uniform value held in %0" : "+v"(v));
    return v;
}

// The entry whose items hold `item`: the last e < n with begin[e] <= item
// (begin ascending, begin[0] = 0 <= item).  All 64 lanes search together: each
// round every lane tests one of 64 evenly spaced candidates, so n <= 4096
// takes two rounds of one load each (a binary search takes log2 n dependent
// loads).
__device__ __forceinline__ uint32_t find_design(const uint64_t *begin, uint32_t n, uint64_t item) {
    const uint32_t lane = threadIdx.x;
    uint32_t lo = 0, cnt = n;  // the entry lies in [lo, lo + cnt)
    while (cnt > 1) {
        const uint32_t stride = (cnt + 63u) >> 6;
        const uint32_t k = lane * stride;
        const bool le = k < cnt && begin[lo + k] <= item;
        const uint32_t c = (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(le));  // >= 1 (lane 0)
        lo += (c - 1u) * stride;
        cnt = min(stride, cnt - (c - 1u) * stride);
    }
    return lo;
}

// Neumaier-compensated sum (FP64, or FP32 for the compensated FP32 policy).
template <class T>
struct CSumT {
    T s = T(0), c = T(0);
    __device__ void add(T x) {
#pragma clang fp contract(off)
        T t = s + x;
        if (fabs(s) >= fabs(x))
            c += (s - t) + x;
        else
            c += (x - t) + s;
        s = t;
    }
    __device__ T value() const { return s + c; }
};
using CSum = CSumT<double>;

__device__ __forceinline__ double plog2p(double p) { return p > 0.0 ? p * log2(p) : 0.0; }
__device__ __forceinline__ float plog2pf(float p) { return p > 0.f ? p * log2f(p) : 0.f; }

// How a wavefront accumulates transmittance per bin (CPD_OPT_ACCUMULATE).
//  * exact (default): round(t 2^22) into u32 LDS columns, folded into a u64
//    LDS row before a column can overflow, flushed with u64 atomics: exact
//    integer sums, independent of order; FP64 compensated entropies.
//  * f32: FP32 LDS columns and FP32 global atomics, FP32 entropies summed
//    plainly -- the naive FP32 estimator.
//  * f32c: FP32 columns folded every 1024 rays into a Neumaier-compensated FP32
//    row; a work item is a whole wavelength, so the row is stored, not added;
//    FP32 compensated entropies.
//  * f64: FP64 LDS columns and FP64 global atomics, FP64 compensated entropies.
template <int A> struct Acc;
// kTScale, kTOff: the scale and offset trace32's transmittance arrives with
// (make_hot's tscale and toff).
template <> struct Acc<CPD_ACC_EXACT> {
    using Cell = uint32_t;
    static constexpr uint32_t kRowBytes = 8;
    // t arrives as t 2^22 + 1/2 (the fused multiply-add that forms it adds
    // the half): truncation rounds it to the nearest integer
    static constexpr float kTScale = 0.5f * (float)kQuantScale, kTOff = 0.5f;
    static __device__ Cell of32(float t) { return (Cell)t; }
    static __device__ Cell of64(double t) { return quant64(t); }
    static __device__ uint32_t fold_every(uint32_t cols) { return fold_rays(cols); }
};
template <> struct Acc<CPD_ACC_F32> {
    using Cell = float;
    static constexpr uint32_t kRowBytes = 4;
    static constexpr float kTScale = 0.5f, kTOff = 0.f;
    static __device__ Cell of32(float t) { return t; }
    static __device__ Cell of64(double t) { return (float)t; }
    static __device__ uint32_t fold_every(uint32_t) { return 0xffffffffu; }
};
template <> struct Acc<CPD_ACC_F32C> {
    using Cell = float;
    static constexpr uint32_t kRowBytes = 8;
    static constexpr float kTScale = 0.5f, kTOff = 0.f;
    static __device__ Cell of32(float t) { return t; }
    static __device__ Cell of64(double t) { return (float)t; }
    static __device__ uint32_t fold_every(uint32_t) { return 1024; }
};
template <> struct Acc<CPD_ACC_F64> {
    using Cell = double;
    static constexpr uint32_t kRowBytes = 8;
    static constexpr float kTScale = 0.5f, kTOff = 0.f;
    static __device__ Cell of32(float t) { return (double)t; }
    static __device__ Cell of64(double t) { return t; }
    static __device__ uint32_t fold_every(uint32_t) { return 0xffffffffu; }
};

// The bin of a design's last ray for one wavelength: an agent-scope store,
// which reaches the memory side before it completes (the resident loop hands
// a finished item to its design's finalizer, possibly on another XCD, without
// writing back its L2).
__device__ __forceinline__ void set_last_bin(uint32_t *p, uint32_t b) {
    __hip_atomic_store(p, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

struct TileCtx {
    const cpd_design *g;
    const TraceDesign *td;
    LambdaConsts lc;
    uint32_t k_begin, k_end, n_rays, nbin, cols;
    void *hist;      // [bin][column] cells of the accumulation policy
    void *row;       // the wavefront's row per bin
    uint32_t *queue;
    uint32_t *last_bin;
    const uint2 *zk;     // banded order: (FP32 z bits, 1-based ray index)
    const double *zs64;  // banded order, FP64
    uint32_t full_end;   // positions 1..full_end: whole banded blocks (rays past n_rays in the last are idle)
    uint32_t last_pos;   // position of ray k = n_rays
    const Chain64 *chain;  // the wavelength's planar faces in FP64 (LDS)
    const LensBand64 *band;  // the wavelength's lens and acceptance constants in FP64 (LDS)
    uint64_t *masks;         // skip masks of the current chunk's bands, then the tile's lost sub-bands (LDS)
    const LambdaConsts *lc_lds;  // lc, as lane 0 stored it (LDS)
};

// FP64 re-trace of `count` (<= 64) queued rays starting at ring slot `head`.
template <int A>
__device__ void drain_queue(const TileCtx &c, uint32_t head, uint32_t count) {
    using Cell = typename Acc<A>::Cell;
    const uint32_t lane = threadIdx.x;
    if (lane < count) {
        const uint32_t i = c.queue[(head + lane) & (kQueueCap - 1)];  // position
        const double y = beam_y64(*c.g, c.zs64[i - 1]);
        // the affine FP64 trace decides nearly every queued ray; the few it
        // leaves (a decision within its 1e-9 margin) take the reference's
        // operation order
        double t;
        uint32_t b;
        if (!trace64_affine(*c.band, *c.chain, *c.g, y, b, t)) b = trace64_chain(*c.g, *c.chain, y, t);
        if (b < c.nbin) atomicAdd((Cell *)c.hist + b * c.cols + (lane & (c.cols - 1)), Acc<A>::of64(t));
        if (i == c.last_pos) set_last_bin(c.last_bin, b);
    }
}

// Adds each bin's columns (skewed, conflict-free order) to the row and clears
// them.  Lane b handles bins b, b + 64, ...
template <int A>
__device__ void fold_columns(void *hist_v, void *row_v, uint32_t nbin, uint32_t cols) {
    using Cell = typename Acc<A>::Cell;
    Cell *hist = (Cell *)hist_v;
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbin; b += 64) {
        Cell *cell = hist + b * cols;
        if (A == CPD_ACC_EXACT) {
            uint64_t sum = 0;
            for (uint32_t q = 0; q < cols; ++q) {
                const uint32_t cc = (q + b) & (cols - 1);
                sum += (uint64_t)cell[cc];
                cell[cc] = 0;
            }
            ((unsigned long long *)row_v)[b] += sum;
        } else if (A == CPD_ACC_F32C) {
            float sum = 0.f;
            for (uint32_t q = 0; q < cols; ++q) {
                const uint32_t cc = (q + b) & (cols - 1);
                sum += (float)cell[cc];
                cell[cc] = 0;
            }
            CSumT<float> r;
            r.s = ((float *)row_v)[2 * b];
            r.c = ((float *)row_v)[2 * b + 1];
            r.add(sum);
            ((float *)row_v)[2 * b] = r.s;
            ((float *)row_v)[2 * b + 1] = r.c;
        }
    }
    __syncthreads();
}

// Chunk layouts.  A chunk of kMaskChunk positions made of four whole blocks
// of kBandRays banded positions is traced interleaved: iteration i takes the
// ranks 16 i .. 16 i + 15 of each block (lane l: block l / 16, rank
// 16 i + l % 16), so its 64 rays span about 1/256 of the beam's heights, as
// if the chunk were one sorted block.  Any other chunk is traced in position
// order, 64 consecutive positions per iteration.  Band i of a chunk is the
// rays of its iteration i.
struct ChunkMap {
    uint32_t chunk, end;  // positions [chunk, end)
    bool inter;           // interleaved layout
};
__device__ __forceinline__ ChunkMap chunk_map(uint32_t chunk, uint32_t end, uint32_t full_end) {
    static_assert(kMaskChunk == 4 * kBandRays && kBandRays == 4096, "four blocks of 4096 per chunk");
    const bool inter = end - chunk == kMaskChunk && ((chunk - 1u) & (kBandRays - 1u)) == 0 &&
                       chunk + kMaskChunk - 1u <= full_end;
    return ChunkMap{chunk, end, inter};
}

// Position of the n-th (from 0) set bit of m (n < popcount(m)).
__device__ __forceinline__ uint32_t nth_set_bit(uint64_t m, uint32_t n) {
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t half = 32; half >= 1; half >>= 1) {
        const uint32_t c = __popcll(m & ((1ull << half) - 1));
        if (n >= c) {
            n -= c;
            m >>= half;
            pos += half;
        }
    }
    return pos;
}

// Sub-bands of a tile: sub-band u (0..255) is the ranks 16 u .. 16 u + 15 of
// every banded block the tile's positions touch, group s (0..63) the ranks
// 64 s .. 64 s + 63.  Each block is sorted by z, so the heights of a group or
// sub-band lie between the FP64 heights of its lowest and highest rank over
// those blocks.  The wavefront classifies the 64 groups (one lane each), then
// the four sub-bands of every mixed group (64 to a pass), and sets bit u of
// `sub` (four words) when band_class proves sub-band u lost: every sub-band
// of a lost group, and those of the mixed groups that test lost.  Interval
// arithmetic only widens with its interval, and every position of a chunk's
// band lies in the sub-bands its ranks fall in, so a band all of whose
// sub-bands are lost is lost (band_masks).
__device__ void tile_band_classes(const TileCtx &c, uint64_t *sub) {
    const uint32_t lane = threadIdx.x;
    const uint32_t b0 = (c.k_begin - 1u) >> 12, b1 = (c.k_end - 2u) >> 12;  // blocks of positions k_begin .. k_end - 1
    double zlo = INFINITY, zhi = -INFINITY;
    for (uint32_t b = b0; b <= b1; ++b) {
        zlo = fmin(zlo, c.zs64[(b << 12) + 64u * lane]);
        zhi = fmax(zhi, c.zs64[(b << 12) + 64u * lane + 63u]);
    }
    const int cls = band_class(*c.band, beam_y64(*c.g, zlo), beam_y64(*c.g, zhi));
    const uint64_t lost = __builtin_amdgcn_ballot_w64(cls == kBandLost);  // bit = group
    const uint64_t mixed = __builtin_amdgcn_ballot_w64(cls == kBandMixed);
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
        const uint64_t v = __builtin_amdgcn_ballot_w64((lost >> (16u * w + (lane >> 2))) & 1u);
        if (lane == 0) sub[w] = v;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    const uint32_t n_sub = 4u * (uint32_t)__popcll(mixed);
    for (uint32_t p = 0; p < n_sub; p += 64) {
        const uint32_t k = p + lane;
        if (k < n_sub) {
            const uint32_t u = 4u * nth_set_bit(mixed, k >> 2) + (k & 3u);
            double lo = INFINITY, hi = -INFINITY;
            for (uint32_t b = b0; b <= b1; ++b) {
                lo = fmin(lo, c.zs64[(b << 12) + 16u * u]);
                hi = fmax(hi, c.zs64[(b << 12) + 16u * u + 15u]);
            }
            if (!band_may_hit(*c.band, beam_y64(*c.g, lo), beam_y64(*c.g, hi)))
                atomicOr((unsigned long long *)&sub[u >> 6], 1ull << (u & 63u));
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

// Skip masks of the bands of a chunk from its tile's lost sub-bands `sub`:
// word g of `masks` holds bands 64 g .. 64 g + 63.  Band i of an interleaved
// chunk is the ranks 16 i .. 16 i + 15 of its four blocks, inside sub-band i;
// a band of 64 consecutive positions in one block covers the sub-bands of its
// ranks.  A band is skipped when every sub-band it covers is lost.  Bands
// reaching past the whole blocks are never skipped.
__device__ void band_masks(const TileCtx &c, const ChunkMap &m, uint64_t *masks, const uint64_t *sub) {
    const uint32_t lane = threadIdx.x;
    const uint32_t words = (m.end - m.chunk + 4095u) >> 12;
#pragma unroll
    for (uint32_t w = 0; w < 4; ++w) {
        if (w < words) {
            bool lost;
            if (m.inter) {
                lost = (sub[w] >> lane) & 1u;
            } else {
                const uint32_t p0 = m.chunk + 64u * (64u * w + lane), p1 = p0 + 63u;
                lost = p0 < m.end && p1 <= c.full_end && ((p0 - 1u) >> 12) == ((p1 - 1u) >> 12);
                if (lost) {
                    const uint32_t u0 = ((p0 - 1u) & 4095u) >> 4, u1 = ((p1 - 1u) & 4095u) >> 4;  // u1 - u0 <= 4
                    for (uint32_t u = u0; u <= u1; ++u) lost = lost && ((sub[u >> 6] >> (u & 63u)) & 1u);
                }
            }
            const uint64_t v = __builtin_amdgcn_ballot_w64(lost);
            if (lane == 0) masks[w] = v;
        }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

// The iterations a chunk traces, in order: those below n_it whose band is not
// skipped (MASKS: the chunk's skip masks, one 64-bit word per 64 iterations,
// in LDS; otherwise every iteration).  The masks are wave-uniform: the walk
// keeps the current word in scalar registers.
template <bool MASKS>
struct LiveWalk {
    const uint64_t *masks;
    uint64_t live;  // untraced iterations of word w still to come
    uint32_t w, words, n_it;
    __device__ __forceinline__ LiveWalk(const uint64_t *m, uint32_t n) : masks(m), live(0), w(0), n_it(n) {
        words = (n + 63u) >> 6;
        if (words) live = word(0);
    }
    __device__ __forceinline__ uint64_t word(uint32_t q) const {
        uint64_t m = 0;
        if (MASKS) {
            const uint64_t v = masks[q];
            // (readfirstlane returns int: each half is taken as u32, so the
            // low half is not sign-extended into the high one)
            m = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
                (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)v);
        }
        const uint32_t left = n_it - 64u * q;
        return ~m & (left >= 64u ? ~0ull : (1ull << left) - 1ull);
    }
    // whether iteration i (< n_it) is traced
    __device__ __forceinline__ bool traced(uint32_t i) const { return (word(i >> 6) >> (i & 63u)) & 1u; }
    // the next traced iteration, n_it when none is left
    __device__ __forceinline__ uint32_t next() {
        while (!live) {
            if (++w >= words) return n_it;
            live = word(w);
        }
        const uint32_t i = 64u * w + (uint32_t)__builtin_ctzll(live);
        live &= live - 1ull;
        return i;
    }
};

// Beam quantile (FP32 bits) and ray index of this lane's position in the
// iteration whose positions start at base (1-based).
__device__ __forceinline__ uint2 beam_zk(const uint2 *__restrict__ zk, uint32_t base, uint32_t lane_off) {
    return zk[base - 1u + lane_off];
}

template <bool COUNT, int A>
__device__ void tile_loop(const TileCtx &c, unsigned long long &ops) {
    using Cell = typename Acc<A>::Cell;
    const uint32_t lane = threadIdx.x;
    const uint32_t col = lane & (c.cols - 1);
    Cell *__restrict__ hist = (Cell *)c.hist;
    const uint32_t k_end = c.k_end, nbin = c.nbin, cols = c.cols, full_end = c.full_end;
    const uint32_t last_pos = (uint32_t)__builtin_amdgcn_readfirstlane(c.last_pos);
    // the lane's column of bin b is (b << cell_shift) bytes past hist_col
    // (cols is a power of two)
    const uint32_t cell_shift = 31u - __builtin_clz(cols) + (sizeof(Cell) == 8 ? 3u : 2u);
    char *__restrict__ hist_col = (char *)(hist + col);
    uint32_t head = 0, tail = 0;  // ring of queued ray indices (wave-uniform)
    const uint32_t fold = Acc<A>::fold_every(cols);
    uint32_t since_fold = 0;
    // The tile runs in chunks of kMaskChunk positions.  Before its rays, the
    // wavefront proves sub-bands of the tile's ranks lost where the FP64 trace
    // loses every ray of them at the lens, its acceptance test or the array
    // ends (tile_band_classes); before each chunk's rays it marks the chunk's
    // bands of 64 positions that lie in lost sub-bands (band_masks), and the
    // ray loop skips them.  Op counting traces every band.
    if (!COUNT) tile_band_classes(c, c.masks + (kMaskChunk >> 12));
    for (uint32_t chunk = c.k_begin; chunk < k_end; chunk += kMaskChunk) {
        const uint32_t chunk_end = min(chunk + kMaskChunk, k_end);
        const ChunkMap cmap = chunk_map(chunk, chunk_end, full_end);
        if (!COUNT) band_masks(c, cmap, c.masks, c.masks + (kMaskChunk >> 12));
        // register copies of the per-item records (made after the band tests,
        // whose FP64 registers they would otherwise share): the loop stores
        // through last_bin, so loads through c.td would be repeated every
        // iteration
        HotRecord hot = to_sgprs(make_hot(*c.td, *c.lc_lds, Acc<A>::kTScale, Acc<A>::kTOff));
        // operands paired with another uniform operand, and the detector
        // stage's constants (which would otherwise be spilled)
        hot.y_mean = in_vgpr(hot.y_mean);
        hot.g0 = in_vgpr(hot.g0);
        hot.lb0 = in_vgpr(hot.lb0);
        hot.lc1 = in_vgpr(hot.lc1);
        hot.lomr2 = in_vgpr(hot.lomr2);
        hot.nd0 = in_vgpr(hot.nd0);
        hot.wdn0 = in_vgpr(hot.wdn0);
        hot.wdd0 = in_vgpr(hot.wdd0);
        hot.ndd0 = in_vgpr(hot.ndd0);
        hot.dlen0 = in_vgpr(hot.dlen0);
        hot.tcoat = in_vgpr(hot.tcoat);
        hot.nbin = in_vgpr(hot.nbin);
        hot.half_len = in_vgpr(hot.half_len);
        hot.slope = in_vgpr(hot.slope);
        hot.inv_slope = in_vgpr(hot.inv_slope);
        hot.bin_size = in_vgpr(hot.bin_size);
        // this lane's position in iteration it is chunk + it * step + lane_off
        const uint32_t step = cmap.inter ? 16u : 64u;
        const uint32_t lane_off = cmap.inter ? ((lane >> 4) * kBandRays + (lane & 15u)) : lane;
        // the iteration and lane of the last ray's position, when in the chunk
        uint32_t last_it = ~0u, last_lane = 0;
        if (last_pos >= chunk && last_pos < chunk_end) {
            const uint32_t off = last_pos - chunk;
            last_it = cmap.inter ? (off & (kBandRays - 1u)) >> 4 : off >> 6;
            last_lane = cmap.inter ? ((off >> 12) << 4) + (off & 15u) : off & 63u;
        }
        const uint32_t n_it = (chunk_end - chunk + 63u) >> 6;
        LiveWalk<!COUNT> walk(c.masks, n_it);
        // the last ray in a skipped band misses with the rest of it
        if (!COUNT && last_it != ~0u && !walk.traced(last_it) && lane == 0) set_last_bin(c.last_bin, nbin);
        // every iteration holds 64 positions of whole banded blocks; a lane
        // whose ray index exceeds n_rays (the last block of a design whose ray
        // count is not a multiple of kBandRays) traces nothing
        const uint2 *__restrict__ zk = c.zk;
        const uint32_t n_rays = c.n_rays;
        uint32_t it = walk.next();
        uint2 zr = beam_zk(zk, chunk + (it < n_it ? it : 0u) * step, lane_off);
        while (it < n_it) {
            // the next traced iteration's beam height is loaded a whole
            // iteration ahead, so the trace never waits for it
            // (loaded unconditionally -- the current iteration's positions
            // stand in past the chunk's end -- so one load is in flight on
            // every path and the trace waits only for the older one)
            const uint32_t next = walk.next();
            const uint2 zr_next = beam_zk(zk, chunk + (next < n_it ? next : it) * step, lane_off);
            const uint32_t base = chunk + it * step;
            const bool active = zr.y <= n_rays;
            const float y = hot.y_mean - hot.beam_w * __uint_as_float(zr.x);
            const Trace32Result r = trace32<COUNT>(hot, y);
            if (COUNT && active) ops += r.ops;
            const bool flagged = active && r.guard;
            if (active && !r.guard && r.hit) atomicAdd((Cell *)(hist_col + (r.bin << cell_shift)), Acc<A>::of32(r.t));
            if (it == last_it && lane == last_lane && !r.guard) set_last_bin(c.last_bin, r.hit ? r.bin : nbin);
            const uint32_t k = base + lane_off;  // position
            const uint64_t mask = __builtin_amdgcn_ballot_w64(flagged);
            if (mask) {
                if (flagged) c.queue[(tail + count_below(mask)) & (kQueueCap - 1)] = k;
                tail += __popcll(mask);
                if (tail - head >= 64) {
                    __builtin_amdgcn_wave_barrier();
                    drain_queue<A>(c, head, 64);
                    head += 64;
                }
            }
            since_fold += 64;
            if (since_fold >= fold) {
                fold_columns<A>(c.hist, c.row, nbin, cols);
                since_fold = 0;
            }
            it = next;
            zr = zr_next;
        }
    }
    if (tail != head) {
        __builtin_amdgcn_wave_barrier();
        drain_queue<A>(c, head, tail - head);
    }
}

// Last fold of an item and flush of its row into the global sums of
// (design, wavelength j).
template <int A>
__device__ void flush_row(void *hist_v, void *row_v, uint32_t nbin, uint32_t cols, unsigned long long *acc_base,
                          uint64_t idx0) {
    using Cell = typename Acc<A>::Cell;
    Cell *hist = (Cell *)hist_v;
    __syncthreads();
    for (uint32_t b = threadIdx.x; b < nbin; b += 64) {
        Cell *cell = hist + b * cols;
        if (A == CPD_ACC_EXACT) {
            unsigned long long *row = (unsigned long long *)row_v;
            uint64_t sum = row[b];
            row[b] = 0;
            for (uint32_t q = 0; q < cols; ++q) {
                const uint32_t cc = (q + b) & (cols - 1);
                sum += (uint64_t)cell[cc];
                cell[cc] = 0;
            }
            if (sum) atomicAdd(acc_base + idx0 + b, (unsigned long long)sum);
        } else if (A == CPD_ACC_F32) {
            float sum = 0.f;
            for (uint32_t q = 0; q < cols; ++q) {
                const uint32_t cc = (q + b) & (cols - 1);
                sum += (float)cell[cc];
                cell[cc] = 0;
            }
            if (sum != 0.f) atomicAdd((float *)(acc_base + idx0 + b), sum);
        } else if (A == CPD_ACC_F32C) {
            float sum = 0.f;
            for (uint32_t q = 0; q < cols; ++q) {
                const uint32_t cc = (q + b) & (cols - 1);
                sum += (float)cell[cc];
                cell[cc] = 0;
            }
            float *row = (float *)row_v;
            CSumT<float> r;
            r.s = row[2 * b];
            r.c = row[2 * b + 1];
            r.add(sum);
            row[2 * b] = row[2 * b + 1] = 0.f;
            *(float *)(acc_base + idx0 + b) = r.value();  // the only item of this wavelength
        } else {
            double sum = 0.0;
            for (uint32_t q = 0; q < cols; ++q) {
                const uint32_t cc = (q + b) & (cols - 1);
                sum += (double)cell[cc];
                cell[cc] = 0;
            }
            if (sum != 0.0) atomicAdd((double *)(acc_base + idx0 + b), sum);
        }
    }
    __syncthreads();
}

// The LDS regions of one trace wavefront (trace_lds_bytes).
struct TraceLds {
    void *hist;
    uint32_t *queue, *row32;
    Chain64 *chain;
    LensBand64 *band;
    uint64_t *masks;
    LambdaConsts *lc_lds;
};
template <int A>
__device__ __forceinline__ TraceLds trace_lds(uint32_t *lds, uint32_t hist_words, uint32_t row_words) {
    using Cell = typename Acc<A>::Cell;
    TraceLds L;
    const uint32_t hist_end = trace_hist_end_words(hist_words, sizeof(Cell));
    L.hist = lds;
    L.queue = lds + hist_end;
    L.row32 = lds + ((hist_end + kQueueCap + 1) & ~1u);
    L.chain = (Chain64 *)((char *)lds + trace_chain_offset(hist_words, row_words, A));
    L.band = (LensBand64 *)(L.chain + 1);
    L.masks = (uint64_t *)(L.band + 1);
    L.lc_lds = (LambdaConsts *)(L.masks + 2u * (kMaskChunk >> 12));
    return L;
}
template <int A>
__device__ __forceinline__ void trace_lds_clear(uint32_t *lds, uint32_t hist_words, uint32_t row_words) {
    using Cell = typename Acc<A>::Cell;
    const TraceLds L = trace_lds<A>(lds, hist_words, row_words);
    const uint32_t lane = threadIdx.x;
    for (uint32_t i = lane; i < hist_words; i += 64) ((Cell *)L.hist)[i] = 0;
    for (uint32_t i = lane; i < row_words * (Acc<A>::kRowBytes / 4); i += 64) L.row32[i] = 0;
    __syncthreads();
}

// Copies a record of 8-byte words from global memory into LDS, 64 lanes at a
// time (the caller synchronises before reading it).
__device__ __forceinline__ void copy_to_lds(void *dst, const void *src, uint32_t bytes) {
    const uint2 *s = (const uint2 *)src;
    uint2 *d = (uint2 *)dst;
    for (uint32_t i = threadIdx.x; i < bytes / 8u; i += 64) d[i] = s[i];
}

// One work item: wavelength sample j (at `lambda`) and ray tile `tile` of the
// design (g, td); its row is added to acc[acc_j + b] and the bin of the
// design's last ray, when the tile holds it, stored at *last_bin_j.  Args
// supplies the beam tables (zk, zs64, inv) and the LDS shape
// (hist_words, row_words); the LDS regions are derived here from the kernel's
// arguments, which keeps them out of the registers live across items.
template <bool COUNT, int A, bool REC, class Args>
__device__ __forceinline__ void run_item(const Args &a, uint32_t *lds, const cpd_design *g, const TraceDesign *td,
                                         uint32_t j, uint32_t tile, double lambda, const LambdaRecord *rec,
                                         unsigned long long *acc, uint64_t acc_j, uint32_t *last_bin_j,
                                         unsigned long long &ops) {
    const TraceLds L = trace_lds<A>(lds, a.hist_words, a.row_words);
    const uint32_t lane = threadIdx.x;
    const uint32_t M = td->n_rays, nbin = td->nbin;
    const uint32_t tile_rays = td->tile_rays, cols = td->hist_cols;
    TileCtx c;
    c.g = g;
    c.td = td;
    if constexpr (REC) {
        // the wavelength's records, built before the launch's rays
        (void)lambda;
        copy_to_lds(L.chain, &rec->chain, sizeof(Chain64));
        copy_to_lds(L.band, &rec->band, sizeof(LensBand64));
        copy_to_lds(L.lc_lds, &rec->lc, sizeof(LambdaConsts));
    } else {
        // every lane computes the same constants; lane 0 stores the records
        (void)rec;
        make_lambda_consts(*g, lambda, c.lc, lane == 0 ? L.chain : nullptr, lane == 0 ? L.band : nullptr);
        if (lane == 0) *L.lc_lds = c.lc;
    }
    __syncthreads();  // the queue drains and the band tests of this item read the records
    const bool lc_ok = REC ? L.lc_lds->ok != 0 && (COUNT || L.lc_lds->dead == 0) : c.lc.ok != 0;
    c.chain = L.chain;
    c.band = L.band;
    c.masks = L.masks;
    c.lc_lds = L.lc_lds;
    c.k_begin = tile * tile_rays + 1;
    // positions 1..P, the rays rounded up to whole banded blocks
    const uint32_t P = (M + kBandRays - 1u) / kBandRays * kBandRays;
    c.k_end = min(c.k_begin + tile_rays, P + 1u);
    c.n_rays = M;
    c.nbin = nbin;
    c.cols = cols;
    c.hist = L.hist;
    c.row = L.row32;
    c.queue = L.queue;
    c.last_bin = last_bin_j;
    c.zk = a.zk;
    c.zs64 = a.zs64;
    c.full_end = P;
    c.last_pos = a.inv[M - 1];
    if (!lc_ok || td->skip) {
        // every ray of this wavelength is lost in the planar faces or proved
        // lost beyond them, or the design is a stand-in for one that failed
        // construction
        if (c.last_pos >= c.k_begin && c.last_pos < c.k_end && lane == 0) set_last_bin(c.last_bin, nbin);
    } else {
        tile_loop<COUNT, A>(c, ops);
        flush_row<A>(L.hist, L.row32, nbin, cols, acc, acc_j);
    }
}

// REC: the wavelength records come from a.lrec (built by lambda_records_kernel
// before the launch) instead of each item building its own.  Six wavefronts
// per SIMD (80 registers; the ray loop holds no spill at that budget, and the
// LDS of 24 wavefronts fits a CU up to 128 bins).
template <bool COUNT, int A, bool REC>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void trace_kernel(BatchArgs a) {
    extern __shared__ __align__(8) uint32_t lds[];
    trace_lds_clear<A>(lds, a.hist_words, a.row_words);
    const uint32_t lane = threadIdx.x;
    unsigned long long ops = 0;
    uint64_t item = blockIdx.x;
    // The wavefront takes items from counter grp (its workgroup's index mod
    // kWorkCounters, which is also its XCD's); slots [next, end) of that
    // counter are its last fetch, slot q being item grp + kWorkCounters q.
    // Slots ascend, so the first item past n_items means the counter has none
    // left and the wavefront moves on to the next counter; it ends when all
    // kWorkCounters are spent.
    const uint32_t per_fetch = a.items_per_fetch > 1 ? a.items_per_fetch : 1u;
    uint32_t grp = blockIdx.x & (kWorkCounters - 1u), spent = 0;
    uint64_t next = 0, end = 0;
    for (;;) {
        if (a.work_counter) {
            item = a.n_items;
            while (spent < kWorkCounters) {
                if (next == end) {
                    uint32_t v = 0;
                    if (lane == 0) v = atomicAdd(a.work_counter + grp * kCounterStride, per_fetch);
                    next = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane(v);
                    end = next + per_fetch;
                }
                const uint64_t cand = grp + (uint64_t)kWorkCounters * next++;
                if (cand < a.n_items) {
                    item = cand;
                    break;
                }
                ++spent;
                grp = (grp + 1u) & (kWorkCounters - 1u);
                next = end = 0;
            }
        }
        if (item >= a.n_items) break;
        const uint32_t e = find_design(a.item_begin, a.n_designs, item);
        const uint32_t d = a.slot ? a.slot[e] : e;
        const cpd_design *g = a.designs + d;
        const TraceDesign *td = a.traces + d;
        const uint32_t n_tiles = td->n_tiles;
        const uint64_t local = item - a.item_begin[e] + (a.item_offset ? a.item_offset[e] : 0);
        const uint32_t j = (uint32_t)(local / n_tiles), tile = (uint32_t)(local % n_tiles);
        const uint64_t lb = a.lb_begin[d] + j;
        const double lambda = REC ? 0.0 : a.wavelengths ? a.wavelengths[j] : lambda_of(*g, qmc_u64(j + 1));
        const LambdaRecord *rec = nullptr;
        if (REC) {
            const uint64_t j0 = a.item_offset ? a.item_offset[e] / n_tiles : 0;  // first sample of the entry
            rec = a.lrec + (a.lrec_begin ? a.lrec_begin[e] + (j - j0) : lb);
        }
        run_item<COUNT, A, REC>(a, lds, g, td, j, tile, lambda, rec, a.acc, a.acc_begin[d] + (uint64_t)j * td->nbin,
                                a.last_bin + lb, ops);
        if (!a.work_counter) break;
    }
    if (COUNT) {
        for (int off = 32; off > 0; off >>= 1) ops += __shfl_xor(ops, off);
        if (lane == 0 && ops) atomicAdd(a.ops, ops);
    }
}

// The wavelength records of a launch (a.lrec): one workgroup per entry, one
// lane per wavelength sample the entry's items cover, each running the FP64
// make_lambda_consts an item would otherwise run on all 64 of its lanes.
__global__ __launch_bounds__(64) void lambda_records_kernel(BatchArgs a) {
    const uint32_t e = blockIdx.x;
    const uint32_t d = a.slot ? a.slot[e] : e;
    const cpd_design &g = a.designs[d];
    const TraceDesign &td = a.traces[d];
    const uint64_t n_items = a.item_begin[e + 1] - a.item_begin[e];
    if (n_items == 0 || td.n_tiles == 0) return;
    const uint64_t off = a.item_offset ? a.item_offset[e] : 0;
    const uint32_t j0 = (uint32_t)(off / td.n_tiles), j1 = (uint32_t)((off + n_items - 1) / td.n_tiles);
    // the heights of every ray the design traces: its positions 1..P are whole
    // blocks sorted by z, so their first and last positions bound them
    const uint32_t P = (td.n_rays + kBandRays - 1u) / kBandRays * kBandRays;
    double ya = INFINITY, yb = -INFINITY;
    for (uint32_t b = 0; b < P; b += kBandRays) {
        const double y0 = beam_y64(g, a.zs64[b]), y1 = beam_y64(g, a.zs64[b + kBandRays - 1u]);
        ya = fmin(ya, fmin(y0, y1));
        yb = fmax(yb, fmax(y0, y1));
    }
    for (uint32_t j = j0 + threadIdx.x; j <= j1; j += 64) {
        const double lambda = a.wavelengths ? a.wavelengths[j] : lambda_of(g, qmc_u64(j + 1));
        LambdaRecord &r = a.lrec[a.lrec_begin ? a.lrec_begin[e] + (j - j0) : a.lb_begin[d] + j];
        LambdaConsts lc;
        make_lambda_consts(g, lambda, lc, &r.chain, &r.band);
        // one band test over every height: a wavelength it proves lost (the
        // lens, its acceptance cosine or the array ends lose every ray) is not
        // traced at all (the items of a launch that counts operations still are)
        if (lc.ok && band_class(r.band, ya, yb) == kBandLost) lc.dead = 1;
        r.lc = lc;
    }
}

// p(d|l) of one accumulator cell (the bin of the last ray divides by M + 1).
template <int A>
__device__ double cell_p(unsigned long long raw, double div) {
    if (A == CPD_ACC_EXACT) return (double)raw * (1.0 / kQuantScale) / div;
    if (A == CPD_ACC_F64) return __longlong_as_double((long long)raw) / div;
    return (double)(__uint_as_float((uint32_t)raw) / (float)div);  // the cell's first four bytes
}

// The cells (j, b), j = j0 .. j0 + n - 1 (n <= kFinalizeBatch), of one bin and
// the last-ray bins of their wavelengths, loaded together before they are
// summed in order (the summation would otherwise wait for each load in turn).
constexpr uint32_t kFinalizeBatch = 16;
struct CellBatch {
    unsigned long long raw[kFinalizeBatch];
    uint32_t last[kFinalizeBatch];
};
__device__ __forceinline__ void load_cells(const unsigned long long *acc, const uint32_t *lb, uint32_t nbin, uint32_t b,
                                           uint32_t j0, uint32_t n, CellBatch &c) {
#pragma unroll
    for (uint32_t q = 0; q < kFinalizeBatch; ++q) {
        if (q < n) {
            c.raw[q] = acc[(uint64_t)(j0 + q) * nbin + b];
            c.last[q] = lb[j0 + q];
        }
    }
}
// One wavefront reduces a design's sums (N wavelengths x nbin cells at acc,
// the last ray's bins at lb) to its information, writes p(d|l) rows when
// `rows` is set, and leaves the cells zero and the last bins at nbin for the
// next use.  Every lane returns the information.
template <int A>
__device__ __forceinline__ double finalize_design(unsigned long long *acc, uint32_t *lb, double *rows, uint32_t N,
                                                  uint32_t M, uint32_t nbin) {
    const uint32_t lane = threadIdx.x;
    double info;
    if (A == CPD_ACC_EXACT || A == CPD_ACC_F64) {
        CSum h_cond, h_det;
        for (uint32_t b = lane; b < nbin; b += 64) {
            CSum pb;
            for (uint32_t j0 = 0; j0 < N; j0 += kFinalizeBatch) {
                const uint32_t n = min(kFinalizeBatch, N - j0);
                CellBatch cb;
                load_cells(acc, lb, nbin, b, j0, n, cb);
#pragma unroll
                for (uint32_t q = 0; q < kFinalizeBatch; ++q) {
                    if (q < n) {
                        const uint64_t idx = (uint64_t)(j0 + q) * nbin + b;
                        const double p = cell_p<A>(cb.raw[q], cb.last[q] == b ? (double)M + 1.0 : (double)M);
                        h_cond.add(plog2p(p));
                        pb.add(p);
                        if (rows) rows[idx] = p;
                        acc[idx] = 0;
                    }
                }
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
        info = (hs + hc) / (double)N - (ds + dc);
    } else if (A == CPD_ACC_F32) {
        float h_cond = 0.f, h_det = 0.f;
        for (uint32_t b = lane; b < nbin; b += 64) {
            float pb = 0.f;
            for (uint32_t j0 = 0; j0 < N; j0 += kFinalizeBatch) {
                const uint32_t n = min(kFinalizeBatch, N - j0);
                CellBatch cb;
                load_cells(acc, lb, nbin, b, j0, n, cb);
#pragma unroll
                for (uint32_t q = 0; q < kFinalizeBatch; ++q) {
                    if (q < n) {
                        const uint64_t idx = (uint64_t)(j0 + q) * nbin + b;
                        const float p = (float)cell_p<A>(cb.raw[q], cb.last[q] == b ? (double)M + 1.0 : (double)M);
                        h_cond += plog2pf(p);
                        pb += p;
                        if (rows) rows[idx] = p;
                        acc[idx] = 0;
                    }
                }
            }
            h_det += plog2pf(pb / (float)N);
        }
        for (int off = 32; off > 0; off >>= 1) {
            h_cond += __shfl_xor(h_cond, off);
            h_det += __shfl_xor(h_det, off);
        }
        info = (double)(h_cond / (float)N - h_det);
    } else {
        CSumT<float> h_cond, h_det;
        for (uint32_t b = lane; b < nbin; b += 64) {
            CSumT<float> pb;
            for (uint32_t j0 = 0; j0 < N; j0 += kFinalizeBatch) {
                const uint32_t n = min(kFinalizeBatch, N - j0);
                CellBatch cb;
                load_cells(acc, lb, nbin, b, j0, n, cb);
#pragma unroll
                for (uint32_t q = 0; q < kFinalizeBatch; ++q) {
                    if (q < n) {
                        const uint64_t idx = (uint64_t)(j0 + q) * nbin + b;
                        const float p = (float)cell_p<A>(cb.raw[q], cb.last[q] == b ? (double)M + 1.0 : (double)M);
                        h_cond.add(plog2pf(p));
                        pb.add(p);
                        if (rows) rows[idx] = p;
                        acc[idx] = 0;
                    }
                }
            }
            h_det.add(plog2pf(pb.value() / (float)N));
        }
        float hs = h_cond.s, hc = h_cond.c, ds = h_det.s, dc = h_det.c;
        for (int off = 32; off > 0; off >>= 1) {
            CSumT<float> x;
            x.s = hs;
            x.c = hc;
            x.add(__shfl_xor(hs, off));
            x.c += __shfl_xor(hc, off);
            hs = x.s;
            hc = x.c;
            CSumT<float> y;
            y.s = ds;
            y.c = dc;
            y.add(__shfl_xor(ds, off));
            y.c += __shfl_xor(dc, off);
            ds = y.s;
            dc = y.c;
        }
        info = (double)((hs + hc) / (float)N - (ds + dc));
    }
    __syncthreads();
    for (uint32_t j = lane; j < N; j += 64) lb[j] = nbin;
    return info;
}

template <int A>
__global__ __launch_bounds__(64) void finalize_kernel(FinalizeArgs a) {
    const uint32_t e = blockIdx.x;
    const uint32_t d = a.slot ? a.slot[e] : e;
    const uint32_t lane = threadIdx.x;
    const cpd_design &g = a.designs[d];
    const TraceDesign &td = a.traces[d];
    const double info = finalize_design<A>(a.acc + a.acc_begin[d], a.last_bin + a.lb_begin[d],
                                           a.rows ? a.rows + a.row_begin[d] : nullptr, td.n_lambda, td.n_rays,
                                           td.nbin);
    if (lane == 0 && a.out) {
        double size, deviation;
        size_and_deviation64(g, size, deviation);
        cpd_fitness f;
        f.size = size;
        f.info = info;
        f.deviation = deviation;
        a.out[a.out_stride_slot ? d : e] = f;
    }
}

// ---------------------------------------------------------------------------
// Device-resident request loop (ResidentArgs in kernels.h).
//
// Chunk c (at most 64 designs) sits at ring position c % chunk_cap.  Its
// items are handed out by kWorkCounters claim words, one per group of
// wavefronts (the workgroup index mod kWorkCounters, which is also its XCD),
// so the wavefronts of the chip do not all meet at one atomic: word g hands
// out items g, g + kWorkCounters, ... and packs (c mod 2^24, items it holds,
// items claimed) as bits 63..40, 39..20, 19..0.  Ingest stores the words
// last, with release, so a wavefront whose atomic add returns chunk c's word
// owns the item of its count in chunk c and sees c's tables, even if it meant
// to claim from an older chunk at that position.  A wavefront tries its own
// group's word first and moves on to the next when it is spent.  A chunk
// holds fewer than 2^18 items, and no count nears 2^20: past a word's total
// only claims that go on to another word add to it, and once every word is
// spent each wavefront adds to each at most once before the head moves on.
// Items per design are counted down
// in device memory; the wavefront that finishes a design's last item
// finalizes it.  Host memory is read only after a system-scope acquire and
// written only with plain stores followed by a system-scope release.
//
// Ordering: the counters are read and updated with relaxed atomics (an
// acquire at agent scope invalidates the wavefront's XCD L2, which costs the
// trace its cached beam tables; with acquiring loads on every claim the loop
// ran at half the throughput).  Data another wavefront wrote is read after one
// acquire fence, paired with that writer's release: after a claim (ingest
// released the chunk's tables and records before its claim word), in
// finalize (every worker released its sums before counting its item), and
// in ingest (the host released its records before `published`, which a
// poller passes on with release once it has acquired it).
constexpr uint32_t kResCountBits = 20;
constexpr unsigned long long kResCountMask = (1ull << kResCountBits) - 1;

__device__ __forceinline__ unsigned long long bcast64(unsigned long long v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}
// Relaxed: an acquire at agent scope invalidates the L2 of the wavefront's
// XCD, so the loop fences only where it needs data another wavefront wrote.
__device__ __forceinline__ unsigned long long agent_load(unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t agent_load(const uint64_t *p) {
    return __hip_atomic_load((uint64_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t agent_load32(const uint32_t *p) {
    return __hip_atomic_load((uint32_t *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A record of 8-byte words in device memory copied into LDS with agent-scope
// loads (the caller synchronises before reading it).
__device__ __forceinline__ void agent_copy_to_lds(void *dst, const void *src, uint32_t bytes) {
    const uint64_t *s = (const uint64_t *)src;
    uint64_t *d = (uint64_t *)dst;
    for (uint32_t i = threadIdx.x; i < bytes / 8u; i += 64) d[i] = agent_load(s + i);
}
static_assert(sizeof(cpd_design) % 16 == 0 && sizeof(TraceDesign) % 8 == 0,
              "record copies of 8-byte words, the trace record 16-byte aligned after the design");
// The LDS copies of the claimed design's records follow the trace regions.
__device__ __forceinline__ cpd_design *resident_design_lds(uint32_t *lds, uint32_t hist_words, uint32_t row_words) {
    return (cpd_design *)((char *)lds + resident_records_offset(hist_words, row_words));
}

// Copies chunk c's staged records (host memory) into the pools, builds its
// item table and publishes its claim word.  Returns false on an inconsistent
// descriptor (the chunk is then published empty and the host told).
__device__ __attribute__((noinline)) bool resident_ingest(const ResidentArgs &a, unsigned long long c) {
    const uint32_t lane = threadIdx.x;
    const uint32_t pos = (uint32_t)(c % a.chunk_cap);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");  // the host's descriptor and records
    const ResidentChunk *ch = a.chunks + pos;
    const unsigned long long seq1 = ch->seq1;
    uint32_t slot0 = ch->slot0, n = ch->n;
    const uint32_t tag = ch->tag;
    const bool ok = seq1 == c + 1 && n >= 1 && n <= 64 && slot0 < a.max_designs && n <= a.max_designs - slot0;
    if (!ok) n = 0;
    constexpr uint32_t WD = sizeof(cpd_design) / 4, WT = sizeof(TraceDesign) / 4;
    for (uint32_t w = lane; w < n * WD; w += 64) {
        const uint32_t q = w / WD, o = w - q * WD;
        ((uint32_t *)(a.designs + slot0 + q))[o] = ((const uint32_t *)&a.staged[slot0 + q].design)[o];
    }
    for (uint32_t w = lane; w < n * WT; w += 64) {
        const uint32_t q = w / WT, o = w - q * WT;
        ((uint32_t *)(a.traces + slot0 + q))[o] = ((const uint32_t *)&a.staged[slot0 + q].trace)[o];
    }
    uint32_t items = 0;
    if (lane < n) {
        const StagedDesign &sd = a.staged[slot0 + lane];
        items = sd.items;
        a.acc_begin[slot0 + lane] = sd.acc_begin;
        a.lb_begin[slot0 + lane] = sd.lb_begin;
        a.items_total[slot0 + lane] = items;
        a.items_done[slot0 + lane] = 0;
        a.slot_tag[slot0 + lane] = tag;
    }
    // inclusive prefix sums of the items over the chunk's designs
    uint32_t x = items;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t v = __shfl_up(x, off);
        if (lane >= off) x += v;
    }
    uint32_t *pre = a.prefix + (size_t)pos * 65;
    if (lane == 0) pre[0] = 0;
    pre[lane + 1] = x;
    const uint32_t total = __shfl(x, 63);
    a.chunk_slot0[pos] = slot0;
    a.chunk_n[pos] = n;
    const bool fits = total < (1u << 18);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane < kWorkCounters) {
        // word `lane`: items lane, lane + kWorkCounters, ... below the total
        const uint32_t held = fits ? (total + kWorkCounters - 1u - lane) / kWorkCounters : 0u;
        const unsigned long long word = ((c & 0xffffffull) << 40) | ((unsigned long long)held << kResCountBits);
        __hip_atomic_store(a.claim_word + (size_t)lane * a.chunk_cap + pos, word, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if (lane == 0) {
        __hip_atomic_store(a.ingested + pos, c + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        if (!ok || !fits) a.host->error = 1;
    }
    return ok && fits;
}

// The finished design in `slot`: information, size and deviation to host
// memory, then the request's tag.
__device__ __attribute__((noinline)) void resident_finalize(const ResidentArgs &a, uint32_t slot) {
    const uint32_t lane = threadIdx.x;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // every item's sums of this design
    const TraceDesign &td = a.traces[slot];
    const double info = finalize_design<CPD_ACC_EXACT>(a.acc + a.acc_begin[slot], a.last_bin + a.lb_begin[slot],
                                                       nullptr, td.n_lambda, td.n_rays, td.nbin);
    if (lane == 0) {
        double size, deviation;
        size_and_deviation64(a.designs[slot], size, deviation);
        ResidentResult *r = a.results + slot;
        r->fit.size = size;
        r->fit.info = info;
        r->fit.deviation = deviation;
    }
    // the result and the cleared sums reach memory before the tag
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
    if (lane == 0) __hip_atomic_store(&a.results[slot].tag, a.slot_tag[slot], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_SYSTEM);
}

// One work item of the resident loop, compiled as a function of its own: its
// registers are allocated apart from the loop's claim and ingest state, which
// inlined leaves the ray loop spilling.  A function's arguments arrive as
// per-lane values, so the wave-uniform ones are taken back into scalar
// registers, and the LDS is reached through the kernel's dynamic shared array
// (declared here as well), so its accesses stay LDS instructions.
struct ResidentTables {
    uint32_t hist_words, row_words;
    const uint2 *zk;
    const double *zs64;
    const uint32_t *inv;
};
__device__ __forceinline__ uint32_t uniform32(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane(v); }
// A wave-uniform pointer to global memory: the cast through the global
// address space lets the compiler use global (not flat) instructions, which
// LDS waits do not wait for.
template <class T> __device__ __forceinline__ T *uniform_ptr(T *p) {
    const uint64_t v = (uint64_t)p;
    using GT = __attribute__((address_space(1))) T;
    GT *q = (GT *)(((uint64_t)uniform32((uint32_t)(v >> 32)) << 32) | uniform32((uint32_t)v));
    return (T *)q;
}
__device__ __attribute__((noinline)) void resident_item(ResidentTables t, uint32_t j, uint32_t tile,
                                                        unsigned long long *acc, uint64_t acc_j, uint32_t *last_bin_j) {
    extern __shared__ __align__(8) uint32_t lds[];
    t.hist_words = uniform32(t.hist_words);
    t.row_words = uniform32(t.row_words);
    t.zk = uniform_ptr(t.zk);
    t.zs64 = uniform_ptr(t.zs64);
    t.inv = uniform_ptr(t.inv);
    j = uniform32(j);
    tile = uniform32(tile);
    acc = uniform_ptr(acc);
    acc_j = ((uint64_t)uniform32((uint32_t)(acc_j >> 32)) << 32) | uniform32((uint32_t)acc_j);
    last_bin_j = uniform_ptr(last_bin_j);
    const cpd_design *g = resident_design_lds(lds, t.hist_words, t.row_words);
    const TraceDesign *td = (const TraceDesign *)(g + 1);
    unsigned long long ops = 0;
    run_item<false, CPD_ACC_EXACT, false>(t, lds, g, td, j, tile, lambda_of(*g, qmc_u64(j + 1)), nullptr, acc, acc_j,
                                          last_bin_j, ops);
}

// Wavefronts 0..kResPollers-1 poll host memory (new chunks, the stop epoch)
// once per loop and copy what they read into device counters; the others read
// only those counters, so the host link carries a handful of polls.
constexpr uint32_t kResPollers = 4;

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(6))) void resident_kernel(ResidentArgs a) {
    extern __shared__ __align__(8) uint32_t lds[];
    trace_lds_clear<CPD_ACC_EXACT>(lds, a.hist_words, a.row_words);
    const uint32_t lane = threadIdx.x;
    const bool poller = blockIdx.x < kResPollers;
    uint32_t idle = 0;
    for (;;) {
        if (poller && lane == 0) {
            const unsigned int stop = __hip_atomic_load(&a.host->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const unsigned long long hp =
                __hip_atomic_load(&a.host->published, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            if (hp > agent_load(a.ctrl + kResPublished * kResLine)) {
                // new chunks: acquire the host's release, then pass it on
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "");
                __hip_atomic_fetch_max(a.ctrl + kResPublished * kResLine, hp, __ATOMIC_RELEASE,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            if (stop > __hip_atomic_load(a.ctrl + kResStop * kResLine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                __hip_atomic_fetch_max(a.ctrl + kResStop * kResLine, (unsigned long long)stop, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
        unsigned long long pub = 0, stop = 0;
        if (lane == 0) {
            pub = agent_load(a.ctrl + kResPublished * kResLine);
            stop = __hip_atomic_load(a.ctrl + kResStop * kResLine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        pub = bcast64(pub);
        if (bcast64(stop) >= a.epoch) break;
        // 1. ingest the oldest chunk nobody has taken
        unsigned long long took = 0;
        if (lane == 0) {
            unsigned long long ing = agent_load(a.ctrl + kResIngest * kResLine);
            if (ing < pub &&
                __hip_atomic_compare_exchange_strong(a.ctrl + kResIngest * kResLine, &ing, ing + 1, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                took = ing + 1;
        }
        took = bcast64(took);
        if (took) {
            resident_ingest(a, took - 1);
            idle = 0;
            continue;
        }
        // 2. claim an item of the head chunk
        // state: 0 nothing, 1 retry now, 2 item (chunk position pos, item x)
        unsigned long long st = 0, pos_x = 0;
        if (lane == 0) {
            unsigned long long h = agent_load(a.ctrl + kResHead * kResLine);
            if (h < pub) {
                const uint32_t pos = (uint32_t)(h % a.chunk_cap);
                const unsigned long long ing = agent_load(a.ingested + pos);
                if (ing > h + 1) {
                    // the position holds a later chunk: chunk h is complete
                    __hip_atomic_compare_exchange_strong(a.ctrl + kResHead * kResLine, &h, h + 1, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    st = 1;
                } else if (ing == h + 1) {
                    // the group's own word first, then the others
                    uint32_t grp = blockIdx.x & (kWorkCounters - 1u);
                    bool spent = true;  // every word of chunk h is spent
                    for (uint32_t t = 0; t < kWorkCounters; ++t, grp = (grp + 1u) & (kWorkCounters - 1u)) {
                        const unsigned long long w = __hip_atomic_fetch_add(
                            a.claim_word + (size_t)grp * a.chunk_cap + pos, 1ull, __ATOMIC_RELAXED,
                            __HIP_MEMORY_SCOPE_AGENT);
                        const unsigned long long x = w & kResCountMask, held = (w >> kResCountBits) & kResCountMask;
                        if (x < held) {
                            st = 2;
                            pos_x = ((unsigned long long)pos << 32) | (grp + (unsigned long long)kWorkCounters * x);
                            break;
                        }
                        if ((w >> 40) != (h & 0xffffffull)) {
                            spent = false;  // the position holds another chunk now
                            break;
                        }
                    }
                    if (st != 2) {
                        if (spent)
                            __hip_atomic_compare_exchange_strong(a.ctrl + kResHead * kResLine, &h, h + 1, __ATOMIC_RELAXED,
                                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        st = 1;
                    }
                }
                // ing < h + 1: chunk h is being ingested
            }
        }
        st = bcast64(st);
        if (st == 1) {
            idle = 0;
            continue;
        }
        if (st == 0) {
            // nothing to claim: back off (at most ~30 us between polls), and
            // exit after idle_limit polls even if the host never asks
            if (++idle > a.idle_limit) break;
            for (uint32_t k = 0, n = idle < 8 ? idle : 8; k < n; ++k) __builtin_amdgcn_s_sleep(127);
            continue;
        }
        idle = 0;
        pos_x = bcast64(pos_x);
        // No acquire fence here: one at agent scope would invalidate the L2 of
        // this wavefront's XCD (the beam tables the other wavefronts read
        // included) at every claim.  What the ingesting wavefront wrote before
        // releasing the claim word -- the chunk's tables and the slot's records
        // -- is read with agent-scope loads instead, which are coherent across
        // the XCDs, and the records are copied into LDS for the item.
        const uint32_t pos = (uint32_t)(pos_x >> 32), x = (uint32_t)pos_x;
        const uint32_t n = agent_load32(a.chunk_n + pos);
        const uint32_t slot0 = agent_load32(a.chunk_slot0 + pos);
        // the design holding item x: the number of designs whose items end at
        // or before x (pre_l: lane's inclusive prefix)
        const uint32_t pre_l = lane < n ? agent_load32(a.prefix + (size_t)pos * 65 + lane + 1) : 0xffffffffu;
        const uint64_t before = __builtin_amdgcn_ballot_w64(lane < n && pre_l <= x);
        const uint32_t q = (uint32_t)__builtin_popcountll(before);
        const uint32_t slot = slot0 + q;
        if (q >= n || slot >= a.max_designs) {
            if (lane == 0) a.host->error = 3;
            continue;
        }
        const uint32_t pre_q = q ? (uint32_t)__shfl(pre_l, (int)q - 1) : 0u;  // items before design q
        cpd_design *g = resident_design_lds(lds, a.hist_words, a.row_words);
        TraceDesign *td = (TraceDesign *)(g + 1);
        agent_copy_to_lds(g, a.designs + slot, sizeof(cpd_design));
        agent_copy_to_lds(td, a.traces + slot, sizeof(TraceDesign));
        uint64_t acc_b = 0, lb_b = 0;
        if (lane == 0) {
            acc_b = agent_load(a.acc_begin + slot);
            lb_b = agent_load(a.lb_begin + slot);
        }
        acc_b = bcast64(acc_b);
        lb_b = bcast64(lb_b);
        __syncthreads();  // the record copies
        const uint32_t local = x - pre_q;
        const uint32_t n_tiles = td->n_tiles;
        const uint32_t j = local / n_tiles, tile = local - j * n_tiles;
        if (j < td->n_lambda) {
            resident_item(ResidentTables{a.hist_words, a.row_words, a.zk, a.zs64, a.inv}, j, tile, a.acc,
                          acc_b + (uint64_t)j * td->nbin, a.last_bin + lb_b + j);
        } else if (lane == 0) {
            a.host->error = 2;
        }
        // The item's outputs are agent-scope atomics (its sums) and an
        // agent-scope store (its last bin), performed on the memory side; once
        // they have completed, the count below can hand the design to its
        // finalizer.  A release fence at agent scope would instead write back
        // this XCD's whole L2 at every item (buffer_wbl2), which stalls every
        // wavefront of the XCD: it halved the loop's throughput.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_s_waitcnt(0);
        unsigned long long last = 0;
        if (lane == 0) {
            const unsigned int done =
                __hip_atomic_fetch_add(a.items_done + slot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
            last = done == agent_load32(a.items_total + slot);
        }
        if (bcast64(last)) resident_finalize(a, slot);
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
    Trace32Result r{td->nbin, 0.f, false, false, 0};
    if (lc.ok) {
        const float y = td->y_mean - td->beam_w * (float)z;
        r = trace32<false>(make_hot(*td, lc), y);
        if (!r.hit) r.bin = td->nbin;
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
                                    uint32_t tile_cap, TraceDesign *traces, const int *status) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    TraceDesign t;
    make_trace_design(designs[i], sampling[i].n_wavelengths, sampling[i].n_rays, t, tile_cap);
    if (status && status[i] != CPD_OK) t.skip = 1;
    traces[i] = t;
}

}  // namespace

namespace {
// One thread per design (construct.h, FP64).
__global__ __launch_bounds__(64) void construct_kernel(const cpd_params *params, uint32_t n, const double *acceptance,
                                                       const cpd_design *standin, cpd_design *out, int *status) {
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    cpd_design d;
    cpd_status st = construct_design(params[i], *acceptance, d);
    if (standin) {
        if (st == CPD_OK && d.bin_count != standin[i].bin_count) st = CPD_ERR_INVALID_ARGUMENT;
        if (st != CPD_OK) d = standin[i];
    }
    out[i] = d;
    status[i] = st;
}
}  // namespace

hipError_t launch_construct(hipStream_t s, const cpd_params *params, uint32_t n, const double *acceptance,
                            const cpd_design *standin, cpd_design *out, int *status) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(construct_kernel, dim3((n + 63) / 64), dim3(64), 0, s, params, n, acceptance, standin, out,
                       status);
    return hipGetLastError();
}

hipError_t launch_build_traces(hipStream_t s, const cpd_design *designs, const cpd_sampling *sampling, uint32_t n,
                               uint32_t tile_cap, TraceDesign *traces, const int *status) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(build_traces_kernel, dim3((n + 63) / 64), dim3(64), 0, s, designs, sampling, n, tile_cap,
                       traces, status);
    return hipGetLastError();
}


namespace {
template <bool COUNT, int A, bool REC>
hipError_t launch_trace_r(hipStream_t s, const BatchArgs &a, uint32_t grid, hipEvent_t start, hipEvent_t stop) {
    const size_t lds = trace_lds_bytes(a.hist_words, a.row_words, A);
    if (!start && !stop)
        // plain launch: the form a stream capture records as a kernel node
        hipLaunchKernelGGL((trace_kernel<COUNT, A, REC>), dim3(grid), dim3(64), lds, s, a);
    else
        hipExtLaunchKernelGGL((trace_kernel<COUNT, A, REC>), dim3(grid), dim3(64), lds, s, start, stop, 0, a);
    return hipGetLastError();
}
template <bool COUNT, int A>
hipError_t launch_trace_t(hipStream_t s, const BatchArgs &a, uint32_t grid, hipEvent_t start, hipEvent_t stop) {
    if (a.lrec) {
        // the wavelength records first (outside the events, which time the trace)
        hipLaunchKernelGGL(lambda_records_kernel, dim3(a.n_designs), dim3(64), 0, s, a);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        return launch_trace_r<COUNT, A, true>(s, a, grid, start, stop);
    }
    // only the batches without record storage (the allocation-free entry
    // point) build the records per item; they do not count operations
    if (COUNT) return hipErrorInvalidValue;
    return launch_trace_r<false, A, false>(s, a, grid, start, stop);
}
template <bool COUNT>
hipError_t launch_trace_c(hipStream_t s, const BatchArgs &a, uint32_t grid, int acc, hipEvent_t start,
                          hipEvent_t stop) {
    switch (acc) {
        case CPD_ACC_F32: return launch_trace_t<COUNT, CPD_ACC_F32>(s, a, grid, start, stop);
        case CPD_ACC_F32C: return launch_trace_t<COUNT, CPD_ACC_F32C>(s, a, grid, start, stop);
        case CPD_ACC_F64: return launch_trace_t<COUNT, CPD_ACC_F64>(s, a, grid, start, stop);
        default: return launch_trace_t<COUNT, CPD_ACC_EXACT>(s, a, grid, start, stop);
    }
}
}  // namespace

size_t trace_lds_bytes(uint32_t hist_words, uint32_t row_words, int acc) {
    return trace_lds_bytes_d(hist_words, row_words, acc);
}

hipError_t launch_trace(hipStream_t s, const BatchArgs &a, uint32_t grid, bool count_ops, int acc, hipEvent_t start,
                        hipEvent_t stop) {
    if (grid == 0 || a.n_items == 0) return hipSuccess;
    return count_ops ? launch_trace_c<true>(s, a, grid, acc, start, stop)
                     : launch_trace_c<false>(s, a, grid, acc, start, stop);
}

hipError_t launch_finalize(hipStream_t s, const FinalizeArgs &a, int acc) {
    if (a.n_designs == 0) return hipSuccess;
    switch (acc) {
        case CPD_ACC_F32: hipLaunchKernelGGL(finalize_kernel<CPD_ACC_F32>, dim3(a.n_designs), dim3(64), 0, s, a); break;
        case CPD_ACC_F32C: hipLaunchKernelGGL(finalize_kernel<CPD_ACC_F32C>, dim3(a.n_designs), dim3(64), 0, s, a); break;
        case CPD_ACC_F64: hipLaunchKernelGGL(finalize_kernel<CPD_ACC_F64>, dim3(a.n_designs), dim3(64), 0, s, a); break;
        default: hipLaunchKernelGGL(finalize_kernel<CPD_ACC_EXACT>, dim3(a.n_designs), dim3(64), 0, s, a); break;
    }
    return hipGetLastError();
}

size_t resident_lds_bytes(uint32_t hist_words, uint32_t row_words) {
    return resident_records_offset(hist_words, row_words) + sizeof(cpd_design) + sizeof(TraceDesign);
}

hipError_t resident_occupancy(size_t lds, int *per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, resident_kernel, 64, lds);
}

hipError_t launch_resident(hipStream_t s, const ResidentArgs &a, uint32_t grid) {
    if (grid == 0) return hipSuccess;
    hipLaunchKernelGGL(resident_kernel, dim3(grid), dim3(64), resident_lds_bytes(a.hist_words, a.row_words), s, a);
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
