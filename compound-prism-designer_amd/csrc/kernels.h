// kernels.h -- device-side argument blocks and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cpd.h"
#include "trace_core.h"

namespace cpd {

// Fixed-point scale of the per-bin transmittance sums of the default (exact)
// accumulation: a ray adds round(t * 2^22) to its bin, so sums are exact
// integers and independent of the order in which rays and tiles are added.
static constexpr double kQuantScale = 4194304.0;
static constexpr int kQueueCap = 128;   // per-wave queue of rays for the FP64 re-trace
// Positions of a chunk, whose bands' skip masks are set at once before their
// rays (one 64-bit mask per 4096 positions in LDS, four words; four more hold
// the tile's lost sub-bands): four blocks of 4096.
static constexpr uint32_t kMaskChunk = 16384;
// The persistent grid's work counters: kWorkCounters of them, kCounterStride
// u32 apart (a 128-byte line each); counter x hands out items x,
// x + kWorkCounters, x + 2 kWorkCounters, ... so that the wavefronts of the
// chip do not all meet at one atomic (one counter for the 1024-design batch
// held it back by a fifth).  A launch clears kWorkCounterWords words.
static constexpr uint32_t kWorkCounters = 8, kCounterStride = 32;
static constexpr uint32_t kWorkCounterWords = kWorkCounters * kCounterStride;
// Entries past the last ray of the FP32 beam tables (z32, zk), kept readable
// (zk's carry a ray index no design reaches).
static constexpr uint32_t kBeamPad = 64;

// One launch's work on the device.  Design records live in a pool
// (designs/traces, indexed by slot); the launch covers n_entries entries,
// entry e being pool slot slot[e] (e itself when slot is null).  The work of
// entry e is launch items [item_begin[e], item_begin[e+1]), which are items
// item_offset[e] + 0, 1, ... of its design (offset 0 when item_offset is
// null).  Item i of a design is wavelength sample j = i / n_tiles and ray
// tile i % n_tiles of tile_rays rays (both from its trace record).
// Per-(design, wavelength, bin) sums live in 8-byte cells at
// acc[acc_begin[slot] + j * nbin + b] (u64, or the FP32 / FP64 value of the
// other accumulation policies); the bin of ray k = n_rays (the last ray,
// which the reference counts once more, src/fitness.rs:20-32) at
// last_bin[lb_begin[slot] + j].
struct BatchArgs {
    const cpd_design *designs;
    const TraceDesign *traces;
    const uint32_t *slot;        // [n_entries] (may be null)
    const uint64_t *item_begin;  // [n_entries + 1]
    const uint64_t *item_offset; // [n_entries] (may be null)
    const uint64_t *acc_begin;   // [pool slots]
    const uint64_t *lb_begin;    // [pool slots]
    unsigned long long *acc;
    uint32_t *last_bin;
    const float *z32;            // beam quantiles z_k = norminv(u_k), k = 1..n, FP32
    const double *z64;           // the same in FP64
    const uint2 *zk;             // banded (within blocks of kBandRays rays sorted by z): (FP32 z bits, ray index k)
    const double *zs64;          // banded, FP64
    const uint32_t *inv;         // 1-based banded position of ray k (index k - 1)
    unsigned long long *ops;     // algorithmic operation count (may be null)
    unsigned int *work_counter;  // persistent scheduling: kWorkCounters counters (null: one item per workgroup)
    uint32_t items_per_fetch;    // consecutive items one counter fetch claims (0: 1)
    const double *wavelengths;   // explicit wavelength per sample j (null: the R1 samples)
    // per-wavelength records, which launch_trace fills before the trace (null:
    // each work item builds its own): those of entry e at
    // lrec[lrec_begin[e] + j - j0], j0 = item_offset[e] / n_tiles the first
    // wavelength sample its items cover, or at lrec[lb_begin[slot] + j] when
    // lrec_begin is null
    LambdaRecord *lrec;
    const uint64_t *lrec_begin;  // [n_entries] (may be null)
    uint32_t n_designs;          // entries
    uint32_t hist_words;         // max over the entries of nbin * hist_cols
    uint32_t row_words;          // max over the entries of nbin
    uint64_t n_items;
};

struct FinalizeArgs {
    const cpd_design *designs;
    const TraceDesign *traces;
    const uint32_t *slot;        // [n_designs] pool slots to finalize (may be null: 0..n-1)
    uint32_t out_stride_slot;    // results indexed by slot (1) or by entry (0)
    const uint64_t *acc_begin;
    const uint64_t *lb_begin;
    unsigned long long *acc;
    uint32_t *last_bin;
    cpd_fitness *out;            // [n_designs] (may be null)
    double *rows;                // p(d|l) rows, [sum n_lambda * nbin] (may be null)
    const uint64_t *row_begin;   // output row offsets (with rows)
    uint32_t n_designs;
};

// ---------------------------------------------------------------------------
// Device-resident request loop (the server's resident mode).
//
// The host stages requests in host memory mapped into the GPU's address
// space (coherent, never cached stale by the GPU) and publishes them as
// chunks of at most 64 designs; a persistent kernel whose wavefronts loop
// until told to stop, or until they find no work for idle_limit polls,
// ingests chunks into the device pools, claims their work items, traces them
// and finalizes each design when its last item is done, writing the result
// and the request's tag back to host memory.  Only plain loads and stores
// touch host memory; the counters the wavefronts share live in device memory.

struct ResidentHost {                  // control block, host memory
    unsigned long long published;     // chunks the host has written (host stores, release)
    unsigned int stop;                // launches with epoch <= stop exit at their next poll
    unsigned int error;               // nonzero: a wavefront met an inconsistent chunk (device stores)
};
struct ResidentChunk {                 // descriptor of chunk c at ring position c % chunk_cap, host memory
    unsigned long long seq1;          // c + 1 (0: never written)
    uint32_t slot0, n;                // pool slots slot0 .. slot0 + n - 1, n <= 64
    uint32_t tag, pad;                // the request's tag, written per slot when a design is done
};
struct alignas(16) StagedDesign {      // one design as the host stages it (host memory, indexed by slot)
    cpd_design design;
    TraceDesign trace;
    uint64_t acc_begin, lb_begin;     // its accumulator and last-bin offsets
    uint32_t items, pad[3];           // work items (wavelengths x tiles)
};
struct ResidentResult {                // per slot, host memory
    cpd_fitness fit;
    uint32_t tag, pad;                // written after fit: the request whose design finished here
};

// device counters: chunks taken for ingest, the head chunk claims come from,
// the newest `published` and stop epoch the polling wavefronts have read;
// counter k at ctrl[k * kResLine] (a 128-byte line each)
enum { kResIngest = 0, kResHead = 1, kResPublished = 2, kResStop = 3, kResCtrlWords = 4 };
static constexpr uint32_t kResLine = 16;

struct ResidentArgs {
    ResidentHost *host;
    const ResidentChunk *chunks;      // [chunk_cap]
    const StagedDesign *staged;       // [max_designs]
    ResidentResult *results;          // [max_designs]
    uint32_t chunk_cap, max_designs;
    // device memory
    cpd_design *designs;
    TraceDesign *traces;
    uint64_t *acc_begin, *lb_begin;
    unsigned long long *acc;
    uint32_t *last_bin;
    unsigned int *items_total, *items_done;  // per slot
    uint32_t *slot_tag;                      // per slot
    unsigned long long *ingested;            // per ring position: c + 1 once chunk c's records are in the pools
    unsigned long long *claim_word;          // [kWorkCounters][chunk_cap]: (c mod 2^24, items, claimed) per group
    uint32_t *chunk_slot0, *chunk_n;         // per ring position
    uint32_t *prefix;                        // per ring position: 65 inclusive item prefix sums over its designs
    unsigned long long *ctrl;                // kRes* counters
    const float *z32;
    const uint2 *zk;
    const double *z64, *zs64;
    const uint32_t *inv;
    uint32_t max_rays;                       // beam table length
    uint32_t hist_words, row_words;
    uint32_t idle_limit;                     // polls without work before a wavefront exits
    uint32_t epoch;                          // this launch's number (from 1)
};
hipError_t launch_resident(hipStream_t s, const ResidentArgs &a, uint32_t grid);
hipError_t resident_occupancy(size_t lds, int *per_cu);
size_t resident_lds_bytes(uint32_t hist_words, uint32_t row_words);

// LDS of one trace wavefront: columns | flagged-ray queue | row; sizes of the
// accumulation policy `acc` (CPD_ACC_*).
CPD_HD uint32_t trace_hist_end_words(uint32_t hist_words, uint32_t cell_bytes) {
    return (hist_words * (cell_bytes / 4) + 1) & ~1u;
}
// Items one fetch of the persistent grid's work counter claims: enough that a
// fetch covers ~16384 rays on average (at most 16), because short rows
// otherwise serialise on the counter; one for long rows, for the balance;
// at most items / (4 grid), so every wavefront still fetches several times;
// and never so many that the u32 counter could wrap.
inline uint32_t items_per_fetch(uint64_t rays, uint64_t items, uint32_t grid) {
    if (items == 0 || grid == 0) return 1;
    const uint64_t per = rays / items;
    uint64_t g = per >= 16384 ? 1 : 16384 / (per ? per : 1);
    if (g > 16) g = 16;
    const uint64_t balance = items / (4 * (uint64_t)grid);
    if (g > balance) g = balance ? balance : 1;
    while (g > 1 && items + (uint64_t)grid * g >= (1ull << 32)) g >>= 1;
    return (uint32_t)g;
}

// Byte offset of the wavefront's chain record (Chain64), which follows the
// histogram, the queue and the row.
CPD_HD uint32_t trace_chain_offset(uint32_t hist_words, uint32_t row_words, int acc) {
    const uint32_t cell = acc == CPD_ACC_F64 ? 8 : 4;
    const uint32_t row = acc == CPD_ACC_F32 ? 4 : 8;
    const uint32_t end = trace_hist_end_words(hist_words, cell);
    return (((end + kQueueCap + 1) & ~1u) * 4 + row_words * row + 7) & ~7u;
}
CPD_HD inline size_t trace_lds_bytes_d(uint32_t hist_words, uint32_t row_words, int acc) {
    return (size_t)trace_chain_offset(hist_words, row_words, acc) + sizeof(Chain64) + sizeof(LensBand64) +
           2 * (kMaskChunk >> 12) * sizeof(uint64_t) + sizeof(LambdaConsts);
}
size_t trace_lds_bytes(uint32_t hist_words, uint32_t row_words, int acc);
// The resident loop's LDS: the trace regions, then copies of the claimed
// design's records (cpd_design, TraceDesign) at this 16-byte aligned offset.
CPD_HD inline size_t resident_records_offset(uint32_t hist_words, uint32_t row_words) {
    return (trace_lds_bytes_d(hist_words, row_words, CPD_ACC_EXACT) + 15) & ~(size_t)15;
}

// start/stop (may be null) are recorded immediately around the kernel; with
// neither, the launch is a plain one that stream capture records.
hipError_t launch_trace(hipStream_t s, const BatchArgs &a, uint32_t grid, bool count_ops, int acc = CPD_ACC_EXACT,
                        hipEvent_t start = nullptr, hipEvent_t stop = nullptr);
hipError_t launch_finalize(hipStream_t s, const FinalizeArgs &a, int acc = CPD_ACC_EXACT);
hipError_t launch_build_traces(hipStream_t s, const cpd_design *designs, const cpd_sampling *sampling, uint32_t n,
                               uint32_t tile_cap, TraceDesign *traces, const int *status = nullptr);
hipError_t launch_zero_u32(hipStream_t s, unsigned int *p, uint32_t n);
// Single rays: wavelength from its CDF value, beam height from its CDF value.
hipError_t launch_propagate(hipStream_t s, const cpd_design *g, const TraceDesign *td, const float *wcdf,
                            const float *rcdf, uint32_t n, uint32_t *bin, float *t);
// Designs built on the device from their parameters (construct_design):
// out[i] and status[i] for i < n, the detector positioned when *acceptance
// (a device scalar) is > 0.  With standin non-null, a design that fails, or
// whose bin count differs from standin[i]'s, is replaced by standin[i], so a
// launch over the outputs stays well formed.
hipError_t launch_construct(hipStream_t s, const cpd_params *params, uint32_t n, const double *acceptance,
                            const cpd_design *standin, cpd_design *out, int *status);

}  // namespace cpd
