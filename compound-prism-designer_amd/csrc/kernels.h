// kernels.h -- device-side argument blocks and kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cpd.h"
#include "trace_core.h"

namespace cpd {

// Fixed-point scale of the per-bin transmittance sums: a ray adds
// round(t * 2^22) to its bin, so sums are exact integers and independent of
// the order in which rays and tiles are added.
static constexpr double kQuantScale = 4194304.0;
static constexpr int kQueueCap = 128;   // per-wave queue of rays for the FP64 re-trace
static constexpr uint32_t kMaxTileRays = 16384;

// One batch on the device.  Designs are laid out as parallel arrays; the work
// of design d is items [item_begin[d], item_begin[d+1]): wavelength sample j
// and ray tile t of tile_rays rays, item = item_begin[d] + j * n_tiles + t.
// Per-(design, wavelength, bin) sums live at acc[acc_begin[d] + j * nbin + b];
// the bin of ray k = n_rays (the last ray, which the reference counts once
// more, src/fitness.rs:20-32) at last_bin[lb_begin[d] + j].
struct BatchArgs {
    const cpd_design *designs;
    const TraceDesign *traces;
    const uint64_t *item_begin;  // [n_designs + 1]
    const uint64_t *acc_begin;   // [n_designs]
    const uint64_t *lb_begin;    // [n_designs]
    unsigned long long *acc;
    uint32_t *last_bin;
    const float *z32;            // beam quantiles z_k = norminv(u_k), k = 1..n, FP32
    const double *z64;           // the same in FP64
    unsigned long long *ops;     // algorithmic operation count (may be null)
    unsigned int *work_counter;  // persistent scheduling (null: one item per workgroup)
    const double *wavelengths;   // explicit wavelength per sample j (null: the R1 samples)
    uint32_t n_designs;
    uint32_t tile_rays;          // multiple of 64, <= kMaxTileRays
    uint32_t cols;               // LDS histogram columns, 64 or 32
    uint32_t nbin_max;
    uint64_t n_items;
};

struct FinalizeArgs {
    const cpd_design *designs;
    const TraceDesign *traces;
    const uint64_t *acc_begin;
    const uint64_t *lb_begin;
    unsigned long long *acc;
    uint32_t *last_bin;
    cpd_fitness *out;            // [n_designs] (may be null)
    double *rows;                // p(d|l) rows, [sum n_lambda * nbin] (may be null)
    const uint64_t *row_begin;   // output row offsets (with rows)
    uint32_t n_designs;
};

size_t trace_lds_bytes(uint32_t nbin_max, uint32_t cols);

hipError_t launch_beam_table(hipStream_t s, uint32_t n, float *z32, double *z64);
// start/stop (may be null) are recorded immediately around the kernel.
hipError_t launch_trace(hipStream_t s, const BatchArgs &a, uint32_t grid, bool count_ops, hipEvent_t start = nullptr,
                        hipEvent_t stop = nullptr);
hipError_t launch_finalize(hipStream_t s, const FinalizeArgs &a);
hipError_t launch_zero_u32(hipStream_t s, unsigned int *p, uint32_t n);
// Single rays: wavelength from its CDF value, beam height from its CDF value.
hipError_t launch_propagate(hipStream_t s, const cpd_design *g, const TraceDesign *td, const float *wcdf,
                            const float *rcdf, uint32_t n, uint32_t *bin, float *t);

}  // namespace cpd
