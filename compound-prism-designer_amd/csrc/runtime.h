// runtime.h -- device context of the fitness evaluator.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <vector>

#include "cpd.h"
#include "trace_core.h"

namespace cpd {

struct RowSlots;

// Grow-only device allocation.
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    cpd_status reserve(size_t want, bool zero, hipStream_t s);
    void release();
    template <class T> T *as() const { return (T *)ptr; }
};

// Grow-only pinned host allocation (staging for one-copy uploads).
struct HostBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    cpd_status reserve(size_t want);
    void release();
};

// Beam quantiles z_k = norminv(u_k), k = 1..n (n a multiple of kBandRays), of
// the R1 sequence from 0.5: in ray order (FP32, FP64), and banded: within
// each block of kBandRays rays sorted by z (so by beam height), as FP64 and
// as (FP32 z bits, k) pairs, with inv[k - 1] the 1-based position of ray k in
// the banded tables.  A wavefront that traces 64 consecutive banded positions
// covers a narrow band of heights, so whole wavefronts miss together; the
// per-bin sums do not depend on the order rays are added in.  A design traces
// whole blocks; in its last one the positions whose ray index exceeds its ray
// count are idle.
struct BeamTables {
    uint32_t n = 0;
    DeviceBuffer z32, z64, zk, zs64, inv;
    cpd_status ensure(uint32_t want, hipStream_t s);  // grow-only, at least 65536 rays
    void release();
};

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;  // serialises calls on this context
    int cu_count = 0;
    int clock_khz = 0;
    char name[256] = {0};
    uint32_t waves_per_cu = 24;   // resident wavefronts per CU of the persistent grid (six per SIMD)
    uint32_t tile_rays_hint = 0;  // rays per work item (0: the default)
    uint32_t order = CPD_ORDER_GIVEN;
    int acc_kind = CPD_ACC_EXACT;  // accumulation policy of fixed-count batches
    int row_trace = 0;             // adaptive rows: 0 per-estimator default, 32 or 64
    bool sem_fp32() const { return row_trace == 32; }
    bool chunk_fp32() const { return row_trace != 64; }
    double row_kernel_ms = 0.0;    // row kernels of the last adaptive run (CPD_OPT_PROFILE)
    bool profile = false;
    bool count_ops = false;
    cpd_run_stats last{};
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};

    BeamTables beam;
    // batch state
    DeviceBuffer stage, acc, last_bin, results, counters, row_buf, prop_buf, build_buf;
    DeviceBuffer lrec;         // per-(design, wavelength) records of a batch (LambdaRecord)
    // adaptive estimators: row descriptors, design records, seeded beam tables
    DeviceBuffer row_stage, design_stage, seed_z32, seed_z64;
    HostBuffer host_stage, host_out;
    HostBuffer row_args_pin, rows_pin;  // adaptive rows: descriptors up, per-bin means and ray counts down

    static cpd_status create(int device, Context **out);
    ~Context();

    cpd_status ensure_beam_table(uint32_t n);
    cpd_status run(const cpd_design *designs, const cpd_sampling *sampling, const std::vector<uint32_t> &valid,
                   bool persistent, const double *wavelengths, cpd_fitness *out, double *rows,
                   unsigned long long *ops_out);
    cpd_status build_designs(const cpd_params *params, size_t n, double acceptance, cpd_design *out,
                             cpd_status *status);
    cpd_status propagate(const cpd_design &g, const float *wcdf, const float *rcdf, size_t n, uint32_t *bin,
                         float *t);
    // adaptive estimators (adaptive.hip)
    cpd_status run_adaptive(const cpd_design *designs, const cpd_sampling *sampling,
                            const std::vector<uint32_t> &valid, cpd_fitness *out);
    cpd_status rows_adaptive(const cpd_design &g, const double *wavelengths, size_t n_w, uint32_t max_m,
                             double *out);
    cpd_status run_gpu_fitness(const cpd_design *designs, const std::vector<uint32_t> &valid, const double *seeds,
                               size_t n_seeds, uint32_t max_n, uint32_t nwarp, uint32_t max_eval, cpd_fitness *out,
                               cpd_status *per_design);
    cpd_status run_rows(bool chunked, bool fp32, const cpd_design *designs_dev, const std::vector<uint32_t> &row_design,
                        const std::vector<double> &row_lambda, const std::vector<uint64_t> &row_out,
                        const std::vector<uint32_t> &row_cap, uint64_t out_n, uint32_t max_rays, uint32_t nbin_max,
                        const double *z64_dev, const float *z32_dev, const double *&rows_host,
                        const uint64_t *&rays_host, RowSlots *dev_slots = nullptr);
    DeviceBuffer fold_buf;     // the GPU estimator's outer fold state
    DeviceBuffer fit_buf;      // glass fits (cpd_fit_sellmeier1_batch)
    cpd_status stage_designs(const cpd_design *designs, const std::vector<uint32_t> &valid,
                             const cpd_design **designs_dev, uint32_t *nbin_max);
    DeviceBuffer recip;        // [k] = RN(1/k) for the row folds' integer divisors
    uint32_t recip_n = 0;
    cpd_status ensure_recip(uint32_t n);
    cpd_status stage_rows(const std::vector<uint32_t> &row_design, const std::vector<double> &row_lambda,
                          const std::vector<uint64_t> &row_out, const std::vector<uint32_t> &row_cap, uint64_t out_n,
                          uint32_t max_rays, uint32_t nbin_max, double max_err_sq, RowSlots &slots);
    cpd_status finish_rows(const RowSlots &slots, uint64_t out_n, const double *&rows_host,
                           const uint64_t *&rays_host);
    // 3D spectrometers (rows3.hip)
    DeviceBuffer design3_stage, samples3, samples3_gauss;  // records; beam samples of fiber and Gaussian rays
    DeviceBuffer entry3;                                    // entrance hits of the staged designs' rays
    uint64_t samples3_key = ~0ull, samples3_gauss_key = ~0ull;  // seeds of the sample tables
    uint32_t samples3_n = 0, samples3_gauss_n = 0;              // rays they hold
    cpd_status ensure_samples3(uint32_t kind, uint64_t seed_fx, uint32_t n, const void **table);
    cpd_status stage_designs3(const cpd_design3 *designs, const std::vector<uint32_t> &valid,
                              const cpd_design3 **designs_dev, uint32_t *nbin_max);
    cpd_status entrance_table3(const cpd_design3 *designs_dev, uint32_t n_designs, const void *fiber_tab,
                               const void *gauss_tab, uint32_t n_rays, const void **table);
    cpd_status run_rows3(bool chunked, const cpd_design3 *designs_dev, const std::vector<uint32_t> &row_design,
                         const std::vector<double> &row_lambda, const std::vector<uint64_t> &row_out,
                         const std::vector<uint32_t> &row_cap, uint64_t out_n, uint32_t max_rays, uint32_t nbin_max,
                         double max_err_sq, const void *fiber_tab, const void *gauss_tab, const void *entry,
                         uint32_t entry_rays, const double *&rows_host, const uint64_t *&rays_host);
    cpd_status run_fitness3(const cpd_design3 *designs, const cpd_sampling *sampling,
                            const std::vector<uint32_t> &valid, cpd_fitness *out);
    cpd_status rows3(const cpd_design3 &g, const double *wavelengths, size_t n_w, uint32_t max_m, uint32_t mode,
                     double *out);
    cpd_status run_gpu_fitness3(const cpd_design3 *designs, const std::vector<uint32_t> &valid, const double *seeds,
                                size_t n_seeds, uint32_t max_n, uint32_t nwarp, uint32_t max_eval, cpd_fitness *out,
                                cpd_status *per_design);
    cpd_status propagate3(const cpd_design3 &g, const double *wcdf, const double *rcdf, size_t n, uint32_t *bin,
                          double *t);
    void apply(const cpd_schedule &s);
    cpd_schedule schedule() const;
    cpd_status autotune(const cpd_design *designs, const cpd_sampling *sampling, size_t n,
                        const std::vector<uint32_t> &valid, cpd_schedule *best, double *best_ms);
};

cpd_status default_context(Context **out);
// The context a C entry point runs on: ctx, or the default context when NULL.
cpd_status resolve_context(cpd_ctx *ctx, Context **out);
cpd_schedule default_schedule();
cpd_status upload_beam_table(hipStream_t s, uint64_t seed_fx, uint32_t n, float *d_z32, double *d_z64);
cpd_status check_schedule(const cpd_schedule &s);
// Relative cost of one work item of a design (traced ray-stages per item).
double item_cost(const cpd_design &g, const cpd_sampling &s, uint32_t tile_cap);
// Indices of `valid` in the order a schedule runs them.
void order_designs(const cpd_design *designs, const cpd_sampling *sampling, std::vector<uint32_t> &valid,
                   uint32_t order, uint32_t tile_cap);
cpd_status check_design_public(const cpd_design &g, const cpd_sampling &s);

}  // namespace cpd
