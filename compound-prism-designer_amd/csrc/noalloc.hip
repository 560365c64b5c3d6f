// noalloc.hip -- cpd_fitness_batch_noalloc: the fixed-count batch path with
// every device buffer it needs (beam tables, staged records and work-item
// prefixes, per-bin accumulators, last-ray bins, results, the work counter)
// inside one file-scope __device__ array, so a call allocates no memory.
//
// The array is the GPU's own copy for this process (each device holds one
// from module load until the process ends).  Host code never passes the
// array's name to a kernel or a copy: it takes the device address once per
// device with hipGetSymbolAddress and uses that.  Host staging is static as
// well, so the calls of this process are serialised by one mutex; calls that
// overlap on one GPU from several processes each use their own copy.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "cpd.h"
#include "cpd_internal.h"
#include "kernels.h"
#include "runtime.h"

namespace cpd {
namespace {

constexpr uint32_t kMaxDesigns = CPD_NOALLOC_MAX_DESIGNS;
constexpr uint32_t kMaxRays = CPD_NOALLOC_MAX_RAYS;
constexpr uint64_t kMaxRows = CPD_NOALLOC_MAX_ROWS;
constexpr uint64_t kMaxCells = CPD_NOALLOC_MAX_CELLS;

constexpr size_t up(size_t v) { return (v + 255) / 256 * 256; }
// staged records, in the order the batch path stages them
constexpr size_t kStDesigns = 0;
constexpr size_t kStTraces = up(kStDesigns + kMaxDesigns * sizeof(cpd_design));
constexpr size_t kStItems = up(kStTraces + kMaxDesigns * sizeof(TraceDesign));
constexpr size_t kStAcc = up(kStItems + (kMaxDesigns + 1) * sizeof(uint64_t));
constexpr size_t kStLb = up(kStAcc + kMaxDesigns * sizeof(uint64_t));
constexpr size_t kStageBytes = up(kStLb + kMaxDesigns * sizeof(uint64_t));
// the scratch array: beam tables | stage | accumulators | last-ray bins | results | counter
// (kMaxRays is a multiple of kBandRays: whole banded blocks)
constexpr size_t kOffZ32 = 0;
constexpr size_t kOffZ64 = up(kOffZ32 + kMaxRays * sizeof(float));
constexpr size_t kOffZk = up(kOffZ64 + kMaxRays * sizeof(double));
constexpr size_t kOffZs64 = up(kOffZk + kMaxRays * sizeof(uint2));
constexpr size_t kOffInv = up(kOffZs64 + kMaxRays * sizeof(double));
constexpr size_t kOffStage = up(kOffInv + kMaxRays * sizeof(uint32_t));
constexpr size_t kOffAcc = up(kOffStage + kStageBytes);
constexpr size_t kOffLastBin = up(kOffAcc + kMaxCells * sizeof(unsigned long long));
constexpr size_t kOffResults = up(kOffLastBin + kMaxRows * sizeof(uint32_t));
constexpr size_t kOffCounter = up(kOffResults + kMaxDesigns * sizeof(cpd_fitness));
constexpr size_t kScratchBytes = up(kOffCounter + kWorkCounterWords * sizeof(unsigned int));
static_assert(kScratchBytes == CPD_NOALLOC_SCRATCH_BYTES, "cpd.h states the scratch size");
static_assert(kMaxRays % kBandRays == 0, "whole banded blocks");

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;                               // the static host staging below (taken before a context's)
unsigned char *g_ptr[kMaxDevices] = {nullptr};  // device address of scratch, per device
alignas(256) unsigned char g_stage[kStageBytes];
cpd_fitness g_results[kMaxDesigns];
uint32_t g_valid[kMaxDesigns];
bool g_ok[kMaxDesigns];
// beam tables as the batch path builds them for 65536 rays (BeamTables::ensure)
double g_z[kMaxRays], g_zs[kMaxRays];
float g_zf[kMaxRays];
uint2 g_zk[kMaxRays];
uint32_t g_order[kMaxRays], g_pos[kMaxRays];

#define NA_TRY(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(CPD_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));     \
    } while (0)

// Once per device, with that device current: the scratch address, and the
// beam tables z_k = norminv(u_k) for k = 1..kMaxRays (in ray order and banded,
// the same values the batch path's tables hold) copied into it.
cpd_status setup(int device, hipStream_t s, unsigned char **out) {
    if (device < 0 || device >= kMaxDevices) return fail(CPD_ERR_DEVICE, "device %d beyond %d", device, kMaxDevices);
    if (g_ptr[device]) {
        *out = g_ptr[device];
        return CPD_OK;
    }
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(CPD_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    unsigned char *base = (unsigned char *)p;
    for (uint32_t i = 0; i < kMaxRays; ++i) {
        g_z[i] = norminv64(qmc_seeded(kHalfFx, (uint64_t)i + 1));
        g_zf[i] = (float)g_z[i];
        g_order[i] = i;
    }
    for (uint32_t b = 0; b + kBandRays <= kMaxRays; b += kBandRays)
        std::sort(g_order + b, g_order + b + kBandRays,
                  [](uint32_t x, uint32_t y) { return g_z[x] < g_z[y] || (g_z[x] == g_z[y] && x < y); });
    for (uint32_t i = 0; i < kMaxRays; ++i) {
        g_zs[i] = g_z[g_order[i]];
        float f = g_zf[g_order[i]];
        uint32_t bits;
        memcpy(&bits, &f, sizeof bits);
        g_zk[i] = uint2{bits, g_order[i] + 1};
        g_pos[g_order[i]] = i + 1;
    }
    NA_TRY(hipMemcpyAsync(base + kOffZ64, g_z, sizeof g_z, hipMemcpyHostToDevice, s));
    NA_TRY(hipMemcpyAsync(base + kOffZ32, g_zf, sizeof g_zf, hipMemcpyHostToDevice, s));
    NA_TRY(hipMemcpyAsync(base + kOffZs64, g_zs, sizeof g_zs, hipMemcpyHostToDevice, s));
    NA_TRY(hipMemcpyAsync(base + kOffZk, g_zk, sizeof g_zk, hipMemcpyHostToDevice, s));
    NA_TRY(hipMemcpyAsync(base + kOffInv, g_pos, sizeof g_pos, hipMemcpyHostToDevice, s));
    NA_TRY(hipStreamSynchronize(s));
    g_ptr[device] = base;
    *out = base;
    return CPD_OK;
}

}  // namespace
}  // namespace cpd

extern "C" cpd_status cpd_fitness_batch_noalloc(cpd_ctx *ctx, const cpd_design *designs, const cpd_sampling *sampling,
                                                size_t n, cpd_fitness *out, cpd_status *per_design) {
    using namespace cpd;
    if (n == 0) return CPD_OK;
    if (!designs || !sampling || !out) return fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > kMaxDesigns)
        return fail(CPD_ERR_INVALID_ARGUMENT, "cpd_fitness_batch_noalloc takes at most %u designs, got %zu", kMaxDesigns, n);
    std::lock_guard<std::mutex> lk_stage(g_mu);
    uint64_t rows = 0, cells = 0;
    for (size_t i = 0; i < n; ++i) {
        cpd_status s = check_design_public(designs[i], sampling[i]);
        if (s == CPD_OK && sampling[i].mode != CPD_MODE_FIXED)
            s = fail(CPD_ERR_INVALID_ARGUMENT, "cpd_fitness_batch_noalloc evaluates fixed-count sampling only");
        if (s == CPD_OK && sampling[i].n_rays > kMaxRays)
            s = fail(CPD_ERR_INVALID_ARGUMENT, "cpd_fitness_batch_noalloc takes at most %u rays per wavelength", kMaxRays);
        if (per_design) per_design[i] = s;
        g_ok[i] = s == CPD_OK;
        if (g_ok[i]) {
            rows += sampling[i].n_wavelengths;
            cells += (uint64_t)sampling[i].n_wavelengths * designs[i].bin_count;
        } else {
            out[i].size = out[i].info = out[i].deviation = NAN;
        }
    }
    if (rows > kMaxRows || cells > kMaxCells)
        return fail(CPD_ERR_INVALID_ARGUMENT,
                    "cpd_fitness_batch_noalloc: %llu wavelength rows and %llu bin cells exceed the scratch maximum "
                    "(%llu rows, %llu cells)",
                    (unsigned long long)rows, (unsigned long long)cells, (unsigned long long)kMaxRows,
                    (unsigned long long)kMaxCells);
    if (rows == 0) return CPD_OK;
    Context *c;
    cpd_status st = resolve_context(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    NA_TRY(hipSetDevice(c->device));
    unsigned char *base;
    if ((st = setup(c->device, c->stream, &base)) != CPD_OK) return st;

    // stage the valid designs: records, trace records, work-item prefixes
    cpd_design *hd = (cpd_design *)(g_stage + kStDesigns);
    TraceDesign *ht = (TraceDesign *)(g_stage + kStTraces);
    uint64_t *hitem = (uint64_t *)(g_stage + kStItems);
    uint64_t *hacc = (uint64_t *)(g_stage + kStAcc);
    uint64_t *hlb = (uint64_t *)(g_stage + kStLb);
    uint32_t nv = 0, hist_words = 0, row_words = 0;
    uint64_t items = 0, acc_n = 0, lb_n = 0, nrays = 0;
    const uint32_t tile_cap = c->acc_kind == CPD_ACC_F32C ? kMaxTileRays : c->tile_rays_hint;
    for (size_t i = 0; i < n; ++i) {
        if (!g_ok[i]) continue;
        const cpd_design &g = designs[i];
        const cpd_sampling &s = sampling[i];
        g_valid[nv] = (uint32_t)i;
        hd[nv] = g;
        make_trace_design(g, s.n_wavelengths, s.n_rays, ht[nv], tile_cap);
        hist_words = std::max(hist_words, ht[nv].hist_cols * g.bin_count);
        row_words = std::max(row_words, g.bin_count);
        hitem[nv] = items;
        hacc[nv] = acc_n;
        hlb[nv] = lb_n;
        items += (uint64_t)s.n_wavelengths * ht[nv].n_tiles;
        nrays += (uint64_t)s.n_wavelengths * s.n_rays;
        acc_n += (uint64_t)s.n_wavelengths * g.bin_count;
        lb_n += s.n_wavelengths;
        ++nv;
    }
    hitem[nv] = items;
    clear_error();
    hipStream_t s = c->stream;
    unsigned long long *acc = (unsigned long long *)(base + kOffAcc);
    unsigned int *counter = (unsigned int *)(base + kOffCounter);
    cpd_fitness *results = (cpd_fitness *)(base + kOffResults);
    NA_TRY(hipMemcpyAsync(base + kOffStage, g_stage, kStageBytes, hipMemcpyHostToDevice, s));
    // the accumulators start at zero (a failed earlier call may have left
    // partial sums) and so does the work counter; fill kernels, no memset
    NA_TRY(launch_zero_u32(s, (unsigned int *)acc, (uint32_t)(2 * acc_n)));
    NA_TRY(launch_zero_u32(s, counter, kWorkCounterWords));

    BatchArgs a{};
    a.designs = (const cpd_design *)(base + kOffStage + kStDesigns);
    a.traces = (const TraceDesign *)(base + kOffStage + kStTraces);
    a.item_begin = (const uint64_t *)(base + kOffStage + kStItems);
    a.acc_begin = (const uint64_t *)(base + kOffStage + kStAcc);
    a.lb_begin = (const uint64_t *)(base + kOffStage + kStLb);
    a.acc = acc;
    a.last_bin = (uint32_t *)(base + kOffLastBin);
    a.z32 = (const float *)(base + kOffZ32);
    a.z64 = (const double *)(base + kOffZ64);
    a.zk = (const uint2 *)(base + kOffZk);
    a.zs64 = (const double *)(base + kOffZs64);
    a.inv = (const uint32_t *)(base + kOffInv);
    a.n_designs = nv;
    a.hist_words = hist_words;
    a.row_words = row_words;
    a.n_items = items;
    a.work_counter = counter;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(items, (uint64_t)c->cu_count * c->waves_per_cu);
    a.items_per_fetch = items_per_fetch(nrays, items, grid);
    NA_TRY(launch_trace(s, a, grid, false, c->acc_kind));

    FinalizeArgs f{};
    f.designs = a.designs;
    f.traces = a.traces;
    f.acc_begin = a.acc_begin;
    f.lb_begin = a.lb_begin;
    f.acc = acc;
    f.last_bin = a.last_bin;
    f.out = results;
    f.n_designs = nv;
    NA_TRY(launch_finalize(s, f, c->acc_kind));
    NA_TRY(hipMemcpyAsync(g_results, results, nv * sizeof(cpd_fitness), hipMemcpyDeviceToHost, s));
    NA_TRY(hipStreamSynchronize(s));
    for (uint32_t q = 0; q < nv; ++q) out[g_valid[q]] = g_results[q];
    return CPD_OK;
}
