// runtime.hip -- device contexts and the GPU entry points of cpd.h.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "cpd.h"
#include "cpd_internal.h"
#include "kernels.h"
#include "runtime.h"

namespace cpd {

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? CPD_ERR_OUT_OF_MEMORY : CPD_ERR_DEVICE, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                                 \
    } while (0)

cpd_status DeviceBuffer::reserve(size_t want, bool zero, hipStream_t s) {
    if (want <= bytes) return CPD_OK;
    if (ptr) {
        HIP_TRY(hipStreamSynchronize(s));
        HIP_TRY(hipFree(ptr));
        ptr = nullptr;
        bytes = 0;
    }
    size_t sz = std::max<size_t>(want, 256);
    HIP_TRY(hipMalloc(&ptr, sz));
    bytes = sz;
    if (zero) HIP_TRY(hipMemsetAsync(ptr, 0, sz, s));
    return CPD_OK;
}

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

cpd_status HostBuffer::reserve(size_t want) {
    if (want <= bytes) return CPD_OK;
    if (ptr) HIP_TRY(hipHostFree(ptr));
    ptr = nullptr;
    bytes = 0;
    size_t sz = std::max<size_t>(want, 4096);
    HIP_TRY(hipHostMalloc(&ptr, sz, hipHostMallocDefault));
    bytes = sz;
    return CPD_OK;
}

void HostBuffer::release() {
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

// Fill the beam-quantile tables for rays k = 1..n (grow-only).
// z_k = norminv(u_k) for k = 1..n of the R1 sequence from seed_fx, computed
// on the host in FP64 (the same code and libm log as every host FP64 path, so
// the device's beam heights are bit-identical to the host's) and copied to
// the device; the copy has completed when this returns.
cpd_status upload_beam_table(hipStream_t s, uint64_t seed_fx, uint32_t n, float *d_z32, double *d_z64) {
    if (n == 0) return CPD_OK;
    std::vector<double> z(n);
    std::vector<float> zf(n);
    for (uint32_t i = 0; i < n; ++i) {
        z[i] = norminv64(qmc_seeded(seed_fx, (uint64_t)i + 1));
        zf[i] = (float)z[i];
    }
    HIP_TRY(hipMemcpyAsync(d_z64, z.data(), (size_t)n * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_z32, zf.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    return CPD_OK;
}

cpd_status BeamTables::ensure(uint32_t want, hipStream_t s) {
    if (want <= n) return CPD_OK;
    // whole banded blocks: a design's last block is traced whole, its rays
    // past n_rays idle
    const uint32_t m = (std::max<uint32_t>(want, 65536) + kBandRays - 1) / kBandRays * kBandRays;
    cpd_status st;
    // the FP32 tables carry kBeamPad entries past ray m (zk's with a ray
    // index no design reaches)
    if ((st = z32.reserve((size_t)(m + kBeamPad) * sizeof(float), false, s)) != CPD_OK) return st;
    if ((st = z64.reserve((size_t)m * sizeof(double), false, s)) != CPD_OK) return st;
    if ((st = zk.reserve((size_t)(m + kBeamPad) * sizeof(uint2), false, s)) != CPD_OK) return st;
    if ((st = zs64.reserve((size_t)m * sizeof(double), false, s)) != CPD_OK) return st;
    if ((st = inv.reserve((size_t)m * sizeof(uint32_t), false, s)) != CPD_OK) return st;
    std::vector<double> z(m), zs(m);
    std::vector<float> zf(m + kBeamPad, 0.f);
    std::vector<uint2> zkh(m + kBeamPad, uint2{0u, 0xffffffffu});
    std::vector<uint32_t> order(m), pos(m);
    for (uint32_t i = 0; i < m; ++i) {
        z[i] = norminv64(qmc_seeded(kHalfFx, (uint64_t)i + 1));
        zf[i] = (float)z[i];
        order[i] = i;
    }
    for (uint32_t b = 0; b + kBandRays <= m; b += kBandRays)
        std::sort(order.begin() + b, order.begin() + b + kBandRays,
                  [&](uint32_t x, uint32_t y) { return z[x] < z[y] || (z[x] == z[y] && x < y); });
    for (uint32_t i = 0; i < m; ++i) {
        zs[i] = z[order[i]];
        float f = zf[order[i]];
        uint32_t bits;
        memcpy(&bits, &f, sizeof bits);
        zkh[i] = uint2{bits, order[i] + 1};
        pos[order[i]] = i + 1;
    }
    HIP_TRY(hipMemcpyAsync(z64.ptr, z.data(), (size_t)m * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(z32.ptr, zf.data(), (size_t)(m + kBeamPad) * sizeof(float), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(zs64.ptr, zs.data(), (size_t)m * sizeof(double), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(zk.ptr, zkh.data(), (size_t)(m + kBeamPad) * sizeof(uint2), hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(inv.ptr, pos.data(), (size_t)m * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIP_TRY(hipStreamSynchronize(s));
    n = m;
    return CPD_OK;
}

void BeamTables::release() {
    z32.release();
    z64.release();
    zk.release();
    zs64.release();
    inv.release();
    n = 0;
}

cpd_status Context::ensure_beam_table(uint32_t n) { return beam.ensure(n, stream); }

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

cpd_status check_design(const cpd_design &g, const cpd_sampling &s) {
    if (g.n_glass < 1 || g.n_glass > CPD_MAX_PRISMS) return fail(CPD_ERR_INVALID_ARGUMENT, "n_glass out of range");
    if (g.bin_count < 1 || g.bin_count > CPD_MAX_BINS) return fail(CPD_ERR_INVALID_ARGUMENT, "bin_count out of range");
    if (s.n_wavelengths < 1 || s.n_rays < 1) return fail(CPD_ERR_INVALID_ARGUMENT, "empty sampling");
    if (s.n_rays > (1u << 30)) return fail(CPD_ERR_INVALID_ARGUMENT, "n_rays above 2^30");
    if (s.mode != CPD_MODE_FIXED && s.mode != CPD_MODE_ADAPTIVE) return fail(CPD_ERR_INVALID_ARGUMENT, "unknown sampling mode");
    if (!(g.det_slope > 0.0)) return fail(CPD_ERR_INVALID_ARGUMENT, "detector slope must be > 0");
    return CPD_OK;
}

}  // namespace

cpd_schedule default_schedule() { return cpd_schedule{0, 16, CPD_ORDER_GIVEN, 0}; }

cpd_status check_schedule(const cpd_schedule &s) {
    if (s.tile_rays > kMaxTileRays || s.tile_rays % 64)
        return fail(CPD_ERR_INVALID_ARGUMENT, "tile rays must be a multiple of 64 in [0, 2^20]");
    if (s.waves_per_cu < 1 || s.waves_per_cu > 64) return fail(CPD_ERR_INVALID_ARGUMENT, "waves per CU must be in 1..64");
    if (s.order != CPD_ORDER_GIVEN && s.order != CPD_ORDER_COST) return fail(CPD_ERR_INVALID_ARGUMENT, "unknown order");
    return CPD_OK;
}

void Context::apply(const cpd_schedule &s) {
    tile_rays_hint = s.tile_rays;
    waves_per_cu = s.waves_per_cu;
    order = s.order;
}

cpd_schedule Context::schedule() const { return cpd_schedule{tile_rays_hint, waves_per_cu, order, 0}; }

// Coordinate search over work order, grid size and tile cap.  Each candidate
// runs the whole batch twice; its time is the faster run's trace + finalize
// time (HIP events).  Any schedule gives the same results, so they are
// discarded.
cpd_status Context::autotune(const cpd_design *designs, const cpd_sampling *sampling, size_t n,
                             const std::vector<uint32_t> &valid, cpd_schedule *best, double *best_ms) {
    const cpd_schedule saved = schedule();
    const bool saved_profile = profile, saved_count = count_ops;
    const cpd_run_stats saved_last = last;
    profile = true;
    count_ops = false;
    std::vector<cpd_fitness> scratch(n);
    cpd_status st = CPD_OK;
    auto time_of = [&](const cpd_schedule &sc, double &ms) -> cpd_status {
        apply(sc);
        ms = INFINITY;
        for (int rep = 0; rep < 2; ++rep) {
            cpd_status s2 = run(designs, sampling, valid, true, nullptr, scratch.data(), nullptr, nullptr);
            if (s2 != CPD_OK) return s2;
            ms = std::min(ms, last.trace_ms + last.finalize_ms);
        }
        return CPD_OK;
    };
    cpd_schedule cur{0, 12, CPD_ORDER_COST, 0};
    double cur_ms;
    st = time_of(cur, cur_ms);
    const uint32_t orders[] = {CPD_ORDER_GIVEN};
    const uint32_t waves[] = {4, 8, 16, 20, 24};
    const uint32_t tiles[] = {16384, 8192, 4096};
    for (uint32_t o : orders) {
        if (st != CPD_OK) break;
        cpd_schedule cand = cur;
        cand.order = o;
        double ms;
        if ((st = time_of(cand, ms)) == CPD_OK && ms < cur_ms) cur = cand, cur_ms = ms;
    }
    for (uint32_t w : waves) {
        if (st != CPD_OK) break;
        cpd_schedule cand = cur;
        cand.waves_per_cu = w;
        double ms;
        if ((st = time_of(cand, ms)) == CPD_OK && ms < cur_ms) cur = cand, cur_ms = ms;
    }
    for (uint32_t t : tiles) {
        if (st != CPD_OK) break;
        cpd_schedule cand = cur;
        cand.tile_rays = t;
        double ms;
        if ((st = time_of(cand, ms)) == CPD_OK && ms < cur_ms) cur = cand, cur_ms = ms;
    }
    apply(saved);
    profile = saved_profile;
    count_ops = saved_count;
    last = saved_last;
    if (st != CPD_OK) return st;
    *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return CPD_OK;
}

double item_cost(const cpd_design &g, const cpd_sampling &s, uint32_t tile_cap) {
    // per ray the trace costs the same for every prism count (the planar
    // faces are hoisted per wavelength); per item the wavelength set-up adds
    // a few hundred FP64 operations per face
    const double per_ray = kOpsBeam + kOpsChain + kOpsLensHit + kOpsLensRefract + kOpsDetector + kOpsHit;
    return per_ray * tile_for(g.bin_count, (s.n_rays + kBandRays - 1) / kBandRays * kBandRays, tile_cap) +
           300.0 * g.n_glass;
}

void order_designs(const cpd_design *designs, const cpd_sampling *sampling, std::vector<uint32_t> &valid,
                   uint32_t order, uint32_t tile_cap) {
    if (order != CPD_ORDER_COST) return;
    // longest-processing-time first: items of the costliest designs are
    // fetched first, so the grid ends on the cheapest ones
    std::vector<std::pair<double, uint32_t>> key;
    key.reserve(valid.size());
    for (uint32_t i : valid) key.emplace_back(-item_cost(designs[i], sampling[i], tile_cap), i);
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<double, uint32_t> &a, const std::pair<double, uint32_t> &b) {
                         return a.first < b.first;
                     });
    for (size_t q = 0; q < key.size(); ++q) valid[q] = key[q].second;
}

// Stage a batch on the device and run trace + finalize.  `valid` lists the
// indices of the designs to run; out/rows receive results for them.
cpd_status Context::run(const cpd_design *designs, const cpd_sampling *sampling, const std::vector<uint32_t> &valid_in,
                        bool persistent, const double *wavelengths, cpd_fitness *out, double *rows,
                        unsigned long long *ops_out) {
    if (valid_in.empty()) return CPD_OK;
    // per-wavelength rows are written in staging order: keep the given order
    const bool reorder = order != CPD_ORDER_GIVEN && rows == nullptr;
    std::vector<uint32_t> ordered;
    if (reorder) {
        ordered = valid_in;
        order_designs(designs, sampling, ordered, order, tile_rays_hint);
    }
    const std::vector<uint32_t> &valid = reorder ? ordered : valid_in;
    const uint32_t n = (uint32_t)valid.size();
    HIP_TRY(hipSetDevice(device));
    uint32_t max_rays = 0;
    for (uint32_t i : valid) max_rays = std::max(max_rays, sampling[i].n_rays);
    if (acc_kind == CPD_ACC_F32C && max_rays > kMaxTileRays)
        return fail(CPD_ERR_INVALID_ARGUMENT, "the compensated FP32 accumulation takes at most 2^20 rays per wavelength");
    // host staging: designs | traces | item_begin | acc_begin | lb_begin | row_begin
    const size_t off_designs = 0;
    const size_t off_traces = align_up(off_designs + n * sizeof(cpd_design), 256);
    const size_t off_items = align_up(off_traces + n * sizeof(TraceDesign), 256);
    const size_t off_acc = align_up(off_items + (n + 1) * sizeof(uint64_t), 256);
    const size_t off_lb = align_up(off_acc + n * sizeof(uint64_t), 256);
    const size_t off_rows = align_up(off_lb + n * sizeof(uint64_t), 256);
    const size_t off_wl = align_up(off_rows + n * sizeof(uint64_t), 256);
    const uint32_t n_wl = wavelengths ? sampling[valid[0]].n_wavelengths : 0;
    const size_t staged = align_up(off_wl + n_wl * sizeof(double), 256);
    cpd_status st = host_stage.reserve(staged);
    if (st != CPD_OK) return st;
    char *h = (char *)host_stage.ptr;
    cpd_design *hd = (cpd_design *)(h + off_designs);
    TraceDesign *ht = (TraceDesign *)(h + off_traces);
    uint64_t *hitem = (uint64_t *)(h + off_items);
    uint64_t *hacc = (uint64_t *)(h + off_acc);
    uint64_t *hlb = (uint64_t *)(h + off_lb);
    uint64_t *hrow = (uint64_t *)(h + off_rows);
    uint64_t items = 0, acc_n = 0, lb_n = 0, rays = 0;
    uint32_t hist_words = 0, row_words = 0, tile = 0;
    for (uint32_t q = 0; q < n; ++q) {
        const cpd_design &g = designs[valid[q]];
        const cpd_sampling &s = sampling[valid[q]];
        hd[q] = g;
        // the compensated FP32 policy stores a wavelength's row from one item
        make_trace_design(g, s.n_wavelengths, s.n_rays, ht[q], acc_kind == CPD_ACC_F32C ? kMaxTileRays : tile_rays_hint);
        hist_words = std::max(hist_words, ht[q].hist_cols * g.bin_count);
        row_words = std::max(row_words, g.bin_count);
        tile = std::max(tile, ht[q].tile_rays);
        hitem[q] = items;
        hacc[q] = acc_n;
        hlb[q] = lb_n;
        hrow[q] = acc_n;
        items += (uint64_t)s.n_wavelengths * ht[q].n_tiles;
        rays += (uint64_t)s.n_wavelengths * s.n_rays;
        acc_n += (uint64_t)s.n_wavelengths * g.bin_count;
        lb_n += s.n_wavelengths;
    }
    hitem[n] = items;
    if (n_wl) memcpy(h + off_wl, wavelengths, n_wl * sizeof(double));
    if (items >= (1ull << 32)) return fail(CPD_ERR_INVALID_ARGUMENT, "batch too large (>= 2^32 work items)");

    if ((st = ensure_beam_table(max_rays)) != CPD_OK) return st;
    if ((st = stage.reserve(staged, false, stream)) != CPD_OK) return st;
    if ((st = acc.reserve(acc_n * sizeof(unsigned long long), true, stream)) != CPD_OK) return st;
    if ((st = last_bin.reserve(lb_n * sizeof(uint32_t), false, stream)) != CPD_OK) return st;
    if ((st = results.reserve(n * sizeof(cpd_fitness), false, stream)) != CPD_OK) return st;
    if ((st = lrec.reserve(lb_n * sizeof(LambdaRecord), false, stream)) != CPD_OK) return st;
    // the operation count, then the work counters from byte 256
    if ((st = counters.reserve(256 + kWorkCounterWords * sizeof(unsigned int), true, stream)) != CPD_OK) return st;
    if (rows && (st = row_buf.reserve(acc_n * sizeof(double), false, stream)) != CPD_OK) return st;

    char *dstage = (char *)stage.ptr;
    HIP_TRY(hipMemcpyAsync(dstage, h, staged, hipMemcpyHostToDevice, stream));
    unsigned long long *dops = (unsigned long long *)counters.ptr;
    unsigned int *dwork = (unsigned int *)((char *)counters.ptr + 256);
    const bool want_ops = ops_out != nullptr || count_ops;
    if (want_ops) HIP_TRY(hipMemsetAsync(dops, 0, sizeof(unsigned long long), stream));

    BatchArgs a{};
    a.designs = (const cpd_design *)(dstage + off_designs);
    a.traces = (const TraceDesign *)(dstage + off_traces);
    a.item_begin = (const uint64_t *)(dstage + off_items);
    a.acc_begin = (const uint64_t *)(dstage + off_acc);
    a.lb_begin = (const uint64_t *)(dstage + off_lb);
    a.acc = acc.as<unsigned long long>();
    a.last_bin = last_bin.as<uint32_t>();
    a.z32 = beam.z32.as<float>();
    a.z64 = beam.z64.as<double>();
    a.zk = beam.zk.as<uint2>();
    a.zs64 = beam.zs64.as<double>();
    a.inv = beam.inv.as<uint32_t>();
    a.ops = want_ops ? dops : nullptr;
    a.wavelengths = n_wl ? (const double *)(dstage + off_wl) : nullptr;
    a.lrec = lrec.as<LambdaRecord>();
    a.n_designs = n;
    a.hist_words = hist_words;
    a.row_words = row_words;
    a.n_items = items;
    uint32_t grid;
    if (persistent) {
        HIP_TRY(launch_zero_u32(stream, dwork, kWorkCounterWords));
        a.work_counter = dwork;
        const uint64_t resident = (uint64_t)cu_count * waves_per_cu;
        grid = (uint32_t)std::min<uint64_t>(items, resident);
        a.items_per_fetch = items_per_fetch(rays, items, grid);
    } else {
        a.work_counter = nullptr;
        grid = (uint32_t)items;
    }
    HIP_TRY(launch_trace(stream, a, grid, want_ops, acc_kind, profile ? ev[0] : nullptr, profile ? ev[1] : nullptr));

    FinalizeArgs f{};
    f.designs = a.designs;
    f.traces = a.traces;
    f.acc_begin = a.acc_begin;
    f.lb_begin = a.lb_begin;
    f.acc = a.acc;
    f.last_bin = a.last_bin;
    f.out = out ? results.as<cpd_fitness>() : nullptr;
    f.rows = rows ? row_buf.as<double>() : nullptr;
    f.row_begin = (const uint64_t *)(dstage + off_rows);
    f.n_designs = n;
    HIP_TRY(launch_finalize(stream, f, acc_kind));
    if (profile) HIP_TRY(hipEventRecord(ev[2], stream));
    if (out) {
        st = host_out.reserve(n * sizeof(cpd_fitness));
        if (st != CPD_OK) return st;
        HIP_TRY(hipMemcpyAsync(host_out.ptr, results.ptr, n * sizeof(cpd_fitness), hipMemcpyDeviceToHost, stream));
    }
    if (rows) HIP_TRY(hipMemcpyAsync(rows, row_buf.ptr, acc_n * sizeof(double), hipMemcpyDeviceToHost, stream));
    unsigned long long ops_count = 0;
    if (want_ops)
        HIP_TRY(hipMemcpyAsync(&ops_count, dops, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (ops_out) *ops_out = ops_count;
    last = cpd_run_stats{};
    last.items = items;
    last.grid = grid;
    last.tile_rays = tile;
    last.ops = ops_count;
    for (uint32_t q = 0; q < n; ++q) last.rays += (uint64_t)sampling[valid[q]].n_wavelengths * sampling[valid[q]].n_rays;
    if (profile) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
        last.trace_ms = ms;
        HIP_TRY(hipEventElapsedTime(&ms, ev[1], ev[2]));
        last.finalize_ms = ms;
    }
    if (out) {
        const cpd_fitness *r = (const cpd_fitness *)host_out.ptr;
        for (uint32_t q = 0; q < n; ++q) out[valid[q]] = r[q];
    }
    return CPD_OK;
}

Context::~Context() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    beam.release();
    stage.release();
    acc.release();
    last_bin.release();
    lrec.release();
    results.release();
    counters.release();
    row_buf.release();
    prop_buf.release();
    build_buf.release();
    fold_buf.release();
    fit_buf.release();
    row_stage.release();
    design_stage.release();
    seed_z32.release();
    seed_z64.release();
    design3_stage.release();
    samples3.release();
    samples3_gauss.release();
    entry3.release();
    recip.release();
    host_stage.release();
    row_args_pin.release();
    rows_pin.release();
    host_out.release();
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

cpd_status Context::create(int device, Context **out) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return fail(CPD_ERR_NO_DEVICE, "no HIP device");
    if (device < 0 || device >= count) return fail(CPD_ERR_NO_DEVICE, "device %d not present (%d devices)", device, count);
    Context *c = new Context();
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return fail(CPD_ERR_DEVICE, "context creation: %s", hipGetErrorString(e));
    }
    for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreate(&c->ev[i]);
    if (e != hipSuccess) {
        delete c;
        return fail(CPD_ERR_DEVICE, "event creation: %s", hipGetErrorString(e));
    }
    c->cu_count = prop.multiProcessorCount;
    c->clock_khz = prop.clockRate;
    snprintf(c->name, sizeof c->name, "%s", prop.name);
    *out = c;
    return CPD_OK;
}

namespace {
std::mutex g_default_mu;
Context *g_default = nullptr;
}  // namespace

cpd_status default_context(Context **out) {
    std::lock_guard<std::mutex> lk(g_default_mu);
    if (!g_default) {
        cpd_status st = Context::create(0, &g_default);
        if (st != CPD_OK) return st;
    }
    *out = g_default;
    return CPD_OK;
}

}  // namespace cpd

using cpd::Context;

cpd_status cpd::resolve_context(cpd_ctx *ctx, Context **out) {
    if (ctx) {
        *out = reinterpret_cast<Context *>(ctx);
        return CPD_OK;
    }
    return cpd::default_context(out);
}

static cpd_status resolve(cpd_ctx *ctx, Context **out) { return cpd::resolve_context(ctx, out); }

extern "C" {

cpd_status cpd_ctx_create(int device, cpd_ctx **ctx_out) {
    if (!ctx_out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    Context *c = nullptr;
    cpd_status st = Context::create(device, &c);
    if (st != CPD_OK) return st;
    *ctx_out = reinterpret_cast<cpd_ctx *>(c);
    return CPD_OK;
}

void cpd_ctx_destroy(cpd_ctx *ctx) { delete reinterpret_cast<Context *>(ctx); }

cpd_status cpd_ctx_info(cpd_ctx *ctx, char *name, size_t name_len, int *cu_count, int *clock_khz) {
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    if (name && name_len) snprintf(name, name_len, "%s", c->name);
    if (cu_count) *cu_count = c->cu_count;
    if (clock_khz) *clock_khz = c->clock_khz;
    return CPD_OK;
}

cpd_status cpd_ctx_set_option(cpd_ctx *ctx, int option, int64_t value) {
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    switch (option) {
        case CPD_OPT_PROFILE: c->profile = value != 0; return CPD_OK;
        case CPD_OPT_COUNT_OPS: c->count_ops = value != 0; return CPD_OK;
        case CPD_OPT_TILE_RAYS:
            if (value < 0 || value > (int64_t)cpd::kMaxTileRays || value % 64)
                return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "tile rays must be a multiple of 64 in [0, 2^20]");
            c->tile_rays_hint = (uint32_t)value;
            return CPD_OK;
        case CPD_OPT_WAVES_PER_CU:
            if (value < 1 || value > 64) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "waves per CU must be in 1..64");
            c->waves_per_cu = (uint32_t)value;
            return CPD_OK;
        case CPD_OPT_ACCUMULATE:
            if (value < CPD_ACC_EXACT || value > CPD_ACC_F64)
                return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "unknown accumulation %lld", (long long)value);
            c->acc_kind = (int)value;
            return CPD_OK;
        case CPD_OPT_ROW_TRACE:
            if (value != 0 && value != 32 && value != 64)
                return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "row trace precision must be 0, 32 or 64");
            c->row_trace = (int)value;
            return CPD_OK;
        case CPD_OPT_ORDER:
            if (value != CPD_ORDER_GIVEN && value != CPD_ORDER_COST)
                return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "unknown order %lld", (long long)value);
            c->order = (uint32_t)value;
            return CPD_OK;
    }
    return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "unknown option %d", option);
}

cpd_status cpd_ctx_set_schedule(cpd_ctx *ctx, const cpd_schedule *schedule) {
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    const cpd_schedule sc = schedule ? *schedule : cpd::default_schedule();
    if ((st = cpd::check_schedule(sc)) != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    c->apply(sc);
    return CPD_OK;
}

cpd_status cpd_autotune(cpd_ctx *ctx, const cpd_design *designs, const cpd_sampling *sampling, size_t n,
                        cpd_schedule *best, double *best_ms) {
    if (!designs || !sampling || !best) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0 || n > (1u << 24)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "batch size must be in 1..2^24");
    std::vector<uint32_t> valid;
    for (size_t i = 0; i < n; ++i)
        if (cpd::check_design_public(designs[i], sampling[i]) == CPD_OK) valid.push_back((uint32_t)i);
    if (valid.empty()) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "no valid design to tune on");
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    return c->autotune(designs, sampling, n, valid, best, best_ms);
}

cpd_status cpd_ctx_last_stats(cpd_ctx *ctx, cpd_run_stats *out) {
    if (!out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    *out = c->last;
    return CPD_OK;
}

cpd_status cpd_fitness_gpu(cpd_ctx *ctx, const cpd_design *design, const cpd_sampling *sampling, cpd_fitness *out) {
    if (!design || !sampling || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    cpd_status st = cpd::check_design_public(*design, *sampling);
    if (st != CPD_OK) return st;
    Context *c;
    if ((st = resolve(ctx, &c)) != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    std::vector<uint32_t> valid{0};
    if (sampling->mode == CPD_MODE_ADAPTIVE) return c->run_adaptive(design, sampling, valid, out);
    return c->run(design, sampling, valid, false, nullptr, out, nullptr, nullptr);
}

cpd_status cpd_fitness_batch(cpd_ctx *ctx, const cpd_design *designs, const cpd_sampling *sampling, size_t n,
                             cpd_fitness *out, cpd_status *per_design) {
    if (n == 0) return CPD_OK;
    if (!designs || !sampling || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > (1u << 24)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "batch above 2^24 designs");
    std::vector<uint32_t> valid;
    valid.reserve(n);
    for (size_t i = 0; i < n; ++i) {
        cpd_status s = cpd::check_design_public(designs[i], sampling[i]);
        if (per_design) per_design[i] = s;
        if (s == CPD_OK) {
            valid.push_back((uint32_t)i);
        } else {
            out[i].size = out[i].info = out[i].deviation = NAN;
        }
    }
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    std::vector<uint32_t> fixed, adaptive;
    for (uint32_t i : valid) (sampling[i].mode == CPD_MODE_ADAPTIVE ? adaptive : fixed).push_back(i);
    st = c->run(designs, sampling, fixed, true, nullptr, out, nullptr, nullptr);
    if (st != CPD_OK) {
        if (per_design)
            for (uint32_t i : valid) per_design[i] = st;
        return st;
    }
    st = c->run_adaptive(designs, sampling, adaptive, out);
    if (st != CPD_OK && per_design)
        for (uint32_t i : adaptive) per_design[i] = st;
    return st;
}

cpd_status cpd_gpu_fitness_batch(cpd_ctx *ctx, const cpd_design *designs, size_t n, const double *seeds,
                                 size_t n_seeds, uint32_t max_n, uint32_t nwarp, uint32_t max_eval,
                                 cpd_fitness *out, cpd_status *per_design) {
    if (n == 0) return CPD_OK;
    if (!designs || !out || (n_seeds && !seeds)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (nwarp == 0) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "nwarp must be >= 1");
    if (n > (1u << 20)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "batch above 2^20 designs");
    if ((uint64_t)max_n * n > (1ull << 28)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "max_n x designs above 2^28 rows");
    const cpd_sampling probe{1, 1, CPD_MODE_ADAPTIVE, 0};
    std::vector<uint32_t> valid;
    for (size_t i = 0; i < n; ++i) {
        cpd_status s = cpd::check_design_public(designs[i], probe);
        if (per_design) per_design[i] = s;
        if (s == CPD_OK)
            valid.push_back((uint32_t)i);
        else
            out[i].size = out[i].info = out[i].deviation = NAN;
    }
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    st = c->run_gpu_fitness(designs, valid, seeds, n_seeds, max_n, nwarp, max_eval, out, per_design);
    if (st != CPD_OK && per_design)
        for (uint32_t i : valid) per_design[i] = st;
    return st;
}

cpd_status cpd_gpu_fitness(cpd_ctx *ctx, const cpd_design *design, const double *seeds, size_t n_seeds,
                           uint32_t max_n, uint32_t nwarp, uint32_t max_eval, cpd_fitness *out) {
    if (!design || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    cpd_status per = CPD_OK;
    cpd_status st = cpd_gpu_fitness_batch(ctx, design, 1, seeds, n_seeds, max_n, nwarp, max_eval, out, &per);
    if (st != CPD_OK) return st;
    if (per == CPD_ERR_NOT_CONVERGED)
        return cpd::fail(CPD_ERR_NOT_CONVERGED, "bin errors above 5e-3 after %zu seeds", n_seeds);
    return per;
}

cpd_status cpd_p_dets_l_wavelength(cpd_ctx *ctx, const cpd_design *design, const double *wavelengths,
                                   size_t n_wavelengths, uint32_t max_m, uint32_t mode, double *out) {
    if (!design || (!wavelengths && n_wavelengths) || (!out && n_wavelengths))
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n_wavelengths == 0) return CPD_OK;
    if (n_wavelengths > (1u << 24)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "too many wavelengths");
    cpd_sampling s{(uint32_t)n_wavelengths, max_m, mode, 0};
    cpd_status st = cpd::check_design_public(*design, s);
    if (st != CPD_OK) return st;
    Context *c;
    if ((st = resolve(ctx, &c)) != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    if (mode == CPD_MODE_ADAPTIVE) return c->rows_adaptive(*design, wavelengths, n_wavelengths, max_m, out);
    std::vector<uint32_t> valid{0};
    return c->run(design, &s, valid, true, wavelengths, nullptr, out, nullptr);
}

cpd_status cpd_propagate_rays(cpd_ctx *ctx, const cpd_design *design, const float *wavelength_cdf,
                              const float *ray_cdf, size_t n, uint32_t *bin_out, float *t_out) {
    if (n == 0) return CPD_OK;
    if (!design || !wavelength_cdf || !ray_cdf || !bin_out || !t_out)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > (1u << 28)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "too many rays");
    cpd_sampling s{1, 1, CPD_MODE_FIXED, 0};
    cpd_status st = cpd::check_design_public(*design, s);
    if (st != CPD_OK) return st;
    Context *c;
    if ((st = resolve(ctx, &c)) != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    return c->propagate(*design, wavelength_cdf, ray_cdf, n, bin_out, t_out);
}

cpd_status cpd_build_designs(cpd_ctx *ctx, const cpd_params *params, size_t n, double acceptance,
                             cpd_design *designs_out, cpd_status *per_design) {
    if (n == 0) return CPD_OK;
    if (!params || !designs_out || !per_design) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n > (1u << 24)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "more than 2^24 designs");
    if (!(acceptance >= 0.0 && acceptance <= 1.0) && acceptance != CPD_POSITION_FROM_EXIT)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "acceptance must be in [0, 1] or CPD_POSITION_FROM_EXIT");
    Context *c;
    cpd_status st = resolve(ctx, &c);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(c->mu);
    return c->build_designs(params, n, acceptance, designs_out, per_design);
}

}  // extern "C"

namespace cpd {

cpd_status check_design_public(const cpd_design &g, const cpd_sampling &s) { return check_design(g, s); }

// Upload of the parameters (one copy from the pinned stage), one thread per
// design, download of the records and statuses.
cpd_status Context::build_designs(const cpd_params *params, size_t n, double acceptance, cpd_design *out,
                                  cpd_status *status) {
    HIP_TRY(hipSetDevice(device));
    const size_t off_p = 256, off_d = align_up(off_p + n * sizeof(cpd_params), 256);
    const size_t off_s = align_up(off_d + n * sizeof(cpd_design), 256);
    const size_t total = align_up(off_s + n * sizeof(int), 256);
    cpd_status st = build_buf.reserve(total, false, stream);
    if (st != CPD_OK) return st;
    if ((st = host_stage.reserve(off_d)) != CPD_OK) return st;
    char *h = (char *)host_stage.ptr;
    memcpy(h, &acceptance, sizeof acceptance);
    memcpy(h + off_p, params, n * sizeof(cpd_params));
    char *d = (char *)build_buf.ptr;
    HIP_TRY(hipMemcpyAsync(d, h, off_d, hipMemcpyHostToDevice, stream));
    HIP_TRY(launch_construct(stream, (const cpd_params *)(d + off_p), (uint32_t)n, (const double *)d, nullptr,
                             (cpd_design *)(d + off_d), (int *)(d + off_s)));
    HIP_TRY(hipMemcpyAsync(out, d + off_d, n * sizeof(cpd_design), hipMemcpyDeviceToHost, stream));
    static_assert(sizeof(cpd_status) == sizeof(int), "status is an int");
    HIP_TRY(hipMemcpyAsync(status, d + off_s, n * sizeof(int), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return CPD_OK;
}

cpd_status Context::propagate(const cpd_design &g, const float *wcdf, const float *rcdf, size_t n, uint32_t *bin,
                              float *t) {
    HIP_TRY(hipSetDevice(device));
    const size_t off_g = 0, off_t = align_up(sizeof(cpd_design), 256);
    const size_t off_w = align_up(off_t + sizeof(TraceDesign), 256);
    const size_t off_r = align_up(off_w + n * sizeof(float), 256);
    const size_t off_b = align_up(off_r + n * sizeof(float), 256);
    const size_t off_o = align_up(off_b + n * sizeof(uint32_t), 256);
    const size_t total = align_up(off_o + n * sizeof(float), 256);
    cpd_status st = prop_buf.reserve(total, false, stream);
    if (st != CPD_OK) return st;
    if ((st = host_stage.reserve(off_b)) != CPD_OK) return st;
    char *h = (char *)host_stage.ptr;
    memcpy(h + off_g, &g, sizeof g);
    TraceDesign td;
    make_trace_design(g, 1, 1, td);
    memcpy(h + off_t, &td, sizeof td);
    memcpy(h + off_w, wcdf, n * sizeof(float));
    memcpy(h + off_r, rcdf, n * sizeof(float));
    char *d = (char *)prop_buf.ptr;
    HIP_TRY(hipMemcpyAsync(d, h, off_b, hipMemcpyHostToDevice, stream));
    HIP_TRY(launch_propagate(stream, (const cpd_design *)(d + off_g), (const TraceDesign *)(d + off_t),
                             (const float *)(d + off_w), (const float *)(d + off_r), (uint32_t)n,
                             (uint32_t *)(d + off_b), (float *)(d + off_o)));
    HIP_TRY(hipMemcpyAsync(bin, d + off_b, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipMemcpyAsync(t, d + off_o, n * sizeof(float), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    return CPD_OK;
}

}  // namespace cpd
