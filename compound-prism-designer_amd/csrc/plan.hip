// plan.hip -- captured generations (cpd_plan_* in cpd.h).
//
// An optimiser evaluates generation after generation of designs whose shape
// does not change: per slot the same sampling and bin count.  Everything that
// depends only on the shape (work-item prefix sums, accumulator offsets, the
// beam tables, the schedule) is staged on the device once, and the per-
// generation work (upload of the FP64 design records, the FP32 trace records
// built from them on the device, the persistent trace, the entropy reduction
// and the download of the results) is captured once in a HIP graph that each
// generation replays.  No memset node is captured: the work counter is
// cleared by a kernel and the reduction leaves the accumulators at zero.
// A generation can also arrive as parameter vectors (cpd_plan_run_params):
// a second graph uploads them and builds the FP64 design records on the
// device (construct_kernel) before the same evaluation, so the host does no
// per-design work at all.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "cpd.h"
#include "cpd_internal.h"
#include "kernels.h"
#include "runtime.h"

namespace cpd {

#define PLAN_TRY(expr)                                                                                 \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? CPD_ERR_OUT_OF_MEMORY : CPD_ERR_DEVICE, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                                 \
    } while (0)

struct Plan {
    int device = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    cpd_schedule sched{};
    uint32_t n = 0;
    std::vector<cpd_design> shape;      // per slot: a valid record of the slot's shape
    std::vector<cpd_sampling> sampling;  // per slot
    std::vector<uint32_t> perm;         // staging position -> slot
    DeviceBuffer d_fixed, d_designs, d_traces, d_acc, d_last_bin, d_results, d_counter, d_lrec;
    BeamTables beam;
    HostBuffer h_designs, h_out;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // the evaluation every generation runs, set up by init
    BatchArgs args{};
    FinalizeArgs fin{};
    uint32_t grid = 0;
    const cpd_sampling *d_sampling = nullptr;
    // generations from parameter vectors: [acceptance | params] uploaded in
    // one copy, the slots' shape records as stand-ins, per-slot statuses
    DeviceBuffer d_params, d_shape, d_status;
    HostBuffer h_params, h_status;
    hipGraph_t graph_p = nullptr;
    hipGraphExec_t exec_p = nullptr;

    cpd_status init(Context *ctx, const cpd_design *shape_in, const cpd_sampling *samp, size_t count,
                    const cpd_schedule *schedule);
    hipError_t capture_evaluation(const int *status = nullptr);
    cpd_status capture_params();
    cpd_status run(const cpd_design *designs, cpd_fitness *out, cpd_status *per_design);
    cpd_status run_params(const cpd_params *params, double acceptance, cpd_fitness *out, cpd_status *per_design);
    ~Plan();
};

Plan::~Plan() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (exec_p) (void)hipGraphExecDestroy(exec_p);
    if (graph_p) (void)hipGraphDestroy(graph_p);
    d_params.release();
    d_shape.release();
    d_status.release();
    h_params.release();
    h_status.release();
    d_fixed.release();
    d_designs.release();
    d_traces.release();
    d_acc.release();
    d_last_bin.release();
    d_lrec.release();
    d_results.release();
    d_counter.release();
    beam.release();
    h_designs.release();
    h_out.release();
    if (stream) (void)hipStreamDestroy(stream);
}

cpd_status Plan::init(Context *ctx, const cpd_design *shape_in, const cpd_sampling *samp, size_t count,
                      const cpd_schedule *schedule) {
    n = (uint32_t)count;
    shape.assign(shape_in, shape_in + count);
    sampling.assign(samp, samp + count);
    std::vector<uint32_t> all(n);
    for (uint32_t i = 0; i < n; ++i) {
        cpd_status st = check_design_public(shape[i], sampling[i]);
        if (st == CPD_OK && sampling[i].mode != CPD_MODE_FIXED)
            st = fail(CPD_ERR_INVALID_ARGUMENT, "plans run fixed-count sampling");
        if (st != CPD_OK) {
            const std::string why = cpd_last_error();
            return fail(st, "plan shape: design %u is not valid: %s", i, why.c_str());
        }
        all[i] = i;
    }
    if (schedule) {
        cpd_status st = check_schedule(*schedule);
        if (st != CPD_OK) return st;
        sched = *schedule;
    } else {
        std::lock_guard<std::mutex> lk(ctx->mu);
        cpd_status st = ctx->autotune(shape.data(), sampling.data(), n, all, &sched, nullptr);
        if (st != CPD_OK) return st;
    }
    perm = all;
    order_designs(shape.data(), sampling.data(), perm, sched.order, sched.tile_rays);

    // shape-only layout, in staging order
    std::vector<cpd_sampling> samp_q(n);
    std::vector<uint64_t> item_begin(n + 1), acc_begin(n), lb_begin(n);
    uint64_t items = 0, acc_n = 0, lb_n = 0, rays = 0;
    uint32_t hist_words = 0, row_words = 0, max_rays = 0;
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t i = perm[q];
        TraceDesign t;
        make_trace_design(shape[i], sampling[i].n_wavelengths, sampling[i].n_rays, t, sched.tile_rays);
        samp_q[q] = sampling[i];
        item_begin[q] = items;
        acc_begin[q] = acc_n;
        lb_begin[q] = lb_n;
        items += (uint64_t)sampling[i].n_wavelengths * t.n_tiles;
        rays += (uint64_t)sampling[i].n_wavelengths * sampling[i].n_rays;
        acc_n += (uint64_t)sampling[i].n_wavelengths * shape[i].bin_count;
        lb_n += sampling[i].n_wavelengths;
        hist_words = std::max(hist_words, t.hist_cols * t.nbin);
        row_words = std::max(row_words, t.nbin);
        max_rays = std::max(max_rays, sampling[i].n_rays);
    }
    item_begin[n] = items;
    if (items >= (1ull << 32)) return fail(CPD_ERR_INVALID_ARGUMENT, "plan too large (>= 2^32 work items)");

    device = ctx->device;
    PLAN_TRY(hipSetDevice(device));
    PLAN_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    // fixed arrays: sampling | item_begin | acc_begin | lb_begin
    const size_t off_samp = 0;
    const size_t off_items = (n * sizeof(cpd_sampling) + 255) / 256 * 256;
    const size_t off_acc = off_items + ((n + 1) * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t off_lb = off_acc + (n * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t fixed_bytes = off_lb + n * sizeof(uint64_t);
    cpd_status st;
    if ((st = d_fixed.reserve(fixed_bytes, false, stream)) != CPD_OK) return st;
    if ((st = d_designs.reserve(n * sizeof(cpd_design), false, stream)) != CPD_OK) return st;
    if ((st = d_traces.reserve(n * sizeof(TraceDesign), false, stream)) != CPD_OK) return st;
    if ((st = d_acc.reserve(acc_n * sizeof(unsigned long long), true, stream)) != CPD_OK) return st;
    if ((st = d_last_bin.reserve(lb_n * sizeof(uint32_t), false, stream)) != CPD_OK) return st;
    if ((st = d_lrec.reserve(lb_n * sizeof(LambdaRecord), false, stream)) != CPD_OK) return st;
    if ((st = d_results.reserve(n * sizeof(cpd_fitness), false, stream)) != CPD_OK) return st;
    if ((st = d_counter.reserve(kWorkCounterWords * sizeof(unsigned int), true, stream)) != CPD_OK) return st;
    if ((st = h_designs.reserve(n * sizeof(cpd_design))) != CPD_OK) return st;
    if ((st = h_out.reserve(n * sizeof(cpd_fitness))) != CPD_OK) return st;
    char *fx = d_fixed.as<char>();
    PLAN_TRY(hipMemcpyAsync(fx + off_samp, samp_q.data(), n * sizeof(cpd_sampling), hipMemcpyHostToDevice, stream));
    PLAN_TRY(hipMemcpyAsync(fx + off_items, item_begin.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                            stream));
    PLAN_TRY(hipMemcpyAsync(fx + off_acc, acc_begin.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    PLAN_TRY(hipMemcpyAsync(fx + off_lb, lb_begin.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice, stream));
    if ((st = beam.ensure(max_rays, stream)) != CPD_OK) return st;
    PLAN_TRY(hipStreamSynchronize(stream));  // host vectors above go out of scope

    BatchArgs &a = args;
    a.designs = d_designs.as<cpd_design>();
    a.traces = d_traces.as<TraceDesign>();
    a.item_begin = (const uint64_t *)(fx + off_items);
    a.acc_begin = (const uint64_t *)(fx + off_acc);
    a.lb_begin = (const uint64_t *)(fx + off_lb);
    a.acc = d_acc.as<unsigned long long>();
    a.last_bin = d_last_bin.as<uint32_t>();
    a.z32 = beam.z32.as<float>();
    a.z64 = beam.z64.as<double>();
    a.zk = beam.zk.as<uint2>();
    a.zs64 = beam.zs64.as<double>();
    a.inv = beam.inv.as<uint32_t>();
    a.work_counter = d_counter.as<unsigned int>();
    a.lrec = d_lrec.as<LambdaRecord>();
    a.n_designs = n;
    a.hist_words = hist_words;
    a.row_words = row_words;
    a.n_items = items;
    grid = (uint32_t)std::min<uint64_t>(items, (uint64_t)ctx->cu_count * sched.waves_per_cu);
    a.items_per_fetch = items_per_fetch(rays, items, grid);
    d_sampling = (const cpd_sampling *)(fx + off_samp);
    FinalizeArgs &f = fin;
    f.designs = a.designs;
    f.traces = a.traces;
    f.acc_begin = a.acc_begin;
    f.lb_begin = a.lb_begin;
    f.acc = a.acc;
    f.last_bin = a.last_bin;
    f.out = d_results.as<cpd_fitness>();
    f.n_designs = n;

    PLAN_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(d_designs.ptr, h_designs.ptr, n * sizeof(cpd_design), hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = capture_evaluation();
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(stream, &g);
    graph = g;
    if (e != hipSuccess) return fail(CPD_ERR_DEVICE, "plan capture: %s", hipGetErrorString(e));
    if (e_end != hipSuccess) return fail(CPD_ERR_DEVICE, "plan capture: %s", hipGetErrorString(e_end));
    PLAN_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    return CPD_OK;
}

// Trace records, cleared work counter, trace, reduction and download of the
// results, recorded on the stream being captured.  With status, the work
// items of designs whose status is not CPD_OK trace nothing.
hipError_t Plan::capture_evaluation(const int *status) {
    hipError_t e = launch_build_traces(stream, args.designs, d_sampling, n, sched.tile_rays, d_traces.as<TraceDesign>(),
                                       status);
    if (e == hipSuccess) e = launch_zero_u32(stream, args.work_counter, kWorkCounterWords);
    if (e == hipSuccess) e = launch_trace(stream, args, grid, false);
    if (e == hipSuccess) e = launch_finalize(stream, fin);
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_out.ptr, d_results.ptr, n * sizeof(cpd_fitness), hipMemcpyDeviceToHost, stream);
    return e;
}

// The parameter-vector graph, captured on first use.
cpd_status Plan::capture_params() {
    const size_t off_p = 256;
    const size_t bytes = off_p + (size_t)n * sizeof(cpd_params);
    cpd_status st;
    if ((st = d_params.reserve(bytes, false, stream)) != CPD_OK) return st;
    if ((st = h_params.reserve(bytes)) != CPD_OK) return st;
    if ((st = d_shape.reserve((size_t)n * sizeof(cpd_design), false, stream)) != CPD_OK) return st;
    if ((st = d_status.reserve((size_t)n * sizeof(int), false, stream)) != CPD_OK) return st;
    if ((st = h_status.reserve((size_t)n * sizeof(int))) != CPD_OK) return st;
    std::vector<cpd_design> standin(n);
    for (uint32_t q = 0; q < n; ++q) standin[q] = shape[perm[q]];
    PLAN_TRY(hipMemcpyAsync(d_shape.ptr, standin.data(), (size_t)n * sizeof(cpd_design), hipMemcpyHostToDevice, stream));
    PLAN_TRY(hipStreamSynchronize(stream));
    char *dp = (char *)d_params.ptr;
    PLAN_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipMemcpyAsync(dp, h_params.ptr, bytes, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess)
        e = launch_construct(stream, (const cpd_params *)(dp + off_p), n, (const double *)dp,
                             d_shape.as<cpd_design>(), d_designs.as<cpd_design>(), d_status.as<int>());
    if (e == hipSuccess) e = capture_evaluation(d_status.as<int>());
    if (e == hipSuccess)
        e = hipMemcpyAsync(h_status.ptr, d_status.ptr, (size_t)n * sizeof(int), hipMemcpyDeviceToHost, stream);
    hipGraph_t g = nullptr;
    const hipError_t e_end = hipStreamEndCapture(stream, &g);
    graph_p = g;
    if (e != hipSuccess) return fail(CPD_ERR_DEVICE, "plan capture: %s", hipGetErrorString(e));
    if (e_end != hipSuccess) return fail(CPD_ERR_DEVICE, "plan capture: %s", hipGetErrorString(e_end));
    PLAN_TRY(hipGraphInstantiate(&exec_p, graph_p, nullptr, nullptr, 0));
    return CPD_OK;
}

cpd_status Plan::run_params(const cpd_params *params, double acceptance, cpd_fitness *out, cpd_status *per_design) {
    std::lock_guard<std::mutex> lk(mu);
    PLAN_TRY(hipSetDevice(device));
    if (!exec_p) {
        cpd_status st = capture_params();
        if (st != CPD_OK) return st;
    }
    char *h = (char *)h_params.ptr;
    memcpy(h, &acceptance, sizeof acceptance);
    cpd_params *hp = (cpd_params *)(h + 256);
    for (uint32_t q = 0; q < n; ++q) hp[q] = params[perm[q]];
    PLAN_TRY(hipGraphLaunch(exec_p, stream));
    PLAN_TRY(hipStreamSynchronize(stream));
    const cpd_fitness *r = (const cpd_fitness *)h_out.ptr;
    const int *stq = (const int *)h_status.ptr;
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t i = perm[q];
        if (per_design) per_design[i] = (cpd_status)stq[q];
        if (stq[q] != CPD_OK)
            out[i].size = out[i].info = out[i].deviation = NAN;
        else
            out[i] = r[q];
    }
    return CPD_OK;
}

cpd_status Plan::run(const cpd_design *designs, cpd_fitness *out, cpd_status *per_design) {
    std::lock_guard<std::mutex> lk(mu);
    cpd_design *h = (cpd_design *)h_designs.ptr;
    std::vector<uint8_t> bad(n, 0);
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t i = perm[q];
        cpd_status s = check_design_public(designs[i], sampling[i]);
        if (s == CPD_OK && designs[i].bin_count != shape[i].bin_count) s = CPD_ERR_INVALID_ARGUMENT;
        if (per_design) per_design[i] = s;
        bad[q] = s != CPD_OK;
        h[q] = bad[q] ? shape[i] : designs[i];  // a stand-in keeps the slot's work well formed
    }
    PLAN_TRY(hipSetDevice(device));
    PLAN_TRY(hipGraphLaunch(exec, stream));
    PLAN_TRY(hipStreamSynchronize(stream));
    const cpd_fitness *r = (const cpd_fitness *)h_out.ptr;
    for (uint32_t q = 0; q < n; ++q) {
        const uint32_t i = perm[q];
        if (bad[q])
            out[i].size = out[i].info = out[i].deviation = NAN;
        else
            out[i] = r[q];
    }
    return CPD_OK;
}

}  // namespace cpd

extern "C" {

cpd_status cpd_plan_create(cpd_ctx *ctx, const cpd_design *shape, const cpd_sampling *sampling, size_t n,
                           const cpd_schedule *schedule, cpd_plan **out) {
    if (!shape || !sampling || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (n == 0 || n > (1u << 24)) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "plan size must be in 1..2^24");
    cpd::Context *c;
    cpd_status st = ctx ? (c = reinterpret_cast<cpd::Context *>(ctx), CPD_OK) : cpd::default_context(&c);
    if (st != CPD_OK) return st;
    cpd::Plan *p = new cpd::Plan();
    st = p->init(c, shape, sampling, n, schedule);
    if (st != CPD_OK) {
        delete p;
        return st;
    }
    *out = reinterpret_cast<cpd_plan *>(p);
    return CPD_OK;
}

cpd_status cpd_plan_run(cpd_plan *plan, const cpd_design *designs, cpd_fitness *out, cpd_status *per_design) {
    if (!plan || !designs || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    return reinterpret_cast<cpd::Plan *>(plan)->run(designs, out, per_design);
}

cpd_status cpd_plan_run_params(cpd_plan *plan, const cpd_params *params, double acceptance, cpd_fitness *out,
                               cpd_status *per_design) {
    if (!plan || !params || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(acceptance >= 0.0 && acceptance <= 1.0) && acceptance != CPD_POSITION_FROM_EXIT)
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "acceptance must be in [0, 1] or CPD_POSITION_FROM_EXIT");
    return reinterpret_cast<cpd::Plan *>(plan)->run_params(params, acceptance, out, per_design);
}

cpd_status cpd_plan_schedule(const cpd_plan *plan, cpd_schedule *out) {
    if (!plan || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = reinterpret_cast<const cpd::Plan *>(plan)->sched;
    return CPD_OK;
}

void cpd_plan_destroy(cpd_plan *plan) { delete reinterpret_cast<cpd::Plan *>(plan); }

}  // extern "C"
