// server.hip -- continuous batching of fitness requests (cpd_server_* in cpd.h).
//
// A host thread owns a HIP stream and three device pools: design records (one
// slot per design), u64 accumulators and last-ray bins.  Requests are
// admitted first come, first served while their pool regions fit.
// Accumulation is exact integer arithmetic, so a design's result does not
// depend on which requests share the GPU with it.
//
// Steps mode: each step gives every admitted request an equal share of the
// step's work items (leftovers go to the oldest), launches one persistent
// trace over all the shares, finalizes the requests whose last item ran,
// copies their results back and frees their regions.  A request arriving
// during a step is admitted before the next one, so its work runs alongside
// the requests already in flight.
//
// Resident mode: the thread stages admitted requests in mapped host memory,
// publishes them in chunks, keeps the resident loop (resident_kernel) running
// while any request is open, and collects results as the loop writes their
// tags; it launches nothing per request.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cpd.h"
#include "cpd_internal.h"
#include "kernels.h"
#include "runtime.h"

namespace cpd {
namespace {

using Clock = std::chrono::steady_clock;

#define SRV_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess) return fail(CPD_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// First-fit allocator of ranges in [0, capacity).
class RangeAlloc {
   public:
    void reset(uint64_t cap) {
        free_.clear();
        if (cap) free_[0] = cap;
    }
    bool alloc(uint64_t n, uint64_t &off) {
        if (n == 0) {
            off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const uint64_t rest = it->second - n, start = it->first + n;
                free_.erase(it);
                if (rest) free_[start] = rest;
                return true;
            }
        }
        return false;
    }
    void release(uint64_t off, uint64_t n) {
        if (n == 0) return;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

   private:
    std::map<uint64_t, uint64_t> free_;
};

struct Request {
    uint64_t id = 0;
    size_t n_total = 0;
    std::vector<cpd_design> designs;  // the valid designs
    std::vector<cpd_sampling> sampling;
    std::vector<TraceDesign> traces;  // FP32 trace records (tile and histogram shape)
    std::vector<uint32_t> index;      // position of each valid design in the submission
    std::vector<uint64_t> items;      // work items per valid design
    std::vector<cpd_status> status;   // per submitted design
    std::vector<cpd_fitness> out;     // per submitted design
    uint64_t total_items = 0, done_items = 0;
    uint32_t cursor = 0;              // next design with work
    uint64_t cursor_item = 0;         // next item of that design
    uint64_t slot0 = 0, acc0 = 0, lb0 = 0, acc_n = 0, lb_n = 0;
    uint32_t max_rays = 0;
    uint32_t tag = 0;                 // resident mode: written per slot when a design is done
    uint64_t chunk0 = 0, n_chunks = 0;  // resident mode: its chunk numbers
    size_t checked = 0;               // resident mode: designs known done
    bool finished = false;
    cpd_status error = CPD_OK;
    Clock::time_point t_submit, t_done;
};

struct Entry {
    uint32_t slot;
    uint64_t offset, count;
    uint64_t n_rec;  // wavelength samples its items cover (its records in the step)
};

}  // namespace

struct Server {
    int device = 0;
    int cu_count = 0;
    hipStream_t stream = nullptr;
    uint32_t step_items = 0;
    uint32_t waves_per_cu = 24;
    uint64_t max_designs = 0, acc_cap = 0, lb_cap = 0;
    DeviceBuffer d_designs, d_traces, d_acc_begin, d_lb_begin, d_acc, d_last_bin, d_results, d_step, d_counter;
    DeviceBuffer d_lrec;  // the wavelength records of a step
    BeamTables beam;
    HostBuffer h_stage, h_results;
    RangeAlloc slots, accs, lbs;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::vector<std::shared_ptr<Request>> active;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> all;
    uint64_t next_id = 1;
    bool stop = false;
    cpd_status fatal = CPD_OK;
    cpd_server_stats stats{};
    std::thread worker;

    // resident mode
    bool resident = false;
    uint32_t max_bins = 0, max_rays_cap = 0, chunk_cap = 0, grid = 0, hist_words = 0;
    uint64_t next_chunk = 0;
    std::vector<uint8_t> chunk_busy;  // ring positions held by open requests
    DeviceBuffer d_items_total, d_items_done, d_slot_tag, d_ingested, d_claim, d_chunk_slot0, d_chunk_n, d_prefix,
        d_ctrl;
    void *h_mapped = nullptr;         // control block | chunks | staged designs | results (mapped host memory)
    ResidentHost *h_ctrl = nullptr;
    ResidentChunk *h_chunks = nullptr;
    StagedDesign *h_staged = nullptr;
    ResidentResult *h_res = nullptr;
    ResidentArgs rargs{};
    hipEvent_t ev_loop = nullptr;
    bool loop_running = false;
    Clock::time_point loop_t0;

    cpd_status init(Context *ctx, const cpd_server_opts *o);
    cpd_status init_resident(const cpd_server_opts *o);
    void run();
    cpd_status step();
    cpd_status run_resident();
    bool admit_resident(Request &r);
    void publish(Request &r);
    cpd_status ensure_loop();
    bool collect_done(Request &r);
    static void chunk_split(const Request &r, std::vector<std::pair<uint32_t, uint32_t>> *out);
    bool admit(Request &r);
    void release(Request &r);
    cpd_status upload(const std::vector<std::shared_ptr<Request>> &fresh);
    cpd_status ensure_tables(uint32_t n);
};

cpd_status Server::init(Context *ctx, const cpd_server_opts *o) {
    device = ctx->device;
    cu_count = ctx->cu_count;
    waves_per_cu = ctx->waves_per_cu;
    const uint64_t budget = o && o->memory_budget ? o->memory_budget : (4ull << 30);
    max_designs = o && o->max_designs ? o->max_designs : 65536;
    // large steps: per-step host work (shares, finalize, copies) is amortised over
    // ~10 ms of device work, and a step only holds the work that is queued
    step_items = o && o->step_items ? o->step_items : (uint32_t)(64 * cu_count * waves_per_cu);
    const uint64_t record_bytes = max_designs * (sizeof(cpd_design) + sizeof(TraceDesign) + 2 * sizeof(uint64_t) +
                                                 sizeof(cpd_fitness));
    if (budget <= record_bytes + (1 << 20))
        return fail(CPD_ERR_INVALID_ARGUMENT, "memory budget below the %llu bytes of design slots",
                    (unsigned long long)record_bytes);
    const uint64_t rest = budget - record_bytes;
    acc_cap = rest / 10;  // 80% of the rest as u64 accumulators
    lb_cap = rest / 20;   // 20% as u32 last-ray bins
    SRV_TRY(hipSetDevice(device));
    SRV_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    cpd_status st;
    if ((st = d_designs.reserve(max_designs * sizeof(cpd_design), false, stream)) != CPD_OK) return st;
    if ((st = d_traces.reserve(max_designs * sizeof(TraceDesign), false, stream)) != CPD_OK) return st;
    if ((st = d_acc_begin.reserve(max_designs * sizeof(uint64_t), false, stream)) != CPD_OK) return st;
    if ((st = d_lb_begin.reserve(max_designs * sizeof(uint64_t), false, stream)) != CPD_OK) return st;
    if ((st = d_results.reserve(max_designs * sizeof(cpd_fitness), false, stream)) != CPD_OK) return st;
    if ((st = d_acc.reserve(acc_cap * sizeof(unsigned long long), true, stream)) != CPD_OK) return st;
    if ((st = d_last_bin.reserve(lb_cap * sizeof(uint32_t), false, stream)) != CPD_OK) return st;
    if ((st = d_counter.reserve(kWorkCounterWords * sizeof(unsigned int), true, stream)) != CPD_OK) return st;
    if ((st = h_results.reserve(max_designs * sizeof(cpd_fitness))) != CPD_OK) return st;
    SRV_TRY(hipStreamSynchronize(stream));
    slots.reset(max_designs);
    accs.reset(acc_cap);
    lbs.reset(lb_cap);
    if (o && o->mode == CPD_SERVER_RESIDENT) {
        if ((st = init_resident(o)) != CPD_OK) return st;
    } else if (o && o->mode != CPD_SERVER_STEPS) {
        return fail(CPD_ERR_INVALID_ARGUMENT, "unknown server mode %u", o->mode);
    }
    worker = std::thread([this] { run(); });
    return CPD_OK;
}

// Resident mode: ring of chunk positions, per-slot counters, mapped host
// memory, the beam tables for max_rays rays and the loop's launch shape.
cpd_status Server::init_resident(const cpd_server_opts *o) {
    resident = true;
    max_bins = o->max_bins ? o->max_bins : 256;
    max_rays_cap = o->max_rays ? o->max_rays : 65536;
    if (max_bins > CPD_MAX_BINS || max_rays_cap > kMaxTileRays)
        return fail(CPD_ERR_INVALID_ARGUMENT, "max_bins <= %d and max_rays <= %u", CPD_MAX_BINS, kMaxTileRays);
    chunk_cap = (uint32_t)std::max<uint64_t>(64, max_designs);
    chunk_busy.assign(chunk_cap, 0);
    hist_words = 1;
    for (uint32_t b = 1; b <= max_bins; ++b) hist_words = std::max(hist_words, hist_cols(b) * b);
    cpd_status st;
    const size_t D = max_designs, C = chunk_cap;
    if ((st = d_items_total.reserve(D * 4, false, stream)) != CPD_OK) return st;
    if ((st = d_items_done.reserve(D * 4, false, stream)) != CPD_OK) return st;
    if ((st = d_slot_tag.reserve(D * 4, false, stream)) != CPD_OK) return st;
    if ((st = d_ingested.reserve(C * 8, true, stream)) != CPD_OK) return st;
    if ((st = d_claim.reserve(C * 8 * kWorkCounters, true, stream)) != CPD_OK) return st;
    if ((st = d_chunk_slot0.reserve(C * 4, false, stream)) != CPD_OK) return st;
    if ((st = d_chunk_n.reserve(C * 4, false, stream)) != CPD_OK) return st;
    if ((st = d_prefix.reserve(C * 65 * 4, false, stream)) != CPD_OK) return st;
    if ((st = d_ctrl.reserve(kResCtrlWords * kResLine * 8, true, stream)) != CPD_OK) return st;
    if ((st = beam.ensure(max_rays_cap, stream)) != CPD_OK) return st;
    const size_t off_chunks = 256, off_staged = off_chunks + (C * sizeof(ResidentChunk) + 255) / 256 * 256;
    const size_t off_res = off_staged + (D * sizeof(StagedDesign) + 255) / 256 * 256;
    const size_t mapped = off_res + D * sizeof(ResidentResult);
    SRV_TRY(hipHostMalloc(&h_mapped, mapped, hipHostMallocMapped | hipHostMallocCoherent));
    memset(h_mapped, 0, mapped);
    h_ctrl = (ResidentHost *)h_mapped;
    h_chunks = (ResidentChunk *)((char *)h_mapped + off_chunks);
    h_staged = (StagedDesign *)((char *)h_mapped + off_staged);
    h_res = (ResidentResult *)((char *)h_mapped + off_res);
    void *dev = nullptr;
    SRV_TRY(hipHostGetDevicePointer(&dev, h_mapped, 0));
    ResidentArgs &a = rargs;
    a.host = (ResidentHost *)dev;
    a.chunks = (const ResidentChunk *)((char *)dev + off_chunks);
    a.staged = (const StagedDesign *)((char *)dev + off_staged);
    a.results = (ResidentResult *)((char *)dev + off_res);
    a.chunk_cap = chunk_cap;
    a.max_designs = (uint32_t)max_designs;
    a.designs = d_designs.as<cpd_design>();
    a.traces = d_traces.as<TraceDesign>();
    a.acc_begin = d_acc_begin.as<uint64_t>();
    a.lb_begin = d_lb_begin.as<uint64_t>();
    a.acc = d_acc.as<unsigned long long>();
    a.last_bin = d_last_bin.as<uint32_t>();
    a.items_total = d_items_total.as<unsigned int>();
    a.items_done = d_items_done.as<unsigned int>();
    a.slot_tag = d_slot_tag.as<uint32_t>();
    a.ingested = d_ingested.as<unsigned long long>();
    a.claim_word = d_claim.as<unsigned long long>();
    a.chunk_slot0 = d_chunk_slot0.as<uint32_t>();
    a.chunk_n = d_chunk_n.as<uint32_t>();
    a.prefix = d_prefix.as<uint32_t>();
    a.ctrl = d_ctrl.as<unsigned long long>();
    a.z32 = beam.z32.as<float>();
    a.zk = beam.zk.as<uint2>();
    a.z64 = beam.z64.as<double>();
    a.zs64 = beam.zs64.as<double>();
    a.inv = beam.inv.as<uint32_t>();
    a.max_rays = beam.n;
    a.hist_words = hist_words;
    a.row_words = max_bins;
    // a safety net: the host retires the loop when no request is open; the
    // wavefronts also exit by themselves after about a second without work
    a.idle_limit = 1u << 15;
    a.epoch = 0;
    const size_t lds = resident_lds_bytes(a.hist_words, a.row_words);
    int per_cu = 0;
    SRV_TRY(resident_occupancy(lds, &per_cu));
    grid = (uint32_t)cu_count * (uint32_t)std::max(1, std::min<int>(per_cu, (int)waves_per_cu));
    SRV_TRY(hipEventCreateWithFlags(&ev_loop, hipEventDisableTiming));
    SRV_TRY(hipStreamSynchronize(stream));
    return CPD_OK;
}

// Chunks of a request: consecutive designs, at most 64 and fewer than 2^18
// work items each (the claim word's fields); (first design, count) pairs.
void Server::chunk_split(const Request &r, std::vector<std::pair<uint32_t, uint32_t>> *out) {
    out->clear();
    uint32_t first = 0, n = 0;
    uint64_t items = 0;
    for (uint32_t q = 0; q < (uint32_t)r.designs.size(); ++q) {
        if (n == 64 || (n && items + r.items[q] >= (1u << 18))) {
            out->emplace_back(first, n);
            first = q;
            n = 0;
            items = 0;
        }
        ++n;
        items += r.items[q];
    }
    if (n) out->emplace_back(first, n);
}

bool Server::admit_resident(Request &r) {
    std::vector<std::pair<uint32_t, uint32_t>> ch;
    chunk_split(r, &ch);
    if (ch.size() > chunk_cap) return false;
    for (size_t i = 0; i < ch.size(); ++i)
        if (chunk_busy[(next_chunk + i) % chunk_cap]) return false;
    if (!admit(r)) return false;
    r.chunk0 = next_chunk;
    r.n_chunks = ch.size();
    for (size_t i = 0; i < ch.size(); ++i) chunk_busy[(next_chunk + i) % chunk_cap] = 1;
    next_chunk += ch.size();
    return true;
}

// Stage an admitted request's designs and chunk descriptors in mapped host
// memory, then publish its chunks.
void Server::publish(Request &r) {
    const size_t n = r.designs.size();
    uint64_t acc = r.acc0, lb = r.lb0;
    for (size_t q = 0; q < n; ++q) {
        StagedDesign &sd = h_staged[r.slot0 + q];
        sd.design = r.designs[q];
        sd.trace = r.traces[q];
        sd.acc_begin = acc;
        sd.lb_begin = lb;
        sd.items = (uint32_t)r.items[q];
        acc += (uint64_t)r.sampling[q].n_wavelengths * r.designs[q].bin_count;
        lb += r.sampling[q].n_wavelengths;
    }
    std::vector<std::pair<uint32_t, uint32_t>> ch;
    chunk_split(r, &ch);
    for (size_t i = 0; i < ch.size(); ++i) {
        const uint64_t c = r.chunk0 + i;
        ResidentChunk &d = h_chunks[c % chunk_cap];
        d.slot0 = (uint32_t)(r.slot0 + ch[i].first);
        d.n = ch[i].second;
        d.tag = r.tag;
        d.pad = 0;
        d.seq1 = c + 1;
    }
    __atomic_store_n(&h_ctrl->published, (unsigned long long)(r.chunk0 + r.n_chunks), __ATOMIC_RELEASE);
}

// Launch the loop unless it is running; true work for it exists whenever a
// request is open, and the loop exits by itself once it finds none.
cpd_status Server::ensure_loop() {
    if (loop_running) {
        const hipError_t e = hipEventQuery(ev_loop);
        if (e == hipErrorNotReady) return CPD_OK;
        SRV_TRY(e);
        loop_running = false;
        std::lock_guard<std::mutex> lk(mu);
        stats.busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - loop_t0).count();
    }
    rargs.epoch++;
    SRV_TRY(launch_resident(stream, rargs, grid));
    SRV_TRY(hipEventRecord(ev_loop, stream));
    loop_running = true;
    loop_t0 = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    stats.steps++;
    return CPD_OK;
}

// Whether every design of r has its tag in the results (designs finish
// roughly in order: the scan resumes where it stopped).
bool Server::collect_done(Request &r) {
    const size_t n = r.designs.size();
    while (r.checked < n &&
           __atomic_load_n(&h_res[r.slot0 + r.checked].tag, __ATOMIC_ACQUIRE) == r.tag)
        ++r.checked;
    return r.checked == n;
}

cpd_status Server::run_resident() {
    SRV_TRY(hipSetDevice(device));
    std::vector<std::shared_ptr<Request>> fresh, act;
    for (;;) {
        fresh.clear();
        {
            std::unique_lock<std::mutex> lk(mu);
            if (active.empty() && queue.empty() && !stop) {
                // idle: the loop exits by itself; wake for work or to note its exit
                cv_work.wait_for(lk, std::chrono::milliseconds(loop_running ? 2 : 50),
                                 [&] { return stop || !queue.empty(); });
            }
            if (stop) break;
            while (!queue.empty()) {
                std::shared_ptr<Request> r = queue.front();
                if (!admit_resident(*r)) break;
                queue.pop_front();
                active.push_back(r);
                fresh.push_back(r);
            }
            act = active;
            stats.queued = (uint32_t)queue.size();
            stats.admitted = (uint32_t)active.size();
        }
        for (auto &r : fresh) publish(*r);
        if (!act.empty()) {
            cpd_status st = ensure_loop();
            if (st != CPD_OK) return st;
        } else if (loop_running) {
            // no open request: retire this launch (its wavefronts exit at their next poll)
            __atomic_store_n(&h_ctrl->stop, rargs.epoch, __ATOMIC_RELEASE);
            if (hipEventQuery(ev_loop) != hipErrorNotReady) {
                loop_running = false;
                std::lock_guard<std::mutex> lk(mu);
                stats.busy_ms += std::chrono::duration<double, std::milli>(Clock::now() - loop_t0).count();
            }
        }
        const unsigned int err = __atomic_load_n(&h_ctrl->error, __ATOMIC_ACQUIRE);
        if (err) return fail(CPD_ERR_DEVICE, "resident loop found an inconsistent chunk (code %u)", err);
        bool any_done = false;
        const Clock::time_point now = Clock::now();
        for (auto &rp : act) {
            Request &r = *rp;
            if (!collect_done(r)) continue;
            std::lock_guard<std::mutex> lk(mu);
            for (size_t q = 0; q < r.designs.size(); ++q) r.out[r.index[q]] = h_res[r.slot0 + q].fit;
            r.finished = true;
            r.t_done = now;
            release(r);
            for (uint64_t c = r.chunk0; c < r.chunk0 + r.n_chunks; ++c) chunk_busy[c % chunk_cap] = 0;
            stats.completed++;
            stats.designs += r.designs.size();
            stats.items += r.total_items;
            any_done = true;
        }
        if (any_done) {
            std::lock_guard<std::mutex> lk(mu);
            std::vector<std::shared_ptr<Request>> keep;
            for (const auto &r : active)
                if (!r->finished) keep.push_back(r);
            active.swap(keep);
            cv_done.notify_all();
        } else if (!act.empty()) {
            std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
    // stop: the loop's wavefronts exit at their next poll
    __atomic_store_n(&h_ctrl->stop, 0xffffffffu, __ATOMIC_RELEASE);
    SRV_TRY(hipStreamSynchronize(stream));
    loop_running = false;
    return CPD_OK;
}

bool Server::admit(Request &r) {
    uint64_t s, a, l;
    if (!slots.alloc(r.designs.size(), s)) return false;
    if (!accs.alloc(r.acc_n, a)) {
        slots.release(s, r.designs.size());
        return false;
    }
    if (!lbs.alloc(r.lb_n, l)) {
        slots.release(s, r.designs.size());
        accs.release(a, r.acc_n);
        return false;
    }
    r.slot0 = s;
    r.acc0 = a;
    r.lb0 = l;
    return true;
}

void Server::release(Request &r) {
    slots.release(r.slot0, r.designs.size());
    accs.release(r.acc0, r.acc_n);
    lbs.release(r.lb0, r.lb_n);
}

cpd_status Server::ensure_tables(uint32_t n) { return beam.ensure(n, stream); }

// Copy the records of newly admitted requests into their pool slots.
cpd_status Server::upload(const std::vector<std::shared_ptr<Request>> &fresh) {
    for (const auto &rp : fresh) {
        Request &r = *rp;
        const size_t n = r.designs.size();
        const size_t bytes = n * (sizeof(cpd_design) + sizeof(TraceDesign) + 2 * sizeof(uint64_t));
        cpd_status st = h_stage.reserve(bytes);
        if (st != CPD_OK) return st;
        SRV_TRY(hipStreamSynchronize(stream));  // staging reuse
        char *h = (char *)h_stage.ptr;
        cpd_design *hd = (cpd_design *)h;
        TraceDesign *ht = (TraceDesign *)(h + n * sizeof(cpd_design));
        uint64_t *ha = (uint64_t *)(h + n * (sizeof(cpd_design) + sizeof(TraceDesign)));
        uint64_t *hl = ha + n;
        uint64_t acc = r.acc0, lb = r.lb0;
        for (size_t q = 0; q < n; ++q) {
            hd[q] = r.designs[q];
            ht[q] = r.traces[q];
            ha[q] = acc;
            hl[q] = lb;
            acc += (uint64_t)r.sampling[q].n_wavelengths * r.designs[q].bin_count;
            lb += r.sampling[q].n_wavelengths;
        }
        SRV_TRY(hipMemcpyAsync(d_designs.as<cpd_design>() + r.slot0, hd, n * sizeof(cpd_design),
                               hipMemcpyHostToDevice, stream));
        SRV_TRY(hipMemcpyAsync(d_traces.as<TraceDesign>() + r.slot0, ht, n * sizeof(TraceDesign),
                               hipMemcpyHostToDevice, stream));
        SRV_TRY(hipMemcpyAsync(d_acc_begin.as<uint64_t>() + r.slot0, ha, n * sizeof(uint64_t),
                               hipMemcpyHostToDevice, stream));
        SRV_TRY(hipMemcpyAsync(d_lb_begin.as<uint64_t>() + r.slot0, hl, n * sizeof(uint64_t),
                               hipMemcpyHostToDevice, stream));
    }
    return CPD_OK;
}

// One scheduling step over the admitted requests.
cpd_status Server::step() {
    std::vector<std::shared_ptr<Request>> fresh, act;
    {
        std::lock_guard<std::mutex> lk(mu);
        while (!queue.empty()) {
            std::shared_ptr<Request> r = queue.front();
            if (!admit(*r)) break;
            queue.pop_front();
            active.push_back(r);
            fresh.push_back(r);
        }
        act = active;
        stats.queued = (uint32_t)queue.size();
        stats.admitted = (uint32_t)active.size();
    }
    if (act.empty()) return CPD_OK;
    const Clock::time_point t0 = Clock::now();
    SRV_TRY(hipSetDevice(device));
    cpd_status st = upload(fresh);
    if (st != CPD_OK) return st;
    uint32_t max_rays = 0;
    for (const auto &r : act) max_rays = std::max(max_rays, r->max_rays);
    if ((st = ensure_tables(max_rays)) != CPD_OK) return st;

    // shares: equal split of the step, leftovers to the oldest requests
    const size_t A = act.size();
    std::vector<uint64_t> take(A, 0);
    uint64_t budget = step_items;
    const uint64_t share = std::max<uint64_t>(1, budget / A);
    for (size_t i = 0; i < A && budget; ++i) {
        take[i] = std::min<uint64_t>({share, act[i]->total_items - act[i]->done_items, budget});
        budget -= take[i];
    }
    for (size_t i = 0; i < A && budget; ++i) {
        const uint64_t more = std::min<uint64_t>(act[i]->total_items - act[i]->done_items - take[i], budget);
        take[i] += more;
        budget -= more;
    }
    std::vector<Entry> entries;
    std::vector<uint32_t> finishing;  // slots to finalize
    std::vector<size_t> done_req;
    uint32_t hist_words = 1, row_words = 1;
    for (size_t i = 0; i < A; ++i) {
        Request &r = *act[i];
        uint64_t left = take[i];
        while (left) {
            const uint64_t avail = r.items[r.cursor] - r.cursor_item;
            const uint64_t t = std::min(avail, left);
            const uint64_t nt = r.traces[r.cursor].n_tiles;
            entries.push_back(Entry{(uint32_t)(r.slot0 + r.cursor), r.cursor_item, t,
                                    (r.cursor_item + t - 1) / nt - r.cursor_item / nt + 1});
            hist_words = std::max(hist_words, r.traces[r.cursor].hist_cols * r.traces[r.cursor].nbin);
            row_words = std::max(row_words, r.traces[r.cursor].nbin);
            left -= t;
            r.cursor_item += t;
            if (r.cursor_item == r.items[r.cursor]) {
                ++r.cursor;
                r.cursor_item = 0;
            }
        }
        r.done_items += take[i];
        if (r.done_items == r.total_items) {
            done_req.push_back(i);
            for (size_t q = 0; q < r.designs.size(); ++q) finishing.push_back((uint32_t)(r.slot0 + q));
        }
    }
    // step arrays: slot[E] | item_begin[E+1] | item_offset[E] | lrec_begin[E] | finalize slots[F]
    const size_t E = entries.size(), F = finishing.size();
    const size_t off_begin = (E * sizeof(uint32_t) + 255) / 256 * 256;
    const size_t off_offset = off_begin + ((E + 1) * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t off_rec = off_offset + (E * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t off_fin = off_rec + (E * sizeof(uint64_t) + 255) / 256 * 256;
    const size_t bytes = off_fin + F * sizeof(uint32_t) + 256;
    SRV_TRY(hipStreamSynchronize(stream));  // staging reuse after uploads
    if ((st = h_stage.reserve(bytes)) != CPD_OK) return st;
    if ((st = d_step.reserve(bytes, false, stream)) != CPD_OK) return st;
    char *h = (char *)h_stage.ptr;
    uint32_t *hslot = (uint32_t *)h;
    uint64_t *hbegin = (uint64_t *)(h + off_begin);
    uint64_t *hoff = (uint64_t *)(h + off_offset);
    uint64_t *hrec = (uint64_t *)(h + off_rec);
    uint32_t *hfin = (uint32_t *)(h + off_fin);
    uint64_t items = 0, n_rec = 0;
    for (size_t e = 0; e < E; ++e) {
        hslot[e] = entries[e].slot;
        hbegin[e] = items;
        hoff[e] = entries[e].offset;
        hrec[e] = n_rec;
        items += entries[e].count;
        n_rec += entries[e].n_rec;
    }
    if ((st = d_lrec.reserve(n_rec * sizeof(LambdaRecord), false, stream)) != CPD_OK) return st;
    hbegin[E] = items;
    for (size_t f = 0; f < F; ++f) hfin[f] = finishing[f];
    char *d = (char *)d_step.ptr;
    SRV_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));

    BatchArgs a{};
    a.designs = d_designs.as<cpd_design>();
    a.traces = d_traces.as<TraceDesign>();
    a.slot = (const uint32_t *)d;
    a.item_begin = (const uint64_t *)(d + off_begin);
    a.item_offset = (const uint64_t *)(d + off_offset);
    a.acc_begin = d_acc_begin.as<uint64_t>();
    a.lb_begin = d_lb_begin.as<uint64_t>();
    a.acc = d_acc.as<unsigned long long>();
    a.last_bin = d_last_bin.as<uint32_t>();
    a.z32 = beam.z32.as<float>();
    a.z64 = beam.z64.as<double>();
    a.zk = beam.zk.as<uint2>();
    a.zs64 = beam.zs64.as<double>();
    a.inv = beam.inv.as<uint32_t>();
    a.work_counter = d_counter.as<unsigned int>();
    a.lrec = d_lrec.as<LambdaRecord>();
    a.lrec_begin = (const uint64_t *)(d + off_rec);
    a.n_designs = (uint32_t)E;
    a.hist_words = hist_words;
    a.row_words = row_words;
    a.n_items = items;
    SRV_TRY(launch_zero_u32(stream, a.work_counter, kWorkCounterWords));
    const uint32_t grid = (uint32_t)std::min<uint64_t>(items, (uint64_t)cu_count * waves_per_cu);
    SRV_TRY(launch_trace(stream, a, grid, false));
    if (F) {
        FinalizeArgs f{};
        f.designs = a.designs;
        f.traces = a.traces;
        f.slot = (const uint32_t *)(d + off_fin);
        f.out_stride_slot = 1;
        f.acc_begin = a.acc_begin;
        f.lb_begin = a.lb_begin;
        f.acc = a.acc;
        f.last_bin = a.last_bin;
        f.out = d_results.as<cpd_fitness>();
        f.n_designs = (uint32_t)F;
        SRV_TRY(launch_finalize(stream, f));
        for (size_t i : done_req) {
            const Request &r = *act[i];
            SRV_TRY(hipMemcpyAsync((cpd_fitness *)h_results.ptr + r.slot0, d_results.as<cpd_fitness>() + r.slot0,
                                   r.designs.size() * sizeof(cpd_fitness), hipMemcpyDeviceToHost, stream));
        }
    }
    SRV_TRY(hipStreamSynchronize(stream));
    const Clock::time_point t1 = Clock::now();
    {
        std::lock_guard<std::mutex> lk(mu);
        for (size_t i : done_req) {
            Request &r = *act[i];
            const cpd_fitness *res = (const cpd_fitness *)h_results.ptr + r.slot0;
            for (size_t q = 0; q < r.designs.size(); ++q) r.out[r.index[q]] = res[q];
            r.finished = true;
            r.t_done = t1;
            release(r);
            stats.completed++;
            stats.designs += r.designs.size();
        }
        if (!done_req.empty()) {
            std::vector<std::shared_ptr<Request>> keep;
            for (const auto &r : active)
                if (!r->finished) keep.push_back(r);
            active.swap(keep);
        }
        stats.steps++;
        stats.items += items;
        stats.busy_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    if (!done_req.empty()) cv_done.notify_all();
    return CPD_OK;
}

void Server::run() {
    if (resident) {
        const cpd_status st = run_resident();
        if (st != CPD_OK) {
            if (h_ctrl) __atomic_store_n(&h_ctrl->stop, 0xffffffffu, __ATOMIC_RELEASE);
            (void)hipStreamSynchronize(stream);
            loop_running = false;
            std::lock_guard<std::mutex> lk(mu);
            fatal = st;
            for (auto &kv : all) {
                Request &r = *kv.second;
                if (!r.finished) {
                    r.finished = true;
                    r.error = st;
                    r.t_done = Clock::now();
                }
            }
            queue.clear();
            active.clear();
            cv_done.notify_all();
        }
        return;
    }
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || !queue.empty() || !active.empty(); });
            if (stop) break;
        }
        const cpd_status st = step();
        if (st != CPD_OK) {
            std::lock_guard<std::mutex> lk(mu);
            fatal = st;
            for (auto &kv : all) {
                Request &r = *kv.second;
                if (!r.finished) {
                    r.finished = true;
                    r.error = st;
                    r.t_done = Clock::now();
                }
            }
            queue.clear();
            active.clear();
            cv_done.notify_all();
            break;
        }
    }
}

}  // namespace cpd

using cpd::Server;

extern "C" {

cpd_status cpd_server_create(cpd_ctx *ctx, const cpd_server_opts *opts, cpd_server **out) {
    if (!out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL output");
    cpd::Context *c = nullptr;
    cpd_status st = ctx ? CPD_OK : cpd::default_context(&c);
    if (st != CPD_OK) return st;
    if (ctx) c = reinterpret_cast<cpd::Context *>(ctx);
    Server *s = new Server();
    st = s->init(c, opts);
    if (st != CPD_OK) {
        delete s;
        return st;
    }
    *out = reinterpret_cast<cpd_server *>(s);
    return CPD_OK;
}

cpd_status cpd_server_submit(cpd_server *srv, const cpd_design *designs, const cpd_sampling *sampling, size_t n,
                             uint64_t *request_id) {
    Server *s = reinterpret_cast<Server *>(srv);
    if (!s || !request_id || (n && (!designs || !sampling)))
        return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    auto r = std::make_shared<cpd::Request>();
    r->n_total = n;
    r->status.assign(n, CPD_OK);
    r->out.assign(n, cpd_fitness{NAN, NAN, NAN});
    for (size_t i = 0; i < n; ++i) {
        const cpd_status ds = cpd::check_design_public(designs[i], sampling[i]);
        if (ds == CPD_OK && sampling[i].mode != CPD_MODE_FIXED) {
            r->status[i] = CPD_ERR_INVALID_ARGUMENT;  // the server runs fixed-count sampling
            continue;
        }
        r->status[i] = ds;
        if (ds != CPD_OK) continue;
        r->index.push_back((uint32_t)i);
        r->designs.push_back(designs[i]);
        r->sampling.push_back(sampling[i]);
        r->traces.emplace_back();
        cpd::make_trace_design(designs[i], sampling[i].n_wavelengths, sampling[i].n_rays, r->traces.back());
        r->items.push_back((uint64_t)sampling[i].n_wavelengths * r->traces.back().n_tiles);
        r->total_items += r->items.back();
        r->acc_n += (uint64_t)sampling[i].n_wavelengths * designs[i].bin_count;
        r->lb_n += sampling[i].n_wavelengths;
        r->max_rays = std::max(r->max_rays, sampling[i].n_rays);
    }
    if (s->resident) {
        for (size_t q = 0; q < r->designs.size(); ++q)
            if (r->designs[q].bin_count > s->max_bins || r->sampling[q].n_rays > s->max_rays_cap ||
                r->items[q] >= (1u << 18)) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->stats.rejected++;
                return cpd::fail(CPD_ERR_CAPACITY,
                                 "resident server: a design has %u bins, %u rays per wavelength, %llu work items "
                                 "(limits %u, %u, 2^18)",
                                 r->designs[q].bin_count, r->sampling[q].n_rays, (unsigned long long)r->items[q],
                                 s->max_bins, s->max_rays_cap);
            }
    }
    if (r->designs.size() > s->max_designs || r->acc_n > s->acc_cap || r->lb_n > s->lb_cap) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.rejected++;
        return cpd::fail(CPD_ERR_CAPACITY, "request needs %zu slots, %llu accumulators: above the server's budget",
                         r->designs.size(), (unsigned long long)r->acc_n);
    }
    r->t_submit = cpd::Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal != CPD_OK) return cpd::fail(s->fatal, "server stopped after a device error");
    r->id = s->next_id++;
    r->tag = (uint32_t)((r->id - 1) % 0xffffffffull) + 1;  // never 0, the results' initial tag
    *request_id = r->id;
    s->all[r->id] = r;
    s->stats.submitted++;
    if (r->designs.empty()) {
        r->finished = true;
        r->t_done = r->t_submit;
        s->cv_done.notify_all();
        return CPD_OK;
    }
    s->queue.push_back(r);
    s->cv_work.notify_one();
    return CPD_OK;
}

static cpd_status collect(Server *s, std::unique_lock<std::mutex> &lk, uint64_t id, int64_t timeout_ms,
                          cpd_fitness *out, cpd_status *per_design, double *latency_ms) {
    auto it = s->all.find(id);
    if (it == s->all.end()) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "unknown request %llu", (unsigned long long)id);
    std::shared_ptr<cpd::Request> r = it->second;
    if (!r->finished && timeout_ms != 0) {
        auto pred = [&] { return r->finished; };
        if (timeout_ms < 0)
            s->cv_done.wait(lk, pred);
        else
            s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    }
    if (!r->finished) return CPD_ERR_NOT_READY;
    if (out) memcpy(out, r->out.data(), r->n_total * sizeof(cpd_fitness));
    if (per_design)
        for (size_t i = 0; i < r->n_total; ++i) per_design[i] = r->error != CPD_OK ? r->error : r->status[i];
    if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    s->all.erase(id);
    return r->error;
}

cpd_status cpd_server_poll(cpd_server *srv, uint64_t request_id, cpd_fitness *out, cpd_status *per_design,
                           double *latency_ms) {
    Server *s = reinterpret_cast<Server *>(srv);
    if (!s) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    return collect(s, lk, request_id, 0, out, per_design, latency_ms);
}

cpd_status cpd_server_wait(cpd_server *srv, uint64_t request_id, int64_t timeout_ms, cpd_fitness *out,
                           cpd_status *per_design, double *latency_ms) {
    Server *s = reinterpret_cast<Server *>(srv);
    if (!s) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    return collect(s, lk, request_id, timeout_ms, out, per_design, latency_ms);
}

cpd_status cpd_server_stats_get(cpd_server *srv, cpd_server_stats *out) {
    Server *s = reinterpret_cast<Server *>(srv);
    if (!s || !out) return cpd::fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->queued = (uint32_t)s->queue.size();
    out->admitted = (uint32_t)s->active.size();
    return CPD_OK;
}

void cpd_server_destroy(cpd_server *srv) {
    Server *s = reinterpret_cast<Server *>(srv);
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (cpd::DeviceBuffer *b : {&s->d_designs, &s->d_traces, &s->d_acc_begin, &s->d_lb_begin, &s->d_acc,
                                 &s->d_last_bin, &s->d_results, &s->d_step, &s->d_counter, &s->d_lrec, &s->d_items_total,
                                 &s->d_items_done, &s->d_slot_tag, &s->d_ingested, &s->d_claim, &s->d_chunk_slot0,
                                 &s->d_chunk_n, &s->d_prefix, &s->d_ctrl})
        b->release();
    if (s->h_mapped) (void)hipHostFree(s->h_mapped);
    if (s->ev_loop) (void)hipEventDestroy(s->ev_loop);
    s->beam.release();
    s->h_stage.release();
    s->h_results.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // extern "C"
