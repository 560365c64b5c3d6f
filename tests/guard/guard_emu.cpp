// guard_emu.cpp -- soundness check of the FP32 trace's error margins, on the
// host (test infrastructure; built by tests/test_guard_soundness.py).
//
// The device traces rays in FP32 (trace32 in csrc/trace_core.h) and re-traces
// in FP64 (trace64_chain, the reference's operation order) every ray whose
// FP32 outcome lies within an error margin of a hit/miss or bin decision.  The
// margins are sound when no ray that trace32 does NOT flag lands in a
// different bin than the FP64 trace.  This harness compiles the same
// trace32 source for the host, with the device's approximate reciprocal and
// square root modelled as the correctly rounded value moved by -1, 0 or +1 ulp
// (drawn per call), and compares every ray of every (design, wavelength) of a
// batch with the FP64 trace: the same beam quantiles z_k = norminv(u_k), the
// same per-wavelength constants (make_lambda_consts, make_hot), y computed in
// FP32 as the kernel does.  It counts unflagged bin mismatches (must be 0),
// flagged rays, hits, and the largest transmittance error of unflagged hits.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <thread>
#include <vector>

#include "cpd.h"
#include "trace_core.h"

namespace {
thread_local uint64_t g_state = 0x9E3779B97F4A7C15ull;
thread_local int g_mode = 0;  // 0: exact; 1: random -1/0/+1 ulp per call

inline int draw() {
    // SplitMix64 step; three outcomes
    uint64_t z = (g_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (int)(z % 3) - 1;
}

inline float perturb(float v) {
    if (g_mode == 0 || !isfinite(v) || v == 0.f) return v;
    const int d = draw();
    if (d > 0) return nextafterf(v, INFINITY);
    if (d < 0) return nextafterf(v, -INFINITY);
    return v;
}
}  // namespace

namespace cpd {
float cpd_emu_rcp(float x) { return perturb(1.0f / x); }
float cpd_emu_sqrt(float x) { return perturb(sqrtf(x)); }
}  // namespace cpd

extern "C" {

// stats[0] rays, [1] flagged, [2] hits (FP64), [3] unflagged bin mismatches,
// [4] unflagged hits with |t32 - t64| > 1e-5 t64, [5] flagged rays whose FP32
// bin differs from the FP64 one (what the margins catch); terr[0] and terr[1]
// the largest absolute and relative transmittance error of unflagged hits.
// mode 0: exact rcp/sqrt, 1: +-1 ulp.
int guard_check(const cpd_design *designs, size_t n, uint32_t n_lambda, uint32_t n_rays, int mode, int threads,
                uint64_t *stats, double *terr) {
    std::vector<double> z64(n_rays);
    std::vector<float> z32(n_rays);
    for (uint32_t k = 0; k < n_rays; ++k) {
        z64[k] = cpd::norminv64(cpd::qmc_u64((uint64_t)k + 1));
        z32[k] = (float)z64[k];
    }
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> tot[6];
    for (auto &t : tot) t = 0;
    std::vector<double> worst(threads > 0 ? threads : 1, 0.0), worst_rel(worst);
    auto work = [&](int tid) {
        g_mode = mode;
        g_state = 0x9E3779B97F4A7C15ull * (uint64_t)(tid + 1);
        uint64_t loc[6] = {0, 0, 0, 0, 0, 0};
        double w = 0.0, wr = 0.0;
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n) break;
            const cpd_design &g = designs[i];
            cpd::TraceDesign td;
            cpd::make_trace_design(g, n_lambda, n_rays, td);
            for (uint32_t j = 1; j <= n_lambda; ++j) {
                const double lambda = cpd::lambda_of(g, cpd::qmc_u64(j));
                cpd::LambdaConsts lc;
                cpd::Chain64 chain;
                cpd::make_lambda_consts(g, lambda, lc, &chain);
                const cpd::HotRecord hot = cpd::make_hot(td, lc);
                for (uint32_t k = 0; k < n_rays; ++k) {
                    double t64 = 0.0;
                    const uint32_t b64 = cpd::trace64(g, lambda, cpd::beam_y64(g, z64[k]), t64);
                    loc[0]++;
                    loc[2] += b64 < g.bin_count;
                    if (!lc.ok) {
                        // the kernel traces nothing for this wavelength: every ray is lost
                        if (b64 < g.bin_count) loc[3]++;
                        continue;
                    }
                    const float y = hot.y_mean - hot.beam_w * z32[k];
                    const cpd::Trace32Result r = cpd::trace32<false>(hot, y);
                    const uint32_t b32 = r.hit ? r.bin : g.bin_count;
                    if (r.guard) {
                        loc[1]++;
                        loc[5] += b32 != b64;
                        continue;
                    }
                    if (b32 != b64) {
                        loc[3]++;
                        continue;
                    }
                    if (b64 < g.bin_count) {
                        const double e = fabs((double)r.t - t64);
                        if (e > w) w = e;
                        if (e / t64 > wr) wr = e / t64;
                        if (e > 1e-5 * t64) loc[4]++;
                    }
                }
            }
        }
        for (int q = 0; q < 6; ++q) tot[q] += loc[q];
        worst[tid] = w;
        worst_rel[tid] = wr;
    };
    const int T = threads > 0 ? threads : 1;
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    for (int q = 0; q < 6; ++q) stats[q] = tot[q];
    terr[0] = terr[1] = 0.0;
    for (size_t t = 0; t < worst.size(); ++t) {
        terr[0] = worst[t] > terr[0] ? worst[t] : terr[0];
        terr[1] = worst_rel[t] > terr[1] ? worst_rel[t] : terr[1];
    }
    return 0;
}

}  // extern "C"

#include <algorithm>

extern "C" {

// The band test (band_class, trace_core.h) over the banded beam order the
// device traces in (within each block of 4096 rays sorted by height), as the
// kernel applies it to a design whose rays form one tile: sub-band u is the
// ranks 16 u .. 16 u + 15 of every block, group s the ranks 64 s .. 64 s + 63,
// each bounded by the heights of its lowest and highest rank over the blocks;
// the 64 groups are classified first, then the sub-bands of the mixed ones.
// The kernel's bands are an iteration's 64 rays (interleaved over the four
// blocks of a whole 16384-position chunk, else 64 consecutive positions); a
// band is skipped when every sub-band its positions' ranks fall in is lost.
// The positions cover the rays rounded up to whole blocks; in the last block
// the positions of rays past n_rays are idle (and count towards the bounds).
// stats[0] bands, [1] bands a wavefront would take past the planar-chain
// test, [2] of those skipped, [3] skipped bands (any) holding an FP64 hit
// (must be 0), [4] bands holding an FP64 hit, [5] wavelengths the test over
// every height proves lost (the trace skips them whole), [6] of those holding
// an FP64 hit (must be 0).
int band_check(const cpd_design *designs, size_t n, uint32_t n_lambda, uint32_t n_rays, int threads,
               uint64_t *stats) {
    const uint32_t full = (n_rays + cpd::kBandRays - 1) / cpd::kBandRays * cpd::kBandRays;
    std::vector<double> z64(full);
    std::vector<uint32_t> order(full);
    for (uint32_t k = 0; k < full; ++k) {
        z64[k] = cpd::norminv64(cpd::qmc_u64((uint64_t)k + 1));
        order[k] = k;
    }
    for (uint32_t b = 0; b < full; b += cpd::kBandRays)
        std::sort(order.begin() + b, order.begin() + b + cpd::kBandRays, [&](uint32_t x, uint32_t y) {
            return z64[x] < z64[y] || (z64[x] == z64[y] && x < y);
        });
    // the kernel's bands (0-based positions): a chunk of 16384 whole banded
    // positions interleaved (band i: ranks 16 i .. 16 i + 15 of each block of
    // 4096), otherwise 64 consecutive positions
    const uint32_t kChunk = 4 * cpd::kBandRays;
    std::vector<uint32_t> band_pos;  // 64 positions per band, in band order
    std::vector<uint32_t> band_chunk;
    for (uint32_t c0 = 0; c0 < full; c0 += kChunk) {
        const bool inter = c0 + kChunk <= full;
        const uint32_t len = inter ? kChunk : full - c0;
        for (uint32_t b = 0; b < len / 64; ++b) {
            for (uint32_t l = 0; l < 64; ++l)
                band_pos.push_back(inter ? c0 + (l >> 4) * cpd::kBandRays + 16 * b + (l & 15) : c0 + 64 * b + l);
            band_chunk.push_back(c0);
        }
    }
    const size_t n_bands = band_chunk.size();
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> tot[7];
    for (auto &t : tot) t = 0;
    auto work = [&]() {
        uint64_t loc[7] = {0, 0, 0, 0, 0, 0, 0};
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n) break;
            const cpd_design &g = designs[i];
            cpd::TraceDesign td;
            cpd::make_trace_design(g, n_lambda, n_rays, td);
            for (uint32_t j = 1; j <= n_lambda; ++j) {
                const double lambda = cpd::lambda_of(g, cpd::qmc_u64(j));
                cpd::LambdaConsts lc;
                cpd::Chain64 chain;
                cpd::LensBand64 band;
                cpd::make_lambda_consts(g, lambda, lc, &chain, &band);
                const cpd::HotRecord hot = cpd::make_hot(td, lc);
                // lowest and highest height of ranks r0 .. r1 over every block
                const uint32_t nblk = full / cpd::kBandRays;
                auto hull = [&](uint32_t r0, uint32_t r1, double &lo, double &hi) {
                    double zl = INFINITY, zh = -INFINITY;
                    for (uint32_t q = 0; q < nblk; ++q) {
                        zl = fmin(zl, z64[order[q * cpd::kBandRays + r0]]);
                        zh = fmax(zh, z64[order[q * cpd::kBandRays + r1]]);
                    }
                    lo = cpd::beam_y64(g, zl);
                    hi = cpd::beam_y64(g, zh);
                };
                bool sub_lost[256];
                for (uint32_t gs = 0; gs < 64; ++gs) {
                    double ya, yb;
                    hull(64 * gs, 64 * gs + 63, ya, yb);
                    const int cls = cpd::band_class(band, ya, yb);
                    for (uint32_t q = 0; q < 4; ++q) {
                        const uint32_t u = 4 * gs + q;
                        sub_lost[u] = cls == cpd::kBandLost;
                        if (cls == cpd::kBandMixed) {
                            hull(16 * u, 16 * u + 15, ya, yb);
                            sub_lost[u] = !cpd::band_may_hit(band, ya, yb);
                        }
                    }
                }
                // the records kernel's test over every height of the design's
                // positions (idle ones included)
                double wa = INFINITY, wb = -INFINITY;
                for (uint32_t q = 0; q < full; ++q) {
                    const double y = cpd::beam_y64(g, z64[q]);
                    wa = fmin(wa, y);
                    wb = fmax(wb, y);
                }
                const bool dead = lc.ok && cpd::band_class(band, wa, wb) == cpd::kBandLost;
                loc[5] += dead;
                bool any_hit = false;
                for (size_t b = 0; b < n_bands; ++b) {
                    loc[0]++;
                    bool chain_pass = false, hit = false;
                    for (uint32_t q = 0; q < 64; ++q) {
                        const uint32_t k = order[band_pos[64 * b + q]];
                        if (k >= n_rays) continue;  // idle position
                        const float y = hot.y_mean - hot.beam_w * (float)z64[k];
                        const bool alive = (y >= hot.ylo) & (y <= hot.yhi);
                        const bool guard = fminf(fabsf(y - hot.ylo), fabsf(y - hot.yhi)) <= hot.ymarg;
                        chain_pass |= alive | guard;
                        double t;
                        hit |= lc.ok && cpd::trace64(g, lambda, cpd::beam_y64(g, z64[k]), t) < g.bin_count;
                    }
                    loc[4] += hit;
                    any_hit |= hit;
                    if (!lc.ok) continue;
                    bool skipped = true;
                    for (uint32_t q = 0; q < 64; ++q) skipped = skipped && sub_lost[(band_pos[64 * b + q] % cpd::kBandRays) / 16];
                    if (chain_pass) {
                        loc[1]++;
                        loc[2] += skipped;
                    }
                    if (skipped) loc[3] += hit;
                }
                if (dead) loc[6] += any_hit;
            }
        }
        for (int q = 0; q < 7; ++q) tot[q] += loc[q];
    };
    const int T = threads > 0 ? threads : 1;
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work);
    work();
    for (auto &th : pool) th.join();
    for (int q = 0; q < 7; ++q) stats[q] = tot[q];
    return 0;
}

}  // extern "C"

extern "C" {

// The affine FP64 trace of the queue drains (trace64_affine) against trace64,
// the reference's operation order, on every ray of every (design,
// wavelength) with a valid planar chain: stats[0] rays, [1] rays it leaves
// to trace64_chain, [2] decided rays whose bin differs from trace64's (must
// be 0), [3] decided hits; terr[0] the largest relative transmittance
// difference of a decided hit.
int affine_check(const cpd_design *designs, size_t n, uint32_t n_lambda, uint32_t n_rays, int threads,
                 uint64_t *stats, double *terr) {
    std::vector<double> z64(n_rays);
    for (uint32_t k = 0; k < n_rays; ++k) z64[k] = cpd::norminv64(cpd::qmc_u64((uint64_t)k + 1));
    std::atomic<size_t> next{0};
    std::atomic<uint64_t> tot[4];
    for (auto &t : tot) t = 0;
    const int T = threads > 0 ? threads : 1;
    std::vector<double> worst(T, 0.0);
    auto work = [&](int tid) {
        uint64_t loc[4] = {0, 0, 0, 0};
        double w = 0.0;
        for (;;) {
            const size_t i = next.fetch_add(1);
            if (i >= n) break;
            const cpd_design &g = designs[i];
            for (uint32_t j = 1; j <= n_lambda; ++j) {
                const double lambda = cpd::lambda_of(g, cpd::qmc_u64(j));
                cpd::LambdaConsts lc;
                cpd::Chain64 chain;
                cpd::LensBand64 band;
                cpd::make_lambda_consts(g, lambda, lc, &chain, &band);
                if (!lc.ok) continue;
                for (uint32_t k = 0; k < n_rays; ++k) {
                    const double y = cpd::beam_y64(g, z64[k]);
                    double t64 = 0.0, ta = 0.0;
                    const uint32_t b64 = cpd::trace64(g, lambda, y, t64);
                    uint32_t ba = 0;
                    loc[0]++;
                    if (!cpd::trace64_affine(band, chain, g, y, ba, ta)) {
                        loc[1]++;
                        continue;
                    }
                    if (ba != b64) {
                        loc[2]++;
                        continue;
                    }
                    if (b64 < g.bin_count) {
                        loc[3]++;
                        const double e = fabs(ta - t64) / t64;
                        if (e > w) w = e;
                    }
                }
            }
        }
        for (int q = 0; q < 4; ++q) tot[q] += loc[q];
        worst[tid] = w;
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < T; ++t) pool.emplace_back(work, t);
    work(0);
    for (auto &th : pool) th.join();
    for (int q = 0; q < 4; ++q) stats[q] = tot[q];
    terr[0] = 0.0;
    for (double v : worst) terr[0] = v > terr[0] ? v : terr[0];
    return 0;
}

}  // extern "C"
