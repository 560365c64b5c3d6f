// outer_fold.h -- host side of the adaptive estimators: the fold of a
// design's rows p(d|lambda) into I(Lambda; D) (src/fitness.rs:79-123), and a
// small parallel loop.
#pragma once

#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "welford.h"

namespace cpd {

// Per-design state of the outer fold over rows (fitness.rs:89-114,
// cuda_fitness.rs:366-380).
struct OuterFold {
    std::vector<W64> p_dets;
    W64 h{0.0, 0.0, 0.0};
    uint64_t rows = 0, rays = 0;
    bool done = false;
    double e2 = kMaxErr * kMaxErr;  // 0: no early exit (every row is folded)

    void reset(uint32_t nbin) {
        p_dets.assign(nbin, W64{0.0, 0.0, 0.0});
        h = W64{0.0, 0.0, 0.0};
        rows = rays = 0;
        done = false;
    }
    // Folds one row; returns whether every bin's mean is precise enough.
    bool fold(const double *row) {
        double h_row = 0.0;
        for (size_t b = 0; b < p_dets.size(); ++b) {
            w_next(p_dets[b], row[b]);
            h_row += plog2p(row[b]);
        }
        w_next(h, h_row);
        ++rows;
        return converged();
    }
    bool converged() const {
        for (const W64 &w : p_dets)
            if (!w_sem_le(w, e2)) return false;
        return true;
    }
    // H(D) - H(D|Lambda) (fitness.rs:115-122)
    double info() const {
        double h_det = 0.0;
        for (const W64 &w : p_dets) h_det = h_det + plog2p(w.mean);
        return h.mean - h_det;
    }
};

// Runs body(i) for i in [0, n) on up to 16 host threads.
template <class F>
void parallel_for(size_t n, F body) {
    const size_t hw = std::max<size_t>(1, std::min<size_t>(16, std::thread::hardware_concurrency()));
    const size_t threads = std::min(hw, n / 64 + 1);
    if (threads <= 1) {
        for (size_t i = 0; i < n; ++i) body(i);
        return;
    }
    std::atomic<size_t> next{0};
    std::vector<std::thread> pool;
    for (size_t t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (size_t i; (i = next.fetch_add(1)) < n;) body(i);
        });
    for (auto &th : pool) th.join();
}

}  // namespace cpd
