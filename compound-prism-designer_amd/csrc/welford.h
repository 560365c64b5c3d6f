// welford.h -- running mean and M2 in FP64 (host and device), in the
// operation order of src/spectrometer/welford.rs, contraction off.
#pragma once

#include <math.h>

#include "trace_core.h"

namespace cpd {

struct W64 {
    double count, mean, m2;
};

// welford.rs:26-32 next_sample
CPD_HD void w_next(W64 &w, double x) {
#pragma clang fp contract(off)
    w.count += 1.0;
    const double delta = x - w.mean;
    w.mean += delta / w.count;
    const double delta2 = x - w.mean;
    w.m2 += delta * delta2;
}

// welford.rs:34-42 skip: pad with zeros up to new_count; M2 uses the updated mean
CPD_HD void w_skip(W64 &w, double new_count) {
#pragma clang fp contract(off)
    if (w.count < new_count) {
        const double zero_count = new_count - w.count;
        const double c = new_count;
        w.mean = (w.count * w.mean) / c;
        w.m2 += w.mean * w.mean * w.count * zero_count / c;
        w.count = c;
    }
}

// x / d for an integer d (1 <= d <= 2^24) from rd = RN(1/d): one fused
// correction of x rd gives the correctly rounded quotient (Markstein's
// theorem: rd within half an ulp of 1/d and x rd within one ulp of x / d),
// i.e. exactly the value x / d has, in three dependent operations instead of
// a division sequence.
CPD_HD double div_int(double x, double d, double rd) {
#pragma clang fp contract(off)
    const double q0 = x * rd;
    return fma(fma(-q0, d, x), rd, q0);
}

// w_skip, w_next and w_add with their divisions by integer counts taken from
// reciprocals: rc = RN(1/new_count) for skip, RN(1/(count + 1)) for next,
// RN(1/(a.count + b.count)) for add.  Bit-identical to the plain forms.
CPD_HD void w_skip_r(W64 &w, double new_count, double rc) {
#pragma clang fp contract(off)
    if (w.count < new_count) {
        const double zero_count = new_count - w.count;
        w.mean = div_int(w.count * w.mean, new_count, rc);
        w.m2 += div_int(w.mean * w.mean * w.count * zero_count, new_count, rc);
        w.count = new_count;
    }
}
CPD_HD void w_next_r(W64 &w, double x, double rc) {
#pragma clang fp contract(off)
    w.count += 1.0;
    const double delta = x - w.mean;
    w.mean += div_int(delta, w.count, rc);
    const double delta2 = x - w.mean;
    w.m2 += delta * delta2;
}
CPD_HD void w_skip_mean_r(W64 &w, double new_count, double rc) {
#pragma clang fp contract(off)
    if (w.count < new_count) {
        w.mean = div_int(w.count * w.mean, new_count, rc);
        w.count = new_count;
    }
}
CPD_HD void w_next_mean_r(W64 &w, double x, double rc) {
#pragma clang fp contract(off)
    w.count += 1.0;
    w.mean += div_int(x - w.mean, w.count, rc);
}

// w_skip and w_next without the m2 update (for folds whose m2 is never read):
// the same count and mean, operation for operation.
CPD_HD void w_skip_mean(W64 &w, double new_count) {
#pragma clang fp contract(off)
    if (w.count < new_count) {
        w.mean = (w.count * w.mean) / new_count;
        w.count = new_count;
    }
}
CPD_HD void w_next_mean(W64 &w, double x) {
#pragma clang fp contract(off)
    w.count += 1.0;
    w.mean += (x - w.mean) / w.count;
}

// welford.rs:55-58: SEM^2 <= e^2 as m2 < e^2 n (n - 1)
CPD_HD bool w_sem_le(const W64 &w, double error_squared) {
#pragma clang fp contract(off)
    return w.m2 < error_squared * (w.count * (w.count - 1.0));
}

// welford.rs:76-85 Add: the parallel combination of two summaries
CPD_HD W64 w_add(const W64 &a, const W64 &b) {
#pragma clang fp contract(off)
    const double count = a.count + b.count;
    const double delta = b.mean - a.mean;
    const double mean = (a.count * a.mean + b.count * b.mean) / count;
    const double m2 = a.m2 + b.m2 + delta * delta * a.count * b.count / count;
    return W64{count, mean, m2};
}
CPD_HD W64 w_add_r(const W64 &a, const W64 &b, double rc) {
#pragma clang fp contract(off)
    const double count = a.count + b.count;
    const double delta = b.mean - a.mean;
    const double mean = div_int(a.count * a.mean + b.count * b.mean, count, rc);
    const double m2 = a.m2 + b.m2 + div_int(delta * delta * a.count * b.count, count, rc);
    return W64{count, mean, m2};
}

// utils.rs:196-202 plog2p
inline double plog2p(double p) { return p > 0.0 ? p * log2(p) : 0.0; }

// The reference's SEM thresholds: max_err = 5e-3, compared squared
// (fitness.rs:17,108-111; kernel.rs:103-104; cuda_fitness.rs:359-360).
static constexpr double kMaxErr = 5e-3;

}  // namespace cpd
