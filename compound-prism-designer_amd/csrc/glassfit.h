// glassfit.h -- the Sellmeier1 -> degree-5 polynomial fit of a glass
// (compound_prism_designer/glasscat.py:114-127, 144-152) in a form that runs
// one glass per GPU thread: least squares over the sample wavelengths by
// Givens rotations applied to one row at a time, so a thread holds only the
// 6 x 6 triangle, its right-hand side and the current row (no m x 6 matrix).
// Columns are scaled by their largest magnitude over [lo, hi] (the monomials
// w^j of positive w peak at an end), as the host's Householder fit scales
// them.  Compiled for host and device.
#pragma once

#include <math.h>
#include <stdint.h>

#include "cpd.h"
#include "trace_core.h"

namespace cpd {

// Sellmeier1.index_of_refraction (glasscat.py:144-152).
CPD_HD double sellmeier1_n(const double b[3], const double c[3], double w) {
#pragma clang fp contract(off)
    const double w2 = w * w;
    return sqrt(1.0 + b[0] * w2 / (w2 - c[0]) + b[1] * w2 / (w2 - c[1]) + b[2] * w2 / (w2 - c[2]));
}

// numpy.linspace(lo, hi, npts)[i]
CPD_HD double fit_wavelength(double lo, double hi, uint32_t npts, uint32_t i) {
#pragma clang fp contract(off)
    const double step = (hi - lo) / (double)(npts - 1);
    return i == npts - 1 ? hi : lo + (double)i * step;
}

// Fits n(w) = c0 w^5 + ... + c5 (Horner order, Glass<F, 6>) to Sellmeier1 at
// npts equally spaced wavelengths in [lo, hi]; 0 on success, -1 when a sample
// is not finite (a pole in range) or the columns are dependent.
CPD_HD int fit_sellmeier1_givens(const double b[3], const double c[3], double lo, double hi, uint32_t npts,
                                 double coeffs[CPD_GLASS_COEFFS], double *chisq) {
#pragma clang fp contract(off)
    constexpr int n = CPD_GLASS_COEFFS;
    double scale[n];
    {
        const double m = fmax(fabs(lo), fabs(hi));
        double p = 1.0;
        for (int j = n - 1; j >= 0; --j) {  // column j holds w^(5 - j)
            scale[j] = p;
            p *= m;
        }
    }
    double R[n][n], z[n];
    for (int i = 0; i < n; ++i) {
        z[i] = 0.0;
        for (int j = 0; j < n; ++j) R[i][j] = 0.0;
    }
    for (uint32_t i = 0; i < npts; ++i) {
        const double w = fit_wavelength(lo, hi, npts, i);
        double y = sellmeier1_n(b, c, w);
        if (!isfinite(y)) return -1;
        double x[n];
        double p = 1.0;
        for (int j = n - 1; j >= 0; --j) {
            x[j] = p / scale[j];
            p *= w;
        }
        // rotate the row into the triangle
        for (int k = 0; k < n; ++k) {
            if (x[k] == 0.0) continue;
            const double r = hypot(R[k][k], x[k]);
            const double cs = R[k][k] / r, sn = x[k] / r;
            R[k][k] = r;
            for (int j = k + 1; j < n; ++j) {
                const double t = cs * R[k][j] + sn * x[j];
                x[j] = cs * x[j] - sn * R[k][j];
                R[k][j] = t;
            }
            const double t = cs * z[k] + sn * y;
            y = cs * y - sn * z[k];
            z[k] = t;
        }
    }
    for (int k = 0; k < n; ++k)
        if (R[k][k] == 0.0) return -1;
    double xs[n];
    for (int k = n - 1; k >= 0; --k) {
        double s = z[k];
        for (int j = k + 1; j < n; ++j) s -= R[k][j] * xs[j];
        xs[k] = s / R[k][k];
    }
    for (int j = 0; j < n; ++j) coeffs[j] = xs[j] / scale[j];
    if (chisq) {
        double s = 0.0;
        for (uint32_t i = 0; i < npts; ++i) {
            const double w = fit_wavelength(lo, hi, npts, i);
            const double r = sellmeier1_n(b, c, w) - glass_n64(coeffs, w);
            s += r * r;
        }
        *chisq = s;
    }
    return 0;
}

}  // namespace cpd
