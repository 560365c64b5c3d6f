// givens_host.cpp -- the device's Sellmeier1 fit (fit_sellmeier1_givens in
// csrc/glassfit.h) compiled for the host, so its least squares is checked on
// the CPU (test infrastructure; built by tests/test_host.py).
#include "glassfit.h"

extern "C" int givens_fit(const double *b, const double *c, double lo, double hi, uint32_t npts, double *coeffs,
                          double *chisq) {
    return cpd::fit_sellmeier1_givens(b, c, lo, hi, npts, coeffs, chisq);
}
