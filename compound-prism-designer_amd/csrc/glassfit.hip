// glassfit.hip -- Sellmeier1 glass fits on the GPU, one glass per thread
// (glassfit.h), for catalogues fitted in one call.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <mutex>
#include <vector>

#include "cpd.h"
#include "cpd_internal.h"
#include "glassfit.h"
#include "runtime.h"

namespace cpd {
namespace {

#define GF_TRY(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(e_ == hipErrorOutOfMemory ? CPD_ERR_OUT_OF_MEMORY : CPD_ERR_DEVICE, "%s: %s", \
                        #expr, hipGetErrorString(e_));                                                 \
    } while (0)

// bc: n records of (B1, B2, B3, C1, C2, C3); out: n records of 6 coefficients
// and the residual sum of squares; status: 0 or CPD_ERR_INVALID_ARGUMENT.
__global__ __launch_bounds__(64) void glass_fit_kernel(const double *bc, uint32_t n, double lo, double hi,
                                                       uint32_t npts, double *out, int *status) {
    const uint32_t i = blockIdx.x * 64 + threadIdx.x;
    if (i >= n) return;
    double coeffs[CPD_GLASS_COEFFS], chisq = 0.0;
    const int r = fit_sellmeier1_givens(bc + 6 * i, bc + 6 * i + 3, lo, hi, npts, coeffs, &chisq);
    for (int j = 0; j < CPD_GLASS_COEFFS; ++j) out[7 * i + j] = r == 0 ? coeffs[j] : NAN;
    out[7 * i + 6] = r == 0 ? chisq : NAN;
    status[i] = r == 0 ? CPD_OK : CPD_ERR_INVALID_ARGUMENT;
}

}  // namespace
}  // namespace cpd

extern "C" cpd_status cpd_fit_sellmeier1_batch(cpd_ctx *ctx, const double *b, const double *c, size_t n, double lo,
                                               double hi, uint32_t npts, double *coeffs_out, double *chisq_out,
                                               cpd_status *per_glass) {
    using namespace cpd;
    if (n == 0) return CPD_OK;
    if (!b || !c || !coeffs_out || !per_glass) return fail(CPD_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!(lo < hi) || npts < CPD_GLASS_COEFFS || npts > 100000)
        return fail(CPD_ERR_INVALID_ARGUMENT, "need lo < hi and 6 <= npts <= 100000");
    if (n > (1u << 24)) return fail(CPD_ERR_INVALID_ARGUMENT, "more than 2^24 glasses");
    Context *cx;
    cpd_status st = resolve_context(ctx, &cx);
    if (st != CPD_OK) return st;
    std::lock_guard<std::mutex> lk(cx->mu);
    GF_TRY(hipSetDevice(cx->device));
    const size_t in_bytes = n * 6 * sizeof(double), out_bytes = n * 7 * sizeof(double);
    const size_t off_out = (in_bytes + 255) / 256 * 256, off_st = off_out + (out_bytes + 255) / 256 * 256;
    const size_t total = off_st + n * sizeof(int);
    if ((st = cx->fit_buf.reserve(total, false, cx->stream)) != CPD_OK) return st;
    if ((st = cx->host_stage.reserve(total)) != CPD_OK) return st;
    double *h = (double *)cx->host_stage.ptr;
    for (size_t i = 0; i < n; ++i) {
        for (int j = 0; j < 3; ++j) h[6 * i + j] = b[3 * i + j];
        for (int j = 0; j < 3; ++j) h[6 * i + 3 + j] = c[3 * i + j];
    }
    char *d = cx->fit_buf.as<char>();
    GF_TRY(hipMemcpyAsync(d, h, in_bytes, hipMemcpyHostToDevice, cx->stream));
    hipLaunchKernelGGL(glass_fit_kernel, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, cx->stream,
                       (const double *)d, (uint32_t)n, lo, hi, npts, (double *)(d + off_out), (int *)(d + off_st));
    GF_TRY(hipGetLastError());
    GF_TRY(hipMemcpyAsync(cx->host_stage.ptr, d + off_out, total - off_out, hipMemcpyDeviceToHost, cx->stream));
    GF_TRY(hipStreamSynchronize(cx->stream));
    const double *o = (const double *)cx->host_stage.ptr;
    const int *s = (const int *)((const char *)cx->host_stage.ptr + (off_st - off_out));
    for (size_t i = 0; i < n; ++i) {
        for (int j = 0; j < CPD_GLASS_COEFFS; ++j) coeffs_out[6 * i + j] = o[7 * i + j];
        if (chisq_out) chisq_out[i] = o[7 * i + 6];
        per_glass[i] = (cpd_status)s[i];
    }
    return CPD_OK;
}
