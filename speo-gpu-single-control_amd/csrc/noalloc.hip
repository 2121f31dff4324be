// noalloc.hip -- speo_evaluate_noalloc: speo_evaluate without any device
// allocation, for callers that cannot allocate (code under hipGraph capture,
// tight serving loops, harnesses that forbid hipMalloc).  The evaluation's
// workspace (per-chunk row sums above 128 columns, per-component values of a
// composition) lives in a fixed file-scope device array; each GPU holds its own
// copy until the process ends, and every call on that GPU shares it.
#include <mutex>

#include "eval.h"

namespace speo {

constexpr size_t kNoAllocScratchFloats = size_t(1) << 20;  // 4 MiB per GPU

static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
float *g_ptr[kMaxDevices] = {};  // the array's device address, per GPU

// The device address of `scratch` on the current GPU, looked up once per GPU
// with that GPU current (host code must not pass the array's name itself).
int scratch_ptr(float **out) {
    int dev = 0;
    SPEO_TRY_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return set_error(SPEO_ERR_ARG, "speo_evaluate_noalloc: device index");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        SPEO_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return set_error(SPEO_ERR_HIP, "speo_evaluate_noalloc: no address for the workspace");
        g_ptr[dev] = static_cast<float *>(p);
    }
    *out = g_ptr[dev];
    return SPEO_OK;
}
}  // namespace

}  // namespace speo

using namespace speo;

extern "C" int speo_evaluate_noalloc(speo_problem p, const float *X, int ld, int n, float *f_out,
                                     int precision, speo_stream_t stream) {
    if (!p || !X || !f_out || n < 0 || ld < p->h.dim)
        return set_error(SPEO_ERR_ARG, "speo_evaluate_noalloc: bad arguments");
    if (precision < SPEO_PREC_FP32 || precision > SPEO_PREC_BF16X6)
        return set_error(SPEO_ERR_ARG, "speo_evaluate_noalloc: unknown precision");
    const size_t need = p->h.generic ? 0 : eval_scratch_floats(p->h.ncomp, p->h.dim, n);
    if (need > kNoAllocScratchFloats)
        return set_error(SPEO_ERR_ARG, "speo_evaluate_noalloc: " + std::to_string(n) + " rows need " +
                                           std::to_string(need) + " workspace floats, above the " +
                                           std::to_string(kNoAllocScratchFloats) + " it holds");
    float *ws = nullptr;
    SPEO_TRY(scratch_ptr(&ws));
    return eval_problem(p->h, X, ld, n, f_out, ws, precision, default_rows_per_block(n, p->h.dim, precision),
                        as_stream(stream));
}
