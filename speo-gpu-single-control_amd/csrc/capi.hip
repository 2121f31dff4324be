// capi.hip -- error state and device helpers of the C ABI.
#include <mutex>
#include <string>

#include "common.h"

namespace speo {

namespace {
thread_local std::string t_last_error;
std::mutex g_deferred_mu;
int g_deferred_code = SPEO_OK;
std::string g_deferred_msg;
}  // namespace

int set_error(int code, const std::string &msg) {
    t_last_error = msg;
    return code;
}

int check_hip(hipError_t err, const char *what) {
    if (err == hipSuccess) return SPEO_OK;
    return set_error(SPEO_ERR_HIP, std::string(what) + ": " + hipGetErrorString(err));
}

int check_launch(const char *kernel) {
    return check_hip(hipGetLastError(), kernel);
}

void defer_error(int code, const std::string &msg) {
    set_error(code, msg);
    std::lock_guard<std::mutex> lk(g_deferred_mu);
    if (g_deferred_code == SPEO_OK) {
        g_deferred_code = code;
        g_deferred_msg = msg;
    }
}

int take_deferred_error() {
    std::lock_guard<std::mutex> lk(g_deferred_mu);
    const int c = g_deferred_code;
    if (c != SPEO_OK) set_error(c, g_deferred_msg);
    g_deferred_code = SPEO_OK;
    g_deferred_msg.clear();
    return c;
}

hipStream_t as_stream(speo_stream_t s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace speo

using namespace speo;

extern "C" {

const char *speo_last_error(void) { return t_last_error.c_str(); }

int speo_device_check(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return set_error(SPEO_ERR_NO_DEVICE, "no HIP device visible");
    int dev = 0;
    SPEO_TRY_HIP(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SPEO_TRY_HIP(hipGetDeviceProperties(&prop, dev));
    if (std::string(prop.gcnArchName).rfind("gfx950", 0) != 0)
        return set_error(SPEO_ERR_NO_DEVICE, std::string("device is ") + prop.gcnArchName +
                                                 ", this library is built for gfx950");
    return SPEO_OK;
}

int SPEO_NextPow2Dim(int dim) { return next_pow2_dim(dim); }

// kernel_EA_CUDA.cu:64-69 + __cudaCheckError (:39-61): launch error, then a
// device synchronisation; here the status is returned instead of exit(-1).
int API_CheckCUDAError(void) {
    int rc = take_deferred_error();
    if (rc != SPEO_OK) return rc;
    SPEO_TRY(check_hip(hipGetLastError(), "API_CheckCUDAError (launch)"));
    return check_hip(hipDeviceSynchronize(), "API_CheckCUDAError (sync)");
}

}  // extern "C"
