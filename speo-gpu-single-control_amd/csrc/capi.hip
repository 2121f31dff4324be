// capi.hip -- error state and device helpers of the C ABI.
#include <cinttypes>
#include <cstdio>
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

// ---- run driver and result record (IslandEA, src/island_EA.cc)

int64_t speo_generations_for_fes(int64_t max_FEs, int64_t current, int64_t per_generation) {
    if (per_generation <= 0 || current < 0 || max_FEs <= current) return 0;
    return (max_FEs - current + per_generation - 1) / per_generation;
}

uint64_t speo_run_seed(int node_ID, int function_ID, int run_ID) {
    return (uint64_t)((int64_t)(node_ID + function_ID) * (int64_t)(run_ID + function_ID));
}

// The numbers as a C++ stream writes a double by default (six significant
// digits, the shorter of fixed and scientific notation): "%g".
int speo_de_parameters(double CR, double F, int strategy_ID, char *buf, int buflen) {
    static const char *const kNames[9] = {"current/1/bin",         "current/2/bin", "current-to-best/1/bin",
                                          "current-to-best/2/bin", "rand/1/bin",    "rand/2/bin",
                                          "best/1/bin",            "best/2/bin",    "current_to_rand/1/bin"};
    if (strategy_ID < 0 || strategy_ID > 8 || !buf || buflen < 1)
        return set_error(SPEO_ERR_ARG, "speo_de_parameters: strategy 0..8 and a buffer are required");
    const int n = std::snprintf(buf, (size_t)buflen, "CR/F=%g/%g_%s", CR, F, kNames[strategy_ID]);
    if (n < 0 || n >= buflen) return set_error(SPEO_ERR_ARG, "speo_de_parameters: buffer too short");
    return n;
}

const char *speo_result_header(void) {
    return "function_ID,run_ID,dim,best_fitness,time_period,communication_percentage,total_FEs,"
           "node_num,GPU_num,island_num,total_pop_size,subisland_size,interval,migration_size,"
           "migration_rate,buffer_capacity,migration_topology,buffer_manage,regroup_option,"
           "configure_EA,EA_parameters";
}

int speo_result_record(const speo_run_record *r, char *buf, int buflen) {
    if (!r || !buf || buflen < 1 || !r->migration_topology || !r->buffer_manage || !r->regroup_option ||
        !r->configure_EA || !r->EA_parameters)
        return set_error(SPEO_ERR_ARG, "speo_result_record: null record, field or buffer");
    const int n = std::snprintf(buf, (size_t)buflen,
                                "%d,%d,%d,%g,%g,%g,%" PRId64 ",%d,%d,%d,%d,%d,%d,%d,%g,%d,%s,%s,%s,%s,%s",
                                r->function_ID, r->run_ID, r->dim, r->best_fitness, r->time_period,
                                r->communication_percentage, r->total_FEs, r->node_num, r->GPU_num,
                                r->island_num, r->total_pop_size, r->subisland_size, r->interval,
                                r->migration_size, r->migration_rate, r->buffer_capacity,
                                r->migration_topology, r->buffer_manage, r->regroup_option, r->configure_EA,
                                r->EA_parameters);
    if (n < 0 || n >= buflen) return set_error(SPEO_ERR_ARG, "speo_result_record: buffer too short");
    return n;
}

}  // extern "C"
