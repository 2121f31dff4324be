// common.h -- shared device/host helpers of the MI355X DE engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>

#include "speo_de.h"

namespace speo {

// ------------------------------------------------------------ error state

// Records `msg` as the calling thread's last error and returns `code`.
int set_error(int code, const std::string &msg);
// Records a HIP failure (if any) and returns SPEO_ERR_HIP, else SPEO_OK.
int check_hip(hipError_t err, const char *what);
// Checks the launch that just happened on this thread.
int check_launch(const char *kernel);
// First error recorded by a void entry point (API_evaluateFitness), cleared on read.
int take_deferred_error();
void defer_error(int code, const std::string &msg);

#define SPEO_TRY_HIP(expr)                                        \
    do {                                                          \
        hipError_t e_ = (expr);                                   \
        if (e_ != hipSuccess) return ::speo::check_hip(e_, #expr); \
    } while (0)

#define SPEO_TRY(expr)                 \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != SPEO_OK) return rc_; \
    } while (0)

// ------------------------------------------------------------- constants

constexpr int kWave = 64;     // CDNA wavefront
constexpr int kMaxR = 5;      // donor draws per individual (kernel_DE_CUDA.cu:6)
constexpr int kMaxDraws = 4096;

enum Purpose : uint32_t {
    kPurposeDonor = 0,  // donor indices and j_rand of (row, generation)
    kPurposeCross = 1,  // crossover draw j of (row, generation)
    kPurposeInit = 2,   // init coordinate j of row
    kPurposeShift = 16, // problem shift vector
    kPurposeGauss = 17, // problem N(0,1) matrix
    kPurposeProblem = 18,
    kPurposeRegroup = 19,      // host draws of sub-island regroups and DE configuration
    kPurposeSurrogate = 32,    // MLP init
    kPurposeSurrBatch = 33,    // training minibatch rows
    kPurposeCandidate = 64     // candidate c >= 1: donors 64 + 2c, crossover 65 + 2c
};

// --------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3"
// (SC'11): ten rounds of the 4x32 Philox bijection with Weyl key schedule.

struct U4 {
    uint32_t x, y, z, w;
};

__host__ __device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c.x;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c.z;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// u32 -> float in (0,1]: ((x >> 8) + 1) * 2^-24, exact in FP32.
__host__ __device__ __forceinline__ float u01(uint32_t x) {
    return (float)((x >> 8) + 1u) * 0x1p-24f;
}

// The crossover test u01(x) < CR as an integer compare: with n = x >> 8,
// u01(x) = (n + 1) 2^-24 exactly, and for an integer m and real t,
// m < t <=> m < ceil(t); so u01(x) < CR <=> n < ceil(CR 2^24) - 1.
// NaN or CR <= 0 never takes (threshold 0), CR >= 1 + 2^-24 always does.
__host__ __device__ inline uint32_t cr_threshold(float CR) {
    if (!(CR > 0.0f)) return 0u;
    const double t = ceil((double)CR * 16777216.0) - 1.0;
    return t >= 16777216.0 ? 16777216u : (uint32_t)t;
}
__host__ __device__ __forceinline__ bool cross_take(uint32_t x, uint32_t threshold) {
    return (x >> 8) < threshold;
}

__host__ __device__ __forceinline__ U4 philox_block(uint64_t seed, uint32_t block,
                                                    uint32_t a, uint32_t b, uint32_t purpose) {
    return philox4x32_10(U4{block, a, b, purpose}, (uint32_t)seed, (uint32_t)(seed >> 32));
}

__host__ __device__ __forceinline__ uint32_t word(const U4 &v, uint32_t i) {
    return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}

// Reflective bound (kernel_DE_CUDA.cu:160-170), capped at 64 reflections then
// clamped; NaN passes through.  The reflection loop sits behind the in-bound
// test so that a wavefront whose lanes are all in bounds skips it.
__host__ __device__ __forceinline__ float reflect_bound(float v, float lo, float hi) {
    int it = 0;
#pragma clang loop unroll(disable)
    do {
        if (v < lo) v = lo + (lo - v);
        if (v > hi) v = hi - (v - hi);
        if (++it >= 64) {
            v = v < lo ? lo : (v > hi ? hi : v);
            break;
        }
    } while ((v < lo) || (v > hi));
    return v;
}

__host__ __device__ __forceinline__ float check_bound(float v, float lo, float hi) {
    if (!((v < lo) || (v > hi))) return v;
    return reflect_bound(v, lo, hi);
}

inline int next_pow2_dim(int dim) {
    int p = 1;
    while (p < dim) p <<= 1;
    return p < 32 ? 32 : p;
}

__host__ __device__ inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

hipStream_t as_stream(speo_stream_t s);

}  // namespace speo
