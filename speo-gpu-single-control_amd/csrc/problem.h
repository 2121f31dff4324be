// problem.h -- device-resident shifted-rotated problem (objective host wrapper).
// Counterpart of the reference's CEC2014_CUDA (CUDA_EA_library/CEC2014_CUDA.h:
// 38-82): owns the shift vectors and rotation matrices on the device, in every
// layout the kernels use, instead of loading CEC text files.
#pragma once

#include <vector>

#include "eval_core.h"

namespace speo {

constexpr int kMaxComp = 8;

// Column weights w_j of the weighted-sphere objectives, sum_j w_j z_j^2:
// elliptic 10^(6 j / (D - 1)), bent cigar (1, 10^6, 10^6, ...), discus
// (10^6, 1, 1, ...).  The FP32 kernels evaluate all three through one path.
bool weighted_sphere(int objective);
// The column that outweighs all others together (index), or -1; *w_out its
// FP64 weight.  column_weights gives it weight 0: the kernels add its term
// in FP64 (CompDev::heavy).
int heavy_column(int objective, int dim, double *w_out);
void column_weights(int objective, int dim, float *w);

struct CompHost {
    int objective = SPEO_OBJ_RASTRIGIN;
    float rate = 1.0f;     // FP32 kernels
    double rate64 = 1.0;   // FP64 verifier (a rate such as 0.0512 is not a float)
    double sigma = 10.0, bias = 0.0, lambda = 1.0;
    float *shift32 = nullptr;   // [dim]
    double *shift64 = nullptr;  // [dim]
    float *B32 = nullptr;       // [dim][ldb32]: B[k][c] = M[c][k]
    uint16_t *Mbf[3] = {nullptr, nullptr, nullptr};  // BF16 parts of M, [dim][ldm16]
    uint16_t *Mh[3] = {nullptr, nullptr, nullptr};   // FP16 parts of 2^14 M
    double *M64 = nullptr;      // [dim][dim]
    // dim <= 128 only: zero-padded layouts of the register-resident-B path
    float *B32s = nullptr;                                   // 128 x 128, fragment order
    uint16_t *Mbfs[3] = {nullptr, nullptr, nullptr};         // [128][round_up(dim,32)]
    uint16_t *Mhs[3] = {nullptr, nullptr, nullptr};
    float *colw = nullptr;      // [dim] bent cigar / discus weights (elliptic: ProblemHost::colw)
    bool unrotated = false;     // M = I: evaluated in FP32 whatever the requested precision,
                                // where the identity GEMM is exact
};

// Per-component parameters of the FP64 path and of the composition combine,
// stored on the device as an array.
struct CompParams {
    int objective;
    double rate;
    double sigma, bias, lambda;
    const float *shift32;
    const double *shift64;
    const double *M64;
};

struct GenProblem;

struct ProblemHost {
    int dim = 0, ncomp = 0;
    int ldb32 = 0, ldm16 = 0;
    float *colw = nullptr;          // [dim] elliptic weights
    CompParams *params_dev = nullptr;  // [ncomp]
    CompHost comp[kMaxComp];
    CompDev view(int c, int precision) const;
    // Objectives outside the fused FP32 epilogues (Rosenbrock, Weierstrass,
    // Griewank, Katsuura, Griewank-Rosenbrock, Schaffer F6, hybrids): the
    // whole problem is evaluated by the FP64 generic kernel (cec_generic.h).
    bool generic = false;
    int *perm[kMaxComp] = {};       // hybrids: 1-based shuffle [dim] on the device
    GenProblem *gen = nullptr;      // host description with device pointers
};

}  // namespace speo

struct speo_problem_s {
    speo::ProblemHost h;
    float *scratch = nullptr;  // eval_problem workspace of speo_evaluate
    size_t scratch_cap = 0;
};
