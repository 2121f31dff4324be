// cec_generic.h -- the CEC2014 objectives outside the fused FP32 epilogues:
// Rosenbrock, Weierstrass, Griewank, Katsuura, Griewank-Rosenbrock, expanded
// Schaffer F6, the hybrid functions (shuffled z cut into segments of different
// objectives) and compositions over any of them (F4, F6, F7, F12, F15-F23,
// F26-F30 of CEC2014.cc:14-145; kernel_CEC2014_CUDA.cu:355-1671).
//
// One kernel per tile of 16 rows: z = M ((x - o) rate) for each component on
// v_mfma_f64_16x16x4f64 with FP64 accumulation into an LDS-resident Z tile,
// then each row's objective by one wave from that tile (segment objectives
// over the shuffled z for hybrids), then the composition weights.  FP64
// throughout: Weierstrass scales z + 1/2 by 3^20 and Katsuura z by 2^32
// before taking a cosine or a fractional part, so an FP32 z (relative error
// ~1e-7) would leave those terms with no correct digits.
#pragma once

#include "problem.h"

namespace speo {

constexpr int kGenMaxDim = 1024;  // Z tile of 16 rows in LDS
constexpr int kGenMaxSeg = 5;
constexpr int kGenRows = 16;

// One segment of a component: objective `obj` (SPEO_OBJ_*, 1..14) over
// `len` consecutive coordinates of the (shuffled) z, scaled by `rate`.
struct GenSeg {
    int obj;
    int len;
    double rate;
};

struct GenComp {
    int nseg;                  // 1: one objective over all coordinates
    GenSeg seg[kGenMaxSeg];
    const int *perm;           // hybrids: 1-based shuffle [dim] (the shuffle_data format), or null
    double rate;               // applied to x - o before the rotation (sr_func's sh_rate)
    int rotate;                // 0: z = (x - o) rate
    const double *o64;         // shift [dim] in FP64, or null: o32
    const float *o32;
    const double *M64;         // M [dim][dim] row-major (z_i = sum_j M_ij y_j), or null: B32
    const float *B32;          // B = M^T: B[k][c] = M[c][k], row stride ldb
    int ldb;
    double lambda, bias, sigma;  // fit = lambda g + bias; composition weight width
};

struct GenProblem {
    int dim, ncomp;
    const float *sigma32;      // per-component sigma on the device (overrides comp.sigma), or null
    GenComp comp[kMaxComp];
};

struct GenLaunch {
    GenProblem P;
    const float *X;            // rows of stride ld
    int ld, n;
    const int *rows;           // optional: launch row i reads row rows[i] - row_base of X
    int row_base;
    float *f32;                // outputs (either may be null)
    double *f64;
    double out_bias;           // added to every output
};

// SPEO_OBJ_HYBRID_1..6 -> the component's segments at `dim` (CEC2014.cc:954-1207):
// lengths ceil(p_i dim) in FP64 and the rest for the last segment.  Returns
// false for another objective.
bool hybrid_segments(int objective, int dim, GenComp &c);

// One component of CEC2014 function func_id (1..30) as the reference defines
// it (CEC2014.cc:14-145, 954-1371): objective, rate, rotation, lambda, bias,
// sigma; component count in *ncomp.  Returns false for an unknown function.
struct CecComp {
    int objective;   // SPEO_OBJ_* (hybrids: SPEO_OBJ_HYBRID_k)
    double rate;
    int rotate;
    double lambda, bias, sigma;
};
bool cec_function(int func_id, int *ncomp, CecComp comp[kMaxComp]);

// True when the problem needs this path (an objective outside 1..8, i.e.
// not one the fused FP32 epilogues take).
bool generic_objective(int objective);

// Evaluates n rows; the launch needs dim <= kGenMaxDim.
int eval_generic(const GenLaunch &p, hipStream_t s);

}  // namespace speo
