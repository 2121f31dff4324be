// engine.h -- device tables shared by the island engine's kernels.
#pragma once

#include "de.h"
#include "problem.h"

namespace speo {

// One problem as the generation kernels see it.
struct FusedProblem {
    int ncomp, dim;
    float lambda[kMaxComp], bias[kMaxComp], sigma[kMaxComp];
    CompDev comp[kMaxComp];
};

// ------------------------------------------------------------ epoch kernel
//
// One workgroup per island runs a whole epoch (the generations between two
// migrations).  Every row has two slots, the current individual and the
// slot the next trial is written to; an accepted trial flips the row's slot
// bit instead of being copied.  The slots live in LDS when the island fits
// ("LDS plan"), else in HBM (slot 0 = X, slot 1 = a scratch population) with
// the trials of a pass of rows staged in LDS for the evaluation.

constexpr int kEpochThreads = 512;
constexpr int kEpochLdsMax = 160 * 1024;

struct EpochPlan {
    int S;          // row stride (floats) of rows in LDS
    int in_lds;     // 1: both slots in LDS; 0: slots in HBM, trials staged per pass
    int pass_rows;  // rows per generate/evaluate pass (multiple of 16)
    int gens_ahead; // generations whose donors are drawn together
    int lds_bytes;  // LDS the island needs
};

struct EpochArgs {
    const IslandDesc *isl;
    const EpochPlan *plan;
    const int *order;          // block -> island
    const FusedProblem *probs;
    float *X, *fX;             // population and fitness (in, out)
    float *Xalt;               // HBM slot 1 of islands without the LDS plan
    float *U, *fU;             // the last generation's trials and their fitness (out)
    float lo, hi;
    uint64_t seed;
    const uint32_t *gen_base;  // generation of the epoch's first generation
    int L;                     // generations in this launch
};

// Plan of one island (dim <= 128, single component); lds_bytes > kEpochLdsMax
// means the island cannot run in the epoch kernel.
EpochPlan epoch_plan(int np, int dim);

int launch_epoch(int precision, const EpochArgs &a, int nblocks, int lds_bytes, hipStream_t s);

}  // namespace speo
