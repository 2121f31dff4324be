// engine.h -- device tables shared by the island engine's kernels.
#pragma once

#include "de.h"
#include "problem.h"

namespace speo {

// One problem as the generation kernels see it.
struct FusedProblem {
    int ncomp, dim;
    int generic;  // evaluated by the generic kernel (cec_generic.h), skipped by k_generation
    float lambda[kMaxComp], bias[kMaxComp], sigma[kMaxComp];
    CompDev comp[kMaxComp];
};

// ------------------------------------------------------------ epoch kernel
//
// k_island_epoch: a workgroup runs a whole epoch (the generations between
// two migrations) of one island.
//   LDS plan: the island's current rows and its trial rows live in LDS; an
//     accepted trial is copied over its row.
// k_island_split: an island split over `parts` workgroups, one generation
//   per launch: every workgroup loads all rows of its island from the input
//   population, generates, evaluates and selects its share and writes the
//   share's rows to the output population (the next launch's input), so the
//   kernel boundary is the only synchronisation between the shares.
//   HBM plan (one workgroup per island, islands too large for LDS): every row
//     has two slots in HBM (slot 0 = X, slot 1 = a scratch population), an
//     accepted trial flips the row's slot bit, and the trials of a pass of
//     rows are staged in LDS for the evaluation.

constexpr int kEpochThreads = 512;
constexpr int kEpochLdsMax = 160 * 1024;
constexpr int kEpochMaxParts = 8;

struct EpochPlan {
    int S;          // row stride (floats) of population rows in LDS
    int ST;         // row stride (floats) of trial rows in LDS
    int in_lds;     // 1: LDS plan; 0: HBM plan
    int pass_rows;  // rows per generate/evaluate pass (multiple of 16)
    int parts;      // workgroups sharing the island (LDS plan only)
    int own_rows;   // rows per share (multiple of 16; the last share may hold fewer)
    int lds_bytes;  // LDS the island needs
};

// Everything a workgroup needs to know about its island, in one record.
struct EpochBlock {
    IslandDesc I;
    EpochPlan pl;
    int share;                 // this workgroup's share of the island's rows
    int objective, kp32;
    uint32_t inv_ng;           // ceil(2^32 / ng), ng = ceil(dim / 4) (0 when ng == 1)
    uint32_t crt;              // cr_threshold(CR)
    float rate, lambda, bias;
    const float *shift, *B32s, *colw;
    const uint16_t *M16s[3];
    HeavyCol heavy;            // heavy.j < 0: none
};

struct EpochArgs {
    const EpochBlock *blocks;  // [workgroups]
    const float *Xin, *fXin;   // population and fitness at the start of the launch
    float *Xout, *fXout;       // at the end (the same buffers when parts == 1)
    float *Xalt;               // HBM slot 1 of islands on the HBM plan
    float *U, *fU;             // the last generation's trials and their fitness (out)
    float lo, hi;
    uint64_t seed;
    const uint32_t *gen_base;  // generation counter; the first generation is *gen_base + gen_offset
    uint32_t gen_offset;
    int L;                     // generations in this launch (1 when parts > 1)
    const int4 *donors;        // [generation][row] donors and j_rand (k_epoch_donors)
    int rows;                  // rows of all islands
    int donor_gen;             // generation g of the launch reads donors[donor_gen + g]
    int write_trials;          // write U / fU (the chunk's last generation)
};

// Plan of one island (dim <= 128, single component) split over `parts`
// workgroups, for a precision with `parts16` 16-bit operand parts (0: FP32);
// lds_bytes > kEpochLdsMax means it does not fit.
EpochPlan epoch_plan(int np, int dim, int parts, int parts16);
// 16-bit operand parts of a precision (0 for FP32).
int precision_parts(int precision);

// split: islands split over workgroups (LDS plan, a.L == 1).
int launch_epoch(int precision, const EpochArgs &a, int nblocks, int lds_bytes, bool split, hipStream_t s);

// Donors and j_rand of generations *gen_base + gen_offset + [0, L) of every
// row into out[g * rows + row]; row_island maps a row to its island.
int epoch_donors(const IslandDesc *isl, const int *row_island, int rows, int L, uint64_t seed,
                 const uint32_t *gen_base, uint32_t gen_offset, int4 *out, hipStream_t s);

}  // namespace speo
