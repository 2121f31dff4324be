// eval.h -- host entry points of the evaluation kernels.
#pragma once

#include "problem.h"

namespace speo {

struct EvalLaunch {
    CompDev comp;
    const float *X;
    int ld, n, dim;
    float *f_out;   // [n]
    float *part;    // [nchunks][2][n] when nchunks > 1
    int nchunks;
    float lambda, bias;
    int raw;        // 1: write g (composition component), 0: lambda * g + bias
    int small;      // dim <= 128 with the padded layouts: register-resident-B path
};

// One component over n rows; `part` is needed when dim > 128.
int eval_component(const CompDev &comp, const float *X, int ld, int n, int dim, float *f_out,
                   float *part, float lambda, float bias, int raw, int precision,
                   int rows_per_block, hipStream_t s);

// Scratch (floats) eval_problem needs for n rows.
size_t eval_scratch_floats(int ncomp, int dim, int n);

// Whole problem (single component or composition) over n rows.
int eval_problem(const ProblemHost &P, const float *X, int ld, int n, float *f_out,
                 float *scratch, int precision, int rows_per_block, hipStream_t s);

// FP64 evaluation, one block per row.
int eval_problem_f64(const ProblemHost &P, const float *X, int ld, int n, double *f_out,
                     hipStream_t s);

int default_rows_per_block(int n, int dim, int precision = SPEO_PREC_FP32);

// k_eval_wide: the 16-bit evaluation of one component wider than 128 columns
// with k_eval<64, precision>'s bits (partial sums into part, finalized by the
// caller); applies when X, the shift and the 16-bit operand rows allow
// 16-byte loads.
bool eval_wide_applies(const CompDev &c, const float *X, int ld, int dim, int precision);
int eval_wide(const CompDev &c, const float *X, int ld, int n, int dim, float *part, int precision,
              hipStream_t s);

// Problems with objectives outside the fused FP32 epilogues (ProblemHost::
// generic): the FP64 generic kernel, FP32 and/or FP64 outputs; rows (device,
// optional) select rows[i] - row_base of X for launch row i.
int eval_problem_generic(const ProblemHost &P, const float *X, int ld, int n, float *f32,
                         double *f64, const int *rows, int row_base, hipStream_t s);

}  // namespace speo
