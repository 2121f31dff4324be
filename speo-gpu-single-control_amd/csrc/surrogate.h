// surrogate.h -- the learned surrogate: an MLP D -> 256 -> 256 -> 1 with SiLU,
// BF16 MFMA operands with FP32 accumulation and FP32 master weights, trained
// online with Adam on a ring archive of true evaluations y = log(1 + f).
// No counterpart in the reference; new in this build.
#pragma once

#include "common.h"

namespace speo {

constexpr int kHidden = 256;
constexpr float kXScale = 0.01f;  // inputs in [-100, 100] -> [-1, 1]

// Flat parameter layout (floats): W1 [256][Kp], b1 [256], W2 [256][256],
// b2 [256], w3 [256], b3 [4 (1 used)].
constexpr int kMaxSplitK = 8;
constexpr int kFusedMaxKp = 128;  // widest input of the fused inference and training kernels
constexpr int kTrainRows = 32;    // minibatch rows per workgroup of the fused training step
constexpr int kTrainPart = 1024;  // floats of partial sums per such workgroup

struct MlpLayout {
    int D = 0, Kp = 0;
    size_t oW1 = 0, ob1 = 0, oW2 = 0, ob2 = 0, ow3 = 0, ob3 = 0, n = 0;
    void set(int dim) {
        D = dim;
        Kp = round_up(dim, 32);
        oW1 = 0;
        ob1 = oW1 + (size_t)kHidden * Kp;
        oW2 = ob1 + kHidden;
        ob2 = oW2 + (size_t)kHidden * kHidden;
        ow3 = ob2 + kHidden;
        ob3 = ow3 + kHidden;
        n = ob3 + 4;
    }
};

struct IslandDesc;
struct SurrRowDesc;

// Device buffers of one MLP with its training workspace.
struct Mlp {
    MlpLayout L;
    float *P = nullptr, *G = nullptr, *m = nullptr, *v = nullptr;  // [n]
    int max_batch = 0;
    float *A1 = nullptr, *H1 = nullptr, *A2 = nullptr, *H2 = nullptr;  // [max_batch][256]
    float *dH1 = nullptr, *dH2 = nullptr;                              // [max_batch][256]
    float *y = nullptr, *dy = nullptr;                                 // [max_batch]
    float *stats = nullptr;  // [0] loss, [1] target mean, [2] target std, [3] rho
    uint16_t *W1b = nullptr, *W2b = nullptr;  // BF16 copies (MFMA fragment order) for inference
    uint16_t *W2Tb = nullptr;                 // BF16 W2 transposed (backward of the fused step)
    bool bf16_stale = true;                   // the BF16 copies lag the FP32 weights
    float *part = nullptr;                    // split-K partial products of the weight gradients
    float *lterm = nullptr;                   // [max_batch] per-row loss terms
    // Fused training step (Kp <= 128): activations transposed, [rows][Bp]
    // BF16 in MFMA fragment order with Bp = round_up(max_batch, 32), and
    // per-row-tile partial sums of the bias and output-layer gradients,
    // [tiles][1024] floats.
    uint16_t *XT = nullptr, *H1T = nullptr, *dH2T = nullptr, *dH1T = nullptr;
    float *tpart = nullptr;
    int Bp = 0;
    int *t_dev = nullptr;  // Adam steps taken (device), read by the update kernels
    ~Mlp();
    int alloc(int dim, int batch);
    int init(uint64_t seed, hipStream_t s);
    // y = MLP(X) for n rows of X (stride ld >= D, raw inputs), batched through
    // the workspace; pred receives n values.
    int forward(const float *X, int ld, int n, float *pred, hipStream_t s);
    // Loss (0.5 mean (y - t)^2) and gradients into G for n <= max_batch rows
    // (the fused step when Kp <= 128, else GEMMs and reductions layer by layer).
    int loss_grad(const float *X, int ld, int n, const float *t, hipStream_t s,
                  bool want_loss = true);
    int loss_grad_unfused(const float *X, int ld, int n, const float *t, hipStream_t s,
                          bool want_loss);
    int adam(float lr, hipStream_t s);
    // One Adam step on a minibatch of the archive (rows AX [slot][Kp], labels
    // Ay): the slots drawn as gather_batch draws them, loss and gradients,
    // Adam, and fresh BF16 copies.  Two kernels when Kp <= 128.
    int train_step(const float *AX, const float *Ay, int n_valid, uint64_t seed, uint32_t round,
                   uint32_t step, int batch, float lr, bool want_loss, float *Xb, float *tb,
                   hipStream_t s);
    // Refreshes the BF16 weight copies from the FP32 master weights.
    int refresh_bf16(hipStream_t s);
    // Fused inference (one kernel): pred[i] = MLP(X[i]); uses the BF16 copies.
    int infer(const float *X, int ld, int n, float *pred, hipStream_t s);
};

// One target row in surrogate mode, rows ordered problem by problem.
struct SurrRowDesc {
    int gr;       // global row
    int island;   // island index
    int problem;  // problem index
    int li;       // row within its island
};

constexpr int kMaxSurrProblems = 16;
struct SurrHeads {
    int qoff[kMaxSurrProblems];  // first q of each problem
    int rows[kMaxSurrProblems];  // targets of each problem
};

struct IslandDesc;
int gen_candidates(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C, const float *X,
                   float lo, float hi, uint64_t seed, const uint32_t *gen_base,
                   uint32_t gen_offset, float *cand, int Kp, hipStream_t s);
int pick_candidates(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C,
                    const float *cand, int Kp, const float *pred, const float *rho,
                    float rho_min, float *U, hipStream_t s);
// Per-problem networks of one fused screening launch.
struct CandInferBatch {
    const uint16_t *W1b[kMaxSurrProblems], *W2b[kMaxSurrProblems];
    const float *P[kMaxSurrProblems];
    int qoff[kMaxSurrProblems], nq[kMaxSurrProblems], first[kMaxSurrProblems];
    int prob[kMaxSurrProblems];  // problem index of network k (rho, scored)
    int nnets;
    size_t ob1, ob2, ow3, ob3;
    const float *rho;  // held-out rho per problem (device)
    float rho_min;
    long long *scored;  // candidates scored per problem (device, accumulated)
    int tiles;          // 64-candidate tiles of all problems
};
// gen_candidates for every problem's targets [qoff[k], qoff[k] + nq[k]) and
// each problem's network on them, in one kernel that never stores the
// candidates: pred[q C + c], Kp <= 128.  A problem whose rho (device) is not
// at least rho_min is skipped on the device (its pick is candidate 0 without
// scores); scored[problem] counts the candidates scored.
int cand_infer(Mlp *const *nets, const int *qoff, const int *nq, const int *prob, int nn,
               const IslandDesc *isl, const SurrRowDesc *rows, int C, const float *X, float lo,
               float hi, uint64_t seed, const uint32_t *gen_base, uint32_t gen_offset,
               const float *rho, float rho_min, long long *scored, float *pred, hipStream_t s);
// pick_candidates with the chosen candidate generated again into U.
int pick_regen(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C, const float *pred,
               const float *rho, float rho_min, const float *X, float lo, float hi, uint64_t seed,
               const uint32_t *gen_base, uint32_t gen_offset, float *U, hipStream_t s);
// Appends every target's trial and log(1 + f) at its problem's next archive
// slots (head_dev[p] on the device), then advances head_dev[p] by the
// problem's rows modulo cap.
int archive_append(const IslandDesc *isl, const SurrRowDesc *rows, int Q, const float *U,
                   const float *fU, float *AX, float *Ay, int cap, int Kp, SurrHeads heads,
                   int nproblems, int *head_dev, hipStream_t s);
// *p += v (one thread; also inside captured graphs).
int add_i32(int *p, int v, hipStream_t s);
int gather_batch(const float *AX, const float *Ay, int n_valid, int Kp, uint64_t seed,
                 uint32_t round, uint32_t step, int batch, float *Xb, float *tb, hipStream_t s);
int gather_holdout(const float *AX, const float *Ay, int H, int Kp, float *Xh, float *yh,
                   hipStream_t s);
int mean_std(const float *y, int n, float *stats, hipStream_t s);
// One training step (Mlp::train_step's, Kp <= 128) of nn networks of equal
// width on their own archives, in two launches per eight networks.
// The minibatch draw uses the device round counter *round_dev; Adam's step
// count is each network's *t_dev + step + 1 (the caller advances t_dev after
// the round).
int train_steps(Mlp *const *nets, const float *const *AX, const float *const *Ay,
                const int *n_valid, const uint64_t *seed, int nn, const uint32_t *round_dev,
                uint32_t step, int batch, float lr, bool want_loss, hipStream_t s);
// Spearman's rho of n <= 4096 values; ranks: workspace of 4096 float2.
int spearman(const float *a, const float *b, int n, float *rho_out, float2 *ranks, hipStream_t s);

}  // namespace speo
