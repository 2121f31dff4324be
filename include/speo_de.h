/*
 * speo_de.h -- C ABI of the MI355X island-model DE/rand/1/bin engine.
 *
 * Two layers, both plain C (raw pointers, sizes, POD structs by value):
 *
 *  1. Drop-in entry points with the reference's names and roles.  The
 *     reference's host classes EA_CUDA / DE_CUDA / CEC2014_CUDA call these
 *     (CUDA_EA_library/EA_CUDA.cc:41-43, DE_CUDA.cc:210-212,
 *     CEC2014_CUDA.cc:292-298).  They keep the reference's population layout
 *     (row stride = next power of two of dim, at least 32; EA_CUDA.cc:197-206)
 *     and take the reference built with GPU_SINGLE_PRECISION (real = float,
 *     config.h:35-37).  The one signature change: cuRAND's device state array
 *     becomes a host-side counter-based Philox4x32-10 state (SPEO_RandState).
 *     Errors are returned (or recorded for the void API_evaluateFitness), never
 *     exit(); API_CheckCUDAError() synchronises and reports them.
 *
 *  2. The native engine: islands of uneven size and dimension held in one
 *     GPU's HBM, a fused select/mutate/crossover/evaluate generation kernel,
 *     ring migration in HBM, FP64 re-verification of each island's best,
 *     split-precision rotation, hipGraph capture and an MLP surrogate.
 *     Every call takes a HIP stream (NULL = the default stream) and is
 *     stream-ordered; nothing synchronises unless its name says so.
 *
 * Status codes: 0 = success, negative = SPEO_ERR_*; speo_last_error() returns
 * the message of the last failure on the calling thread.
 */
#ifndef SPEO_DE_H_
#define SPEO_DE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPEO_EXPORT __attribute__((visibility("default")))

typedef struct ihipStream_t *speo_stream_t; /* == hipStream_t */

enum {
    SPEO_OK = 0,
    SPEO_ERR_ARG = -1,         /* invalid argument (size, pointer, layout)   */
    SPEO_ERR_HIP = -2,         /* a HIP runtime call or kernel launch failed */
    SPEO_ERR_UNSUPPORTED = -3, /* valid in the reference, not built here      */
    SPEO_ERR_NO_DEVICE = -4,   /* no usable gfx950 device                     */
    SPEO_ERR_ALLOC = -5        /* device or host allocation failed            */
};

/* ==================================================================== *
 *  1. Drop-in layer (reference names and struct layouts, real = float)  *
 * ==================================================================== */

typedef float real;

/* CUDA_EA_library/EA_CUDA.h:21-25 */
typedef struct {
    real *elements;      /* [island_size][next_pow2(dim)], padding = 0 */
    real *fitness_value; /* [island_size] */
} GPU_Population;

/* CUDA_EA_library/EA_CUDA.h:26-31 */
typedef struct {
    int island_size;
    int subisland_num;
    int *permutated_index; /* device, [island_size] */
} GPU_IslandInfo;

/* CUDA_EA_library/EA_CUDA.h:33-39 */
typedef struct {
    real CR;
    real F;
    int strategy_ID; /* 0..8, CUDA_EA_library/DE_CUDA.cc:36-53; 4 = rand/1/bin */
    int pop_size;
} DEInfo;

/* CUDA_EA_library/config.h:48-59 */
typedef struct {
    int dim;
    int function_ID;
    int run_ID;
    int max_base_FEs;
    int seed;
    int running_time;
    int computing_time;
    real max_bound;
    real min_bound;
} ProblemInfo;

/* Replaces `curandState *d_rand_states` (EA_CUDA.h:61).  Counter-based:
 * draw k of individual i in generation g is Philox4x32-10 with key = seed
 * and counter = {k/4, row_base + i, g, purpose}.  The library advances
 * `generation` once per API_DE_GenerateNewPopulation call. */
typedef struct {
    uint64_t seed;
    uint32_t generation;
    uint32_t row_base;
} SPEO_RandState;

/* Row stride of the drop-in layout: next power of two >= dim, at least 32
 * (CUDA_EA_library/EA_CUDA.cc:197-206). */
SPEO_EXPORT int SPEO_NextPow2Dim(int dim);

/* Replaces API_Initilize (CUDA_EA_library/EA_CUDA.h:41-42,
 * kernel_EA_CUDA.cu:121-134): seeds *rand_state from problem_info.seed and
 * fills d_population.elements with min + u*(max-min), u in (0,1]. */
SPEO_EXPORT int API_Initilize(GPU_Population d_population, SPEO_RandState *rand_state,
                              GPU_IslandInfo GPU_island_info, ProblemInfo problem_info);

/* Replaces API_CheckCUDAError (CUDA_EA_library/EA_CUDA.h:43-44,
 * kernel_EA_CUDA.cu:64-69): synchronises the device and returns the first
 * recorded or pending error (0 if none), instead of exit(-1). */
SPEO_EXPORT int API_CheckCUDAError(void);

/* Replaces API_DE_GenerateNewPopulation (CUDA_EA_library/DE_CUDA.cc:2-3,
 * kernel_DE_CUDA.cu:291-302): one DE generation of d_population into
 * d_candidate.  d_best_individual_ID (device, [subisland_num]) is filled with
 * each sub-island's best slot (used by strategies 2, 3, 6, 7).  d_DE_info is
 * a device array of subisland_num entries.  Every individual is processed,
 * also when island_size is not a multiple of 16. */
SPEO_EXPORT int API_DE_GenerateNewPopulation(GPU_Population d_candidate,
                                             GPU_Population d_population,
                                             int *d_best_individual_ID,
                                             GPU_IslandInfo d_island_info, DEInfo *d_DE_info,
                                             SPEO_RandState *rand_state,
                                             ProblemInfo problem_info);

/* Replaces API_DE_SelectSurvival (CUDA_EA_library/DE_CUDA.cc:4-5,
 * kernel_DE_CUDA.cu:329-338): slot k takes the candidate when its fitness is
 * strictly smaller. */
SPEO_EXPORT int API_DE_SelectSurvival(GPU_Population d_population, GPU_Population d_candidate,
                                      GPU_IslandInfo GPU_island_info, ProblemInfo problem_info);

/* Replaces API_evaluateFitness (CUDA_EA_library/CEC2014_CUDA.h:35-37,
 * kernel_CEC2014_CUDA.cu:1672-1797) for every function_ID 1..30, in the
 * reference's layouts: d_original [pop_size][next_pow2(dim)], d_shift
 * [ncomp][next_pow2(dim)], d_rotated_data = M transposed and row-padded,
 * [ncomp][next_pow2(dim)][dim] (CEC2014_CUDA.cc:255), d_shuffle_data the
 * 1-based shuffle of the hybrids (padded to next_pow2(dim) per component),
 * d_sigma the compositions' widths (read).  IDs 1, 2, 3, 5, 8, 9, 10, 11, 13
 * and 14 run as one fused FP32 MFMA kernel (above 128 columns its per-chunk
 * row sums use d_rotated_elements as workspace); IDs 4, 6, 7, 12 and 15-30
 * take the FP64 generic kernel (dim <= 1024; values rounded to real).  As in
 * the reference, IDs 1-16 ignore `bias` and IDs 17-30 add it.  The other
 * scratch arguments (d_shuffled_elements, d_tmp_fitness, d_wi) are accepted
 * and unused.  IDs outside 1..30 record SPEO_ERR_UNSUPPORTED, reported by
 * API_CheckCUDAError(). */
SPEO_EXPORT void API_evaluateFitness(real *d_fitness_value, real *d_shuffled_elements,
                                     real *d_rotated_elements, real *d_original_elements,
                                     real *d_shift_data, real *d_rotated_data,
                                     int *d_shuffle_data, real *d_tmp_fitness, real *d_wi,
                                     real *d_sigma, int pop_size, int problem_dim, real bias,
                                     int function_ID);

/* ==================================================================== *
 *  2. Native engine                                                     *
 * ==================================================================== */

SPEO_EXPORT const char *speo_last_error(void);
SPEO_EXPORT int speo_device_check(void); /* SPEO_OK if a gfx950 device is current */

/* ---- objectives ---------------------------------------------------- */

/* g(z) of one component, z = rate * M (x - o):
 *   ELLIPTIC    sum_j 10^(6 j / (D - 1)) z_j^2          CEC2014.cc:481-497
 *   RASTRIGIN   sum_j z_j^2 - 10 cos(2 pi z_j) + 10     CEC2014.cc:668-683
 *   ACKLEY      -20 e^(-0.2 sqrt(|z|^2 / D)) - e^(mean cos(2 pi z_j)) + 20 + e
 *                                                       CEC2014.cc:590-612
 *   BENT_CIGAR  z_0^2 + 10^6 sum_{j >= 1} z_j^2         CEC2014.cc:499-516
 *   DISCUS      10^6 z_0^2 + sum_{j >= 1} z_j^2         CEC2014.cc:518-530
 *   HAPPYCAT    |r2 - D|^(1/4) + (r2 / 2 + s) / D + 1/2   CEC2014.cc:907-929
 *   HGBAT       |r2^2 - s^2|^(1/2) + (r2 / 2 + s) / D + 1/2   CEC2014.cc:931-953
 *               with u_j = z_j - 1, r2 = sum_j u_j^2, s = sum_j u_j.
 *   SCHWEFEL    418.9828872724338 D - sum_j g(z_j + 420.9687462275036), the
 *               modified Schwefel function folded beyond |v| = 500
 *                                                       CEC2014.cc:705-738
 * SPEO_MODE_EPOCH takes objectives 1-5. */
enum {
    SPEO_OBJ_ELLIPTIC = 1,
    SPEO_OBJ_RASTRIGIN = 2,
    SPEO_OBJ_ACKLEY = 3,
    SPEO_OBJ_BENT_CIGAR = 4,
    SPEO_OBJ_DISCUS = 5,
    SPEO_OBJ_HAPPYCAT = 6,
    SPEO_OBJ_HGBAT = 7,
    SPEO_OBJ_SCHWEFEL = 8,
    /* Evaluated with an FP64 rotation and FP64 terms (any precision setting;
     * SPEO_MODE_PLAIN only; dim <= 1024):                                     */
    SPEO_OBJ_ROSENBROCK = 9,   /* y = z + 1: sum 100 (y_i^2 - y_{i+1})^2 + (y_i - 1)^2
                                                           CEC2014.cc:552-568 */
    SPEO_OBJ_WEIERSTRASS = 10, /* sum_i sum_k 0.5^k cos(2 pi 3^k (z_i + 1/2)) - D c,
                                  k = 0..20                CEC2014.cc:615-644 */
    SPEO_OBJ_GRIEWANK = 11,    /* 1 + |z|^2 / 4000 - prod cos(z_i / sqrt(i + 1))
                                                           CEC2014.cc:646-666 */
    SPEO_OBJ_KATSUURA = 12,    /* 10/D^2 prod (1 + (i+1) sum_j |2^j z_i - round(2^j z_i)| 2^-j)^(10/D^1.2)
                                  - 10/D^2                 CEC2014.cc:739-765 */
    SPEO_OBJ_GRIE_ROSEN = 13,  /* Griewank of Rosenbrock pairs, y = z + 1, wrap-around
                                                           CEC2014.cc:842-868 */
    SPEO_OBJ_ESCAFFER6 = 14,   /* expanded Schaffer F6, wrap-around pairs
                                                           CEC2014.cc:870-905 */
    /* Hybrid functions hf01-hf06 (CEC2014.cc:954-1207): z = M (x - o), shuffled
     * by the component's permutation, cut into segments of ceil(p_i D)
     * coordinates, each its own objective at its own rate, summed.  They need a
     * shuffle (speo_problem_create_ex).                                        */
    SPEO_OBJ_HYBRID_1 = 17, /* Schwefel 0.3, Rastrigin 0.3, elliptic 0.4           */
    SPEO_OBJ_HYBRID_2 = 18, /* bent cigar 0.3, HGBat 0.3, Rastrigin 0.4            */
    SPEO_OBJ_HYBRID_3 = 19, /* Griewank 0.2, Weierstrass 0.2, Rosenbrock 0.3, Schaffer F6 0.3 */
    SPEO_OBJ_HYBRID_4 = 20, /* HGBat 0.2, discus 0.2, Griewank-Rosenbrock 0.3, Rastrigin 0.3 */
    SPEO_OBJ_HYBRID_5 = 21, /* Schaffer F6 0.1, HGBat 0.2, Rosenbrock 0.2, Schwefel 0.2, elliptic 0.3 */
    SPEO_OBJ_HYBRID_6 = 22  /* Katsuura 0.1, HappyCat 0.2, Griewank-Rosenbrock 0.2, Schwefel 0.2,
                               Ackley 0.2                                            */
};

/* Precision of the rotation GEMM (accumulation is always FP32). */
enum {
    SPEO_PREC_FP32 = 0,   /* FP32 operands, v_mfma_f32_16x16x4_f32          */
    SPEO_PREC_BF16 = 1,   /* BF16 operands                                   */
    SPEO_PREC_FP16 = 2,   /* FP16 operands (x and M scaled to FP16 range)    */
    SPEO_PREC_BF16X3 = 3, /* split BF16: hi*hi + hi*lo + lo*hi               */
    SPEO_PREC_FP16X3 = 4, /* split FP16, same three products                 */
    SPEO_PREC_BF16X6 = 5  /* three-part BF16 split, six products (~FP32)     */
};

typedef struct speo_problem_s *speo_problem;

/*
 * A shifted-rotated problem of dimension `dim` with `ncomp` components.
 * ncomp == 1: f(x) = lambda_0 * g_0(rate * M_0 (x - o_0)) + bias_0.
 * ncomp > 1: the composition of CEC2014.cc:1481-1519 (cf_cal) over
 * components c with sigma[c], bias[c], lambda[c].
 * shift: [ncomp][dim] (FP64; stored on the device in FP32 and FP64),
 * rotation: [ncomp][dim][dim] row-major, z = M (x - o); NULL = identity.
 * sigma/bias/lambda may be NULL (10 / 0 / 1).  rate: NULL = 1.
 */
SPEO_EXPORT int speo_problem_create(int dim, int ncomp, const int *objectives,
                                    const double *shift, const double *rotation,
                                    const double *sigma, const double *bias,
                                    const double *lambda, const double *rate,
                                    speo_problem *out);
/* Same, for components that may be hybrid functions: shuffle [ncomp][dim]
 * holds each component's 1-based permutation of 1..dim (the reference's
 * shuffle_data format, CEC2014.cc:440-458), read for hybrid components only;
 * NULL when there are none. */
SPEO_EXPORT int speo_problem_create_ex(int dim, int ncomp, const int *objectives,
                                       const double *shift, const double *rotation,
                                       const int *shuffle, const double *sigma,
                                       const double *bias, const double *lambda,
                                       const double *rate, speo_problem *out);
SPEO_EXPORT int speo_problem_destroy(speo_problem p);
SPEO_EXPORT int speo_problem_dim(speo_problem p);
SPEO_EXPORT int speo_problem_ncomp(speo_problem p); /* components (1 unless a composition) */

/* Seeded problem data: o_j = fl32(-80 + 160 u), M = Haar orthogonal (QR of a
 * Philox/Box-Muller N(0,1) matrix, sign-fixed so diag(R) > 0).  Host-side,
 * FP64.  shift_out [dim], rotation_out [dim][dim] (may be NULL). */
SPEO_EXPORT int speo_generate_problem_data(uint64_t seed, int dim, double *shift_out,
                                           double *rotation_out);
SPEO_EXPORT uint64_t speo_problem_seed(uint64_t base_seed, uint32_t index);

/* Reads the reference's problem data files, as CEC2014::Initilize
 * (CEC2014.cc:353-420) and CEC2014_CUDA::Initilize (CEC2014_CUDA.cc:195-277):
 * <dir>/M_<func_id>_D<dim>.txt holds ncomp * dim * dim matrix values and
 * <dir>/shift_data_<func_id>.txt one line per component of which the first
 * dim values are taken.  Host only.  shift_out [ncomp][dim], rotation_out
 * [ncomp][dim][dim] (either may be NULL).  SPEO_ERR_ARG with the file name
 * when a file is missing or holds too few values. */
SPEO_EXPORT int speo_load_cec_data(const char *dir, int func_id, int dim, int ncomp,
                                   double *shift_out, double *rotation_out);

/* The problem of CEC2014 function `func_id` (1..30) at `dim` from those files,
 * with the reference's objectives, rates, rotation flags and composition
 * parameters (CEC2014.cc:14-145, 954-1372; kernel_CEC2014_CUDA.cu:1099-1671);
 * the hybrids, and the compositions built on them (F29, F30), also read
 * <dir>/shuffle_data_<func_id>_D<dim>.txt (ncomp * dim 1-based indices,
 * CEC2014.cc:440-458).  Values exclude the function's bias 100 * func_id
 * (API_evaluateFitness takes it as its `bias` argument). */
SPEO_EXPORT int speo_problem_create_cec(const char *dir, int func_id, int dim, speo_problem *out);

/* f_out[i] = f(X[i]) for n rows of stride ld (device pointers), FP32
 * evaluation with the rotation GEMM in `precision`.  Rows must have ld >= dim. */
SPEO_EXPORT int speo_evaluate(speo_problem p, const float *X, int ld, int n, float *f_out,
                              int precision, speo_stream_t stream);

/* Kernel of speo_evaluate's 16-bit evaluations (BF16, FP16 and their
 * splits) of problems wider than 128 columns on at least 1024 rows, where the
 * arithmetic is that of 64-row tiles: enable 1 (default) a kernel with
 * 256-column tiles, 0 the 64-row-tile kernel, -1 leaves the setting.  Both
 * give the same bits; the switch is for tests and measurements.  Process-wide;
 * returns the previous setting, and *launches (optional) the number of
 * launches of the 256-column kernel so far. */
SPEO_EXPORT int speo_wide_eval(int enable, long long *launches);

/* FP64 evaluation on the device (re-verification path). */
SPEO_EXPORT int speo_evaluate_f64(speo_problem p, const float *X, int ld, int n, double *f_out,
                                  speo_stream_t stream);

/* speo_evaluate without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * workspace (per-chunk row sums above 128 columns, per-component values of a
 * composition: 2 ceil(dim/128) n + ncomp n floats, the first term only when
 * dim > 128, the second only when ncomp > 1) is a fixed 4 MiB (2^20 floats)
 * device array per GPU, held until the process ends; larger inputs are
 * rejected with SPEO_ERR_ARG before anything is launched.  Every call on one
 * GPU shares that array, which keeps the last call's contents: calls that
 * overlap on one GPU (other streams, other threads) race.  Same results as
 * speo_evaluate, bit for bit. */
SPEO_EXPORT int speo_evaluate_noalloc(speo_problem p, const float *X, int ld, int n,
                                      float *f_out, int precision, speo_stream_t stream);

/* ---- DE operators on caller-owned buffers (device pointers) ---------- */

/* x = lo + u (hi - lo) for rows [0, n) of stride ld (padding set to 0);
 * draw j of row i is (seed, purpose 2, row_base + i, 0, j). */
SPEO_EXPORT int speo_init_population(float *X, int ld, int n, int dim, float lo, float hi,
                                     uint64_t seed, uint32_t row_base, speo_stream_t stream);

/*
 * DE generation for one population of n rows split into subisland_num
 * equal sub-islands through perm (NULL = identity), with per-sub-island
 * de_info (device, [subisland_num]) and best (device, [subisland_num] or
 * NULL).  Row i writes slot perm[i] of U.  Optional test outputs (device,
 * may be NULL): r_out [n][5] donor slots, jrand_out [n].  Optional injected
 * draws (device, may be NULL): inj_r [n][n_inj] replace the first n_inj donor
 * draws of each row, inj_cr [n][dim] replace the crossover draws.
 */
SPEO_EXPORT int speo_de_generate(float *U, const float *X, int ld, int n, int dim,
                                 int subisland_num, const int *perm, const DEInfo *de_info,
                                 const int *best, float lo, float hi, SPEO_RandState rng,
                                 const float *inj_r, int n_inj, const float *inj_cr, int *r_out,
                                 int *jrand_out, speo_stream_t stream);

/* Strict-"<" greedy selection of n rows; accept_out (device, [n]) optional. */
SPEO_EXPORT int speo_de_select(float *X, float *fX, const float *U, const float *fU, int ld,
                               int n, int dim, uint8_t *accept_out, speo_stream_t stream);

/* Per-segment argmin of fitness (ties to the lower row, NaN last):
 * segments are rows [off[s], off[s+1]) (device), best_out [nseg] (device). */
SPEO_EXPORT int speo_segment_argmin(const float *f, const int *off, int nseg, int *best_out,
                                    speo_stream_t stream);

/* ---- the island engine ---------------------------------------------- */

enum {
    SPEO_MODE_PLAIN = 0, /* generate, evaluate, select: three launches per generation */
    SPEO_MODE_FUSED = 1, /* one launch per generation: select(g-1)+generate(g)+evaluate(g) */
    SPEO_MODE_EPOCH = 2  /* one launch per epoch (the generations between migrations):
                          * one workgroup per island keeps its population on chip.
                          * Needs single-component objectives with dim <= 128.      */
};

typedef struct {
    int num_islands;
    const int *island_size;    /* host [num_islands]: NP of each island (>= 8)      */
    const int *island_problem; /* host [num_islands]: index into problems[]          */
    int num_problems;
    const speo_problem *problems; /* host [num_problems]                             */
    const float *island_F;     /* host [num_islands] or NULL (then F for all)       */
    const float *island_CR;    /* host [num_islands] or NULL (then CR for all)      */
    float F, CR;
    float min_bound, max_bound;
    uint64_t seed;
    int interval;        /* migrate when generation % interval == 0 (0 = never)     */
    int migration_size;  /* best-k sent along the ring of islands of one problem    */
    int precision;       /* SPEO_PREC_*                                             */
    int mode;            /* SPEO_MODE_*                                             */
    int use_graph;       /* capture the generation loop between migrations in a hipGraph */
    int rows_per_block;  /* 0 = autotuned; else 16, 32 or 64                        */
    /* Learned surrogate (no counterpart in the reference).  0 candidates = off.
     * When on, every target gets `surrogate_candidates` trials (candidate 0 is
     * the plain DE trial); a per-problem MLP trained on the archive of true
     * evaluations ranks them and only the best-ranked one is truly evaluated.
     * Until the MLP is trained, and whenever its Spearman rho on held-out
     * archive points is below surrogate_rho_min, candidate 0 is used (plain
     * DE).  All islands must then share one dimension, and mode must be
     * SPEO_MODE_PLAIN.  With use_graph, once every network screens and every
     * archive is full, each round (the generations up to a training point and
     * the training) is captured once and replayed. */
    int surrogate_candidates;     /* e.g. 8                                       */
    int surrogate_train_interval; /* generations between training rounds (10)    */
    int surrogate_archive;        /* ring archive rows per problem (65536)        */
    int surrogate_steps;          /* Adam steps per training round                */
    int surrogate_batch;          /* minibatch rows                               */
    float surrogate_lr;           /* Adam learning rate (1e-3)                    */
    float surrogate_rho_min;      /* minimum held-out Spearman rho (0.5)          */
    /* Epoch mode: workgroups per island (0 = automatic: islands are split while
     * there are fewer islands than compute units; 1, 2, 4 or 8 = fixed).  A
     * split island runs one launch per generation: every workgroup reads the
     * whole island from HBM, evolves its share of the rows and writes it back;
     * the launch boundary is the only synchronisation (graph-capturable). */
    int epoch_parts;
    /* The reference island's sub-islands and DE configuration (IslandInfo,
     * config.h:62-84; DE_CUDA::Initilize / ConfigureEA, DE_CUDA.cc:57-119;
     * Migrate::RegroupIslands at every migration point, migrate.cc:59,
     * 147-186).  Anything but the defaults needs SPEO_MODE_PLAIN without
     * graphs; the drawn sizes, permutations and configurations follow
     * speo_island_regroup. */
    int subisland_size;  /* rows per sub-island (>= 4, divides every island); 0 = whole island */
    int de_strategy;     /* 0 = rand/1/bin; SPEO_DE_STRATEGY(id) for the reference's id 0..8  */
    int configure_EA;    /* SPEO_CONFIGURE_*                                                  */
    int regroup_option;  /* SPEO_REGROUP_*; the dynamic options need islands of 4 * 2^k rows  */
} speo_engine_config;

/* Sub-island options of the reference's island (IslandInfo::regroup_option and
 * configure_EA, config.h:62-84, src/main.cc:62-63, 104-117). */
enum {
    SPEO_REGROUP_STATIC_ORDERED = 0,  /* "static_and_ordered" (the reference's default)  */
    SPEO_REGROUP_STATIC_RANDOM = 1,   /* "static_and_random": new permutation           */
    SPEO_REGROUP_DYNAMIC_RANDOM = 2,  /* "dynamic_and_random": new size and permutation */
    SPEO_REGROUP_DYNAMIC_ORDERED = 3  /* "dynamic_and_ordered": new size               */
};
enum {
    SPEO_CONFIGURE_CONSTANT = 0, /* "constant": every sub-island CR, F, strategy of the config */
    SPEO_CONFIGURE_DYNAMIC = 1   /* "dynamic": CR in {0.1, 0.9}, strategy in {4, 6} drawn per
                                    sub-island (DE_CUDA.cc:92-103)                           */
};
/* speo_engine_config.de_strategy from the reference's strategy ID 0..8 */
#define SPEO_DE_STRATEGY(reference_id) ((reference_id) + 1)

typedef struct speo_engine_s *speo_engine;

typedef struct {
    int64_t generations;   /* generations run                        */
    int64_t evaluations;   /* true objective evaluations (FEs)       */
    int64_t migrations;    /* migration events                        */
    int64_t graph_launches;
    int rows_per_block;    /* tile in use                            */
    int64_t surrogate_predictions; /* candidates scored by the surrogate (a problem
                                      whose held-out rho is below the threshold has
                                      its candidates skipped, not scored) */
    int64_t surrogate_rounds;      /* training rounds run               */
    int epoch_parts;               /* workgroups per island (epoch mode) */
} speo_engine_stats;

/* Allocates the populations, initialises them (Philox) and evaluates them
 * (counted as FEs, island_EA.cc:81).  Synchronous. */
SPEO_EXPORT int speo_engine_create(const speo_engine_config *cfg, speo_stream_t stream,
                                   speo_engine *out);
SPEO_EXPORT int speo_engine_destroy(speo_engine e);

/* Runs `generations` generations; after generation g (counted from 0 over the
 * engine's life) islands migrate when g % interval == 0 (island_EA.cc:93-99).
 * Asynchronous on the engine's stream. */
SPEO_EXPORT int speo_engine_run(speo_engine e, int generations);
/* Captures and instantiates every hipGraph that speo_engine_run(e, generations)
 * would launch next (use_graph engines), without running anything, so that a
 * timed or latency-critical run that follows launches only.  Enqueues no
 * work; no effect without graphs or with the surrogate. */
SPEO_EXPORT int speo_engine_prepare(speo_engine e, int generations);
SPEO_EXPORT int speo_engine_sync(speo_engine e);

/* Applies any pending selection, then per island: best row, its FP32 fitness
 * and its FP64 re-verified fitness.  Host outputs [num_islands], each may be
 * NULL.  Synchronous. */
SPEO_EXPORT int speo_engine_best(speo_engine e, int *row_out, float *f32_out, double *f64_out);

/* Host copies of the current population (pending selection applied):
 * X_out [total_rows][max_dim] (row r of island i holds dim_i values, rest 0),
 * f_out [total_rows].  Synchronous. */
SPEO_EXPORT int speo_engine_read(speo_engine e, float *X_out, float *f_out);
/* Host copies of the most recent trial vectors and their fitness, same
 * layout as speo_engine_read; call before anything that applies a pending
 * selection (speo_engine_best / _read / _write, or a migration after the last
 * generation run), which in SPEO_MODE_FUSED clears the trial fitness.
 * Synchronous. */
SPEO_EXPORT int speo_engine_read_trials(speo_engine e, float *U_out, float *fU_out);
/* Overwrites the population and fitness from host arrays (same layout). */
SPEO_EXPORT int speo_engine_write(speo_engine e, const float *X_in, const float *f_in);
/* Counters of the engine; with the surrogate on it waits for the engine's
 * stream (the count of scored candidates is kept on the device). */
SPEO_EXPORT int speo_engine_get_stats(speo_engine e, speo_engine_stats *out);
SPEO_EXPORT int speo_engine_total_rows(speo_engine e);
SPEO_EXPORT int speo_engine_max_dim(speo_engine e);

/* Runs `generations` generations eagerly with HIP events recorded on the
 * engine's stream around every launch and returns the average duration (ms)
 * of each kernel role per generation: [0] generate, [1] evaluate, [2] select
 * (plain mode), [3] fused generation (fused mode), [4] the epoch kernel and
 * [5] the donor draws ahead of it (epoch mode; launch durations divided by
 * the generations they cover); unused roles are 0.  avg_ms_out holds 6
 * doubles.  Migration follows the same schedule as speo_engine_run.
 * Synchronous. */
SPEO_EXPORT int speo_engine_profile(speo_engine e, int generations, double *avg_ms_out);

/* Per problem: the surrogate's last held-out Spearman rho (-1 before the
 * first training round), whether it currently screens candidates, and its
 * last training loss.  Host outputs [num_problems], each may be NULL.
 * Synchronous. */
SPEO_EXPORT int speo_engine_surrogate_stats(speo_engine e, float *rho_out, int *active_out,
                                            float *loss_out);

/* Island `island`'s current sub-island partition: *subisland_num_out, the
 * permutation perm_out [island_size] (row perm[i] plays member i; members
 * [s S, (s+1) S) form sub-island s) and info_out [subisland_num] (each may be
 * NULL).  Synchronous. */
SPEO_EXPORT int speo_engine_subislands(speo_engine e, int island, int *subisland_num_out,
                                       int *perm_out, DEInfo *info_out);

/* One regroup event of the reference's island, host side: event 0 is the
 * initial configuration (DE_CUDA::Initilize -> ConfigureEA), event k > 0 the
 * k-th migration point (Migrate::RegroupIslands, migrate.cc:147-186, then
 * ConfigureEA for the dynamic options).  Draws: word k of Philox4x32-10 with
 * key = seed and counter {k / 4, island, event, 19}; RandIntUnif(a, b) =
 * a + w % (b - a + 1), no draw when a == b (src/random.cc:14-19); the
 * permutation is Random::Permutate (src/random.cc:29-53).  perm [np] and
 * info_out [np / 4] are in/out: the current partition and configuration on
 * entry, the new ones on return; base holds the constant CR, F, strategy. */
SPEO_EXPORT int speo_island_regroup(uint64_t seed, int island, int event, int regroup_option,
                                    int configure_EA, int np, int subisland_num, int *perm,
                                    const DEInfo *base, DEInfo *info_out,
                                    int *subisland_num_out);

/* The surrogate's state of problem `problem` (host copies; each output may be
 * NULL): *cands_out candidates per target in the last generation (1 before
 * the first training round, else surrogate_candidates), cand_out
 * [rows][cands][round_up(dim, 32)] those candidates (rows = the problem's rows
 * in island order), pred_out [rows][cands] their scores (the network's
 * log(1 + f) predictions; unused before training), archive_X
 * [surrogate_archive][round_up(dim, 32)] and archive_y [surrogate_archive]
 * the ring archive of true evaluations (y = log(1 + f)), *head_out its next
 * slot and *appended_out the rows ever appended.  Candidates are written to
 * memory only from the first call on (until then each generation scores its
 * candidates without storing them, with the same results), so cand_out holds
 * those of the last generation run after an earlier call.  Synchronous. */
SPEO_EXPORT int speo_engine_surrogate_dump(speo_engine e, int problem, int *cands_out,
                                           float *cand_out, float *pred_out, float *archive_X,
                                           float *archive_y, int *head_out,
                                           int64_t *appended_out);
/* Overwrites the archive's labels of `problem` (archive_y [surrogate_archive]),
 * e.g. to test the held-out correlation check.  Synchronous. */
SPEO_EXPORT int speo_engine_surrogate_set_labels(speo_engine e, int problem,
                                                 const float *archive_y);

/* Device pointers of the engine's buffers (for profiling and tests). */
SPEO_EXPORT int speo_engine_buffers(speo_engine e, float **X, float **fX, int64_t *elem_off,
                                    int *row_off, int *ld);

/* ---- the surrogate MLP ------------------------------------------------ *
 * No counterpart in the reference (new in this build).  An MLP
 * dim -> 256 -> 256 -> 1 with SiLU on inputs x / 100, BF16 MFMA operands with
 * FP32 accumulation and FP32 master weights.  Flat parameter layout:
 * W1 [256][round_up(dim, 32)] (padding columns 0), b1 [256], W2 [256][256],
 * b2 [256], w3 [256], b3 [4] (one used).  These entry points expose one
 * network for tests and tools; the engine owns its own networks. */
typedef struct speo_mlp_s *speo_mlp;

SPEO_EXPORT int speo_mlp_create(int dim, int max_batch, uint64_t seed, speo_mlp *out);
SPEO_EXPORT int speo_mlp_destroy(speo_mlp m);
SPEO_EXPORT int64_t speo_mlp_num_params(speo_mlp m);
/* Host copies of the flat parameters: get_out receives them, set_in (if not
 * NULL) replaces them first and resets the Adam state.  Synchronous. */
SPEO_EXPORT int speo_mlp_params(speo_mlp m, float *get_out, const float *set_in);
/* pred[i] = MLP(X[i]) for n rows (device pointers, stride ld >= dim). */
SPEO_EXPORT int speo_mlp_forward(speo_mlp m, const float *X, int ld, int n, float *pred,
                                 speo_stream_t stream);
/* Same prediction through the fused single-kernel inference path (BF16 weight
 * copies refreshed from the master weights first). */
SPEO_EXPORT int speo_mlp_infer(speo_mlp m, const float *X, int ld, int n, float *pred,
                               speo_stream_t stream);
/* Loss 0.5 mean((y - t')^2) with t' = (t - t_mean) / t_std over n <= max_batch
 * rows (device X, t) and its gradient (host, flat layout).  Synchronous. */
SPEO_EXPORT int speo_mlp_loss_grad(speo_mlp m, const float *X, int ld, int n, const float *t,
                                   float t_mean, float t_std, float *loss_out, float *grad_out);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on the last gradient.  Synchronous. */
SPEO_EXPORT int speo_mlp_adam_step(speo_mlp m, float lr);
/* The engine's training step: a minibatch of `batch` rows drawn from the
 * archive's first n_valid slots (device AX [slot][round_up(dim, 32)], labels
 * Ay [slot]; slot of row b = floor(u n_valid) with u the u01 of word b % 4 of
 * Philox block {b / 4, round, step, 33} under `seed`, clamped to n_valid - 1,
 * minus one when slot % 8 == 7), labels standardised by t_mean / t_std, loss,
 * gradients and one Adam step in two kernels.  *loss_out (host, may be NULL)
 * receives the loss.  Synchronous. */
SPEO_EXPORT int speo_mlp_train_step(speo_mlp m, const float *AX, const float *Ay, int n_valid,
                                    uint64_t seed, uint32_t round, uint32_t step, int batch,
                                    float t_mean, float t_std, float lr, float *loss_out);
/* Spearman rank correlation (average ranks) of device arrays a, b of n <=
 * 4096 values into *rho_out (device). */
SPEO_EXPORT int speo_spearman(const float *a, const float *b, int n, float *rho_out,
                              speo_stream_t stream);

/* ---- run driver and result record ------------------------------------- *
 * The evaluation-budget loop and the result row of the reference's island
 * driver (IslandEA::Execute / PrintResult, src/island_EA.cc:78-105, 146-171;
 * header line :23), for runs whose output must line up row for row with the
 * reference's SPEO_GPU.csv. */

/* Generations the budget loop runs: while current < max_FEs, one generation
 * of per_generation evaluations (island_EA.cc:86-104; the initial population
 * counts, :81).  0 when current >= max_FEs; negative arguments give 0. */
SPEO_EXPORT int64_t speo_generations_for_fes(int64_t max_FEs, int64_t current, int64_t per_generation);
/* Runs the engine until its evaluation count (speo_engine_stats.evaluations,
 * all islands) reaches max_FEs, migrating on the engine's interval;
 * *generations_out (may be NULL) receives the generations run.  The
 * reference's budget is per GPU island, max_base_FEs * dim / island_num; an
 * engine holding every island is given max_base_FEs * dim.  Asynchronous, as
 * speo_engine_run. */
SPEO_EXPORT int speo_engine_run_fes(speo_engine e, int64_t max_FEs, int *generations_out);
/* The reference's per-run seed (src/main.cc:138). */
SPEO_EXPORT uint64_t speo_run_seed(int node_ID, int function_ID, int run_ID);
/* "CR/F=<CR>/<F>_<strategy>" as DE_CUDA::GetParameters (DE_CUDA.cc:18-56);
 * writes at most buflen bytes including the terminator and returns the
 * length, or SPEO_ERR_ARG for an unknown strategy or a short buffer. */
SPEO_EXPORT int speo_de_parameters(double CR, double F, int strategy_ID, char *buf, int buflen);

/* One result row: the 21 columns of the reference's record, in its order. */
typedef struct speo_run_record {
    int function_ID, run_ID, dim;
    double best_fitness, time_period, communication_percentage;
    int64_t total_FEs;
    int node_num, GPU_num, island_num, total_pop_size, subisland_size, interval, migration_size;
    double migration_rate;
    int buffer_capacity;
    const char *migration_topology, *buffer_manage, *regroup_option, *configure_EA,
        *EA_parameters;
} speo_run_record;

/* Runs until the evaluation count reaches max_FEs, as speo_engine_run_fes,
 * and measures it: *seconds_out is the wall time from the call to the
 * completion of the last generation, *comm_seconds_out the part spent at
 * migration points (migration and regroup run between synchronisations, as
 * the reference brackets MigrateOut / MigrateIn with MPI_Wtime,
 * island_EA.cc:94-104).  Results equal speo_engine_run_fes's.  Synchronous. */
SPEO_EXPORT int speo_engine_run_fes_timed(speo_engine e, int64_t max_FEs, int *generations_out,
                                          double *seconds_out, double *comm_seconds_out);

/* The reference's function x run sweep on this GPU (ConstructAndExecuteIslandModule,
 * src/main.cc:124-150; IslandEA::Initilize / Execute / PrintResult,
 * src/island_EA.cc:35-171): for every function and run, the seed
 * speo_run_seed(node_ID, f, r), the CEC2014 problem from data_dir, an engine of
 * island_num islands of island_size rows, the evaluation budget
 * max_base_FEs * dim over all islands, the FP64-verified best, the measured
 * time and communication share, and one row appended to record_file (created
 * with the header line when absent, island_EA.cc:13-31).  Strings of the
 * record: migration_topology "ring", buffer_manage "none" (ring migration of
 * the best rows needs no immigrant buffer), regroup_option and configure_EA
 * as the reference spells them. */
typedef struct {
    const char *data_dir;      /* the reference's input_data directory            */
    const char *record_file;   /* CSV path; NULL = no file                        */
    int dim;
    const int *function_IDs;   /* host [num_functions], 1..30                      */
    int num_functions;
    const int *run_IDs;        /* host [num_runs]                                  */
    int num_runs;
    int node_ID, node_num, GPU_num;
    int island_num, island_size;
    int max_base_FEs;
    int interval, migration_size;
    double migration_rate;     /* recorded only                                    */
    int buffer_capacity;       /* recorded only                                    */
    int subisland_size, configure_EA, regroup_option;
    int mode, precision, use_graph;
    float min_bound, max_bound;
} speo_driver_config;

/* rows_out [num_functions * num_runs] (may be NULL) receives the records; their
 * string fields point to static storage.  Synchronous. */
SPEO_EXPORT int speo_island_driver(const speo_driver_config *cfg, speo_run_record *rows_out);

/* The header line of the record file (no newline). */
SPEO_EXPORT const char *speo_result_header(void);
/* The row as the reference's stream output writes it (numbers with six
 * significant digits, %g), without the newline; returns the length, or
 * SPEO_ERR_ARG for a null argument, a null string field or a short buffer. */
SPEO_EXPORT int speo_result_record(const speo_run_record *r, char *buf, int buflen);

#ifdef __cplusplus
}
#endif

#endif /* SPEO_DE_H_ */
