// engine.hip -- the island engine: islands of uneven size and dimension in one
// GPU's HBM, evolved by DE/rand/1/bin with ring migration.
//
// Reference counterparts: the generation loop of IslandEA::Execute
// (src/island_EA.cc:78-113), DE_CUDA::Run (CUDA_EA_library/DE_CUDA.cc:200-216),
// EA_CUDA::FindBestIndividual (EA_CUDA.cc:104-120) and the MPI migration of
// src/migrate.cc:46-144, which here becomes a ring among islands in HBM.
//
// One kernel, k_generation, covers the four per-generation roles over a table
// of (island, row tile) work items, so every launch spans all islands:
//   kModeGen     trials U <- X                       (plain engine, launch 1)
//   kModeEval    f(U) through the MFMA evaluation    (plain engine, launch 2)
//   kModeSelect  X, fX <- U, fU where fU < fX        (plain engine, launch 3)
//   kModeFused   select(g-1) + generate(g) + evaluate(g) in one launch.
// The fused mode needs no grid-wide barrier: the selection of generation g-1
// is not applied before the donors are read; instead a donor row r is read as
// U_prev[r] when fU_prev[r] < fX_prev[r] and as X[r] otherwise, and X[r] is
// only written for accepted rows (which nobody reads as X), while U and the
// fitness arrays are double-buffered.  The result equals the plain engine's.
#pragma clang fp contract(off)

#include <sys/stat.h>

#include <chrono>
#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <map>
#include <memory>
#include <mutex>
#include <tuple>
#include <vector>

#include "engine.h"
#include "eval.h"
#include "eval_f64.h"
#include "surrogate.h"

namespace speo {

enum GenMode { kModeGen = 0, kModeEval = 1, kModeSelect = 2, kModeFused = 3 };

// A block's work item: rows [row0, row0 + nrows) of an island, with a copy of
// the island's descriptor (immutable after set-up), so a block reads its work
// in one round trip instead of two dependent ones.
struct TileDesc {
    IslandDesc I;
    int island, row0, nrows, pad;
};

struct GenerationArgs {
    const IslandDesc *isl;
    const TileDesc *tiles;
    const FusedProblem *probs;
    float *X;
    const float *fX_prev;
    float *fX_new;
    const float *U_prev;
    float *fU_prev;
    float *U_new;
    float *fU_new;
    const float *eval_src;
    float *eval_out;
    float lo, hi;
    uint64_t seed;
    const uint32_t *gen_base;
    uint32_t gen_offset;
    int trial_in_lds;
    int clear_trials;
};

constexpr int kTLD = kSmallD;  // widest dimension whose rows stay in LDS

// LDS of one block: [trial rows RT x kTS] [gval kMaxComp x RT] [GEMM region];
// with rows in LDS the GEMM region is the small-D path's (scratch + row
// reduction), otherwise the K-chunked path's.
template <int RT, int PREC>
constexpr int generation_lds_bytes(bool trial_in_lds) {
    return trial_in_lds ? RT * kTS * 4 + kMaxComp * RT * 4 +
                              small_scratch_bytes<RT>(Prec<PREC>::parts) + 10 * RT * 4
                        : kMaxComp * RT * 4 + gemm_lds_bytes<RT, PREC>();
}

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4(float *p, float4 v) { *reinterpret_cast<float4 *>(p) = v; }
__device__ __forceinline__ float comp4(const float4 &v, int q) {
    return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}

// Row tile 32: three waves per SIMD (at most 168 registers).  Held to four
// waves (128 registers) the fused FP32 kernel spills 44 registers and runs
// 9 % slower on an uneven batch of 256 islands, D 10..100 (with 13 spilled
// values, before the FP64 heavy-column step was added, four waves were 7 %
// faster).  The other row tiles keep the compiler's allocation.
template <int RT, int PREC, int MODE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RT == 32 ? 3 : 1))) void k_generation(
    GenerationArgs a) {
    extern __shared__ __align__(16) char lds[];
    const TileDesc t = a.tiles[blockIdx.x];
    const IslandDesc I = t.I;
    const int tid = threadIdx.x;

    if constexpr (MODE == kModeSelect) {
        // 16 lanes per row: trials (U_prev, fU_prev) into (X, fX_new).
        for (int pass = 0; pass < RT / 16; ++pass) {
            const int lr = pass * 16 + (tid >> 4), l = tid & 15;
            if (lr >= t.nrows) continue;
            const int li = t.row0 + lr, gr = I.row_off + li;
            const float fu = a.fU_prev[gr], fx = a.fX_new[gr];
            const bool acc = fu < fx;
            if (acc) {
                const float *u = a.U_prev + I.elem_off + (size_t)li * I.ld;
                float *x = a.X + I.elem_off + (size_t)li * I.ld;
                for (int g = l; g < (I.ld >> 2); g += 16) st4(x + 4 * g, ld4(u + 4 * g));
            }
            if (l == 0) {
                if (acc) a.fX_new[gr] = fu;
                if (a.clear_trials) a.fU_prev[gr] = INFINITY;
            }
        }
        return;
    }

    const bool lds_trial = (MODE == kModeFused || MODE == kModeEval) && a.trial_in_lds;
    float *trial = reinterpret_cast<float *>(lds);
    float *gval = trial + (lds_trial ? RT * kTS : 0);
    char *gemm_lds = reinterpret_cast<char *>(gval + kMaxComp * RT);

    if constexpr (MODE == kModeGen || MODE == kModeFused) {
        const uint32_t gen = (a.gen_base ? *a.gen_base : 0u) + a.gen_offset;
        const int ngroups = (I.dim + 3) >> 2;
        const uint32_t crt = cr_threshold(I.CR);
        // Two passes (rows lr and lr + 16 of the tile, 16 lanes per row) at a
        // time: the parent-selection loads of both rows, then their row
        // loads, are in flight together.
        struct RowCtx {
            bool valid, own_acc;
            int lr, li, jr;
            uint32_t gr;
            float fx_own, fu_own;
            const float *p0, *p1, *p2, *own;
        };
        const int l = tid & 15;
        auto prep = [&](int lr) {
            RowCtx c{};
            c.lr = lr;
            c.valid = lr < t.nrows;
            if (!c.valid) return c;
            c.li = t.row0 + lr;
            c.gr = (uint32_t)(I.row_off + c.li);
            DonorStream ds(a.seed, c.gr, gen);
            int r[kMaxR];
            draw_donors(ds, nullptr, 0, I.np, I.dim, r, c.jr);
            // Row `k` of the island as the parent population of this generation.
            auto parent = [&](int k) -> const float * {
                const int g = I.row_off + k;
                if (MODE == kModeFused && a.fU_prev[g] < a.fX_prev[g])
                    return a.U_prev + I.elem_off + (size_t)k * I.ld;
                return a.X + I.elem_off + (size_t)k * I.ld;
            };
            c.p0 = parent(r[0]);
            c.p1 = parent(r[1]);
            c.p2 = parent(r[2]);
            c.own_acc = false;
            c.fx_own = c.fu_own = 0.f;
            if (MODE == kModeFused) {
                c.fx_own = a.fX_prev[c.gr];
                c.fu_own = a.fU_prev[c.gr];
                c.own_acc = c.fu_own < c.fx_own;
            }
            c.own = c.own_acc ? a.U_prev + I.elem_off + (size_t)c.li * I.ld
                              : a.X + I.elem_off + (size_t)c.li * I.ld;
            return c;
        };
        // trial coordinates 4 g .. 4 g + 3 of a prepared row
        auto make = [&](const RowCtx &c, int g, const float4 &e0, const float4 &e1, const float4 &e2,
                        const float4 &xo) {
            const U4 w = philox_block(a.seed, (uint32_t)g, c.gr, gen, kPurposeCross);
            float v[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = 4 * g + q;
                if (j < I.dim) {
                    const bool take = (j == c.jr) || cross_take(word(w, q), crt);
                    const float m = take ? fmaf(I.F, comp4(e1, q) - comp4(e2, q), comp4(e0, q))
                                         : comp4(xo, q);
                    v[q] = check_bound(m, a.lo, a.hi);
                } else {
                    v[q] = 0.f;
                }
            }
            const float4 out{v[0], v[1], v[2], v[3]};
            st4(a.U_new + I.elem_off + (size_t)c.li * I.ld + 4 * g, out);
            if (lds_trial) st4(trial + c.lr * kTS + 4 * g, out);
            if (MODE == kModeFused && c.own_acc) st4(a.X + I.elem_off + (size_t)c.li * I.ld + 4 * g, xo);
        };
        for (int pass = 0; pass < RT / 16; pass += 2) {
            const int lrA = pass * 16 + (tid >> 4), lrB = lrA + 16;
            const bool hasB = pass + 1 < RT / 16;
            const RowCtx A = prep(lrA);
            const RowCtx B = hasB ? prep(lrB) : RowCtx{};
            for (int g = l; g < ngroups; g += 16) {
                float4 a0, a1, a2, ao, b0, b1, b2, bo;
                if (A.valid) {
                    a0 = ld4(A.p0 + 4 * g);
                    a1 = ld4(A.p1 + 4 * g);
                    a2 = ld4(A.p2 + 4 * g);
                    ao = ld4(A.own + 4 * g);
                }
                if (B.valid) {
                    b0 = ld4(B.p0 + 4 * g);
                    b1 = ld4(B.p1 + 4 * g);
                    b2 = ld4(B.p2 + 4 * g);
                    bo = ld4(B.own + 4 * g);
                }
                if (A.valid) make(A, g, a0, a1, a2, ao);
                else if (lds_trial) st4(trial + lrA * kTS + 4 * g, float4{0.f, 0.f, 0.f, 0.f});
                if (B.valid) make(B, g, b0, b1, b2, bo);
                else if (hasB && lds_trial) st4(trial + lrB * kTS + 4 * g, float4{0.f, 0.f, 0.f, 0.f});
            }
            if (MODE == kModeFused && l == 0) {
                if (A.valid) a.fX_new[A.gr] = A.own_acc ? A.fu_own : A.fx_own;
                if (B.valid) a.fX_new[B.gr] = B.own_acc ? B.fu_own : B.fx_own;
            }
        }
    }

    if constexpr (MODE == kModeEval || MODE == kModeFused) {
        const FusedProblem *P = a.probs + I.problem;
        if (MODE == kModeEval && P->generic) return;  // the generic kernel evaluates this island
        const int ncomp = P->ncomp, dim = I.dim;
        if (MODE == kModeEval && lds_trial) {
            // stage the rows to evaluate into the LDS tile (zero rows past nrows)
            const float *srcp = a.eval_src + I.elem_off + (size_t)t.row0 * I.ld;
            const int ngroups = (dim + 3) >> 2;
            for (int e = tid; e < RT * ngroups; e += 256) {
                const int r = e / ngroups, g = e % ngroups;
                st4(trial + r * kTS + 4 * g,
                    r < t.nrows ? ld4(srcp + (size_t)r * I.ld + 4 * g) : float4{0.f, 0.f, 0.f, 0.f});
            }
        }
        __syncthreads();
        RowSource src;
        if (lds_trial)
            src = RowSource{trial, kTS, t.nrows};
        else if (MODE == kModeEval)
            src = RowSource{a.eval_src + I.elem_off + (size_t)t.row0 * I.ld, I.ld, t.nrows};
        else
            src = RowSource{a.U_new + I.elem_off + (size_t)t.row0 * I.ld, I.ld, t.nrows};
        for (int c = 0; c < ncomp; ++c) {
            const CompDev comp = P->comp[c];
            if (lds_trial) {
                float *red = reinterpret_cast<float *>(gemm_lds + small_scratch_bytes<RT>(Prec<PREC>::parts));
                float *o1 = red + 8 * RT, *o2 = o1 + RT;
                float s1[RT / 16][4], s2[RT / 16][4];
                gemm_small<RT, PREC>(comp, trial, dim, gemm_lds, s1, s2);
                reduce_rows_small<RT>(s1, s2, red, o1, o2);
                if (tid < t.nrows) gval[c * RT + tid] = finish_objective_ext(comp.objective, o1[tid], o2[tid], dim);
            } else {
                float *red = reinterpret_cast<float *>(gemm_lds);
                float *o1 = red + 8 * RT, *o2 = o1 + RT;
                float s1[4] = {0.f, 0.f, 0.f, 0.f}, s2[4] = {0.f, 0.f, 0.f, 0.f};
                for (int n0 = 0; n0 < dim; n0 += kNC)
                    gemm_chunk<RT, PREC, false>(comp, src, dim, n0, gemm_lds, s1, s2);
                reduce_rows<RT>(s1, s2, red, o1, o2);
                if (tid < t.nrows) gval[c * RT + tid] = finish_objective_ext(comp.objective, o1[tid], o2[tid], dim);
            }
            if (comp.heavy.j >= 0) {
                // the heavy column's FP64 term, one wave per row
                __syncthreads();
                const int wave = tid >> 6, lane = tid & 63;
                for (int r = wave; r < t.nrows; r += 4) {
                    const double h = heavy_term(comp.heavy, src.rows + (size_t)r * src.ld, dim, lane);
                    if (lane == 0) gval[c * RT + r] = with_heavy(gval[c * RT + r], (float)h);
                }
                __syncthreads();
            }
        }
        if (tid < t.nrows) {
            float f;
            if (ncomp == 1) {
                f = fmaf(P->lambda[0], gval[tid], P->bias[0]);
            } else {
                // cf_cal weights (CEC2014.cc:1481-1519) in the log domain.
                const float *x = src.rows + (size_t)tid * src.ld;
                // (components at distance 0 share the weight equally, as
                // cf_cal's w = 1e99 each)
                float lw[kMaxComp], lwmax = -INFINITY;
                bool zero[kMaxComp];
                bool any_zero = false;
                for (int c = 0; c < ncomp; ++c) {
                    const float *o = P->comp[c].shift;
                    float d = 0.f;
                    for (int j = 0; j < dim; ++j) {
                        const float v = x[j] - o[j];
                        d = fmaf(v, v, d);
                    }
                    zero[c] = d == 0.f;
                    any_zero = any_zero || zero[c];
                    const float s = P->sigma[c];
                    lw[c] = -0.5f * logf(d) - d / (2.0f * (float)dim * s * s);
                    lwmax = fmaxf(lwmax, lw[c]);
                }
                float wsum = 0.f, acc = 0.f;
                for (int c = 0; c < ncomp; ++c) {
                    const float w = any_zero ? (zero[c] ? 1.f : 0.f) : expf(lw[c] - lwmax);
                    wsum += w;
                    acc = fmaf(w, fmaf(P->lambda[c], gval[c * RT + tid], P->bias[c]), acc);
                }
                f = acc / wsum;
            }
            float *out = MODE == kModeEval ? a.eval_out : a.fU_new;
            out[I.row_off + t.row0 + tid] = f;
        }
    }
}

__global__ void k_fill(float *p, int n, float v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_add_u32(uint32_t *p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) p[0] += v;
}
__global__ void k_eval_f64_rows(const CompParams *const *comps, const int *ncomp, const int *dims,
                                const IslandDesc *isl, const int *rows, const float *X,
                                double *out) {
    extern __shared__ __align__(16) char lds_raw[];
    const IslandDesc I = isl[blockIdx.x];
    const int li = rows[blockIdx.x] - I.row_off;
    const double f = eval_row_f64(comps[I.problem], ncomp[I.problem], dims[I.problem],
                                  X + I.elem_off + (size_t)li * I.ld,
                                  reinterpret_cast<double *>(lds_raw));
    if (threadIdx.x == 0) out[blockIdx.x] = f;
}

// ------------------------------------------------------------------ launch

template <int RT, int PREC, int MODE>
static int launch_generation(const GenerationArgs &a, int ntiles, hipStream_t s) {
    const int lds = (MODE == kModeSelect || MODE == kModeGen)
                        ? 0
                        : generation_lds_bytes<RT, PREC>(a.trial_in_lds != 0);
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_generation<RT, PREC, MODE>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  std::max(generation_lds_bytes<RT, PREC>(true),
                                           generation_lds_bytes<RT, PREC>(false)));
    });
    hipLaunchKernelGGL((k_generation<RT, PREC, MODE>), dim3(ntiles), dim3(256), lds, s, a);
    return check_launch("k_generation");
}

template <int RT, int PREC>
static int launch_mode(int mode, const GenerationArgs &a, int ntiles, hipStream_t s) {
    switch (mode) {
    case kModeGen: return launch_generation<RT, PREC, kModeGen>(a, ntiles, s);
    case kModeEval: return launch_generation<RT, PREC, kModeEval>(a, ntiles, s);
    case kModeSelect: return launch_generation<RT, SPEO_PREC_FP32, kModeSelect>(a, ntiles, s);
    default: return launch_generation<RT, PREC, kModeFused>(a, ntiles, s);
    }
}

template <int RT>
static int launch_prec(int prec, int mode, const GenerationArgs &a, int ntiles, hipStream_t s) {
    if (mode == kModeGen || mode == kModeSelect) return launch_mode<RT, SPEO_PREC_FP32>(mode, a, ntiles, s);
    switch (prec) {
    case SPEO_PREC_FP32: return launch_mode<RT, SPEO_PREC_FP32>(mode, a, ntiles, s);
    case SPEO_PREC_BF16: return launch_mode<RT, SPEO_PREC_BF16>(mode, a, ntiles, s);
    case SPEO_PREC_FP16: return launch_mode<RT, SPEO_PREC_FP16>(mode, a, ntiles, s);
    case SPEO_PREC_BF16X3: return launch_mode<RT, SPEO_PREC_BF16X3>(mode, a, ntiles, s);
    case SPEO_PREC_FP16X3: return launch_mode<RT, SPEO_PREC_FP16X3>(mode, a, ntiles, s);
    case SPEO_PREC_BF16X6: return launch_mode<RT, SPEO_PREC_BF16X6>(mode, a, ntiles, s);
    default: return set_error(SPEO_ERR_ARG, "unknown precision");
    }
}

static int launch_any(int rt, int prec, int mode, const GenerationArgs &a, int ntiles,
                      hipStream_t s) {
    switch (rt) {
    case 16: return launch_prec<16>(prec, mode, a, ntiles, s);
    case 32: return launch_prec<32>(prec, mode, a, ntiles, s);
    case 64: return launch_prec<64>(prec, mode, a, ntiles, s);
    default: return set_error(SPEO_ERR_ARG, "rows_per_block must be 16, 32 or 64");
    }
}

// ------------------------------------------------------------ tuning cache
//
// The autotuner's choices, persisted across processes: one "key value" line
// per choice in $SPEO_TUNE_CACHE (empty: no file), else
// $HOME/.cache/speo_de/autotune-v1.txt.  Keys name the mode, the precision
// and every island's (dimension, NP rounded up to a power of two).

namespace {
std::mutex g_tune_mu;

std::string tune_path() {
    if (const char *e = std::getenv("SPEO_TUNE_CACHE")) return e;
    const char *h = std::getenv("HOME");
    return h ? std::string(h) + "/.cache/speo_de/autotune-v1.txt" : std::string();
}

bool tune_lookup(const std::string &key, int *value) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string path = tune_path();
    if (path.empty()) return false;
    std::ifstream f(path);
    std::string k;
    int v = 0;
    bool found = false;
    while (f >> k >> v)
        if (k == key) {
            *value = v;  // the last line for a key wins
            found = true;
        }
    return found;
}

void tune_store(const std::string &key, int value) {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    const std::string path = tune_path();
    if (path.empty()) return;
    for (size_t i = 1; i < path.size(); ++i)  // create the directories on the way
        if (path[i] == '/') (void)::mkdir(path.substr(0, i).c_str(), 0755);
    std::ofstream f(path, std::ios::app);
    if (f) f << key << ' ' << value << '\n';
}
}  // namespace

// ------------------------------------------------------------------ engine

struct Engine {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    speo_engine_config cfg{};
    int I = 0, total_rows = 0, max_dim = 0, max_ld = 0;
    int64_t total_elems = 0;
    std::vector<IslandDesc> isl;
    std::vector<int> row_off, dst;
    std::vector<speo_problem> problems;
    bool all_small = true;  // every dim <= kTLD
    bool any_generic = false;  // some problem needs the generic kernel (plain mode only)
    // The reference island's sub-islands and DE configuration (plain mode):
    // per island the sub-island count, permutation and DEInfo per sub-island.
    bool configured = false;
    int regroup_events = 0;
    // measured communication (run_fes_timed): migrations (and regroups) run
    // eagerly between synchronisations and their wall time is summed, as the
    // reference's MPI_Wtime brackets around MigrateOut / MigrateIn
    // (island_EA.cc:94-104)
    bool time_comm = false;
    double comm_seconds = 0.0;
    int timed_point(bool mig, bool point);
    std::vector<int> subs_h;
    std::vector<std::vector<int>> perm_h;
    std::vector<std::vector<DEInfo>> info_h;
    std::vector<DEInfo> base_h;
    std::vector<int> doff;           // island -> first DEInfo / best slot
    int *perm_dev = nullptr, *best_sub_dev = nullptr;
    DEInfo *info_dev = nullptr;
    int configure_setup();
    int regroup(bool initial);
    int generate_configured(int gen_offset, hipStream_t s);
    int subislands(int island, int *subs_out, int *perm_out, DEInfo *info_out);
    int eval_generic_islands(const float *src, float *out, hipStream_t s);

    IslandDesc *isl_dev = nullptr;
    int *row_off_dev = nullptr, *dst_dev = nullptr, *best_buf = nullptr, *worst_buf = nullptr;
    int *best_row_dev = nullptr;
    double *best_f64_dev = nullptr;
    FusedProblem *probs_dev = nullptr;
    const CompParams **params_tab_dev = nullptr;
    int *ncomp_dev = nullptr, *dims_dev = nullptr;
    std::map<int, TileDesc *> tiles_dev;  // per rows-per-block
    std::map<int, int> ntiles;
    uint32_t *gen_base = nullptr;

    float *X = nullptr, *fX[2] = {nullptr, nullptr}, *U[2] = {nullptr, nullptr};
    float *fU[2] = {nullptr, nullptr};
    int p = 0;
    bool pending = false;

    int rt = 32;
    int64_t gens = 0, fes = 0, migrations = 0, graph_launches = 0;

    // learned surrogate (config.surrogate_candidates > 0)
    struct SurrProblem {
        Mlp net;
        int qoff = 0, rows = 0, head = 0;
        int64_t appended = 0;
        bool trained = false;
    };
    int C = 0, Kp = 0, Q = 0;
    std::vector<std::unique_ptr<SurrProblem>> surr;
    SurrRowDesc *srows_dev = nullptr;
    float *cand = nullptr, *pred = nullptr, *rho_dev = nullptr, *AX = nullptr, *Ay = nullptr;
    long long *scored_dev = nullptr;  // candidates scored per problem by the one-kernel screening
    float *Xb = nullptr, *tb = nullptr, *Xh = nullptr, *yh = nullptr, *predh = nullptr;
    float2 *ranks_ws = nullptr;  // Spearman workspace
    int *heads_dev = nullptr;         // next archive slot per problem (device)
    uint32_t *round_dev = nullptr;    // training rounds run (device; keys the minibatch draws)
    bool any_trained = false;
    int last_Ce = 1;  // candidates per target in the last surrogate generation
    bool retain_cand = false;  // candidates also written to memory (after a dump request)
    int64_t surr_pred = 0, surr_rounds = 0;
    int surrogate_setup();
    int surrogate_generation(int gen_offset, hipStream_t s);
    int enqueue_surrogate_round(int TI, int mig_pos, hipStream_t s);
    void surrogate_round_host(int TI);
    bool surrogate_round_replayable() const;
    int surrogate_train(hipStream_t s);
    std::map<std::tuple<int, int, int>, hipGraphExec_t> graphs;
    bool graph_failed = false;

    // epoch mode: per-island plans, block order, LDS of a launch
    EpochBlock *blocks_dev = nullptr;
    int epoch_lds = 0, epoch_parts = 1, epoch_blocks = 0;
    float *Xnext = nullptr;  // split islands: the population of the next generation
    int ep_par = 0;          // split islands: buffers swapped an odd number of times
    int4 *donor_buf = nullptr;    // [donor_gens][rows] donors drawn ahead of the epoch launches
    int *row_island_dev = nullptr;
    int donor_gens = 0;
    int epoch_setup();
    int epoch_plan_all(int parts, std::vector<EpochPlan> *plans, int *lds);
    int launch_epoch_chunk(int L, hipStream_t s);
    void swap_epoch_buffers();

    ~Engine();
    int create(const speo_engine_config &c, hipStream_t s);
    int upload_tables();
    int build_tiles(int rt_);
    int launch(int mode, int gen_offset, int rt_, hipStream_t s, int clear = 0);
    int enqueue_chunk(int L, bool migrate, hipStream_t s);
    int finalize_pending(hipStream_t s);
    int migrate(hipStream_t s);
    // dry: capture and instantiate the chunks' graphs, launch nothing (host
    // state is advanced as if launched; prepare() restores it)
    int run(int generations, bool dry = false);
    int prepare(int generations);
    int profile(int generations, double *avg_ms);
    int profile_epoch(int generations, double *avg_ms);
    // when set, launch_epoch_chunk records an event pair around every launch:
    // (role 4 = epoch kernel, role 5 = donor draws, events, generations covered)
    struct ProfEvent {
        int role;
        hipEvent_t a, b;
        int gens;
    };
    std::vector<ProfEvent> *prof = nullptr;
    int prof_begin(int role, int gens, hipStream_t s);
    int prof_end(hipStream_t s);
    int next_chunk(int left, bool *mig) const;
    int autotune();
    // epoch mode: workgroups per island measured over the candidates on
    // throwaway engines of the same configuration (cfg.epoch_parts == 0)
    bool tuning_child = false;
    int tuned_from_cache = 0;  // bit 0: row tile, bit 1: epoch parts
    std::string tune_key(const char *what) const;
    int tune_epoch_parts();
    int best(int *row_out, float *f32_out, double *f64_out);
    int read(float *Xo, float *fo);
    int read_rows(const float *src, const float *fsrc, float *Xo, float *fo);
    int write(const float *Xi, const float *fi);
};

Engine::~Engine() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : tiles_dev) (void)hipFree(kv.second);
    void *bufs[] = {isl_dev, row_off_dev, dst_dev, best_buf, worst_buf, best_row_dev,
                    best_f64_dev, probs_dev, params_tab_dev, ncomp_dev, dims_dev, gen_base,
                    blocks_dev, Xnext, donor_buf, row_island_dev,
                    X, fX[0], fX[1], U[0], U[1], fU[0], fU[1], srows_dev, cand, pred, rho_dev, scored_dev,
                    AX, Ay, Xb, tb, Xh, yh, predh, ranks_ws, heads_dev, round_dev, perm_dev,
                    best_sub_dev, info_dev};
    for (void *b : bufs) (void)hipFree(b);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

template <typename T>
static int upload_vec(T **dst, const std::vector<T> &src) {
    SPEO_TRY_HIP(hipMalloc(reinterpret_cast<void **>(dst), sizeof(T) * std::max<size_t>(src.size(), 1)));
    if (!src.empty())
        SPEO_TRY_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return SPEO_OK;
}

int Engine::build_tiles(int rt_) {
    if (tiles_dev.count(rt_)) return SPEO_OK;
    std::vector<TileDesc> tiles;
    for (int i = 0; i < I; ++i)
        for (int r0 = 0; r0 < isl[(size_t)i].np; r0 += rt_)
            tiles.push_back(TileDesc{isl[(size_t)i], i, r0, std::min(rt_, isl[(size_t)i].np - r0), 0});
    // Heaviest tiles first (rows x dim^2, ties in island order): blocks are
    // dispatched in this order, so with uneven islands the last wave of blocks
    // holds the cheap tiles.  Tiles are independent: the order changes no result.
    auto cost = [&](const TileDesc &t) { return (int64_t)t.nrows * t.I.dim * t.I.dim; };
    std::stable_sort(tiles.begin(), tiles.end(),
                     [&](const TileDesc &x, const TileDesc &y) { return cost(x) > cost(y); });
    TileDesc *d = nullptr;
    SPEO_TRY(upload_vec(&d, tiles));
    tiles_dev[rt_] = d;
    ntiles[rt_] = (int)tiles.size();
    return SPEO_OK;
}

int Engine::upload_tables() {
    SPEO_TRY(upload_vec(&isl_dev, isl));
    SPEO_TRY(upload_vec(&row_off_dev, row_off));
    SPEO_TRY(upload_vec(&dst_dev, dst));
    std::vector<FusedProblem> fp(problems.size());
    std::vector<const CompParams *> tab(problems.size());
    std::vector<int> nc(problems.size()), dm(problems.size());
    for (size_t q = 0; q < problems.size(); ++q) {
        const ProblemHost &P = problems[q]->h;
        FusedProblem f{};
        f.ncomp = P.ncomp;
        f.dim = P.dim;
        f.generic = P.generic ? 1 : 0;
        any_generic = any_generic || P.generic;
        for (int c = 0; c < P.ncomp; ++c) {
            f.lambda[c] = (float)P.comp[c].lambda;
            f.bias[c] = (float)P.comp[c].bias;
            f.sigma[c] = (float)P.comp[c].sigma;
            f.comp[c] = P.view(c, cfg.precision);
        }
        fp[q] = f;
        tab[q] = P.params_dev;
        nc[q] = P.ncomp;
        dm[q] = P.dim;
    }
    SPEO_TRY(upload_vec(&probs_dev, fp));
    SPEO_TRY(upload_vec(&params_tab_dev, tab));
    SPEO_TRY(upload_vec(&ncomp_dev, nc));
    SPEO_TRY(upload_vec(&dims_dev, dm));
    const int k = std::max(cfg.migration_size, 1);
    SPEO_TRY_HIP(hipMalloc(&best_buf, sizeof(int) * (size_t)I * k));
    SPEO_TRY_HIP(hipMalloc(&worst_buf, sizeof(int) * (size_t)I * k));
    SPEO_TRY_HIP(hipMalloc(&best_row_dev, sizeof(int) * (size_t)I));
    SPEO_TRY_HIP(hipMalloc(&best_f64_dev, sizeof(double) * (size_t)I));
    SPEO_TRY_HIP(hipMalloc(&gen_base, sizeof(uint32_t)));
    SPEO_TRY_HIP(hipMemset(gen_base, 0, sizeof(uint32_t)));
    return SPEO_OK;
}

int Engine::create(const speo_engine_config &c, hipStream_t s) {
    cfg = c;
    if (c.num_islands < 1 || !c.island_size || !c.island_problem || c.num_problems < 1 ||
        !c.problems)
        return set_error(SPEO_ERR_ARG, "speo_engine_create: islands and problems are required");
    if (c.mode != SPEO_MODE_PLAIN && c.mode != SPEO_MODE_FUSED && c.mode != SPEO_MODE_EPOCH)
        return set_error(SPEO_ERR_ARG, "speo_engine_create: unknown mode");
    if (c.precision < SPEO_PREC_FP32 || c.precision > SPEO_PREC_BF16X6)
        return set_error(SPEO_ERR_ARG, "speo_engine_create: unknown precision");
    if (c.migration_size < 0 || c.interval < 0 || !(c.min_bound < c.max_bound))
        return set_error(SPEO_ERR_ARG, "speo_engine_create: bad migration or bounds");
    if (c.rows_per_block != 0 && c.rows_per_block != 16 && c.rows_per_block != 32 &&
        c.rows_per_block != 64)
        return set_error(SPEO_ERR_ARG, "speo_engine_create: rows_per_block must be 0, 16, 32, 64");
    problems.assign(c.problems, c.problems + c.num_problems);
    for (auto *q : problems) {
        if (!q) return set_error(SPEO_ERR_ARG, "speo_engine_create: null problem");
        if (c.precision != SPEO_PREC_FP32 && !q->h.generic)
            for (int k = 0; k < q->h.ncomp; ++k)
                if (q->h.comp[k].unrotated)
                    return set_error(SPEO_ERR_UNSUPPORTED,
                                     "speo_engine_create: a 16-bit rotation precision would round "
                                     "an unrotated component's z; use SPEO_PREC_FP32 for it");
    }
    I = c.num_islands;
    for (int i = 0; i < I; ++i) {
        const int pr = c.island_problem[i];
        if (pr >= 0 && pr < c.num_problems && problems[(size_t)pr]->h.dim > kTLD) all_small = false;
    }
    // When the fused kernel reads its trial rows back from global memory
    // (dim > kTLD), rows start on 128-byte lines so that no cache line holds
    // rows of two tiles (two blocks).
    const int ld_align = all_small ? 4 : 32;
    int64_t eoff = 0;
    row_off.assign((size_t)I + 1, 0);
    for (int i = 0; i < I; ++i) {
        const int np = c.island_size[i], pr = c.island_problem[i];
        if (np < 8 || pr < 0 || pr >= c.num_problems || np < 2 * c.migration_size)
            return set_error(SPEO_ERR_ARG, "speo_engine_create: island size must be >= 8 and "
                                           ">= 2 * migration_size; problem index in range");
        IslandDesc d{};
        d.elem_off = eoff;
        d.row_off = row_off[(size_t)i];
        d.np = np;
        d.dim = problems[(size_t)pr]->h.dim;
        d.ld = round_up(d.dim, ld_align);
        d.problem = pr;
        d.F = c.island_F ? c.island_F[i] : c.F;
        d.CR = c.island_CR ? c.island_CR[i] : c.CR;
        isl.push_back(d);
        eoff += (int64_t)np * d.ld;
        row_off[(size_t)i + 1] = row_off[(size_t)i] + np;
        max_dim = std::max(max_dim, d.dim);
        max_ld = std::max(max_ld, d.ld);
    }
    total_rows = row_off[(size_t)I];
    total_elems = eoff;
    // Ring per problem: island i sends to the next island (by index) that
    // solves the same problem; islands alone on their problem send nothing.
    dst.assign((size_t)I, -1);
    for (int i = 0; i < I; ++i) {
        for (int k = 1; k < I; ++k) {
            const int j = (i + k) % I;
            if (c.island_problem[j] == c.island_problem[i]) {
                dst[(size_t)i] = j;
                break;
            }
        }
    }
    if (s) {
        stream = s;
    } else {
        SPEO_TRY_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        own_stream = true;
    }
    SPEO_TRY(upload_tables());
    if (any_generic && (c.mode != SPEO_MODE_PLAIN || c.surrogate_candidates > 0))
        return set_error(SPEO_ERR_UNSUPPORTED,
                         "speo_engine_create: objectives 9-22 run in SPEO_MODE_PLAIN without the surrogate");
    const size_t E = (size_t)total_elems, R = (size_t)total_rows;
    SPEO_TRY_HIP(hipMalloc(&X, sizeof(float) * E));
    for (int b = 0; b < 2; ++b) {
        SPEO_TRY_HIP(hipMalloc(&U[b], sizeof(float) * E));
        SPEO_TRY_HIP(hipMalloc(&fX[b], sizeof(float) * R));
        SPEO_TRY_HIP(hipMalloc(&fU[b], sizeof(float) * R));
        SPEO_TRY_HIP(hipMemsetAsync(U[b], 0, sizeof(float) * E, stream));
    }
    for (int i = 0; i < I; ++i) {
        const IslandDesc &d = isl[(size_t)i];
        SPEO_TRY(init_population(X + d.elem_off, d.ld, d.np, d.dim, c.min_bound, c.max_bound,
                                 c.seed, (uint32_t)d.row_off, stream));
    }
    for (int b = 0; b < 2; ++b) {
        hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, stream, fU[b], total_rows, INFINITY);
        SPEO_TRY(check_launch("k_fill"));
    }
    configured = c.subisland_size != 0 || c.de_strategy != 0 || c.configure_EA != 0 ||
                 c.regroup_option != 0;
    if (configured) {
        if (c.mode != SPEO_MODE_PLAIN || c.use_graph || c.surrogate_candidates > 0)
            return set_error(SPEO_ERR_UNSUPPORTED, "speo_engine_create: sub-islands, strategies and "
                                                   "regroups run in SPEO_MODE_PLAIN without graphs "
                                                   "or the surrogate");
        if (c.de_strategy < 0 || c.de_strategy > 9 || c.configure_EA < 0 || c.configure_EA > 1 ||
            c.regroup_option < 0 || c.regroup_option > 3 || c.subisland_size < 0 ||
            (c.subisland_size > 0 && c.subisland_size < 4))
            return set_error(SPEO_ERR_ARG, "speo_engine_create: bad sub-island options");
        for (int i = 0; i < I; ++i) {
            const int np = isl[(size_t)i].np;
            if (c.subisland_size > 0 && np % c.subisland_size)
                return set_error(SPEO_ERR_ARG, "speo_engine_create: subisland_size must divide "
                                               "every island size");
            if ((c.regroup_option == SPEO_REGROUP_DYNAMIC_RANDOM ||
                 c.regroup_option == SPEO_REGROUP_DYNAMIC_ORDERED) &&
                ((np & (np - 1)) != 0 || np < 4))
                return set_error(SPEO_ERR_ARG, "speo_engine_create: dynamic regroups need islands "
                                               "of 4 * 2^k rows");
        }
    }
    if (c.surrogate_candidates > 0 && c.mode != SPEO_MODE_PLAIN)
        return set_error(SPEO_ERR_UNSUPPORTED, "speo_engine_create: the surrogate runs in "
                                               "SPEO_MODE_PLAIN");
    if (cfg.mode == SPEO_MODE_EPOCH && cfg.epoch_parts == 0 && !tuning_child)
        SPEO_TRY(tune_epoch_parts());
    if (cfg.mode == SPEO_MODE_EPOCH) SPEO_TRY(epoch_setup());
    rt = c.rows_per_block;
    if (rt == 0) SPEO_TRY(autotune());
    if (c.surrogate_candidates > 0) SPEO_TRY(surrogate_setup());
    if (configured) SPEO_TRY(configure_setup());
    SPEO_TRY(build_tiles(rt));
    // Initial evaluation (island_EA.cc:81 counts it as island_size FEs).
    GenerationArgs a{};
    a.isl = isl_dev;
    a.tiles = tiles_dev[rt];
    a.probs = probs_dev;
    a.eval_src = X;
    a.eval_out = fX[0];
    a.trial_in_lds = all_small ? 1 : 0;
    SPEO_TRY(launch_any(rt, cfg.precision, kModeEval, a, ntiles[rt], stream));
    SPEO_TRY(eval_generic_islands(X, fX[0], stream));
    hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, stream, fX[1], total_rows, INFINITY);
    SPEO_TRY(check_launch("k_fill"));
    fes = total_rows;
    SPEO_TRY_HIP(hipStreamSynchronize(stream));
    return SPEO_OK;
}

// Plans of all islands for `parts` workgroups per island; fails when one
// does not fit.
int Engine::epoch_plan_all(int parts, std::vector<EpochPlan> *plans, int *lds) {
    plans->assign((size_t)I, EpochPlan{});
    *lds = 0;
    for (int i = 0; i < I; ++i) {
        const IslandDesc &d = isl[(size_t)i];
        const EpochPlan p = epoch_plan(d.np, d.dim, parts, precision_parts(cfg.precision));
        if (p.lds_bytes > kEpochLdsMax) return SPEO_ERR_UNSUPPORTED;
        if (parts > 1 && d.np < 16 * parts) return SPEO_ERR_UNSUPPORTED;
        (*plans)[(size_t)i] = p;
        *lds = std::max(*lds, p.lds_bytes);
    }
    return SPEO_OK;
}

// Plans the epoch kernel.  With fewer islands than compute units an island
// is split over `parts` workgroups, one generation per launch
// (cfg.epoch_parts = 0: the largest power of two up to 8 with islands x parts
// <= compute units whose plans fit in LDS; else as configured); otherwise one
// workgroup per island runs whole epochs.  Workgroup b serves island b % I,
// share b / I.
int Engine::epoch_setup() {
    for (int i = 0; i < I; ++i) {
        const ProblemHost &P = problems[(size_t)isl[(size_t)i].problem]->h;
        if (P.ncomp != 1 || P.dim > kSmallD)
            return set_error(SPEO_ERR_UNSUPPORTED, "epoch mode needs single-component objectives "
                                                   "with dim <= 128");
        if (P.view(0, cfg.precision).objective > SPEO_OBJ_ACKLEY)
            return set_error(SPEO_ERR_UNSUPPORTED, "epoch mode: objectives 1-5 only");
        if (isl[(size_t)i].np > 8192)
            return set_error(SPEO_ERR_UNSUPPORTED, "epoch mode: islands of at most 8192 rows");
        // the epoch kernel reads an island's rows as one block of np * 4 ceil(dim / 4) floats
        if (isl[(size_t)i].ld != round_up(P.dim, 4))
            return set_error(SPEO_ERR_UNSUPPORTED, "epoch mode: unexpected population row stride");
    }
    int dev = 0, cus = 0;
    SPEO_TRY_HIP(hipGetDevice(&dev));
    SPEO_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (cfg.epoch_parts < 0 || cfg.epoch_parts > kEpochMaxParts ||
        (cfg.epoch_parts & (cfg.epoch_parts - 1)) != 0)
        return set_error(SPEO_ERR_ARG, "epoch_parts must be 0 (automatic), 1, 2, 4 or 8");
    int parts = cfg.epoch_parts;
    if (parts == 0) {
        parts = 1;
        while (parts < kEpochMaxParts && (int64_t)I * parts * 2 <= cus) parts *= 2;
    }
    std::vector<EpochPlan> plans;
    int lds = 0;
    while (epoch_plan_all(parts, &plans, &lds) != SPEO_OK) {
        if (parts == 1) return set_error(SPEO_ERR_UNSUPPORTED, "epoch mode: island too large for one workgroup");
        if (cfg.epoch_parts > 1)
            return set_error(SPEO_ERR_UNSUPPORTED, "epoch mode: islands cannot be split that far");
        parts /= 2;
    }
    epoch_parts = parts;
    epoch_lds = lds;
    epoch_blocks = I * parts;
    std::vector<EpochBlock> blocks((size_t)epoch_blocks);
    for (int b = 0; b < epoch_blocks; ++b) {
        const int i = b % I;
        const IslandDesc &d = isl[(size_t)i];
        const ProblemHost &P = problems[(size_t)d.problem]->h;
        const CompDev c = P.view(0, cfg.precision);
        EpochBlock &e = blocks[(size_t)b];
        e.I = d;
        e.pl = plans[(size_t)i];
        e.share = b / I;
        e.objective = c.objective;
        e.kp32 = c.kp32;
        const uint64_t ng = (uint64_t)(d.dim + 3) / 4;
        e.inv_ng = ng > 1 ? (uint32_t)((0x100000000ull + ng - 1) / ng) : 0u;
        e.crt = cr_threshold(d.CR);
        e.rate = c.rate;
        e.lambda = (float)P.comp[0].lambda;
        e.bias = (float)P.comp[0].bias;
        e.shift = c.shift;
        e.B32s = c.B32s;
        e.colw = c.colw;
        for (int q = 0; q < 3; ++q) e.M16s[q] = c.M16s[q];
        e.heavy = c.heavy;
    }
    SPEO_TRY(upload_vec(&blocks_dev, blocks));
    std::vector<int> row_island((size_t)total_rows);
    for (int i = 0; i < I; ++i)
        for (int r = 0; r < isl[(size_t)i].np; ++r) row_island[(size_t)(isl[(size_t)i].row_off + r)] = i;
    SPEO_TRY(upload_vec(&row_island_dev, row_island));
    donor_gens = (int)std::max<int64_t>(1, std::min<int64_t>(64, (int64_t(1) << 26) / total_rows));
    SPEO_TRY_HIP(hipMalloc(&donor_buf, sizeof(int4) * (size_t)donor_gens * (size_t)total_rows));
    if (parts > 1) {
        SPEO_TRY_HIP(hipMalloc(&Xnext, sizeof(float) * (size_t)total_elems));
        SPEO_TRY_HIP(hipMemcpy(Xnext, X, sizeof(float) * (size_t)total_elems, hipMemcpyDeviceToDevice));
    }
    return SPEO_OK;
}

void Engine::swap_epoch_buffers() {
    std::swap(X, Xnext);
    std::swap(fX[0], fX[1]);
    ep_par ^= 1;
}

// Chunks of at most donor_gens generations: the donors of the chunk are drawn
// first, then one launch runs the chunk (one workgroup per island) or one
// launch per generation runs it (split islands, swapping the population
// buffers).
int Engine::launch_epoch_chunk(int L, hipStream_t s) {
    EpochArgs a{};
    a.blocks = blocks_dev;
    a.Xalt = U[1];
    a.U = U[0];
    a.fU = fU[0];
    a.lo = cfg.min_bound;
    a.hi = cfg.max_bound;
    a.seed = cfg.seed;
    a.gen_base = gen_base;
    a.donors = donor_buf;
    a.rows = total_rows;
    for (int g0 = 0; g0 < L; g0 += donor_gens) {
        const int Ls = std::min(donor_gens, L - g0);
        SPEO_TRY(prof_begin(5, Ls, s));
        SPEO_TRY(epoch_donors(isl_dev, row_island_dev, total_rows, Ls, cfg.seed, gen_base,
                              (uint32_t)g0, donor_buf, s));
        SPEO_TRY(prof_end(s));
        if (epoch_parts == 1) {
            a.Xin = X;
            a.Xout = X;
            a.fXin = fX[0];
            a.fXout = fX[0];
            a.gen_offset = (uint32_t)g0;
            a.donor_gen = 0;
            a.L = Ls;
            a.write_trials = 1;
            SPEO_TRY(prof_begin(4, Ls, s));
            SPEO_TRY(launch_epoch(cfg.precision, a, epoch_blocks, epoch_lds, epoch_parts > 1, s));
            SPEO_TRY(prof_end(s));
            continue;
        }
        SPEO_TRY(prof_begin(4, Ls, s));  // one event pair around the sub-chunk's launches
        for (int g = 0; g < Ls; ++g) {
            a.Xin = X;
            a.Xout = Xnext;
            a.fXin = fX[0];
            a.fXout = fX[1];
            a.gen_offset = (uint32_t)(g0 + g);
            a.donor_gen = g;
            a.L = 1;
            a.write_trials = g0 + g == L - 1;
            SPEO_TRY(launch_epoch(cfg.precision, a, epoch_blocks, epoch_lds, epoch_parts > 1, s));
            swap_epoch_buffers();
        }
        SPEO_TRY(prof_end(s));
    }
    return SPEO_OK;
}

int Engine::surrogate_setup() {
    C = cfg.surrogate_candidates;
    const int D = isl[0].dim;
    for (const auto &d : isl)
        if (d.dim != D) return set_error(SPEO_ERR_ARG, "surrogate: all islands must share one dimension");
    if (C > 32 || cfg.num_problems > kMaxSurrProblems || cfg.surrogate_archive < 1024 ||
        cfg.surrogate_batch < 16 || cfg.surrogate_steps < 0 || cfg.surrogate_train_interval < 1 ||
        !(cfg.surrogate_lr > 0.f))
        return set_error(SPEO_ERR_ARG, "surrogate: need <= 32 candidates, <= 16 problems, archive "
                                       ">= 1024, batch >= 16, train interval >= 1, lr > 0");
    Kp = round_up(D, 32);
    std::vector<SurrRowDesc> rows;
    for (int p = 0; p < cfg.num_problems; ++p) {
        auto sp = std::make_unique<SurrProblem>();
        sp->qoff = (int)rows.size();
        for (int i = 0; i < I; ++i)
            if (isl[(size_t)i].problem == p)
                for (int li = 0; li < isl[(size_t)i].np; ++li)
                    rows.push_back(SurrRowDesc{isl[(size_t)i].row_off + li, i, p, li});
        sp->rows = (int)rows.size() - sp->qoff;
        if (sp->rows > 0) {
            const int mb = std::max({cfg.surrogate_batch, sp->rows * C, 1024});
            SPEO_TRY(sp->net.alloc(D, mb));
            SPEO_TRY(sp->net.init(cfg.seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(p + 1)), stream));
        }
        surr.push_back(std::move(sp));
    }
    Q = (int)rows.size();
    SPEO_TRY(upload_vec(&srows_dev, rows));
    const size_t P = (size_t)cfg.num_problems, cap = (size_t)cfg.surrogate_archive;
    SPEO_TRY_HIP(hipMalloc(&cand, sizeof(float) * (size_t)Q * C * Kp));
    SPEO_TRY_HIP(hipMalloc(&pred, sizeof(float) * (size_t)Q * C));
    SPEO_TRY(upload_vec(&rho_dev, std::vector<float>(P, -1.0f)));
    SPEO_TRY(upload_vec(&scored_dev, std::vector<long long>(P, 0)));
    SPEO_TRY_HIP(hipMalloc(&AX, sizeof(float) * P * cap * Kp));
    SPEO_TRY_HIP(hipMalloc(&Ay, sizeof(float) * P * cap));
    SPEO_TRY_HIP(hipMalloc(&Xb, sizeof(float) * (size_t)cfg.surrogate_batch * Kp));
    SPEO_TRY_HIP(hipMalloc(&tb, sizeof(float) * (size_t)cfg.surrogate_batch));
    SPEO_TRY_HIP(hipMalloc(&Xh, sizeof(float) * 1024 * (size_t)Kp));
    SPEO_TRY_HIP(hipMalloc(&yh, sizeof(float) * 1024));
    SPEO_TRY_HIP(hipMalloc(&predh, sizeof(float) * 1024));
    SPEO_TRY_HIP(hipMalloc(&ranks_ws, sizeof(float2) * 4096));
    SPEO_TRY_HIP(hipMalloc(&heads_dev, sizeof(int) * kMaxSurrProblems));
    SPEO_TRY_HIP(hipMemset(heads_dev, 0, sizeof(int) * kMaxSurrProblems));
    SPEO_TRY_HIP(hipMalloc(&round_dev, sizeof(uint32_t)));
    SPEO_TRY_HIP(hipMemset(round_dev, 0, sizeof(uint32_t)));
    return SPEO_OK;
}

// One surrogate-screened generation: C candidates per target (only candidate
// 0 before the first training round), MLP scores, pick, true evaluation of
// the picked trial, archive append, selection.
int Engine::surrogate_generation(int gen_offset, hipStream_t s) {
    const int Ce = any_trained ? C : 1;
    last_Ce = Ce;
    const float rho_min = any_trained ? cfg.surrogate_rho_min : INFINITY;
    const float lo = cfg.min_bound, hi = cfg.max_bound;
    if (retain_cand || Kp > kFusedMaxKp) {
        // candidates written to memory, scored, the pick copied
        SPEO_TRY(gen_candidates(isl_dev, srows_dev, Q, Ce, X, lo, hi, cfg.seed, gen_base,
                                (uint32_t)gen_offset, cand, Kp, s));
        if (any_trained) {
            for (auto &sp : surr)
                if (sp->rows > 0)
                    SPEO_TRY(sp->net.infer(cand + (size_t)sp->qoff * Ce * Kp, Kp, sp->rows * Ce,
                                           pred + (size_t)sp->qoff * Ce, s));
            surr_pred += (int64_t)Q * Ce;
        }
        SPEO_TRY(pick_candidates(isl_dev, srows_dev, Q, Ce, cand, Kp, pred, rho_dev, rho_min, U[0], s));
    } else {
        // candidates generated and scored in one kernel per problem, the pick
        // generated again into U: the same trials without the candidate buffer
        if (any_trained) {
            std::vector<Mlp *> nets;
            std::vector<int> qo, nq, pi;
            for (size_t p = 0; p < surr.size(); ++p)
                if (surr[p]->rows > 0) {
                    nets.push_back(&surr[p]->net);
                    qo.push_back(surr[p]->qoff);
                    nq.push_back(surr[p]->rows);
                    pi.push_back((int)p);
                }
            SPEO_TRY(cand_infer(nets.data(), qo.data(), nq.data(), pi.data(), (int)nets.size(), isl_dev,
                                srows_dev, Ce, X, lo, hi, cfg.seed, gen_base, (uint32_t)gen_offset,
                                rho_dev, rho_min, scored_dev, pred, s));
        }
        SPEO_TRY(pick_regen(isl_dev, srows_dev, Q, Ce, pred, rho_dev, rho_min, X, lo, hi, cfg.seed,
                            gen_base, (uint32_t)gen_offset, U[0], s));
    }
    SPEO_TRY(launch(kModeEval, gen_offset, rt, s));
    SurrHeads h{};
    for (size_t p = 0; p < surr.size(); ++p) {
        h.qoff[p] = surr[p]->qoff;
        h.rows[p] = surr[p]->rows;
    }
    SPEO_TRY(archive_append(isl_dev, srows_dev, Q, U[0], fU[0], AX, Ay, cfg.surrogate_archive, Kp,
                            h, (int)surr.size(), heads_dev, s));
    for (auto &sp : surr) {  // host copies of the device heads
        sp->head = (int)(((int64_t)sp->head + sp->rows) % cfg.surrogate_archive);
        sp->appended += sp->rows;
    }
    return launch(kModeSelect, gen_offset, rt, s);
}

int Engine::surrogate_train(hipStream_t s) {
    const size_t cap = (size_t)cfg.surrogate_archive;
    std::vector<size_t> todo;
    for (size_t p = 0; p < surr.size(); ++p) {
        SurrProblem &sp = *surr[p];
        if (sp.rows == 0) continue;
        const int n_valid = (int)std::min<int64_t>(sp.appended, (int64_t)cap);
        if (n_valid < std::max(cfg.surrogate_batch, 512)) continue;
        SPEO_TRY(mean_std(Ay + p * cap, n_valid, sp.net.stats, s));
        todo.push_back(p);
    }
    auto seed_of = [&](size_t p) { return cfg.seed ^ (0xD1B54A32D192ED03ull * (p + 1)); };
    auto n_valid_of = [&](size_t p) { return (int)std::min<int64_t>(surr[p]->appended, (int64_t)cap); };
    if (!todo.empty() && Kp <= kFusedMaxKp) {
        // every problem's network takes its step in the same two launches
        std::vector<Mlp *> nets;
        std::vector<const float *> ax, ay;
        std::vector<int> nv;
        std::vector<uint64_t> seeds;
        for (size_t p : todo) {
            nets.push_back(&surr[p]->net);
            ax.push_back(AX + p * cap * Kp);
            ay.push_back(Ay + p * cap);
            nv.push_back(n_valid_of(p));
            seeds.push_back(seed_of(p));
        }
        for (int step = 0; step < cfg.surrogate_steps; ++step)
            SPEO_TRY(train_steps(nets.data(), ax.data(), ay.data(), nv.data(), seeds.data(),
                                 (int)todo.size(), round_dev, (uint32_t)step, cfg.surrogate_batch,
                                 cfg.surrogate_lr, step == cfg.surrogate_steps - 1, s));
        for (Mlp *n : nets) SPEO_TRY(add_i32(n->t_dev, cfg.surrogate_steps, s));
    } else {
        for (size_t p : todo)
            for (int step = 0; step < cfg.surrogate_steps; ++step)
                SPEO_TRY(surr[p]->net.train_step(AX + p * cap * Kp, Ay + p * cap, n_valid_of(p),
                                                 seed_of(p), (uint32_t)surr_rounds, (uint32_t)step,
                                                 cfg.surrogate_batch, cfg.surrogate_lr,
                                                 step == cfg.surrogate_steps - 1, Xb, tb, s));
    }
    for (size_t p : todo) {
        SurrProblem &sp = *surr[p];
        if (sp.net.bf16_stale) SPEO_TRY(sp.net.refresh_bf16(s));
        const int H = std::min(1024, n_valid_of(p) / 8);
        SPEO_TRY(gather_holdout(AX + p * cap * Kp, Ay + p * cap, H, Kp, Xh, yh, s));
        SPEO_TRY(sp.net.infer(Xh, Kp, H, predh, s));
        SPEO_TRY(spearman(predh, yh, H, rho_dev + p, ranks_ws, s));
        sp.trained = true;
        any_trained = true;
    }
    ++surr_rounds;
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, s, round_dev, 1u);
    return check_launch("k_add_u32");
}

int Engine::launch(int mode, int gen_offset, int rt_, hipStream_t s, int clear) {
    GenerationArgs a{};
    a.isl = isl_dev;
    a.tiles = tiles_dev[rt_];
    a.probs = probs_dev;
    a.X = X;
    a.lo = cfg.min_bound;
    a.hi = cfg.max_bound;
    a.seed = cfg.seed;
    a.gen_base = gen_base;
    a.gen_offset = (uint32_t)gen_offset;
    a.clear_trials = clear;
    switch (mode) {
    case kModeGen:
        a.U_new = U[0];
        break;
    case kModeEval:
        a.eval_src = U[0];
        a.eval_out = fU[0];
        a.trial_in_lds = all_small ? 1 : 0;
        break;
    case kModeSelect:
        if (cfg.mode == SPEO_MODE_FUSED) {  // apply the pending selection
            a.U_prev = U[p];
            a.fU_prev = fU[p];
            a.fX_new = fX[p];
        } else {
            a.U_prev = U[0];
            a.fU_prev = fU[0];
            a.fX_new = fX[0];
        }
        break;
    default:
        a.fX_prev = fX[p];
        a.U_prev = U[p];
        a.fU_prev = fU[p];
        a.fX_new = fX[1 - p];
        a.U_new = U[1 - p];
        a.fU_new = fU[1 - p];
        a.trial_in_lds = all_small ? 1 : 0;
        break;
    }
    return launch_any(rt_, cfg.precision, mode, a, ntiles[rt_], s);
}

// ------------------------------------------------- sub-islands (plain mode)

int Engine::configure_setup() {
    doff.assign((size_t)I + 1, 0);
    subs_h.assign((size_t)I, 1);
    perm_h.resize((size_t)I);
    info_h.resize((size_t)I);
    base_h.resize((size_t)I);
    for (int i = 0; i < I; ++i) {
        const IslandDesc &d = isl[(size_t)i];
        doff[(size_t)i + 1] = doff[(size_t)i] + d.np / 4;
        subs_h[(size_t)i] = cfg.subisland_size > 0 ? d.np / cfg.subisland_size : 1;
        perm_h[(size_t)i].resize((size_t)d.np);
        for (int r = 0; r < d.np; ++r) perm_h[(size_t)i][(size_t)r] = r;
        info_h[(size_t)i].assign((size_t)(d.np / 4), DEInfo{});
        // DE_CUDA::Initilize: CR 0.9, F 0.5, strategy 4 (DE_CUDA.cc:61-63);
        // here the engine's per-island F, CR and strategy
        base_h[(size_t)i] = DEInfo{d.CR, d.F, cfg.de_strategy ? cfg.de_strategy - 1 : 4, -1};
    }
    SPEO_TRY_HIP(hipMalloc(&perm_dev, sizeof(int) * (size_t)total_rows));
    SPEO_TRY_HIP(hipMalloc(&info_dev, sizeof(DEInfo) * (size_t)std::max(1, doff[(size_t)I])));
    SPEO_TRY_HIP(hipMalloc(&best_sub_dev, sizeof(int) * (size_t)std::max(1, doff[(size_t)I])));
    return regroup(true);
}

// One regroup event of every island (event 0: the initial ConfigureEA),
// then the new partitions and configurations uploaded.
int Engine::regroup(bool initial) {
    const int event = initial ? 0 : ++regroup_events;
    const bool noop = !initial && cfg.regroup_option == SPEO_REGROUP_STATIC_ORDERED;
    if (noop) return SPEO_OK;
    SPEO_TRY_HIP(hipStreamSynchronize(stream));  // the previous upload's host arrays
    for (int i = 0; i < I; ++i) {
        const IslandDesc &d = isl[(size_t)i];
        int subs = subs_h[(size_t)i];
        SPEO_TRY(speo_island_regroup(cfg.seed, i, event, cfg.regroup_option, cfg.configure_EA,
                                     d.np, subs, perm_h[(size_t)i].data(), &base_h[(size_t)i],
                                     info_h[(size_t)i].data(), &subs));
        subs_h[(size_t)i] = subs;
        SPEO_TRY_HIP(hipMemcpy(perm_dev + d.row_off, perm_h[(size_t)i].data(), sizeof(int) * d.np,
                               hipMemcpyHostToDevice));
        SPEO_TRY_HIP(hipMemcpy(info_dev + doff[(size_t)i], info_h[(size_t)i].data(),
                               sizeof(DEInfo) * (size_t)subs, hipMemcpyHostToDevice));
    }
    return SPEO_OK;
}

// The generation of every island through the reference's operator kernel
// (sub-islands, permutation, per sub-island DEInfo, strategies 0-8), one
// launch per island; strategies that use the best member get each
// sub-island's best slot first (global_DE_FindBestIndividualInIsland).
int Engine::generate_configured(int gen_offset, hipStream_t s) {
    const uint32_t gen = (uint32_t)(gens + gen_offset);
    for (int i = 0; i < I; ++i) {
        const IslandDesc &d = isl[(size_t)i];
        const int subs = subs_h[(size_t)i], S = d.np / subs;
        bool need_best = false;
        for (int q = 0; q < subs; ++q) {
            const int st = info_h[(size_t)i][(size_t)q].strategy_ID;
            need_best = need_best || st == 2 || st == 3 || st == 6 || st == 7;
        }
        int *best = best_sub_dev + doff[(size_t)i];
        if (need_best) SPEO_TRY(subisland_best(fX[0] + d.row_off, perm_dev + d.row_off, S, subs, best, s));
        GenArgs a{U[0] + d.elem_off, X + d.elem_off, d.ld, d.np, d.dim, S, perm_dev + d.row_off,
                  info_dev + doff[(size_t)i], need_best ? best : nullptr, cfg.min_bound,
                  cfg.max_bound, cfg.seed, gen, (uint32_t)d.row_off, nullptr, 0, nullptr, nullptr,
                  nullptr};
        SPEO_TRY(de_generate(a, s));
    }
    return SPEO_OK;
}

int Engine::subislands(int island, int *subs_out, int *perm_out, DEInfo *info_out) {
    if (island < 0 || island >= I) return set_error(SPEO_ERR_ARG, "speo_engine_subislands: bad island");
    const int np = isl[(size_t)island].np;
    if (!configured) {
        if (subs_out) *subs_out = 1;
        if (perm_out)
            for (int r = 0; r < np; ++r) perm_out[r] = r;
        if (info_out) info_out[0] = DEInfo{isl[(size_t)island].CR, isl[(size_t)island].F, 4, np};
        return SPEO_OK;
    }
    SPEO_TRY_HIP(hipStreamSynchronize(stream));
    const int subs = subs_h[(size_t)island];
    if (subs_out) *subs_out = subs;
    if (perm_out) std::copy(perm_h[(size_t)island].begin(), perm_h[(size_t)island].end(), perm_out);
    if (info_out)
        for (int q = 0; q < subs; ++q) info_out[q] = info_h[(size_t)island][(size_t)q];
    return SPEO_OK;
}

// Islands whose problem the generic kernel evaluates: one launch per island.
int Engine::eval_generic_islands(const float *src, float *out, hipStream_t s) {
    if (!any_generic) return SPEO_OK;
    for (int i = 0; i < I; ++i) {
        const IslandDesc &d = isl[(size_t)i];
        const ProblemHost &P = problems[(size_t)d.problem]->h;
        if (!P.generic) continue;
        SPEO_TRY(eval_problem_generic(P, src + d.elem_off, d.ld, d.np, out + d.row_off, nullptr,
                                      nullptr, 0, s));
    }
    return SPEO_OK;
}

int Engine::finalize_pending(hipStream_t s) {
    if (cfg.mode != SPEO_MODE_FUSED || !pending) return SPEO_OK;
    SPEO_TRY(launch(kModeSelect, 0, rt, s, 1));
    pending = false;
    return SPEO_OK;
}

int Engine::migrate(hipStream_t s) {
    if (cfg.migration_size <= 0) return SPEO_OK;
    SPEO_TRY(finalize_pending(s));
    const int cur = cfg.mode == SPEO_MODE_FUSED ? p : 0;
    return migrate_ring(X, fX[cur], isl_dev, row_off_dev, dst_dev, best_buf, worst_buf, I,
                        cfg.migration_size, s);
}

// L generations, then (optionally) migration, then gen_base += L.  The parity
// `p` and `pending` advance on the host as the launches are enqueued.
int Engine::enqueue_chunk(int L, bool do_migrate, hipStream_t s) {
    if (cfg.mode == SPEO_MODE_EPOCH) {
        SPEO_TRY(launch_epoch_chunk(L, s));
    } else {
        for (int g = 0; g < L; ++g) {
            if (C > 0) {
                SPEO_TRY(surrogate_generation(g, s));
            } else if (cfg.mode == SPEO_MODE_FUSED) {
                SPEO_TRY(launch(kModeFused, g, rt, s));
                p = 1 - p;
                pending = true;
            } else {
                if (configured)
                    SPEO_TRY(generate_configured(g, s));
                else
                    SPEO_TRY(launch(kModeGen, g, rt, s));
                SPEO_TRY(launch(kModeEval, g, rt, s));
                SPEO_TRY(eval_generic_islands(U[0], fU[0], s));
                SPEO_TRY(launch(kModeSelect, g, rt, s));
            }
        }
    }
    if (do_migrate) SPEO_TRY(migrate(s));
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, s, gen_base, (uint32_t)L);
    return check_launch("k_add_u32");
}

// Generations are counted from 0; migration follows generation g when
// g % interval == 0 (island_EA.cc:93-99).  Returns the generations up to and
// including the next migration point, or `left` when it lies beyond.
int Engine::next_chunk(int left, bool *mig) const {
    *mig = false;
    if (cfg.interval > 0 && cfg.migration_size > 0) {
        const int64_t g0 = gens;
        const int64_t next = (g0 % cfg.interval == 0) ? g0 : (g0 / cfg.interval + 1) * cfg.interval;
        const int64_t to_mig = next - g0 + 1;
        if (to_mig <= left) {
            *mig = true;
            return (int)to_mig;
        }
    }
    return left;
}

// A migration point of a timed run: the work before it completes, then the
// migration (and, for configured engines, the regroup) runs eagerly and its
// wall time until completion is added to comm_seconds.
int Engine::timed_point(bool mig, bool point) {
    SPEO_TRY_HIP(hipStreamSynchronize(stream));
    const auto t0 = std::chrono::steady_clock::now();
    if (mig) SPEO_TRY(migrate(stream));
    if (point && configured) SPEO_TRY(regroup(false));
    SPEO_TRY_HIP(hipStreamSynchronize(stream));
    comm_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return SPEO_OK;
}

int Engine::prepare(int generations) {
    if (C > 0 || !cfg.use_graph || graph_failed) return SPEO_OK;
    const int64_t g0 = gens, f0 = fes, m0 = migrations, gl0 = graph_launches;
    const int p0 = p, ep0 = ep_par;
    const bool pend0 = pending;
    const int rc = run(generations, true);
    gens = g0;
    fes = f0;
    migrations = m0;
    graph_launches = gl0;
    p = p0;
    pending = pend0;
    if (ep_par != ep0) swap_epoch_buffers();
    return rc;
}

// A surrogate round: the train_interval generations up to a training point
// (migration after generation mig_pos when >= 0), the generation counter, then
// the training.  The same launches surrogate generations and training enqueue
// one by one, with the generations' counters offset within the round.
int Engine::enqueue_surrogate_round(int TI, int mig_pos, hipStream_t s) {
    for (int k = 0; k < TI; ++k) {
        SPEO_TRY(surrogate_generation(k, s));
        if (k == mig_pos) SPEO_TRY(migrate(s));
    }
    hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, s, gen_base, (uint32_t)TI);
    SPEO_TRY(check_launch("k_add_u32"));
    return surrogate_train(s);
}

// The host-side counters a replayed surrogate round advances (a captured
// round advanced them once while it was enqueued).
void Engine::surrogate_round_host(int TI) {
    const int64_t cap = cfg.surrogate_archive;
    for (int k = 0; k < TI; ++k) {
        last_Ce = any_trained ? C : 1;
        // (the one-kernel screening counts its candidates on the device)
        if (any_trained && (retain_cand || Kp > kFusedMaxKp)) surr_pred += (int64_t)Q * last_Ce;
        for (auto &sp : surr) {
            sp->head = (int)(((int64_t)sp->head + sp->rows) % cap);
            sp->appended += sp->rows;
        }
    }
    ++surr_rounds;
}

// Rounds are replayable once every network screens and every archive is full:
// then no launch argument changes from one round to the next.
bool Engine::surrogate_round_replayable() const {
    if (!any_trained || Kp > kFusedMaxKp || cfg.surrogate_archive < std::max(cfg.surrogate_batch, 512))
        return false;
    for (const auto &sp : surr)
        if (sp->rows > 0 && (!sp->trained || sp->appended < cfg.surrogate_archive)) return false;
    return true;
}

int Engine::run(int generations, bool dry) {
    if (C > 0) {
        const int TI = cfg.surrogate_train_interval;
        const bool mig_on = cfg.interval > 0 && cfg.migration_size > 0;
        int left = generations;
        while (left > 0) {
            if (cfg.use_graph && !graph_failed && !dry && gens % TI == 0 && left >= TI &&
                surrogate_round_replayable()) {
                int mig_pos = -1, nmig = 0;
                for (int k = 0; k < TI && mig_on; ++k)
                    if ((gens + k) % cfg.interval == 0) {
                        mig_pos = k;
                        ++nmig;
                    }
                if (nmig <= 1) {
                    const auto key = std::make_tuple(-TI, mig_pos, retain_cand ? 1 : 0);
                    auto it = graphs.find(key);
                    bool captured = false;
                    if (it == graphs.end()) {
                        hipGraph_t g = nullptr;
                        hipGraphExec_t ge = nullptr;
                        SPEO_TRY_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                        const int rc = enqueue_surrogate_round(TI, mig_pos, stream);
                        hipError_t e = hipStreamEndCapture(stream, &g);
                        if (rc != SPEO_OK || e != hipSuccess ||
                            hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
                            // the host counters advanced during the attempt: this
                            // engine cannot continue consistently
                            if (g) (void)hipGraphDestroy(g);
                            (void)hipGetLastError();
                            graph_failed = true;
                            return rc != SPEO_OK ? rc
                                                 : set_error(SPEO_ERR_HIP, "surrogate round capture failed");
                        }
                        (void)hipGraphDestroy(g);
                        it = graphs.emplace(key, ge).first;
                        captured = true;
                    }
                    SPEO_TRY_HIP(hipGraphLaunch(it->second, stream));
                    ++graph_launches;
                    if (!captured) surrogate_round_host(TI);
                    gens += TI;
                    fes += (int64_t)TI * total_rows;
                    migrations += nmig;
                    left -= TI;
                    continue;
                }
            }
            const bool mig = mig_on && gens % cfg.interval == 0;
            SPEO_TRY(enqueue_chunk(1, mig, stream));
            ++gens;
            fes += total_rows;
            if (mig) ++migrations;
            if (gens % TI == 0) SPEO_TRY(surrogate_train(stream));
            --left;
        }
        return SPEO_OK;
    }
    if (configured && !dry) {
        // generation by generation: after generation g with g % interval == 0
        // the ring migration and then the regroup of every island
        // (island_EA.cc:94-99; migrate.cc:59)
        for (int g = 0; g < generations; ++g) {
            const bool point = cfg.interval > 0 && gens % cfg.interval == 0;
            const bool mig = point && cfg.migration_size > 0;
            SPEO_TRY(enqueue_chunk(1, mig && !time_comm, stream));
            ++gens;
            fes += total_rows;
            if (mig) ++migrations;
            if (time_comm && point) {
                SPEO_TRY(timed_point(mig, true));
            } else if (point) {
                SPEO_TRY(regroup(false));
            }
        }
        return SPEO_OK;
    }
    int left = generations;
    while (left > 0) {
        bool mig_point = false;
        const int L = next_chunk(left, &mig_point);
        const bool mig = mig_point && !(time_comm && !dry);  // timed: migrate after the chunk
        const bool use_graph = cfg.use_graph && !graph_failed;
        if (use_graph) {
            const auto key = std::make_tuple(L, cfg.mode == SPEO_MODE_FUSED ? p : ep_par,
                                             (mig ? 1 : 0) | (pending ? 2 : 0));
            auto it = graphs.find(key);
            const int p_before = p;
            const bool pend_before = pending;
            const int ep_before = ep_par;
            if (it == graphs.end()) {
                hipGraph_t g = nullptr;
                hipGraphExec_t ge = nullptr;
                SPEO_TRY_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
                const int rc = enqueue_chunk(L, mig, stream);
                hipError_t e = hipStreamEndCapture(stream, &g);
                if (rc != SPEO_OK || e != hipSuccess ||
                    hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
                    if (g) (void)hipGraphDestroy(g);
                    (void)hipGetLastError();
                    graph_failed = true;  // fall back to eager launches
                    p = p_before;
                    pending = pend_before;
                    if (ep_par != ep_before) swap_epoch_buffers();
                    continue;
                }
                (void)hipGraphDestroy(g);
                it = graphs.emplace(key, ge).first;
                // enqueue_chunk advanced the host-side parity during capture
                p = p_before;
                pending = pend_before;
                if (ep_par != ep_before) swap_epoch_buffers();
            }
            if (!dry) SPEO_TRY_HIP(hipGraphLaunch(it->second, stream));
            ++graph_launches;
            if (cfg.mode == SPEO_MODE_FUSED) {
                if (L % 2) p = 1 - p;
                pending = L > 0 ? true : pending;
                if (mig && cfg.migration_size > 0) pending = false;
            }
            if (cfg.mode == SPEO_MODE_EPOCH && epoch_parts > 1 && L % 2) swap_epoch_buffers();
        } else if (!dry) {
            SPEO_TRY(enqueue_chunk(L, mig, stream));
        }
        if (mig_point && !mig) SPEO_TRY(timed_point(true, false));
        gens += L;
        fes += (int64_t)L * total_rows;
        if (mig_point) ++migrations;
        left -= L;
    }
    return SPEO_OK;
}

int Engine::profile(int generations, double *avg_ms) {
    if (cfg.mode == SPEO_MODE_EPOCH) return profile_epoch(generations, avg_ms);
    const int roles = 6;
    std::vector<hipEvent_t> ev;
    std::vector<int> role;
    auto record = [&](int r) -> int {
        hipEvent_t a, b;
        SPEO_TRY_HIP(hipEventCreate(&a));
        SPEO_TRY_HIP(hipEventCreate(&b));
        ev.push_back(a);
        ev.push_back(b);
        role.push_back(r);
        return SPEO_OK;
    };
    for (int g = 0; g < generations; ++g) {
        const int64_t gi = gens;
        const int modes[3] = {kModeGen, kModeEval, kModeSelect};
        if (cfg.mode == SPEO_MODE_FUSED) {
            SPEO_TRY(record(3));
            SPEO_TRY_HIP(hipEventRecord(ev[ev.size() - 2], stream));
            SPEO_TRY(launch(kModeFused, 0, rt, stream));
            SPEO_TRY_HIP(hipEventRecord(ev.back(), stream));
            p = 1 - p;
            pending = true;
        } else {
            for (int m = 0; m < 3; ++m) {
                SPEO_TRY(record(m));
                SPEO_TRY_HIP(hipEventRecord(ev[ev.size() - 2], stream));
                if (modes[m] == kModeGen && configured)
                    SPEO_TRY(generate_configured(0, stream));
                else
                    SPEO_TRY(launch(modes[m], 0, rt, stream));
                if (modes[m] == kModeEval) SPEO_TRY(eval_generic_islands(U[0], fU[0], stream));
                SPEO_TRY_HIP(hipEventRecord(ev.back(), stream));
            }
        }
        const bool point = cfg.interval > 0 && gi % cfg.interval == 0;
        const bool mig = point && cfg.migration_size > 0;
        if (mig) SPEO_TRY(migrate(stream));
        hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, stream, gen_base, 1u);
        SPEO_TRY(check_launch("k_add_u32"));
        ++gens;
        fes += total_rows;
        if (mig) ++migrations;
        if (point && configured) SPEO_TRY(regroup(false));
    }
    SPEO_TRY_HIP(hipStreamSynchronize(stream));
    double sum[roles] = {0, 0, 0, 0, 0, 0};
    int cnt[roles] = {0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < role.size(); ++i) {
        float ms = 0.f;
        SPEO_TRY_HIP(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        sum[role[i]] += ms;
        cnt[role[i]] += 1;
    }
    for (auto e : ev) (void)hipEventDestroy(e);
    for (int r = 0; r < roles; ++r) avg_ms[r] = cnt[r] ? sum[r] / cnt[r] : 0.0;
    return SPEO_OK;
}

// Epoch mode: epochs on the run schedule, HIP events around every epoch
// launch; role 4 is the launch duration per generation.
int Engine::prof_begin(int role, int gens_covered, hipStream_t s) {
    if (!prof) return SPEO_OK;
    ProfEvent p{role, nullptr, nullptr, gens_covered};
    SPEO_TRY_HIP(hipEventCreate(&p.a));
    SPEO_TRY_HIP(hipEventCreate(&p.b));
    prof->push_back(p);
    SPEO_TRY_HIP(hipEventRecord(p.a, s));
    return SPEO_OK;
}

int Engine::prof_end(hipStream_t s) {
    if (!prof) return SPEO_OK;
    SPEO_TRY_HIP(hipEventRecord(prof->back().b, s));
    return SPEO_OK;
}

int Engine::profile_epoch(int generations, double *avg_ms) {
    for (int r = 0; r < 6; ++r) avg_ms[r] = 0.0;
    std::vector<ProfEvent> events;
    prof = &events;
    int64_t total_gens = 0;
    int left = generations, rc = SPEO_OK;
    while (left > 0 && rc == SPEO_OK) {
        bool mig = false;
        const int L = next_chunk(left, &mig);
        if ((rc = launch_epoch_chunk(L, stream)) != SPEO_OK) break;
        if (mig && (rc = migrate(stream)) != SPEO_OK) break;
        hipLaunchKernelGGL(k_add_u32, dim3(1), dim3(64), 0, stream, gen_base, (uint32_t)L);
        if ((rc = check_launch("k_add_u32")) != SPEO_OK) break;
        total_gens += L;
        gens += L;
        fes += (int64_t)L * total_rows;
        if (mig) ++migrations;
        left -= L;
    }
    prof = nullptr;
    if (rc == SPEO_OK) rc = check_hip(hipStreamSynchronize(stream), "hipStreamSynchronize");
    double sum[6] = {0, 0, 0, 0, 0, 0};
    for (const ProfEvent &p : events) {
        float ms = 0.f;
        if (rc == SPEO_OK) rc = check_hip(hipEventElapsedTime(&ms, p.a, p.b), "hipEventElapsedTime");
        sum[p.role] += ms;
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    if (total_gens > 0)
        for (int r = 4; r < 6; ++r) avg_ms[r] = sum[r] / (double)total_gens;
    return rc;
}

// Times the generation kernel with each row tile on scratch copies of the
// state and keeps the fastest.  Results are cached per configuration shape.
int Engine::autotune() {
    static std::mutex mu;
    static std::map<std::vector<int>, int> cache;
    std::vector<int> key{cfg.precision, cfg.mode};
    for (const auto &d : isl) {
        key.push_back(d.np);
        key.push_back(d.dim);
    }
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            rt = it->second;
            return SPEO_OK;
        }
    }
    const std::string pkey = tune_key("rt");
    int stored = 0;
    if (tune_lookup(pkey, &stored) && (stored == 16 || stored == 32 || stored == 64)) {
        rt = stored;
        tuned_from_cache |= 1;
        std::lock_guard<std::mutex> lk(mu);
        cache[key] = rt;
        return SPEO_OK;
    }
    float *saveX = X, *savefX[2] = {fX[0], fX[1]}, *saveU[2] = {U[0], U[1]},
          *savefU[2] = {fU[0], fU[1]};
    const size_t E = (size_t)total_elems, R = (size_t)total_rows;
    float *tX = nullptr, *tf[4] = {nullptr, nullptr, nullptr, nullptr}, *tU[2] = {nullptr, nullptr};
    SPEO_TRY_HIP(hipMalloc(&tX, sizeof(float) * E));
    SPEO_TRY_HIP(hipMalloc(&tU[0], sizeof(float) * E));
    SPEO_TRY_HIP(hipMalloc(&tU[1], sizeof(float) * E));
    for (auto &b : tf) SPEO_TRY_HIP(hipMalloc(&b, sizeof(float) * R));
    SPEO_TRY_HIP(hipMemcpyAsync(tX, saveX, sizeof(float) * E, hipMemcpyDeviceToDevice, stream));
    SPEO_TRY_HIP(hipMemsetAsync(tU[0], 0, sizeof(float) * E, stream));
    SPEO_TRY_HIP(hipMemsetAsync(tU[1], 0, sizeof(float) * E, stream));
    for (auto &b : tf) {
        hipLaunchKernelGGL(k_fill, dim3(256), dim3(256), 0, stream, b, total_rows, INFINITY);
        SPEO_TRY(check_launch("k_fill"));
    }
    X = tX;
    fX[0] = tf[0];
    fX[1] = tf[1];
    fU[0] = tf[2];
    fU[1] = tf[3];
    U[0] = tU[0];
    U[1] = tU[1];
    hipEvent_t e0, e1;
    SPEO_TRY_HIP(hipEventCreate(&e0));
    SPEO_TRY_HIP(hipEventCreate(&e1));
    float best_ms = 1e30f;
    int best_rt = 32;
    const int saved_p = p;
    for (int cand : {16, 32, 64}) {
        SPEO_TRY(build_tiles(cand));
        const int mode = cfg.mode == SPEO_MODE_FUSED ? kModeFused : kModeEval;
        for (int rep = 0; rep < 2; ++rep) SPEO_TRY(launch(mode, rep, cand, stream));
        SPEO_TRY_HIP(hipEventRecord(e0, stream));
        const int reps = 6;
        for (int rep = 0; rep < reps; ++rep) {
            SPEO_TRY(launch(mode, rep, cand, stream));
            if (mode == kModeFused) p = 1 - p;
        }
        SPEO_TRY_HIP(hipEventRecord(e1, stream));
        SPEO_TRY_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        SPEO_TRY_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best_ms) {
            best_ms = ms;
            best_rt = cand;
        }
    }
    p = saved_p;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    X = saveX;
    fX[0] = savefX[0];
    fX[1] = savefX[1];
    U[0] = saveU[0];
    U[1] = saveU[1];
    fU[0] = savefU[0];
    fU[1] = savefU[1];
    (void)hipFree(tX);
    (void)hipFree(tU[0]);
    (void)hipFree(tU[1]);
    for (auto &b : tf) (void)hipFree(b);
    rt = best_rt;
    tune_store(pkey, rt);
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = rt;
    return SPEO_OK;
}

std::string Engine::tune_key(const char *what) const {
    std::vector<std::pair<int, int>> dn;
    for (const auto &d : isl) {
        int b = 1;
        while (b < d.np) b <<= 1;
        dn.emplace_back(d.dim, b);
    }
    std::sort(dn.begin(), dn.end());
    std::ostringstream k;
    k << what << ":m" << cfg.mode << ":p" << cfg.precision << ":n" << I;
    for (size_t i = 0; i < dn.size();) {  // runs of equal (dim, bucket) as dim/bucket*count
        size_t j = i;
        while (j < dn.size() && dn[j] == dn[i]) ++j;
        k << ':' << dn[i].first << '/' << dn[i].second << '*' << (j - i);
        i = j;
    }
    return k.str();
}

// Workgroups per island for epoch mode: every feasible power of two up to 8
// (islands x parts at most four per compute unit) timed on a throwaway
// engine of this configuration over 24 generations after a warm-up that
// captures the graphs; the fastest is kept and persisted.
int Engine::tune_epoch_parts() {
    const std::string key = tune_key("epoch_parts");
    int stored = 0;
    if (tune_lookup(key, &stored) && (stored == 1 || stored == 2 || stored == 4 || stored == 8)) {
        cfg.epoch_parts = stored;
        tuned_from_cache |= 2;
        return SPEO_OK;
    }
    int dev = 0, cus = 0;
    SPEO_TRY_HIP(hipGetDevice(&dev));
    SPEO_TRY_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    std::vector<int> cands;
    for (int parts = 1; parts <= kEpochMaxParts; parts *= 2) {
        std::vector<EpochPlan> plans;
        int lds = 0;
        if ((int64_t)I * parts > 4 * (int64_t)cus) break;
        if (epoch_plan_all(parts, &plans, &lds) == SPEO_OK) cands.push_back(parts);
    }
    if (cands.empty()) return SPEO_OK;  // epoch_setup reports the reason
    int best = cands[0];
    if (cands.size() > 1) {
        double best_s = 1e30;
        for (int parts : cands) {
            speo_engine_config c2 = cfg;
            c2.epoch_parts = parts;
            c2.use_graph = 1;
            c2.rows_per_block = cfg.rows_per_block ? cfg.rows_per_block : 32;
            auto child = std::make_unique<Engine>();
            child->tuning_child = true;
            SPEO_TRY(child->create(c2, stream));
            SPEO_TRY(child->run(24));
            SPEO_TRY_HIP(hipStreamSynchronize(stream));
            const auto t0 = std::chrono::steady_clock::now();
            SPEO_TRY(child->run(24));
            SPEO_TRY_HIP(hipStreamSynchronize(stream));
            const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (dt < best_s) {
                best_s = dt;
                best = parts;
            }
        }
    }
    cfg.epoch_parts = best;
    tune_store(key, best);
    return SPEO_OK;
}

int Engine::best(int *row_out, float *f32_out, double *f64_out) {
    SPEO_TRY(finalize_pending(stream));
    const int cur = cfg.mode == SPEO_MODE_FUSED ? p : 0;
    SPEO_TRY(segment_argmin(fX[cur], row_off_dev, I, best_row_dev, stream));
    const size_t lds = sizeof(double) * ((size_t)max_dim + 16);
    if (lds > 48 * 1024)
        SPEO_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eval_f64_rows),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_eval_f64_rows, dim3(I), dim3(256), lds, stream, params_tab_dev,
                       ncomp_dev, dims_dev, isl_dev, best_row_dev, X, best_f64_dev);
    SPEO_TRY(check_launch("k_eval_f64_rows"));
    for (int i = 0; i < I && any_generic; ++i) {  // generic problems: the FP64 generic kernel
        const IslandDesc &d = isl[(size_t)i];
        const ProblemHost &P = problems[(size_t)d.problem]->h;
        if (P.generic)
            SPEO_TRY(eval_problem_generic(P, X + d.elem_off, d.ld, 1, nullptr, best_f64_dev + i,
                                          best_row_dev + i, d.row_off, stream));
    }
    std::vector<int> rows((size_t)I);
    std::vector<double> f64((size_t)I);
    std::vector<float> fall((size_t)total_rows);
    SPEO_TRY_HIP(hipMemcpyAsync(rows.data(), best_row_dev, sizeof(int) * I, hipMemcpyDeviceToHost, stream));
    SPEO_TRY_HIP(hipMemcpyAsync(f64.data(), best_f64_dev, sizeof(double) * I, hipMemcpyDeviceToHost, stream));
    if (f32_out)
        SPEO_TRY_HIP(hipMemcpyAsync(fall.data(), fX[cur], sizeof(float) * total_rows,
                                    hipMemcpyDeviceToHost, stream));
    SPEO_TRY_HIP(hipStreamSynchronize(stream));
    for (int i = 0; i < I; ++i) {
        if (row_out) row_out[i] = rows[(size_t)i];
        if (f32_out) f32_out[i] = fall[(size_t)rows[(size_t)i]];
        if (f64_out) f64_out[i] = f64[(size_t)i];
    }
    return SPEO_OK;
}

int Engine::read(float *Xo, float *fo) {
    SPEO_TRY(finalize_pending(stream));
    const int cur = cfg.mode == SPEO_MODE_FUSED ? p : 0;
    return read_rows(X, fX[cur], Xo, fo);
}

int Engine::read_rows(const float *src, const float *fsrc, float *Xo, float *fo) {
    if (Xo) {
        std::vector<float> h((size_t)total_elems);
        SPEO_TRY_HIP(hipMemcpyAsync(h.data(), src, sizeof(float) * h.size(), hipMemcpyDeviceToHost, stream));
        SPEO_TRY_HIP(hipStreamSynchronize(stream));
        for (const auto &d : isl)
            for (int r = 0; r < d.np; ++r)
                for (int j = 0; j < max_dim; ++j)
                    Xo[(size_t)(d.row_off + r) * max_dim + j] =
                        j < d.dim ? h[(size_t)d.elem_off + (size_t)r * d.ld + j] : 0.f;
    }
    if (fo) {
        SPEO_TRY_HIP(hipMemcpyAsync(fo, fsrc, sizeof(float) * total_rows, hipMemcpyDeviceToHost, stream));
        SPEO_TRY_HIP(hipStreamSynchronize(stream));
    }
    return SPEO_OK;
}

int Engine::write(const float *Xi, const float *fi) {
    SPEO_TRY(finalize_pending(stream));
    const int cur = cfg.mode == SPEO_MODE_FUSED ? p : 0;
    if (Xi) {
        std::vector<float> h((size_t)total_elems, 0.f);
        for (const auto &d : isl)
            for (int r = 0; r < d.np; ++r)
                for (int j = 0; j < d.dim; ++j)
                    h[(size_t)d.elem_off + (size_t)r * d.ld + j] = Xi[(size_t)(d.row_off + r) * max_dim + j];
        SPEO_TRY_HIP(hipMemcpyAsync(X, h.data(), sizeof(float) * h.size(), hipMemcpyHostToDevice, stream));
        SPEO_TRY_HIP(hipStreamSynchronize(stream));
    }
    if (fi) {
        SPEO_TRY_HIP(hipMemcpyAsync(fX[cur], fi, sizeof(float) * total_rows, hipMemcpyHostToDevice, stream));
        SPEO_TRY_HIP(hipStreamSynchronize(stream));
    }
    return SPEO_OK;
}

}  // namespace speo

struct speo_engine_s {
    speo::Engine e;
};

using namespace speo;

extern "C" {

int speo_engine_create(const speo_engine_config *cfg, speo_stream_t stream, speo_engine *out) {
    if (!cfg || !out) return set_error(SPEO_ERR_ARG, "speo_engine_create: null argument");
    *out = nullptr;
    auto *h = new (std::nothrow) speo_engine_s;
    if (!h) return set_error(SPEO_ERR_ALLOC, "speo_engine_create: out of host memory");
    const int rc = h->e.create(*cfg, as_stream(stream));
    if (rc != SPEO_OK) {
        const std::string msg = speo_last_error();
        delete h;
        return set_error(rc, msg);
    }
    *out = h;
    return SPEO_OK;
}

int speo_engine_destroy(speo_engine e) {
    delete e;
    return SPEO_OK;
}

int speo_engine_run(speo_engine e, int generations) {
    if (!e || generations < 0) return set_error(SPEO_ERR_ARG, "speo_engine_run: bad arguments");
    return e->e.run(generations);
}

// IslandEA::Execute's budget loop (src/island_EA.cc:86-104) over the engine's
// evaluation count; speo_engine_run splits the generations at the migration
// points as the reference migrates at generation % interval == 0.
int speo_engine_run_fes(speo_engine e, int64_t max_FEs, int *generations_out) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_run_fes: null engine");
    const int64_t g = speo_generations_for_fes(max_FEs, e->e.fes, e->e.total_rows);
    if (g > 0x7fffffff) return set_error(SPEO_ERR_ARG, "speo_engine_run_fes: budget beyond 2^31 generations");
    if (generations_out) *generations_out = (int)g;
    return e->e.run((int)g);
}

int speo_engine_run_fes_timed(speo_engine e, int64_t max_FEs, int *generations_out,
                              double *seconds_out, double *comm_seconds_out) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_run_fes_timed: null engine");
    Engine &E = e->e;
    const int64_t g = speo_generations_for_fes(max_FEs, E.fes, E.total_rows);
    if (g > 0x7fffffff) return set_error(SPEO_ERR_ARG, "speo_engine_run_fes_timed: budget beyond 2^31 generations");
    if (generations_out) *generations_out = (int)g;
    SPEO_TRY_HIP(hipStreamSynchronize(E.stream));
    E.time_comm = true;
    E.comm_seconds = 0.0;
    const auto t0 = std::chrono::steady_clock::now();
    int rc = E.run((int)g);
    if (rc == SPEO_OK) rc = check_hip(hipStreamSynchronize(E.stream), "speo_engine_run_fes_timed");
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    E.time_comm = false;
    if (seconds_out) *seconds_out = dt;
    if (comm_seconds_out) *comm_seconds_out = E.comm_seconds;
    return rc;
}

int speo_engine_subislands(speo_engine e, int island, int *subisland_num_out, int *perm_out,
                           DEInfo *info_out) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_subislands: null engine");
    return e->e.subislands(island, subisland_num_out, perm_out, info_out);
}

int speo_engine_prepare(speo_engine e, int generations) {
    if (!e || generations < 0) return set_error(SPEO_ERR_ARG, "speo_engine_prepare: bad arguments");
    return e->e.prepare(generations);
}

int speo_engine_sync(speo_engine e) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_sync: null engine");
    return check_hip(hipStreamSynchronize(e->e.stream), "speo_engine_sync");
}

int speo_engine_best(speo_engine e, int *row_out, float *f32_out, double *f64_out) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_best: null engine");
    return e->e.best(row_out, f32_out, f64_out);
}

int speo_engine_read(speo_engine e, float *X_out, float *f_out) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_read: null engine");
    return e->e.read(X_out, f_out);
}

int speo_engine_read_trials(speo_engine e, float *U_out, float *fU_out) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_read_trials: null engine");
    const int cur = e->e.cfg.mode == SPEO_MODE_FUSED ? e->e.p : 0;
    return e->e.read_rows(e->e.U[cur], e->e.fU[cur], U_out, fU_out);
}

int speo_engine_write(speo_engine e, const float *X_in, const float *f_in) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_write: null engine");
    return e->e.write(X_in, f_in);
}

int speo_engine_get_stats(speo_engine e, speo_engine_stats *out) {
    if (!e || !out) return set_error(SPEO_ERR_ARG, "speo_engine_get_stats: null argument");
    out->generations = e->e.gens;
    out->evaluations = e->e.fes;
    out->migrations = e->e.migrations;
    out->graph_launches = e->e.graph_launches;
    out->rows_per_block = e->e.rt;
    out->surrogate_predictions = e->e.surr_pred;
    if (e->e.scored_dev) {
        // the one-kernel screening's count lives on the device
        const size_t P = (size_t)e->e.cfg.num_problems;
        std::vector<long long> scored(P, 0);
        SPEO_TRY_HIP(hipStreamSynchronize(e->e.stream));
        SPEO_TRY_HIP(hipMemcpy(scored.data(), e->e.scored_dev, sizeof(long long) * P, hipMemcpyDeviceToHost));
        for (long long v : scored) out->surrogate_predictions += v;
    }
    out->surrogate_rounds = e->e.surr_rounds;
    out->epoch_parts = e->e.cfg.mode == SPEO_MODE_EPOCH ? e->e.epoch_parts : 0;
    return SPEO_OK;
}

int speo_engine_surrogate_stats(speo_engine e, float *rho_out, int *active_out, float *loss_out) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_surrogate_stats: null engine");
    Engine &E = e->e;
    const int P = E.cfg.num_problems;
    std::vector<float> rho((size_t)P, -1.f), loss((size_t)P, 0.f);
    if (E.C > 0) {
        SPEO_TRY_HIP(hipStreamSynchronize(E.stream));
        SPEO_TRY_HIP(hipMemcpy(rho.data(), E.rho_dev, sizeof(float) * P, hipMemcpyDeviceToHost));
        for (int p = 0; p < P; ++p)
            if (E.surr[(size_t)p]->rows > 0 && E.surr[(size_t)p]->trained)
                SPEO_TRY_HIP(hipMemcpy(&loss[(size_t)p], E.surr[(size_t)p]->net.stats, sizeof(float),
                                       hipMemcpyDeviceToHost));
    }
    for (int p = 0; p < P; ++p) {
        if (rho_out) rho_out[p] = rho[(size_t)p];
        if (active_out)
            active_out[p] = (E.C > 0 && E.surr[(size_t)p]->trained && rho[(size_t)p] >= E.cfg.surrogate_rho_min) ? 1 : 0;
        if (loss_out) loss_out[p] = loss[(size_t)p];
    }
    return SPEO_OK;
}

int speo_engine_surrogate_dump(speo_engine e, int problem, int *cands_out, float *cand_out,
                               float *pred_out, float *archive_X, float *archive_y, int *head_out,
                               int64_t *appended_out) {
    if (!e || e->e.C <= 0 || problem < 0 || problem >= e->e.cfg.num_problems)
        return set_error(SPEO_ERR_ARG, "speo_engine_surrogate_dump: no surrogate or bad problem");
    Engine &E = e->e;
    const auto &sp = *E.surr[(size_t)problem];
    const size_t Kp = (size_t)E.Kp, Ce = (size_t)E.last_Ce, cap = (size_t)E.cfg.surrogate_archive;
    SPEO_TRY_HIP(hipStreamSynchronize(E.stream));
    if (cands_out) *cands_out = E.last_Ce;
    if (cand_out)
        SPEO_TRY_HIP(hipMemcpy(cand_out, E.cand + (size_t)sp.qoff * Ce * Kp,
                               sizeof(float) * (size_t)sp.rows * Ce * Kp, hipMemcpyDeviceToHost));
    if (pred_out)
        SPEO_TRY_HIP(hipMemcpy(pred_out, E.pred + (size_t)sp.qoff * Ce, sizeof(float) * (size_t)sp.rows * Ce,
                               hipMemcpyDeviceToHost));
    if (archive_X)
        SPEO_TRY_HIP(hipMemcpy(archive_X, E.AX + (size_t)problem * cap * Kp, sizeof(float) * cap * Kp,
                               hipMemcpyDeviceToHost));
    if (archive_y)
        SPEO_TRY_HIP(hipMemcpy(archive_y, E.Ay + (size_t)problem * cap, sizeof(float) * cap,
                               hipMemcpyDeviceToHost));
    if (head_out) *head_out = sp.head;
    if (appended_out) *appended_out = sp.appended;
    E.retain_cand = true;  // later generations keep their candidates for the next dump
    return SPEO_OK;
}

int speo_engine_surrogate_set_labels(speo_engine e, int problem, const float *archive_y) {
    if (!e || e->e.C <= 0 || problem < 0 || problem >= e->e.cfg.num_problems || !archive_y)
        return set_error(SPEO_ERR_ARG, "speo_engine_surrogate_set_labels: no surrogate or bad arguments");
    Engine &E = e->e;
    const size_t cap = (size_t)E.cfg.surrogate_archive;
    SPEO_TRY_HIP(hipStreamSynchronize(E.stream));
    SPEO_TRY_HIP(hipMemcpy(E.Ay + (size_t)problem * cap, archive_y, sizeof(float) * cap,
                           hipMemcpyHostToDevice));
    return SPEO_OK;
}

int speo_engine_profile(speo_engine e, int generations, double *avg_ms_out) {
    if (!e || generations < 1 || !avg_ms_out)
        return set_error(SPEO_ERR_ARG, "speo_engine_profile: bad arguments");
    return e->e.profile(generations, avg_ms_out);
}

int speo_engine_total_rows(speo_engine e) { return e ? e->e.total_rows : 0; }
int speo_engine_max_dim(speo_engine e) { return e ? e->e.max_dim : 0; }

int speo_engine_buffers(speo_engine e, float **X, float **fX, int64_t *elem_off, int *row_off,
                        int *ld) {
    if (!e) return set_error(SPEO_ERR_ARG, "speo_engine_buffers: null engine");
    if (X) *X = e->e.X;
    if (fX) *fX = e->e.fX[e->e.cfg.mode == SPEO_MODE_FUSED ? e->e.p : 0];
    for (int i = 0; i < e->e.I; ++i) {
        if (elem_off) elem_off[i] = e->e.isl[(size_t)i].elem_off;
        if (row_off) row_off[i] = e->e.isl[(size_t)i].row_off;
        if (ld) ld[i] = e->e.isl[(size_t)i].ld;
    }
    return SPEO_OK;
}

}  // extern "C"
