// de_kernels.hip -- DE operator kernels and the drop-in DE entry points.
//
// Operator math follows the reference GPU kernels exactly (bit for bit, given
// the same uniforms):
//   donor draw with redraw quirks        kernel_DE_CUDA.cu:183-216
//   j_rand without "== dim" guard        kernel_DE_CUDA.cu:218-222
//   crossover test j == j_rand || u < CR kernel_DE_CUDA.cu:229-241
//   strategies 0-8                       kernel_DE_CUDA.cu:243-282
//   reflective bound                     kernel_DE_CUDA.cu:160-170
//   strict "<" selection                 kernel_DE_CUDA.cu:305-327
//   uniform init                         kernel_EA_CUDA.cu:91-118
// F*(a-b) terms are fused into the running sum with fmaf (nvcc's contraction
// of the reference expression); nothing else may contract.
#pragma clang fp contract(off)

#include "de.h"

namespace speo {

// ------------------------------------------------------------------ init

__global__ void k_init_population(float *X, int ld, int n, int dim, float lo, float hi,
                                  uint64_t seed, uint32_t row_base) {
    const int groups = ld >> 2;
    const int64_t total = (int64_t)n * groups;
    const float span = hi - lo;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int row = (int)(t / groups), g = (int)(t % groups);
        const U4 w = philox_block(seed, (uint32_t)g, row_base + (uint32_t)row, 0u, kPurposeInit);
        const int j0 = g * 4;
        float4 v;
        v.x = j0 + 0 < dim ? fmaf(u01(w.x), span, lo) : 0.f;
        v.y = j0 + 1 < dim ? fmaf(u01(w.y), span, lo) : 0.f;
        v.z = j0 + 2 < dim ? fmaf(u01(w.z), span, lo) : 0.f;
        v.w = j0 + 3 < dim ? fmaf(u01(w.w), span, lo) : 0.f;
        *reinterpret_cast<float4 *>(X + (size_t)row * ld + j0) = v;
    }
}

int init_population(float *X, int ld, int n, int dim, float lo, float hi, uint64_t seed,
                    uint32_t row_base, hipStream_t s) {
    if (n <= 0) return SPEO_OK;
    const int64_t total = (int64_t)n * (ld >> 2);
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_init_population, dim3(blocks), dim3(256), 0, s, X, ld, n, dim, lo, hi,
                       seed, row_base);
    return check_launch("k_init_population");
}

// -------------------------------------------------------------- generate

__device__ __forceinline__ float mutant(int st, float F, const float *X, int ld, int j, int sel,
                                        const int *r, int best) {
#define E(i) X[(size_t)(i) * ld + j]
    switch (st) {
    case 0: return fmaf(F, E(r[0]) - E(r[1]), E(sel));
    case 1: return fmaf(F, E(r[2]) - E(r[3]), fmaf(F, E(r[0]) - E(r[1]), E(sel)));
    case 2: return fmaf(F, E(r[0]) - E(r[1]), fmaf(F, E(best) - E(sel), E(sel)));
    case 3:
        return fmaf(F, E(r[1]) - E(r[2]),
                    fmaf(F, E(r[0]) - E(r[1]), fmaf(F, E(best) - E(sel), E(sel))));
    case 4: return fmaf(F, E(r[1]) - E(r[2]), E(r[0]));
    case 5: return fmaf(F, E(r[3]) - E(r[4]), fmaf(F, E(r[1]) - E(r[2]), E(r[0])));
    case 6: return fmaf(F, E(r[0]) - E(r[1]), E(best));
    case 7: return fmaf(F, E(r[2]) - E(r[3]), fmaf(F, E(r[0]) - E(r[1]), E(best)));
    case 8: return fmaf(F, E(r[1]) - E(r[2]), fmaf(F, E(r[0]) - E(sel), E(sel)));
    default: return E(sel);
    }
#undef E
}

// 16 lanes per row, 16 rows per 256-thread block.  The donor draw is uniform
// within a row's lanes (each lane repeats it, as the reference's 32 lanes all
// read lane 0's state, kernel_DE_CUDA.cu:184-185).
template <bool INJ>
__global__ __launch_bounds__(256) void k_de_generate(GenArgs a) {
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 16 + (tid >> 4);
    const int l = tid & 15;
    if (i >= a.n) return;
    const uint32_t row = a.row_base + (uint32_t)i;

    DonorStream ds(a.seed, row, a.gen);
    const float *inj = nullptr;
    if (INJ && a.inj_r) inj = a.inj_r + (size_t)i * a.n_inj;
    int r[kMaxR];
    int jr;
    draw_donors(ds, inj, INJ ? a.n_inj : 0, a.S, a.dim, r, jr);

    const int isl = i / a.S;
    const int sel = a.perm ? a.perm[i] : i;
#pragma unroll
    for (int k = 0; k < kMaxR; ++k) r[k] = a.perm ? a.perm[r[k] + isl * a.S] : r[k] + isl * a.S;
    if (l == 0) {
        if (a.r_out)
            for (int k = 0; k < kMaxR; ++k) a.r_out[(size_t)i * kMaxR + k] = r[k];
        if (a.jrand_out) a.jrand_out[i] = jr;
    }
    const DEInfo info = a.de_info[isl];
    const int b = a.best ? a.best[isl] : sel;
    const int ngroups = (a.dim + 3) >> 2;
    for (int g = l; g < ngroups; g += 16) {
        const U4 w = philox_block(a.seed, (uint32_t)g, row, a.gen, kPurposeCross);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = g * 4 + q;
            if (j < a.dim) {
                const float u = (INJ && a.inj_cr) ? a.inj_cr[(size_t)i * a.dim + j] : u01(word(w, q));
                const bool take = (j == jr) || (u < info.CR);
                const float v = take ? mutant(info.strategy_ID, info.F, a.X, a.ld, j, sel, r, b)
                                     : a.X[(size_t)sel * a.ld + j];
                a.U[(size_t)sel * a.ld + j] = check_bound(v, a.lo, a.hi);
            }
        }
    }
}

int de_generate(const GenArgs &a, hipStream_t s) {
    if (a.n <= 0) return SPEO_OK;
    const dim3 grid((a.n + 15) / 16), block(256);
    if (a.inj_r || a.inj_cr)
        hipLaunchKernelGGL(k_de_generate<true>, grid, block, 0, s, a);
    else
        hipLaunchKernelGGL(k_de_generate<false>, grid, block, 0, s, a);
    return check_launch("k_de_generate");
}

// ---------------------------------------------------------------- select

__global__ void k_de_select(float *X, float *fX, const float *U, float *fU, int ld, int n,
                            int dim, uint8_t *accept_out, int clear_trials) {
    const int tid = threadIdx.x;
    const int i = blockIdx.x * 16 + (tid >> 4);
    const int l = tid & 15;
    if (i >= n) return;
    const float fu = fU[i], fx = fX[i];
    const bool acc = fu < fx;
    if (acc)
        for (int j = l; j < dim; j += 16) X[(size_t)i * ld + j] = U[(size_t)i * ld + j];
    if (l == 0) {
        if (acc) fX[i] = fu;
        if (accept_out) accept_out[i] = acc ? 1 : 0;
        if (clear_trials) fU[i] = INFINITY;
    }
}

int de_select(float *X, float *fX, const float *U, float *fU, int ld, int n, int dim,
              uint8_t *accept_out, int clear_trials, hipStream_t s) {
    if (n <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_de_select, dim3((n + 15) / 16), dim3(256), 0, s, X, fX, U, fU, ld, n,
                       dim, accept_out, clear_trials);
    return check_launch("k_de_select");
}

// ---------------------------------------------------------------- argmin

__device__ __forceinline__ bool before(float a, int ia, float b, int ib) {
    const bool an = isnan(a), bn = isnan(b);
    if (an != bn) return bn;
    if (!an && a != b) return a < b;
    return ia < ib;
}

// Block argmin of (f[slot], slot) over members t in [0, count): slot =
// perm ? perm[base + t] : base + t.
__device__ void block_argmin(const float *f, const int *perm, int base, int count,
                             int *best_out) {
    __shared__ float sf[4];
    __shared__ int si[4];
    float bf = NAN;
    int bi = 0x7fffffff;
    for (int t = threadIdx.x; t < count; t += blockDim.x) {
        const int slot = perm ? perm[base + t] : base + t;
        const float v = f[slot];
        if (bi == 0x7fffffff || before(v, slot, bf, bi)) {
            bf = v;
            bi = slot;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        const float of = __shfl_xor(bf, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (oi != 0x7fffffff && (bi == 0x7fffffff || before(of, oi, bf, bi))) {
            bf = of;
            bi = oi;
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        sf[wave] = bf;
        si[wave] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            if (si[w] != 0x7fffffff && (bi == 0x7fffffff || before(sf[w], si[w], bf, bi))) {
                bf = sf[w];
                bi = si[w];
            }
        *best_out = bi;
    }
}

__global__ void k_segment_argmin(const float *f, const int *off, int *best_out) {
    const int s = blockIdx.x;
    block_argmin(f, nullptr, off[s], off[s + 1] - off[s], best_out + s);
}

__global__ void k_subisland_best(const float *f, const int *perm, int S, int *best_out) {
    block_argmin(f, perm, blockIdx.x * S, S, best_out + blockIdx.x);
}

int subisland_best(const float *f, const int *perm, int S, int subs, int *best_out, hipStream_t s) {
    if (subs <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_subisland_best, dim3(subs), dim3(256), 0, s, f, perm, S, best_out);
    return check_launch("k_subisland_best");
}

int segment_argmin(const float *f, const int *off, int nseg, int *best_out, hipStream_t s) {
    if (nseg <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_segment_argmin, dim3(nseg), dim3(256), 0, s, f, off, best_out);
    return check_launch("k_segment_argmin");
}

// ------------------------------------------------------------- migration

// Block per island: rank of each row in the order (f ascending, row
// ascending, NaN last); the first k rows are its best, the last k its worst.
// The island's fitness is compared from LDS (islands of up to 4096 rows; larger
// ones from global memory).
constexpr int kRankLds = 4096;

// before(a, ia, b, ib) without branches (the rank loop runs it n^2 times).
__device__ __forceinline__ int before_nb(float a, int ia, float b, int ib) {
    const bool an = a != a, bn = b != b, idx = ia < ib;
    const bool neither = !an && !bn;
    return (int)((bn && !an) | (an && bn && idx) | (neither && ((a < b) | ((a == b) && idx))));
}
__global__ void k_rank_islands(const float *f, const int *row_off, int k, int *best_out,
                               int *worst_out) {
    __shared__ float fs[kRankLds];
    const int isl = blockIdx.x;
    const int r0 = row_off[isl], np = row_off[isl + 1] - r0;
    const bool staged = np <= kRankLds;
    if (staged)
        for (int r = threadIdx.x; r < np; r += blockDim.x) fs[r] = f[r0 + r];
    __syncthreads();
    auto rank_rows = [&](const auto *fi) {
        for (int r = threadIdx.x; r < np; r += blockDim.x) {
            const float fr = fi[r];
            int rank = 0;
#pragma unroll 8
            for (int q = 0; q < np; ++q) rank += before_nb(fi[q], r0 + q, fr, r0 + r);
            if (rank < k) best_out[isl * k + rank] = r0 + r;
            if (rank >= np - k) worst_out[isl * k + (np - 1 - rank)] = r0 + r;
        }
    };
    // two loops, so that the staged one reads LDS with LDS instructions
    if (staged)
        rank_rows(fs);
    else
        rank_rows(f + r0);
}

// Block per (island, migrant): island isl's t-th best row replaces the t-th
// worst row of island dst[isl].
__global__ void k_migrate_apply(float *X, float *fX, const IslandDesc *isl_desc, const int *dst,
                                const int *best, const int *worst, int k) {
    const int isl = blockIdx.x / k, t = blockIdx.x % k;
    const int d = dst[isl];
    if (d < 0) return;
    const IslandDesc S = isl_desc[isl], T = isl_desc[d];
    const int sr = best[isl * k + t] - S.row_off, tr = worst[d * k + t] - T.row_off;
    const float *src = X + S.elem_off + (size_t)sr * S.ld;
    float *out = X + T.elem_off + (size_t)tr * T.ld;
    for (int j = threadIdx.x; j < S.dim; j += blockDim.x) out[j] = src[j];
    if (threadIdx.x == 0) fX[T.row_off + tr] = fX[S.row_off + sr];
}

int migrate_ring(float *X, float *fX, const IslandDesc *isl_desc_dev, const int *row_off_dev,
                 const int *dst_dev, int *best_buf, int *worst_buf, int I, int k,
                 hipStream_t s) {
    if (I <= 0 || k <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_rank_islands, dim3(I), dim3(256), 0, s, fX, row_off_dev, k, best_buf,
                       worst_buf);
    SPEO_TRY(check_launch("k_rank_islands"));
    hipLaunchKernelGGL(k_migrate_apply, dim3(I * k), dim3(128), 0, s, X, fX, isl_desc_dev, dst_dev,
                       best_buf, worst_buf, k);
    return check_launch("k_migrate_apply");
}

}  // namespace speo

// ======================================================================
//  C ABI
// ======================================================================
using namespace speo;

extern "C" {

int speo_init_population(float *X, int ld, int n, int dim, float lo, float hi, uint64_t seed,
                         uint32_t row_base, speo_stream_t stream) {
    if (!X || n < 0 || dim < 1 || ld < dim || (ld & 3) || (reinterpret_cast<uintptr_t>(X) & 15))
        return set_error(SPEO_ERR_ARG, "speo_init_population: need ld >= dim, ld % 4 == 0 and "
                                       "a 16-byte aligned X");
    return init_population(X, ld, n, dim, lo, hi, seed, row_base, as_stream(stream));
}

int speo_de_generate(float *U, const float *X, int ld, int n, int dim, int subisland_num,
                     const int *perm, const DEInfo *de_info, const int *best, float lo, float hi,
                     SPEO_RandState rng, const float *inj_r, int n_inj, const float *inj_cr,
                     int *r_out, int *jrand_out, speo_stream_t stream) {
    if (!U || !X || !de_info || n < 0 || dim < 1 || ld < dim || subisland_num < 1 ||
        n % subisland_num || (n > 0 && n / subisland_num < 4) || n_inj < 0 ||
        (inj_r && n_inj == 0))
        return set_error(SPEO_ERR_ARG, "speo_de_generate: need ld >= dim and equal sub-islands "
                                       "of at least 4 rows");
    GenArgs a{U, X, ld, n, dim, n > 0 ? n / subisland_num : 1, perm, de_info, best, lo, hi,
              rng.seed, rng.generation, rng.row_base, inj_r, n_inj, inj_cr, r_out, jrand_out};
    return de_generate(a, as_stream(stream));
}

int speo_de_select(float *X, float *fX, const float *U, const float *fU, int ld, int n, int dim,
                   uint8_t *accept_out, speo_stream_t stream) {
    if (!X || !fX || !U || !fU || n < 0 || ld < dim)
        return set_error(SPEO_ERR_ARG, "speo_de_select: bad arguments");
    return de_select(X, fX, U, const_cast<float *>(fU), ld, n, dim, accept_out, 0,
                     as_stream(stream));
}

int speo_segment_argmin(const float *f, const int *off, int nseg, int *best_out,
                        speo_stream_t stream) {
    if (!f || !off || !best_out || nseg < 0)
        return set_error(SPEO_ERR_ARG, "speo_segment_argmin: bad arguments");
    return segment_argmin(f, off, nseg, best_out, as_stream(stream));
}

// ---- drop-in layer ----------------------------------------------------

int API_Initilize(GPU_Population d_population, SPEO_RandState *rand_state,
                  GPU_IslandInfo info, ProblemInfo pi) {
    if (!rand_state || !d_population.elements || info.island_size < 0 || pi.dim < 1)
        return set_error(SPEO_ERR_ARG, "API_Initilize: bad arguments");
    // global_SetupRandomState(d_rand_states, problem_info.seed), kernel_EA_CUDA.cu:128
    rand_state->seed = (uint64_t)(uint32_t)pi.seed;
    rand_state->generation = 0;
    rand_state->row_base = 0;
    const int ld = next_pow2_dim(pi.dim);
    return init_population(d_population.elements, ld, info.island_size, pi.dim, pi.min_bound,
                           pi.max_bound, rand_state->seed, rand_state->row_base, nullptr);
}

int API_DE_GenerateNewPopulation(GPU_Population d_candidate, GPU_Population d_population,
                                 int *d_best_individual_ID, GPU_IslandInfo info,
                                 DEInfo *d_DE_info, SPEO_RandState *rand_state,
                                 ProblemInfo pi) {
    const int n = info.island_size, subs = info.subisland_num;
    if (!rand_state || !d_DE_info || !d_candidate.elements || !d_population.elements ||
        !d_population.fitness_value || n <= 0 || subs < 1 || n % subs || n / subs < 4 ||
        pi.dim < 1)
        return set_error(SPEO_ERR_ARG, "API_DE_GenerateNewPopulation: need equal sub-islands "
                                       "of at least 4 individuals");
    const int S = n / subs;
    // global_DE_FindBestIndividualInIsland (kernel_DE_CUDA.cu:120-158), with
    // each sub-island's own offset into permutated_index.
    if (d_best_individual_ID) {
        hipLaunchKernelGGL(k_subisland_best, dim3(subs), dim3(256), 0, nullptr,
                           d_population.fitness_value, info.permutated_index, S,
                           d_best_individual_ID);
        SPEO_TRY(check_launch("k_subisland_best"));
    }
    GenArgs a{d_candidate.elements, d_population.elements, next_pow2_dim(pi.dim), n, pi.dim, S,
              info.permutated_index, d_DE_info, d_best_individual_ID, pi.min_bound,
              pi.max_bound, rand_state->seed, rand_state->generation, rand_state->row_base,
              nullptr, 0, nullptr, nullptr, nullptr};
    SPEO_TRY(de_generate(a, nullptr));
    rand_state->generation++;
    return SPEO_OK;
}

int API_DE_SelectSurvival(GPU_Population d_population, GPU_Population d_candidate,
                          GPU_IslandInfo info, ProblemInfo pi) {
    if (!d_population.elements || !d_population.fitness_value || !d_candidate.elements ||
        !d_candidate.fitness_value || info.island_size < 0 || pi.dim < 1)
        return set_error(SPEO_ERR_ARG, "API_DE_SelectSurvival: bad arguments");
    return de_select(d_population.elements, d_population.fitness_value, d_candidate.elements,
                     d_candidate.fitness_value, next_pow2_dim(pi.dim), info.island_size, pi.dim,
                     nullptr, 0, nullptr);
}

}  // extern "C"
