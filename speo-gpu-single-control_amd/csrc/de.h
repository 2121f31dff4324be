// de.h -- DE operator building blocks shared by the standalone kernels and
// the fused generation kernel.
#pragma once

#include "common.h"

namespace speo {

// One island of the engine's batch, as the kernels see it.
struct IslandDesc {
    int64_t elem_off;  // first element of row 0 in X / U
    int row_off;       // first row in the fitness arrays
    int np, dim, ld;   // rows, coordinates, row stride (ld % 4 == 0)
    int problem;       // index into the problem table
    float F, CR;
};

struct GenArgs {
    float *U;
    const float *X;
    int ld, n, dim;
    int S;  // sub-island size
    const int *perm;
    const DEInfo *de_info;
    const int *best;
    float lo, hi;
    uint64_t seed;
    uint32_t gen, row_base;
    const float *inj_r;
    int n_inj;
    const float *inj_cr;
    int *r_out, *jrand_out;
};

// Sequential donor-phase draws of one row: draw k is word k%4 of Philox block
// {k/4, row, gen, purpose}; the first n_inj draws may be injected.  Blocks 0
// and 1 (the eight draws that cover the five donors and j_rand unless a row
// redraws more than twice) are computed up front, independently of each other.
struct DonorStream {
    uint64_t seed;
    uint32_t row, gen, k, cached, purpose;
    U4 w0, w1, cw;
    __host__ __device__ DonorStream(uint64_t s, uint32_t r, uint32_t g, uint32_t pur = kPurposeDonor)
        : seed(s), row(r), gen(g), k(0), cached(0xffffffffu), purpose(pur), cw{0, 0, 0, 0} {
        w0 = philox_block(s, 0u, r, g, pur);
        w1 = philox_block(s, 1u, r, g, pur);
    }
    __host__ __device__ __forceinline__ float next(const float *inj, int n_inj) {
        const uint32_t kk = k++;
        if (inj && (int)kk < n_inj) return inj[kk];
        if (kk < 8u) return u01(word(kk < 4u ? w0 : w1, kk & 3u));
        const uint32_t blk = kk >> 2;
        if (blk != cached) {
            cw = philox_block(seed, blk, row, gen, purpose);
            cached = blk;
        }
        return u01(word(cw, kk & 3u));
    }
};

// (int)(u S), clamped to [0, S] so that injected draws outside (0, 1] cannot
// index outside the sub-island; u in (0, 1] is unaffected.
__host__ __device__ __forceinline__ int scale_draw(float u, int S) {
    const int v = (int)(u * (float)S);
    return v < 0 ? 0 : (v > S ? S : v);
}

// kernel_DE_CUDA.cu:187-222: five sub-island-local donors, each redrawn while
// it equals an earlier one (earlier ones are not rechecked after a redraw),
// "== S" decremented before and after the redraw loop, then j_rand = (int)(u D)
// with no "== D" guard.  The target is not excluded.  Written with one draw
// site per donor: j walks the earlier donors, a collision redraws and stays
// on j, so the sequence of draws is that of the nested loops.
__host__ __device__ __forceinline__ void draw_donors(DonorStream &s, const float *inj, int n_inj, int S,
                                            int D, int (&r)[kMaxR], int &jr) {
#pragma unroll
    for (int i = 0; i < kMaxR; ++i) {
        int v = 0, j = 0;
        bool draw = true, first = true;
        for (;;) {
            if (draw) {
                v = scale_draw(s.next(inj, n_inj), S);
                if (first && v == S) v--;
                first = false;
                draw = false;
            }
            if (j >= i) break;
            int rj = r[0];
#pragma unroll
            for (int q = 1; q < kMaxR - 1; ++q)
                if (q == j) rj = r[q];
            if (v == rj && s.k < (uint32_t)kMaxDraws)
                draw = true;
            else
                ++j;
        }
        if (v == S) v--;
        r[i] = v;
    }
    jr = (int)(s.next(inj, n_inj) * (float)D);
}

int init_population(float *X, int ld, int n, int dim, float lo, float hi, uint64_t seed,
                    uint32_t row_base, hipStream_t s);
int de_generate(const GenArgs &a, hipStream_t s);
int de_select(float *X, float *fX, const float *U, float *fU, int ld, int n, int dim,
              uint8_t *accept_out, int clear_trials, hipStream_t s);
int segment_argmin(const float *f, const int *off, int nseg, int *best_out, hipStream_t s);
// Best slot of each of `subs` sub-islands of size S (members perm[s S ..]),
// ties to the lower slot, NaN last.
int subisland_best(const float *f, const int *perm, int S, int subs, int *best_out, hipStream_t s);
int migrate_ring(float *X, float *fX, const IslandDesc *isl_desc_dev, const int *row_off_dev,
                 const int *dst_dev, int *best_buf, int *worst_buf, int I, int k,
                 hipStream_t s);

}  // namespace speo
