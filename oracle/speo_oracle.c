/*
 * speo_oracle.c -- CPU oracle for the island-model DE/rand/1/bin hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links or calls
 * this file; only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker (or as the timed CPU baseline).
 *
 * What it restates (reference paths relative to SPEO-GPU-single-control/):
 *   - DE operator math of the GPU kernels, in FP32 (the reference's
 *     GPU_SINGLE_PRECISION branch):
 *       donor draw with its redraw quirks   CUDA_EA_library/kernel_DE_CUDA.cu:183-216
 *       j_rand without "== dim" guard       CUDA_EA_library/kernel_DE_CUDA.cu:218-222
 *       binomial crossover test             CUDA_EA_library/kernel_DE_CUDA.cu:229-241
 *       strategies 0-8 (4 = rand/1/bin)     CUDA_EA_library/kernel_DE_CUDA.cu:243-282
 *       reflective bound                    CUDA_EA_library/kernel_DE_CUDA.cu:160-170
 *       greedy strict "<" selection         CUDA_EA_library/kernel_DE_CUDA.cu:305-327
 *       uniform population init             CUDA_EA_library/kernel_EA_CUDA.cu:91-118
 *   - objectives in FP64, following the reference CPU code:
 *       shift / rotate                      CUDA_EA_library/CEC2014.cc:1373-1393
 *       elliptic                            CUDA_EA_library/CEC2014.cc:481-497
 *       Ackley                              CUDA_EA_library/CEC2014.cc:590-612
 *       Rastrigin                           CUDA_EA_library/CEC2014.cc:668-683
 *       bent cigar                          CUDA_EA_library/CEC2014.cc:499-516
 *       discus                              CUDA_EA_library/CEC2014.cc:518-530
 *       Schwefel (modified)                 CUDA_EA_library/CEC2014.cc:705-738
 *       HappyCat                            CUDA_EA_library/CEC2014.cc:907-929
 *       HGBat                               CUDA_EA_library/CEC2014.cc:931-953
 *       Rosenbrock                          CUDA_EA_library/CEC2014.cc:552-568
 *       Weierstrass                         CUDA_EA_library/CEC2014.cc:615-644
 *       Griewank                            CUDA_EA_library/CEC2014.cc:646-666
 *       Katsuura                            CUDA_EA_library/CEC2014.cc:739-765
 *       Griewank-Rosenbrock                 CUDA_EA_library/CEC2014.cc:842-868
 *       expanded Schaffer F6                CUDA_EA_library/CEC2014.cc:870-905
 *       hybrid functions hf01-hf06          CUDA_EA_library/CEC2014.cc:954-1207
 *       compositions cf01-cf08              CUDA_EA_library/CEC2014.cc:1209-1371
 *       F1-F30 dispatch                     CUDA_EA_library/CEC2014.cc:14-145
 *       composition weights (cf_cal)        CUDA_EA_library/CEC2014.cc:1481-1519
 *   - Philox4x32-10 (Salmon et al., "Parallel random numbers: as easy as
 *     1, 2, 3", SC'11), replacing cuRAND XORWOW, which is closed and is not
 *     restated.  The mapping of draws to (purpose, individual, generation,
 *     index) is this project's own and is shared with the HIP kernels.
 *
 * Pinning (tests/test_oracle_*.py): Philox against the Random123 known-answer
 * triplets and the open-source rocRAND header compiled on the CPU; DE operators
 * against hand-worked cases, forced edge draws and invariants; objectives
 * against closed forms and against the reference's own CEC2014.cc compiled
 * from its sources (oracle/_ref), via committed golden vectors.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_EXPORT __attribute__((visibility("default")))

/* ---------------------------------------------------------------- Philox */

ORACLE_EXPORT void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2],
                                        uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        if (round < 9) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* u32 -> float in (0,1]: the top 24 bits plus one, times 2^-24.  Every value
 * is exactly representable, so 1.0f occurs (probability 2^-24), as with the
 * reference's (0,1] uniforms, and CPU and GPU agree bit for bit. */
ORACLE_EXPORT float oracle_u01(uint32_t x) {
    return (float)((x >> 8) + 1u) * 0x1p-24f;
}

/* Draw `idx` of stream (purpose, a, b): counter = {idx/4, a, b, purpose},
 * key = seed; word idx%4 of the block. */
ORACLE_EXPORT float oracle_draw(uint64_t seed, uint32_t purpose, uint32_t a, uint32_t b,
                                uint32_t idx) {
    uint32_t ctr[4] = {idx >> 2, a, b, purpose};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t out[4];
    oracle_philox4x32_10(ctr, key, out);
    return oracle_u01(out[idx & 3u]);
}

enum { PURPOSE_DONOR = 0, PURPOSE_CROSS = 1, PURPOSE_INIT = 2 };

/* ----------------------------------------------------------- DE operators */

/* Reflective bound, kernel_DE_CUDA.cu:160-170.  The loop is capped at 64
 * reflections and then clamped; finite inputs within (1+4)x the range never
 * reach the cap.  NaN passes through unchanged, as in the reference. */
ORACLE_EXPORT float oracle_check_bound(float v, float lo, float hi) {
    int it = 0;
    while ((v < lo) || (v > hi)) {
        if (v < lo) v = lo + (lo - v);
        if (v > hi) v = hi - (v - hi);
        if (++it >= 64) { v = v < lo ? lo : (v > hi ? hi : v); break; }
    }
    return v;
}

#define MAX_R 5
#define MAX_DRAWS 4096

typedef struct {
    uint64_t seed;
    uint32_t row, gen;
    uint32_t k;              /* draws consumed */
    const float *inj;        /* injected draws for this row, or NULL */
    int n_inj;
} draw_stream;

static float next_draw(draw_stream *s) {
    uint32_t k = s->k++;
    if (s->inj && (int)k < s->n_inj) return s->inj[k];
    return oracle_draw(s->seed, PURPOSE_DONOR, s->row, s->gen, k);
}

/* (int)(u S) clamped to [0, S]: injected draws outside (0,1] stay in range;
 * u in (0,1] is unaffected. */
static int scale_draw(float u, int S) {
    float p = u * (float)S;
    if (!(p >= 0.0f)) return 0;             /* negative or NaN */
    if (p >= (float)S) return p > (float)S ? S : (int)p;
    return (int)p;
}

/* Donor indices and j_rand for one individual, kernel_DE_CUDA.cu:183-222.
 * S = sub-island size; r[] are sub-island-local (before the permutation). */
ORACLE_EXPORT void oracle_draw_donors(uint64_t seed, uint32_t row, uint32_t gen, int S, int D,
                                      const float *inj, int n_inj, int r_out[MAX_R],
                                      int *jrand_out, int *draws_used) {
    draw_stream s = {seed, row, gen, 0, inj, n_inj};
    int r[MAX_R] = {0};
    for (int i = 0; i < MAX_R; ++i) {
        r[i] = scale_draw(next_draw(&s), S);
        if (r[i] == S) r[i]--;
        for (int j = 0; j < i; ++j) {
            while (r[i] == r[j] && s.k < MAX_DRAWS) r[i] = scale_draw(next_draw(&s), S);
        }
        if (r[i] == S) r[i]--;
    }
    int jr = (int)(next_draw(&s) * (float)D);
    for (int i = 0; i < MAX_R; ++i) r_out[i] = r[i];
    *jrand_out = jr;
    if (draws_used) *draws_used = (int)s.k;
}

/* Mutant coordinate for strategy 0-8, kernel_DE_CUDA.cu:243-277.  The sums
 * are evaluated left to right with each F*(a-b) fused into the running sum
 * (fmaf), the contraction nvcc applies to the reference expression. */
static float mutant(int strategy, float F, const float *X, int ld, int j, int sel, const int *r,
                    int best) {
#define E(i) X[(size_t)(i) * ld + j]
    float v;
    switch (strategy) {
    case 0: v = fmaf(F, E(r[0]) - E(r[1]), E(sel)); break;
    case 1: v = fmaf(F, E(r[2]) - E(r[3]), fmaf(F, E(r[0]) - E(r[1]), E(sel))); break;
    case 2: v = fmaf(F, E(r[0]) - E(r[1]), fmaf(F, E(best) - E(sel), E(sel))); break;
    case 3:
        v = fmaf(F, E(r[1]) - E(r[2]),
                 fmaf(F, E(r[0]) - E(r[1]), fmaf(F, E(best) - E(sel), E(sel))));
        break;
    case 4: v = fmaf(F, E(r[1]) - E(r[2]), E(r[0])); break;
    case 5: v = fmaf(F, E(r[3]) - E(r[4]), fmaf(F, E(r[1]) - E(r[2]), E(r[0]))); break;
    case 6: v = fmaf(F, E(r[0]) - E(r[1]), E(best)); break;
    case 7: v = fmaf(F, E(r[2]) - E(r[3]), fmaf(F, E(r[0]) - E(r[1]), E(best))); break;
    case 8: v = fmaf(F, E(r[1]) - E(r[2]), fmaf(F, E(r[0]) - E(sel), E(sel))); break;
    default: v = E(sel); break;
    }
#undef E
    return v;
}

/*
 * One DE generation over one population of NP rows split into
 * `subisland_num` sub-islands through `perm` (kernel_DE_CUDA.cu:172-289):
 * row i is processed by "thread" i; it writes the trial of slot perm[i].
 * The random stream of row i is (seed, row_base + i, gen).  Per-sub-island
 * parameters CR/F/strategy come from cr[], f[], strategy[]; best[] holds the
 * per-sub-island best slot for strategies 2, 3, 6, 7.
 * Optional outputs (may be NULL): r_out[NP*5] (after the permutation),
 * jrand_out[NP], mask_out[NP*D] (1 = coordinate taken from the mutant).
 * Injected draws: inj_r[NP*n_inj] replace the first n_inj donor draws of each
 * row; inj_cr[NP*D] replace the crossover draws.
 * Returns 0, or -1 on invalid arguments.
 */
ORACLE_EXPORT int oracle_de_generate(float *U, const float *X, int ld, int NP, int D,
                                     int subisland_num, const int *perm, const float *cr,
                                     const float *f, const int *strategy, const int *best,
                                     float lo, float hi, uint64_t seed, uint32_t gen,
                                     uint32_t row_base, const float *inj_r, int n_inj,
                                     const float *inj_cr, int *r_out, int *jrand_out,
                                     unsigned char *mask_out) {
    if (NP <= 0 || D <= 0 || ld < D || subisland_num <= 0 || NP % subisland_num) return -1;
    int S = NP / subisland_num;
    if (S < 4) return -1;
    for (int i = 0; i < NP; ++i) {
        int sel = perm ? perm[i] : i;
        int isl = i / S;
        int r[MAX_R], jr;
        oracle_draw_donors(seed, row_base + (uint32_t)i, gen, S, D,
                           inj_r ? inj_r + (size_t)i * n_inj : NULL, n_inj, r, &jr, NULL);
        for (int k = 0; k < MAX_R; ++k) r[k] = perm ? perm[r[k] + isl * S] : r[k] + isl * S;
        if (r_out) for (int k = 0; k < MAX_R; ++k) r_out[i * MAX_R + k] = r[k];
        if (jrand_out) jrand_out[i] = jr;
        float F = f[isl], CR = cr[isl];
        int st = strategy ? strategy[isl] : 4;
        int b = best ? best[isl] : sel;
        for (int j = 0; j < D; ++j) {
            float u = inj_cr ? inj_cr[(size_t)i * D + j]
                             : oracle_draw(seed, PURPOSE_CROSS, row_base + (uint32_t)i, gen,
                                           (uint32_t)j);
            int take = (j == jr) || (u < CR);
            float v = take ? mutant(st, F, X, ld, j, sel, r, b) : X[(size_t)sel * ld + j];
            if (mask_out) mask_out[(size_t)i * D + j] = (unsigned char)take;
            U[(size_t)sel * ld + j] = oracle_check_bound(v, lo, hi);
        }
    }
    return 0;
}

/* Greedy selection, kernel_DE_CUDA.cu:305-327: strict "<", so ties and NaN
 * keep the parent.  Returns the number of accepted trials. */
ORACLE_EXPORT int oracle_de_select(float *X, float *fX, const float *U, const float *fU, int ld,
                                   int NP, int D, unsigned char *accept_out) {
    int n = 0;
    for (int i = 0; i < NP; ++i) {
        int acc = fU[i] < fX[i];
        if (accept_out) accept_out[i] = (unsigned char)acc;
        if (acc) {
            memcpy(X + (size_t)i * ld, U + (size_t)i * ld, sizeof(float) * (size_t)D);
            fX[i] = fU[i];
            ++n;
        }
    }
    return n;
}

/* Uniform init, kernel_EA_CUDA.cu:91-118: x = lo + u*(hi-lo), u in (0,1]. */
ORACLE_EXPORT void oracle_init_population(float *X, int ld, int NP, int D, float lo, float hi,
                                          uint64_t seed, uint32_t row_base) {
    for (int i = 0; i < NP; ++i)
        for (int j = 0; j < D; ++j) {
            float u = oracle_draw(seed, PURPOSE_INIT, row_base + (uint32_t)i, 0u, (uint32_t)j);
            X[(size_t)i * ld + j] = fmaf(u, hi - lo, lo);
        }
}

/* ------------------------------------------------------------- objectives */

enum {
    OBJ_ELLIPTIC = 1,
    OBJ_RASTRIGIN = 2,
    OBJ_ACKLEY = 3,
    OBJ_BENT_CIGAR = 4,
    OBJ_DISCUS = 5,
    OBJ_HAPPYCAT = 6,
    OBJ_HGBAT = 7,
    OBJ_SCHWEFEL = 8,
    OBJ_ROSENBROCK = 9,
    OBJ_WEIERSTRASS = 10,
    OBJ_GRIEWANK = 11,
    OBJ_KATSUURA = 12,
    OBJ_GRIE_ROSEN = 13,
    OBJ_ESCAFFER6 = 14
};

static const double ORACLE_PI = 3.1415926535897932384626433832795029;
static const double ORACLE_E = 2.7182818284590452353602874713526625;

/* z = M (x - o), CEC2014.cc:1373-1393 (xrot[i] = sum_j x_j M[i*nx+j]). */
ORACLE_EXPORT void oracle_shift_rotate(const double *x, const double *o, const double *M, int D,
                                       double *z, double *scratch) {
    for (int j = 0; j < D; ++j) scratch[j] = x[j] - o[j];
    for (int i = 0; i < D; ++i) {
        double acc = 0.0;
        for (int j = 0; j < D; ++j) acc = acc + scratch[j] * M[(size_t)i * D + j];
        z[i] = acc;
    }
}

/* CEC2014.cc:488-496 */
ORACLE_EXPORT double oracle_elliptic(const double *z, int D) {
    double f = 0.0;
    for (int i = 0; i < D; ++i) f += pow(10.0, 6.0 * i / (D - 1)) * z[i] * z[i];
    return f;
}

/* CEC2014.cc:679-682 */
ORACLE_EXPORT double oracle_rastrigin(const double *z, int D) {
    double f = 0.0;
    for (int i = 0; i < D; ++i) f += (z[i] * z[i] - 10.0 * cos(2.0 * ORACLE_PI * z[i]) + 10.0);
    return f;
}

/* CEC2014.cc:604-611 */
ORACLE_EXPORT double oracle_ackley(const double *z, int D) {
    double s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < D; ++i) {
        s1 += z[i] * z[i];
        s2 += cos(2.0 * ORACLE_PI * z[i]);
    }
    s1 = -0.2 * sqrt(s1 / (double)D);
    s2 /= (double)D;
    return ORACLE_E - 20.0 * exp(s1) + 20.0 - exp(s2);
}

/* CEC2014.cc:507-515: z_0^2, then 10^6 z_i^2 for i >= 1, in index order. */
ORACLE_EXPORT double oracle_bent_cigar(const double *z, int D) {
    double f = z[0] * z[0];
    for (int i = 1; i < D; ++i) f += pow(10.0, 6.0) * z[i] * z[i];
    return f;
}

/* CEC2014.cc:521-529: 10^6 z_0^2, then z_i^2 for i >= 1. */
ORACLE_EXPORT double oracle_discus(const double *z, int D) {
    double f = pow(10.0, 6.0) * z[0] * z[0];
    for (int i = 1; i < D; ++i) f += z[i] * z[i];
    return f;
}

/* CEC2014.cc:919-928: u_i = z_i - 1, r2 = sum u_i^2, s = sum u_i,
 * |r2 - D|^(2 alpha) + (r2 / 2 + s) / D + 1/2 with alpha = 1/8. */
ORACLE_EXPORT double oracle_happycat(const double *z, int D) {
    double r2 = 0.0, s = 0.0;
    for (int i = 0; i < D; ++i) {
        const double u = z[i] - 1.0;
        r2 += u * u;
        s += u;
    }
    return pow(fabs(r2 - D), 2.0 * (1.0 / 8.0)) + (0.5 * r2 + s) / D + 0.5;
}

/* CEC2014.cc:943-952: the same sums, |r2^2 - s^2|^(2 alpha) + ... with
 * alpha = 1/4. */
ORACLE_EXPORT double oracle_hgbat(const double *z, int D) {
    double r2 = 0.0, s = 0.0;
    for (int i = 0; i < D; ++i) {
        const double u = z[i] - 1.0;
        r2 += u * u;
        s += u;
    }
    return pow(fabs(pow(r2, 2.0) - pow(s, 2.0)), 2.0 * (1.0 / 4.0)) + (0.5 * r2 + s) / D + 0.5;
}

/* CEC2014.cc:717-737: v = z + 420.9687462275036; inside [-500, 500] the
 * term -v sin(sqrt|v|); outside, |v| folded back by fmod into
 * m = 500 - fmod(|v|, 500) with -m sin(sqrt m) for v > 500 and
 * +m sin(sqrt m) for v < -500, plus ((|v| - 500) / 100)^2 / D; then
 * 418.9828872724338 D. */
ORACLE_EXPORT double oracle_schwefel(const double *z, int D) {
    double f = 0.0;
    for (int i = 0; i < D; ++i) {
        const double v = z[i] + 4.209687462275036e+002;
        if (v > 500) {
            f -= (500.0 - fmod(v, 500)) * sin(pow(500.0 - fmod(v, 500), 0.5));
            const double t = (v - 500.0) / 100;
            f += t * t / D;
        } else if (v < -500) {
            f -= (-500.0 + fmod(fabs(v), 500)) * sin(pow(500.0 - fmod(fabs(v), 500), 0.5));
            const double t = (v + 500.0) / 100;
            f += t * t / D;
        } else {
            f -= v * sin(pow(fabs(v), 0.5));
        }
    }
    return f + 4.189828872724338e+002 * D;
}

/* CEC2014.cc:552-568 with the fixed shift of 1 its callers pass (F4 at
 * :32, hf03 / hf05 at :1066, :1151, cf01 at :1217): y = z + 1, then
 * sum_{i < D-1} 100 (y_i^2 - y_{i+1})^2 + (y_i - 1)^2. */
ORACLE_EXPORT double oracle_rosenbrock(const double *z, int D) {
    double f = 0.0;
    for (int i = 0; i < D - 1; ++i) {
        const double y0 = z[i] + 1.0, y1 = z[i + 1] + 1.0;
        const double t1 = y0 * y0 - y1, t2 = y0 - 1.0;
        f += 100.0 * t1 * t1 + t2 * t2;
    }
    return f;
}

/* CEC2014.cc:615-644: a = 0.5, b = 3, k_max = 20;
 * sum_i sum_k a^k cos(2 pi b^k (z_i + 0.5)) - D sum_k a^k cos(2 pi b^k 0.5),
 * each term formed as pow(a, k) * cos(2.0 * PI * pow(b, k) * (z + 0.5)). */
ORACLE_EXPORT double oracle_weierstrass(const double *z, int D) {
    double f = 0.0, c = 0.0;
    for (int k = 0; k <= 20; ++k) c += pow(0.5, k) * cos(2.0 * ORACLE_PI * pow(3.0, k) * 0.5);
    for (int i = 0; i < D; ++i) {
        double s = 0.0;
        for (int k = 0; k <= 20; ++k)
            s += pow(0.5, k) * cos(2.0 * ORACLE_PI * pow(3.0, k) * (z[i] + 0.5));
        f += s;
    }
    return f - D * c;
}

/* CEC2014.cc:646-666: 1 + sum z_i^2 / 4000 - prod cos(z_i / sqrt(i + 1)). */
ORACLE_EXPORT double oracle_griewank(const double *z, int D) {
    double s = 0.0, p = 1.0;
    for (int i = 0; i < D; ++i) {
        s += z[i] * z[i];
        p *= cos(z[i] / sqrt(1.0 + i));
    }
    return 1.0 + s / 4000.0 - p;
}

/* CEC2014.cc:739-765: t_i = sum_{j=1..32} |2^j z_i - floor(2^j z_i + 1/2)| / 2^j,
 * f = 10/D^2 prod_i (1 + (i + 1) t_i)^(10 / D^1.2) - 10/D^2. */
ORACLE_EXPORT double oracle_katsuura(const double *z, int D) {
    double f = 1.0;
    const double e = 10.0 / pow(1.0 * D, 1.2);
    for (int i = 0; i < D; ++i) {
        double t = 0.0;
        for (int j = 1; j <= 32; ++j) {
            const double s = pow(2.0, j), v = s * z[i];
            t += fabs(v - floor(v + 0.5)) / s;
        }
        f *= pow(1.0 + (i + 1) * t, e);
    }
    const double c = 10.0 / D / D;
    return f * c - c;
}

/* Griewank of a Rosenbrock term, t -> t^2 / 4000 - cos t + 1. */
static double grie_of_rosen(double a, double b) {
    const double t1 = a * a - b, t2 = a - 1.0;
    const double t = 100.0 * t1 * t1 + t2 * t2;
    return (t * t) / 4000.0 - cos(t) + 1.0;
}

/* CEC2014.cc:842-868 with the fixed shift of 1 of its callers (F15 at :76,
 * hf04 / hf06 at :1109, :1196, cf06 at :1324): y = z + 1, the pairs
 * (y_i, y_{i+1}) for i < D-1, then the wrap-around pair (y_{D-1}, y_0). */
ORACLE_EXPORT double oracle_grie_rosen(const double *z, int D) {
    double f = 0.0;
    for (int i = 0; i < D - 1; ++i) f += grie_of_rosen(z[i] + 1.0, z[i + 1] + 1.0);
    f += grie_of_rosen(z[D - 1] + 1.0, z[0] + 1.0);
    return f;
}

/* One Schaffer F6 term of r2 = a^2 + b^2. */
static double scaffer6(double r2) {
    double t1 = sin(sqrt(r2));
    t1 = t1 * t1;
    const double t2 = 1.0 + 0.001 * r2;
    return 0.5 + (t1 - 0.5) / (t2 * t2);
}

/* CEC2014.cc:870-905: pairs (z_i, z_{i+1}) for i < D-1 and the wrap-around
 * pair (z_{D-1}, z_0); a single coordinate gives the term of z_0^2. */
ORACLE_EXPORT double oracle_escaffer6(const double *z, int D) {
    if (D == 1) return scaffer6(z[0] * z[0]);
    double f = 0.0;
    for (int i = 0; i < D - 1; ++i) f += scaffer6(z[i] * z[i] + z[i + 1] * z[i + 1]);
    f += scaffer6(z[D - 1] * z[D - 1] + z[0] * z[0]);
    return f;
}

ORACLE_EXPORT double oracle_basic(int objective, const double *z, int D) {
    switch (objective) {
    case OBJ_ROSENBROCK: return oracle_rosenbrock(z, D);
    case OBJ_WEIERSTRASS: return oracle_weierstrass(z, D);
    case OBJ_GRIEWANK: return oracle_griewank(z, D);
    case OBJ_KATSUURA: return oracle_katsuura(z, D);
    case OBJ_GRIE_ROSEN: return oracle_grie_rosen(z, D);
    case OBJ_ESCAFFER6: return oracle_escaffer6(z, D);
    case OBJ_ELLIPTIC: return oracle_elliptic(z, D);
    case OBJ_BENT_CIGAR: return oracle_bent_cigar(z, D);
    case OBJ_DISCUS: return oracle_discus(z, D);
    case OBJ_HAPPYCAT: return oracle_happycat(z, D);
    case OBJ_HGBAT: return oracle_hgbat(z, D);
    case OBJ_SCHWEFEL: return oracle_schwefel(z, D);
    case OBJ_RASTRIGIN: return oracle_rastrigin(z, D);
    case OBJ_ACKLEY: return oracle_ackley(z, D);
    default: return NAN;
    }
}

/* Shifted-rotated objective for n points (rows of x with stride ldx). */
ORACLE_EXPORT void oracle_eval(int objective, const double *x, int ldx, int n, const double *o,
                               const double *M, int D, double *f_out) {
    double *z = (double *)malloc(sizeof(double) * 2 * (size_t)D);
    for (int p = 0; p < n; ++p) {
        oracle_shift_rotate(x + (size_t)p * ldx, o, M, D, z, z + D);
        f_out[p] = oracle_basic(objective, z, D);
    }
    free(z);
}

/* Same, from FP32 points (the device population), evaluated in FP64. */
ORACLE_EXPORT void oracle_eval_f32(int objective, const float *x, int ldx, int n, const double *o,
                                   const double *M, int D, double *f_out) {
    double *xd = (double *)malloc(sizeof(double) * (size_t)D);
    for (int p = 0; p < n; ++p) {
        for (int j = 0; j < D; ++j) xd[j] = (double)x[(size_t)p * ldx + j];
        oracle_eval(objective, xd, D, 1, o, M, D, f_out + p);
    }
    free(xd);
}

/*
 * Composition of `ncomp` shifted-rotated components, cf_cal
 * (CEC2014.cc:1481-1519): fit_i = lambda_i * g_i(M_i (x - o_i)) + bias_i,
 * w_i = d_i^-1/2 exp(-d_i / (2 D sigma_i^2)) with d_i = |x - o_i|^2
 * (w_i = 1e99 when d_i == 0), all-zero weights fall back to uniform, and
 * f = sum_i w_i / sum_j w_j * fit_i.  os: [ncomp][D], Ms: [ncomp][D][D].
 */
/* Same, with each component's z = M_c ((x - o_c) rate_c), the order of
 * sr_func (CEC2014.cc:1373-1393); rate NULL = 1. */
ORACLE_EXPORT void oracle_composition_rated(const float *x, int ldx, int n, int ncomp,
                                            const int *objectives, const double *os,
                                            const double *Ms, const double *sigma,
                                            const double *bias, const double *lambda,
                                            const double *rate, int D, double *f_out) {
    double *xd = (double *)malloc(sizeof(double) * (size_t)D * 3);
    double *z = xd + D, *tmp = xd + 2 * D;
    double w[16], fit[16];
    for (int p = 0; p < n; ++p) {
        for (int j = 0; j < D; ++j) xd[j] = (double)x[(size_t)p * ldx + j];
        double wmax = 0.0, wsum = 0.0;
        for (int c = 0; c < ncomp; ++c) {
            const double *o = os + (size_t)c * D;
            const double *M = Ms + (size_t)c * D * D;
            const double r = rate ? rate[c] : 1.0;
            for (int j = 0; j < D; ++j) tmp[j] = (xd[j] - o[j]) * r;
            for (int i = 0; i < D; ++i) {
                double acc = 0.0;
                for (int j = 0; j < D; ++j) acc = acc + tmp[j] * M[(size_t)i * D + j];
                z[i] = acc;
            }
            fit[c] = lambda[c] * oracle_basic(objectives[c], z, D) + bias[c];
            double d = 0.0;
            for (int j = 0; j < D; ++j) d += pow(xd[j] - o[j], 2.0);
            if (d != 0.0)
                w[c] = pow(1.0 / d, 0.5) * exp(-d / 2.0 / D / pow(sigma[c], 2.0));
            else
                w[c] = 1.0e99;
            if (w[c] > wmax) wmax = w[c];
        }
        for (int c = 0; c < ncomp; ++c) wsum += w[c];
        if (wmax == 0.0) {
            for (int c = 0; c < ncomp; ++c) w[c] = 1.0;
            wsum = ncomp;
        }
        double fv = 0.0;
        for (int c = 0; c < ncomp; ++c) fv += w[c] / wsum * fit[c];
        f_out[p] = fv;
    }
    free(xd);
}

ORACLE_EXPORT void oracle_composition(const float *x, int ldx, int n, int ncomp,
                                      const int *objectives, const double *os, const double *Ms,
                                      const double *sigma, const double *bias,
                                      const double *lambda, int D, double *f_out) {
    oracle_composition_rated(x, ldx, n, ncomp, objectives, os, Ms, sigma, bias, lambda, NULL, D,
                             f_out);
}

/* ------------------------------------------- CEC2014 F1-F30 (CEC2014.cc) */

/* sr_func (CEC2014.cc:1395-1435): y = (x - o) * rate, then z = M y when
 * rotated (rotatefunc, :1382-1393, z_i = sum_j y_j M[i D + j]). */
static void sr(const double *x, const double *o, const double *M, double rate, int rotate, int D,
               double *z, double *tmp) {
    for (int j = 0; j < D; ++j) tmp[j] = (x[j] - o[j]) * rate;
    if (!rotate) {
        for (int j = 0; j < D; ++j) z[j] = tmp[j];
        return;
    }
    for (int i = 0; i < D; ++i) {
        double acc = 0.0;
        for (int j = 0; j < D; ++j) acc = acc + tmp[j] * M[(size_t)i * D + j];
        z[i] = acc;
    }
}

/* The six hybrid functions hf01-hf06 (CEC2014.cc:954-1207): fractions of
 * the dimension, each segment's objective and the rate it scales its
 * segment by (sr_func with s_flag = r_flag = 0 inside each segment). */
typedef struct {
    int nseg;
    double p[5];
    int obj[5];
    double rate[5];
} hybrid_def;

static const hybrid_def HYBRID[6] = {
    {3, {0.3, 0.3, 0.4}, {OBJ_SCHWEFEL, OBJ_RASTRIGIN, OBJ_ELLIPTIC},
     {1000.0 / 100.0, 5.12 / 100.0, 1.0}},                                    /* :954-991  */
    {3, {0.3, 0.3, 0.4}, {OBJ_BENT_CIGAR, OBJ_HGBAT, OBJ_RASTRIGIN},
     {1.0, 5.0 / 100.0, 5.12 / 100.0}},                                       /* :993-1032 */
    {4, {0.2, 0.2, 0.3, 0.3}, {OBJ_GRIEWANK, OBJ_WEIERSTRASS, OBJ_ROSENBROCK, OBJ_ESCAFFER6},
     {600.0 / 100.0, 0.5 / 100.0, 2.048 / 100.0, 1.0}},                       /* :1034-1075 */
    {4, {0.2, 0.2, 0.3, 0.3}, {OBJ_HGBAT, OBJ_DISCUS, OBJ_GRIE_ROSEN, OBJ_RASTRIGIN},
     {5.0 / 100.0, 1.0, 5.0 / 100.0, 5.12 / 100.0}},                          /* :1077-1117 */
    {5, {0.1, 0.2, 0.2, 0.2, 0.3},
     {OBJ_ESCAFFER6, OBJ_HGBAT, OBJ_ROSENBROCK, OBJ_SCHWEFEL, OBJ_ELLIPTIC},
     {1.0, 5.0 / 100.0, 2.048 / 100.0, 1000.0 / 100.0, 1.0}},                 /* :1119-1162 */
    {5, {0.1, 0.2, 0.2, 0.2, 0.2},
     {OBJ_KATSUURA, OBJ_HAPPYCAT, OBJ_GRIE_ROSEN, OBJ_SCHWEFEL, OBJ_ACKLEY},
     {5.0 / 100.0, 5.0 / 100.0, 5.0 / 100.0, 1000.0 / 100.0, 1.0}},           /* :1164-1207 */
};

/* Segment lengths of a hybrid: ceil(p_i D) in FP64 for all but the last,
 * which takes the rest (CEC2014.cc:961-967). */
ORACLE_EXPORT int oracle_hybrid_segments(int h, int D, int *len_out) {
    if (h < 1 || h > 6) return -1;
    const hybrid_def *H = &HYBRID[h - 1];
    int used = 0;
    for (int i = 0; i < H->nseg - 1; ++i) {
        len_out[i] = (int)ceil(H->p[i] * D);
        used += len_out[i];
    }
    len_out[H->nseg - 1] = D - used;
    return H->nseg;
}

/* Hybrid h (1..6) at x: z = M (x - o) (rate 1), y_i = z[S_i - 1] (S is
 * 1-based, as in the shuffle_data files), then the segments' objectives of
 * y_seg * rate_seg, summed (CEC2014.cc:970-990). */
static double hybrid(int h, const double *x, const double *o, const double *M, const int *S,
                     int D, double *z, double *tmp) {
    const hybrid_def *H = &HYBRID[h - 1];
    int len[5];
    oracle_hybrid_segments(h, D, len);
    sr(x, o, M, 1.0, 1, D, z, tmp);
    for (int i = 0; i < D; ++i) tmp[i] = z[S[i] - 1];
    double f = 0.0;
    int start = 0;
    for (int s = 0; s < H->nseg; ++s) {
        for (int i = 0; i < len[s]; ++i) z[i] = tmp[start + i] * H->rate[s];
        f += oracle_basic(H->obj[s], z, len[s]);
        start += len[s];
    }
    return f;
}

/* cf_cal (CEC2014.cc:1481-1519) over fit[] (bias added first). */
static double cf_combine(const double *x, const double *os, int D, const double *delta,
                         const double *bias, double *fit, int n) {
    double w[8], wmax = 0.0, wsum = 0.0;
    for (int c = 0; c < n; ++c) {
        fit[c] += bias[c];
        double d = 0.0;
        for (int j = 0; j < D; ++j) d += pow(x[j] - os[(size_t)c * D + j], 2.0);
        w[c] = d != 0.0 ? pow(1.0 / d, 0.5) * exp(-d / 2.0 / D / pow(delta[c], 2.0)) : 1.0e99;
        if (w[c] > wmax) wmax = w[c];
    }
    for (int c = 0; c < n; ++c) wsum += w[c];
    if (wmax == 0.0) {
        for (int c = 0; c < n; ++c) w[c] = 1.0;
        wsum = n;
    }
    double f = 0.0;
    for (int c = 0; c < n; ++c) f += w[c] / wsum * fit[c];
    return f;
}

/* Components of the compositions cf01-cf06 (CEC2014.cc:1209-1339): objective,
 * rate, rotation flag, and the scaling num * g / den applied before cf_cal. */
typedef struct {
    int obj;
    double rate;
    int rotate;
    double num, den;
} comp_def;

typedef struct {
    int n;
    comp_def c[5];
    double delta[5], bias[5];
} cf_def;

static const cf_def CF[6] = {
    {5,
     {{OBJ_ROSENBROCK, 2.048 / 100.0, 1, 10000, 1e+4}, {OBJ_ELLIPTIC, 1, 1, 10000, 1e+10},
      {OBJ_BENT_CIGAR, 1, 1, 10000, 1e+30}, {OBJ_DISCUS, 1, 1, 10000, 1e+10},
      {OBJ_ELLIPTIC, 1, 0, 10000, 1e+10}},
     {10, 20, 30, 40, 50}, {0, 100, 200, 300, 400}},                         /* cf01 :1209 */
    {3,
     {{OBJ_SCHWEFEL, 1000.0 / 100.0, 0, 1, 1}, {OBJ_RASTRIGIN, 5.12 / 100.0, 1, 1, 1},
      {OBJ_HGBAT, 5.0 / 100.0, 1, 1, 1}},
     {20, 20, 20}, {0, 100, 200}},                                            /* cf02 :1235 */
    {3,
     {{OBJ_SCHWEFEL, 1000.0 / 100.0, 1, 1000, 4e+3}, {OBJ_RASTRIGIN, 5.12 / 100.0, 1, 1000, 1e+3},
      {OBJ_ELLIPTIC, 1, 1, 1000, 1e+10}},
     {10, 30, 50}, {0, 100, 200}},                                            /* cf03 :1251 */
    {5,
     {{OBJ_SCHWEFEL, 1000.0 / 100.0, 1, 1000, 4e+3}, {OBJ_HAPPYCAT, 5.0 / 100.0, 1, 1000, 1e+3},
      {OBJ_ELLIPTIC, 1, 1, 1000, 1e+10}, {OBJ_WEIERSTRASS, 0.5 / 100.0, 1, 1000, 400},
      {OBJ_GRIEWANK, 600.0 / 100.0, 1, 1000, 100}},
     {10, 10, 10, 10, 10}, {0, 100, 200, 300, 400}},                         /* cf04 :1269 */
    {5,
     {{OBJ_HGBAT, 5.0 / 100.0, 1, 10000, 1000}, {OBJ_RASTRIGIN, 5.12 / 100.0, 1, 10000, 1e+3},
      {OBJ_SCHWEFEL, 1000.0 / 100.0, 1, 10000, 4e+3},
      {OBJ_WEIERSTRASS, 0.5 / 100.0, 1, 10000, 400}, {OBJ_ELLIPTIC, 1, 1, 10000, 1e+10}},
     {10, 10, 10, 20, 20}, {0, 100, 200, 300, 400}},                         /* cf05 :1293 */
    {5,
     {{OBJ_GRIE_ROSEN, 5.0 / 100.0, 1, 10000, 4e+3}, {OBJ_HAPPYCAT, 5.0 / 100.0, 1, 10000, 1e+3},
      {OBJ_SCHWEFEL, 1000.0 / 100.0, 1, 10000, 4e+3}, {OBJ_ESCAFFER6, 1, 1, 10000, 2e+7},
      {OBJ_ELLIPTIC, 1, 1, 10000, 1e+10}},
     {10, 20, 30, 40, 50}, {0, 100, 200, 300, 400}},                         /* cf06 :1317 */
};

/* CEC2014 function func_id (1..30) as CEC2014::EvaluateFitness
 * (CEC2014.cc:14-145) without the 100 * func_id it adds and removes:
 *   F1-F16   the basic functions with the rates and rotation flags of :17-79;
 *   F17-F22  hf01-hf06 (:80-97);   F23-F28  cf01-cf06 (:98-115);
 *   F29, F30 cf07 / cf08 (:116-121): hf01-03 / hf04-06 per component, each on
 *            its own o, M and shuffle, weighted by cf_cal with delta {10, 30, 50}
 *            and bias {0, 100, 200} (:1341-1371).
 * os [ncomp][D], Ms [ncomp][D][D], SS [ncomp][D] (1-based; NULL unless
 * F17-F22, F29, F30).  Returns 0, or -1 for an unknown func_id. */
ORACLE_EXPORT int oracle_cec(int func_id, const double *x, int ldx, int n, int D,
                             const double *os, const double *Ms, const int *SS, double *f_out) {
    static const struct {
        int obj;
        double rate;
        int rotate;
    } BASIC[16] = {{OBJ_ELLIPTIC, 1, 1},           {OBJ_BENT_CIGAR, 1, 1},
                   {OBJ_DISCUS, 1, 1},             {OBJ_ROSENBROCK, 2.048 / (double)100, 1},
                   {OBJ_ACKLEY, 1, 1},             {OBJ_WEIERSTRASS, 0.5 / (double)100, 1},
                   {OBJ_GRIEWANK, 6, 1},           {OBJ_RASTRIGIN, 5.12 / (double)100, 0},
                   {OBJ_RASTRIGIN, 5.12 / (double)100, 1}, {OBJ_SCHWEFEL, 10, 0},
                   {OBJ_SCHWEFEL, 10, 1},          {OBJ_KATSUURA, 5 / (double)100, 1},
                   {OBJ_HAPPYCAT, 5 / (double)100, 1}, {OBJ_HGBAT, 5 / (double)100, 1},
                   {OBJ_GRIE_ROSEN, 5 / (double)100, 1}, {OBJ_ESCAFFER6, 1, 1}};
    if (func_id < 1 || func_id > 30) return -1;
    double *buf = (double *)malloc(sizeof(double) * 2 * (size_t)D);
    double *z = buf, *tmp = buf + D;
    for (int p = 0; p < n; ++p) {
        const double *x_p = x + (size_t)p * ldx;
        double f;
        if (func_id <= 16) {
            sr(x_p, os, Ms, BASIC[func_id - 1].rate, BASIC[func_id - 1].rotate, D, z, tmp);
            f = oracle_basic(BASIC[func_id - 1].obj, z, D);
        } else if (func_id <= 22) {
            f = hybrid(func_id - 16, x_p, os, Ms, SS, D, z, tmp);
        } else if (func_id <= 28) {
            const cf_def *C = &CF[func_id - 23];
            double fit[5];
            for (int c = 0; c < C->n; ++c) {
                sr(x_p, os + (size_t)c * D, Ms + (size_t)c * D * D, C->c[c].rate, C->c[c].rotate,
                   D, z, tmp);
                fit[c] = C->c[c].num * oracle_basic(C->c[c].obj, z, D) / C->c[c].den;
            }
            f = cf_combine(x_p, os, D, C->delta, C->bias, fit, C->n);
        } else {
            static const double delta[3] = {10, 30, 50}, bias[3] = {0, 100, 200};
            const int h0 = func_id == 29 ? 1 : 4;
            double fit[3];
            for (int c = 0; c < 3; ++c)
                fit[c] = hybrid(h0 + c, x_p, os + (size_t)c * D, Ms + (size_t)c * D * D,
                                SS + (size_t)c * D, D, z, tmp);
            f = cf_combine(x_p, os, D, delta, bias, fit, 3);
        }
        f_out[p] = f;
    }
    free(buf);
    return 0;
}

/* Components per function: 5 for F23, F26-F28; 3 for F24, F25, F29, F30;
 * else 1 (CEC2014.cc:304-330). */
ORACLE_EXPORT int oracle_cec_ncomp(int func_id) {
    if (func_id == 23 || (func_id >= 26 && func_id <= 28)) return 5;
    if (func_id == 24 || func_id == 25 || func_id == 29 || func_id == 30) return 3;
    return 1;
}

/* ------------------------------------------------------- island driver */

/*
 * Ring migration among islands held in one population array (replaces the
 * reference's MPI migration, src/migrate.cc:46-144).  Islands are ordered by
 * the total order "ascending fitness, then ascending row", NaN after every
 * number.  Island i's first k rows in that order (its best) replace the last k
 * rows (its worst) of island dst[i] (dst[i] < 0: island i sends nothing), all
 * chosen from the state before any copy.  Fitness travels with the row, so
 * dst[i] must solve the same problem (objective instance and dimension) as i.
 * Islands: rows off[i]..off[i+1]-1 of X (stride ld) and fX, dimension dims[i].
 */
static int before(float a, int ia, float b, int ib) {
    int an = isnan(a), bn = isnan(b);
    if (an != bn) return bn;        /* a number sorts before NaN */
    if (!an && a != b) return a < b;
    return ia < ib;
}

ORACLE_EXPORT void oracle_migrate_ring(float *X, float *fX, int ld, const int *off,
                                       const int *dims, const int *dst, int I, int k) {
    int total = off[I];
    float *Xs = (float *)malloc(sizeof(float) * (size_t)total * ld);
    float *fs = (float *)malloc(sizeof(float) * (size_t)total);
    memcpy(Xs, X, sizeof(float) * (size_t)total * ld);
    memcpy(fs, fX, sizeof(float) * (size_t)total);
    int *best = (int *)malloc(sizeof(int) * (size_t)I * k);
    int *worst = (int *)malloc(sizeof(int) * (size_t)I * k);
    for (int i = 0; i < I; ++i) {
        int n = off[i + 1] - off[i];
        char *used_b = (char *)calloc((size_t)n, 1), *used_w = (char *)calloc((size_t)n, 1);
        for (int t = 0; t < k; ++t) {
            int bi = -1, wi = -1;
            for (int r = 0; r < n; ++r) {
                int g = off[i] + r;
                if (!used_b[r] && (bi < 0 || before(fs[g], g, fs[off[i] + bi], off[i] + bi))) bi = r;
                if (!used_w[r] && (wi < 0 || before(fs[off[i] + wi], off[i] + wi, fs[g], g))) wi = r;
            }
            used_b[bi] = 1; used_w[wi] = 1;
            best[i * k + t] = off[i] + bi;
            worst[i * k + t] = off[i] + wi;
        }
        free(used_b); free(used_w);
    }
    for (int i = 0; i < I; ++i) {
        if (dst[i] < 0) continue;
        for (int t = 0; t < k; ++t) {
            int s = best[i * k + t], d = worst[dst[i] * k + t];
            memcpy(X + (size_t)d * ld, Xs + (size_t)s * ld, sizeof(float) * (size_t)dims[i]);
            fX[d] = fs[s];
        }
    }
    free(Xs); free(fs); free(best); free(worst);
}
