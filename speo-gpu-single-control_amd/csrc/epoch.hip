// epoch.hip -- the epoch kernel: one workgroup per island runs all the
// generations between two migrations without leaving the chip.
//
// Reference counterpart: the generation loop of IslandEA::Execute
// (src/island_EA.cc:84-92) around DE_CUDA::Run (CUDA_EA_library/DE_CUDA.cc:
// 200-216), whose three launches per generation (generate, evaluate, select)
// and their HBM round trips become phases of one workgroup separated by
// barriers.  Per generation and island:
//   1. donors: one lane per row draws the five donors and j_rand
//      (kernel_DE_CUDA.cu:187-222) into LDS;
//   2. trials: (row, 4-coordinate group) items, DE/rand/1/bin with the
//      crossover draws of Philox block {group, row, generation, 1}
//      (kernel_DE_CUDA.cu:139-158), written to the row's free slot;
//   3. evaluation: z = rate M (x - o) on MFMA with the rotation's B fragments
//      held in registers for the whole epoch, per-coordinate terms reduced by
//      wave shuffles (kernel_CEC2014_CUDA.cu:147-181, 254-286, 389-442,
//      557-587);
//   4. selection: strict "<" (kernel_DE_CUDA.cu:124-136) flips the row's slot
//      bit.
// The arithmetic of every step is the per-generation kernels' (the column
// split of the rotation over four waves and the order of every sum are the
// same), so an epoch launch reproduces the plain engine bit for bit.
#pragma clang fp contract(off)

#include <mutex>

#include "engine.h"

namespace speo {

namespace {

__device__ __forceinline__ float4 ld4(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4(float *p, float4 v) { *reinterpret_cast<float4 *>(p) = v; }
__device__ __forceinline__ float comp4(const float4 &v, int q) {
    return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}

constexpr int kFixedLds = 128 * 4 + 128 * 4;  // shift vector, zero pad after the rows
// per pass row: donor offsets, j_rand + trial row, A row, partials
constexpr int kPassRowBytes = 16 + 8 + 4 + 32;


int lds_row_stride(int dim) {
    int S = round_up(dim, 4);
    if (S % 32 == 0) S += 4;  // consecutive rows start in different LDS banks
    return S;
}

// Row stride of trial rows (the evaluation's A operand): the FP32 evaluation
// reads K steps in blocks of four, up to column 16 ceil(dim / 16) - 1, so the
// rows carry zeroed padding to a multiple of 16 columns (+4 for the banks).
int lds_trial_stride(int dim) { return round_up(round_up(dim, 4), 16) + 4; }

// LDS of k_island_split: fX, fU, flags of the share [own], shift [128],
// donor offsets [own] int4, j_rand [own], partials [16][own], rows [NPp][S],
// trial rows and their shifted, scaled copies [own][ST] each.
// Accumulators of column tiles 4..7 handed to the waves of tiles 0..3: four
// waves x two 16-row tiles x 64 lanes x 4 floats.
constexpr int kSplitXchBytes = 4 * 2 * 64 * 16;

// The A operand region holds the FP32 copies (parts16 == 0) or the 16-bit
// parts of every element [parts16][own][kp32 + 8].
int split_a_bytes(int dim, int own, int parts16) {
    return parts16 == 0 ? own * lds_trial_stride(dim) * 4 : parts16 * own * (round_up(dim, 32) + 8) * 2;
}
int split_lds_bytes(int np, int dim, int own, int parts16) {
    const int NPp = round_up(np, 16), S = lds_row_stride(dim), ST = lds_trial_stride(dim);
    return own * (4 + 4 + 4 + 16 + 4 + 8 * 4) + 128 * 4 + kSplitXchBytes + (NPp * S + own * ST) * 4 +
           split_a_bytes(dim, own, parts16);
}

}  // namespace

int precision_parts(int precision) {
    switch (precision) {
    case SPEO_PREC_BF16: return Prec<SPEO_PREC_BF16>::parts;
    case SPEO_PREC_FP16: return Prec<SPEO_PREC_FP16>::parts;
    case SPEO_PREC_BF16X3: return Prec<SPEO_PREC_BF16X3>::parts;
    case SPEO_PREC_FP16X3: return Prec<SPEO_PREC_FP16X3>::parts;
    case SPEO_PREC_BF16X6: return Prec<SPEO_PREC_BF16X6>::parts;
    default: return 0;
    }
}

EpochPlan epoch_plan(int np, int dim, int parts, int parts16) {
    EpochPlan p{};
    p.S = lds_row_stride(dim);
    p.ST = lds_trial_stride(dim);
    p.parts = parts;
    const int NPp = round_up(np, 16);
    const int own = parts > 1 ? round_up((np + parts - 1) / parts, 16) : NPp;
    p.own_rows = own;
    const int fixed = NPp * 8 + round_up(NPp, 16) + kFixedLds;  // fX, fU, flags, shift, pad
    int in_lds = fixed + own * kPassRowBytes + (NPp * p.S + own * p.ST) * 4;
    if (parts > 1) in_lds = split_lds_bytes(np, dim, own, parts16);  // the split kernel's layout
    if (in_lds <= kEpochLdsMax) {
        p.in_lds = 1;
        p.pass_rows = own;
        p.lds_bytes = in_lds;
        return p;
    }
    if (parts > 1) {
        p.lds_bytes = kEpochLdsMax + 1;
        return p;
    }
    int pr = (kEpochLdsMax - fixed) / (kPassRowBytes + 4 * p.ST);
    pr = std::min(pr / 16 * 16, NPp);
    p.in_lds = 0;
    p.pass_rows = pr;
    p.lds_bytes = pr >= 16 ? fixed + pr * (kPassRowBytes + 4 * p.ST) : kEpochLdsMax + 1;
    return p;
}

// Donors and j_rand of generations [gen0, gen0 + L) for every row: one
// (generation, row) per lane, out[g * rows + row] = {r0, r1, r2, j_rand}
// (island-local).  They do not depend on the population, so a whole chunk is
// drawn ahead of the epoch launches.
__global__ void k_epoch_donors(const IslandDesc *isl, const int *row_island, int rows, int L,
                               uint64_t seed, const uint32_t *gen_base, uint32_t gen_offset,
                               int4 *out) {
    const int64_t total = (int64_t)rows * L;
    const uint32_t gen0 = *gen_base + gen_offset;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int g = (int)(t / rows), row = (int)(t - (int64_t)g * rows);
        const IslandDesc I = isl[row_island[row]];
        DonorStream ds(seed, (uint32_t)row, gen0 + (uint32_t)g);
        int r[kMaxR], jr;
        draw_donors(ds, nullptr, 0, I.np, I.dim, r, jr);
        out[t] = int4{r[0], r[1], r[2], jr};
    }
}

int epoch_donors(const IslandDesc *isl, const int *row_island, int rows, int L, uint64_t seed,
                 const uint32_t *gen_base, uint32_t gen_offset, int4 *out, hipStream_t s) {
    const int64_t total = (int64_t)rows * L;
    const int blocks = (int)std::min<int64_t>((total + 255) / 256, 4096);
    hipLaunchKernelGGL(k_epoch_donors, dim3(blocks), dim3(256), 0, s, isl, row_island, rows, L, seed,
                       gen_base, gen_offset, out);
    return check_launch("k_epoch_donors");
}

// Per-coordinate terms of one 16-row tile's two column tiles, summed over the
// 16 lanes of each row (gemm_small's epilogue with the objective fixed): the
// lane's terms of column tile t0 then t0 + 4, then the 16 lanes.
template <int OBJ>
__device__ __forceinline__ void tile_terms(const floatx4 &acc0, const floatx4 &acc1, bool c0, bool c1,
                                           float w0, float w1, float unscale, float (&s1)[4],
                                           float (&s2)[4]) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        s1[q] = 0.f;
        s2[q] = 0.f;
    }
    if (c0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) add_terms(OBJ, acc0[q] * unscale, w0, s1[q], s2[q]);
    }
    if (c1) {
#pragma unroll
        for (int q = 0; q < 4; ++q) add_terms(OBJ, acc1[q] * unscale, w1, s1[q], s2[q]);
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        s1[q] = sum16(s1[q]);
        if (OBJ == SPEO_OBJ_ACKLEY) s2[q] = sum16(s2[q]);
    }
}

// One workgroup per island, all generations of an epoch (islands split over
// several workgroups run k_island_split instead, one generation per launch).
template <int PREC>
__global__ __launch_bounds__(kEpochThreads) void k_island_epoch(EpochArgs a) {
    extern __shared__ __align__(16) char lds[];
    constexpr int P = Prec<PREC>::parts;
    constexpr bool H = Prec<PREC>::half;
    const EpochBlock blk = a.blocks[blockIdx.x];
    const IslandDesc &I = blk.I;
    const EpochPlan &pl = blk.pl;
    const int share = blk.share;
    const float lambda = blk.lambda, bias = blk.bias;
    const int NP = I.np, NPp = (NP + 15) & ~15, D = I.dim, S = pl.S, ST = pl.ST, PR = pl.pass_rows;
    const bool in_lds = pl.in_lds != 0;
    // the workgroup's share of the rows: [own0, own0 + nown)
    const int OWN = in_lds ? pl.own_rows : NPp;
    const int own0 = in_lds ? min(NP, share * OWN) : 0;
    const int nown = in_lds ? min(NP, own0 + OWN) - own0 : NP;
    const int ng = (D + 3) >> 2;                        // float4 groups per row
    // it / ng as umulhi(it, ceil(2^32 / ng)): exact for it < 2^32 / ng (it < 2^18 here)
    const uint32_t inv_ng = blk.inv_ng;
    auto div_ng = [&](int it) -> int { return ng > 1 ? (int)__umulhi((uint32_t)it, inv_ng) : it; };
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int wq = wave & 3, grp = wave >> 2, lr = lane & 15, lk = lane >> 4;
    const uint32_t crt = blk.crt;
    const uint32_t gen0 = *a.gen_base + a.gen_offset;
    // The launch arguments and descriptor fields the prologue needs, and the
    // generation counter, as one batch of scalar loads with one wait (the
    // compiler would otherwise issue each at its first use and wait for it
    // there: one memory latency per argument).
    asm volatile("" ::"s"(a.Xin), "s"(a.fXin), "s"(a.donors), "s"(a.rows), "s"(a.donor_gen), "s"(gen0),
                 "s"(blk.B32s), "s"(blk.shift), "s"(blk.colw), "s"(I.elem_off), "s"(I.row_off), "s"(NP),
                 "s"(D), "s"(pl.S));

    // LDS: fX[NPp] fU[NPp] flags[NPp] shift[128] doff[PR] tinfo[PR] aoff[PR]
    //      part[8][PR] rows (+128 floats)
    float *fXs = reinterpret_cast<float *>(lds);
    float *fUs = fXs + NPp;
    uint8_t *flags = reinterpret_cast<uint8_t *>(fUs + NPp);  // HBM plan: bit 0 current slot, bit 1
                                                              // last trial's; LDS plan: accepted
    float *sh = reinterpret_cast<float *>(flags + ((NPp + 15) & ~15));
    int4 *doff = reinterpret_cast<int4 *>(sh + 128);     // rows of e0, e1, e2 and the target
    int2 *tinfo = reinterpret_cast<int2 *>(doff + PR);     // j_rand, trial row
    int *aoff = reinterpret_cast<int *>(tinfo + PR);       // LDS row of the trial to evaluate
    float *part = reinterpret_cast<float *>(aoff + PR);
    // LDS plan: current rows [NPp][S], then the share's trial rows [OWN][ST];
    // HBM plan: the pass's trials [PR][ST]
    float *rows = part + 8 * PR;
    const int trial0 = NPp * S;

    // Row addresses: in LDS a float offset into `rows`; in HBM 2 * (r * ld) + slot.
    // HBM plan (one share, Xin == Xout): slot 0 is the population itself
    float *gX = a.Xout + I.elem_off, *gA = a.Xalt + I.elem_off;
    const float *gXin = a.Xin + I.elem_off;
    auto row_code = [&](int slot, int r) -> int {
        return in_lds ? (slot ? trial0 + (r - own0) * ST : r * S) : 2 * r * I.ld + slot;
    };
    auto grow = [&](int code) -> float * { return ((code & 1) ? gA : gX) + (code >> 1); };

    // Prologue: every global load is issued before the first barrier (the
    // rotation's register fragments, a first batch of the population, the
    // shift, the fitness), LDS is zeroed meanwhile, and the stores follow.
    //
    // The rotation's fragments of this wave's column tiles t0, t0 + 4, held
    // for the whole epoch (the four waves of a group split the columns); the
    // FP32 ones come from the fragment-ordered B32s as 16-byte loads, so that
    // with the population batch they stay within a wave's outstanding loads.
    const int t0 = wq, t1 = wq + 4;
    const bool v0 = t0 * 16 < D, v1 = t1 * 16 < D;
    const bool c0 = t0 * 16 + lr < D, c1 = t1 * 16 + lr < D;
    const bool ell = blk.objective == SPEO_OBJ_ELLIPTIC;
    const int KS = P == 0 ? ng : (blk.kp32 >> 5);
    typedef typename std::conditional<H, halfx8, bf16x8>::type vec8;
    typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
    constexpr int NB32 = P == 0 ? 32 : 1;
    constexpr int NB16 = P == 0 ? 1 : P;
    float b0[NB32], b1[NB32], shv[NB32];
    vec8 bb[2][4][NB16];
    if constexpr (P == 0) {
        load_b_frag(blk.B32s, t0, v0, KS, lane, b0);
        load_b_frag(blk.B32s, t1, v1, KS, lane, b1);
    } else {
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    const bool ok = (t == 0 ? v0 : v1) && ks < KS;
                    bb[t][ks][p] = ok ? gld(reinterpret_cast<const vec8 *>(
                                            blk.M16s[p] + (size_t)((t == 0 ? t0 : t1) * 16 + lr) * blk.kp32 +
                                            ks * 32 + lk * 8))
                                      : vec8{};
                }
    }
    const float w0 = (ell && c0) ? gld(blk.colw + t0 * 16 + lr) : 1.0f;
    const float w1 = (ell && c1) ? gld(blk.colw + t1 * 16 + lr) : 1.0f;
    // The island's rows are one contiguous block of NP * 4 ng floats (the row
    // stride is round_up(D, 4) when no island is wider than 128).  They are
    // read through a buffer resource over that block: offsets past it return
    // zeros, so the batch needs no bounds branches and no per-load index
    // arithmetic; with the LDS row stride equal to D (D % 4 == 0 and D not a
    // multiple of 32) the stores are linear too.
    constexpr int kBatch = 16;
    const int total = in_lds ? NP * ng : 0;
    const bool linear = (D & 3) == 0 && S == D;
    floatx4 pre[kBatch];
    {
        // (a load of batch b is skipped when the whole batch lies past the
        // block: a wave-uniform branch around it, no value merged after it)
        const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(gXin), (short)0, total * 16, kBufferWord3);
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
            if (b * kEpochThreads < total)
                pre[b] = __builtin_bit_cast(
                    floatx4, __builtin_amdgcn_raw_buffer_load_b128(xr, (tid + b * kEpochThreads) * 16, 0, 0));
    }
    const float shift_k = tid < 128 && tid < D ? gld(blk.shift + tid) : 0.f;
    const float fx_r = tid < NP ? a.fXin[I.row_off + tid] : INFINITY;
    // donors of the first pass of the first generation
    const int nr_first = min(PR, nown);
    int4 d_first = int4{0, 0, 0, 0};
    if (tid < nr_first) d_first = a.donors[(size_t)a.donor_gen * a.rows + I.row_off + own0 + tid];
    if (tid < 128) sh[tid] = shift_k;
    for (int r = tid; r < NPp; r += kEpochThreads) {
        fXs[r] = r == tid ? fx_r : (r < NP ? a.fXin[I.row_off + r] : INFINITY);
        fUs[r] = INFINITY;
        flags[r] = 0;
    }
    // Trial rows: columns [4 ng, ST) are zeroed once (trials write the first
    // 4 ng).  Rows land in LDS with their padding columns zeroed; nothing
    // else in LDS is read before it is written.
    {
        const int pad = ST - 4 * ng, nt = in_lds ? OWN : PR, t0f = in_lds ? trial0 : 0;
        for (int e = tid; e < nt * pad; e += kEpochThreads) {
            const int r = e / pad;
            rows[t0f + r * ST + 4 * ng + (e - r * pad)] = 0.f;
        }
    }
    auto put_row_group = [&](int it, float4 x) {
        if (linear) {
            st4(rows + 4 * it, x);
            return;
        }
        const int r = div_ng(it), q4 = it - r * ng, j = 4 * q4;
        if (j + 1 >= D) x.y = 0.f;
        if (j + 2 >= D) x.z = 0.f;
        if (j + 3 >= D) x.w = 0.f;
        st4(rows + (size_t)r * S + 4 * q4, x);
    };
#pragma unroll
    for (int b = 0; b < kBatch; ++b) {
        const int it = tid + b * kEpochThreads;
        if (it < total) put_row_group(it, float4{pre[b][0], pre[b][1], pre[b][2], pre[b][3]});
    }
    for (int it = tid + kBatch * kEpochThreads; it < total; it += kEpochThreads)
        put_row_group(it, ld4(gXin + 4 * it));
    __syncthreads();
    if constexpr (P == 0) {
#pragma unroll
        for (int ks = 0; ks < 32; ++ks) shv[ks] = sh[ks * 4 + lk];
    }

    const int L = a.L;
    const int npass = (nown + PR - 1) / PR;
    for (int g = 0; g < L; ++g) {
        const uint32_t gen = gen0 + (uint32_t)g;
        // 1. donors and j_rand of this generation, drawn ahead of the launch
        const int4 *dn = a.donors + (size_t)(a.donor_gen + g) * a.rows + I.row_off;
        for (int pass = 0; pass < npass; ++pass) {
            const int p0 = own0 + pass * PR;
            const int nr = min(PR, own0 + nown - p0);
            // 2. the pass's rows: rows of the donors and the target, trial row
            for (int rl = tid; rl < nr; rl += kEpochThreads) {
                const int r = p0 + rl;
                const int4 d = (g == 0 && p0 == own0 && rl == tid) ? d_first : dn[r];
                if (in_lds) {
                    doff[rl] = int4{d.x * S, d.y * S, d.z * S, r * S};
                    tinfo[rl] = int2{d.w, trial0 + (r - own0) * ST};
                    aoff[rl] = trial0 + (r - own0) * ST;
                } else {
                    const int cur = flags[r] & 1;
                    doff[rl] = int4{row_code(flags[d.x] & 1, d.x), row_code(flags[d.y] & 1, d.y),
                                    row_code(flags[d.z] & 1, d.z), row_code(cur, r)};
                    tinfo[rl] = int2{d.w, row_code(cur ^ 1, r)};
                    aoff[rl] = rl * ST;
                }
            }
            __syncthreads();
            // 3. trials (the LDS plan and the HBM plan as separate loops, so that
            // each compiles to plain LDS or global accesses).  Two items per
            // lane and iteration, the operands of both fetched before either
            // is made.
            struct Operands {
                int rl, q4;
                int2 ti;  // j_rand, trial row
                float4 e0, e1, e2, xo;
            };
            auto trials = [&](auto row_of, auto store) {
                const int n_items = nr * ng;
                auto fetch = [&](int it) {
                    Operands o;
                    o.rl = div_ng(it);
                    o.q4 = it - o.rl * ng;
                    const int4 d = doff[o.rl];
                    o.ti = tinfo[o.rl];
                    o.e0 = ld4(row_of(d.x) + 4 * o.q4);
                    o.e1 = ld4(row_of(d.y) + 4 * o.q4);
                    o.e2 = ld4(row_of(d.z) + 4 * o.q4);
                    o.xo = ld4(row_of(d.w) + 4 * o.q4);
                    return o;
                };
                // the four coordinates branch-free; the reflective bound only
                // when a lane holds a coordinate out of bounds
                auto make = [&](const Operands &o, const U4 &w) {
                    float v[4];
                    bool out = false;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int j = 4 * o.q4 + q;
                        // (bitwise, not short-circuit, operators: selects, no branches)
                        const bool take = (j == o.ti.x) | cross_take(word(w, q), crt);
                        const float mt = fmaf(I.F, comp4(o.e1, q) - comp4(o.e2, q), comp4(o.e0, q));
                        v[q] = j < D ? (take ? mt : comp4(o.xo, q)) : 0.f;
                        out |= (j < D) & ((v[q] < a.lo) | (v[q] > a.hi));
                    }
                    if (out) {
#pragma unroll
                        for (int q = 0; q < 4; ++q)
                            if (4 * o.q4 + q < D) v[q] = check_bound(v[q], a.lo, a.hi);
                    }
                    return float4{v[0], v[1], v[2], v[3]};
                };
                for (int it = tid; it < n_items; it += 2 * kEpochThreads) {
                    const int it2 = it + kEpochThreads;
                    const bool two = it2 < n_items;
                    const Operands A = fetch(it), B = fetch(two ? it2 : it);
                    const U4 wa = philox_block(a.seed, (uint32_t)A.q4, (uint32_t)(I.row_off + p0 + A.rl), gen,
                                               kPurposeCross);
                    const U4 wb = philox_block(a.seed, (uint32_t)B.q4, (uint32_t)(I.row_off + p0 + B.rl), gen,
                                               kPurposeCross);
                    const float4 va = make(A, wa), vb = make(B, wb);
                    store(A.rl, A.ti.y, A.q4, va);
                    if (two) store(B.rl, B.ti.y, B.q4, vb);
                }
            };
            if (in_lds) {
                trials([&](int code) -> const float * { return rows + code; },
                       [&](int, int code, int q4, float4 v) { st4(rows + code + 4 * q4, v); });
            } else {
                trials([&](int code) -> const float * { return grow(code); },
                       [&](int rl, int code, int q4, float4 v) {
                           st4(rows + (size_t)rl * ST + 4 * q4, v);
                           st4(grow(code) + 4 * q4, v);
                       });
            }
            __syncthreads();
            // 4. evaluation of the pass's trials, 16-row tiles, two tiles at a time
            const int ntile = (nr + 15) >> 4;
            for (int tile = grp; tile < ntile; tile += 2) {
                const float *arow = rows + aoff[min(tile * 16 + lr, nr - 1)];
                floatx4 acc0 = floatx4{0.f, 0.f, 0.f, 0.f}, acc1 = floatx4{0.f, 0.f, 0.f, 0.f};
                if constexpr (P == 0) {
                    // k steps in blocks of four (A read at fixed offsets; the
                    // steps of the last block past KS read the trial rows'
                    // zeroed padding, B = 0 there, and add exact zeros).
                    const float rate = blk.rate;
                    const float *ap = arow + lk;
                    float av[32];
#pragma unroll
                    for (int kb = 0; kb < 32; kb += 4) {
                        if (kb >= KS) break;
#pragma unroll
                        for (int u = 0; u < 4; ++u) av[kb + u] = (ap[(kb + u) * 4] - shv[kb + u]) * rate;
                    }
                    if (v1) {
#pragma unroll
                        for (int kb = 0; kb < 32; kb += 4) {
                            if (kb >= KS) break;
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[kb + u], b0[kb + u], acc0, 0, 0, 0);
                                acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[kb + u], b1[kb + u], acc1, 0, 0, 0);
                            }
                        }
                    } else if (v0) {
#pragma unroll
                        for (int kb = 0; kb < 32; kb += 4) {
                            if (kb >= KS) break;
#pragma unroll
                            for (int u = 0; u < 4; ++u)
                                acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(av[kb + u], b0[kb + u], acc0, 0, 0, 0);
                        }
                    }
                } else {
                    const float scale = blk.rate * (H ? kF16ScaleA : 1.0f);
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        if (ks < KS) {
                            const int kb = ks * 32 + lk * 8;
                            float xv[8];
#pragma unroll
                            for (int h = 0; h < 2; ++h) {
                                const float4 x4 = ld4(arow + kb + 4 * h);
                                xv[4 * h + 0] = x4.x;
                                xv[4 * h + 1] = x4.y;
                                xv[4 * h + 2] = x4.z;
                                xv[4 * h + 3] = x4.w;
                            }
                            u16x8 pb[NB16];
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const int k = kb + e;
                                float rem = k < D ? (xv[e] - sh[k]) * scale : 0.f;
#pragma unroll
                                for (int p = 0; p < P; ++p) {
                                    const uint16_t bits = to_bits(rem, H);
                                    pb[p][e] = bits;
                                    rem -= from_bits(bits, H);
                                }
                            }
                            vec8 av[NB16];
#pragma unroll
                            for (int p = 0; p < P; ++p) av[p] = __builtin_bit_cast(vec8, pb[p]);
#pragma unroll
                            for (int t = 0; t < 2; ++t) {
                                if (t == 0 ? !v0 : !v1) continue;
                                floatx4 &acc = t == 0 ? acc0 : acc1;
#pragma unroll
                                for (int sum = P - 1; sum >= 0; --sum) {
#pragma unroll
                                    for (int pa = 0; pa <= sum; ++pa) {
                                        const int pbi = sum - pa;
                                        if (pa < P && pbi < P) {
                                            if constexpr (H)
                                                acc = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                                    av[pa], bb[t][ks][pbi], acc, 0, 0, 0);
                                            else
                                                acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                                    av[pa], bb[t][ks][pbi], acc, 0, 0, 0);
                                        }
                                    }
                                }
                            }
                        }
                    }
                }
                const float unscale = H ? kF16Unscale : 1.0f;
                float s1[4], s2[4];
                switch (blk.objective) {
                case SPEO_OBJ_ELLIPTIC:
                    tile_terms<SPEO_OBJ_ELLIPTIC>(acc0, acc1, c0, c1, w0, w1, unscale, s1, s2);
                    break;
                case SPEO_OBJ_RASTRIGIN:
                    tile_terms<SPEO_OBJ_RASTRIGIN>(acc0, acc1, c0, c1, w0, w1, unscale, s1, s2);
                    break;
                default:
                    tile_terms<SPEO_OBJ_ACKLEY>(acc0, acc1, c0, c1, w0, w1, unscale, s1, s2);
                    break;
                }
                if (lr == 0) {
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int rl = tile * 16 + lk * 4 + q;
                        part[wq * PR + rl] = s1[q];
                        part[(4 + wq) * PR + rl] = s2[q];
                    }
                }
            }
            __syncthreads();
            if (blk.heavy.j >= 0) {
                // the heavy column's FP64 term, one wave per row, parked in
                // the row's fU slot until the combine replaces it
                for (int rl = wave; rl < nr; rl += kEpochThreads / 64) {
                    const double h = heavy_term(blk.heavy, rows + aoff[rl], D, lane);
                    if (lane == 0) fUs[p0 + rl] = (float)h;
                }
                __syncthreads();
            }
            // the four column groups' partial sums, in column-group order
            for (int rl = tid; rl < nr; rl += kEpochThreads) {
                float s1 = 0.f, s2 = 0.f;
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    s1 += part[w * PR + rl];
                    s2 += part[(4 + w) * PR + rl];
                }
                if (blk.heavy.j >= 0) s1 = with_heavy(s1, fUs[p0 + rl]);
                fUs[p0 + rl] = fmaf(lambda, finish_objective(blk.objective, s1, s2, D), bias);
            }
            __syncthreads();
        }
        // 5. selection, strict "<"
        if (in_lds) {
            // the share's accepted trials are copied over their rows
            for (int r = own0 + tid; r < own0 + nown; r += kEpochThreads) {
                const bool acc = fUs[r] < fXs[r];
                if (acc) fXs[r] = fUs[r];
                flags[r] = acc ? 1 : 0;
            }
            __syncthreads();
            for (int it = tid; it < nown * ng; it += kEpochThreads) {
                const int rl = div_ng(it), q4 = it - rl * ng, r = own0 + rl;
                if (flags[r]) st4(rows + r * S + 4 * q4, ld4(rows + trial0 + rl * ST + 4 * q4));
            }
            __syncthreads();
        } else {
            // an accepted trial's slot becomes the row's current one
            for (int r = tid; r < NP; r += kEpochThreads) {
                const float fu = fUs[r];
                const int cur = flags[r] & 1, ts = cur ^ 1;
                const bool acc = fu < fXs[r];
                if (acc) fXs[r] = fu;
                flags[r] = (uint8_t)((acc ? ts : cur) | (ts << 1));
            }
            __syncthreads();
        }
    }

    // the share's rows (and, at the end of a chunk, its last trials) back to HBM
    float *gU = a.U + I.elem_off;
    for (int it = tid; it < nown * ng; it += kEpochThreads) {
        const int rl = div_ng(it), q4 = it - rl * ng, r = own0 + rl;
        float4 cur, tri;
        if (in_lds) {
            tri = ld4(rows + trial0 + rl * ST + 4 * q4);
            cur = ld4(rows + r * S + 4 * q4);
        } else {
            const int fl = flags[r];
            cur = ld4(grow(row_code(fl & 1, r)) + 4 * q4);
            tri = ld4(grow(row_code((fl >> 1) & 1, r)) + 4 * q4);
        }
        st4(gX + (size_t)r * I.ld + 4 * q4, cur);
        if (a.write_trials) st4(gU + (size_t)r * I.ld + 4 * q4, tri);
    }
    for (int r = own0 + tid; r < own0 + nown; r += kEpochThreads) {
        a.fXout[I.row_off + r] = fXs[r];
        if (a.write_trials) a.fU[I.row_off + r] = fUs[r];
    }
}


// One generation of islands split over several workgroups.  1024 threads:
// 16 waves, two per SIMD more than the general kernel, to hide the LDS and
// memory latency of the short, single-generation launch.
//   prologue: the island's rows (linear buffer loads), the share's fitness,
//     donors and the shift; waves 0..7 load the rotation fragments of their
//     column tile (t = wave) for the evaluation;
//   trials: one item (row, four coordinates) per lane; the trial row and the
//     evaluation's A operand go to LDS: the shifted, scaled copy (u - o) *
//     rate in FP32, or its 16-bit parts, formed as the general kernel and
//     gemm_small form them;
//   evaluation: waves 0..7, column tile t = wave, the 16-row tiles of the
//     share two at a time; waves 4..7 hand their accumulators to waves 0..3
//     through LDS, which form the terms of tiles t and t + 4 in the order of
//     every other kernel (the lane's terms of t, then of t + 4, then the 16
//     lanes);
//   combine (column groups 0..3 in order), selection, write-back of the
//   selected row.
constexpr int kSplitThreads = 1024;
typedef __attribute__((address_space(3))) void *split_lds_ptr;
// s_waitcnt vmcnt(0), expcnt and lgkmcnt left alone (gfx9 encoding: vmcnt in
// bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8)
constexpr int kWaitVmcnt0 = (7 << 4) | (15 << 8);

template <int PREC>
__global__ __launch_bounds__(kSplitThreads) void k_island_split(EpochArgs a) {
    extern __shared__ __align__(16) char lds[];
    constexpr int P = Prec<PREC>::parts;
    constexpr bool H = Prec<PREC>::half;
    typedef typename std::conditional<H, halfx8, bf16x8>::type vec8;
    const EpochBlock blk = a.blocks[blockIdx.x];
    const IslandDesc &I = blk.I;
    const EpochPlan &pl = blk.pl;
    const float lambda = blk.lambda, bias = blk.bias;
    const int NP = I.np, NPp = (NP + 15) & ~15, D = I.dim, S = pl.S, ST = pl.ST;
    const int OWN = pl.own_rows;
    const int own0 = min(NP, blk.share * OWN);
    const int nown = min(NP, own0 + OWN) - own0;
    const int ng = (D + 3) >> 2;
    const uint32_t inv_ng = blk.inv_ng;
    auto div_ng = [&](int it) -> int { return ng > 1 ? (int)__umulhi((uint32_t)it, inv_ng) : it; };
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lr = lane & 15, lk = lane >> 4;
    const uint32_t crt = blk.crt;
    const uint32_t gen = *a.gen_base + a.gen_offset;
    asm volatile("" ::"s"(a.Xin), "s"(a.fXin), "s"(a.donors), "s"(a.rows), "s"(a.donor_gen), "s"(gen),
                 "s"(blk.B32s), "s"(blk.shift), "s"(blk.colw), "s"(I.elem_off), "s"(I.row_off), "s"(NP),
                 "s"(D), "s"(pl.S));

    float *fXs = reinterpret_cast<float *>(lds);   // [OWN] the share's rows only
    float *fUs = fXs + OWN;
    int *acc_f = reinterpret_cast<int *>(fUs + OWN);
    float *sh = reinterpret_cast<float *>(acc_f + OWN);
    int4 *doff = reinterpret_cast<int4 *>(sh + 128);
    int *jr = reinterpret_cast<int *>(doff + OWN);
    float *part = reinterpret_cast<float *>(jr + OWN);  // [8][OWN]
    floatx4 *xch = reinterpret_cast<floatx4 *>(part + 8 * OWN);  // [4][2][64]
    float *rows = part + 8 * OWN + kSplitXchBytes / 4;   // [NPp][S]
    float *trial = rows + NPp * S;                        // [OWN][ST]
    float *arow = trial + OWN * ST;                       // FP32: [OWN][ST]
    uint16_t *a16 = reinterpret_cast<uint16_t *>(arow);   // 16-bit: [P][OWN][LDA]
    const int kp32 = blk.kp32, LDA = kp32 + 8;

    // ---- prologue: every global load first
    const float *gXin = a.Xin + I.elem_off;
    constexpr int kBatch = 8;
    const int total = NP * ng;
    // With an LDS row stride of D the island is one linear block in HBM and
    // in LDS: copied straight into LDS, 1 KB per wave instruction (lane l's
    // 16 bytes land at the base + 16 l), with no registers or LDS stores on
    // the way.  Otherwise rows go through registers to their padded slots.
    const bool linear = (D & 3) == 0 && S == D;
    floatx4 pre[kBatch];
    if (linear) {
        const int nchunk = (total + 63) >> 6;
        for (int c = wave; c < nchunk; c += kSplitThreads / 64) {
            const int it = c * 64 + lane;
            if (it < total)
                __builtin_amdgcn_global_load_lds(gXin + 4 * it, (split_lds_ptr)(rows + c * 256), 16, 0, 0);
        }
    } else {
        const __amdgpu_buffer_rsrc_t xr = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(gXin), (short)0, total * 16, kBufferWord3);
#pragma unroll
        for (int b = 0; b < kBatch; ++b)
            if (b * kSplitThreads < total)
                pre[b] = __builtin_bit_cast(
                    floatx4, __builtin_amdgcn_raw_buffer_load_b128(xr, (tid + b * kSplitThreads) * 16, 0, 0));
    }
    const float shift_k = tid < 128 && tid < D ? gld(blk.shift + tid) : 0.f;
    float fx_r;
    int4 d_r;
    if (tid < nown) {
        fx_r = a.fXin[I.row_off + own0 + tid];
        d_r = a.donors[(size_t)a.donor_gen * a.rows + I.row_off + own0 + tid];
    }
    // rotation fragments and column weights of the evaluating waves
    const int t = wave & 7;
    const bool evaluates = wave < 8;
    const bool vt = t * 16 < D;
    const int KS = P == 0 ? ng : kp32 >> 5;
    constexpr int NB32 = P == 0 ? 32 : 1;
    constexpr int NB16 = P == 0 ? 1 : P;
    float b[NB32];
    vec8 bb[4][NB16];
    // the pair of column tiles (tl, tl + 4) whose terms this wave forms
    const int tl = t & 3;
    const bool cl = tl * 16 + lr < D, ch = (tl + 4) * 16 + lr < D;
    float wl = 1.0f, wh = 1.0f;
    if (evaluates && blk.objective == SPEO_OBJ_ELLIPTIC) {
        if (cl) wl = gld(blk.colw + tl * 16 + lr);
        if (ch) wh = gld(blk.colw + (tl + 4) * 16 + lr);
    }
    // The crossover draws of the lane's first trial item depend only on the
    // launch: they are computed while the loads are in flight.
    const int it0 = tid, rl0 = div_ng(it0), q40 = it0 - rl0 * ng;
    const U4 w0 = philox_block(a.seed, (uint32_t)q40, (uint32_t)(I.row_off + own0 + rl0), gen, kPurposeCross);
    asm volatile("" ::"v"(w0.x), "v"(w0.y), "v"(w0.z), "v"(w0.w));  // (computed here, not at its use)

    // ---- stores
    if (tid < 128) sh[tid] = shift_k;
    if (tid < nown) {
        fXs[tid] = fx_r;
        const int r = own0 + tid;
        doff[tid] = int4{d_r.x * S, d_r.y * S, d_r.z * S, r * S};
        jr[tid] = d_r.w;
    }
    auto put_row_group = [&](int it, float4 x) {
        const int r = div_ng(it), q4 = it - r * ng, j = 4 * q4;
        if (j + 1 >= D) x.y = 0.f;
        if (j + 2 >= D) x.z = 0.f;
        if (j + 3 >= D) x.w = 0.f;
        st4(rows + r * S + 4 * q4, x);
    };
    if (!linear) {
#pragma unroll
        for (int bb = 0; bb < kBatch; ++bb) {
            const int it = tid + bb * kSplitThreads;
            if (it < total) put_row_group(it, float4{pre[bb][0], pre[bb][1], pre[bb][2], pre[bb][3]});
        }
        for (int it = tid + kBatch * kSplitThreads; it < total; it += kSplitThreads)
            put_row_group(it, ld4(gXin + 4 * it));
    }
    if constexpr (P == 0) {
        // the shifted copies' columns [4 ng, ST) are read by the evaluation's
        // blocks of four K steps and stay zero
        const int pad = ST - 4 * ng;
        for (int e = tid; e < OWN * pad; e += kSplitThreads) {
            const int r = e / pad;
            arow[r * ST + 4 * ng + (e - r * pad)] = 0.f;
        }
    } else {
        // the parts' columns [4 ng, kp32) are zero
        const int pad = kp32 - 4 * ng;
        for (int e = tid; e < P * OWN * pad; e += kSplitThreads) {
            const int r = e / pad;  // part * OWN + row
            a16[r * LDA + 4 * ng + (e - r * pad)] = 0;
        }
    }
    // the copies into LDS count as outstanding vector-memory loads: every one
    // of this wave's has landed before the barrier
    if (linear) __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
    __syncthreads();

    // the rotation fragments, needed from the evaluation on: issued now, they
    // load while the trials are made rather than ahead of the population
    if (evaluates) {
        if constexpr (P == 0) {
            load_b_frag(blk.B32s, t, vt, KS, lane, b);
        } else {
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                for (int p = 0; p < P; ++p)
                    if (vt && ks < KS)
                        bb[ks][p] = gld(reinterpret_cast<const vec8 *>(
                            blk.M16s[p] + (size_t)(t * 16 + lr) * kp32 + ks * 32 + lk * 8));
        }
    }

    // ---- trials: one item per lane
    const float rate = blk.rate;
    for (int it = tid; it < nown * ng; it += kSplitThreads) {
        const bool first = it == it0;
        const int rl = first ? rl0 : div_ng(it), q4 = first ? q40 : it - rl * ng;
        const int4 o = doff[rl];
        const int jrand = jr[rl];
        const float4 e0 = ld4(rows + o.x + 4 * q4), e1 = ld4(rows + o.y + 4 * q4);
        const float4 e2 = ld4(rows + o.z + 4 * q4), xo = ld4(rows + o.w + 4 * q4);
        const float4 sv = ld4(sh + 4 * q4);
        const U4 w = first ? w0
                           : philox_block(a.seed, (uint32_t)q4, (uint32_t)(I.row_off + own0 + rl), gen, kPurposeCross);
        float v[4];
        bool out = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int j = 4 * q4 + q;
            const bool take = (j == jrand) | cross_take(word(w, q), crt);
            const float mt = fmaf(I.F, comp4(e1, q) - comp4(e2, q), comp4(e0, q));
            v[q] = j < D ? (take ? mt : comp4(xo, q)) : 0.f;
            out |= (j < D) & ((v[q] < a.lo) | (v[q] > a.hi));
        }
        if (out) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (4 * q4 + q < D) v[q] = check_bound(v[q], a.lo, a.hi);
        }
        st4(trial + rl * ST + 4 * q4, float4{v[0], v[1], v[2], v[3]});
        if constexpr (P == 0) {
            st4(arow + rl * ST + 4 * q4,
                float4{(v[0] - sv.x) * rate, (v[1] - sv.y) * rate, (v[2] - sv.z) * rate, (v[3] - sv.w) * rate});
        } else {
            const float scale = rate * (H ? kF16ScaleA : 1.0f);
            typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
            u16x4 pb[P];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float rem = 4 * q4 + q < D ? (v[q] - comp4(sv, q)) * scale : 0.f;
#pragma unroll
                for (int pp = 0; pp < P; ++pp) {
                    const uint16_t bits = to_bits(rem, H);
                    pb[pp][q] = bits;
                    rem -= from_bits(bits, H);
                }
            }
#pragma unroll
            for (int pp = 0; pp < P; ++pp)
                *reinterpret_cast<u16x4 *>(a16 + (pp * OWN + rl) * LDA + 4 * q4) = pb[pp];
        }
    }
    __syncthreads();

    // ---- evaluation: waves 0..7, column tile t, the 16-row tiles two at a time
    const int ntile = (nown + 15) >> 4;
    for (int tile = 0; tile < ntile; tile += 2) {  // every wave: the barriers are block-wide
        const bool two = tile + 1 < ntile;
        floatx4 acc0 = floatx4{0.f, 0.f, 0.f, 0.f}, acc1 = floatx4{0.f, 0.f, 0.f, 0.f};
        if (evaluates) {
            const int r0 = min(tile * 16 + lr, nown - 1), r1 = min(tile * 16 + 16 + lr, nown - 1);
            if (P == 0 && vt) {
                const float *ap0 = arow + r0 * ST + lk;
                const float *ap1 = arow + r1 * ST + lk;
#pragma unroll
                for (int kb = 0; kb < 32; kb += 4) {
                    if (kb >= KS) break;
                    float a0[4], a1[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        a0[u] = ap0[(kb + u) * 4];
                        a1[u] = ap1[(kb + u) * 4];
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        acc0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0[u], b[kb + u], acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1[u], b[kb + u], acc1, 0, 0, 0);
                    }
                }
            }
            if constexpr (P > 0) {
                if (vt) {
                    // the parts' cross products, in gemm_small's order
#pragma unroll
                    for (int ks = 0; ks < 4; ++ks) {
                        if (ks >= KS) break;
                        vec8 a0[P], a1[P];
#pragma unroll
                        for (int pp = 0; pp < P; ++pp) {
                            a0[pp] = *reinterpret_cast<const vec8 *>(a16 + (pp * OWN + r0) * LDA + ks * 32 + lk * 8);
                            a1[pp] = *reinterpret_cast<const vec8 *>(a16 + (pp * OWN + r1) * LDA + ks * 32 + lk * 8);
                        }
#pragma unroll
                        for (int sum = P - 1; sum >= 0; --sum) {
#pragma unroll
                            for (int pa = 0; pa <= sum; ++pa) {
                                const int pbi = sum - pa;
                                if (pa < P && pbi < P) {
                                    if constexpr (H) {
                                        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a0[pa], bb[ks][pbi], acc0, 0, 0, 0);
                                        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(a1[pa], bb[ks][pbi], acc1, 0, 0, 0);
                                    } else {
                                        acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0[pa], bb[ks][pbi], acc0, 0, 0, 0);
                                        acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[pa], bb[ks][pbi], acc1, 0, 0, 0);
                                    }
                                }
                            }
                        }
                    }
                }
            }
            // waves 4..7 hand over the first row tile's accumulator, waves
            // 0..3 the second's: each wave then forms one row tile's terms
            if (wave >= 4) xch[((wave - 4) * 2 + 0) * 64 + lane] = acc0;
            else xch[(wave * 2 + 1) * 64 + lane] = acc1;
        }
        __syncthreads();
        if (evaluates && (wave < 4 || two)) {
            const int i = wave >> 2;  // row tile tile + i
            const floatx4 lo = wave < 4 ? acc0 : xch[((wave - 4) * 2 + 1) * 64 + lane];
            const floatx4 hi = wave < 4 ? xch[(wave * 2 + 0) * 64 + lane] : acc1;
            const float unscale = H ? kF16Unscale : 1.0f;
            float s1[4], s2[4];
            switch (blk.objective) {
            case SPEO_OBJ_ELLIPTIC:
                tile_terms<SPEO_OBJ_ELLIPTIC>(lo, hi, cl, ch, wl, wh, unscale, s1, s2);
                break;
            case SPEO_OBJ_RASTRIGIN:
                tile_terms<SPEO_OBJ_RASTRIGIN>(lo, hi, cl, ch, wl, wh, unscale, s1, s2);
                break;
            default:
                tile_terms<SPEO_OBJ_ACKLEY>(lo, hi, cl, ch, wl, wh, unscale, s1, s2);
                break;
            }
            if (lr == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int rl = (tile + i) * 16 + lk * 4 + q;
                    part[tl * OWN + rl] = s1[q];
                    part[(4 + tl) * OWN + rl] = s2[q];
                }
            }
        }
        __syncthreads();
    }

    if (blk.heavy.j >= 0) {
        // the heavy column's FP64 term, one wave per row, parked in the row's
        // fU slot until the combine replaces it
        for (int rl = wave; rl < nown; rl += kSplitThreads / 64) {
            const double h = heavy_term(blk.heavy, trial + rl * ST, D, lane);
            if (lane == 0) fUs[rl] = (float)h;
        }
        __syncthreads();
    }
    // ---- combine (column groups 0..3 in order) and selection, strict "<"
    if (tid < nown) {
        float s1 = 0.f, s2 = 0.f;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            s1 += part[w * OWN + tid];
            s2 += part[(4 + w) * OWN + tid];
        }
        if (blk.heavy.j >= 0) s1 = with_heavy(s1, fUs[tid]);
        const float fu = fmaf(lambda, finish_objective(blk.objective, s1, s2, D), bias);
        const float fx = fXs[tid];
        const bool acc = fu < fx;
        fUs[tid] = fu;
        fXs[tid] = acc ? fu : fx;
        acc_f[tid] = acc ? 1 : 0;
    }
    __syncthreads();

    // ---- write-back: the selected rows, and at the end of a chunk the trials
    float *gX = a.Xout + I.elem_off, *gU = a.U + I.elem_off;
    for (int it = tid; it < nown * ng; it += kSplitThreads) {
        const int rl = div_ng(it), q4 = it - rl * ng, r = own0 + rl;
        const float4 tri = ld4(trial + rl * ST + 4 * q4);
        const float4 cur = acc_f[rl] ? tri : ld4(rows + r * S + 4 * q4);
        st4(gX + r * I.ld + 4 * q4, cur);
        if (a.write_trials) st4(gU + r * I.ld + 4 * q4, tri);
    }
    if (tid < nown) {
        a.fXout[I.row_off + own0 + tid] = fXs[tid];
        if (a.write_trials) a.fU[I.row_off + own0 + tid] = fUs[tid];
    }
}

template <int PREC>
static int launch_epoch_t(const EpochArgs &a, int nblocks, int lds, hipStream_t s) {
    static std::once_flag once;
    static hipError_t attr = hipSuccess;
    std::call_once(once, [] {
        attr = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_island_epoch<PREC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kEpochLdsMax);
    });
    if (attr != hipSuccess) return check_hip(attr, "hipFuncSetAttribute(k_island_epoch)");
    hipLaunchKernelGGL((k_island_epoch<PREC>), dim3(nblocks), dim3(kEpochThreads), lds, s, a);
    return check_launch("k_island_epoch");
}

template <int PREC>
static int launch_split_t(const EpochArgs &a, int nblocks, int lds, hipStream_t s) {
    static std::once_flag once;
    static hipError_t attr = hipSuccess;
    std::call_once(once, [] {
        attr = hipFuncSetAttribute(reinterpret_cast<const void *>(&k_island_split<PREC>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, kEpochLdsMax);
    });
    if (attr != hipSuccess) return check_hip(attr, "hipFuncSetAttribute(k_island_split)");
    hipLaunchKernelGGL((k_island_split<PREC>), dim3(nblocks), dim3(kSplitThreads), lds, s, a);
    return check_launch("k_island_split");
}

template <int PREC>
static int launch_epoch_p(const EpochArgs &a, int nblocks, int lds, bool split, hipStream_t s) {
    return split ? launch_split_t<PREC>(a, nblocks, lds, s) : launch_epoch_t<PREC>(a, nblocks, lds, s);
}

int launch_epoch(int precision, const EpochArgs &a, int nblocks, int lds_bytes, bool split, hipStream_t s) {
    if (a.L < 1) return set_error(SPEO_ERR_ARG, "epoch launch needs at least one generation");
    if (split && a.L != 1) return set_error(SPEO_ERR_ARG, "split epoch launches run one generation");
    if (lds_bytes > kEpochLdsMax) return set_error(SPEO_ERR_ARG, "epoch launch: LDS plan too large");
    switch (precision) {
    case SPEO_PREC_FP32: return launch_epoch_p<SPEO_PREC_FP32>(a, nblocks, lds_bytes, split, s);
    case SPEO_PREC_BF16: return launch_epoch_p<SPEO_PREC_BF16>(a, nblocks, lds_bytes, split, s);
    case SPEO_PREC_FP16: return launch_epoch_p<SPEO_PREC_FP16>(a, nblocks, lds_bytes, split, s);
    case SPEO_PREC_BF16X3: return launch_epoch_p<SPEO_PREC_BF16X3>(a, nblocks, lds_bytes, split, s);
    case SPEO_PREC_FP16X3: return launch_epoch_p<SPEO_PREC_FP16X3>(a, nblocks, lds_bytes, split, s);
    case SPEO_PREC_BF16X6: return launch_epoch_p<SPEO_PREC_BF16X6>(a, nblocks, lds_bytes, split, s);
    default: return set_error(SPEO_ERR_ARG, "unknown precision");
    }
}

}  // namespace speo
