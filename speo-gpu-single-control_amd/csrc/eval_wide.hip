// eval_wide.hip -- the 16-bit evaluation of wide problems (dim > 128).
//
// The arithmetic of k_eval's 64-row form (gemm_chunk<64, PREC>): for every
// element of Z the same v_mfma_f32_16x16x32 chain over K in steps of 32, the
// operand parts formed and multiplied in the same order; per row and
// 128-column chunk of Z the terms summed lane by lane over the chunk's eight
// column tiles in order, then across the 16 lanes by sum16.  The partial sums
// it writes are therefore bit-identical to k_eval<64, PREC>'s.  What changes
// is the tiling, for throughput on large row counts:
//  * a block of 512 threads owns RB rows x 256 columns (two chunks): wave w
//    owns rows [WR (w % 4), +WR) of chunk w / 4, WR / 16 x 8 accumulator
//    tiles, so each operand fragment read from LDS feeds 8 (A) or WR / 16 (B)
//    MFMA steps;
//  * operands copied from global memory straight into LDS (global_load_lds,
//    16 bytes per lane, no registers on the way) in MFMA fragment order: one
//    16 x 32 fragment is 64 lanes x 16 bytes with lane l at byte 16 l, so
//    every fragment read is one conflict-free ds_read_b128 per lane.  The
//    rows of X arrive as FP32 and each wave forms its A operand parts
//    ((x - o) rate, split into BF16 / FP16 parts) as it reads them;
//  * K in stages of 32 over three LDS buffers (two when the six-product
//    split needs the room): stage s + 2 is requested while stage s is
//    multiplied, one barrier per stage;
//  * the block order is XCD-aware: the column pairs of one row block run on
//    the same XCD (workgroup b on XCD b mod 8), so the block's rows of X are
//    fetched from HBM once into that XCD's L2 and shared by its column pairs.
#include <mutex>
#include <type_traits>

#include "eval.h"

namespace speo {

namespace {

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef float floatx8 __attribute__((ext_vector_type(8)));

template <int PREC> struct Wide {
    static constexpr int P = Prec<PREC>::parts;
    static constexpr bool H = Prec<PREC>::half;
    static constexpr int NW = 8;                            // waves per block
    static constexpr int NT = 64 * NW;                      // threads per block
    static constexpr int RG = NW / 2;                       // row groups (two column chunks)
    static constexpr int WR = P == 1 ? 64 : 32;             // rows per wave
    static constexpr int RB = RG * WR;                      // rows per block
    static constexpr int RTW = WR / 16;                     // row tiles per wave
    static constexpr int RTB = RB / 16;                     // row tiles per block
    static constexpr int A_BYTES = RTB * 2 * 1024;          // FP32 rows of a stage: two 1 KB halves per row tile
    static constexpr int B_BYTES = P * 16 * 1024;           // operand fragments of 16 column tiles per part
    static constexpr int STAGE = A_BYTES + B_BYTES;
    // three stage buffers where they fit (two for the three-part split).
    // Four-wave blocks, two per CU with two buffers each, measured 14 %
    // slower for BF16: their copies are requested one stage ahead instead of
    // two and load more bytes per MFMA step.
    static constexpr int NBUF = 3 * STAGE <= 144 * 1024 ? 3 : 2;
    static constexpr int LDS = NBUF * STAGE;
    static constexpr int NA = RTB * 2 / NW;                 // A copies per wave and stage
    static constexpr int NB = P * 16 / NW;                  // B copies per wave and stage
    static constexpr int NLD = NA + NB;
    static_assert(RTB * 2 % NW == 0 && 16 % NW == 0, "copies must tile the waves");
    static_assert(2 * RB * 2 * 4 <= LDS, "row sums must fit in the staging area");
};

// s_waitcnt vmcnt(N), the other counters left alone (gfx9 encoding: vmcnt
// in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8).
template <int N> __device__ __forceinline__ void wait_vm() {
    static_assert(N >= 0 && N < 64, "vmcnt range");
    __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | (((N >> 4) & 3) << 14));
}

// A workgroup barrier that neither waits for outstanding memory operations
// (the LDS copies of later stages stay in flight) nor lets the compiler move
// memory accesses across it.
__device__ __forceinline__ void stage_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

typedef __attribute__((address_space(3))) void *lds_ptr;

constexpr int kWideLdsMax = 160 * 1024;  // LDS of one CU

struct WideArgs {
    CompDev comp;
    const float *X;
    int ld, n, dim, nchunks;
    float *part;  // [nchunks][2][n]
};

// Per lane, the terms of the wave's eight column tiles in order into the row
// sums of its RTW row tiles (gemm_chunk's epilogue with the objective fixed;
// a function rather than a lambda, so that the accumulators stay in
// registers).
template <int OBJ, int RTW>
__device__ __forceinline__ void wide_terms(const floatx4 (&acc)[RTW][8], const float *colw, int n0,
                                           int dim, int lr, float unscale, float (&s1)[RTW][4],
                                           float (&s2)[RTW][4]) {
#pragma unroll
    for (int t = 0; t < 8; ++t) {
        const int col = n0 + t * 16 + lr;
        if (col < dim) {
            const float w = OBJ == SPEO_OBJ_ELLIPTIC ? colw[col] : 1.0f;
#pragma unroll
            for (int i = 0; i < RTW; ++i)
#pragma unroll
                for (int q = 0; q < 4; ++q) add_terms(OBJ, acc[i][t][q] * unscale, w, s1[i][q], s2[i][q]);
        }
    }
}

template <int PREC>
__global__ __launch_bounds__(Wide<PREC>::NT) __attribute__((amdgpu_waves_per_eu(1))) void k_eval_wide(
    WideArgs p) {
    using W = Wide<PREC>;
    constexpr int P = W::P;
    constexpr bool H = W::H;
    typedef typename std::conditional<H, halfx8, bf16x8>::type vec8;
    extern __shared__ __align__(16) char lds[];

    const int nrb = (p.n + W::RB - 1) / W::RB, ncp = (p.nchunks + 1) / 2;
    const int b = blockIdx.x, xcd = b & 7, j = b >> 3;
    const int cp = j % ncp, rb = (j / ncp) * 8 + xcd;
    if (rb >= nrb) return;  // whole block, before any barrier
    const int row0 = rb * W::RB;
    const CompDev &c = p.comp;
    const int dim = p.dim;

    // wave and everything derived from it are wave-uniform (scalar branches)
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int rg = wave % W::RG, cg = wave / W::RG;
    const int n0 = (2 * cp + cg) * kNC;  // this wave's chunk of Z columns
    const float sa = H ? kF16ScaleA : 1.0f;

    // The shift, zero-padded to a multiple of 32, stays in LDS after the
    // stage buffers.
    float *shs = reinterpret_cast<float *>(lds + W::LDS);
    for (int k = tid; k < round_up(dim, 32); k += W::NT) shs[k] = k < dim ? c.shift[k] : 0.f;
    __syncthreads();

    // The copies of stage s into buffer s % NBUF.  Wave w issues A copies
    // f = w + NW j (row tile f / 2, coordinates 8 lk + 4 (f % 2) .. + 3 of
    // its lane's row) and B copies f = w + NW j (part f / 16, column tile
    // f % 16: coordinates 8 lk .. + 7 of its lane's column of M).  Copies
    // past the last row, the dimension or the last column read a valid
    // address instead (rows are clamped, the rest read the operand's first
    // bytes); the A values are masked when formed, and B's stray columns
    // and coordinates meet zero A values or columns that are not summed.
    // Rows have a stride of a multiple of 4 floats and M16 rows one of 8, so
    // a group that starts below dim lies inside its row.
    auto issue = [&](int s) __attribute__((always_inline)) {
        char *st = lds + (s % W::NBUF) * W::STAGE;
        const int k0 = s * 32;
#pragma unroll
        for (int j = 0; j < W::NA; ++j) {
            const int f = wave + W::NW * j;
            const int r = min(row0 + (f >> 1) * 16 + lr, p.n - 1), k = k0 + 8 * lk + 4 * (f & 1);
            const float *src = p.X + (size_t)r * p.ld + (k < dim ? k : 0);
            __builtin_amdgcn_global_load_lds(src, (lds_ptr)(st + f * 1024), 16, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < W::NB; ++j) {
            const int f = wave + W::NW * j, col = 2 * cp * kNC + (f & 15) * 16 + lr, k = k0 + 8 * lk;
            const bool ok = col < dim && k < dim;
            const uint16_t *src = c.M16[(W::NW * j) >> 4] + (ok ? (size_t)col * c.ldm16 + k : 0);
            __builtin_amdgcn_global_load_lds(src, (lds_ptr)(st + W::A_BYTES + f * 1024), 16, 0, 0);
        }
    };

    floatx4 acc[W::RTW][8];
#pragma unroll
    for (int i = 0; i < W::RTW; ++i)
#pragma unroll
        for (int t = 0; t < 8; ++t) acc[i][t] = floatx4{0.f, 0.f, 0.f, 0.f};

    // One stage: this wave's A operand parts formed from the FP32 rows in
    // LDS ((x - o) rate sa as gemm_chunk forms it, split into parts; MASK
    // zeroes entries past the last row or the dimension, UNIT drops the
    // multiplication by a rate of 1, which is exact), its B fragments read,
    // and every tile multiplied (a column tile past dim holds stray values
    // and is not summed).
    const float rate = c.rate;
    auto compute = [&](int s, auto mask, auto unit) __attribute__((always_inline)) {
        constexpr bool MASK = decltype(mask)::value, UNIT = decltype(unit)::value;
        const char *st = lds + (s % W::NBUF) * W::STAGE;
        const int k = s * 32 + 8 * lk;
        const floatx4 o0 = *reinterpret_cast<const floatx4 *>(shs + k);
        const floatx4 o1 = *reinterpret_cast<const floatx4 *>(shs + k + 4);
        const floatx8 ov = __builtin_shufflevector(o0, o1, 0, 1, 2, 3, 4, 5, 6, 7);
        vec8 a[W::RTW][P], bb[8][P];
#pragma unroll
        for (int i = 0; i < W::RTW; ++i) {
            const int rt = rg * W::RTW + i;
            const floatx4 x0 = *reinterpret_cast<const floatx4 *>(st + (rt * 2 + 0) * 1024 + lane * 16);
            const floatx4 x1 = *reinterpret_cast<const floatx4 *>(st + (rt * 2 + 1) * 1024 + lane * 16);
            const bool rv = row0 + rt * 16 + lr < p.n;
            // eight coordinates as one vector: the conversions pair up into
            // packed converts with no packing instructions after them
            const floatx8 d = __builtin_shufflevector(x0, x1, 0, 1, 2, 3, 4, 5, 6, 7) - ov;
            floatx8 rem = UNIT ? d * sa : d * rate * sa;
            if constexpr (MASK) {
#pragma unroll
                for (int q = 0; q < 8; ++q)
                    if (!(rv & (k + q < dim))) rem[q] = 0.f;
            }
#pragma unroll
            for (int pp = 0; pp < P; ++pp) {
                a[i][pp] = __builtin_convertvector(rem, vec8);
                if (pp + 1 < P) rem -= __builtin_convertvector(a[i][pp], floatx8);
            }
        }
        const char *bs = st + W::A_BYTES;
#pragma unroll
        for (int t = 0; t < 8; ++t)
#pragma unroll
            for (int pp = 0; pp < P; ++pp)
                bb[t][pp] = *reinterpret_cast<const vec8 *>(bs + ((pp * 16 + cg * 8 + t) * 64 + lane) * 16);
#pragma unroll
        for (int t = 0; t < 8; ++t) {
#pragma unroll
            for (int i = 0; i < W::RTW; ++i) {
                // small products first, then hi*hi (gemm_chunk's order)
#pragma unroll
                for (int sum = P - 1; sum >= 0; --sum) {
#pragma unroll
                    for (int pa = 0; pa <= sum; ++pa) {
                        const int pb = sum - pa;
                        if (pa < P && pb < P) {
                            if constexpr (H)
                                acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[i][pa], bb[t][pb],
                                                                                   acc[i][t], 0, 0, 0);
                            else
                                acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i][pa], bb[t][pb],
                                                                                    acc[i][t], 0, 0, 0);
                        }
                    }
                }
            }
        }
    };

    // The copies of stages 0 .. NBUF - 2 are issued up front.  At stage s a
    // wave waits for its own copies of stage s (those of the stages after it
    // that were issued, NBUF - 2 of them, may stay in flight); after the
    // barrier every wave's copies of stage s have landed and every wave has
    // finished stage s - 1, whose buffer takes stage s + NBUF - 1.
    const int nst = (dim + 31) >> 5;
#pragma unroll
    for (int s = 0; s < W::NBUF - 1; ++s)
        if (s < nst) issue(s);
    auto run = [&](auto mask, auto unit) __attribute__((always_inline)) {
        for (int s = 0; s < nst - 1; ++s) {
            if (W::NBUF == 3)
                wait_vm<W::NLD>();
            else
                wait_vm<0>();
            stage_barrier();
            if (s + W::NBUF - 1 < nst) issue(s + W::NBUF - 1);
            compute(s, mask, unit);
        }
        wait_vm<0>();
        stage_barrier();
        compute(nst - 1, std::true_type{}, std::false_type{});  // coordinates past dim
    };
    if (row0 + W::RB > p.n)
        run(std::true_type{}, std::false_type{});  // rows past the last
    else if (c.rate == 1.0f)
        run(std::false_type{}, std::true_type{});
    else
        run(std::false_type{}, std::false_type{});
    __syncthreads();  // the stage buffers become the row-sum area

    // Terms of this wave's columns, in gemm_chunk's order: per lane, the
    // chunk's column tiles in order, then the 16-lane butterfly.
    const float unscale = H ? kF16Unscale : 1.0f;
    float s1[W::RTW][4], s2[W::RTW][4];
#pragma unroll
    for (int i = 0; i < W::RTW; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) s1[i][q] = s2[i][q] = 0.f;
    // the objective dispatch of with_objective
    switch (c.objective) {
    case SPEO_OBJ_ELLIPTIC: wide_terms<SPEO_OBJ_ELLIPTIC>(acc, c.colw, n0, dim, lr, unscale, s1, s2); break;
    case SPEO_OBJ_RASTRIGIN: wide_terms<SPEO_OBJ_RASTRIGIN>(acc, c.colw, n0, dim, lr, unscale, s1, s2); break;
    case SPEO_OBJ_ACKLEY: wide_terms<SPEO_OBJ_ACKLEY>(acc, c.colw, n0, dim, lr, unscale, s1, s2); break;
    case SPEO_OBJ_SCHWEFEL: wide_terms<SPEO_OBJ_SCHWEFEL>(acc, c.colw, n0, dim, lr, unscale, s1, s2); break;
    default: wide_terms<SPEO_OBJ_HAPPYCAT>(acc, c.colw, n0, dim, lr, unscale, s1, s2); break;
    }
    float *red = reinterpret_cast<float *>(lds);  // [2 chunks][RB rows][2]
#pragma unroll
    for (int i = 0; i < W::RTW; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float a1 = sum16(s1[i][q]), a2 = sum16(s2[i][q]);
            if (lr == 0) {
                const int r = rg * W::WR + i * 16 + lk * 4 + q;
                // 0 + s: k_eval's one-term reduction across waves
                red[(cg * W::RB + r) * 2 + 0] = 0.f + a1;
                red[(cg * W::RB + r) * 2 + 1] = 0.f + a2;
            }
        }
    __syncthreads();
    if (c.heavy.j >= 0 && cp == 0) {
        // the heavy column's FP64 term into chunk 0's sums, one wave per row
        for (int r = wave; r < W::RB && row0 + r < p.n; r += W::NW) {
            const double hv = heavy_term(c.heavy, p.X + (size_t)(row0 + r) * p.ld, dim, lane);
            if (lane == 0) red[r * 2] = with_heavy(red[r * 2], (float)hv);
        }
        __syncthreads();
    }
    for (int idx = tid; idx < 2 * W::RB; idx += W::NT) {
        const int g = idx / W::RB, r = idx % W::RB, chunk = 2 * cp + g, row = row0 + r;
        if (chunk < p.nchunks && row < p.n) {
            p.part[((size_t)chunk * 2 + 0) * p.n + row] = red[idx * 2 + 0];
            p.part[((size_t)chunk * 2 + 1) * p.n + row] = red[idx * 2 + 1];
        }
    }
}

template <int PREC> int launch_wide(const WideArgs &a, hipStream_t s) {
    using W = Wide<PREC>;
    static std::once_flag once;
    std::call_once(once, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eval_wide<PREC>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, kWideLdsMax);
    });
    const int nrb = (a.n + W::RB - 1) / W::RB, ncp = (a.nchunks + 1) / 2;
    const int blocks = round_up(nrb, 8) * ncp;
    hipLaunchKernelGGL(k_eval_wide<PREC>, dim3(blocks), dim3(W::NT), W::LDS + 4 * round_up(a.dim, 32), s,
                       a);
    return check_launch("k_eval_wide");
}

bool aligned16(const void *q) { return (reinterpret_cast<uintptr_t>(q) & 15) == 0; }

}  // namespace

bool eval_wide_applies(const CompDev &c, const float *X, int ld, int dim, int precision) {
    int parts = 0;
    switch (precision) {
    case SPEO_PREC_BF16: case SPEO_PREC_FP16: parts = 1; break;
    case SPEO_PREC_BF16X3: case SPEO_PREC_FP16X3: parts = 2; break;
    case SPEO_PREC_BF16X6: parts = 3; break;
    default: return false;
    }
    const int stages = parts == 1 ? Wide<SPEO_PREC_BF16>::LDS
                       : parts == 2 ? Wide<SPEO_PREC_BF16X3>::LDS
                                    : Wide<SPEO_PREC_BF16X6>::LDS;
    if (dim <= kNC || (ld & 3) != 0 || !aligned16(X) || (c.ldm16 & 7) != 0 ||
        stages + 4 * (size_t)round_up(dim, 32) > (size_t)kWideLdsMax)
        return false;
    for (int q = 0; q < parts; ++q)
        if (c.M16[q] == nullptr || !aligned16(c.M16[q])) return false;
    return true;
}

int eval_wide(const CompDev &c, const float *X, int ld, int n, int dim, float *part, int precision,
              hipStream_t s) {
    const WideArgs a{c, X, ld, n, dim, (dim + kNC - 1) / kNC, part};
    switch (precision) {
    case SPEO_PREC_BF16: return launch_wide<SPEO_PREC_BF16>(a, s);
    case SPEO_PREC_FP16: return launch_wide<SPEO_PREC_FP16>(a, s);
    case SPEO_PREC_BF16X3: return launch_wide<SPEO_PREC_BF16X3>(a, s);
    case SPEO_PREC_FP16X3: return launch_wide<SPEO_PREC_FP16X3>(a, s);
    case SPEO_PREC_BF16X6: return launch_wide<SPEO_PREC_BF16X6>(a, s);
    default: return set_error(SPEO_ERR_ARG, "eval_wide: not a 16-bit precision");
    }
}

}  // namespace speo
