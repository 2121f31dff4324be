// eval_core.h -- device building blocks of the shifted-rotated objective.
//
// z = rate * M (x - o) for a tile of RT rows, as the GEMM Z = A B with
// A = rate (X - 1 o^T) (staged in LDS with the shift subtracted) and
// B = M^T, on MFMA:
//   FP32:        v_mfma_f32_16x16x4_f32, A and B as FP32
//   16-bit:      v_mfma_f32_16x16x32_{bf16,f16}, operands split into 1-3 parts
//                (a = a0 + a1 [+ a2]) and the cross products summed in FP32.
// Then per row sum_j g(z_j) (elliptic / Rastrigin / Ackley) with the per
// coordinate terms reduced across the 16 lanes that hold one row's columns.
//
// MFMA 16x16 operand layout (lane l of 64): A row l%16, B column l%16,
// k = (l/16) (FP32, one value) or 8*(l/16)+0..7 (16-bit, eight values);
// the 4 accumulator values of lane l are rows 4*(l/16)+i, column l%16.
//
// Reference counterparts: global_ShiftAndRotation (kernel_CEC2014_CUDA.cu:
// 147-181, shift fused into the A-tile load at 161), global_f1 / f5 / f8
// (254-286, 389-442, 557-587) and device_ParallelSum (6-26).
#pragma once

#include <type_traits>

#include "common.h"

namespace speo {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// A load through the global address space.  Pointers read from device tables
// are generic to the compiler, which then emits flat loads: those also count
// against the LDS/scalar wait counter, so the waits that follow serialise
// them with LDS traffic.  T must be a scalar or an ext_vector_type (a class
// type's copy constructor would take the value back through a generic
// reference).
template <typename T> __device__ __forceinline__ T gld(const T *p) {
    return *(const __attribute__((address_space(1))) T *)p;
}
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));

constexpr int kNC = 128;  // columns of Z per N chunk (8 MFMA tiles)
constexpr int kKC = 64;   // K per LDS chunk

// FP16 operands are scaled into FP16's normal range: A by 2^7 (|x-o| <= 2^8
// maps below 2^15), M by 2^14 (|M_ij| <= 1); Z is rescaled by 2^-21.
constexpr float kF16ScaleA = 128.0f;
constexpr float kF16ScaleB = 16384.0f;
constexpr float kF16Unscale = 1.0f / (128.0f * 16384.0f);

template <int PREC> struct Prec;
template <> struct Prec<SPEO_PREC_FP32> {
    static constexpr int parts = 0;
    static constexpr bool half = false;
};
template <> struct Prec<SPEO_PREC_BF16> {
    static constexpr int parts = 1;
    static constexpr bool half = false;
};
template <> struct Prec<SPEO_PREC_FP16> {
    static constexpr int parts = 1;
    static constexpr bool half = true;
};
template <> struct Prec<SPEO_PREC_BF16X3> {
    static constexpr int parts = 2;
    static constexpr bool half = false;
};
template <> struct Prec<SPEO_PREC_FP16X3> {
    static constexpr int parts = 2;
    static constexpr bool half = true;
};
template <> struct Prec<SPEO_PREC_BF16X6> {
    static constexpr int parts = 3;
    static constexpr bool half = false;
};

// A weighted-sphere column that outweighs all others together (discus's z_0;
// elliptic and bent cigar at small D): its weight is 0 in the FP32 weight
// table and w z_j^2 is added in FP64 (heavy_term), since any FP32 rotation
// loses the digits of a z_j that cancels.  j < 0: none.
struct HeavyCol {
    int j;
    double w;
    const double *m64;   // row j of M (FP64), or null: m32
    const float *m32;    // M[j][k] at m32[k * ld]
    int ld;
    const double *o64;   // shift (FP64), or null: o32
    const float *o32;
    double rate;
};

// One objective component as the kernels see it.
struct CompDev {
    int objective;           // SPEO_OBJ_*
    float rate;              // z = rate * M (x - o)
    const float *shift;      // [dim] FP32
    const float *B32;        // FP32: B[k][c] = M[c][k], row stride ldb32
    int ldb32;
    const uint16_t *M16[3];  // 16-bit parts of M, row-major [c][k], row stride ldm16
    int ldm16;
    int is_half;             // M16 holds FP16 (scaled by 2^14) instead of BF16
    const float *colw;       // elliptic column weights 10^(6j/(D-1)), [dim]
    // Small-D layouts (dim <= 128), zero padded, or null:
    // B32s: B = M^T (B[k][c] = M[c][k], zero padded to 128 x 128) in MFMA
    // fragment order: float4 q of lane l (lr = l % 16, lk = l / 16) of column
    // tile t sits at float offset ((t * 8 + q) * 64 + l) * 4 and holds
    // B[16 q + 4 e + lk][16 t + lr] for e = 0..3, i.e. the b operands of K
    // steps ks = 4 q .. 4 q + 3 of v_mfma_f32_16x16x4f32 (k = 4 ks + lk).
    const float *B32s;
    const uint16_t *M16s[3];    // [128][kp32]: 16-bit parts of M (FP16 scaled by 2^14)
    int kp32;                   // round_up(dim, 32)
    HeavyCol heavy;             // heavy.j < 0: none
};

constexpr int kSmallD = 128;  // widest dimension of the register-resident-B path
constexpr int kTS = 132;      // LDS row stride (floats) of staged rows, 16-byte aligned

// LDS footprint (bytes) of the GEMM tile for RT rows.
template <int RT, int PREC> constexpr int gemm_lds_bytes() {
    return Prec<PREC>::parts == 0
               ? (RT * (kKC + 2) + kKC * (kNC + 16)) * 4
               : Prec<PREC>::parts * (RT * (kKC + 8) + kNC * (kKC + 8)) * 2;
}

__device__ __forceinline__ uint16_t to_bits(float v, bool half) {
    if (half) {
        _Float16 h = (_Float16)v;
        return __builtin_bit_cast(uint16_t, h);
    }
    __bf16 b = (__bf16)v;
    return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ float from_bits(uint16_t v, bool half) {
    if (half) return (float)__builtin_bit_cast(_Float16, v);
    return (float)__builtin_bit_cast(__bf16, v);
}

// DPP lane permutation within rows of 16 lanes (GCN/CDNA DPP controls:
// quad_perm 0x00-0xFF, row_mirror 0x140, row_half_mirror 0x141).
template <int CTRL> __device__ __forceinline__ float dpp(float v) {
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                                 0xF, 0xF, false));
}

// Sum over the 16 lanes holding one row (lanes differing in bits 0..3); every
// lane of the row receives it.  The butterfly pairs lanes i^1, i^2 (quad
// permutations), then i and 7-i within each half row, then i and 15-i: after
// each step all lanes of the group hold one value, so every lane gets the
// xor-butterfly sum ((v0+v1)+(v2+v3)) + ... bit for bit.
__device__ __forceinline__ float sum16(float v) {
    v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp<0x141>(v);  // row_half_mirror
    v += dpp<0x140>(v);  // row_mirror
    return v;
}

// sin^2(pi z).  r = z - rint(z) lies in [-1/2, 1/2] and is exact, and
// sin^2(pi z) = sin^2(pi r) = u Q(u) with u = r^2, so no quadrant logic is
// needed.  Q is a degree-6 polynomial fitted (least squares on Chebyshev
// nodes, reweighted towards minimax relative error) to sin^2(pi sqrt(u)) / u
// on [0, 1/4]; its constant term is pi^2.  With FP32 coefficients and Horner
// steps the relative error stays below 2.3e-7 over the whole range.
__host__ __device__ __forceinline__ float sin2pi(float z) {
    const float r = z - rintf(z);
    const float u = r * r;
    float q = 0x1.7a242cp-1f;
    q = fmaf(q, u, -0x1.f48a3cp+1f);
    q = fmaf(q, u, 0x1.a6966cp+3f);
    q = fmaf(q, u, -0x1.e1f24cp+4f);
    q = fmaf(q, u, 0x1.55d3b8p+5f);
    q = fmaf(q, u, -0x1.03c1f0p+5f);
    q = fmaf(q, u, 0x1.3bd3ccp+3f);
    return q * u;
}

// sin x for 0 <= x <= 23 (the Schwefel terms take the sine of a square root
// of at most 500): x = k pi + r with k = rint(x / pi) <= 7 and |r| <= pi / 2,
// reduced with pi = 3.140625 + 9.6765358979e-4 (k * 3.140625 is exact and each
// step is one fma), then sin r by its Taylor polynomial through r^11, whose
// remainder is below (pi/2)^13 / 13! = 5.7e-8, negated for odd k.  Small
// enough to inline into the unrolled epilogues, where the general sinf's
// large-argument reduction is not.
__host__ __device__ __forceinline__ float sin_small(float x) {
    const float k = rintf(x * 0.318309886183790672f);
    float r = fmaf(-k, 3.140625f, x);
    r = fmaf(-k, 9.67653589793e-4f, r);
    const float r2 = r * r;
    float p = -1.0f / 39916800.0f;
    p = fmaf(p, r2, 1.0f / 362880.0f);
    p = fmaf(p, r2, -1.0f / 5040.0f);
    p = fmaf(p, r2, 1.0f / 120.0f);
    p = fmaf(p, r2, -1.0f / 6.0f);
    const float s = fmaf(p * r2, r, r);
    return (static_cast<int>(k) & 1) ? -s : s;
}

// Per-coordinate terms.  s1/s2 accumulate, per row:
//   elliptic:  s1 += w_j z^2
//   Rastrigin: s1 += z^2 + 20 sin^2(pi z)   (== z^2 - 10 cos(2 pi z) + 10)
//   Ackley:    s1 += z^2, s2 += sin^2(pi z) (cos(2 pi z) - 1 == -2 sin^2(pi z))
//   HappyCat, HGBat: s1 += (z - 1)^2, s2 += z - 1
//   Schwefel: v = z + 420.9687462275036; s1 -= v sin(sqrt|v|) for |v| <= 500,
//             else with m = 500 - fmod(|v|, 500): s1 -= sign(v) m sin(sqrt m),
//             s2 += ((|v| - 500) / 100)^2 (divided by D when finished)
__device__ __forceinline__ void add_terms(int objective, float z, float w, float &s1,
                                          float &s2) {
    if (objective == SPEO_OBJ_ELLIPTIC) {
        s1 = fmaf(w * z, z, s1);
    } else if (objective == SPEO_OBJ_SCHWEFEL) {
        const float v = z + 420.9687462275036f;
        const float av = fabsf(v);
        if (av <= 500.0f) {
            s1 -= v * sin_small(sqrtf(av));
        } else {
            const float m = 500.0f - fmodf(av, 500.0f);
            s1 -= copysignf(m, v) * sin_small(sqrtf(m));
            const float t = (av - 500.0f) * 0.01f;
            s2 = fmaf(t, t, s2);
        }
    } else if (objective == SPEO_OBJ_HAPPYCAT || objective == SPEO_OBJ_HGBAT) {
        const float u = z - 1.0f;
        s1 = fmaf(u, u, s1);
        s2 += u;
    } else {
        const float s2z = sin2pi(z);
        if (objective == SPEO_OBJ_RASTRIGIN) {
            s1 += fmaf(z, z, 20.0f * s2z);
        } else {
            s1 = fmaf(z, z, s1);
            s2 += s2z;
        }
    }
}

// Calls f(std::integral_constant<int, OBJ>{}) for the runtime objective, so
// that a loop of add_terms over a tile is compiled once per objective with
// its branches resolved, instead of testing the objective at every element.
// HappyCat and HGBat share their terms.
template <typename F>
__device__ __forceinline__ void with_objective(int objective, F &&f) {
    switch (objective) {
    case SPEO_OBJ_ELLIPTIC: f(std::integral_constant<int, SPEO_OBJ_ELLIPTIC>{}); break;
    case SPEO_OBJ_RASTRIGIN: f(std::integral_constant<int, SPEO_OBJ_RASTRIGIN>{}); break;
    case SPEO_OBJ_ACKLEY: f(std::integral_constant<int, SPEO_OBJ_ACKLEY>{}); break;
    case SPEO_OBJ_SCHWEFEL: f(std::integral_constant<int, SPEO_OBJ_SCHWEFEL>{}); break;
    default: f(std::integral_constant<int, SPEO_OBJ_HAPPYCAT>{}); break;
    }
}

// Objective value from the two row sums.  Ackley in the cancellation-free
// form -20 expm1(-0.2 r) - e expm1(mean(cos) - 1) of
// e - 20 exp(-0.2 r) + 20 - exp(mean(cos)) (CEC2014.cc:604-611).
__device__ __forceinline__ float finish_objective(int objective, float s1, float s2, int dim) {
    if (objective == SPEO_OBJ_ACKLEY) {
        const float r = sqrtf(s1 / (float)dim);
        return -20.0f * expm1f(-0.2f * r) - 2.71828182845904523536f * expm1f(-2.0f * s2 / (float)dim);
    }
    return s1;
}

// finish_objective plus HappyCat and HGBat (CEC2014.cc:907-953; the GPU forms
// global_f11 / global_f12, kernel_CEC2014_CUDA.cu:702-808) and Schwefel
// (CEC2014.cc:705-738; global_f9, kernel_CEC2014_CUDA.cu:588-648), which the
// evaluation and fused-generation kernels take; r2^2 - s^2 is formed as
// (r2 - s)(r2 + s).
__device__ __forceinline__ float finish_objective_ext(int objective, float s1, float s2, int dim) {
    const float D = (float)dim;
    if (objective == SPEO_OBJ_HAPPYCAT)
        return sqrtf(sqrtf(fabsf(s1 - D))) + (0.5f * s1 + s2) / D + 0.5f;
    if (objective == SPEO_OBJ_HGBAT)
        return sqrtf(fabsf((s1 - s2) * (s1 + s2))) + (0.5f * s1 + s2) / D + 0.5f;
    if (objective == SPEO_OBJ_SCHWEFEL) return fmaf(418.9828872724338f, D, s1 + s2 / D);
    return finish_objective(objective, s1, s2, dim);
}

// w z_j^2 of a heavy column for the row x (FP32 values,
// global or LDS), z_j = sum_k M_jk (x_k - o_k) rate in FP64: lane-strided
// partial sums, then the xor butterfly, so every kernel gets the same bits.
// All 64 lanes of the wave call it; the value is returned on every lane.
__device__ __forceinline__ double heavy_term(const HeavyCol &c, const float *x, int dim, int lane) {
    double z = 0.0;
    for (int k = lane; k < dim; k += 64) {
        const double m = c.m64 ? c.m64[k] : (double)c.m32[(size_t)k * c.ld];
        const double o = c.o64 ? c.o64[k] : (double)c.o32[k];
        z = fma(m, ((double)x[k] - o) * c.rate, z);
    }
    for (int off = 32; off > 0; off >>= 1) z += __shfl_xor(z, off, 64);
    return c.w * z * z;
}

// The FP32 row sum with the heavy column's term added (the term rounded to
// FP32 first, so that kernels that park it in an FP32 slot of LDS get the
// same bits).
__device__ __forceinline__ float with_heavy(float s1, float h) {
    return (float)((double)s1 + (double)h);
}

// Rows of the A tile: row r of the tile is `rows + r * ld` (global or LDS),
// r < nrows; coordinates j < dim.
struct RowSource {
    const float *rows;
    int ld;
    int nrows;
};

// Vector staging of one K chunk of gemm_chunk, for rows in global memory that
// are 16-byte aligned with a row stride of a multiple of 4 floats: every
// 16-byte load of the chunk is issued before the LDS operands are formed
// (scalar loads only for a group of four that straddles the end of K).
// FP32 (ChunkF32) splits this in two: load() fills registers, store() writes
// LDS, and gemm_chunk stores chunk k and then loads chunk k + 1, whose loads
// are in flight during chunk k's MFMA work.  The 16-bit form
// (stage_chunk_16) loads and stores each chunk in turn: holding its P-fold B
// operand across the MFMA loop measured slower than the overlap.
// The values are those of the scalar loops: A = (x - o) * rate (* sa) for
// r < nrows, k < kc, else 0; B from the operand's zero-padded rows (B32 rows
// padded to a multiple of 4 floats, M16 rows to a multiple of 8).
template <int RT>
struct ChunkF32 {
    static constexpr int LDA = kKC + 2, LDB = kNC + 16;
    static constexpr int GA = RT * kKC / 4 / 256, GB = kKC * kNC / 4 / 256;
    floatx4 xa[GA], oa[GA], xb[GB];
    __device__ __forceinline__ void load(const CompDev &c, const RowSource &src, int k0, int kc, int n0,
                                         int ncols, int tid) {
#pragma unroll
        for (int i = 0; i < GA; ++i) {
            const int e = tid + i * 256, r = e / (kKC / 4), k = (e % (kKC / 4)) * 4;
            const bool full = r < src.nrows && k + 3 < kc;
            xa[i] = gld(reinterpret_cast<const floatx4 *>(full ? src.rows + (size_t)r * src.ld + k0 + k : src.rows));
            oa[i] = gld(reinterpret_cast<const floatx4 *>(full ? c.shift + k0 + k : c.shift));
        }
#pragma unroll
        for (int i = 0; i < GB; ++i) {
            const int e = tid + i * 256, k = e / (kNC / 4), n = (e % (kNC / 4)) * 4;
            const bool ok = k < kc && n < ncols;
            xb[i] = gld(reinterpret_cast<const floatx4 *>(c.B32 + (ok ? (size_t)(k0 + k) * c.ldb32 + n0 + n : 0)));
        }
    }
    __device__ __forceinline__ void store(const CompDev &c, const RowSource &src, int k0, int kc, int ncols,
                                          float *As, float *Bs, int tid) const {
#pragma unroll
        for (int i = 0; i < GA; ++i) {
            const int e = tid + i * 256, r = e / (kKC / 4), k = (e % (kKC / 4)) * 4;
            const bool full = r < src.nrows && k + 3 < kc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                float v;
                if (full)
                    v = (xa[i][q] - oa[i][q]) * c.rate;
                else
                    v = (r < src.nrows && k + q < kc)
                            ? (src.rows[(size_t)r * src.ld + k0 + k + q] - c.shift[k0 + k + q]) * c.rate
                            : 0.f;
                As[r * LDA + k + q] = v;
            }
        }
#pragma unroll
        for (int i = 0; i < GB; ++i) {
            const int e = tid + i * 256, k = e / (kNC / 4), n = (e % (kNC / 4)) * 4;
            const bool ok = k < kc && n < ncols;
            *reinterpret_cast<floatx4 *>(Bs + k * LDB + n) = ok ? xb[i] : floatx4{0.f, 0.f, 0.f, 0.f};
        }
    }
};

template <int RT, int P, bool H>
__device__ __forceinline__ void stage_chunk_16(const CompDev &c, const RowSource &src, int k0, int kc, int n0,
                                               int ncols, uint16_t *As, uint16_t *Bs, int tid) {
    constexpr int LDA = kKC + 8, LDB = kKC + 8;
    constexpr int GA = RT * kKC / 4 / 256, GB = kNC * kKC / 8 / 256;
    typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const float sa = H ? kF16ScaleA : 1.0f;
    floatx4 xa[GA], oa[GA];
    u32x4 xb[P][GB];
#pragma unroll
    for (int i = 0; i < GA; ++i) {
        const int e = tid + i * 256, r = e / (kKC / 4), k = (e % (kKC / 4)) * 4;
        const bool full = r < src.nrows && k + 3 < kc;
        xa[i] = gld(reinterpret_cast<const floatx4 *>(full ? src.rows + (size_t)r * src.ld + k0 + k : src.rows));
        oa[i] = gld(reinterpret_cast<const floatx4 *>(full ? c.shift + k0 + k : c.shift));
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int i = 0; i < GB; ++i) {
            const int e = tid + i * 256, n = e / (kKC / 8), k = (e % (kKC / 8)) * 8;
            const bool ok = k < kc && n < ncols;
            xb[p][i] = gld(reinterpret_cast<const u32x4 *>(c.M16[p] + (ok ? (size_t)(n0 + n) * c.ldm16 + k0 + k : 0)));
        }
#pragma unroll
    for (int i = 0; i < GA; ++i) {
        const int e = tid + i * 256, r = e / (kKC / 4), k = (e % (kKC / 4)) * 4;
        const bool full = r < src.nrows && k + 3 < kc;
        u16x4 h[P];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float v;
            if (full)
                v = (xa[i][q] - oa[i][q]) * c.rate * sa;
            else
                v = (r < src.nrows && k + q < kc)
                        ? (src.rows[(size_t)r * src.ld + k0 + k + q] - c.shift[k0 + k + q]) * c.rate * sa
                        : 0.f;
            float rem = v;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const uint16_t b = to_bits(rem, H);
                h[p][q] = b;
                rem -= from_bits(b, H);
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) *reinterpret_cast<u16x4 *>(As + (p * RT + r) * LDA + k) = h[p];
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
        for (int i = 0; i < GB; ++i) {
            const int e = tid + i * 256, n = e / (kKC / 8), k = (e % (kKC / 8)) * 8;
            const bool ok = k < kc && n < ncols;
            *reinterpret_cast<u32x4 *>(Bs + (p * kNC + n) * LDB + k) = ok ? xb[p][i] : u32x4{0u, 0u, 0u, 0u};
        }
}

// ------------------------------------------------------------------ GEMM
//
// Computes the Z columns [n0, n0 + kNC) of the tile's RT rows for component
// `c` and adds their terms to s1[i], s2[i] (row 4*(lane/16)+i of this wave's
// row tile; only lanes with lane%16 == 0 hold the full sums after the final
// sum16, done by the caller).  All 256 threads of the block must call it.
// src.rows is in global memory.
// `lds` must hold gemm_lds_bytes<RT, PREC>() bytes.
// VS: vector staging (ChunkF32 / stage_chunk_16) where the rows allow it.  Its registers
// count in every kernel that contains the call, so a kernel whose other paths
// need them (the generation kernel's small-D phases) keeps the element-wise
// form.
template <int RT, int PREC, bool VS = true>
__device__ __forceinline__ void gemm_chunk(const CompDev &c, RowSource src, int dim, int n0,
                                           char *lds, float (&s1)[4], float (&s2)[4]) {
    constexpr int NSPLIT = 4 / (RT / 16);  // waves sharing one 16-row tile
    constexpr int TPW = 8 / NSPLIT;        // 16-column tiles per wave
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int rt = wave / NSPLIT, wn = wave % NSPLIT;
    const int lr = lane & 15, lk = lane >> 4;

    floatx4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};

    const int ncols = min(kNC, dim - n0);
    // rows and B operand aligned for the vector staging (uniform over the
    // block): FP32 reads B32 rows as float4 (stride a multiple of 4 floats),
    // the 16-bit forms read M16 rows as groups of 8 (stride a multiple of 8).
    // The drop-in evaluation hands the caller's B with a row stride of dim,
    // which at dim = 10, 30, 50 takes the element-wise form.
    const bool b_vec = Prec<PREC>::parts == 0
                           ? (c.ldb32 & 3) == 0 && (reinterpret_cast<uintptr_t>(c.B32) & 15) == 0
                           : (c.ldm16 & 7) == 0 && (reinterpret_cast<uintptr_t>(c.M16[0]) & 15) == 0 &&
                                 (Prec<PREC>::parts < 2 || (reinterpret_cast<uintptr_t>(c.M16[1]) & 15) == 0) &&
                                 (Prec<PREC>::parts < 3 || (reinterpret_cast<uintptr_t>(c.M16[2]) & 15) == 0);
    const bool vec = VS && b_vec && (src.ld & 3) == 0 && (reinterpret_cast<uintptr_t>(src.rows) & 15) == 0;

    if constexpr (Prec<PREC>::parts == 0) {
        float *As = reinterpret_cast<float *>(lds);  // [RT][kKC+2]
        float *Bs = As + RT * (kKC + 2);             // [kKC][kNC+16]
        constexpr int LDA = kKC + 2, LDB = kNC + 16;
        ChunkF32<RT> stg;
        if (VS && vec) stg.load(c, src, 0, min(kKC, dim), n0, ncols, tid);
        for (int k0 = 0; k0 < dim; k0 += kKC) {
            const int kc = min(kKC, dim - k0);
            __syncthreads();
            if (VS && vec) {
                stg.store(c, src, k0, kc, ncols, As, Bs, tid);
                if (k0 + kKC < dim) stg.load(c, src, k0 + kKC, min(kKC, dim - k0 - kKC), n0, ncols, tid);
            } else {
                for (int e = tid; e < RT * kKC; e += 256) {
                    const int r = e / kKC, k = e % kKC;
                    float v = 0.f;
                    if (r < src.nrows && k < kc)
                        v = (src.rows[(size_t)r * src.ld + k0 + k] - c.shift[k0 + k]) * c.rate;
                    As[r * LDA + k] = v;
                }
                for (int e = tid; e < kKC * kNC; e += 256) {
                    const int k = e / kNC, n = e % kNC;
                    float v = 0.f;
                    if (k < kc && n < ncols) v = c.B32[(size_t)(k0 + k) * c.ldb32 + n0 + n];
                    Bs[k * LDB + n] = v;
                }
            }
            __syncthreads();
            const int ksteps = (kc + 3) >> 2;
            for (int ks = 0; ks < ksteps; ++ks) {
                const float a = As[(rt * 16 + lr) * LDA + ks * 4 + lk];
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    const int tile = wn + t * NSPLIT;
                    if (tile * 16 < ncols) {
                        const float b = Bs[(ks * 4 + lk) * LDB + tile * 16 + lr];
                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[t], 0, 0, 0);
                    }
                }
            }
        }
    } else {
        constexpr int P = Prec<PREC>::parts;
        constexpr bool H = Prec<PREC>::half;
        constexpr int LDA = kKC + 8, LDB = kKC + 8;
        uint16_t *As = reinterpret_cast<uint16_t *>(lds);  // [P][RT][LDA]
        uint16_t *Bs = As + P * RT * LDA;                  // [P][kNC][LDB] (B^T: column-major K)
        const float sa = H ? kF16ScaleA : 1.0f;
        for (int k0 = 0; k0 < dim; k0 += kKC) {
            const int kc = min(kKC, dim - k0);
            __syncthreads();
            if (VS && vec) {
                stage_chunk_16<RT, P, H>(c, src, k0, kc, n0, ncols, As, Bs, tid);
            } else {
                for (int e = tid; e < RT * kKC; e += 256) {
                    const int r = e / kKC, k = e % kKC;
                    float v = 0.f;
                    if (r < src.nrows && k < kc)
                        v = (src.rows[(size_t)r * src.ld + k0 + k] - c.shift[k0 + k]) * c.rate * sa;
                    float rem = v;
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        const uint16_t b = to_bits(rem, H);
                        As[(p * RT + r) * LDA + k] = b;
                        rem -= from_bits(b, H);
                    }
                }
                for (int e = tid; e < kNC * kKC; e += 256) {
                    const int n = e / kKC, k = e % kKC;
                    const bool ok = (k < kc && n < ncols);
#pragma unroll
                    for (int p = 0; p < P; ++p)
                        Bs[(p * kNC + n) * LDB + k] =
                            ok ? c.M16[p][(size_t)(n0 + n) * c.ldm16 + k0 + k] : (uint16_t)0;
                }
            }
            __syncthreads();
            const int ksteps = (kc + 31) >> 5;
            for (int ks = 0; ks < ksteps; ++ks) {
                const int kb = ks * 32 + lk * 8;
                typedef typename std::conditional<H, halfx8, bf16x8>::type vec8;
                vec8 a[P];
#pragma unroll
                for (int p = 0; p < P; ++p)
                    a[p] = *reinterpret_cast<const vec8 *>(&As[(p * RT + rt * 16 + lr) * LDA + kb]);
#pragma unroll
                for (int t = 0; t < TPW; ++t) {
                    const int tile = wn + t * NSPLIT;
                    if (tile * 16 < ncols) {
                        vec8 b[P];
#pragma unroll
                        for (int p = 0; p < P; ++p)
                            b[p] = *reinterpret_cast<const vec8 *>(
                                &Bs[(p * kNC + tile * 16 + lr) * LDB + kb]);
                        // small products first, then hi*hi
#pragma unroll
                        for (int sum = P - 1; sum >= 0; --sum) {
#pragma unroll
                            for (int pa = 0; pa <= sum; ++pa) {
                                const int pb = sum - pa;
                                if (pa < P && pb < P) {
                                    if constexpr (H)
                                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                                            a[pa], b[pb], acc[t], 0, 0, 0);
                                    else
                                        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                            a[pa], b[pb], acc[t], 0, 0, 0);
                                }
                            }
                        }
                    }
                }
            }
        }
    }

    // epilogue: terms of this wave's columns
    const float unscale = Prec<PREC>::half ? kF16Unscale : 1.0f;
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int col = n0 + (wn + t * NSPLIT) * 16 + lr;
        if (col < dim) {
            const float w = (c.objective == SPEO_OBJ_ELLIPTIC) ? c.colw[col] : 1.0f;
            with_objective(c.objective, [&](auto obj) {
#pragma unroll
                for (int i = 0; i < 4; ++i) add_terms(decltype(obj)::value, acc[t][i] * unscale, w, s1[i], s2[i]);
            });
        }
    }
}

// Word 3 of a gfx9 buffer resource (raw buffer, 32-bit data format).
constexpr int kBufferWord3 = 0x00020000;

// The FP32 b fragments of column tile t for K steps ks < KS, rounded up to a
// block of four, from the fragment-ordered B32s (zero past the dimension).
// Blocks past that, and every block of a tile past the dimension, are left
// undefined: callers use b[ks] only for ks < round_up(KS, 4) of valid tiles.
// The 16-byte loads go through a buffer resource over the tile, behind
// wave-uniform branches, and nothing merges a loaded value with another one
// (which would make the wave wait for that load right after issuing it).
__device__ __forceinline__ void load_b_frag(const float *B32s, int t, bool valid, int KS, int lane,
                                            float (&b)[32]) {
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float *>(B32s) + (size_t)t * 2048, (short)0, 8192, kBufferWord3);
#pragma unroll
    for (int q = 0; q < 8; ++q) {
        if (valid && q * 4 < KS) {
            const floatx4 f =
                __builtin_bit_cast(floatx4, __builtin_amdgcn_raw_buffer_load_b128(r, lane * 16 + q * 1024, 0, 0));
#pragma unroll
            for (int e = 0; e < 4; ++e) b[4 * q + e] = f[e];
        }
    }
}

// ------------------------------------------------------- small-D path
//
// dim <= 128.  The tile's RT rows are in LDS (rows[r * kTS + k], FP32,
// unshifted; rows >= nrows and columns >= dim hold zeros).  Wave w owns the
// 16-column tiles w and w + 4 for all RT rows; its B fragments come straight
// from the padded global layouts into registers (no LDS staging, no barrier
// inside the K loop).  `scratch` holds the shifted rows' 16-bit parts for the
// 16-bit precisions: P * RT * (kp32 + 8) uint16, and the shift: 128 floats.
// Output: s1[i][q], s2[i][q] for row 16 i + 4 (lane / 16) + q of the tile,
// summed over this wave's columns and the lane's 16-lane group.
template <int RT> constexpr int small_scratch_bytes(int parts) {
    return 128 * 4 + parts * RT * (kSmallD + 8) * 2;
}

template <int RT, int PREC>
__device__ __forceinline__ void gemm_small(const CompDev &c, const float *rows, int dim,
                                           char *scratch, float (&s1)[RT / 16][4],
                                           float (&s2)[RT / 16][4]) {
    constexpr int NRT = RT / 16;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    float *sh = reinterpret_cast<float *>(scratch);  // shift, then rate-scaled in place
    floatx4 acc[NRT][2];
#pragma unroll
    for (int i = 0; i < NRT; ++i) {
        acc[i][0] = floatx4{0.f, 0.f, 0.f, 0.f};
        acc[i][1] = floatx4{0.f, 0.f, 0.f, 0.f};
    }
    const int t0 = wave, t1 = wave + 4;
    const bool v0 = t0 * 16 < dim, v1 = t1 * 16 < dim;

    if constexpr (Prec<PREC>::parts == 0) {
        const int kp4 = (dim + 3) & ~3, KS = kp4 >> 2;
        // B fragments into registers first (their latency overlaps the shift staging)
        float b0[32], b1[32];
        load_b_frag(c.B32s, t0, v0, KS, lane, b0);
        load_b_frag(c.B32s, t1, v1, KS, lane, b1);
        for (int k = tid; k < kSmallD; k += 256) sh[k] = k < dim ? gld(c.shift + k) : 0.f;
        __syncthreads();
        const float rate = c.rate;
#pragma unroll
        for (int ks = 0; ks < 32; ++ks) {
            if (ks < KS) {
                const int k = ks * 4 + lk;
                const float o = sh[k];
#pragma unroll
                for (int i = 0; i < NRT; ++i) {
                    const float a = (rows[(i * 16 + lr) * kTS + k] - o) * rate;
                    if (v0) acc[i][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b0[ks], acc[i][0], 0, 0, 0);
                    if (v1) acc[i][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b1[ks], acc[i][1], 0, 0, 0);
                }
            }
        }
    } else {
        constexpr int P = Prec<PREC>::parts;
        constexpr bool H = Prec<PREC>::half;
        typedef typename std::conditional<H, halfx8, bf16x8>::type vec8;
        const int kp32 = c.kp32, KS = kp32 >> 5, LDA = kSmallD + 8;
        uint16_t *As = reinterpret_cast<uint16_t *>(sh + kSmallD);  // [P][RT][LDA]
        for (int k = tid; k < kSmallD; k += 256) sh[k] = k < dim ? gld(c.shift + k) : 0.f;
        __syncthreads();
        const float scale = c.rate * (H ? kF16ScaleA : 1.0f);
        for (int e = tid; e < RT * kp32; e += 256) {
            const int r = e / kp32, k = e % kp32;
            float rem = k < dim ? (rows[r * kTS + k] - sh[k]) * scale : 0.f;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const uint16_t b = to_bits(rem, H);
                As[(p * RT + r) * LDA + k] = b;
                rem -= from_bits(b, H);
            }
        }
        __syncthreads();
        for (int ks = 0; ks < KS; ++ks) {
            const int kb = ks * 32 + lk * 8;
            vec8 b[2][P];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                b[0][p] = v0 ? gld(reinterpret_cast<const vec8 *>(c.M16s[p] + (size_t)(t0 * 16 + lr) * kp32 + kb)) : vec8{};
                b[1][p] = v1 ? gld(reinterpret_cast<const vec8 *>(c.M16s[p] + (size_t)(t1 * 16 + lr) * kp32 + kb)) : vec8{};
            }
#pragma unroll
            for (int i = 0; i < NRT; ++i) {
                vec8 a[P];
#pragma unroll
                for (int p = 0; p < P; ++p)
                    a[p] = *reinterpret_cast<const vec8 *>(&As[(p * RT + i * 16 + lr) * LDA + kb]);
#pragma unroll
                for (int t = 0; t < 2; ++t) {
                    if (t == 0 ? !v0 : !v1) continue;
#pragma unroll
                    for (int sum = P - 1; sum >= 0; --sum) {
#pragma unroll
                        for (int pa = 0; pa <= sum; ++pa) {
                            const int pb = sum - pa;
                            if (pa < P && pb < P) {
                                if constexpr (H)
                                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[pa], b[t][pb], acc[i][t], 0, 0, 0);
                                else
                                    acc[i][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[pa], b[t][pb], acc[i][t], 0, 0, 0);
                            }
                        }
                    }
                }
            }
        }
    }
    const float unscale = Prec<PREC>::half ? kF16Unscale : 1.0f;
#pragma unroll
    for (int i = 0; i < NRT; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            s1[i][q] = 0.f;
            s2[i][q] = 0.f;
        }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = (t == 0 ? t0 : t1) * 16 + lr;
        if (col < dim) {
            const float w = (c.objective == SPEO_OBJ_ELLIPTIC) ? c.colw[col] : 1.0f;
            with_objective(c.objective, [&](auto obj) {
#pragma unroll
                for (int i = 0; i < NRT; ++i)
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        add_terms(decltype(obj)::value, acc[i][t][q] * unscale, w, s1[i][q], s2[i][q]);
            });
        }
    }
#pragma unroll
    for (int i = 0; i < NRT; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            s1[i][q] = sum16(s1[i][q]);
            s2[i][q] = sum16(s2[i][q]);
        }
}

// Per-row sums of gemm_small's per-wave partials (fixed order over waves).
// red: 2 * 4 * RT floats of LDS; out1/out2: RT floats each.
template <int RT>
__device__ __forceinline__ void reduce_rows_small(const float (&s1)[RT / 16][4],
                                                  const float (&s2)[RT / 16][4], float *red,
                                                  float *out1, float *out2) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    __syncthreads();
    if ((lane & 15) == 0) {
#pragma unroll
        for (int i = 0; i < RT / 16; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = i * 16 + (lane >> 4) * 4 + q;
                red[(wave * RT + r) * 2 + 0] = s1[i][q];
                red[(wave * RT + r) * 2 + 1] = s2[i][q];
            }
    }
    __syncthreads();
    if (tid < RT) {
        float a = 0.f, b = 0.f;
        for (int w = 0; w < 4; ++w) {
            a += red[(w * RT + tid) * 2 + 0];
            b += red[(w * RT + tid) * 2 + 1];
        }
        out1[tid] = a;
        out2[tid] = b;
    }
    __syncthreads();
}

// Reduces s1/s2 of all waves of the block into per-row sums:
// out1[r], out2[r] for r < RT (LDS, caller-provided, >= 2 * 4 * RT floats of
// scratch in `red`).  Deterministic order.  All threads call it.
template <int RT>
__device__ __forceinline__ void reduce_rows(float (&s1)[4], float (&s2)[4], float *red,
                                           float *out1, float *out2) {
    constexpr int NSPLIT = 4 / (RT / 16);
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int rt = wave / NSPLIT, wn = wave % NSPLIT;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        s1[i] = sum16(s1[i]);
        s2[i] = sum16(s2[i]);
    }
    __syncthreads();
    if ((lane & 15) == 0) {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = rt * 16 + (lane >> 4) * 4 + i;
            red[(wn * RT + r) * 2 + 0] = s1[i];
            red[(wn * RT + r) * 2 + 1] = s2[i];
        }
    }
    __syncthreads();
    if (tid < RT) {
        float a = 0.f, b = 0.f;
        for (int w = 0; w < NSPLIT; ++w) {
            a += red[(w * RT + tid) * 2 + 0];
            b += red[(w * RT + tid) * 2 + 1];
        }
        out1[tid] = a;
        out2[tid] = b;
    }
    __syncthreads();
}

}  // namespace speo
