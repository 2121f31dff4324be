// surrogate.hip -- HIP kernels of the surrogate MLP (forward, backward, Adam),
// its archive and rank statistics, and the candidate-screening kernels.
//
// GEMMs run on v_mfma_f32_16x16x32_bf16: FP32 operands are rounded to BF16
// (round to nearest even) while staged to LDS, accumulation is FP32.  The
// output layer (256 -> 1), the loss, the SiLU derivatives and Adam run in
// FP32 on the vector ALUs.
#include "surrogate.h"

#include <algorithm>
#include <atomic>

#include "de.h"
#include "eval_core.h"

namespace speo {

enum { kEpiStore = 0, kEpiBiasSilu = 1, kEpiMulDsilu = 2 };

struct GemmArgs {
    int M, N, K;
    const float *A;
    long a_m, a_k;  // A(m, k) = A[m * a_m + k * a_k]
    const float *B;
    long b_k, b_n;  // B(k, n) = B[k * b_k + n * b_n]
    float a_scale, b_scale;
    float *C;
    long c_m;
    const float *bias;  // [N]
    float *P;           // pre-activations: written by kEpiBiasSilu, read by kEpiMulDsilu
    long p_m;
    int epi;
    int kc;             // K per split (blockIdx.z); C of split z at C + z * c_split
    long c_split;
    int vec_a, vec_b;   // 16-byte loads along the unit-stride index are aligned
    int bytes_a, bytes_b;  // extent of A and B in bytes (the buffer range of their loads)
};

// Whether 16-byte loads along the unit-stride index of a matrix are aligned.
static int vec_ok(const float *p, long s_row, long s_k) {
    const long other = s_k == 1 ? s_row : s_k;
    return (reinterpret_cast<uintptr_t>(p) % 16 == 0) && (other % 4 == 0) && (s_row == 1 || s_k == 1);
}

// Bytes spanned by the elements (r, k), r < rows, k < K, of a strided matrix
// (0 when it does not fit the 32-bit buffer range).
static int extent_bytes(long rows, long K, long s_row, long s_k) {
    const long e = ((rows - 1) * s_row + (K - 1) * s_k + 1) * 4;
    return e > 0 && e < 0x7fffffffL ? (int)e : 0;
}

// sigmoid(a) = 1 / (1 + 2^(-a log2(e))) on the hardware base-2 exponential
// (v_exp_f32) and reciprocal (v_rcp_f32), both within 1 ulp: about 1e-6
// relative error for |a| < 20, far below the BF16 rounding of the
// activations, at a fifth of the instructions of expf and an IEEE divide.
// Large |a| saturate to 0 and 1 (2^(+big) = inf, 1 / inf = 0).
__device__ __forceinline__ float sigmoidf_(float a) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(a * -1.44269504088896341f));
}
__device__ __forceinline__ float siluf_(float a) { return a * sigmoidf_(a); }
__device__ __forceinline__ float dsiluf_(float a) {
    const float s = sigmoidf_(a);
    return s * (1.0f + a * (1.0f - s));
}

// 32W x 32W tile, 4 waves as 2 x 2, each wave 16W x 16W = W x W MFMA 16x16
// tiles (W = 2: 64 x 64; W = 1: 32 x 32 for the small GEMMs, four times the
// workgroups and a quarter of each one's staging and epilogue), K staged
// 128 at a time, so the MLP's K <= 256 costs one or two global-load round
// trips per tile; split z of the K range
// [z kc, (z + 1) kc) writes its own C.  Operand tiles are staged with the
// unit-stride index fastest, so the global reads coalesce for either layout.
// Every output element accumulates its K blocks in the same order for any W.
template <int W>
__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
    constexpr int TM = 32 * W, TN = 32 * W, TK = 128, LDK = TK + 8;
    __shared__ __align__(16) uint16_t As[TM * LDK];
    __shared__ __align__(16) uint16_t Bs[TN * LDK];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = wave >> 1, wn = wave & 1, lr = lane & 15, lk = lane >> 4;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int kbeg = blockIdx.z * g.kc, kend = min(g.K, kbeg + g.kc);
    float *C = g.C + (long)blockIdx.z * g.c_split;
    const bool a_kfast = g.a_k == 1, b_kfast = g.b_k == 1;
    floatx4 acc[W][W];
#pragma unroll
    for (int i = 0; i < W; ++i)
#pragma unroll
        for (int j = 0; j < W; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // One operand tile: `rows` x TK (K fast) or TK x `rows` (rows fast) from a
    // global matrix whose unit-stride index is the fast one.  Every load of
    // the tile is issued before any is used: a 16-byte load per group of four
    // that lies inside the tile (when the matrix allows them), 4-byte loads
    // for the groups that straddle its edge (rare: a second pass) or for the
    // whole tile otherwise; elements outside the tile's rows or the split's
    // K range are 0.  Buffer loads, their range the matrix extent.
    auto stage = [&](const float *src, long s_row, long s_k, bool kfast, bool vec, int bytes, int row0,
                     int nrows_total, int k0, float scale, uint16_t *dst) __attribute__((always_inline)) {
        constexpr int NI = TM * TK / 4 / 256;  // float4 groups per thread
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(src), (short)0, bytes, kBufferWord3);
        const long s_q = kfast ? s_k : s_row;  // stride between the group's four elements
        auto elem_ok = [&](int gr, int gk, int q) {
            return kfast ? (gr < nrows_total && gk + q < kend) : (gr + q < nrows_total && gk < kend);
        };
        auto elem_off = [&](int gr, int gk, int q) -> uint32_t {
            return elem_ok(gr, gk, q) ? (uint32_t)(((long)gr * s_row + (long)gk * s_k + q * s_q) * 4) : 0u;
        };
        auto ld1 = [&](uint32_t off) {
            return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, off, 0, 0));
        };
        // the group of thread slot i: tile row r, tile column k
        auto slot = [&](int i, int &r, int &k) {
            const int e4 = tid + i * 256;
            r = kfast ? e4 / (TK / 4) : (e4 % (TM / 4)) * 4;
            k = kfast ? (e4 % (TK / 4)) * 4 : e4 / (TM / 4);
        };
        // raw loads first (no use of a loaded value until all are issued)
        float v[NI][4];
        if (vec) {
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                int r, k;
                slot(i, r, k);
                const int gr = row0 + r, gk = k0 + k;
                const bool full = elem_ok(gr, gk, 3) && elem_ok(gr, gk, 0);
                const floatx4 x = __builtin_bit_cast(
                    floatx4, __builtin_amdgcn_raw_buffer_load_b128(rs, full ? elem_off(gr, gk, 0) : 0u, 0, 0));
#pragma unroll
                for (int q = 0; q < 4; ++q) v[i][q] = x[q];
            }
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                int r, k;
                slot(i, r, k);
                const int gr = row0 + r, gk = k0 + k;
                if (elem_ok(gr, gk, 0) && !elem_ok(gr, gk, 3)) {  // straddles the tile's edge
#pragma unroll
                    for (int q = 0; q < 4; ++q) v[i][q] = ld1(elem_off(gr, gk, q));
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < NI; ++i) {
                int r, k;
                slot(i, r, k);
                const int gr = row0 + r, gk = k0 + k;
#pragma unroll
                for (int q = 0; q < 4; ++q) v[i][q] = ld1(elem_off(gr, gk, q));
            }
        }
        // then the elements outside the tile set to 0, BF16, LDS
#pragma unroll
        for (int i = 0; i < NI; ++i) {
            int r, k;
            slot(i, r, k);
            const int gr = row0 + r, gk = k0 + k;
            uint16_t h[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) h[q] = to_bits((elem_ok(gr, gk, q) ? v[i][q] : 0.f) * scale, false);
            if (kfast) {
                typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
                *reinterpret_cast<u16x4 *>(&dst[r * LDK + k]) = u16x4{h[0], h[1], h[2], h[3]};
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q) dst[(r + q) * LDK + k] = h[q];
            }
        }
    };
    for (int k0 = kbeg; k0 < kend; k0 += TK) {
        __syncthreads();
        stage(g.A, g.a_m, g.a_k, a_kfast, g.vec_a != 0, g.bytes_a, m0, g.M, k0, g.a_scale, As);
        stage(g.B, g.b_n, g.b_k, b_kfast, g.vec_b != 0, g.bytes_b, n0, g.N, k0, g.b_scale, Bs);
        __syncthreads();
#pragma unroll
        for (int ks = 0; ks < TK; ks += 32) {
            bf16x8 a[W], b[W];
#pragma unroll
            for (int i = 0; i < W; ++i) {
                a[i] = *reinterpret_cast<const bf16x8 *>(&As[(wm * 16 * W + i * 16 + lr) * LDK + ks + lk * 8]);
                b[i] = *reinterpret_cast<const bf16x8 *>(&Bs[(wn * 16 * W + i * 16 + lr) * LDK + ks + lk * 8]);
            }
#pragma unroll
            for (int i = 0; i < W; ++i)
#pragma unroll
                for (int j = 0; j < W; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
    }
#pragma unroll
    for (int i = 0; i < W; ++i)
#pragma unroll
        for (int j = 0; j < W; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int m = m0 + wm * 16 * W + i * 16 + lk * 4 + q;
                const int n = n0 + wn * 16 * W + j * 16 + lr;
                if (m >= g.M || n >= g.N) continue;
                float v = acc[i][j][q];
                if (g.epi == kEpiBiasSilu) {
                    v += g.bias ? g.bias[n] : 0.f;
                    if (g.P) g.P[(long)m * g.p_m + n] = v;
                    v = siluf_(v);
                } else if (g.epi == kEpiMulDsilu) {
                    v *= dsiluf_(g.P[(long)m * g.p_m + n]);
                }
                C[(long)m * g.c_m + n] = v;
            }
}

// C[m][n] = sum over the splits of part[z][m][n], in split order.
__global__ void k_splitk_reduce(const float *part, int splits, int M, int N, float *C, long c_m) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)M * N) return;
    float v = 0.f;
    for (int z = 0; z < splits; ++z) v += part[(long)z * M * N + i];
    C[(i / N) * c_m + i % N] = v;
}

// The operands' buffer ranges (their extents must fit 31 bits).
static int set_extents(GemmArgs &g) {
    g.bytes_a = extent_bytes(g.M, g.K, g.a_m, g.a_k);
    g.bytes_b = extent_bytes(g.N, g.K, g.b_n, g.b_k);
    if (g.bytes_a == 0 || g.bytes_b == 0) return set_error(SPEO_ERR_ARG, "mlp gemm: operand larger than 2 GiB");
    return SPEO_OK;
}

// 64 x 64 tiles when they give the device at least one workgroup per compute
// unit (256), else 32 x 32.
static int launch_gemm(const GemmArgs &g, int splits, hipStream_t s) {
    const long tiles64 = (long)((g.N + 63) / 64) * ((g.M + 63) / 64) * splits;
    if (tiles64 >= 256) {
        dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, splits);
        hipLaunchKernelGGL(k_gemm_bf16<2>, grid, dim3(256), 0, s, g);
    } else {
        dim3 grid((g.N + 31) / 32, (g.M + 31) / 32, splits);
        hipLaunchKernelGGL(k_gemm_bf16<1>, grid, dim3(256), 0, s, g);
    }
    return check_launch("k_gemm_bf16");
}

static int gemm(const GemmArgs &g0, hipStream_t s) {
    if (g0.M <= 0 || g0.N <= 0) return SPEO_OK;
    GemmArgs g = g0;
    g.kc = g.K;
    g.c_split = 0;
    g.vec_a = vec_ok(g.A, g.a_m, g.a_k);
    g.vec_b = vec_ok(g.B, g.b_n, g.b_k);
    SPEO_TRY(set_extents(g));
    return launch_gemm(g, 1, s);
}

// K split into chunks of 128 (plain store epilogue only), partial products
// reduced in a fixed order: the weight gradients, whose K is the batch.
static int gemm_splitk(const GemmArgs &g0, float *part, hipStream_t s) {
    if (g0.M <= 0 || g0.N <= 0) return SPEO_OK;
    const int splits = std::min(kMaxSplitK, (g0.K + 127) / 128);
    if (splits <= 1) return gemm(g0, s);
    GemmArgs g = g0;
    g.vec_a = vec_ok(g.A, g.a_m, g.a_k);
    g.vec_b = vec_ok(g.B, g.b_n, g.b_k);
    SPEO_TRY(set_extents(g));
    g.kc = (g.K + splits - 1) / splits;
    g.kc = (g.kc + 31) / 32 * 32;
    g.C = part;
    g.c_m = g.N;
    g.c_split = (long)g.M * g.N;
    SPEO_TRY(launch_gemm(g, splits, s));
    const long total = (long)g.M * g.N;
    hipLaunchKernelGGL(k_splitk_reduce, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, part,
                       splits, g.M, g.N, g0.C, g0.c_m);
    return check_launch("k_splitk_reduce");
}

// y[b] = h2[b] . w3 + b3 (one wave per row); optional dy = (y - t') / B with
// t' = (t - mean) / std from stats[1..2], and the per-row loss terms
// 0.5 (y - t')^2 / B into lterm (summed in a fixed order when the loss is wanted).
__global__ void k_mlp_out(const float *H2, const float *w3, const float *b3, int n, float *y,
                          const float *t, float *dy, const float *stats, float *lterm) {
    const int row = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    float acc = 0.f;
    for (int j = lane; j < kHidden; j += 64) acc = fmaf(H2[(long)row * kHidden + j], w3[j], acc);
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) {
        const float yy = acc + b3[0];
        y[row] = yy;
        if (t) {
            const float tt = (t[row] - stats[1]) / stats[2];
            dy[row] = (yy - tt) / (float)n;
            lterm[row] = 0.5f * (yy - tt) * (yy - tt) / (float)n;
        }
    }
}

// dH2[b][j] = dy[b] w3[j] silu'(A2[b][j])
__global__ void k_mlp_dh2(const float *dy, const float *w3, const float *A2, int n, float *dH2) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * kHidden) return;
    const int b = (int)(i / kHidden), j = (int)(i % kHidden);
    dH2[i] = dy[b] * w3[j] * dsiluf_(A2[i]);
}

// out[j] = sum_b M[b][j] (* w[b] if w): 64 columns per block, 16 row groups
// (rows g, g + 16, ...) summed in row order, then the groups in order.
__global__ __launch_bounds__(1024) void k_colsum(const float *Mx, const float *w, int n, int cols,
                                                float *out) {
    __shared__ float red[16][64];
    const int c = blockIdx.x * 64 + (threadIdx.x & 63), grp = threadIdx.x >> 6;
    float acc = 0.f;
    if (c < cols) {
        // rows grp, grp + 16, ... in order; 32 of them loaded before they are added
        int b = grp;
        for (; b + 16 * 31 < n; b += 16 * 32) {
            float x[32], y[32];
#pragma unroll
            for (int u = 0; u < 32; ++u) {
                x[u] = Mx[(long)(b + 16 * u) * cols + c];
                y[u] = w ? w[b + 16 * u] : 1.0f;
            }
#pragma unroll
            for (int u = 0; u < 32; ++u) acc = fmaf(x[u], y[u], acc);
        }
        for (; b < n; b += 16) acc = fmaf(Mx[(long)b * cols + c], w ? w[b] : 1.0f, acc);
    }
    red[grp][threadIdx.x & 63] = acc;
    __syncthreads();
    if (grp == 0 && c < cols) {
        float v = 0.f;
        for (int q = 0; q < 16; ++q) v += red[q][threadIdx.x];
        out[c] = v;
    }
}

static int colsum(const float *Mx, const float *w, int n, int cols, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_colsum, dim3((cols + 63) / 64), dim3(1024), 0, s, Mx, w, n, cols, out);
    return check_launch("k_colsum");
}

__global__ void k_sum(const float *v, int n, float *out) {
    __shared__ float sh[256];
    float acc = 0.f;
    for (int i = threadIdx.x; i < n; i += 256) acc += v[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

// Adam with bias correction (Kingma & Ba, 2015) on parameter i with gradient
// g; c1 = 1 - 0.9^t, c2 = 1 - 0.999^t.  Returns the new value.  Shared by
// k_adam and the fused training step, so both give the same bits.
__device__ __forceinline__ float adam_update(float *P, float *m, float *v, size_t i, float g,
                                             float lr, float c1, float c2) {
    const float mi = 0.9f * m[i] + 0.1f * g;
    const float vi = 0.999f * v[i] + 0.001f * g * g;
    m[i] = mi;
    v[i] = vi;
    const float p = P[i] - lr * (mi / c1) / (sqrtf(vi / c2) + 1e-8f);
    P[i] = p;
    return p;
}

// Bias corrections of Adam step t (1-based).
__device__ __forceinline__ void adam_corrections(int t, float &c1, float &c2) {
    c1 = 1.0f - powf(0.9f, (float)t);
    c2 = 1.0f - powf(0.999f, (float)t);
}

// Step *t_dev + 1 on every parameter.
__global__ void k_adam(float *P, const float *G, float *m, float *v, size_t n, float lr,
                       const int *t_dev) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float c1, c2;
    adam_corrections(*t_dev + 1, c1, c2);
    (void)adam_update(P, m, v, i, G[i], lr, c1, c2);
}

__global__ void k_add_i32(int *p, int v) { *p += v; }

int add_i32(int *p, int v, hipStream_t s) {
    hipLaunchKernelGGL(k_add_i32, dim3(1), dim3(1), 0, s, p, v);
    return check_launch("k_add_i32");
}

// He-uniform init from Philox: W1 ~ U(+-sqrt(6/D)), W2 ~ U(+-sqrt(6/256)),
// w3 ~ 0.1 U(+-sqrt(6/256)), biases 0; W1 padding columns 0.
__global__ void k_mlp_init(float *P, MlpLayout L, uint64_t seed) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L.n) return;
    const U4 w = philox_block(seed, (uint32_t)(i >> 2), 0u, 0u, kPurposeSurrogate);
    const float u = 2.0f * u01(word(w, (uint32_t)(i & 3))) - 1.0f;
    float v = 0.f;
    if (i < L.ob1) {
        const int col = (int)(i % L.Kp);
        v = col < L.D ? u * sqrtf(6.0f / (float)L.D) : 0.f;
    } else if (i >= L.oW2 && i < L.ob2) {
        v = u * sqrtf(6.0f / (float)kHidden);
    } else if (i >= L.ow3 && i < L.ob3) {
        v = 0.1f * u * sqrtf(6.0f / (float)kHidden);
    }
    P[i] = v;
}

__global__ void k_zero(float *p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.f;
}

static int zero(float *p, size_t n, hipStream_t s) {
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, s, p, n);
    return check_launch("k_zero");
}

// ------------------------------------------------------------------ Mlp

Mlp::~Mlp() {
    void *b[] = {P,   G,   m,     v,   A1,  H1,   A2,   H2,   dH1,   dH2,  y,     dy,
                 stats, W1b, W2b, part, lterm, W2Tb, XT, H1T, dH2T, dH1T, tpart, t_dev};
    for (void *p : b) (void)hipFree(p);
}

int Mlp::alloc(int dim, int batch) {
    L.set(dim);
    max_batch = batch;
    for (float **p : {&P, &G, &m, &v}) SPEO_TRY_HIP(hipMalloc(p, sizeof(float) * L.n));
    for (float **p : {&A1, &H1, &A2, &H2, &dH1, &dH2})
        SPEO_TRY_HIP(hipMalloc(p, sizeof(float) * (size_t)batch * kHidden));
    SPEO_TRY_HIP(hipMalloc(&y, sizeof(float) * batch));
    SPEO_TRY_HIP(hipMalloc(&dy, sizeof(float) * batch));
    SPEO_TRY_HIP(hipMalloc(&lterm, sizeof(float) * batch));
    SPEO_TRY_HIP(hipMalloc(&stats, sizeof(float) * 8));
    SPEO_TRY_HIP(hipMalloc(&t_dev, sizeof(int)));
    SPEO_TRY_HIP(hipMemset(t_dev, 0, sizeof(int)));
    SPEO_TRY_HIP(hipMemset(G, 0, sizeof(float) * L.n));
    SPEO_TRY_HIP(hipMemset(stats, 0, sizeof(float) * 8));
    const int splits = std::min(kMaxSplitK, (batch + 127) / 128);
    SPEO_TRY_HIP(hipMalloc(&part, sizeof(float) * (size_t)std::max(splits, 1) * kHidden *
                                      (size_t)std::max(kHidden, L.Kp)));
    if (L.Kp <= kFusedMaxKp) {
        Bp = round_up(batch, kTrainRows);
        SPEO_TRY_HIP(hipMalloc(&XT, sizeof(uint16_t) * (size_t)L.Kp * Bp));
        for (uint16_t **p : {&H1T, &dH2T, &dH1T})
            SPEO_TRY_HIP(hipMalloc(p, sizeof(uint16_t) * (size_t)kHidden * Bp));
        SPEO_TRY_HIP(hipMalloc(&tpart, sizeof(float) * (size_t)(Bp / kTrainRows) * kTrainPart));
    }
    return SPEO_OK;
}

int Mlp::init(uint64_t seed, hipStream_t s) {
    bf16_stale = true;
    hipLaunchKernelGGL(k_mlp_init, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, P, L, seed);
    SPEO_TRY(check_launch("k_mlp_init"));
    SPEO_TRY_HIP(hipMemsetAsync(m, 0, sizeof(float) * L.n, s));
    SPEO_TRY_HIP(hipMemsetAsync(v, 0, sizeof(float) * L.n, s));
    const float st[8] = {0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    SPEO_TRY_HIP(hipMemcpyAsync(stats, st, sizeof(st), hipMemcpyHostToDevice, s));
    SPEO_TRY_HIP(hipStreamSynchronize(s));
    SPEO_TRY_HIP(hipMemset(t_dev, 0, sizeof(int)));
    return SPEO_OK;
}

static GemmArgs layer1(const Mlp &N, const float *X, int ld, int n) {
    GemmArgs g{};
    g.M = n; g.N = kHidden; g.K = N.L.D;
    g.A = X; g.a_m = ld; g.a_k = 1; g.a_scale = kXScale;
    g.B = N.P + N.L.oW1; g.b_k = 1; g.b_n = N.L.Kp; g.b_scale = 1.f;
    g.C = N.H1; g.c_m = kHidden; g.bias = N.P + N.L.ob1; g.P = N.A1; g.p_m = kHidden;
    g.epi = kEpiBiasSilu;
    return g;
}

static GemmArgs layer2(const Mlp &N, int n) {
    GemmArgs g{};
    g.M = n; g.N = kHidden; g.K = kHidden;
    g.A = N.H1; g.a_m = kHidden; g.a_k = 1; g.a_scale = 1.f;
    g.B = N.P + N.L.oW2; g.b_k = 1; g.b_n = kHidden; g.b_scale = 1.f;
    g.C = N.H2; g.c_m = kHidden; g.bias = N.P + N.L.ob2; g.P = N.A2; g.p_m = kHidden;
    g.epi = kEpiBiasSilu;
    return g;
}

int Mlp::forward(const float *X, int ld, int n, float *pred, hipStream_t s) {
    for (int r0 = 0; r0 < n; r0 += max_batch) {
        const int nb = std::min(max_batch, n - r0);
        SPEO_TRY(gemm(layer1(*this, X + (size_t)r0 * ld, ld, nb), s));
        SPEO_TRY(gemm(layer2(*this, nb), s));
        hipLaunchKernelGGL(k_mlp_out, dim3((nb + 3) / 4), dim3(256), 0, s, H2, P + L.ow3,
                           P + L.ob3, nb, pred + r0, nullptr, nullptr, nullptr, nullptr);
        SPEO_TRY(check_launch("k_mlp_out"));
    }
    return SPEO_OK;
}

int Mlp::loss_grad_unfused(const float *X, int ld, int n, const float *t, hipStream_t s,
                           bool want_loss) {
    if (n > max_batch) return set_error(SPEO_ERR_ARG, "mlp: batch larger than the workspace");
    SPEO_TRY(gemm(layer1(*this, X, ld, n), s));
    SPEO_TRY(gemm(layer2(*this, n), s));
    hipLaunchKernelGGL(k_mlp_out, dim3((n + 3) / 4), dim3(256), 0, s, H2, P + L.ow3, P + L.ob3, n,
                       y, t, dy, stats, lterm);
    SPEO_TRY(check_launch("k_mlp_out"));
    if (want_loss) {
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, lterm, n, stats);
        SPEO_TRY(check_launch("k_sum"));
    }
    hipLaunchKernelGGL(k_mlp_dh2, dim3((unsigned)(((size_t)n * kHidden + 255) / 256)), dim3(256),
                       0, s, dy, P + L.ow3, A2, n, dH2);
    SPEO_TRY(check_launch("k_mlp_dh2"));
    // g_w3 = H2^T dy, g_b3 = sum dy, g_b2 = colsum dH2
    SPEO_TRY(colsum(H2, dy, n, kHidden, G + L.ow3, s));
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, s, dy, n, G + L.ob3);
    SPEO_TRY(check_launch("k_sum"));
    SPEO_TRY(colsum(dH2, nullptr, n, kHidden, G + L.ob2, s));
    // g_W2 = dH2^T H1
    GemmArgs g{};
    g.M = kHidden; g.N = kHidden; g.K = n;
    g.A = dH2; g.a_m = 1; g.a_k = kHidden; g.a_scale = 1.f;
    g.B = H1; g.b_k = kHidden; g.b_n = 1; g.b_scale = 1.f;
    g.C = G + L.oW2; g.c_m = kHidden; g.epi = kEpiStore;
    SPEO_TRY(gemm_splitk(g, part, s));
    // dH1 = (dH2 W2) * silu'(A1)
    g = GemmArgs{};
    g.M = n; g.N = kHidden; g.K = kHidden;
    g.A = dH2; g.a_m = kHidden; g.a_k = 1; g.a_scale = 1.f;
    g.B = P + L.oW2; g.b_k = kHidden; g.b_n = 1; g.b_scale = 1.f;
    g.C = dH1; g.c_m = kHidden; g.P = A1; g.p_m = kHidden; g.epi = kEpiMulDsilu;
    SPEO_TRY(gemm(g, s));
    SPEO_TRY(colsum(dH1, nullptr, n, kHidden, G + L.ob1, s));
    // g_W1 = dH1^T (X / 100); columns >= D are never written (zero since alloc)
    g = GemmArgs{};
    g.M = kHidden; g.N = L.D; g.K = n;
    g.A = dH1; g.a_m = 1; g.a_k = kHidden; g.a_scale = 1.f;
    g.B = X; g.b_k = ld; g.b_n = 1; g.b_scale = kXScale;
    g.C = G + L.oW1; g.c_m = L.Kp; g.epi = kEpiStore;
    return gemm_splitk(g, part, s);
}

int Mlp::adam(float lr, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((L.n + 255) / 256)), dim3(256), 0, s, P, G, m, v,
                       L.n, lr, t_dev);
    bf16_stale = true;
    SPEO_TRY(check_launch("k_adam"));
    return add_i32(t_dev, 1, s);
}

// ------------------------------------------------------ fused inference

// BF16 weight copies are stored in MFMA B-fragment order: for B(k, n) of a
// layer (n = output neuron in 16-column tiles t, k in steps of 32), lane
// l = n % 16 + 16 ((k % 32) / 8) of tile t, step ks holds k % 8 = 0..7 at
// ((t KS + ks) 64 + l) 8, so one wave's fragment is 1 KB contiguous.
__host__ __device__ __forceinline__ size_t frag_index(int n, int k, int KS) {
    return ((size_t)((n >> 4) * KS + (k >> 5)) * 64 + (n & 15) + 16 * ((k & 31) >> 3)) * 8 + (k & 7);
}

// dst = BF16 of a [256][K] row-major matrix as B(k, n) = src[n][k]
// (transposed: B(k, n) = src[k][n], K = 256), in fragment order.
__global__ void k_to_frag(const float *src, uint16_t *dst, int K, int transposed) {
    const int n = blockIdx.x, KS = K / 32;
    for (int k = threadIdx.x; k < K; k += blockDim.x) {
        const float v = transposed ? src[(size_t)k * kHidden + n] : src[(size_t)n * K + k];
        dst[frag_index(n, k, KS)] = to_bits(v, false);
    }
}

// The network on a tile of 64 rows whose scaled BF16 inputs are in Xs
// ([64][KP + 8]), 16 waves: wave w owns hidden columns [16w, 16w + 16) of both
// layers for all 64 rows (4 MFMA 16x16 tiles); its weight fragments (W1 and
// W2, frag_index order, 16 bytes per lane) come from mlp_prefetch.  H1 (BF16)
// in Hs; the output layer reduced across lanes and waves in
// LDS.  pred[r] for rows r < nrows of the tile.  BF16 operands, FP32
// accumulation, FP32 SiLU and output layer.
template <int KP> struct MlpFrags {
    bf16x8 w1[KP / 32], w2[kHidden / 32];
    float bias1, bias2, wout, bias3;
};

// A wave's weight fragments and column parameters for mlp_tile64 (issued
// early, so that their loads overlap whatever the block does meanwhile).
template <int KP>
__device__ __forceinline__ void mlp_prefetch(MlpFrags<KP> &f, const uint16_t *W1b, const float *b1,
                                             const uint16_t *W2b, const float *b2, const float *w3,
                                             const float *b3) {
    constexpr int K1 = KP / 32, K2 = kHidden / 32;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, col = wave * 16 + (lane & 15);
#pragma unroll
    for (int ks = 0; ks < K1; ++ks)
        f.w1[ks] = *reinterpret_cast<const bf16x8 *>(&W1b[((size_t)(wave * K1 + ks) * 64 + lane) * 8]);
#pragma unroll
    for (int ks = 0; ks < K2; ++ks)
        f.w2[ks] = *reinterpret_cast<const bf16x8 *>(&W2b[((size_t)(wave * K2 + ks) * 64 + lane) * 8]);
    f.bias1 = b1[col];
    f.bias2 = b2[col];
    f.wout = w3[col];
    f.bias3 = b3[0];
}

template <int KP>
__device__ __forceinline__ void mlp_tile64(const uint16_t *Xs, uint16_t *Hs, float (*red)[64],
                                           const MlpFrags<KP> &f, int nrows, float *pred) {
    constexpr int TM = 64, LX = KP + 8, LH = kHidden + 8, K1 = KP / 32, K2 = kHidden / 32;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int col = wave * 16 + lr;
    const bf16x8 *w1 = f.w1, *w2 = f.w2;
    const float bias1 = f.bias1, bias2 = f.bias2, wout = f.wout, bias3 = f.bias3;
    floatx4 acc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) acc[i] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K1; ++ks) {
        const int kb = ks * 32 + lk * 8;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                *reinterpret_cast<const bf16x8 *>(&Xs[(i * 16 + lr) * LX + kb]), w1[ks], acc[i], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Hs[(i * 16 + lk * 4 + q) * LH + col] = to_bits(siluf_(acc[i][q] + bias1), false);
            acc[i][q] = 0.f;
        }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < K2; ++ks) {
        const int kb = ks * 32 + lk * 8;
#pragma unroll
        for (int i = 0; i < 4; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                *reinterpret_cast<const bf16x8 *>(&Hs[(i * 16 + lr) * LH + kb]), w2[ks], acc[i], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float v = sum16(siluf_(acc[i][q] + bias2) * wout);
            if (lr == 0) red[wave][i * 16 + lk * 4 + q] = v;
        }
    __syncthreads();
    if (tid < TM && tid < nrows) {
        float y = 0.f;
#pragma unroll
        for (int w = 0; w < kHidden / 16; ++w) y += red[w][tid];
        pred[tid] = y + bias3;
    }
}


// 64 rows of X per block (1024 threads): the rows scaled and rounded to
// BF16 in LDS, then mlp_tile64.
template <int KP>
__global__ __launch_bounds__(1024) void k_mlp_infer(const float *X, int ld, int n, int D,
                                                    const uint16_t *W1b, const float *b1,
                                                    const uint16_t *W2b, const float *b2,
                                                    const float *w3, const float *b3,
                                                    float *pred) {
    constexpr int TM = 64, LX = KP + 8, LH = kHidden + 8;
    __shared__ __align__(16) uint16_t Xs[TM * LX];
    __shared__ __align__(16) uint16_t Hs[TM * LH];
    __shared__ float red[kHidden / 16][TM];
    const int tid = threadIdx.x;
    const int row0 = blockIdx.x * TM;
    MlpFrags<KP> f;
    mlp_prefetch<KP>(f, W1b, b1, W2b, b2, w3, b3);
    typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
    if ((ld & 3) == 0 && ld >= KP && (reinterpret_cast<uintptr_t>(X) & 15) == 0) {
        // 16-byte loads, all of the tile issued before the first is converted
        constexpr int G = (TM * KP / 4 + 1023) / 1024;
        floatx4 x[G];
#pragma unroll
        for (int i = 0; i < G; ++i) {
            const int e = tid + i * 1024, r = e / (KP / 4), k = (e % (KP / 4)) * 4;
            x[i] = e < TM * KP / 4 && row0 + r < n
                       ? gld(reinterpret_cast<const floatx4 *>(X + (size_t)(row0 + r) * ld + k))
                       : floatx4{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
        for (int i = 0; i < G; ++i) {
            const int e = tid + i * 1024, r = e / (KP / 4), k = (e % (KP / 4)) * 4;
            if (e >= TM * KP / 4) continue;
            u16x4 h;
#pragma unroll
            for (int q = 0; q < 4; ++q) h[q] = to_bits(k + q < D ? x[i][q] * kXScale : 0.f, false);
            *reinterpret_cast<u16x4 *>(&Xs[r * LX + k]) = h;
        }
    } else {
        for (int e = tid; e < TM * KP; e += 1024) {
            const int r = e / KP, k = e % KP;
            float v = 0.f;
            if (row0 + r < n && k < D) v = X[(size_t)(row0 + r) * ld + k] * kXScale;
            Xs[r * LX + k] = to_bits(v, false);
        }
    }
    __syncthreads();
    mlp_tile64<KP>(Xs, Hs, red, f, n - row0, pred + row0);
}

int Mlp::refresh_bf16(hipStream_t s) {
    if (!W1b) {
        SPEO_TRY_HIP(hipMalloc(&W1b, sizeof(uint16_t) * (size_t)kHidden * L.Kp));
        SPEO_TRY_HIP(hipMalloc(&W2b, sizeof(uint16_t) * (size_t)kHidden * kHidden));
        SPEO_TRY_HIP(hipMalloc(&W2Tb, sizeof(uint16_t) * (size_t)kHidden * kHidden));
    }
    hipLaunchKernelGGL(k_to_frag, dim3(kHidden), dim3(128), 0, s, P + L.oW1, W1b, L.Kp, 0);
    hipLaunchKernelGGL(k_to_frag, dim3(kHidden), dim3(128), 0, s, P + L.oW2, W2b, kHidden, 0);
    hipLaunchKernelGGL(k_to_frag, dim3(kHidden), dim3(128), 0, s, P + L.oW2, W2Tb, kHidden, 1);
    bf16_stale = false;
    return check_launch("k_to_frag");
}

int Mlp::infer(const float *X, int ld, int n, float *pred, hipStream_t s) {
    if (n <= 0) return SPEO_OK;
    if (bf16_stale || !W1b) SPEO_TRY(refresh_bf16(s));
    const dim3 grid((n + 63) / 64), block(1024);
    const float *b1 = P + L.ob1, *b2 = P + L.ob2, *w3 = P + L.ow3, *b3 = P + L.ob3;
    switch (L.Kp) {
    case 32: hipLaunchKernelGGL(k_mlp_infer<32>, grid, block, 0, s, X, ld, n, L.D, W1b, b1, W2b, b2, w3, b3, pred); break;
    case 64: hipLaunchKernelGGL(k_mlp_infer<64>, grid, block, 0, s, X, ld, n, L.D, W1b, b1, W2b, b2, w3, b3, pred); break;
    case 96: hipLaunchKernelGGL(k_mlp_infer<96>, grid, block, 0, s, X, ld, n, L.D, W1b, b1, W2b, b2, w3, b3, pred); break;
    case 128: hipLaunchKernelGGL(k_mlp_infer<128>, grid, block, 0, s, X, ld, n, L.D, W1b, b1, W2b, b2, w3, b3, pred); break;
    default: return forward(X, ld, n, pred, s);  // wide inputs: the unfused path
    }
    return check_launch("k_mlp_infer");
}

// ------------------------------------------------------ fused training step
//
// One Adam step in two launches instead of a dozen: k_train_rows runs the
// forward and backward passes of 32 minibatch rows per workgroup up to the
// activation gradients, and k_train_update forms the weight gradients as
// GEMMs over the whole minibatch and applies Adam to them.  Same rounding
// points as the unfused path (BF16 operands, FP32 accumulation, FP32 SiLU,
// output layer, loss and Adam); the sums run in other fixed orders.

typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

struct TrainRowsArgs {
    const float *X;        // direct rows (stride ld) and targets, or null:
    int ld;                //   then minibatch slots of the archive AX / Ay
    const float *t;
    const float *AX, *Ay;  // archive rows [slot][Kp], labels [slot]
    int n_valid;
    uint64_t seed;
    uint32_t round, step;  // the minibatch draw's counters (round: *round_dev when set)
    const uint32_t *round_dev;
    int n, D;
    const float *stats;    // [1] mean, [2] std of the labels
    const uint16_t *W1b, *W2b, *W2Tb;
    const float *b1, *b2, *w3, *b3;
    uint16_t *XT, *H1T, *dH2T, *dH1T;  // [rows][Bp] in fragment order: k = minibatch row
    int Bp;
    float *part;           // [tile][kTrainPart]: db1 | db2 | dw3 | (db3, loss)
};

// Rows [32 w, 32 w + 32) of the minibatch, 16 waves: wave v owns hidden
// columns [16 v, 16 v + 16) of every layer for the 32 rows (two MFMA 16x16
// tiles).  A wave's weight fragments (W1, W2 and W2^T for its columns, 20
// 16-byte loads per lane) are all issued at the start, so the kernel waits
// on one L2 round trip for them instead of one per K step.  Rows >= n are
// zero inputs with dy = 0, so they add nothing to any gradient.
// Networks trained in one launch (one per problem), each a run of blocks.
constexpr int kMaxTrainNets = 8;
struct TrainRowsBatch {
    TrainRowsArgs net[kMaxTrainNets];
    int tiles;  // row tiles per network
};

template <int KP>
__global__ __launch_bounds__(1024) void k_train_rows(TrainRowsBatch batch) {
    const TrainRowsArgs &a = batch.net[blockIdx.x / batch.tiles];
    constexpr int TM = kTrainRows, LX = KP + 8, LH = kHidden + 8, NW = kHidden / 16;
    constexpr int K1 = KP / 32, K2 = kHidden / 32;
    __shared__ __align__(16) uint16_t Xs[TM * LX];
    __shared__ __align__(16) uint16_t Hs[TM * LH];  // H1
    __shared__ __align__(16) uint16_t Gs[TM * LH];  // dH2
    __shared__ __align__(16) uint16_t Ds[TM * LH];  // dH1
    __shared__ float red[NW][TM];
    __shared__ float ts[TM], dys[TM], lts[TM];
    __shared__ int slots[TM];
    const int tile = blockIdx.x % batch.tiles;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int row0 = tile * TM;
    const int col = wave * 16 + lr;  // this lane's hidden column in every layer
    float *part = a.part + (size_t)tile * kTrainPart;
    // weight fragments first: B(k, n = col) for K steps of 32 (8 per lane)
    bf16x8 w1[K1], w2[K2], w2t[K2];
#pragma unroll
    for (int ks = 0; ks < K1; ++ks)
        w1[ks] = *reinterpret_cast<const bf16x8 *>(&a.W1b[((size_t)(wave * K1 + ks) * 64 + lane) * 8]);
#pragma unroll
    for (int ks = 0; ks < K2; ++ks) {
        w2[ks] = *reinterpret_cast<const bf16x8 *>(&a.W2b[((size_t)(wave * K2 + ks) * 64 + lane) * 8]);
        w2t[ks] = *reinterpret_cast<const bf16x8 *>(&a.W2Tb[((size_t)(wave * K2 + ks) * 64 + lane) * 8]);
    }
    const float bias1 = a.b1[col], bias2 = a.b2[col], wout = a.w3[col], bias3 = a.b3[0];
    if (tid < TM) {
        const int b = row0 + tid;
        int slot = b;
        float t = 0.f;
        if (b < a.n) {
            if (a.AX) {  // the slot gather_batch draws for row b
                const uint32_t round = a.round_dev ? *a.round_dev : a.round;
                const U4 w = philox_block(a.seed, (uint32_t)(b >> 2), round, a.step, kPurposeSurrBatch);
                slot = (int)(u01(word(w, (uint32_t)(b & 3))) * (float)a.n_valid);
                slot = slot >= a.n_valid ? a.n_valid - 1 : slot;
                if (slot % 8 == 7) slot -= 1;
                t = a.Ay[slot];
            } else {
                t = a.t[b];
            }
            t = (t - a.stats[1]) / a.stats[2];
        }
        slots[tid] = slot;
        ts[tid] = t;
    }
    __syncthreads();
    {
        const float *src = a.AX ? a.AX : a.X;
        const size_t ld = a.AX ? (size_t)KP : (size_t)a.ld;
        constexpr int G = TM * KP / 1024;
        float x[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int e = tid + g * 1024, r = e / KP, k = e % KP;
            x[g] = (row0 + r < a.n && k < a.D) ? src[(size_t)slots[r] * ld + k] : 0.f;
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int e = tid + g * 1024, r = e / KP, k = e % KP;
            Xs[r * LX + k] = to_bits(x[g] * kXScale, false);
        }
    }
    __syncthreads();
    // transposed copy of a [TM][stride] BF16 tile: element (c, row0 + r) of
    // T, a [ncols][Bp] matrix as B-fragment operand (k = minibatch row), in
    // fragment order (frag_index), 8 consecutive rows per 16-byte store
    auto store_t = [&](const uint16_t *S, int stride, int ncols, uint16_t *T) __attribute__((always_inline)) {
        for (int e = tid; e < ncols * (TM / 8); e += 1024) {
            const int c = e / (TM / 8), r8 = (e % (TM / 8)) * 8;
            u16x8 v;
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = S[(r8 + u) * stride + c];
            *reinterpret_cast<u16x8 *>(T + frag_index(c, row0 + r8, a.Bp / 32)) = v;
        }
    };
    floatx4 acc[2];
    // layer 1
    acc[0] = acc[1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K1; ++ks) {
        const int kb = ks * 32 + lk * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                *reinterpret_cast<const bf16x8 *>(&Xs[(i * 16 + lr) * LX + kb]), w1[ks], acc[i], 0, 0, 0);
    }
    float d1[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float pre = acc[i][q] + bias1;
            d1[i][q] = dsiluf_(pre);
            Hs[(i * 16 + lk * 4 + q) * LH + col] = to_bits(siluf_(pre), false);
        }
    __syncthreads();
    // layer 2
    acc[0] = acc[1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K2; ++ks) {
        const int kb = ks * 32 + lk * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                *reinterpret_cast<const bf16x8 *>(&Hs[(i * 16 + lr) * LH + kb]), w2[ks], acc[i], 0, 0, 0);
    }
    float d2[2][4], h2[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float pre = acc[i][q] + bias2, h = siluf_(pre);
            d2[i][q] = dsiluf_(pre);
            h2[i][q] = h;
            const float v = sum16(h * wout);
            if (lr == 0) red[wave][i * 16 + lk * 4 + q] = v;
        }
    __syncthreads();
    if (tid < TM) {
        float dy = 0.f, lt = 0.f;
        if (row0 + tid < a.n) {
            float y = 0.f;
#pragma unroll
            for (int w = 0; w < NW; ++w) y += red[w][tid];
            y += bias3;
            const float e = y - ts[tid];
            dy = e / (float)a.n;
            lt = 0.5f * e * e / (float)a.n;
        }
        dys[tid] = dy;
        lts[tid] = lt;
    }
    __syncthreads();
    // dH2 = dy w3 silu'(A2); partial sums of db2 = sum dH2 and dw3 = sum H2 dy
    // over the tile's rows (this lane's 8 rows, then the four lane groups)
    {
        float pb2 = 0.f, pw3 = 0.f;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int r = i * 16 + lk * 4 + q;
                const float dy = dys[r];
                const float g = dy * wout * d2[i][q];
                Gs[r * LH + col] = to_bits(g, false);
                pb2 += g;
                pw3 = fmaf(h2[i][q], dy, pw3);
            }
        pb2 += __shfl_xor(pb2, 16, 64);
        pb2 += __shfl_xor(pb2, 32, 64);
        pw3 += __shfl_xor(pw3, 16, 64);
        pw3 += __shfl_xor(pw3, 32, 64);
        if (lk == 0) {
            part[kHidden + col] = pb2;
            part[2 * kHidden + col] = pw3;
        }
    }
    if (wave == 0) {  // db3 and loss partials: the tile's rows by an xor butterfly
        float sdy = lane < TM ? dys[lane] : 0.f, sl = lane < TM ? lts[lane] : 0.f;
        for (int off = 1; off < 64; off <<= 1) {
            sdy += __shfl_xor(sdy, off, 64);
            sl += __shfl_xor(sl, off, 64);
        }
        if (lane == 0) {
            part[3 * kHidden] = sdy;
            part[3 * kHidden + 1] = sl;
        }
    }
    __syncthreads();
    // dH1 = (dH2 W2) silu'(A1)
    acc[0] = acc[1] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < K2; ++ks) {
        const int kb = ks * 32 + lk * 8;
#pragma unroll
        for (int i = 0; i < 2; ++i)
            acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                *reinterpret_cast<const bf16x8 *>(&Gs[(i * 16 + lr) * LH + kb]), w2t[ks], acc[i], 0, 0, 0);
    }
    {
        float pb1 = 0.f;
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float g = acc[i][q] * d1[i][q];
                Ds[(i * 16 + lk * 4 + q) * LH + col] = to_bits(g, false);
                pb1 += g;
            }
        pb1 += __shfl_xor(pb1, 16, 64);
        pb1 += __shfl_xor(pb1, 32, 64);
        if (lk == 0) part[col] = pb1;
    }
    __syncthreads();
    // the four transposed activation tiles, stored last so that no barrier
    // above waits for them
    store_t(Xs, LX, KP, a.XT);
    store_t(Hs, LH, kHidden, a.H1T);
    store_t(Gs, LH, kHidden, a.dH2T);
    store_t(Ds, LH, kHidden, a.dH1T);
}

struct TrainUpdateArgs {
    const uint16_t *XT, *H1T, *dH2T, *dH1T;
    int Bp, K;      // K = 32 * tiles: the minibatch rows the GEMMs sum over
    const float *part;
    int tiles;
    float *P, *G, *m, *v;  // G: gradients out (or null); Adam when adam != 0
    MlpLayout L;
    uint16_t *W1b, *W2b, *W2Tb;
    int adam;
    float lr;
    const int *t_dev;  // Adam step of this update: *t_dev + t_off + 1
    int t_off;
    float *stats;   // [0] = loss when want_loss
    int want_loss;
};

// Blocks [0, 64): 32 x 32 tiles of dW2 = dH2^T H1; then (256 / 32) (Kp / 32)
// tiles of dW1 = dH1^T (X / 100); the last block: b1, b2, w3, b3 from the
// row tiles' partial sums (in tile order).  A GEMM block's four waves take
// every fourth 32-row step of the minibatch, and their 32 x 32 results are
// added in wave order.  Each block then writes the gradient and/or applies
// Adam to its parameters and refreshes their BF16 copies.
struct TrainUpdateBatch {
    TrainUpdateArgs net[kMaxTrainNets];
    int blocks;  // blocks per network
};

__global__ __launch_bounds__(256) void k_train_update(TrainUpdateBatch batch) {
    const TrainUpdateArgs &a = batch.net[blockIdx.x / batch.blocks];
    float c1 = 1.f, c2 = 1.f;
    if (a.adam) adam_corrections(*a.t_dev + a.t_off + 1, c1, c2);
    __shared__ float red[4][32][33];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    constexpr int nW2 = (kHidden / 32) * (kHidden / 32);
    const int kt = a.L.Kp / 32, nW1 = (kHidden / 32) * kt;
    const int bid = blockIdx.x % batch.blocks;
    if (bid == nW2 + nW1) {
        // tile partials in tile order, eight tiles' loads in flight at a time
        const int c = tid;
        float g1 = 0.f, g2 = 0.f, g3 = 0.f;
        for (int w0 = 0; w0 < a.tiles; w0 += 8) {
            float v1[8], v2[8], v3[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const float *pw = a.part + (size_t)min(w0 + u, a.tiles - 1) * kTrainPart;
                v1[u] = pw[c];
                v2[u] = pw[kHidden + c];
                v3[u] = pw[2 * kHidden + c];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (w0 + u < a.tiles) {
                    g1 += v1[u];
                    g2 += v2[u];
                    g3 += v3[u];
                }
        }
        const size_t o1 = a.L.ob1 + c, o2 = a.L.ob2 + c, o3 = a.L.ow3 + c;
        if (a.G) {
            a.G[o1] = g1;
            a.G[o2] = g2;
            a.G[o3] = g3;
        }
        if (a.adam) {
            (void)adam_update(a.P, a.m, a.v, o1, g1, a.lr, c1, c2);
            (void)adam_update(a.P, a.m, a.v, o2, g2, a.lr, c1, c2);
            (void)adam_update(a.P, a.m, a.v, o3, g3, a.lr, c1, c2);
        }
        if (tid == 0) {
            float gb = 0.f, loss = 0.f;
            for (int w0 = 0; w0 < a.tiles; w0 += 8) {
                float vb[8], vl[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const float *pw = a.part + (size_t)min(w0 + u, a.tiles - 1) * kTrainPart;
                    vb[u] = pw[3 * kHidden];
                    vl[u] = pw[3 * kHidden + 1];
                }
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (w0 + u < a.tiles) {
                        gb += vb[u];
                        loss += vl[u];
                    }
            }
            if (a.G) a.G[a.L.ob3] = gb;
            if (a.adam) (void)adam_update(a.P, a.m, a.v, a.L.ob3, gb, a.lr, c1, c2);
            if (a.want_loss) a.stats[0] = loss;
        }
        return;
    }
    const bool w2 = bid < nW2;
    const int t = w2 ? bid : bid - nW2, ncolt = w2 ? kHidden / 32 : kt;
    const int i0 = (t / ncolt) * 32, j0 = (t % ncolt) * 32;
    const uint16_t *A = w2 ? a.dH2T : a.dH1T;
    const uint16_t *B = w2 ? a.H1T : a.XT;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    // wave w takes steps w, w + 4, ... of 32 rows; eight steps' fragments are
    // loaded before the first is used (steps past K read as zero)
    const int steps = a.K / 32, kbs = a.Bp / 32;
    const bf16x8 zero8 = __builtin_bit_cast(bf16x8, u16x8{0, 0, 0, 0, 0, 0, 0, 0});
    for (int base = wave; base < steps; base += 32) {
        bf16x8 av[8][2], bv[8][2];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int st = base + 4 * u, kb = st * 32 + lk * 8;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                av[u][i] = st < steps ? *reinterpret_cast<const bf16x8 *>(A + frag_index(i0 + i * 16 + lr, kb, kbs))
                                      : zero8;
                bv[u][i] = st < steps ? *reinterpret_cast<const bf16x8 *>(B + frag_index(j0 + i * 16 + lr, kb, kbs))
                                      : zero8;
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[u][i], bv[u][j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) red[wave][i * 16 + lk * 4 + q][j * 16 + lr] = acc[i][j][q];
    __syncthreads();
    const int mr = tid >> 3;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const int nc = (tid & 7) * 4 + e;
        const float g = ((red[0][mr][nc] + red[1][mr][nc]) + red[2][mr][nc]) + red[3][mr][nc];
        const int I = i0 + mr, J = j0 + nc;
        const size_t off = w2 ? a.L.oW2 + (size_t)I * kHidden + J : a.L.oW1 + (size_t)I * a.L.Kp + J;
        if (a.G) a.G[off] = g;
        if (a.adam) {
            const float p = adam_update(a.P, a.m, a.v, off, g, a.lr, c1, c2);
            const uint16_t h = to_bits(p, false);
            if (w2) {
                a.W2b[frag_index(I, J, kHidden / 32)] = h;
                a.W2Tb[frag_index(J, I, kHidden / 32)] = h;
            } else {
                a.W1b[frag_index(I, J, a.L.Kp / 32)] = h;
            }
        }
    }
}

// One fused step of nn networks with equal Kp and n: rows (X, ld, t) or
// archive slots (AX, Ay, ...) per network, gradients into G (grad) and/or
// Adam (lr > 0); two launches for all of them.
static int fused_steps(Mlp *const *nets, const TrainRowsArgs *r0, int nn, bool grad, float lr,
                       int t_off, bool want_loss, hipStream_t s) {
    if (nn < 1 || nn > kMaxTrainNets) return set_error(SPEO_ERR_ARG, "mlp: 1..8 networks per step");
    const int Kp = nets[0]->L.Kp, n = r0[0].n;
    const int tiles = (n + kTrainRows - 1) / kTrainRows;
    const int blocks = (kHidden / 32) * (kHidden / 32) + (kHidden / 32) * (Kp / 32) + 1;
    TrainRowsBatch rb{};
    TrainUpdateBatch ub{};
    rb.tiles = tiles;
    ub.blocks = blocks;
    for (int k = 0; k < nn; ++k) {
        Mlp &N = *nets[k];
        if (N.L.Kp != Kp || r0[k].n != n) return set_error(SPEO_ERR_ARG, "mlp: networks of one step differ");
        TrainRowsArgs &r = rb.net[k];
        r = r0[k];
        r.D = N.L.D;
        r.stats = N.stats;
        r.W1b = N.W1b;
        r.W2b = N.W2b;
        r.W2Tb = N.W2Tb;
        r.b1 = N.P + N.L.ob1;
        r.b2 = N.P + N.L.ob2;
        r.w3 = N.P + N.L.ow3;
        r.b3 = N.P + N.L.ob3;
        r.XT = N.XT;
        r.H1T = N.H1T;
        r.dH2T = N.dH2T;
        r.dH1T = N.dH1T;
        r.Bp = N.Bp;
        r.part = N.tpart;
        TrainUpdateArgs &u = ub.net[k];
        u.XT = N.XT;
        u.H1T = N.H1T;
        u.dH2T = N.dH2T;
        u.dH1T = N.dH1T;
        u.Bp = N.Bp;
        u.K = tiles * kTrainRows;
        u.part = N.tpart;
        u.tiles = tiles;
        u.P = N.P;
        u.G = grad ? N.G : nullptr;
        u.m = N.m;
        u.v = N.v;
        u.L = N.L;
        u.W1b = N.W1b;
        u.W2b = N.W2b;
        u.W2Tb = N.W2Tb;
        u.adam = lr > 0.f ? 1 : 0;
        u.lr = lr;
        u.t_dev = N.t_dev;
        u.t_off = t_off;
        u.stats = N.stats;
        u.want_loss = want_loss ? 1 : 0;
    }
    const dim3 grid(tiles * nn), block(1024);
    switch (Kp) {
    case 32: hipLaunchKernelGGL(k_train_rows<32>, grid, block, 0, s, rb); break;
    case 64: hipLaunchKernelGGL(k_train_rows<64>, grid, block, 0, s, rb); break;
    case 96: hipLaunchKernelGGL(k_train_rows<96>, grid, block, 0, s, rb); break;
    case 128: hipLaunchKernelGGL(k_train_rows<128>, grid, block, 0, s, rb); break;
    default: return set_error(SPEO_ERR_ARG, "mlp: fused step needs Kp <= 128");
    }
    SPEO_TRY(check_launch("k_train_rows"));
    hipLaunchKernelGGL(k_train_update, dim3(blocks * nn), dim3(256), 0, s, ub);
    return check_launch("k_train_update");
}

// One step of one network; with Adam (lr > 0) its step count advances by one.
static int fused_step(Mlp &N, const TrainRowsArgs &r0, bool grad, float lr, bool want_loss,
                      hipStream_t s) {
    Mlp *net = &N;
    SPEO_TRY(fused_steps(&net, &r0, 1, grad, lr, 0, want_loss, s));
    return lr > 0.f ? add_i32(N.t_dev, 1, s) : SPEO_OK;
}

int train_steps(Mlp *const *nets, const float *const *AX, const float *const *Ay,
                const int *n_valid, const uint64_t *seed, int nn, const uint32_t *round_dev,
                uint32_t step, int batch, float lr, bool want_loss, hipStream_t s) {
    for (int k0 = 0; k0 < nn; k0 += kMaxTrainNets) {
        const int m = std::min(kMaxTrainNets, nn - k0);
        TrainRowsArgs r[kMaxTrainNets] = {};
        for (int k = 0; k < m; ++k) {
            Mlp &N = *nets[k0 + k];
            if (batch > N.max_batch || N.L.Kp > kFusedMaxKp)
                return set_error(SPEO_ERR_ARG, "mlp: batch or width beyond the fused step");
            if (N.bf16_stale) SPEO_TRY(N.refresh_bf16(s));
            r[k].AX = AX[k0 + k];
            r[k].Ay = Ay[k0 + k];
            r[k].n_valid = n_valid[k0 + k];
            r[k].seed = seed[k0 + k];
            r[k].round_dev = round_dev;
            r[k].step = step;
            r[k].n = batch;
        }
        SPEO_TRY(fused_steps(nets + k0, r, m, false, lr, (int)step, want_loss, s));
    }
    return SPEO_OK;
}

int Mlp::loss_grad(const float *X, int ld, int n, const float *t, hipStream_t s, bool want_loss) {
    if (n > max_batch) return set_error(SPEO_ERR_ARG, "mlp: batch larger than the workspace");
    if (L.Kp > kFusedMaxKp) return loss_grad_unfused(X, ld, n, t, s, want_loss);
    if (bf16_stale) SPEO_TRY(refresh_bf16(s));
    TrainRowsArgs r{};
    r.X = X;
    r.ld = ld;
    r.t = t;
    r.n = n;
    return fused_step(*this, r, true, 0.f, want_loss, s);
}

int Mlp::train_step(const float *AX, const float *Ay, int n_valid, uint64_t seed, uint32_t round,
                    uint32_t step, int batch, float lr, bool want_loss, float *Xb, float *tb,
                    hipStream_t s) {
    if (batch > max_batch) return set_error(SPEO_ERR_ARG, "mlp: batch larger than the workspace");
    if (L.Kp > kFusedMaxKp) {  // wide inputs: gather, the unfused gradients, Adam
        SPEO_TRY(gather_batch(AX, Ay, n_valid, L.Kp, seed, round, step, batch, Xb, tb, s));
        SPEO_TRY(loss_grad_unfused(Xb, L.Kp, batch, tb, s, want_loss));
        return adam(lr, s);
    }
    if (bf16_stale) SPEO_TRY(refresh_bf16(s));
    TrainRowsArgs r{};
    r.AX = AX;
    r.Ay = Ay;
    r.n_valid = n_valid;
    r.seed = seed;
    r.round = round;
    r.step = step;
    r.n = batch;
    return fused_step(*this, r, false, lr, want_loss, s);
}

// ------------------------------------------------------------- statistics

// Spearman rank correlation of a[0..n) and b[0..n) (average ranks for ties),
// n <= 4096.  k_spearman_ranks: the whole array in LDS, 32 values per block,
// the values below and equal to value i counted by kRankLanes lanes, each
// over its own stretch of the array, and the integer counts added across
// those lanes (exact, so the split does not change them); it writes the
// centred ranks.  k_spearman_corr: one block sums their products in a fixed
// order.
constexpr int kRankLanes = 8;
__global__ __launch_bounds__(256) void k_spearman_ranks(const float *a, const float *b, int n,
                                                        float2 *ranks) {
    __shared__ float as[4096], bs[4096];
    for (int i = threadIdx.x; i < n; i += 256) {
        as[i] = a[i];
        bs[i] = b[i];
    }
    __syncthreads();
    const int i = (blockIdx.x * 256 + threadIdx.x) / kRankLanes, part = threadIdx.x % kRankLanes;
    const bool valid = i < n;
    const int span = (n + kRankLanes - 1) / kRankLanes, j0 = part * span, j1 = min(n, j0 + span);
    int la = 0, ea = 0, lb = 0, eb = 0;
    const float ai = valid ? as[i] : 0.f, bi = valid ? bs[i] : 0.f;
    for (int j = j0; j < j1; ++j) {
        const float aj = as[j], bj = bs[j];
        la += aj < ai;
        ea += aj == ai;
        lb += bj < bi;
        eb += bj == bi;
    }
#pragma unroll
    for (int off = 1; off < kRankLanes; off <<= 1) {
        la += __shfl_xor(la, off, 64);
        ea += __shfl_xor(ea, off, 64);
        lb += __shfl_xor(lb, off, 64);
        eb += __shfl_xor(eb, off, 64);
    }
    if (!valid || part != 0) return;
    const float mean = 0.5f * (float)(n - 1);  // ranks are multiples of 1/2: exact in FP32
    ranks[i] = float2{(float)la + 0.5f * (float)(ea - 1) - mean, (float)lb + 0.5f * (float)(eb - 1) - mean};
}

__global__ __launch_bounds__(1024) void k_spearman_corr(const float2 *ranks, int n, float *rho_out) {
    __shared__ double sh[3][1024];
    double sab = 0.0, saa = 0.0, sbb = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const double ra = ranks[i].x, rb = ranks[i].y;
        sab += ra * rb;
        saa += ra * ra;
        sbb += rb * rb;
    }
    sh[0][threadIdx.x] = sab;
    sh[1][threadIdx.x] = saa;
    sh[2][threadIdx.x] = sbb;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; ++q) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double den = sqrt(sh[1][0] * sh[2][0]);
        *rho_out = den > 0 ? (float)(sh[0][0] / den) : 0.f;
    }
}

int spearman(const float *a, const float *b, int n, float *rho_out, float2 *ranks, hipStream_t s) {
    if (n < 2 || n > 4096) return set_error(SPEO_ERR_ARG, "spearman: need 2 <= n <= 4096");
    hipLaunchKernelGGL(k_spearman_ranks, dim3((n * kRankLanes + 255) / 256), dim3(256), 0, s, a, b, n, ranks);
    SPEO_TRY(check_launch("k_spearman_ranks"));
    hipLaunchKernelGGL(k_spearman_corr, dim3(1), dim3(1024), 0, s, ranks, n, rho_out);
    return check_launch("k_spearman_corr");
}

// ------------------------------------------------------ candidate screening

__device__ __forceinline__ float4 ld4s(const float *p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ float pick4(const float4 &v, int q) {
    return q == 0 ? v.x : (q == 1 ? v.y : (q == 2 ? v.z : v.w));
}

// Candidate c of a target row: its donor rows, j_rand and crossover stream.
// Candidate 0 uses the plain DE streams (purposes 0 / 1), so it is exactly
// the plain engine's trial; candidate c >= 1 uses purposes 64 + 2c / 65 + 2c.
struct CandSrc {
    const float *p0, *p1, *p2, *own;
    int jr, dim;
    float F, CR;
    uint32_t pc, gr, gen;
};

__device__ __forceinline__ CandSrc cand_src(const IslandDesc &I, const SurrRowDesc &rd, int c,
                                            const float *X, uint64_t seed, uint32_t gen) {
    const uint32_t pd = c == 0 ? (uint32_t)kPurposeDonor : (uint32_t)(kPurposeCandidate + 2 * c);
    DonorStream ds(seed, (uint32_t)rd.gr, gen, pd);
    int r[kMaxR], jr;
    draw_donors(ds, nullptr, 0, I.np, I.dim, r, jr);
    const float *base = X + I.elem_off;
    CandSrc cs;
    cs.p0 = base + (size_t)r[0] * I.ld;
    cs.p1 = base + (size_t)r[1] * I.ld;
    cs.p2 = base + (size_t)r[2] * I.ld;
    cs.own = base + (size_t)rd.li * I.ld;
    cs.jr = jr;
    cs.dim = I.dim;
    cs.F = I.F;
    cs.CR = I.CR;
    cs.pc = c == 0 ? (uint32_t)kPurposeCross : (uint32_t)(kPurposeCandidate + 2 * c + 1);
    cs.gr = (uint32_t)rd.gr;
    cs.gen = gen;
    return cs;
}

// Coordinates 4g .. 4g + 3 of the candidate (0 past the dimension).
__device__ __forceinline__ float4 cand_group(const CandSrc &cs, int g, uint64_t seed, float lo,
                                             float hi) {
#pragma clang fp contract(off)
    if (4 * g >= cs.dim) return float4{0.f, 0.f, 0.f, 0.f};
    const U4 w = philox_block(seed, (uint32_t)g, cs.gr, cs.gen, cs.pc);
    const float4 e0 = ld4s(cs.p0 + 4 * g), e1 = ld4s(cs.p1 + 4 * g), e2 = ld4s(cs.p2 + 4 * g);
    const float4 xo = ld4s(cs.own + 4 * g);
    float v[4];
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) {
        const int j = 4 * g + qq;
        if (j < cs.dim) {
            const bool take = (j == cs.jr) || (u01(word(w, qq)) < cs.CR);
            const float m = take ? fmaf(cs.F, pick4(e1, qq) - pick4(e2, qq), pick4(e0, qq)) : pick4(xo, qq);
            v[qq] = check_bound(m, lo, hi);
        } else {
            v[qq] = 0.f;
        }
    }
    return float4{v[0], v[1], v[2], v[3]};
}

// 16 lanes per (target q, candidate c): the candidate's Kp columns.
__global__ __launch_bounds__(256) void k_gen_candidates(const IslandDesc *isl,
                                                        const SurrRowDesc *rows, int Q, int C,
                                                        const float *X, float lo, float hi,
                                                        uint64_t seed, const uint32_t *gen_base,
                                                        uint32_t gen_offset, float *cand, int Kp) {
    const int item = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (item >= Q * C) return;
    const int q = item / C, c = item % C;
    const SurrRowDesc rd = rows[q];
    const IslandDesc I = isl[rd.island];
    const uint32_t gen = (gen_base ? *gen_base : 0u) + gen_offset;
    const CandSrc cs = cand_src(I, rd, c, X, seed, gen);
    float *out = cand + (size_t)item * Kp;
    for (int g = l; g < (Kp >> 2); g += 16)
        *reinterpret_cast<float4 *>(out + 4 * g) = cand_group(cs, g, seed, lo, hi);
}

// The candidates of each problem's targets generated straight into LDS (64
// per tile, 16 lanes per candidate) and scored by that problem's network
// (mlp_tile64): the same predictions as k_gen_candidates followed by
// k_mlp_infer, without writing the candidates to memory.  All problems in one
// launch: problem k owns tiles [first[k], first[k + 1]).  One block per CU
// walks the tiles (b, b + grid, ...), so a wave loads its weight fragments
// once per problem instead of once per 64 candidates.  mlp_tile64 reads Xs
// before its first barrier and Hs before its second, and the next tile's
// candidates are written after both, so the tiles need no barrier between
// them.
template <int KP>
__global__ __launch_bounds__(1024) void k_cand_infer(CandInferBatch B, const IslandDesc *isl,
                                                     const SurrRowDesc *rows, int C, const float *X,
                                                     float lo, float hi, uint64_t seed,
                                                     const uint32_t *gen_base, uint32_t gen_offset,
                                                     float *pred) {
    constexpr int TM = 64, LX = KP + 8, LH = kHidden + 8;
    __shared__ __align__(16) uint16_t Xs[TM * LX];
    __shared__ __align__(16) uint16_t Hs[TM * LH];
    __shared__ float red[kHidden / 16][TM];
    const int r = threadIdx.x >> 4, l = threadIdx.x & 15;  // candidate r of the tile, lane l
    const uint32_t gen = (gen_base ? *gen_base : 0u) + gen_offset;
    typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));
    MlpFrags<KP> f;
    int kf = -1;  // the network whose fragments f holds
    for (int tile = blockIdx.x; tile < B.tiles; tile += gridDim.x) {
        int k = 0;
        while (k + 1 < B.nnets && tile >= B.first[k + 1]) ++k;
        // a problem below the rho threshold picks candidate 0 without scores
        // (every thread reads the same rho, so the whole block skips the tile)
        if (!(B.rho[B.prob[k]] >= B.rho_min)) continue;
        if (threadIdx.x == 0 && tile == B.first[k]) B.scored[B.prob[k]] += (long long)B.nq[k] * C;
        const int qoff = B.qoff[k], n = B.nq[k] * C;
        if (k != kf) {
            const float *P = B.P[k];
            mlp_prefetch<KP>(f, B.W1b[k], P + B.ob1, B.W2b[k], P + B.ob2, P + B.ow3, P + B.ob3);
            kf = k;
        }
        const int item0 = (tile - B.first[k]) * TM, item = item0 + r;
        if (item < n) {
            const SurrRowDesc rd = rows[qoff + item / C];
            const CandSrc cs = cand_src(isl[rd.island], rd, item % C, X, seed, gen);
            for (int g = l; g < KP / 4; g += 16) {
                const float4 v = cand_group(cs, g, seed, lo, hi);
                *reinterpret_cast<u16x4 *>(&Xs[r * LX + 4 * g]) =
                    u16x4{to_bits(v.x * kXScale, false), to_bits(v.y * kXScale, false),
                          to_bits(v.z * kXScale, false), to_bits(v.w * kXScale, false)};
            }
        } else {
            for (int g = l; g < KP / 4; g += 16)
                *reinterpret_cast<u16x4 *>(&Xs[r * LX + 4 * g]) = u16x4{0, 0, 0, 0};
        }
        __syncthreads();
        mlp_tile64<KP>(Xs, Hs, red, f, n - item0, pred + (size_t)qoff * C + item0);
    }
}

// One resident block of k_cand_infer per compute unit of the current device.
static int cand_infer_blocks() {
    static std::atomic<int> cached[16];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 16) return 256;
    int v = cached[dev].load(std::memory_order_relaxed);
    if (v == 0) {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        cached[dev].store(cus, std::memory_order_relaxed);
        v = cus;
    }
    return v;
}

int cand_infer(Mlp *const *nets, const int *qoff, const int *nq, const int *prob, int nn,
               const IslandDesc *isl, const SurrRowDesc *rows, int C, const float *X, float lo,
               float hi, uint64_t seed, const uint32_t *gen_base, uint32_t gen_offset,
               const float *rho, float rho_min, long long *scored, float *pred, hipStream_t s) {
    if (nn < 1 || nn > kMaxSurrProblems) return set_error(SPEO_ERR_ARG, "cand_infer: 1..16 problems");
    CandInferBatch B{};
    B.nnets = nn;
    int blocks = 0;
    const MlpLayout &L = nets[0]->L;
    if (L.Kp > kFusedMaxKp) return set_error(SPEO_ERR_ARG, "mlp: fused screening needs Kp <= 128");
    for (int k = 0; k < nn; ++k) {
        Mlp &N = *nets[k];
        if (N.L.Kp != L.Kp) return set_error(SPEO_ERR_ARG, "cand_infer: networks differ in width");
        if (N.bf16_stale || !N.W1b) SPEO_TRY(N.refresh_bf16(s));
        B.W1b[k] = N.W1b;
        B.W2b[k] = N.W2b;
        B.P[k] = N.P;
        B.qoff[k] = qoff[k];
        B.nq[k] = nq[k];
        B.prob[k] = prob[k];
        B.first[k] = blocks;
        blocks += (nq[k] * C + 63) / 64;
    }
    B.ob1 = L.ob1;
    B.ob2 = L.ob2;
    B.ow3 = L.ow3;
    B.ob3 = L.ob3;
    B.rho = rho;
    B.rho_min = rho_min;
    B.scored = scored;
    if (blocks == 0) return SPEO_OK;
    B.tiles = blocks;
    const dim3 grid(std::min(blocks, cand_infer_blocks())), block(1024);
#define SPEO_CAND_INFER(KP)                                                                          \
    hipLaunchKernelGGL(k_cand_infer<KP>, grid, block, 0, s, B, isl, rows, C, X, lo, hi, seed, gen_base, \
                       gen_offset, pred)
    switch (L.Kp) {
    case 32: SPEO_CAND_INFER(32); break;
    case 64: SPEO_CAND_INFER(64); break;
    case 96: SPEO_CAND_INFER(96); break;
    default: SPEO_CAND_INFER(128); break;
    }
#undef SPEO_CAND_INFER
    return check_launch("k_cand_infer");
}

int gen_candidates(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C, const float *X,
                   float lo, float hi, uint64_t seed, const uint32_t *gen_base,
                   uint32_t gen_offset, float *cand, int Kp, hipStream_t s) {
    const int items = Q * C;
    if (items <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_gen_candidates, dim3((items + 15) / 16), dim3(256), 0, s, isl, rows, Q,
                       C, X, lo, hi, seed, gen_base, gen_offset, cand, Kp);
    return check_launch("k_gen_candidates");
}

// 16 lanes per target: the candidate with the lowest prediction (ties to the
// lower index, NaN predictions skipped) while the problem's surrogate is
// active, else candidate 0; copied into the target's trial row of U.
__global__ void k_pick_candidates(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C,
                                  const float *cand, int Kp, const float *pred, const float *rho,
                                  float rho_min, float *U) {
    const int q = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (q >= Q) return;
    const SurrRowDesc rd = rows[q];
    const IslandDesc I = isl[rd.island];
    int best = 0;
    if (rho[rd.problem] >= rho_min) {
        float bv = pred[(size_t)q * C];
        for (int c = 1; c < C; ++c) {
            const float v = pred[(size_t)q * C + c];
            if (v < bv || (isnan(bv) && !isnan(v))) {
                bv = v;
                best = c;
            }
        }
    }
    const float *src = cand + ((size_t)q * C + best) * Kp;
    float *dst = U + I.elem_off + (size_t)rd.li * I.ld;
    for (int g = l; g < (I.ld >> 2); g += 16)
        *reinterpret_cast<float4 *>(dst + 4 * g) = *reinterpret_cast<const float4 *>(src + 4 * g);
}

// The same pick, the chosen candidate generated again into U instead of
// copied (its coordinates are a function of the counters and X only).
__global__ void k_pick_regen(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C,
                             const float *pred, const float *rho, float rho_min, const float *X,
                             float lo, float hi, uint64_t seed, const uint32_t *gen_base,
                             uint32_t gen_offset, float *U) {
    const int q = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (q >= Q) return;
    const SurrRowDesc rd = rows[q];
    const IslandDesc I = isl[rd.island];
    int best = 0;
    if (rho[rd.problem] >= rho_min) {
        float bv = pred[(size_t)q * C];
        for (int c = 1; c < C; ++c) {
            const float v = pred[(size_t)q * C + c];
            if (v < bv || (isnan(bv) && !isnan(v))) {
                bv = v;
                best = c;
            }
        }
    }
    const uint32_t gen = (gen_base ? *gen_base : 0u) + gen_offset;
    const CandSrc cs = cand_src(I, rd, best, X, seed, gen);
    float *dst = U + I.elem_off + (size_t)rd.li * I.ld;
    for (int g = l; g < (I.ld >> 2); g += 16)
        *reinterpret_cast<float4 *>(dst + 4 * g) = cand_group(cs, g, seed, lo, hi);
}

int pick_regen(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C, const float *pred,
               const float *rho, float rho_min, const float *X, float lo, float hi, uint64_t seed,
               const uint32_t *gen_base, uint32_t gen_offset, float *U, hipStream_t s) {
    if (Q <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_pick_regen, dim3((Q + 15) / 16), dim3(256), 0, s, isl, rows, Q, C, pred,
                       rho, rho_min, X, lo, hi, seed, gen_base, gen_offset, U);
    return check_launch("k_pick_regen");
}

int pick_candidates(const IslandDesc *isl, const SurrRowDesc *rows, int Q, int C,
                    const float *cand, int Kp, const float *pred, const float *rho,
                    float rho_min, float *U, hipStream_t s) {
    if (Q <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_pick_candidates, dim3((Q + 15) / 16), dim3(256), 0, s, isl, rows, Q, C,
                       cand, Kp, pred, rho, rho_min, U);
    return check_launch("k_pick_candidates");
}

// Archive slot of target q: (head[p] + q - qoff[p]) % cap of its problem p;
// x is stored with Kp columns (zero padded) and y = log(1 + f).
__global__ void k_archive_append(const IslandDesc *isl, const SurrRowDesc *rows, int Q,
                                 const float *U, const float *fU, float *AX, float *Ay, int cap,
                                 int Kp, SurrHeads heads, const int *head_dev) {
    const int q = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int l = threadIdx.x & 15;
    if (q >= Q) return;
    const SurrRowDesc rd = rows[q];
    const IslandDesc I = isl[rd.island];
    const int p = rd.problem;
    const int slot = (int)(((int64_t)head_dev[p] + q - heads.qoff[p]) % cap);
    float *dst = AX + ((size_t)p * cap + slot) * Kp;
    const float *src = U + I.elem_off + (size_t)rd.li * I.ld;
    for (int j = l; j < Kp; j += 16) dst[j] = j < I.dim ? src[j] : 0.f;
    if (l == 0) Ay[(size_t)p * cap + slot] = log1pf(fmaxf(fU[rd.gr], 0.f));
}

__global__ void k_archive_advance(int *head_dev, SurrHeads heads, int nproblems, int cap) {
    const int p = threadIdx.x;
    if (p < nproblems) head_dev[p] = (int)(((int64_t)head_dev[p] + heads.rows[p]) % cap);
}

int archive_append(const IslandDesc *isl, const SurrRowDesc *rows, int Q, const float *U,
                   const float *fU, float *AX, float *Ay, int cap, int Kp, SurrHeads heads,
                   int nproblems, int *head_dev, hipStream_t s) {
    if (Q <= 0) return SPEO_OK;
    hipLaunchKernelGGL(k_archive_append, dim3((Q + 15) / 16), dim3(256), 0, s, isl, rows, Q, U, fU,
                       AX, Ay, cap, Kp, heads, head_dev);
    SPEO_TRY(check_launch("k_archive_append"));
    hipLaunchKernelGGL(k_archive_advance, dim3(1), dim3(kMaxSurrProblems), 0, s, head_dev, heads,
                       nproblems, cap);
    return check_launch("k_archive_advance");
}

// Minibatch of the valid archive slots [0, n_valid), skipping the held-out
// slots (slot % 8 == 7, mapped to slot - 1).
__global__ void k_gather_batch(const float *AX, const float *Ay, int n_valid, int Kp,
                               uint64_t seed, uint32_t round, uint32_t step, int batch, float *Xb,
                               float *tb) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (b >= batch) return;
    const U4 w = philox_block(seed, (uint32_t)(b >> 2), round, step, kPurposeSurrBatch);
    int slot = (int)(u01(word(w, (uint32_t)(b & 3))) * (float)n_valid);
    slot = slot >= n_valid ? n_valid - 1 : slot;
    if (slot % 8 == 7) slot -= 1;
    for (int j = lane; j < Kp; j += 64) Xb[(size_t)b * Kp + j] = AX[(size_t)slot * Kp + j];
    if (lane == 0) tb[b] = Ay[slot];
}

int gather_batch(const float *AX, const float *Ay, int n_valid, int Kp, uint64_t seed,
                 uint32_t round, uint32_t step, int batch, float *Xb, float *tb, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_batch, dim3((batch + 3) / 4), dim3(256), 0, s, AX, Ay, n_valid, Kp,
                       seed, round, step, batch, Xb, tb);
    return check_launch("k_gather_batch");
}

// Held-out points: slots 8h + 7, h < H.
__global__ void k_gather_holdout(const float *AX, const float *Ay, int H, int Kp, float *Xh,
                                 float *yh) {
    const int h = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (h >= H) return;
    const size_t slot = (size_t)h * 8 + 7;
    for (int j = lane; j < Kp; j += 64) Xh[(size_t)h * Kp + j] = AX[slot * Kp + j];
    if (lane == 0) yh[h] = Ay[slot];
}

int gather_holdout(const float *AX, const float *Ay, int H, int Kp, float *Xh, float *yh,
                   hipStream_t s) {
    hipLaunchKernelGGL(k_gather_holdout, dim3((H + 3) / 4), dim3(256), 0, s, AX, Ay, H, Kp, Xh, yh);
    return check_launch("k_gather_holdout");
}

// stats[1] = mean(y), stats[2] = max(std(y), 1e-3) over n values (one block,
// FP64 accumulation).
__global__ __launch_bounds__(1024) void k_mean_std(const float *y, int n, float *stats) {
    __shared__ double s1[1024], s2[1024];
    double a = 0.0, b = 0.0;
    // thread t sums y[t], y[t + 1024], ... in that order; eight loads are
    // issued before their sums instead of one round trip per value
    constexpr int kB = 8;
    int i = threadIdx.x;
    for (; i + (kB - 1) * 1024 < n; i += kB * 1024) {
        float v[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) v[u] = y[i + u * 1024];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            a += v[u];
            b += (double)v[u] * v[u];
        }
    }
    for (; i < n; i += 1024) {
        a += y[i];
        b += (double)y[i] * y[i];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            s1[threadIdx.x] += s1[threadIdx.x + s];
            s2[threadIdx.x] += s2[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double m = s1[0] / n;
        const double var = fmax(s2[0] / n - m * m, 0.0);
        stats[1] = (float)m;
        stats[2] = (float)fmax(sqrt(var), 1e-3);
    }
}

int mean_std(const float *y, int n, float *stats, hipStream_t s) {
    hipLaunchKernelGGL(k_mean_std, dim3(1), dim3(1024), 0, s, y, n, stats);
    return check_launch("k_mean_std");
}

}  // namespace speo

// ======================================================================
//  C ABI (test and tooling access to one MLP)
// ======================================================================
using namespace speo;

struct speo_mlp_s {
    Mlp m;
};

extern "C" {

int speo_mlp_create(int dim, int max_batch, uint64_t seed, speo_mlp *out) {
    if (!out || dim < 1 || max_batch < 1) return set_error(SPEO_ERR_ARG, "speo_mlp_create: bad arguments");
    *out = nullptr;
    auto *h = new (std::nothrow) speo_mlp_s;
    if (!h) return set_error(SPEO_ERR_ALLOC, "speo_mlp_create: out of host memory");
    int rc = h->m.alloc(dim, max_batch);
    if (rc == SPEO_OK) rc = h->m.init(seed, nullptr);
    if (rc != SPEO_OK) {
        delete h;
        return rc;
    }
    *out = h;
    return SPEO_OK;
}

int speo_mlp_destroy(speo_mlp h) {
    delete h;
    return SPEO_OK;
}

int64_t speo_mlp_num_params(speo_mlp h) { return h ? (int64_t)h->m.L.n : 0; }

int speo_mlp_params(speo_mlp h, float *get_out, const float *set_in) {
    if (!h) return set_error(SPEO_ERR_ARG, "speo_mlp_params: null handle");
    if (set_in) {
        SPEO_TRY_HIP(hipMemcpy(h->m.P, set_in, sizeof(float) * h->m.L.n, hipMemcpyHostToDevice));
        SPEO_TRY_HIP(hipMemset(h->m.m, 0, sizeof(float) * h->m.L.n));
        SPEO_TRY_HIP(hipMemset(h->m.v, 0, sizeof(float) * h->m.L.n));
        SPEO_TRY_HIP(hipMemset(h->m.t_dev, 0, sizeof(int)));
        h->m.bf16_stale = true;
    }
    if (get_out) SPEO_TRY_HIP(hipMemcpy(get_out, h->m.P, sizeof(float) * h->m.L.n, hipMemcpyDeviceToHost));
    return SPEO_OK;
}

int speo_mlp_forward(speo_mlp h, const float *X, int ld, int n, float *pred,
                     speo_stream_t stream) {
    if (!h || !X || !pred || n < 0 || ld < h->m.L.D)
        return set_error(SPEO_ERR_ARG, "speo_mlp_forward: bad arguments");
    return h->m.forward(X, ld, n, pred, as_stream(stream));
}

int speo_mlp_infer(speo_mlp h, const float *X, int ld, int n, float *pred,
                   speo_stream_t stream) {
    if (!h || !X || !pred || n < 0 || ld < h->m.L.D)
        return set_error(SPEO_ERR_ARG, "speo_mlp_infer: bad arguments");
    SPEO_TRY(h->m.refresh_bf16(as_stream(stream)));
    return h->m.infer(X, ld, n, pred, as_stream(stream));
}

int speo_mlp_loss_grad(speo_mlp h, const float *X, int ld, int n, const float *t,
                       float t_mean, float t_std, float *loss_out, float *grad_out) {
    if (!h || !X || !t || n < 1 || ld < h->m.L.D || !(t_std > 0))
        return set_error(SPEO_ERR_ARG, "speo_mlp_loss_grad: bad arguments");
    const float st[3] = {0.f, t_mean, t_std};
    SPEO_TRY_HIP(hipMemcpy(h->m.stats, st, sizeof(st), hipMemcpyHostToDevice));
    SPEO_TRY(h->m.loss_grad(X, ld, n, t, nullptr, loss_out != nullptr));
    SPEO_TRY_HIP(hipDeviceSynchronize());
    if (loss_out) SPEO_TRY_HIP(hipMemcpy(loss_out, h->m.stats, sizeof(float), hipMemcpyDeviceToHost));
    if (grad_out)
        SPEO_TRY_HIP(hipMemcpy(grad_out, h->m.G, sizeof(float) * h->m.L.n, hipMemcpyDeviceToHost));
    return SPEO_OK;
}

int speo_mlp_adam_step(speo_mlp h, float lr) {
    if (!h) return set_error(SPEO_ERR_ARG, "speo_mlp_adam_step: null handle");
    SPEO_TRY(h->m.adam(lr, nullptr));
    return check_hip(hipDeviceSynchronize(), "speo_mlp_adam_step");
}

int speo_mlp_train_step(speo_mlp h, const float *AX, const float *Ay, int n_valid, uint64_t seed,
                        uint32_t round, uint32_t step, int batch, float t_mean, float t_std,
                        float lr, float *loss_out) {
    if (!h || !AX || !Ay || n_valid < 8 || batch < 1 || batch > h->m.max_batch || !(t_std > 0) ||
        !(lr > 0))
        return set_error(SPEO_ERR_ARG, "speo_mlp_train_step: bad arguments");
    const float st[3] = {0.f, t_mean, t_std};
    SPEO_TRY_HIP(hipMemcpy(h->m.stats, st, sizeof(st), hipMemcpyHostToDevice));
    float *Xb = nullptr, *tb = nullptr;
    if (h->m.L.Kp > kFusedMaxKp) {  // the unfused fallback gathers into buffers
        SPEO_TRY_HIP(hipMalloc(&Xb, sizeof(float) * (size_t)batch * h->m.L.Kp));
        SPEO_TRY_HIP(hipMalloc(&tb, sizeof(float) * (size_t)batch));
    }
    int rc = h->m.train_step(AX, Ay, n_valid, seed, round, step, batch, lr, loss_out != nullptr, Xb,
                             tb, nullptr);
    if (rc == SPEO_OK) rc = check_hip(hipDeviceSynchronize(), "speo_mlp_train_step");
    (void)hipFree(Xb);
    (void)hipFree(tb);
    if (rc == SPEO_OK && loss_out)
        rc = check_hip(hipMemcpy(loss_out, h->m.stats, sizeof(float), hipMemcpyDeviceToHost),
                       "speo_mlp_train_step");
    return rc;
}

int speo_spearman(const float *a, const float *b, int n, float *rho_out, speo_stream_t stream) {
    if (!a || !b || !rho_out) return set_error(SPEO_ERR_ARG, "speo_spearman: null argument");
    float2 *ranks = nullptr;
    SPEO_TRY_HIP(hipMalloc(&ranks, sizeof(float2) * 4096));
    int rc = spearman(a, b, n, rho_out, ranks, as_stream(stream));
    if (rc == SPEO_OK) rc = check_hip(hipStreamSynchronize(as_stream(stream)), "speo_spearman");
    (void)hipFree(ranks);
    return rc;
}

}  // extern "C"
