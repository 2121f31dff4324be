// eval_kernels.hip -- fused shift + rotation-GEMM + objective kernels, the FP64
// re-verification kernel and the composition combine.
//
// Replaces the reference's per-objective launch chains (API_evaluateFitness,
// kernel_CEC2014_CUDA.cu:1672-1797: ShiftAndRotation -> global_fN, plus
// CalW/CombineFitness for compositions) by one kernel per N chunk of Z.
#include <algorithm>
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "cec_generic.h"
#include "eval.h"
#include "eval_f64.h"

namespace speo {

// ------------------------------------------------------------------ kernels

template <int RT, int PREC> constexpr int eval_small_lds_bytes() {
    return RT * kTS * 4 + small_scratch_bytes<RT>(Prec<PREC>::parts) + 10 * RT * 4;
}

template <int RT, int PREC>
__global__ __launch_bounds__(256) void k_eval(EvalLaunch p) {
    extern __shared__ __align__(16) char lds[];
    const int row0 = blockIdx.x * RT;
    const int chunk = blockIdx.y;
    const int nrows = min(RT, p.n - row0);
    float *o1, *o2;
    if (p.small) {
        float *rows = reinterpret_cast<float *>(lds);
        char *scratch = lds + RT * kTS * 4;
        float *red = reinterpret_cast<float *>(scratch + small_scratch_bytes<RT>(Prec<PREC>::parts));
        o1 = red + 8 * RT;
        o2 = o1 + RT;
        const int kp4 = (p.dim + 3) & ~3;
        if ((p.ld & 3) == 0 && (reinterpret_cast<uintptr_t>(p.X) & 15) == 0) {
            // 16-byte loads, all of the tile issued before the first is
            // stored; columns >= dim set to 0 (the row's padding may hold
            // anything)
            constexpr int G = RT * (kSmallD / 4) / 256;  // groups per thread at most
            const int ng = kp4 >> 2;
            floatx4 x[G];
#pragma unroll
            for (int i = 0; i < G; ++i) {
                const int e = threadIdx.x + i * 256, r = e / ng, g = e - r * ng;
                const bool ok = e < RT * ng && r < nrows;
                x[i] = gld(reinterpret_cast<const floatx4 *>(p.X + (ok ? (size_t)(row0 + r) * p.ld + 4 * g : 0)));
            }
#pragma unroll
            for (int i = 0; i < G; ++i) {
                const int e = threadIdx.x + i * 256, r = e / ng, g = e - r * ng;
                if (e >= RT * ng) break;
                floatx4 v = x[i];
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (r >= nrows || 4 * g + q >= p.dim) v[q] = 0.f;
                *reinterpret_cast<floatx4 *>(rows + r * kTS + 4 * g) = v;
            }
        } else {
            for (int e = threadIdx.x; e < RT * kp4; e += 256) {
                const int r = e / kp4, k = e % kp4;
                rows[r * kTS + k] = (r < nrows && k < p.dim) ? p.X[(size_t)(row0 + r) * p.ld + k] : 0.f;
            }
        }
        __syncthreads();
        float s1[RT / 16][4], s2[RT / 16][4];
        gemm_small<RT, PREC>(p.comp, rows, p.dim, scratch, s1, s2);
        reduce_rows_small<RT>(s1, s2, red, o1, o2);
    } else {
        RowSource src{p.X + (size_t)row0 * p.ld, p.ld, nrows};
        float s1[4] = {0.f, 0.f, 0.f, 0.f}, s2[4] = {0.f, 0.f, 0.f, 0.f};
        gemm_chunk<RT, PREC>(p.comp, src, p.dim, chunk * kNC, lds, s1, s2);
        float *red = reinterpret_cast<float *>(lds);
        o1 = red + 4 * RT * 2;
        o2 = o1 + RT;
        reduce_rows<RT>(s1, s2, red, o1, o2);
    }
    if (p.comp.heavy.j >= 0 && chunk == 0) {
        // the heavy column's FP64 term, one wave per row, into the row sums
        // (the row staging area is free after the reductions' barriers)
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int r = wave; r < nrows; r += 4) {
            const double h = heavy_term(p.comp.heavy, p.X + (size_t)(row0 + r) * p.ld, p.dim, lane);
            if (lane == 0) o1[r] = with_heavy(o1[r], (float)h);
        }
        __syncthreads();
    }
    const int t = threadIdx.x;
    if (t < RT && row0 + t < p.n) {
        const int row = row0 + t;
        if (p.nchunks == 1) {
            const float g = finish_objective_ext(p.comp.objective, o1[t], o2[t], p.dim);
            p.f_out[row] = p.raw ? g : fmaf(p.lambda, g, p.bias);
        } else {
            p.part[((size_t)chunk * 2 + 0) * p.n + row] = o1[t];
            p.part[((size_t)chunk * 2 + 1) * p.n + row] = o2[t];
        }
    }
}

__global__ void k_eval_finalize(const float *part, int nchunks, int n, int dim, int objective,
                                float lambda, float bias, int raw, float *f_out) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    float a = 0.f, b = 0.f;
    for (int c = 0; c < nchunks; ++c) {
        a += part[((size_t)c * 2 + 0) * n + row];
        b += part[((size_t)c * 2 + 1) * n + row];
    }
    const float g = finish_objective_ext(objective, a, b, dim);
    f_out[row] = raw ? g : fmaf(lambda, g, bias);
}

// Composition (cf_cal, CEC2014.cc:1481-1519) in FP32: one wave per row.
// Weights in the log domain, lw_c = -log(d_c)/2 - d_c / (2 D sigma_c^2),
// normalised by their maximum, so no weight underflows before the ratio is
// formed; components with d_c == 0 share all the weight equally (cf_cal gives
// each of them w = 1e99, beside which every other weight vanishes).
// One wave per row: the row's squared distance to each component's shift
// (per-lane partial sums, then a butterfly over the 64 lanes), the cf_cal
// weights in the log domain and the weighted sum.  When the row stride, X and
// every shift allow 16-byte loads and dim % 4 == 0, lane l takes coordinates
// 4 (l + 64 i) .. + 3 and the row (dim <= 1024) is read once into registers
// for all components; otherwise coordinates l + 64 i, one at a time.
__global__ void k_comp_combine(const CompParams *comps, int ncomp, int dim, const float *X,
                               int ld, int n, const float *g, float *f_out) {
    constexpr int kGroups = 4;  // float4 groups per lane held in registers
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (wave >= n) return;
    const float *x = X + (size_t)wave * ld;
    const int ng = dim >> 2;
    bool vec = (dim & 3) == 0 && (ld & 3) == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0 &&
               ng <= 64 * kGroups;
    for (int c = 0; c < ncomp; ++c) vec = vec && (reinterpret_cast<uintptr_t>(comps[c].shift32) & 15) == 0;
    float4 xr[kGroups];
    if (vec) {
#pragma unroll
        for (int i = 0; i < kGroups; ++i) {
            const int q = lane + 64 * i;
            xr[i] = q < ng ? *reinterpret_cast<const float4 *>(x + 4 * q) : float4{0.f, 0.f, 0.f, 0.f};
        }
    }
    float lw[kMaxComp];
    bool zero[kMaxComp];
    bool any_zero = false;
    float lwmax = -INFINITY;
    for (int c = 0; c < ncomp; ++c) {
        const float *o = comps[c].shift32;
        float d = 0.f;
        if (vec) {
#pragma unroll
            for (int i = 0; i < kGroups; ++i) {
                const int q = lane + 64 * i;
                if (q < ng) {
                    const float4 ov = *reinterpret_cast<const float4 *>(o + 4 * q);
                    const float v0 = xr[i].x - ov.x, v1 = xr[i].y - ov.y, v2 = xr[i].z - ov.z,
                                v3 = xr[i].w - ov.w;
                    d = fmaf(v0, v0, d);
                    d = fmaf(v1, v1, d);
                    d = fmaf(v2, v2, d);
                    d = fmaf(v3, v3, d);
                }
            }
        } else {
            for (int j = lane; j < dim; j += 64) {
                const float v = x[j] - o[j];
                d = fmaf(v, v, d);
            }
        }
        for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
        zero[c] = d == 0.f;
        any_zero = any_zero || zero[c];
        const float s = (float)comps[c].sigma;
        lw[c] = -0.5f * logf(d) - d / (2.0f * (float)dim * s * s);
        lwmax = fmaxf(lwmax, lw[c]);
    }
    if (lane == 0) {
        float wsum = 0.f, acc = 0.f;
        for (int c = 0; c < ncomp; ++c) {
            const float w = any_zero ? (zero[c] ? 1.f : 0.f) : expf(lw[c] - lwmax);
            const float fit = fmaf((float)comps[c].lambda, g[(size_t)c * n + wave], (float)comps[c].bias);
            wsum += w;
            acc = fmaf(w, fit, acc);
        }
        f_out[wave] = acc / wsum;
    }
}

// k_comp_combine's 16-byte path for many rows: the shifts staged in LDS once
// per block, kCombRows rows per wave with the next row's loads (X and the
// component values g) in flight while the current row is combined.  Lane c
// (c < ncomp) forms component c's weight and value, and lane 0 sums them in
// component order: the same arithmetic as k_comp_combine, result for result.
constexpr int kCombRows = 4;
__global__ __launch_bounds__(256) void k_comp_combine_rows(const CompParams *comps, int ncomp, int dim,
                                                           const float *X, int ld, int n,
                                                           const float *g, float *f_out) {
    constexpr int kGroups = 4;
    extern __shared__ __align__(16) float osh[];  // [ncomp][dim]
    __shared__ float sig[kMaxComp], lam[kMaxComp], bia[kMaxComp];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ng = dim >> 2;
    for (int c = 0; c < ncomp; ++c)
        for (int q = tid; q < ng; q += 256)
            reinterpret_cast<float4 *>(osh + (size_t)c * dim)[q] =
                reinterpret_cast<const float4 *>(comps[c].shift32)[q];
    if (tid < ncomp) {
        sig[tid] = (float)comps[tid].sigma;
        lam[tid] = (float)comps[tid].lambda;
        bia[tid] = (float)comps[tid].bias;
    }
    const int row0 = (blockIdx.x * 4 + wave) * kCombRows;
    auto load = [&](int r, float4 (&xr)[kGroups], float &gv) {
#pragma unroll
        for (int i = 0; i < kGroups; ++i) {
            const int q = lane + 64 * i;
            xr[i] = r < n && q < ng ? *reinterpret_cast<const float4 *>(X + (size_t)r * ld + 4 * q)
                                    : float4{0.f, 0.f, 0.f, 0.f};
        }
        gv = r < n && lane < ncomp ? g[(size_t)lane * n + r] : 0.f;
    };
    float4 cur[kGroups], nxt[kGroups];
    float gcur, gnxt;
    load(row0, cur, gcur);
    __syncthreads();
    for (int k = 0; k < kCombRows; ++k) {
        const int r = row0 + k;
        if (r >= n) break;  // wave-uniform
        load(r + 1, nxt, gnxt);
        float my_lw = -INFINITY, lwmax = -INFINITY;
        bool my_zero = false, any_zero = false;
        for (int c = 0; c < ncomp; ++c) {
            const float *o = osh + (size_t)c * dim;
            float d = 0.f;
#pragma unroll
            for (int i = 0; i < kGroups; ++i) {
                const int q = lane + 64 * i;
                if (q < ng) {
                    const float4 ov = *reinterpret_cast<const float4 *>(o + 4 * q);
                    const float v0 = cur[i].x - ov.x, v1 = cur[i].y - ov.y, v2 = cur[i].z - ov.z,
                                v3 = cur[i].w - ov.w;
                    d = fmaf(v0, v0, d);
                    d = fmaf(v1, v1, d);
                    d = fmaf(v2, v2, d);
                    d = fmaf(v3, v3, d);
                }
            }
            for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
            const float s = sig[c];
            const float lw = -0.5f * logf(d) - d / (2.0f * (float)dim * s * s);
            any_zero = any_zero || d == 0.f;
            lwmax = fmaxf(lwmax, lw);
            if (lane == c) {
                my_lw = lw;
                my_zero = d == 0.f;
            }
        }
        float w = 0.f, fit = 0.f;
        if (lane < ncomp) {
            w = any_zero ? (my_zero ? 1.f : 0.f) : expf(my_lw - lwmax);
            fit = fmaf(lam[lane], gcur, bia[lane]);
        }
        float wsum = 0.f, acc = 0.f;
        for (int c = 0; c < ncomp; ++c) {
            const float wc = __shfl(w, c, 64), fc = __shfl(fit, c, 64);
            wsum += wc;
            acc = fmaf(wc, fc, acc);
        }
        if (lane == 0) f_out[r] = acc / wsum;
#pragma unroll
        for (int i = 0; i < kGroups; ++i) cur[i] = nxt[i];
        gcur = gnxt;
    }
}

__global__ __launch_bounds__(256) void k_eval_f64(const CompParams *comps, int ncomp, int dim,
                                                  const float *X, int ld, int n, double *f_out) {
    extern __shared__ __align__(16) char lds_raw[];
    const int row = blockIdx.x;
    if (row >= n) return;
    const double f = eval_row_f64(comps, ncomp, dim, X + (size_t)row * ld,
                                  reinterpret_cast<double *>(lds_raw));
    if (threadIdx.x == 0) f_out[row] = f;
}

// ------------------------------------------------------------------ launch

// k_eval_wide for the 16-bit evaluations that k_eval<64> would run.
static std::atomic<int> g_wide_eval{1};
static std::atomic<long long> g_wide_launches{0};

template <int RT, int PREC>
static int launch_rt_prec(const EvalLaunch &p, int n, int nchunks, hipStream_t s) {
    static std::once_flag once;
    constexpr int lds_max = std::max(gemm_lds_bytes<RT, PREC>(), eval_small_lds_bytes<RT, PREC>());
    std::call_once(once, [] {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eval<RT, PREC>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds_max);
    });
    const int lds = p.small ? eval_small_lds_bytes<RT, PREC>() : gemm_lds_bytes<RT, PREC>();
    dim3 grid((n + RT - 1) / RT, nchunks);
    hipLaunchKernelGGL((k_eval<RT, PREC>), grid, dim3(256), lds, s, p);
    return check_launch("k_eval");
}

template <int RT>
static int launch_rt(const EvalLaunch &p, int precision, int n, int nchunks, hipStream_t s) {
    switch (precision) {
    case SPEO_PREC_FP32: return launch_rt_prec<RT, SPEO_PREC_FP32>(p, n, nchunks, s);
    case SPEO_PREC_BF16: return launch_rt_prec<RT, SPEO_PREC_BF16>(p, n, nchunks, s);
    case SPEO_PREC_FP16: return launch_rt_prec<RT, SPEO_PREC_FP16>(p, n, nchunks, s);
    case SPEO_PREC_BF16X3: return launch_rt_prec<RT, SPEO_PREC_BF16X3>(p, n, nchunks, s);
    case SPEO_PREC_FP16X3: return launch_rt_prec<RT, SPEO_PREC_FP16X3>(p, n, nchunks, s);
    case SPEO_PREC_BF16X6: return launch_rt_prec<RT, SPEO_PREC_BF16X6>(p, n, nchunks, s);
    default: return set_error(SPEO_ERR_ARG, "unknown precision mode");
    }
}

int eval_component(const CompDev &comp, const float *X, int ld, int n, int dim, float *f_out,
                   float *part, float lambda, float bias, int raw, int precision,
                   int rows_per_block, hipStream_t s) {
    if (n <= 0) return SPEO_OK;
    const int nchunks = (dim + kNC - 1) / kNC;
    if (nchunks > 1 && part == nullptr) return set_error(SPEO_ERR_ARG, "eval: missing partial buffer");
    const int small = (dim <= kSmallD && comp.B32s != nullptr) ? 1 : 0;
    EvalLaunch p{comp, X, ld, n, dim, f_out, part, nchunks, lambda, bias, raw, small};
    int rc;
    if (!small && rows_per_block == 64 && g_wide_eval.load(std::memory_order_relaxed) &&
        eval_wide_applies(comp, X, ld, dim, precision)) {
        // the same bits as k_eval<64, precision>, on wider tiles
        rc = eval_wide(comp, X, ld, n, dim, part, precision, s);
        g_wide_launches.fetch_add(1, std::memory_order_relaxed);
    } else switch (rows_per_block) {
    case 16: rc = launch_rt<16>(p, precision, n, nchunks, s); break;
    case 32: rc = launch_rt<32>(p, precision, n, nchunks, s); break;
    case 64: rc = launch_rt<64>(p, precision, n, nchunks, s); break;
    default: return set_error(SPEO_ERR_ARG, "rows_per_block must be 16, 32 or 64");
    }
    if (rc != SPEO_OK || nchunks == 1) return rc;
    hipLaunchKernelGGL(k_eval_finalize, dim3((n + 255) / 256), dim3(256), 0, s, part, nchunks, n,
                       dim, comp.objective, lambda, bias, raw, f_out);
    return check_launch("k_eval_finalize");
}

int eval_problem_generic(const ProblemHost &P, const float *X, int ld, int n, float *f32,
                         double *f64, const int *rows, int row_base, hipStream_t s) {
    GenLaunch g{};
    g.P = *P.gen;
    g.X = X;
    g.ld = ld;
    g.n = n;
    g.rows = rows;
    g.row_base = row_base;
    g.f32 = f32;
    g.f64 = f64;
    return eval_generic(g, s);
}

size_t eval_scratch_floats(int ncomp, int dim, int n) {
    const int nchunks = (dim + kNC - 1) / kNC;
    size_t part = nchunks > 1 ? (size_t)nchunks * 2 * n : 0;
    size_t comp = ncomp > 1 ? (size_t)ncomp * n : 0;
    return part + comp;
}

int eval_problem(const ProblemHost &P, const float *X, int ld, int n, float *f_out,
                 float *scratch, int precision, int rows_per_block, hipStream_t s) {
    if (n <= 0) return SPEO_OK;
    if (P.generic) return eval_problem_generic(P, X, ld, n, f_out, nullptr, nullptr, 0, s);
    const int nchunks = (P.dim + kNC - 1) / kNC;
    float *part = nchunks > 1 ? scratch : nullptr;
    float *gbuf = scratch + (nchunks > 1 ? (size_t)nchunks * 2 * n : 0);
    // an unrotated component's z = (x - o) rate is exact in FP32; a 16-bit
    // identity GEMM would round it to the operand precision
    auto prec_of = [&](int c) { return P.comp[c].unrotated ? (int)SPEO_PREC_FP32 : precision; };
    if (P.ncomp == 1) {
        const CompHost &c = P.comp[0];
        return eval_component(P.view(0, prec_of(0)), X, ld, n, P.dim, f_out, part,
                              (float)c.lambda, (float)c.bias, 0, prec_of(0), rows_per_block, s);
    }
    for (int c = 0; c < P.ncomp; ++c)
        SPEO_TRY(eval_component(P.view(c, prec_of(c)), X, ld, n, P.dim, gbuf + (size_t)c * n, part,
                                1.f, 0.f, 1, prec_of(c), rows_per_block, s));
    bool rows = (P.dim & 3) == 0 && P.dim <= 1024 && (ld & 3) == 0 &&
                (reinterpret_cast<uintptr_t>(X) & 15) == 0 && n >= 4096;
    for (int c = 0; c < P.ncomp; ++c) rows = rows && (reinterpret_cast<uintptr_t>(P.comp[c].shift32) & 15) == 0;
    if (rows) {
        hipLaunchKernelGGL(k_comp_combine_rows, dim3((n + 4 * kCombRows - 1) / (4 * kCombRows)), dim3(256),
                           sizeof(float) * (size_t)P.ncomp * P.dim, s, P.params_dev, P.ncomp, P.dim, X, ld,
                           n, gbuf, f_out);
        return check_launch("k_comp_combine_rows");
    }
    const int waves_per_block = 4;
    hipLaunchKernelGGL(k_comp_combine, dim3((n + waves_per_block - 1) / waves_per_block),
                       dim3(64 * waves_per_block), 0, s, P.params_dev, P.ncomp, P.dim, X, ld, n,
                       gbuf, f_out);
    return check_launch("k_comp_combine");
}

int eval_problem_f64(const ProblemHost &P, const float *X, int ld, int n, double *f_out,
                     hipStream_t s) {
    if (n <= 0) return SPEO_OK;
    if (P.generic) return eval_problem_generic(P, X, ld, n, nullptr, f_out, nullptr, 0, s);
    const size_t lds = sizeof(double) * ((size_t)P.dim + 16);
    if (lds > 48 * 1024)
        SPEO_TRY_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eval_f64),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_eval_f64, dim3(n), dim3(256), lds, s, P.params_dev, P.ncomp, P.dim, X,
                       ld, n, f_out);
    return check_launch("k_eval_f64");
}

int default_rows_per_block(int n, int dim, int precision) {
    // Enough blocks to cover the 256 CUs; larger tiles share B staging.
    const int nchunks = (dim + kNC - 1) / kNC;
    // 16-bit evaluations wider than one chunk: 64-row arithmetic from 1024
    // rows on, which k_eval_wide runs on its own wider tiles
    if (precision != SPEO_PREC_FP32 && nchunks > 1 && n >= 1024) return 64;
    if ((long)((n + 63) / 64) * nchunks >= 512) return 64;
    if ((long)((n + 31) / 32) * nchunks >= 256) return 32;
    return 16;
}

// ------------------------------------------------- drop-in API_evaluateFitness

namespace {
struct DimTables {
    float *colw[3] = {nullptr, nullptr, nullptr};  // elliptic, bent cigar, discus weights
    float *identity = nullptr;
};
std::mutex g_tables_mu;
std::unordered_map<int, DimTables> g_tables;

int dim_tables(int dim, DimTables &out) {
    std::lock_guard<std::mutex> lk(g_tables_mu);
    auto it = g_tables.find(dim);
    if (it != g_tables.end()) {
        out = it->second;
        return SPEO_OK;
    }
    DimTables t;
    std::vector<float> w((size_t)dim), I((size_t)dim * dim, 0.f);
    for (int j = 0; j < dim; ++j) I[(size_t)j * dim + j] = 1.f;
    const int kinds[3] = {SPEO_OBJ_ELLIPTIC, SPEO_OBJ_BENT_CIGAR, SPEO_OBJ_DISCUS};
    for (int k = 0; k < 3; ++k) {
        column_weights(kinds[k], dim, w.data());
        SPEO_TRY_HIP(hipMalloc(&t.colw[k], sizeof(float) * dim));
        SPEO_TRY_HIP(hipMemcpy(t.colw[k], w.data(), sizeof(float) * dim, hipMemcpyHostToDevice));
    }
    SPEO_TRY_HIP(hipMalloc(&t.identity, sizeof(float) * dim * dim));
    SPEO_TRY_HIP(hipMemcpy(t.identity, I.data(), sizeof(float) * dim * dim, hipMemcpyHostToDevice));
    g_tables[dim] = t;
    out = t;
    return SPEO_OK;
}
}  // namespace

// The reference's CEC2014_CUDA (CEC2014_CUDA.cc:68-128, 160-289) hands every
// call its device buffers: d_shift_data [ncomp][next_pow2(dim)],
// d_rotated_data [ncomp][next_pow2(dim)][dim] (B = M^T, row-padded),
// d_shuffle_data [ncomp][next_pow2(dim)] (1-based), d_sigma [ncomp] and the
// scratch rows d_rotated_elements [pop_size][next_pow2(dim)].
static int dropin_generic(real *f, real *X, real *shift, real *B, int *shuffle, real *sigma,
                          int n, int dim, real bias, int fid) {
    int nc = 0;
    CecComp cc[kMaxComp];
    if (!cec_function(fid, &nc, cc)) return set_error(SPEO_ERR_UNSUPPORTED, "unknown function");
    const int Dp = next_pow2_dim(dim);
    GenLaunch g{};
    g.P.dim = dim;
    g.P.ncomp = nc;
    g.P.sigma32 = nc > 1 ? sigma : nullptr;
    for (int c = 0; c < nc; ++c) {
        GenComp &G = g.P.comp[c];
        if (cc[c].objective >= SPEO_OBJ_HYBRID_1) {
            if (!shuffle) return set_error(SPEO_ERR_ARG, "missing shuffle data");
            hybrid_segments(cc[c].objective, dim, G);
            G.perm = shuffle + (size_t)c * Dp;
        } else {
            G.nseg = 1;
            G.seg[0] = GenSeg{cc[c].objective, dim, 1.0};
        }
        if (cc[c].rotate && !B) return set_error(SPEO_ERR_ARG, "missing rotation matrix");
        G.rate = cc[c].rate;
        G.rotate = cc[c].rotate;
        G.o32 = shift + (size_t)c * Dp;
        G.B32 = B ? B + (size_t)c * Dp * dim : nullptr;
        G.ldb = dim;
        G.lambda = cc[c].lambda;
        G.bias = cc[c].bias;
        G.sigma = cc[c].sigma;
    }
    g.X = X;
    g.ld = Dp;
    g.n = n;
    g.f32 = f;
    // F1-F16 ignore the call's bias; the hybrid and composition chains add it
    // in CombineFitness (kernel_CEC2014_CUDA.cu:205, 1227, 1447).
    g.out_bias = fid >= 17 ? bias : 0.0;
    return eval_generic(g, nullptr);
}

}  // namespace speo

using namespace speo;

extern "C" int speo_wide_eval(int enable, long long *launches) {
    if (launches) *launches = g_wide_launches.load();
    return enable < 0 ? g_wide_eval.load() : g_wide_eval.exchange(enable ? 1 : 0);
}

extern "C" void API_evaluateFitness(real *d_fitness_value, real *, real *d_rotated_elements,
                                    real *d_original_elements, real *d_shift_data,
                                    real *d_rotated_data, int *d_shuffle_data, real *,
                                    real *, real *d_sigma, int pop_size, int problem_dim,
                                    real bias, int function_ID) {
    if (pop_size <= 0 || problem_dim < 2 || !d_fitness_value || !d_original_elements ||
        !d_shift_data) {
        defer_error(SPEO_ERR_ARG, "API_evaluateFitness: invalid arguments");
        return;
    }
    if (function_ID < 1 || function_ID > 30) {
        defer_error(SPEO_ERR_UNSUPPORTED, "API_evaluateFitness: function_ID " +
                                              std::to_string(function_ID) + " is not in 1..30");
        return;
    }
    CompDev c{};
    bool rotate = true;
    int weights = 0;  // DimTables::colw entry of the weighted-sphere objectives
    switch (function_ID) {
    case 1: c.objective = SPEO_OBJ_ELLIPTIC; c.rate = 1.0f; break;       // F1, kernel_CEC2014_CUDA.cu:1105
    // F2 bent cigar and F3 discus (:1111, :1117; global_f2 / global_f3 at
    // :287-352): the weighted sphere with their own column weights.
    case 2: c.objective = SPEO_OBJ_ELLIPTIC; c.rate = 1.0f; weights = 1; break;
    case 3: c.objective = SPEO_OBJ_ELLIPTIC; c.rate = 1.0f; weights = 2; break;
    case 5: c.objective = SPEO_OBJ_ACKLEY; c.rate = 1.0f; break;         // F5, :1130
    case 8: c.objective = SPEO_OBJ_RASTRIGIN; c.rate = 0.0512f; rotate = false; break;  // F8, :1150
    case 9: c.objective = SPEO_OBJ_RASTRIGIN; c.rate = 0.0512f; break;   // F9, :1156
    case 10: c.objective = SPEO_OBJ_SCHWEFEL; c.rate = 10.0f; rotate = false; break;  // F10, :1162
    case 11: c.objective = SPEO_OBJ_SCHWEFEL; c.rate = 10.0f; break;     // F11, :1168
    case 13: c.objective = SPEO_OBJ_HAPPYCAT; c.rate = 0.05f; break;     // F13, :1180
    case 14: c.objective = SPEO_OBJ_HGBAT; c.rate = 0.05f; break;        // F14, :1186
    default: {
        // F4, F6, F7, F12, F15, F16, the hybrids F17-F22 and the compositions
        // F23-F30 (:1123-1671): the FP64 generic kernel.
        const int rc = dropin_generic(d_fitness_value, d_original_elements, d_shift_data,
                                      d_rotated_data, d_shuffle_data, d_sigma, pop_size,
                                      problem_dim, bias, function_ID);
        if (rc != SPEO_OK) defer_error(rc, std::string("API_evaluateFitness: ") + speo_last_error());
        return;
    }
    }
    if (rotate && !d_rotated_data) {
        defer_error(SPEO_ERR_ARG, "API_evaluateFitness: missing rotation matrix");
        return;
    }
    DimTables t;
    if (dim_tables(problem_dim, t) != SPEO_OK) {
        defer_error(SPEO_ERR_HIP, "API_evaluateFitness: table allocation failed");
        return;
    }
    c.shift = d_shift_data;
    c.B32 = rotate ? d_rotated_data : t.identity;
    c.ldb32 = problem_dim;  // reference layout [next_pow2(dim)][dim] (CEC2014_CUDA.cc:255)
    c.colw = t.colw[weights];
    c.heavy.j = -1;
    if (c.objective == SPEO_OBJ_ELLIPTIC) {
        static const int kinds[3] = {SPEO_OBJ_ELLIPTIC, SPEO_OBJ_BENT_CIGAR, SPEO_OBJ_DISCUS};
        c.heavy.j = heavy_column(kinds[weights], problem_dim, &c.heavy.w);
        c.heavy.m32 = c.B32 + (c.heavy.j >= 0 ? c.heavy.j : 0);  // column j of B = row j of M
        c.heavy.ld = problem_dim;
        c.heavy.o32 = d_shift_data;
        c.heavy.rate = 1.0;
    }
    const int ld = next_pow2_dim(problem_dim);
    // Above 128 columns the per-chunk row sums go to the caller's rotated-row
    // buffer (pop_size * next_pow2(dim) floats, at least 2 * chunks * pop_size).
    float *part = nullptr;
    if (problem_dim > kNC) {
        if (!d_rotated_elements) {
            defer_error(SPEO_ERR_ARG, "API_evaluateFitness: dim > 128 needs d_rotated_elements");
            return;
        }
        part = d_rotated_elements;
    }
    const int rc = eval_component(c, d_original_elements, ld, pop_size, problem_dim,
                                  d_fitness_value, part, 1.f, 0.f, 0, SPEO_PREC_FP32,
                                  default_rows_per_block(pop_size, problem_dim), nullptr);
    if (rc != SPEO_OK) defer_error(rc, speo_last_error());
}
