// cec_generic.hip -- FP64 evaluation of the CEC2014 objectives that the fused
// FP32 epilogues do not take (see cec_generic.h).
//
// Reference counterparts: the basic kernels global_f4, f6, f7, f10, f13, f14
// (kernel_CEC2014_CUDA.cu:355-387, 443-556, 649-701, 809-976), the hybrid
// chains HybirdFunc1-6 (:1204-1415: ShiftAndRotation, Shuffle, one kernel per
// segment, CombineFitness) and the compositions CompFunc1-8 (:1416-1671:
// ShiftAndRotation per component, CalW, CombineFitness).  The arithmetic
// follows the reference's FP64 CPU forms (CEC2014.cc:481-1519), which define
// the values; here one launch does all of it for a tile of 16 rows.
#include <algorithm>
#include <cmath>
#include <mutex>

#include "cec_generic.h"

namespace speo {

typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace {

constexpr double kPi = 3.1415926535897932384626433832795029;
constexpr double kE = 2.7182818284590452353602874713526625;
constexpr int kAS = 65;  // LDS row stride (doubles) of a 64-wide A chunk

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_prod(double v) {
    for (int off = 32; off > 0; off >>= 1) v *= __shfl_xor(v, off, 64);
    return v;
}

// Coordinate i of a segment: rate * z[j], j = start + i, through the shuffle
// (1-based, as the reference's global_Shuffle reads it, kernel_CEC2014_CUDA.cu:144).
struct Seg {
    const double *z;
    const int *perm;
    int start, len;
    double rate;
    __device__ __forceinline__ double y(int i) const {
        int j = start + i;
        if (perm) j = perm[j] - 1;
        return rate * z[j];
    }
};

__device__ __forceinline__ double grie_of_rosen(double a, double b) {
    const double t1 = a * a - b, t2 = a - 1.0;
    const double t = 100.0 * t1 * t1 + t2 * t2;
    return (t * t) / 4000.0 - cos(t) + 1.0;
}

__device__ __forceinline__ double scaffer6(double r2) {
    double t1 = sin(sqrt(r2));
    t1 = t1 * t1;
    const double t2 = 1.0 + 0.001 * r2;
    return 0.5 + (t1 - 0.5) / (t2 * t2);
}

// The objective of one segment (FP64, CEC2014.cc forms), every lane of the
// wave taking coordinates lane, lane + 64, ...; the wave's value is
// returned on every lane.
__device__ double seg_value(int obj, const Seg &s, int lane) {
    const int n = s.len;
    double a = 0.0, b = 0.0;
    switch (obj) {
    case SPEO_OBJ_ELLIPTIC:  // :481-497
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            a += pow(10.0, 6.0 * i / (n - 1)) * y * y;
        }
        return wave_sum(a);
    case SPEO_OBJ_BENT_CIGAR:  // :499-516
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            a += (i == 0 ? 1.0 : 1.0e6) * y * y;
        }
        return wave_sum(a);
    case SPEO_OBJ_DISCUS:  // :518-530
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            a += (i == 0 ? 1.0e6 : 1.0) * y * y;
        }
        return wave_sum(a);
    case SPEO_OBJ_ROSENBROCK:  // :552-568, fixed shift 1
        for (int i = lane; i < n - 1; i += 64) {
            const double y0 = s.y(i) + 1.0, y1 = s.y(i + 1) + 1.0;
            const double t1 = y0 * y0 - y1, t2 = y0 - 1.0;
            a += 100.0 * t1 * t1 + t2 * t2;
        }
        return wave_sum(a);
    case SPEO_OBJ_ACKLEY:  // :590-612
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            a += y * y;
            b += cos(2.0 * kPi * y);
        }
        a = wave_sum(a);
        b = wave_sum(b);
        return kE - 20.0 * exp(-0.2 * sqrt(a / (double)n)) + 20.0 - exp(b / (double)n);
    case SPEO_OBJ_WEIERSTRASS: {  // :615-644, a = 0.5, b = 3, k_max = 20
        // a^k and b^k as running products: 2^-k and 3^k (< 2^53) are exact in
        // FP64, so these are the values pow(a, k), pow(b, k) return
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            double t = 0.0, ak = 1.0, bk = 1.0;
            for (int k = 0; k <= 20; ++k, ak *= 0.5, bk *= 3.0) t += ak * cos(2.0 * kPi * bk * (y + 0.5));
            a += t;
        }
        double c = 0.0, ak = 1.0, bk = 1.0;
        for (int k = 0; k <= 20; ++k, ak *= 0.5, bk *= 3.0) c += ak * cos(2.0 * kPi * bk * 0.5);
        return wave_sum(a) - n * c;
    }
    case SPEO_OBJ_GRIEWANK: {  // :646-666
        double p = 1.0;
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            a += y * y;
            p *= cos(y / sqrt(1.0 + i));
        }
        return 1.0 + wave_sum(a) / 4000.0 - wave_prod(p);
    }
    case SPEO_OBJ_RASTRIGIN:  // :668-683
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            a += (y * y - 10.0 * cos(2.0 * kPi * y) + 10.0);
        }
        return wave_sum(a);
    case SPEO_OBJ_SCHWEFEL:  // :705-738
        for (int i = lane; i < n; i += 64) {
            const double v = s.y(i) + 4.209687462275036e+002;
            if (v > 500.0) {
                const double m = 500.0 - fmod(v, 500.0);
                a -= m * sin(sqrt(m));
                const double t = (v - 500.0) / 100.0;
                a += t * t / n;
            } else if (v < -500.0) {
                const double m = 500.0 - fmod(fabs(v), 500.0);
                a += m * sin(sqrt(m));
                const double t = (v + 500.0) / 100.0;
                a += t * t / n;
            } else {
                a -= v * sin(sqrt(fabs(v)));
            }
        }
        return wave_sum(a) + 4.189828872724338e+002 * n;
    case SPEO_OBJ_KATSUURA: {  // :739-765
        const double e = 10.0 / pow(1.0 * n, 1.2);
        double p = 1.0;
        for (int i = lane; i < n; i += 64) {
            const double y = s.y(i);
            double t = 0.0;
            for (int j = 1; j <= 32; ++j) {
                const double sc = ldexp(1.0, j), v = sc * y;
                t += fabs(v - floor(v + 0.5)) / sc;
            }
            p *= pow(1.0 + (i + 1) * t, e);
        }
        const double c = 10.0 / n / n;
        return wave_prod(p) * c - c;
    }
    case SPEO_OBJ_HAPPYCAT:  // :907-929
    case SPEO_OBJ_HGBAT:     // :931-953
        for (int i = lane; i < n; i += 64) {
            const double u = s.y(i) - 1.0;
            a += u * u;
            b += u;
        }
        a = wave_sum(a);
        b = wave_sum(b);
        if (obj == SPEO_OBJ_HAPPYCAT) return pow(fabs(a - n), 0.25) + (0.5 * a + b) / n + 0.5;
        return pow(fabs(a * a - b * b), 0.5) + (0.5 * a + b) / n + 0.5;
    case SPEO_OBJ_GRIE_ROSEN:  // :842-868, fixed shift 1, wrap-around pair
        for (int i = lane; i < n; i += 64)
            a += grie_of_rosen(s.y(i) + 1.0, s.y(i + 1 < n ? i + 1 : 0) + 1.0);
        return wave_sum(a);
    case SPEO_OBJ_ESCAFFER6:  // :870-905, wrap-around pair; one coordinate alone
        if (n == 1) {
            const double y = s.y(0);
            return scaffer6(y * y);
        }
        for (int i = lane; i < n; i += 64) {
            const double y0 = s.y(i), y1 = s.y(i + 1 < n ? i + 1 : 0);
            a += scaffer6(y0 * y0 + y1 * y1);
        }
        return wave_sum(a);
    default:
        return __builtin_nan("");
    }
}

__device__ __forceinline__ double shift_at(const GenComp &c, int k) {
    return c.o64 ? c.o64[k] : (double)c.o32[k];
}
__device__ __forceinline__ double rot_at(const GenComp &c, int dim, int col, int k) {
    return c.M64 ? c.M64[(size_t)col * dim + k] : (double)c.B32[(size_t)k * c.ldb + col];
}

}  // namespace

// LDS: Z tile [16][ldz] doubles (ldz = round_up(dim, 16) + 1), an A chunk
// [16][65], and per component the rows' fit and squared distance.
static size_t generic_lds_bytes(int dim) {
    const size_t ldz = (size_t)round_up(dim, 16) + 1;
    return sizeof(double) * (kGenRows * ldz + kGenRows * kAS + 2 * kMaxComp * kGenRows);
}

// NQ: 16-column tiles of Z per wave (dim <= 64 NQ).
template <int NQ>
__global__ __launch_bounds__(256) void k_eval_generic(GenLaunch p) {
    extern __shared__ __align__(16) char lds_raw[];
    const int dim = p.P.dim;
    const int ldz = round_up(dim, 16) + 1;
    double *Zs = reinterpret_cast<double *>(lds_raw);
    double *As = Zs + kGenRows * ldz;
    double *gv = As + kGenRows * kAS;  // [comp][row] fit
    double *dv = gv + kMaxComp * kGenRows;  // [comp][row] |x - o_c|^2
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int lr = lane & 15, lk = lane >> 4;
    const int row0 = blockIdx.x * kGenRows;
    const int nrows = min(kGenRows, p.n - row0);
    auto xrow = [&](int r) -> const float * {
        const int i = row0 + r;
        const int li = p.rows ? p.rows[i] - p.row_base : i;
        return p.X + (size_t)li * p.ld;
    };
    const int T = (dim + 15) >> 4;  // 16-column tiles of Z; wave w owns w, w + 4, ...

    for (int ci = 0; ci < p.P.ncomp; ++ci) {
        const GenComp &C = p.P.comp[ci];
        if (C.rotate) {
            doublex4 acc[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = doublex4{0.0, 0.0, 0.0, 0.0};
            for (int k0 = 0; k0 < dim; k0 += 64) {
                __syncthreads();
                for (int e = tid; e < kGenRows * 64; e += 256) {
                    const int r = e >> 6, k = k0 + (e & 63);
                    double v = 0.0;
                    if (r < nrows && k < dim) v = ((double)xrow(r)[k] - shift_at(C, k)) * C.rate;
                    As[r * kAS + (e & 63)] = v;
                }
                __syncthreads();
                const int ksteps = (min(64, dim - k0) + 3) >> 2;
                for (int ks = 0; ks < ksteps; ++ks) {
                    const int k = k0 + 4 * ks + lk;
                    const double a = As[lr * kAS + 4 * ks + lk];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int t = wave + 4 * q;
                        if (t < T) {
                            const int col = 16 * t + lr;
                            const double b = (k < dim && col < dim) ? rot_at(C, dim, col, k) : 0.0;
                            acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
                        }
                    }
                }
            }
            // accumulator element i of lane l of v_mfma_f64_16x16x4f64: row
            // l / 16 + 4 i, column l % 16
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int t = wave + 4 * q, col = 16 * t + lr;
                if (t < T && col < dim)
#pragma unroll
                    for (int i = 0; i < 4; ++i) Zs[(lk + 4 * i) * ldz + col] = acc[q][i];
            }
        } else {
            __syncthreads();
            for (int e = tid; e < kGenRows * dim; e += 256) {
                const int r = e / dim, k = e - r * dim;
                Zs[r * ldz + k] = r < nrows ? ((double)xrow(r)[k] - shift_at(C, k)) * C.rate : 0.0;
            }
        }
        __syncthreads();
        for (int r = wave; r < nrows; r += 4) {
            const double *z = Zs + r * ldz;
            double g = 0.0;
            int start = 0;
            for (int s = 0; s < C.nseg; ++s) {
                const Seg sv{z, C.perm, start, C.seg[s].len, C.seg[s].rate};
                g += seg_value(C.seg[s].obj, sv, lane);
                start += C.seg[s].len;
            }
            double d = 0.0;
            if (p.P.ncomp > 1) {
                const float *x = xrow(r);
                for (int j = lane; j < dim; j += 64) {
                    const double v = (double)x[j] - shift_at(C, j);
                    d += v * v;
                }
                d = wave_sum(d);
            }
            if (lane == 0) {
                gv[ci * kGenRows + r] = C.lambda * g + C.bias;
                dv[ci * kGenRows + r] = d;
            }
        }
    }
    __syncthreads();
    if (tid < nrows) {
        const int nc = p.P.ncomp;
        double f;
        if (nc == 1) {
            f = gv[tid];
        } else {
            // cf_cal (CEC2014.cc:1481-1519)
            double w[kMaxComp], wmax = 0.0, wsum = 0.0;
            for (int c = 0; c < nc; ++c) {
                const double d = dv[c * kGenRows + tid];
                const double sg = p.P.sigma32 ? (double)p.P.sigma32[c] : p.P.comp[c].sigma;
                w[c] = d != 0.0 ? pow(1.0 / d, 0.5) * exp(-d / 2.0 / dim / (sg * sg)) : 1.0e99;
                wmax = fmax(wmax, w[c]);
            }
            for (int c = 0; c < nc; ++c) wsum += w[c];
            if (wmax == 0.0) {
                for (int c = 0; c < nc; ++c) w[c] = 1.0;
                wsum = nc;
            }
            f = 0.0;
            for (int c = 0; c < nc; ++c) f += w[c] / wsum * gv[c * kGenRows + tid];
        }
        f += p.out_bias;
        if (p.f32) p.f32[row0 + tid] = (float)f;
        if (p.f64) p.f64[row0 + tid] = f;
    }
}

int eval_generic(const GenLaunch &p, hipStream_t s) {
    if (p.n <= 0) return SPEO_OK;
    if (p.P.dim < 1 || p.P.dim > kGenMaxDim || p.P.ncomp < 1 || p.P.ncomp > kMaxComp)
        return set_error(SPEO_ERR_UNSUPPORTED, "generic evaluation: dim must be in [1, 1024]");
    for (int c = 0; c < p.P.ncomp; ++c) {
        const GenComp &C = p.P.comp[c];
        int total = 0;
        for (int q = 0; q < C.nseg; ++q) total += C.seg[q].len;
        if (C.nseg < 1 || C.nseg > kGenMaxSeg || total != p.P.dim)
            return set_error(SPEO_ERR_ARG, "generic evaluation: segments do not cover the dimension");
        if (!C.o64 && !C.o32) return set_error(SPEO_ERR_ARG, "generic evaluation: missing shift");
        if (C.rotate && !C.M64 && !C.B32)
            return set_error(SPEO_ERR_ARG, "generic evaluation: missing rotation");
    }
    const dim3 grid((p.n + kGenRows - 1) / kGenRows);
    const size_t lds = generic_lds_bytes(p.P.dim);
    if (p.P.dim <= 128) {
        hipLaunchKernelGGL(k_eval_generic<2>, grid, dim3(256), lds, s, p);
    } else if (p.P.dim <= 256) {
        hipLaunchKernelGGL(k_eval_generic<4>, grid, dim3(256), lds, s, p);
    } else {
        static std::once_flag once;
        std::call_once(once, [] {
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_eval_generic<kGenMaxDim / 64>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      (int)generic_lds_bytes(kGenMaxDim));
        });
        hipLaunchKernelGGL(k_eval_generic<kGenMaxDim / 64>, grid, dim3(256), lds, s, p);
    }
    return check_launch("k_eval_generic");
}

// ------------------------------------------------------- function tables

bool generic_objective(int objective) {
    return objective > SPEO_OBJ_SCHWEFEL;
}

bool hybrid_segments(int objective, int dim, GenComp &c) {
    struct Def {
        int nseg;
        double p[kGenMaxSeg];
        int obj[kGenMaxSeg];
        double rate[kGenMaxSeg];
    };
    // hf01-hf06, CEC2014.cc:954-1207 (the GPU chains HybirdFunc1-6,
    // kernel_CEC2014_CUDA.cu:1204-1415, use the same objectives and rates)
    static const Def defs[6] = {
        {3, {0.3, 0.3, 0.4}, {SPEO_OBJ_SCHWEFEL, SPEO_OBJ_RASTRIGIN, SPEO_OBJ_ELLIPTIC},
         {1000.0 / 100.0, 5.12 / 100.0, 1.0}},
        {3, {0.3, 0.3, 0.4}, {SPEO_OBJ_BENT_CIGAR, SPEO_OBJ_HGBAT, SPEO_OBJ_RASTRIGIN},
         {1.0, 5.0 / 100.0, 5.12 / 100.0}},
        {4, {0.2, 0.2, 0.3, 0.3},
         {SPEO_OBJ_GRIEWANK, SPEO_OBJ_WEIERSTRASS, SPEO_OBJ_ROSENBROCK, SPEO_OBJ_ESCAFFER6},
         {600.0 / 100.0, 0.5 / 100.0, 2.048 / 100.0, 1.0}},
        {4, {0.2, 0.2, 0.3, 0.3},
         {SPEO_OBJ_HGBAT, SPEO_OBJ_DISCUS, SPEO_OBJ_GRIE_ROSEN, SPEO_OBJ_RASTRIGIN},
         {5.0 / 100.0, 1.0, 5.0 / 100.0, 5.12 / 100.0}},
        {5, {0.1, 0.2, 0.2, 0.2, 0.3},
         {SPEO_OBJ_ESCAFFER6, SPEO_OBJ_HGBAT, SPEO_OBJ_ROSENBROCK, SPEO_OBJ_SCHWEFEL, SPEO_OBJ_ELLIPTIC},
         {1.0, 5.0 / 100.0, 2.048 / 100.0, 1000.0 / 100.0, 1.0}},
        {5, {0.1, 0.2, 0.2, 0.2, 0.2},
         {SPEO_OBJ_KATSUURA, SPEO_OBJ_HAPPYCAT, SPEO_OBJ_GRIE_ROSEN, SPEO_OBJ_SCHWEFEL, SPEO_OBJ_ACKLEY},
         {5.0 / 100.0, 5.0 / 100.0, 5.0 / 100.0, 1000.0 / 100.0, 1.0}},
    };
    const int h = objective - SPEO_OBJ_HYBRID_1;
    if (h < 0 || h >= 6) return false;
    const Def &d = defs[h];
    c.nseg = d.nseg;
    int used = 0;
    for (int s = 0; s < d.nseg; ++s) {
        // ceil(p D) in FP64, CEC2014.cc:961-967; the last segment takes the rest
        const int len = s + 1 < d.nseg ? (int)std::ceil(d.p[s] * dim) : dim - used;
        c.seg[s] = GenSeg{d.obj[s], len, d.rate[s]};
        used += len;
    }
    return true;
}

bool cec_function(int func_id, int *ncomp, CecComp comp[kMaxComp]) {
    // F1-F16 (CEC2014.cc:17-79; kernel_CEC2014_CUDA.cu:1105-1203)
    static const CecComp basic[16] = {
        {SPEO_OBJ_ELLIPTIC, 1.0, 1, 1, 0, 10},           {SPEO_OBJ_BENT_CIGAR, 1.0, 1, 1, 0, 10},
        {SPEO_OBJ_DISCUS, 1.0, 1, 1, 0, 10},             {SPEO_OBJ_ROSENBROCK, 2.048 / 100.0, 1, 1, 0, 10},
        {SPEO_OBJ_ACKLEY, 1.0, 1, 1, 0, 10},             {SPEO_OBJ_WEIERSTRASS, 0.5 / 100.0, 1, 1, 0, 10},
        {SPEO_OBJ_GRIEWANK, 6.0, 1, 1, 0, 10},           {SPEO_OBJ_RASTRIGIN, 5.12 / 100.0, 0, 1, 0, 10},
        {SPEO_OBJ_RASTRIGIN, 5.12 / 100.0, 1, 1, 0, 10}, {SPEO_OBJ_SCHWEFEL, 10.0, 0, 1, 0, 10},
        {SPEO_OBJ_SCHWEFEL, 10.0, 1, 1, 0, 10},          {SPEO_OBJ_KATSUURA, 5.0 / 100.0, 1, 1, 0, 10},
        {SPEO_OBJ_HAPPYCAT, 5.0 / 100.0, 1, 1, 0, 10},   {SPEO_OBJ_HGBAT, 5.0 / 100.0, 1, 1, 0, 10},
        {SPEO_OBJ_GRIE_ROSEN, 5.0 / 100.0, 1, 1, 0, 10}, {SPEO_OBJ_ESCAFFER6, 1.0, 1, 1, 0, 10}};
    if (func_id >= 1 && func_id <= 16) {
        *ncomp = 1;
        comp[0] = basic[func_id - 1];
        return true;
    }
    if (func_id >= 17 && func_id <= 22) {  // hf01-hf06 (CEC2014.cc:80-97)
        *ncomp = 1;
        comp[0] = CecComp{SPEO_OBJ_HYBRID_1 + (func_id - 17), 1.0, 1, 1, 0, 10};
        return true;
    }
    // Compositions cf01-cf08 (CEC2014.cc:1209-1371; CompFunc1-8,
    // kernel_CEC2014_CUDA.cu:1416-1671): lambda is the scaling
    // 10000 * fit / 1e+4 etc. applied before cf_cal adds the bias.
    const double b5[5] = {0, 100, 200, 300, 400};
    auto set = [&](int n, const CecComp *src) {
        *ncomp = n;
        for (int c = 0; c < n; ++c) comp[c] = src[c];
        return true;
    };
    switch (func_id) {
    case 23: {
        const CecComp c[5] = {{SPEO_OBJ_ROSENBROCK, 2.048 / 100.0, 1, 10000 / 1e+4, b5[0], 10},
                              {SPEO_OBJ_ELLIPTIC, 1.0, 1, 10000 / 1e+10, b5[1], 20},
                              {SPEO_OBJ_BENT_CIGAR, 1.0, 1, 10000 / 1e+30, b5[2], 30},
                              {SPEO_OBJ_DISCUS, 1.0, 1, 10000 / 1e+10, b5[3], 40},
                              {SPEO_OBJ_ELLIPTIC, 1.0, 0, 10000 / 1e+10, b5[4], 50}};
        return set(5, c);
    }
    case 24: {
        const CecComp c[3] = {{SPEO_OBJ_SCHWEFEL, 10.0, 0, 1, 0, 20},
                              {SPEO_OBJ_RASTRIGIN, 5.12 / 100.0, 1, 1, 100, 20},
                              {SPEO_OBJ_HGBAT, 5.0 / 100.0, 1, 1, 200, 20}};
        return set(3, c);
    }
    case 25: {
        const CecComp c[3] = {{SPEO_OBJ_SCHWEFEL, 10.0, 1, 1000 / 4e+3, 0, 10},
                              {SPEO_OBJ_RASTRIGIN, 5.12 / 100.0, 1, 1000 / 1e+3, 100, 30},
                              {SPEO_OBJ_ELLIPTIC, 1.0, 1, 1000 / 1e+10, 200, 50}};
        return set(3, c);
    }
    case 26: {
        const CecComp c[5] = {{SPEO_OBJ_SCHWEFEL, 10.0, 1, 1000 / 4e+3, b5[0], 10},
                              {SPEO_OBJ_HAPPYCAT, 5.0 / 100.0, 1, 1000 / 1e+3, b5[1], 10},
                              {SPEO_OBJ_ELLIPTIC, 1.0, 1, 1000 / 1e+10, b5[2], 10},
                              {SPEO_OBJ_WEIERSTRASS, 0.5 / 100.0, 1, 1000 / 400.0, b5[3], 10},
                              {SPEO_OBJ_GRIEWANK, 6.0, 1, 1000 / 100.0, b5[4], 10}};
        return set(5, c);
    }
    case 27: {
        const CecComp c[5] = {{SPEO_OBJ_HGBAT, 5.0 / 100.0, 1, 10000 / 1000.0, b5[0], 10},
                              {SPEO_OBJ_RASTRIGIN, 5.12 / 100.0, 1, 10000 / 1e+3, b5[1], 10},
                              {SPEO_OBJ_SCHWEFEL, 10.0, 1, 10000 / 4e+3, b5[2], 10},
                              {SPEO_OBJ_WEIERSTRASS, 0.5 / 100.0, 1, 10000 / 400.0, b5[3], 20},
                              {SPEO_OBJ_ELLIPTIC, 1.0, 1, 10000 / 1e+10, b5[4], 20}};
        return set(5, c);
    }
    case 28: {
        const CecComp c[5] = {{SPEO_OBJ_GRIE_ROSEN, 5.0 / 100.0, 1, 10000 / 4e+3, b5[0], 10},
                              {SPEO_OBJ_HAPPYCAT, 5.0 / 100.0, 1, 10000 / 1e+3, b5[1], 20},
                              {SPEO_OBJ_SCHWEFEL, 10.0, 1, 10000 / 4e+3, b5[2], 30},
                              {SPEO_OBJ_ESCAFFER6, 1.0, 1, 10000 / 2e+7, b5[3], 40},
                              {SPEO_OBJ_ELLIPTIC, 1.0, 1, 10000 / 1e+10, b5[4], 50}};
        return set(5, c);
    }
    case 29:
    case 30: {
        const int h0 = func_id == 29 ? SPEO_OBJ_HYBRID_1 : SPEO_OBJ_HYBRID_4;
        const CecComp c[3] = {{h0, 1.0, 1, 1, 0, 10}, {h0 + 1, 1.0, 1, 1, 100, 30},
                              {h0 + 2, 1.0, 1, 1, 200, 50}};
        return set(3, c);
    }
    default:
        return false;
    }
}

}  // namespace speo
