// problem.hip -- problem objects and the seeded problem generator.
//
// The reference reads o and M from CEC text files and keeps M transposed and
// row-padded on the device (CUDA_EA_library/CEC2014_CUDA.cc:160-289).  Here a
// problem is created from caller-provided FP64 data (or from the seeded
// generator below) and stored in every layout the kernels read:
//   shift32/shift64, B32 = M^T (FP32 GEMM operand), BF16 and 2^14-scaled FP16
//   hi/lo/lo2 splits of M (16-bit MFMA operands), M64 (FP64 re-verification),
//   and the elliptic column weights.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cec_generic.h"
#include "eval.h"

namespace speo {

static uint16_t bf16_rne(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
static double bf16_val(uint16_t b) {
    uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return (double)f;
}
static uint16_t f16_rne(float f) {
    _Float16 h = (_Float16)f;
    uint16_t b;
    std::memcpy(&b, &h, 2);
    return b;
}
static double f16_val(uint16_t b) {
    _Float16 h;
    std::memcpy(&h, &b, 2);
    return (double)(float)h;
}

bool weighted_sphere(int objective) {
    return objective == SPEO_OBJ_ELLIPTIC || objective == SPEO_OBJ_BENT_CIGAR ||
           objective == SPEO_OBJ_DISCUS;
}

// CEC2014.cc:488-496 (elliptic), 507-515 (bent cigar), 521-529 (discus).
static double weight64(int objective, int dim, int j) {
    if (objective == SPEO_OBJ_BENT_CIGAR) return j == 0 ? 1.0 : 1.0e6;
    if (objective == SPEO_OBJ_DISCUS) return j == 0 ? 1.0e6 : 1.0;
    return std::pow(10.0, 6.0 * (double)j / (dim > 1 ? dim - 1 : 1));
}

int heavy_column(int objective, int dim, double *w_out) {
    if (!weighted_sphere(objective) || dim < 2) return -1;
    int jmax = 0;
    double total = 0.0;
    for (int j = 0; j < dim; ++j) {
        total += weight64(objective, dim, j);
        if (weight64(objective, dim, j) > weight64(objective, dim, jmax)) jmax = j;
    }
    const double wj = weight64(objective, dim, jmax);
    if (!(wj > total - wj)) return -1;
    if (w_out) *w_out = wj;
    return jmax;
}

void column_weights(int objective, int dim, float *w) {
    for (int j = 0; j < dim; ++j) w[j] = (float)weight64(objective, dim, j);
    const int h = heavy_column(objective, dim, nullptr);
    if (h >= 0) w[h] = 0.0f;  // added in FP64 by the kernels (CompDev::heavy)
}

CompDev ProblemHost::view(int c, int precision) const {
    const CompHost &h = comp[c];
    const bool half = precision == SPEO_PREC_FP16 || precision == SPEO_PREC_FP16X3;
    CompDev d{};
    // The kernels' elliptic path is the weighted sphere with the table in colw.
    d.objective = weighted_sphere(h.objective) ? SPEO_OBJ_ELLIPTIC : h.objective;
    d.rate = h.rate;
    d.shift = h.shift32;
    d.B32 = h.B32;
    d.ldb32 = ldb32;
    for (int p = 0; p < 3; ++p) d.M16[p] = half ? h.Mh[p] : h.Mbf[p];
    d.ldm16 = ldm16;
    d.is_half = half ? 1 : 0;
    d.colw = h.colw ? h.colw : colw;
    d.B32s = h.B32s;
    for (int p = 0; p < 3; ++p) d.M16s[p] = half ? h.Mhs[p] : h.Mbfs[p];
    d.kp32 = round_up(dim, 32);
    d.heavy.j = heavy_column(h.objective, dim, &d.heavy.w);
    d.heavy.m64 = d.heavy.j >= 0 ? h.M64 + (size_t)d.heavy.j * dim : nullptr;
    d.heavy.o64 = h.shift64;
    d.heavy.o32 = h.shift32;
    d.heavy.rate = h.rate64;
    return d;
}

template <typename T>
static int upload(T **dst, const std::vector<T> &src) {
    SPEO_TRY_HIP(hipMalloc(reinterpret_cast<void **>(dst), sizeof(T) * src.size()));
    SPEO_TRY_HIP(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
    return SPEO_OK;
}

static void free_problem(ProblemHost &P) {
    for (int c = 0; c < kMaxComp; ++c) {
        CompHost &h = P.comp[c];
        (void)hipFree(h.shift32);
        (void)hipFree(h.shift64);
        (void)hipFree(h.B32);
        (void)hipFree(h.M64);
        (void)hipFree(h.B32s);
        for (int p = 0; p < 3; ++p) {
            (void)hipFree(h.Mbf[p]);
            (void)hipFree(h.Mh[p]);
            (void)hipFree(h.Mbfs[p]);
            (void)hipFree(h.Mhs[p]);
        }
        (void)hipFree(h.colw);
        h = CompHost{};
    }
    for (int c = 0; c < kMaxComp; ++c) {
        (void)hipFree(P.perm[c]);
        P.perm[c] = nullptr;
    }
    delete P.gen;
    P.gen = nullptr;
    P.generic = false;
    (void)hipFree(P.colw);
    (void)hipFree(P.params_dev);
    P.colw = nullptr;
    P.params_dev = nullptr;
}

static bool known_objective(int o) {
    return (o >= SPEO_OBJ_ELLIPTIC && o <= SPEO_OBJ_ESCAFFER6) ||
           (o >= SPEO_OBJ_HYBRID_1 && o <= SPEO_OBJ_HYBRID_6);
}

static bool is_identity(const double *M, int dim) {
    for (int i = 0; i < dim; ++i)
        for (int j = 0; j < dim; ++j)
            if (M[(size_t)i * dim + j] != (i == j ? 1.0 : 0.0)) return false;
    return true;
}

static int create_problem(int dim, int ncomp, const int *objectives, const double *shift,
                          const double *rotation, const int *shuffle, const double *sigma,
                          const double *bias, const double *lambda, const double *rate,
                          ProblemHost &P) {
    if (dim < 2 || ncomp < 1 || ncomp > kMaxComp || !objectives || !shift)
        return set_error(SPEO_ERR_ARG, "speo_problem_create: need dim >= 2, 1 <= ncomp <= 8, "
                                       "objectives and shift");
    bool generic = false;
    for (int c = 0; c < ncomp; ++c) {
        if (!known_objective(objectives[c]))
            return set_error(SPEO_ERR_ARG, "speo_problem_create: unknown objective");
        generic = generic || generic_objective(objectives[c]);
        if (objectives[c] >= SPEO_OBJ_HYBRID_1) {
            if (!shuffle)
                return set_error(SPEO_ERR_ARG, "speo_problem_create: a hybrid objective needs a shuffle");
            std::vector<char> seen((size_t)dim, 0);
            for (int j = 0; j < dim; ++j) {
                const int v = shuffle[(size_t)c * dim + j];
                if (v < 1 || v > dim || seen[(size_t)v - 1])
                    return set_error(SPEO_ERR_ARG, "speo_problem_create: the shuffle of a hybrid "
                                                   "component must be a permutation of 1..dim");
                seen[(size_t)v - 1] = 1;
            }
        }
    }
    if (generic && dim > kGenMaxDim)
        return set_error(SPEO_ERR_UNSUPPORTED, "speo_problem_create: these objectives take dim <= 1024");
    P.dim = dim;
    P.ncomp = ncomp;
    P.ldb32 = round_up(dim, 4);
    P.ldm16 = round_up(dim, 8);
    const size_t D = (size_t)dim;
    std::vector<float> w(D);
    column_weights(SPEO_OBJ_ELLIPTIC, dim, w.data());
    SPEO_TRY(upload(&P.colw, w));
    std::vector<CompParams> params((size_t)ncomp);
    for (int c = 0; c < ncomp; ++c) {
        CompHost &h = P.comp[c];
        h.objective = objectives[c];
        if (h.objective == SPEO_OBJ_BENT_CIGAR || h.objective == SPEO_OBJ_DISCUS) {
            column_weights(h.objective, dim, w.data());
            SPEO_TRY(upload(&h.colw, w));
        }
        h.rate = rate ? (float)rate[c] : 1.0f;
        h.rate64 = rate ? rate[c] : 1.0;
        h.sigma = sigma ? sigma[c] : 10.0;
        h.bias = bias ? bias[c] : 0.0;
        h.lambda = lambda ? lambda[c] : 1.0;
        std::vector<double> M(D * D, 0.0);
        if (rotation)
            std::memcpy(M.data(), rotation + (size_t)c * D * D, sizeof(double) * D * D);
        else
            for (size_t i = 0; i < D; ++i) M[i * D + i] = 1.0;
        h.unrotated = is_identity(M.data(), dim);
        std::vector<double> o(shift + (size_t)c * D, shift + (size_t)(c + 1) * D);
        std::vector<float> o32(D);
        for (size_t j = 0; j < D; ++j) o32[j] = (float)o[j];
        SPEO_TRY(upload(&h.shift32, o32));
        SPEO_TRY(upload(&h.shift64, o));
        SPEO_TRY(upload(&h.M64, M));
        std::vector<float> B((size_t)dim * P.ldb32, 0.f);
        for (size_t k = 0; k < D; ++k)
            for (size_t cc = 0; cc < D; ++cc) B[k * P.ldb32 + cc] = (float)M[cc * D + k];
        SPEO_TRY(upload(&h.B32, B));
        std::vector<uint16_t> pb[3], ph[3];
        for (int p = 0; p < 3; ++p) {
            pb[p].assign((size_t)dim * P.ldm16, 0);
            ph[p].assign((size_t)dim * P.ldm16, 0);
        }
        for (size_t r = 0; r < D; ++r)
            for (size_t k = 0; k < D; ++k) {
                double rb = M[r * D + k], rh = M[r * D + k] * (double)kF16ScaleB;
                for (int p = 0; p < 3; ++p) {
                    const uint16_t b = bf16_rne((float)rb);
                    const uint16_t hh = f16_rne((float)rh);
                    pb[p][r * P.ldm16 + k] = b;
                    ph[p][r * P.ldm16 + k] = hh;
                    rb -= bf16_val(b);
                    rh -= f16_val(hh);
                }
            }
        for (int p = 0; p < 3; ++p) {
            SPEO_TRY(upload(&h.Mbf[p], pb[p]));
            SPEO_TRY(upload(&h.Mh[p], ph[p]));
        }
        if (dim <= kSmallD) {
            const size_t kp32 = (size_t)round_up(dim, 32);
            // B = M^T zero padded to 128 x 128, in MFMA fragment order (CompDev::B32s)
            std::vector<float> Bs((size_t)kSmallD * kSmallD, 0.f);
            for (size_t t = 0; t < 8; ++t)
                for (size_t q = 0; q < 8; ++q)
                    for (size_t l = 0; l < 64; ++l)
                        for (size_t e = 0; e < 4; ++e) {
                            const size_t k = 16 * q + 4 * e + l / 16, cc = 16 * t + l % 16;
                            if (k < D && cc < D) Bs[((t * 8 + q) * 64 + l) * 4 + e] = (float)M[cc * D + k];
                        }
            SPEO_TRY(upload(&h.B32s, Bs));
            for (int p = 0; p < 3; ++p) {
                std::vector<uint16_t> sb((size_t)kSmallD * kp32, 0), sh((size_t)kSmallD * kp32, 0);
                for (size_t r = 0; r < D; ++r)
                    for (size_t k = 0; k < D; ++k) {
                        sb[r * kp32 + k] = pb[p][r * P.ldm16 + k];
                        sh[r * kp32 + k] = ph[p][r * P.ldm16 + k];
                    }
                SPEO_TRY(upload(&h.Mbfs[p], sb));
                SPEO_TRY(upload(&h.Mhs[p], sh));
            }
        }
        params[(size_t)c] = CompParams{h.objective, h.rate64, h.sigma, h.bias, h.lambda,
                                       h.shift32, h.shift64, h.M64};
    }
    SPEO_TRY(upload(&P.params_dev, params));
    if (generic) {
        P.generic = true;
        P.gen = new (std::nothrow) GenProblem{};
        if (!P.gen) return set_error(SPEO_ERR_ALLOC, "speo_problem_create: out of host memory");
        P.gen->dim = dim;
        P.gen->ncomp = ncomp;
        for (int c = 0; c < ncomp; ++c) {
            const CompHost &h = P.comp[c];
            GenComp &g = P.gen->comp[c];
            if (objectives[c] >= SPEO_OBJ_HYBRID_1) {
                std::vector<int> perm(shuffle + (size_t)c * dim, shuffle + (size_t)(c + 1) * dim);
                SPEO_TRY(upload(&P.perm[c], perm));
                hybrid_segments(objectives[c], dim, g);
                g.perm = P.perm[c];
            } else {
                g.nseg = 1;
                g.seg[0] = GenSeg{objectives[c], dim, 1.0};
            }
            g.rate = h.rate64;
            g.rotate = rotation && !is_identity(rotation + (size_t)c * D * D, dim) ? 1 : 0;
            g.o64 = h.shift64;
            g.M64 = h.M64;
            g.lambda = h.lambda;
            g.bias = h.bias;
            g.sigma = h.sigma;
        }
    }
    return SPEO_OK;
}

// ------------------------------------------------------------- generator

static double u53(uint32_t a, uint32_t b) {
    const uint64_t k = (uint64_t)(a >> 5) * (1ull << 26) + (uint64_t)(b >> 6) + 1ull;
    return (double)k * 0x1p-53;
}

// Householder QR of the row-major D x D matrix G; returns Q (row-major) with
// columns sign-fixed so that diag(R) > 0, which makes Q Haar-distributed.
static void haar_from_gauss(const std::vector<double> &G, int D, std::vector<double> &Q_out) {
    const size_t n = (size_t)D;
    std::vector<double> A(n * n), Q(n * n, 0.0), V(n * n, 0.0), rdiag(n);
    for (size_t i = 0; i < n; ++i)  // column-major copy: A[col*n + row]
        for (size_t j = 0; j < n; ++j) A[j * n + i] = G[i * n + j];
    for (size_t k = 0; k < n; ++k) {
        double *a = &A[k * n];
        if (k + 1 == n) {
            rdiag[k] = a[k];
            break;
        }
        double norm = 0.0;
        for (size_t i = k; i < n; ++i) norm += a[i] * a[i];
        norm = std::sqrt(norm);
        const double alpha = a[k] >= 0.0 ? -norm : norm;
        double *v = &V[k * n];
        for (size_t i = k; i < n; ++i) v[i] = a[i];
        v[k] -= alpha;
        double vn = 0.0;
        for (size_t i = k; i < n; ++i) vn += v[i] * v[i];
        vn = std::sqrt(vn);
        rdiag[k] = alpha;
        if (vn == 0.0) continue;
        for (size_t i = k; i < n; ++i) v[i] /= vn;
        for (size_t j = k; j < n; ++j) {
            double *col = &A[j * n];
            double dot = 0.0;
            for (size_t i = k; i < n; ++i) dot += v[i] * col[i];
            dot *= 2.0;
            for (size_t i = k; i < n; ++i) col[i] -= dot * v[i];
        }
    }
    for (size_t i = 0; i < n; ++i) Q[i * n + i] = 1.0;  // column-major identity
    for (size_t kk = n - 1; kk-- > 0;) {
        const double *v = &V[kk * n];
        for (size_t j = 0; j < n; ++j) {
            double *col = &Q[j * n];
            double dot = 0.0;
            for (size_t i = kk; i < n; ++i) dot += v[i] * col[i];
            dot *= 2.0;
            for (size_t i = kk; i < n; ++i) col[i] -= dot * v[i];
        }
    }
    Q_out.assign(n * n, 0.0);
    for (size_t j = 0; j < n; ++j) {
        const double s = rdiag[j] < 0.0 ? -1.0 : 1.0;
        for (size_t i = 0; i < n; ++i) Q_out[i * n + j] = s * Q[j * n + i];
    }
}

}  // namespace speo

using namespace speo;

extern "C" {

uint64_t speo_problem_seed(uint64_t base_seed, uint32_t index) {
    const U4 w = philox_block(base_seed, index, 0, 0, kPurposeProblem);
    return (uint64_t)w.x | ((uint64_t)w.y << 32);
}

int speo_generate_problem_data(uint64_t seed, int dim, double *shift_out, double *rotation_out) {
    if (dim < 1 || (!shift_out && !rotation_out))
        return set_error(SPEO_ERR_ARG, "speo_generate_problem_data: bad arguments");
    if (shift_out) {
        for (int j = 0; j < dim; ++j) {
            const U4 w = philox_block(seed, (uint32_t)(j / 2), (uint32_t)dim, 0, kPurposeShift);
            const double u = (j % 2 == 0) ? u53(w.x, w.y) : u53(w.z, w.w);
            shift_out[j] = (double)(float)(-80.0 + 160.0 * u);
        }
    }
    if (rotation_out) {
        const size_t n = (size_t)dim;
        std::vector<double> G(n * n), Q;
        const double two_pi = 6.283185307179586476925286766559;
        for (size_t e = 0; e < n * n; ++e) {
            const U4 w = philox_block(seed, (uint32_t)(e / 2), (uint32_t)dim, 0, kPurposeGauss);
            const double r = std::sqrt(-2.0 * std::log(u53(w.x, w.y)));
            const double th = two_pi * u53(w.z, w.w);
            G[e] = (e % 2 == 0) ? r * std::cos(th) : r * std::sin(th);
        }
        haar_from_gauss(G, dim, Q);
        std::memcpy(rotation_out, Q.data(), sizeof(double) * n * n);
    }
    return SPEO_OK;
}

int speo_problem_create(int dim, int ncomp, const int *objectives, const double *shift,
                        const double *rotation, const double *sigma, const double *bias,
                        const double *lambda, const double *rate, speo_problem *out) {
    return speo_problem_create_ex(dim, ncomp, objectives, shift, rotation, nullptr, sigma, bias,
                                  lambda, rate, out);
}

int speo_problem_create_ex(int dim, int ncomp, const int *objectives, const double *shift,
                           const double *rotation, const int *shuffle, const double *sigma,
                           const double *bias, const double *lambda, const double *rate,
                           speo_problem *out) {
    if (!out) return set_error(SPEO_ERR_ARG, "speo_problem_create: null output");
    *out = nullptr;
    auto *p = new (std::nothrow) speo_problem_s;
    if (!p) return set_error(SPEO_ERR_ALLOC, "speo_problem_create: out of host memory");
    const int rc = create_problem(dim, ncomp, objectives, shift, rotation, shuffle, sigma, bias,
                                  lambda, rate, p->h);
    if (rc != SPEO_OK) {
        free_problem(p->h);
        delete p;
        return rc;
    }
    *out = p;
    return SPEO_OK;
}

// CEC2014.cc:353-420: fscanf("%lf") over the matrix file; per component the
// first dim values of its shift line, the rest of the line skipped.
int speo_load_cec_data(const char *dir, int func_id, int dim, int ncomp, double *shift_out,
                       double *rotation_out) {
    if (!dir || dim < 1 || ncomp < 1 || ncomp > kMaxComp)
        return set_error(SPEO_ERR_ARG, "speo_load_cec_data: bad arguments");
    const std::string base(dir);
    if (rotation_out) {
        const std::string name = base + "/M_" + std::to_string(func_id) + "_D" + std::to_string(dim) + ".txt";
        FILE *f = std::fopen(name.c_str(), "r");
        if (!f) return set_error(SPEO_ERR_ARG, "speo_load_cec_data: cannot open " + name);
        const size_t n = (size_t)ncomp * dim * dim;
        size_t got = 0;
        while (got < n && std::fscanf(f, "%lf", &rotation_out[got]) == 1) ++got;
        std::fclose(f);
        if (got < n) return set_error(SPEO_ERR_ARG, "speo_load_cec_data: too few values in " + name);
    }
    if (shift_out) {
        const std::string name = base + "/shift_data_" + std::to_string(func_id) + ".txt";
        FILE *f = std::fopen(name.c_str(), "r");
        if (!f) return set_error(SPEO_ERR_ARG, "speo_load_cec_data: cannot open " + name);
        bool ok = true;
        for (int c = 0; c < ncomp && ok; ++c) {
            for (int j = 0; j < dim && ok; ++j) ok = std::fscanf(f, "%lf", &shift_out[(size_t)c * dim + j]) == 1;
            if (ok && c + 1 < ncomp) {
                int ch;
                while ((ch = std::fgetc(f)) != EOF && ch != '\n') {
                }
            }
        }
        std::fclose(f);
        if (!ok) return set_error(SPEO_ERR_ARG, "speo_load_cec_data: too few values in " + name);
    }
    return SPEO_OK;
}

// The shuffle of a hybrid function (CEC2014.cc:440-458, CEC2014_CUDA.cc:
// 275-279): <dir>/shuffle_data_<func_id>_D<dim>.txt, ncomp * dim 1-based indices.
static int load_cec_shuffle(const char *dir, int func_id, int dim, int ncomp, int *out) {
    const std::string name = std::string(dir) + "/shuffle_data_" + std::to_string(func_id) + "_D" +
                             std::to_string(dim) + ".txt";
    FILE *f = std::fopen(name.c_str(), "r");
    if (!f) return set_error(SPEO_ERR_ARG, "speo_problem_create_cec: cannot open " + name);
    const size_t n = (size_t)ncomp * dim;
    size_t got = 0;
    while (got < n && std::fscanf(f, "%d", &out[got]) == 1) ++got;
    std::fclose(f);
    if (got < n) return set_error(SPEO_ERR_ARG, "speo_problem_create_cec: too few values in " + name);
    return SPEO_OK;
}

int speo_problem_create_cec(const char *dir, int func_id, int dim, speo_problem *out) {
    if (!out) return set_error(SPEO_ERR_ARG, "speo_problem_create_cec: null output");
    *out = nullptr;
    int nc = 0;
    CecComp cc[kMaxComp];
    if (!cec_function(func_id, &nc, cc))
        return set_error(SPEO_ERR_ARG, "speo_problem_create_cec: function " + std::to_string(func_id) +
                                           " is not in 1..30");
    if (dim < 2) return set_error(SPEO_ERR_ARG, "speo_problem_create_cec: dim < 2");
    const size_t D = (size_t)dim;
    std::vector<double> o(nc * D), M(nc * D * D);
    bool any_rot = false, any_hybrid = false;
    for (int c = 0; c < nc; ++c) {
        any_rot = any_rot || cc[c].rotate;
        any_hybrid = any_hybrid || cc[c].objective >= SPEO_OBJ_HYBRID_1;
    }
    SPEO_TRY(speo_load_cec_data(dir, func_id, dim, nc, o.data(), any_rot ? M.data() : nullptr));
    std::vector<int> S(any_hybrid ? nc * D : 0);
    if (any_hybrid) SPEO_TRY(load_cec_shuffle(dir, func_id, dim, nc, S.data()));
    int objs[kMaxComp];
    double rate[kMaxComp], sigma[kMaxComp], bias[kMaxComp], lam[kMaxComp];
    for (int c = 0; c < nc; ++c) {
        objs[c] = cc[c].objective;
        rate[c] = cc[c].rate;
        sigma[c] = cc[c].sigma;
        bias[c] = cc[c].bias;
        lam[c] = cc[c].lambda;
        if (!cc[c].rotate) {  // sr_func with r_flag = 0: no rotation for this component
            std::fill(M.begin() + c * D * D, M.begin() + (c + 1) * D * D, 0.0);
            for (size_t i = 0; i < D; ++i) M[c * D * D + i * D + i] = 1.0;
        }
    }
    return speo_problem_create_ex(dim, nc, objs, o.data(), any_rot ? M.data() : nullptr,
                                  any_hybrid ? S.data() : nullptr, sigma, bias, lam, rate, out);
}

int speo_problem_destroy(speo_problem p) {
    if (!p) return SPEO_OK;
    (void)hipDeviceSynchronize();
    free_problem(p->h);
    (void)hipFree(p->scratch);
    delete p;
    return SPEO_OK;
}

int speo_problem_dim(speo_problem p) { return p ? p->h.dim : 0; }
int speo_problem_ncomp(speo_problem p) { return p ? p->h.ncomp : 0; }

int speo_evaluate(speo_problem p, const float *X, int ld, int n, float *f_out, int precision,
                  speo_stream_t stream) {
    if (!p || !X || !f_out || n < 0 || ld < p->h.dim)
        return set_error(SPEO_ERR_ARG, "speo_evaluate: bad arguments");
    if (precision < SPEO_PREC_FP32 || precision > SPEO_PREC_BF16X6)
        return set_error(SPEO_ERR_ARG, "speo_evaluate: unknown precision");
    const size_t need = eval_scratch_floats(p->h.ncomp, p->h.dim, n);
    if (need > p->scratch_cap) {
        SPEO_TRY_HIP(hipStreamSynchronize(as_stream(stream)));
        (void)hipFree(p->scratch);
        p->scratch = nullptr;
        p->scratch_cap = 0;
        SPEO_TRY_HIP(hipMalloc(&p->scratch, sizeof(float) * need));
        p->scratch_cap = need;
    }
    return eval_problem(p->h, X, ld, n, f_out, p->scratch, precision,
                        default_rows_per_block(n, p->h.dim, precision), as_stream(stream));
}

int speo_evaluate_f64(speo_problem p, const float *X, int ld, int n, double *f_out,
                      speo_stream_t stream) {
    if (!p || !X || !f_out || n < 0 || ld < p->h.dim)
        return set_error(SPEO_ERR_ARG, "speo_evaluate_f64: bad arguments");
    return eval_problem_f64(p->h, X, ld, n, f_out, as_stream(stream));
}

}  // extern "C"
