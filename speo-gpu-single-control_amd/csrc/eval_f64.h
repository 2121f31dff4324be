// eval_f64.h -- FP64 objective of one row by one 256-thread block
// (re-verification path), following the reference's FP64 CPU forms:
// shift/rotate CEC2014.cc:1373-1393, elliptic 481-497, bent cigar 499-516,
// discus 518-530, Ackley 590-612, Rastrigin 668-683, Schwefel 705-738,
// HappyCat 907-929, HGBat
// 931-953, composition weights cf_cal 1481-1519.
#pragma once

#include "problem.h"

namespace speo {

// `lds` holds dim + 16 doubles.  All 256 threads call it; the value is
// returned on every thread (valid on thread 0 for certain).
__device__ inline double eval_row_f64(const CompParams *comps, int ncomp, int dim,
                                      const float *x, double *lds) {
    double *a = lds;
    double *red = lds + dim;  // [4][3]
    __shared__ double bcast;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const double PI = 3.1415926535897932384626433832795029;
    const double E = 2.7182818284590452353602874713526625;
    double fit[kMaxComp], d2[kMaxComp];
    for (int c = 0; c < ncomp; ++c) {
        const CompParams cp = comps[c];
        __syncthreads();
        for (int j = tid; j < dim; j += 256) a[j] = ((double)x[j] - cp.shift64[j]) * cp.rate;
        __syncthreads();
        double s1 = 0.0, s2 = 0.0, dd = 0.0;
        for (int i = wave; i < dim; i += 4) {
            double z = 0.0;
            const double *m = cp.M64 + (size_t)i * dim;
            for (int j = lane; j < dim; j += 64) z = fma(m[j], a[j], z);
            for (int off = 32; off > 0; off >>= 1) z += __shfl_xor(z, off, 64);
            if (lane == 0) {
                if (cp.objective == SPEO_OBJ_ELLIPTIC) {
                    s1 += pow(10.0, 6.0 * i / (dim - 1)) * z * z;
                } else if (cp.objective == SPEO_OBJ_BENT_CIGAR) {
                    s1 += (i == 0 ? 1.0 : 1.0e6) * z * z;
                } else if (cp.objective == SPEO_OBJ_DISCUS) {
                    s1 += (i == 0 ? 1.0e6 : 1.0) * z * z;
                } else if (cp.objective == SPEO_OBJ_SCHWEFEL) {
                    const double v = z + 4.209687462275036e+002;
                    if (v > 500.0) {
                        const double m = 500.0 - fmod(v, 500.0);
                        s1 -= m * sin(sqrt(m));
                        s1 += (v - 500.0) / 100.0 * ((v - 500.0) / 100.0) / dim;
                    } else if (v < -500.0) {
                        const double m = 500.0 - fmod(fabs(v), 500.0);
                        s1 += m * sin(sqrt(m));
                        s1 += (v + 500.0) / 100.0 * ((v + 500.0) / 100.0) / dim;
                    } else {
                        s1 -= v * sin(sqrt(fabs(v)));
                    }
                } else if (cp.objective == SPEO_OBJ_HAPPYCAT || cp.objective == SPEO_OBJ_HGBAT) {
                    s1 += (z - 1.0) * (z - 1.0);
                    s2 += z - 1.0;
                } else if (cp.objective == SPEO_OBJ_RASTRIGIN) {
                    s1 += (z * z - 10.0 * cos(2.0 * PI * z) + 10.0);
                } else {
                    s1 += z * z;
                    s2 += cos(2.0 * PI * z);
                }
            }
        }
        for (int j = tid; j < dim; j += 256) {
            const double v = (double)x[j] - cp.shift64[j];
            dd += v * v;
        }
        for (int off = 32; off > 0; off >>= 1) dd += __shfl_xor(dd, off, 64);
        __syncthreads();
        if (lane == 0) {
            red[wave * 3 + 0] = s1;
            red[wave * 3 + 1] = s2;
            red[wave * 3 + 2] = dd;
        }
        __syncthreads();
        double S1 = 0.0, S2 = 0.0, DD = 0.0;
        for (int w = 0; w < 4; ++w) {
            S1 += red[w * 3 + 0];
            S2 += red[w * 3 + 1];
            DD += red[w * 3 + 2];
        }
        double g;
        if (cp.objective == SPEO_OBJ_ACKLEY) {
            const double r = -0.2 * sqrt(S1 / (double)dim);
            g = E - 20.0 * exp(r) + 20.0 - exp(S2 / (double)dim);
        } else if (cp.objective == SPEO_OBJ_SCHWEFEL) {
            g = S1 + 4.189828872724338e+002 * dim;
        } else if (cp.objective == SPEO_OBJ_HAPPYCAT) {
            g = pow(fabs(S1 - dim), 0.25) + (0.5 * S1 + S2) / dim + 0.5;
        } else if (cp.objective == SPEO_OBJ_HGBAT) {
            g = pow(fabs(S1 * S1 - S2 * S2), 0.5) + (0.5 * S1 + S2) / dim + 0.5;
        } else {
            g = S1;
        }
        fit[c] = cp.lambda * g + cp.bias;
        d2[c] = DD;
    }
    double f;
    if (ncomp == 1) {
        f = fit[0];
    } else {
        double w[kMaxComp], wmax = 0.0, wsum = 0.0;
        for (int c = 0; c < ncomp; ++c) {
            const double s = comps[c].sigma;
            w[c] = d2[c] != 0.0 ? sqrt(1.0 / d2[c]) * exp(-d2[c] / 2.0 / dim / (s * s)) : 1.0e99;
            wmax = fmax(wmax, w[c]);
        }
        for (int c = 0; c < ncomp; ++c) wsum += w[c];
        if (wmax == 0.0) {
            for (int c = 0; c < ncomp; ++c) w[c] = 1.0;
            wsum = ncomp;
        }
        f = 0.0;
        for (int c = 0; c < ncomp; ++c) f += w[c] / wsum * fit[c];
    }
    __syncthreads();
    if (tid == 0) bcast = f;
    __syncthreads();
    return bcast;
}

}  // namespace speo
