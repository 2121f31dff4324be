// ref_driver.cc -- drives the reference's CPU objective code
// (CUDA_EA_library/CEC2014.cc, compiled from its own source by oracle/Makefile)
// for validation and for the CPU baseline.  TEST INFRASTRUCTURE ONLY.
//
// The reference loads its shift vector and rotation matrix from
// input_data/shift_data_<id>.txt and input_data/M_<id>_D<dim>.txt relative to
// the working directory (CEC2014.cc:353-420); callers run this program from a
// directory holding those files (the reference's own, or ones written in the
// same text format).
//
//   eval:  ref_driver eval <func_id> <dim> <n>      < n*dim float64 > n float64
//          (CEC2014::EvaluateFitness(double*), CEC2014.cc:14-145)
//   time:  ref_driver time <func_id> <dim> <seconds> < 1*dim float64
//          prints "<evals> <seconds>" after evaluating the point repeatedly.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "CEC2014.h"

int main(int argc, char **argv) {
    if (argc != 5) {
        std::fprintf(stderr, "usage: %s eval|time <func_id> <dim> <n|seconds>\n", argv[0]);
        return 2;
    }
    const char *mode = argv[1];
    int fid = std::atoi(argv[2]);
    int dim = std::atoi(argv[3]);
    CEC2014 cec;
    cec.Initilize(fid, dim);
    if (std::strcmp(mode, "eval") == 0) {
        int n = std::atoi(argv[4]);
        std::vector<double> x((size_t)dim);
        for (int p = 0; p < n; ++p) {
            if (std::fread(x.data(), sizeof(double), (size_t)dim, stdin) != (size_t)dim) return 3;
            double f = cec.EvaluateFitness(x.data());
            std::fwrite(&f, sizeof(double), 1, stdout);
        }
    } else {
        double budget = std::atof(argv[4]);
        std::vector<double> x((size_t)dim);
        if (std::fread(x.data(), sizeof(double), (size_t)dim, stdin) != (size_t)dim) return 3;
        long evals = 0;
        double sink = 0.0;
        auto t0 = std::chrono::steady_clock::now();
        double el = 0.0;
        do {
            for (int r = 0; r < 64; ++r) {
                x[(size_t)(evals % dim)] += 1e-9;  // defeat hoisting of the call
                sink += cec.EvaluateFitness(x.data());
                ++evals;
            }
            el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        } while (el < budget);
        std::printf("%ld %.6f %g\n", evals, el, sink);
    }
    cec.Unitilize();
    return 0;
}
