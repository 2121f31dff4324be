// Host build of the library's donor draw and reflective bound (de.h,
// common.h), driven from stdin so that a test can compare them with the
// oracle on the CPU.  Input lines:
//   D <seed> <row> <gen> <S> <D> <n_inj> <inj_0> ... <inj_{n_inj-1}>
//   B <v as u32 bits> <lo> <hi>
//   T <CR as u32 bits>          (crossover threshold)
//   S <z as u32 bits>           (sin^2(pi z) of the objective terms)
//   N <x as u32 bits>           (sin x, 0 <= x <= 23, of the Schwefel terms)
// Output per line: "r0 r1 r2 r3 r4 jr draws", or one u32 (bits of a float, or
// the threshold).
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "de.h"
#include "eval_core.h"

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        std::istringstream in(line);
        std::string kind;
        in >> kind;
        if (kind == "D") {
            unsigned long long seed;
            uint32_t row, gen;
            int S, D, n;
            in >> seed >> row >> gen >> S >> D >> n;
            std::vector<float> inj((size_t)n);
            for (int i = 0; i < n; ++i) {
                uint32_t bits;
                in >> bits;
                std::memcpy(&inj[(size_t)i], &bits, 4);
            }
            speo::DonorStream ds(seed, row, gen);
            int r[speo::kMaxR], jr;
            speo::draw_donors(ds, n ? inj.data() : nullptr, n, S, D, r, jr);
            std::printf("%d %d %d %d %d %d %u\n", r[0], r[1], r[2], r[3], r[4], jr, ds.k);
        } else if (kind == "T") {
            uint32_t bits;
            in >> bits;
            float cr;
            std::memcpy(&cr, &bits, 4);
            std::printf("%u\n", speo::cr_threshold(cr));
        } else if (kind == "S") {
            uint32_t bits;
            in >> bits;
            float z;
            std::memcpy(&z, &bits, 4);
            const float o = speo::sin2pi(z);
            std::memcpy(&bits, &o, 4);
            std::printf("%u\n", bits);
        } else if (kind == "N") {
            uint32_t bits;
            in >> bits;
            float x;
            std::memcpy(&x, &bits, 4);
            const float o = speo::sin_small(x);
            std::memcpy(&bits, &o, 4);
            std::printf("%u\n", bits);
        } else if (kind == "B") {
            uint32_t bits;
            float lo, hi;
            in >> bits >> lo >> hi;
            float v;
            std::memcpy(&v, &bits, 4);
            const float o = speo::check_bound(v, lo, hi);
            std::memcpy(&bits, &o, 4);
            std::printf("%u\n", bits);
        }
    }
    return 0;
}
