// Test-only: prints Philox4x32-10 outputs from the open-source rocRAND engine
// (rocrand_philox4x32_10.h, compiled for the host) for counters/keys read
// from stdin, so the oracle's Philox can be compared word for word.
#include <cstdio>
#include <hip/hip_runtime.h>
#include <rocrand/rocrand_kernel.h>

struct Engine : rocrand_device::philox4x32_10_engine {
    uint4 rounds(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    Engine e;
    while (std::scanf("%u %u %u %u %u %u", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = e.rounds(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%u %u %u %u\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
