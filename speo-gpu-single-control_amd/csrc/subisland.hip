// subisland.hip -- the reference island's sub-island partition and per
// sub-island DE configuration, host side.
//
// Inside one island the reference splits the rows into subisland_num groups of
// equal size through a permutation (permutated_index); donors are drawn within
// a group and each group has its own DEInfo (CR, F, strategy).  At every
// migration point (generation % interval == 0) Migrate::RegroupIslands
// (src/migrate.cc:147-186) may redraw the group size (4 << RandIntUnif(0,
// max_index)), the permutation (Random::Permutate, src/random.cc:29-53, a
// partial Fisher-Yates) and, through DE_CUDA::ConfigureEA
// (CUDA_EA_library/DE_CUDA.cc:89-119), each group's CR in {0.1, 0.9} and
// strategy in {4, 6}.  The reference draws from the C library's rand(); here
// every draw is a Philox4x32-10 word of (seed, island, event), so the whole
// schedule is reproducible and restated by the oracle draw for draw.
#include <vector>

#include "common.h"

namespace speo {

namespace {

// The event's draw stream: word k of Philox block {k / 4, island, event,
// kPurposeRegroup}; RandIntUnif(a, b) = a + w % (b - a + 1), and no draw when
// a == b (src/random.cc:14-19).
struct HostDraws {
    uint64_t seed;
    uint32_t island, event, k = 0;
    int rand_int(int a, int b) {
        if (a == b) return a;
        const U4 w = philox_block(seed, k >> 2, island, event, kPurposeRegroup);
        const uint32_t v = word(w, k & 3u);
        ++k;
        return a + (int)(v % (uint32_t)(b - a + 1));
    }
};

// Random::Permutate(n, n): take a random remaining index, move the last
// remaining one into its place (src/random.cc:29-53).
void permutate(HostDraws &d, int n, int *out) {
    std::vector<int> g((size_t)n);
    for (int i = 0; i < n; ++i) g[(size_t)i] = i;
    int c = 0;
    for (int i = n; i > 0; --i) {
        const int t = d.rand_int(0, i - 1);
        out[c++] = g[(size_t)t];
        g[(size_t)t] = g[(size_t)i - 1];
    }
}

}  // namespace

}  // namespace speo

using namespace speo;

extern "C" int speo_island_regroup(uint64_t seed, int island, int event, int regroup_option,
                                   int configure_EA, int np, int subisland_num, int *perm,
                                   const DEInfo *base, DEInfo *info_out, int *subisland_num_out) {
    if (np < 4 || !perm || !base || !info_out || !subisland_num_out || island < 0 || event < 0 ||
        subisland_num < 1 || np % subisland_num || np / subisland_num < 4 ||
        regroup_option < SPEO_REGROUP_STATIC_ORDERED || regroup_option > SPEO_REGROUP_DYNAMIC_ORDERED ||
        (configure_EA != SPEO_CONFIGURE_CONSTANT && configure_EA != SPEO_CONFIGURE_DYNAMIC))
        return set_error(SPEO_ERR_ARG, "speo_island_regroup: bad arguments");
    HostDraws d{seed, (uint32_t)island, (uint32_t)event};
    int subs = subisland_num;
    bool configure = event == 0;  // DE_CUDA::Initilize calls ConfigureEA (DE_CUDA.cc:73)
    if (event > 0) {
        const bool dynamic = regroup_option == SPEO_REGROUP_DYNAMIC_RANDOM ||
                             regroup_option == SPEO_REGROUP_DYNAMIC_ORDERED;
        if (dynamic) {  // migrate.cc:151-159, 167-175
            int max_index = 0, p = np / 4;
            while (p > 1) {
                p /= 2;
                ++max_index;
            }
            const int size = 4 << d.rand_int(0, max_index);
            if (np % size)
                return set_error(SPEO_ERR_ARG, "speo_island_regroup: a dynamic regroup needs an "
                                               "island size of 4 * 2^k");
            subs = np / size;
            configure = true;
        }
        if (regroup_option == SPEO_REGROUP_DYNAMIC_RANDOM ||
            regroup_option == SPEO_REGROUP_STATIC_RANDOM)
            permutate(d, np, perm);
    }
    const int S = np / subs;
    for (int s = 0; s < subs; ++s) {
        DEInfo &o = info_out[s];
        if (configure && configure_EA == SPEO_CONFIGURE_DYNAMIC) {  // DE_CUDA.cc:92-103
            static const double CR[2] = {0.1, 0.9};
            static const int ST[2] = {4, 6};
            o.CR = (real)CR[d.rand_int(0, 1)];
            o.F = base->F;
            o.strategy_ID = ST[d.rand_int(0, 1)];
        } else if (configure) {  // "constant", DE_CUDA.cc:104-113
            o = *base;
        } else {                 // no ConfigureEA this event: the groups keep theirs
            o = info_out[s];
        }
        o.pop_size = S;
    }
    *subisland_num_out = subs;
    return SPEO_OK;
}
