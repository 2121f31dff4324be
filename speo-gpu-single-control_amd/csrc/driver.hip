// driver.hip -- the reference's island driver on one GPU: the function x run
// sweep, the evaluation-budget loop with measured time and communication, and
// the result file (src/main.cc:124-150; src/island_EA.cc:13-171).  Host C++
// over the library's own C ABI.
#include <chrono>
#include <cstdio>
#include <string>
#include <vector>

#include "common.h"

namespace {

const char *regroup_name(int o) {
    static const char *const n[4] = {"static_and_ordered", "static_and_random", "dynamic_and_random",
                                     "dynamic_and_ordered"};
    return (o >= 0 && o < 4) ? n[o] : "static_and_ordered";
}

// IslandEA::CheckAndCreatRecordFile (island_EA.cc:13-31): the header line
// when the file does not exist yet.
int open_record(const char *path) {
    if (!path) return SPEO_OK;
    if (FILE *f = std::fopen(path, "r")) {
        std::fclose(f);
        return SPEO_OK;
    }
    FILE *f = std::fopen(path, "w");
    if (!f) return speo::set_error(SPEO_ERR_ARG, std::string("speo_island_driver: cannot create ") + path);
    std::fprintf(f, "%s\n", speo_result_header());
    std::fclose(f);
    return SPEO_OK;
}

struct Guard {
    speo_problem p = nullptr;
    speo_engine e = nullptr;
    ~Guard() {
        if (e) speo_engine_destroy(e);
        if (p) speo_problem_destroy(p);
    }
};

}  // namespace

extern "C" int speo_island_driver(const speo_driver_config *c, speo_run_record *rows_out) {
    using speo::set_error;
    if (!c || !c->data_dir || !c->function_IDs || !c->run_IDs || c->num_functions < 0 ||
        c->num_runs < 0 || c->island_num < 1 || c->island_size < 8 || c->max_base_FEs < 0 ||
        c->dim < 2)
        return set_error(SPEO_ERR_ARG, "speo_island_driver: bad configuration");
    SPEO_TRY(open_record(c->record_file));
    static char params[64];
    int row = 0;
    for (int fi = 0; fi < c->num_functions; ++fi) {
        for (int ri = 0; ri < c->num_runs; ++ri, ++row) {
            const int fid = c->function_IDs[fi], run = c->run_IDs[ri];
            Guard g;
            SPEO_TRY(speo_problem_create_cec(c->data_dir, fid, c->dim, &g.p));
            std::vector<int> sizes((size_t)c->island_num, c->island_size), pidx((size_t)c->island_num, 0);
            speo_engine_config ec{};
            ec.num_islands = c->island_num;
            ec.island_size = sizes.data();
            ec.island_problem = pidx.data();
            ec.num_problems = 1;
            ec.problems = &g.p;
            ec.F = 0.5f;   // DE_CUDA::Initilize, DE_CUDA.cc:61-63
            ec.CR = 0.9f;
            ec.min_bound = c->min_bound;
            ec.max_bound = c->max_bound;
            ec.seed = speo_run_seed(c->node_ID, fid, run);  // main.cc:138
            ec.interval = c->interval;
            ec.migration_size = c->migration_size;
            ec.precision = c->precision;
            ec.mode = c->mode;
            ec.use_graph = c->use_graph;
            ec.subisland_size = c->subisland_size;
            ec.configure_EA = c->configure_EA;
            ec.regroup_option = c->regroup_option;
            SPEO_TRY(speo_engine_create(&ec, nullptr, &g.e));
            // island_EA.cc:82-86: every island's share of max_base_FEs * dim
            const int64_t total_FEs = (int64_t)c->max_base_FEs * c->dim;
            int gens = 0;
            double secs = 0.0, comm = 0.0;
            SPEO_TRY(speo_engine_run_fes_timed(g.e, total_FEs, &gens, &secs, &comm));
            std::vector<double> f64((size_t)c->island_num);
            SPEO_TRY(speo_engine_best(g.e, nullptr, nullptr, f64.data()));
            double best = f64[0];  // MergeResults: the smallest over the islands
            for (double v : f64) best = v < best ? v : best;
            speo_run_record r{};
            r.function_ID = fid;
            r.run_ID = run;
            r.dim = c->dim;
            r.best_fitness = best;
            r.time_period = secs;
            r.communication_percentage = secs > 0.0 ? comm / secs : 0.0;
            r.total_FEs = total_FEs;
            r.node_num = c->node_num;
            r.GPU_num = c->GPU_num;
            r.island_num = c->island_num;
            r.total_pop_size = c->island_num * c->island_size;
            // PrintResult: -1 under the dynamic regroups (island_EA.cc:162-164)
            const bool dyn = c->regroup_option == SPEO_REGROUP_DYNAMIC_RANDOM ||
                             c->regroup_option == SPEO_REGROUP_DYNAMIC_ORDERED;
            r.subisland_size = dyn ? -1 : (c->subisland_size > 0 ? c->subisland_size : c->island_size);
            r.interval = c->interval;
            r.migration_size = c->migration_size;
            r.migration_rate = c->migration_rate;
            r.buffer_capacity = c->buffer_capacity;
            r.migration_topology = "ring";
            r.buffer_manage = "none";
            r.regroup_option = regroup_name(c->regroup_option);
            r.configure_EA = c->configure_EA == SPEO_CONFIGURE_DYNAMIC ? "dynamic" : "constant";
            const int plen = speo_de_parameters(ec.CR, ec.F, 4, params, (int)sizeof(params));
            if (plen < 0) return plen;
            r.EA_parameters = params;
            if (c->record_file) {
                char line[1024];
                const int n = speo_result_record(&r, line, (int)sizeof(line));
                if (n < 0) return n;
                FILE *f = std::fopen(c->record_file, "a");
                if (!f) return set_error(SPEO_ERR_ARG, std::string("speo_island_driver: cannot append to ") +
                                                           c->record_file);
                std::fprintf(f, "%s\n", line);
                std::fclose(f);
            }
            if (rows_out) rows_out[row] = r;
        }
    }
    return SPEO_OK;
}
