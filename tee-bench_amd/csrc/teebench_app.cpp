// Plain harness with the reference's command line (native.cpp:48-130,
// App/Lib/commons.c:13-198): generates R and S with the reference's
// generators and seeds, runs the selected join on the MI355X and prints the
// reference's log lines, ending with "Matches = N".
//
//   -a ALG   RHO|RHT|RJ|RHO_atomic|PHT|NPO_st|CHT|RSM|PSM|MWAY (default PHT)
//   -n N     threads (accepted; the GPU path does not use it)
//   -r/-s N  tuples of R / S          -x/-y MB  size of R / S in MB
//   -d cache-fit|cache-exceed|L       -l SEL    selectivity of S in %
//   -z SKEW  Zipf factor of S         -t/-u F   relations from files
//   --sort-r / --sort-s               sort R / S by key before the join
//   --learned                         learned CDF partitioner (GPU option)
//   --materialize                     materialise (key, Rpayload, Spayload)
#include <getopt.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/tbj_gen.h"
#include "../../include/teebench/joins.h"
#include "teebench_log.h"

namespace {

struct Args {
    algorithm_t* algorithm = &teebench_algorithms[0];
    uint64_t r_size = 2097152, s_size = 2097152;
    uint32_t r_seed = 11111, s_seed = 22222;
    int nthreads = 2, selectivity = 100;
    double skew = 0;
    int sort_r = 0, sort_s = 0, learned = 0, materialize = 0;
    std::string r_path, s_path;
};

void usage_exit(int code) { exit(code); }

Args parse(int argc, char** argv) {
    Args a;
    static int sort_r = 0, sort_s = 0, learned = 0, materialize = 0;
    static struct option longopts[] = {{"sort-r", no_argument, &sort_r, 1},
                                       {"sort-s", no_argument, &sort_s, 1},
                                       {"learned", no_argument, &learned, 1},
                                       {"materialize", no_argument, &materialize, 1},
                                       {"r-path", required_argument, 0, 't'},
                                       {"s-path", required_argument, 0, 'u'},
                                       {0, 0, 0, 0}};
    bool def_dataset = false, def_table = false;
    int c, idx;
    while ((c = getopt_long(argc, argv, "a:c:d:e:l:n:r:s:t:u:x:y:z:hv", longopts, &idx)) != -1) {
        switch (c) {
            case 0:
                break;
            case 'a': {
                algorithm_t* found = nullptr;
                for (algorithm_t* t = teebench_algorithms; t->name[0]; t++)
                    if (!strcmp(optarg, t->name)) found = t;
                if (!found) {
                    // the reference exits with success here (commons.c:67-70)
                    teebench_log(TB_ERROR, "Join algorithm not found: %s", optarg);
                    exit(EXIT_SUCCESS);
                }
                a.algorithm = found;
                break;
            }
            case 'd':
                if (def_table) {
                    teebench_log(TB_ERROR, "Select a predefined dataset OR specify tables sizes");
                    exit(EXIT_FAILURE);
                }
                def_dataset = true;
                if (!strcmp(optarg, "cache-fit")) {
                    a.r_size = (10ull << 20) / 8;
                    a.s_size = (40ull << 20) / 8;
                } else if (!strcmp(optarg, "cache-exceed")) {
                    a.r_size = (100ull << 20) / 8;
                    a.s_size = (400ull << 20) / 8;
                } else if (!strcmp(optarg, "L")) {
                    a.r_size = 50000000;
                    a.s_size = 200000000;
                } else {
                    teebench_log(TB_ERROR, "Unrecognized dataset: %s", optarg);
                    exit(EXIT_FAILURE);
                }
                break;
            case 'h':
                usage_exit(EXIT_SUCCESS);
                break;
            case 'l':
                a.selectivity = atoi(optarg);
                if (a.selectivity < 0 || a.selectivity > 100) {
                    teebench_log(TB_ERROR, "Selectivity should be between 0 and 100");
                    exit(EXIT_FAILURE);
                }
                break;
            case 'n':
                a.nthreads = atoi(optarg);
                break;
            case 'r':
            case 's':
            case 'x':
            case 'y': {
                if (def_dataset) {
                    teebench_log(TB_ERROR, "Select a predefined dataset OR specify tables sizes");
                    exit(EXIT_FAILURE);
                }
                uint64_t v = strtoull(optarg, nullptr, 10);
                if (c == 'x' || c == 'y') v = (v << 20) / 8;
                if (c == 'r' || c == 'x') a.r_size = v;
                else a.s_size = v;
                def_table = true;
                break;
            }
            case 't':
                a.r_path = optarg;
                break;
            case 'u':
                a.s_path = optarg;
                break;
            case 'z':
                a.skew = atof(optarg);
                break;
            default:
                break;
        }
    }
    a.sort_r = sort_r;
    a.sort_s = sort_s;
    a.learned = learned;
    a.materialize = materialize;
    return a;
}

// create_relation_from_file (generator.cpp:290-351) through the library's
// loader: one integer key per line, payload 0, shuffled with the seeded
// rand() stream.
bool from_file(const std::string& path, uint32_t seed, std::vector<tbj_tuple>& out) {
    const int64_t n = tbj_gen_from_file(seed, path.c_str(), nullptr, 0, TBJ_PAYLOAD_ZERO);
    if (n < 0) {
        teebench_log(TB_ERROR, "Could not open file %s", path.c_str());
        return false;
    }
    out.resize((size_t)n);
    tbj_gen_from_file(seed, path.c_str(), out.data(), out.size(), TBJ_PAYLOAD_ZERO);
    return true;
}

void sort_rel(std::vector<tbj_tuple>& v) {
    std::sort(v.begin(), v.end(),
              [](const tbj_tuple& a, const tbj_tuple& b) { return a.key < b.key; });
}

}  // namespace

int main(int argc, char** argv) {
    teebench_seconds_since_start();
    teebench_log(TB_INFO, "Welcome from native!");
    Args a = parse(argc, argv);
    teebench_set_partitioner(a.learned);
    teebench_set_materialize(a.materialize);
    teebench_log(TB_DBG, "Number of threads = %d (N/A for every algorithm)", a.nthreads);

    std::vector<tbj_tuple> R, S;
    if (!a.r_path.empty()) {
        teebench_log(TB_INFO, "Build relation R from file %s", a.r_path.c_str());
        if (!from_file(a.r_path, a.r_seed, R)) return EXIT_FAILURE;
    } else {
        teebench_log(TB_INFO, "Build relation R with size = %.2lf MB (%lu tuples)",
                     8.0 * a.r_size / (1 << 20), (unsigned long)a.r_size);
        R.resize(a.r_size);
        tbj_gen_pk(a.r_seed, a.r_size, R.data(), TBJ_PAYLOAD_ZERO, 0);
    }
    if (a.sort_r) sort_rel(R);
    if (!a.s_path.empty()) {
        teebench_log(TB_INFO, "Build relation S from file %s", a.s_path.c_str());
        if (!from_file(a.s_path, a.s_seed, S)) return EXIT_FAILURE;
    } else {
        teebench_log(TB_INFO, "Build relation S with size = %.2lf MB (%lu tuples)",
                     8.0 * a.s_size / (1 << 20), (unsigned long)a.s_size);
        S.resize(a.s_size);
        // native.cpp:107-118: -z takes precedence over -l
        if (a.skew > 0) {
            tbj_gen_zipf(a.s_seed, a.s_size, (uint32_t)R.size(), a.skew, S.data(),
                         TBJ_PAYLOAD_ZERO, 0);
        } else if (a.selectivity != 100) {
            teebench_log(TB_INFO, "Table S selectivity = %d", a.selectivity);
            const uint32_t maxid =
                a.selectivity != 0 ? (uint32_t)(100 * (uint64_t)R.size() / a.selectivity) : 0;
            tbj_gen_fk_sel(a.s_seed, a.s_size, maxid, S.data(), TBJ_PAYLOAD_ZERO, 0);
        } else {
            tbj_gen_fk(a.s_seed, a.s_size, (int64_t)R.size(), S.data(), TBJ_PAYLOAD_ZERO, 0);
        }
    }
    if (a.sort_s) sort_rel(S);

    relation_t tr{(row_t*)R.data(), R.size(), 0, a.sort_r};
    relation_t ts{(row_t*)S.data(), S.size(), 0, a.sort_s};
    teebench_log(TB_INFO, "Running algorithm %s", a.algorithm->name);
    const double t0 = teebench_seconds_since_start();
    result_t* res = a.algorithm->join(&tr, &ts, a.nthreads);
    teebench_log(TB_INFO, "Total join runtime: %.2fs", teebench_seconds_since_start() - t0);
    teebench_log(TB_INFO, "Matches = %lu", (unsigned long)res->totalresults);
    teebench_free_result(res);
    return 0;
}
