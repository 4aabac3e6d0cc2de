/*
 * CPU oracle command line -- TEST INFRASTRUCTURE ONLY.
 *
 * Same flags as the reference binary (util/argparse.h:27-46) plus --m, --nn
 * (runtime M and N) and --sched-mode {auto,fixed}.  Prints the reference's
 * 16-field CSV line (main.cu:331-335); the last field is the wall time of the
 * timed region (main.cu:210-319) in milliseconds.
 */
#include "oracle.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

int main(int argc, char **argv) {
    std::string mtxfile, outmtxfile, mode = "auto";
    int maxiter = 0, n = 0, sched = 0, v = 0, M = 8, N = 2;
    for (int i = 1; i < argc - 1; ++i) {
        if (!strcmp(argv[i], "--mtxfile")) mtxfile = argv[i + 1];
        else if (!strcmp(argv[i], "--outmtxfile")) outmtxfile = argv[i + 1];
        else if (!strcmp(argv[i], "--maxiter")) maxiter = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--n")) n = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--sched")) sched = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--v")) v = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--m")) M = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--nn")) N = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--sched-mode")) mode = argv[i + 1];
    }
    (void)n;
    orc_coo coo{};
    int rc = orc_mtx_read(mtxfile.c_str(), M, &coo);
    if (rc != ORC_OK) {
        fprintf(stderr, "oracle: cannot load %s (status %d)\n", mtxfile.c_str(), rc);
        return 1;
    }
    if (coo.nnz == 0) { std::cout << std::endl; return 0; }
    if (mode == "auto") {
        int32_t best = 0;
        rc = orc_find_best_sched(&coo, N, M, 4, &best);
        if (rc != ORC_OK) { fprintf(stderr, "oracle: find_best failed %d\n", rc); return 1; }
        sched = best;
    }
    orc_hybrid_stats st;
    auto t0 = std::chrono::steady_clock::now();
    rc = orc_hybrid(&coo, N, M, v, maxiter, sched, &st);
    auto t1 = std::chrono::steady_clock::now();
    if (rc != ORC_OK) { fprintf(stderr, "oracle: hybrid failed %d\n", rc); return 1; }
    float ms = std::chrono::duration<float, std::milli>(t1 - t0).count();
    if (!outmtxfile.empty()) orc_mtx_write(outmtxfile.c_str(), &coo);
    std::cout << st.nrows << ", " << st.nnz << ", " << st.density << ", "
              << st.init_size << ", " << st.final_size << ", "
              << st.init_cnt << ", " << st.final_cnt << ", " << st.improve_rate << ", "
              << st.init_mb << ", " << st.final_mb << ", " << st.init_vnm << ", " << st.final_vnm << ", "
              << st.nm_rounds << ", " << st.vnm_rounds << ", " << st.total_rounds << ", " << ms
              << std::endl;
    orc_coo_free(&coo);
    return 0;
}
