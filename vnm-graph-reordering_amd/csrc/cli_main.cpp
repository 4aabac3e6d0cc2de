// Drop-in `spmm` command line (main.cu:339-352, util/argparse.h:6-60):
//   spmm --mtxfile IN --outmtxfile OUT --maxiter ROUNDS --n BCOLS --sched S --v V
// Same flags, defaults (strings empty, ints 0) and 16-field CSV line.  Extra
// flags: --m M and --nn N (the reference fixes N = 2, M = 8 at compile time,
// mtx/Mtx.h:15-16), --sched-mode auto|fixed (auto = find_best_2toX over
// schedules 0-3 as main.cu:348 does, ties to the lowest index; fixed uses
// --sched), --device D.  Errors print the library message and exit 1.
#include "../../include/vnm.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

static int fail(const char *what) {
    fprintf(stderr, "spmm: %s: %s\n", what, vnm_last_error());
    return 1;
}

int main(int argc, const char **argv) {
    std::string mtxfile, outmtxfile, mode = "auto";
    int maxiter = 0, n = 0, sched = 0, v = 0, M = 8, N = 2, device = 0;
    for (int i = 1; i < argc; ++i) {
        if (i == argc - 1) break;
        if (!strcmp(argv[i], "--mtxfile")) mtxfile = argv[i + 1];
        else if (!strcmp(argv[i], "--outmtxfile")) outmtxfile = argv[i + 1];
        else if (!strcmp(argv[i], "--maxiter")) maxiter = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--n")) n = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--sched")) sched = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--v")) v = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--m")) M = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--nn")) N = atoi(argv[i + 1]);
        else if (!strcmp(argv[i], "--sched-mode")) mode = argv[i + 1];
        else if (!strcmp(argv[i], "--device")) device = atoi(argv[i + 1]);
    }
    (void)n;  // parsed, unused (as in main.cu)
    vnm_coo coo{};
    if (vnm_mtx_read(mtxfile.c_str(), M, &coo) != VNM_OK) return fail("load");
    if (coo.nnz == 0) {  // main.cu:204
        std::cout << std::endl;
        vnm_coo_free(&coo);
        return 0;
    }
    vnm_ctx *ctx = nullptr;
    if (vnm_ctx_create(device, &ctx) != VNM_OK) return fail("device");
    vnm_graph *g = nullptr;
    if (vnm_graph_create(ctx, &coo, M, &g) != VNM_OK) return fail("upload");
    vnm_coo_free(&coo);
    if (mode == "auto") {
        int32_t best = 0;
        if (vnm_find_best_sched(g, N, 4, &best) != VNM_OK) return fail("find_best_2toX");
        sched = best;
    }
    vnm_hybrid_stats st;
    if (vnm_hybrid(g, N, v, maxiter, sched, &st) != VNM_OK) return fail("hybrid");
    if (!outmtxfile.empty()) {
        vnm_coo out{};
        if (vnm_graph_download(g, &out) != VNM_OK) return fail("download");
        if (vnm_mtx_write(outmtxfile.c_str(), &out) != VNM_OK) return fail("write");
        vnm_coo_free(&out);
    }
    std::cout << st.nrows << ", " << st.nnz << ", " << st.density << ", " << st.init_size << ", "
              << st.final_size << ", " << st.init_cnt << ", " << st.final_cnt << ", "
              << st.improve_rate << ", " << st.init_mb << ", " << st.final_mb << ", "
              << st.init_vnm << ", " << st.final_vnm << ", " << st.nm_rounds << ", "
              << st.vnm_rounds << ", " << st.total_rounds << ", " << st.milliseconds << std::endl;
    vnm_graph_destroy(g);
    vnm_ctx_destroy(ctx);
    return 0;
}
