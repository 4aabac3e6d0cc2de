/*
 * Driver around the reference's host-only headers, compiled from the
 * reference sources where they lie (see oracle/refbuild/Makefile).  TEST
 * INFRASTRUCTURE ONLY: it produces golden vectors for the oracle.
 *
 *   ref_driver load <file.mtx> <M>   -> "nrows ncols nnz" then one
 *                                        "row col %a" line per entry, as
 *                                        readMtx + Mtx::init_from_file leave
 *                                        the host COO (mtx/readMtx.hpp:354-425,
 *                                        mtx/Mtx.h:98-112)
 *   ref_driver pq <M>                 -> reads score vectors (2*M*M ints per
 *                                        line) from stdin and prints the pop
 *                                        order of the SwapPairList built as in
 *                                        reorder/reorder.h:521-532
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <iostream>
#include <queue>
#include <string>
#include <vector>

#include "mtx/readMtx.hpp"
#include "reorder/swappair.h"

typedef std::priority_queue<SwapPair *, std::vector<SwapPair *>,
                            std::function<bool(SwapPair *, SwapPair *)>> PQ;

int main(int argc, char **argv) {
    if (argc >= 4 && !strcmp(argv[1], "load")) {
        std::vector<int> row, col;
        std::vector<float> val;
        int nrows, ncols, nnz;
        readMtx(argv[2], &row, &col, &val, &nrows, &ncols, &nnz, 0, false);
        int aligned = atoi(argv[3]);
        if (nrows % aligned != 0) {  /* the padding step of Mtx::init_from_file */
            nrows = ((nrows + aligned - 1) / aligned) * aligned;
            ncols = ((ncols + aligned - 1) / aligned) * aligned;
        }
        printf("%d %d %zu\n", nrows, ncols, val.size());
        for (size_t i = 0; i < val.size(); ++i) printf("%d %d %a\n", row[i], col[i], (double)val[i]);
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "pq")) {
        int M = atoi(argv[2]);
        std::vector<int> s(2 * M * M);
        while (true) {
            for (int i = 0; i < 2 * M * M; ++i)
                if (scanf("%d", &s[i]) != 1) return 0;
            std::vector<SwapPair> store;
            store.reserve(M * M);
            PQ pq(Comparator);
            for (int i = 0; i < M * M; ++i) {
                store.emplace_back(i / M, M + i % M, s[2 * i], s[2 * i + 1]);
                pq.push(&store.back());
            }
            bool first = true;
            while (!pq.empty()) {
                SwapPair *p = pq.top();
                int idx = (int)(p - &store[0]);
                printf(first ? "%d" : " %d", idx);
                first = false;
                pq.pop();
            }
            printf("\n");
        }
    }
    fprintf(stderr, "usage: ref_driver load <file> <M> | pq <M>\n");
    return 2;
}
