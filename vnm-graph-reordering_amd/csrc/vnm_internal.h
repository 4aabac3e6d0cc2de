// Internal declarations shared by the host C++ and HIP sources of libvnm.
#pragma once

#include "../../include/vnm.h"

#include <cstdint>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace vnm {

// Raw tuples as a reader or generator produced them.
struct Tuples {
    int32_t nrows = 0, ncols = 0;
    bool symmetric = false;
    std::vector<int32_t> row, col;
    std::vector<float> val;
};

// Host COO in the reference's order.
struct HostCoo {
    int32_t nrows = 0, ncols = 0;
    std::vector<int32_t> row, col;
    std::vector<float> val;
};

// errors (thread-local message)
int set_error(int code, const char *fmt, ...);
const char *last_error();

// host I/O (mtx.cpp)
// threads: worker threads of the host passes (0 = the machine's, at most 16)
int host_threads(int threads);
int read_mtx_tuples(const char *path, Tuples &t, int threads);
void loader_order(Tuples &t, int M, HostCoo &out, int threads);
int write_mtx(const char *path, const HostCoo &c, int threads);
int banner_symmetric(const char *path, bool *sym);
std::string cache_name(const char *path, bool symmetric);
bool cache_load(const std::string &name, const struct stat &src, HostCoo &c);
void cache_store(const std::string &name, const struct stat &src, bool symmetric, const HostCoo &c);
int coo_export(const HostCoo &c, vnm_coo *out);
void coo_import(const vnm_coo *in, HostCoo &c);
float quant_cpp_write(float v);
float quant_py_roundtrip(float v);

// generators (gen.cpp)
int gen_er(int64_t n, int64_t k, uint64_t seed, float lo, float hi, Tuples &out);
int gen_rmat(int scale, int64_t undirected_edges, double a, double b, double c,
             uint64_t seed, Tuples &out);
int gen_chung_lu(int64_t n, double gamma, double mean_degree, uint64_t seed, Tuples &out);
void gen_normal(uint64_t seed, int64_t count, float *out);

}  // namespace vnm
