// MatrixMarket reader/writer and the loader's ordering rules.
//
// Reader: banner and size line as mm_read_banner / mm_read_mtx_crd_size
// (mtx/mmio.hpp:128-228); entries as readTuples (mtx/readMtx.hpp:84-147):
// "real" values parse like fscanf("%f"), "integer" like "%d" then a float
// cast, "pattern" entries get 1.0.  The text is mapped once and parsed with
// strtol/strtof, which convert exactly as the scanf family does.
// Ordering: symmetric mirroring, (row, col) sort, duplicate removal and the
// padding of Mtx::init_from_file (mtx/readMtx.hpp:152-219,419; mtx/Mtx.h:105-112).
// Writer: Mtx::write_to_file (mtx/Mtx.h:412-435).
#include "vnm_internal.h"

#include <algorithm>
#include <cctype>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

namespace vnm {

namespace {
thread_local std::string g_err;
}

int set_error(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char *last_error() { return g_err.c_str(); }

namespace {

struct Mapped {
    const char *p = nullptr;
    size_t n = 0;
    int fd = -1;
    ~Mapped() {
        if (p && n) munmap((void *)p, n);
        if (fd >= 0) close(fd);
    }
};

// Skip whitespace (as scanf's %d/%f conversions do).
inline const char *skip_ws(const char *p, const char *e) {
    while (p < e && isspace((unsigned char)*p)) ++p;
    return p;
}

}  // namespace

int read_mtx_tuples(const char *path, Tuples &t) {
    Mapped m;
    m.fd = open(path, O_RDONLY);
    if (m.fd < 0) return set_error(VNM_E_IO, "cannot open %s", path);
    struct stat st;
    if (fstat(m.fd, &st) != 0) return set_error(VNM_E_IO, "cannot stat %s", path);
    m.n = (size_t)st.st_size;
    if (m.n == 0) return set_error(VNM_E_FORMAT, "%s: empty file", path);
    void *mp = mmap(nullptr, m.n, PROT_READ, MAP_PRIVATE, m.fd, 0);
    if (mp == MAP_FAILED) return set_error(VNM_E_IO, "cannot map %s", path);
    m.p = (const char *)mp;
    const char *p = m.p, *e = m.p + m.n;

    // banner line
    const char *nl = (const char *)memchr(p, '\n', e - p);
    std::string line(p, nl ? nl : e);
    p = nl ? nl + 1 : e;
    char banner[1100], mtx[1100], crd[1100], dtype[1100], storage[1100];
    if (line.size() > 1024 ||
        sscanf(line.c_str(), "%s %s %s %s %s", banner, mtx, crd, dtype, storage) != 5)
        return set_error(VNM_E_FORMAT, "%s: could not process Matrix Market banner", path);
    for (char *q : {mtx, crd, dtype, storage})
        for (; *q; ++q) *q = (char)tolower(*q);
    if (strncmp(banner, "%%MatrixMarket", 14) != 0)
        return set_error(VNM_E_FORMAT, "%s: missing %%%%MatrixMarket banner", path);
    if (strcmp(mtx, "matrix") != 0 || strcmp(crd, "coordinate") != 0)
        return set_error(VNM_E_FORMAT, "%s: only 'matrix coordinate' is supported", path);
    int field;
    if (!strcmp(dtype, "real")) field = 0;
    else if (!strcmp(dtype, "integer")) field = 1;
    else if (!strcmp(dtype, "pattern")) field = 2;
    else return set_error(VNM_E_FORMAT, "%s: unsupported field '%s'", path, dtype);
    if (!strcmp(storage, "general")) t.symmetric = false;
    else if (!strcmp(storage, "symmetric")) t.symmetric = true;
    else return set_error(VNM_E_FORMAT, "%s: unsupported symmetry '%s'", path, storage);

    // size line: first line not starting with '%'
    while (p < e && *p == '%') {
        nl = (const char *)memchr(p, '\n', e - p);
        p = nl ? nl + 1 : e;
    }
    if (p >= e) return set_error(VNM_E_FORMAT, "%s: missing size line", path);
    char *q;
    long nr = strtol(p, &q, 10);
    if (q == p) return set_error(VNM_E_FORMAT, "%s: bad size line", path);
    p = q;
    long nc = strtol(p, &q, 10);
    if (q == p) return set_error(VNM_E_FORMAT, "%s: bad size line", path);
    p = q;
    long nv = strtol(p, &q, 10);
    if (q == p) return set_error(VNM_E_FORMAT, "%s: bad size line", path);
    p = q;
    if (nr < 0 || nc < 0 || nv < 0 || nr > INT32_MAX || nc > INT32_MAX || nv > INT32_MAX)
        return set_error(VNM_E_FORMAT, "%s: bad dimensions", path);
    t.nrows = (int32_t)nr;
    t.ncols = (int32_t)nc;
    t.row.resize(nv);
    t.col.resize(nv);
    t.val.resize(nv);
    // strtof/strtol need a terminated buffer for the last token: parse from
    // a bounded copy of the tail when the mapping ends without whitespace.
    std::string tail;
    long i = 0;
    for (; i < nv; ++i) {
        p = skip_ws(p, e);
        if (p >= e) break;  // "Not enough rows": the reference keeps what it read
        if (e - p < 64 && tail.empty()) {
            tail.assign(p, e);
            tail.push_back('\n');
            e = tail.data() + tail.size();
            p = tail.data();
        }
        long r = strtol(p, &q, 10);
        if (q == p) return set_error(VNM_E_FORMAT, "%s: bad entry %ld", path, i + 1);
        p = q;
        long c = strtol(p, &q, 10);
        if (q == p) return set_error(VNM_E_FORMAT, "%s: bad entry %ld", path, i + 1);
        p = q;
        float v = 1.0f;
        if (field == 0) {
            v = strtof(p, &q);
            if (q == p) return set_error(VNM_E_FORMAT, "%s: bad value at entry %ld", path, i + 1);
            p = q;
        } else if (field == 1) {
            long iv = strtol(p, &q, 10);
            if (q == p) return set_error(VNM_E_FORMAT, "%s: bad value at entry %ld", path, i + 1);
            p = q;
            v = (float)(int)iv;
        }
        if (r < 1 || r > nr || c < 1 || c > nc)
            return set_error(VNM_E_FORMAT, "%s: entry %ld (%ld, %ld) out of range", path, i + 1, r, c);
        t.row[i] = (int32_t)(r - 1);
        t.col[i] = (int32_t)(c - 1);
        t.val[i] = v;
    }
    t.row.resize(i);
    t.col.resize(i);
    t.val.resize(i);
    return VNM_OK;
}

namespace {

struct Ent { int32_t r, c; float v; };

// customSort (mtx/readMtx.hpp:56-82): std::sort on (row, col) only.  When a
// position repeats, libstdc++'s introsort order of the repeats decides which
// value survives; the same sort over the same sequence reproduces it.
void sort_rc(std::vector<Ent> &e) {
    std::sort(e.begin(), e.end(), [](const Ent &a, const Ent &b) {
        if (a.r == b.r) return a.c < b.c;
        return a.r < b.r;
    });
}

}  // namespace

void loader_order(Tuples &t, int M, HostCoo &out) {
    size_t n0 = t.row.size();
    std::vector<Ent> e;
    e.reserve(t.symmetric ? 2 * n0 : n0);
    for (size_t i = 0; i < n0; ++i) e.push_back(Ent{t.row[i], t.col[i], t.val[i]});
    if (t.symmetric)
        for (size_t i = 0; i < n0; ++i)
            if (t.col[i] != t.row[i]) e.push_back(Ent{t.col[i], t.row[i], t.val[i]});
    out.nrows = t.nrows;
    out.ncols = t.ncols;
    out.row.clear();
    out.col.clear();
    out.val.clear();
    if (!e.empty()) {
        sort_rc(e);
        // Duplicates: positions are compacted, the value array is not, so
        // entry k keeps the k-th sorted value (mtx/readMtx.hpp:197-218).
        size_t n = e.size();
        out.row.reserve(n);
        out.col.reserve(n);
        for (size_t i = 0; i < n; ++i) {
            if (i > 0 && e[i].r == e[i - 1].r && e[i].c == e[i - 1].c) continue;
            out.row.push_back(e[i].r);
            out.col.push_back(e[i].c);
        }
        size_t k = out.row.size();
        out.val.resize(k);
        for (size_t i = 0; i < k; ++i) out.val[i] = e[i].v;
        // readMtx sorts once more (mtx/readMtx.hpp:419); positions are
        // unique and already ordered, so that sort only keeps the order.
    }
    if (M > 0 && out.nrows % M != 0) {  // mtx/Mtx.h:106-109
        out.nrows = ((out.nrows + M - 1) / M) * M;
        out.ncols = ((out.ncols + M - 1) / M) * M;
    }
}

int write_mtx(const char *path, const HostCoo &c) {
    FILE *f = fopen(path, "w");
    if (!f) return set_error(VNM_E_IO, "cannot open %s for writing", path);
    std::vector<char> buf(1 << 20);
    setvbuf(f, buf.data(), _IOFBF, buf.size());
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(f, "%d %d %d\n", c.nrows, c.ncols, (int)c.row.size());
    for (size_t i = 0; i < c.row.size(); ++i)
        fprintf(f, "%d %d %.2f\n", c.row[i] + 1, c.col[i] + 1, (double)c.val[i]);
    if (fclose(f) != 0) return set_error(VNM_E_IO, "write to %s failed", path);
    return VNM_OK;
}

int coo_export(const HostCoo &c, vnm_coo *out) {
    size_t n = c.row.size();
    out->nrows = c.nrows;
    out->ncols = c.ncols;
    out->nnz = (int64_t)n;
    out->row = (int32_t *)malloc(std::max<size_t>(n, 1) * sizeof(int32_t));
    out->col = (int32_t *)malloc(std::max<size_t>(n, 1) * sizeof(int32_t));
    out->val = (float *)malloc(std::max<size_t>(n, 1) * sizeof(float));
    if (!out->row || !out->col || !out->val) {
        free(out->row); free(out->col); free(out->val);
        out->row = out->col = nullptr;
        out->val = nullptr;
        return set_error(VNM_E_OOM, "host allocation of %zu entries failed", n);
    }
    if (n) {
        memcpy(out->row, c.row.data(), n * 4);
        memcpy(out->col, c.col.data(), n * 4);
        memcpy(out->val, c.val.data(), n * 4);
    }
    return VNM_OK;
}

void coo_import(const vnm_coo *in, HostCoo &c) {
    c.nrows = in->nrows;
    c.ncols = in->ncols;
    c.row.assign(in->row, in->row + in->nnz);
    c.col.assign(in->col, in->col + in->nnz);
    c.val.assign(in->val, in->val + in->nnz);
}

// Value of an entry after the C++ writer's "%.2f" is read back with "%f".
float quant_cpp_write(float v) {
    char buf[64];
    snprintf(buf, sizeof buf, "%.2f", (double)v);
    return strtof(buf, nullptr);
}

// Value after the temp.mtx round trip through the VENOM script: "%.2f" text,
// read as float64 (reordering.py:41), written by mmwrite(precision=1) as
// "%.1e" (reordering.py:181), read back with "%f" (mtx/readMtx.hpp:115).
float quant_py_roundtrip(float v) {
    char a[64], b[64];
    snprintf(a, sizeof a, "%.2f", (double)v);
    double d = strtod(a, nullptr);
    snprintf(b, sizeof b, "%.1e", d);
    return strtof(b, nullptr);
}

}  // namespace vnm

using namespace vnm;

extern "C" {

const char *vnm_last_error(void) { return last_error(); }

int vnm_mtx_read(const char *path, int M, vnm_coo *out) {
    if (!path || !out || M < 1) return set_error(VNM_E_ARG, "vnm_mtx_read: bad arguments");
    Tuples t;
    int rc = read_mtx_tuples(path, t);
    if (rc != VNM_OK) return rc;
    HostCoo c;
    loader_order(t, M, c);
    return coo_export(c, out);
}

int vnm_mtx_write(const char *path, const vnm_coo *coo) {
    if (!path || !coo) return set_error(VNM_E_ARG, "vnm_mtx_write: bad arguments");
    HostCoo c;
    coo_import(coo, c);
    return write_mtx(path, c);
}

int vnm_coo_from_tuples(int32_t nrows, int32_t ncols, int symmetric, int64_t n,
                        const int32_t *row, const int32_t *col, const float *val, int M,
                        vnm_coo *out) {
    if (!out || n < 0 || M < 1 || (n > 0 && (!row || !col || !val)))
        return set_error(VNM_E_ARG, "vnm_coo_from_tuples: bad arguments");
    Tuples t;
    t.nrows = nrows;
    t.ncols = ncols;
    t.symmetric = symmetric != 0;
    t.row.assign(row, row + n);
    t.col.assign(col, col + n);
    t.val.assign(val, val + n);
    for (int64_t i = 0; i < n; ++i)
        if (row[i] < 0 || row[i] >= nrows || col[i] < 0 || col[i] >= ncols)
            return set_error(VNM_E_ARG, "vnm_coo_from_tuples: entry %lld out of range", (long long)i);
    HostCoo c;
    loader_order(t, M, c);
    return coo_export(c, out);
}

void vnm_coo_free(vnm_coo *c) {
    if (!c) return;
    free(c->row);
    free(c->col);
    free(c->val);
    c->row = c->col = nullptr;
    c->val = nullptr;
    c->nnz = 0;
}

int vnm_gen_er(int64_t n, int64_t k, uint64_t seed, float lo, float hi, int M, vnm_coo *out) {
    if (!out || M < 1) return set_error(VNM_E_ARG, "vnm_gen_er: bad arguments");
    Tuples t;
    int rc = gen_er(n, k, seed, lo, hi, t);
    if (rc != VNM_OK) return set_error(rc, "vnm_gen_er: bad sizes");
    HostCoo c;
    loader_order(t, M, c);
    return coo_export(c, out);
}

int vnm_gen_rmat(int scale, int64_t und, double a, double b, double c, uint64_t seed, int M,
                 vnm_coo *out) {
    if (!out || M < 1) return set_error(VNM_E_ARG, "vnm_gen_rmat: bad arguments");
    Tuples t;
    int rc = gen_rmat(scale, und, a, b, c, seed, t);
    if (rc != VNM_OK) return set_error(rc, "vnm_gen_rmat: bad sizes");
    HostCoo h;
    loader_order(t, M, h);
    return coo_export(h, out);
}

int vnm_gen_chung_lu(int64_t n, double gamma, double mean_degree, uint64_t seed, int M,
                     vnm_coo *out) {
    if (!out || M < 1) return set_error(VNM_E_ARG, "vnm_gen_chung_lu: bad arguments");
    Tuples t;
    int rc = gen_chung_lu(n, gamma, mean_degree, seed, t);
    if (rc != VNM_OK) return set_error(rc, "vnm_gen_chung_lu: bad sizes");
    HostCoo h;
    loader_order(t, M, h);
    return coo_export(h, out);
}

void vnm_gen_normal(uint64_t seed, int64_t count, float *out) { gen_normal(seed, count, out); }

}  // extern "C"
