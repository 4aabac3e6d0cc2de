// MatrixMarket reader/writer and the loader's ordering rules.
//
// Reader: banner and size line as mm_read_banner / mm_read_mtx_crd_size
// (mtx/mmio.hpp:128-228); entries as readTuples (mtx/readMtx.hpp:84-147):
// "real" values parse like fscanf("%f"), "integer" like "%d" then a float
// cast, "pattern" entries get 1.0.  The text is mapped once and parsed with
// strtol/strtof, which convert exactly as the scanf family does.  Large
// files are parsed by several threads, each over a whitespace-aligned part
// of the text; an entry's fields are found by their token index, so an
// entry may span lines as it may for scanf.  A token that the sequential
// scan would split differently (one strtol/strtof does not consume it
// whole) sends the whole file to the sequential scan, so both give the
// same tuples and the same first error.
// Ordering: symmetric mirroring, (row, col) sort, duplicate removal and the
// padding of Mtx::init_from_file (mtx/readMtx.hpp:152-219,419; mtx/Mtx.h:105-112).
// Writer: Mtx::write_to_file (mtx/Mtx.h:412-435).
#include "vnm_internal.h"

#include <algorithm>
#include <atomic>
#include <cctype>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <system_error>
#include <thread>
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

// Worker threads for a host pass: `threads` if positive, else the machine's
// hardware threads, at most 16.
int host_threads(int threads) {
    if (threads > 0) return std::min(threads, 64);
    const unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

// Run f(k) for k in [0, T) on up to T threads (the caller's thread takes
// k = 0, and any k whose thread could not be started).
template <typename F>
void parallel_for(int T, F f) {
    if (T <= 1) {
        for (int k = 0; k < T; ++k) f(k);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(T - 1);
    int started = 1;
    try {
        for (; started < T; ++started) th.emplace_back(f, started);
    } catch (const std::system_error &) {
    }
    f(0);
    for (int k = started; k < T; ++k) f(k);
    for (auto &x : th) x.join();
}

namespace {

// The sequential scan (readTuples, mtx/readMtx.hpp:84-147).
int parse_serial(const char *path, const char *p, const char *e, int field, long nr, long nc, long nv,
                 Tuples &t) {
    char *q;
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

// strtol of a whole token of at most 18 digits after an optional sign (the
// value is then exact in 64 bits); false for anything else.
inline bool fast_long(const char *s, size_t len, long &v) {
    size_t i = 0;
    bool neg = false;
    if (len && (s[0] == '-' || s[0] == '+')) {
        neg = s[0] == '-';
        i = 1;
    }
    if (i == len || len - i > 18) return false;
    long x = 0;
    for (; i < len; ++i) {
        const unsigned d = (unsigned)(s[i] - '0');
        if (d > 9) return false;
        x = x * 10 + (long)d;
    }
    v = neg ? -x : x;
    return true;
}

// strtof of a whole token "[sign]digits[.digits]" whose digits form an
// integer m <= 2^24 with k <= 10 of them after the point: m and 10^k are
// exact floats, so the one rounding of m / 10^k is strtof's correctly
// rounded result (Clinger's fast path, 1990); false for anything else.
inline bool fast_float(const char *s, size_t len, float &v) {
    static const float p10[11] = {1e0f, 1e1f, 1e2f, 1e3f, 1e4f, 1e5f, 1e6f, 1e7f, 1e8f, 1e9f, 1e10f};
    size_t i = 0;
    bool neg = false;
    if (len && (s[0] == '-' || s[0] == '+')) {
        neg = s[0] == '-';
        i = 1;
    }
    uint64_t m = 0;
    int digits = 0, frac = -1;
    for (; i < len; ++i) {
        const char ch = s[i];
        if (ch == '.') {
            if (frac >= 0) return false;
            frac = 0;
            continue;
        }
        const unsigned d = (unsigned)(ch - '0');
        if (d > 9) return false;
        if (++digits > 18) return false;
        m = m * 10 + d;
        if (frac >= 0) ++frac;
    }
    if (digits == 0 || m > (1u << 24)) return false;
    const int k = frac < 0 ? 0 : frac;
    if (k > 10) return false;
    const float x = (float)m / p10[k];
    v = neg ? -x : x;
    return true;
}

constexpr int PARSE_FALLBACK = 1;   // parse_parallel: use the sequential scan
constexpr long PARALLEL_BYTES = 1 << 20;  // smaller texts are scanned by one thread
constexpr int TOKEN_MAX = 63;       // longest token the parallel scan copies

// The same scan on T threads.  Token g (counted over the whole text) is
// field g % W of entry g / W.  The first event in the sequential order wins:
// a token that does not parse (at its position), or an entry out of range
// (after its last token).
int parse_parallel(const char *path, const char *p, const char *e, int field, long nr, long nc, long nv,
                   int T, Tuples &t) {
    const int W = field == 2 ? 2 : 3;
    const int64_t need = (int64_t)nv * W;
    // parts of the text, each starting at a token start
    std::vector<const char *> cut(T + 1);
    cut[0] = p;
    cut[T] = e;
    for (int k = 1; k < T; ++k) {
        const char *x = p + (size_t)(e - p) * k / T;
        if (x < cut[k - 1]) x = cut[k - 1];
        while (x < e && !isspace((unsigned char)*x)) ++x;
        cut[k] = x;
    }
    std::vector<int64_t> ntok(T + 1, 0);
    parallel_for(T, [&](int k) {
        int64_t n = 0;
        bool in = false;
        for (const char *x = cut[k]; x < cut[k + 1]; ++x) {
            const bool w = isspace((unsigned char)*x);
            n += !w && !in;
            in = !w;
        }
        ntok[k + 1] = n;
    });
    for (int k = 0; k < T; ++k) ntok[k + 1] += ntok[k];
    const int64_t total = ntok[T];
    const int64_t used = std::min(total, need);
    const long nent = (long)(used / W);  // whole entries present
    // indices in range are stored 0-based, others as -1 with their value
    // kept for the message (the first two per part: the first bad entry
    // holds the first one or two of the whole text)
    std::vector<std::vector<std::pair<int64_t, long>>> oor(T);
    t.row.resize(nent);
    t.col.resize(nent);
    t.val.resize(nent);
    std::atomic<int> fallback{0};
    std::vector<int64_t> bad_tok(T, INT64_MAX);
    parallel_for(T, [&](int k) {
        int64_t g = ntok[k];
        const char *x = cut[k], *end = cut[k + 1];
        char buf[TOKEN_MAX + 1];
        while (g < used) {
            while (x < end && isspace((unsigned char)*x)) ++x;
            if (x >= end) break;
            const char *s0 = x;
            while (x < end && !isspace((unsigned char)*x)) ++x;
            const size_t len = (size_t)(x - s0);
            if (len > TOKEN_MAX) {
                fallback.store(1, std::memory_order_relaxed);
                return;
            }
            memcpy(buf, s0, len);
            buf[len] = 0;
            char *q;
            const int64_t i = g / W;
            const int f = (int)(g % W);
            if (f == 2 && field == 0) {
                float v;
                if (!fast_float(buf, len, v)) {
                    v = strtof(buf, &q);
                    if (q == buf) { bad_tok[k] = g; return; }
                    if (q != buf + len) {  // the sequential scan would split this token
                        fallback.store(1, std::memory_order_relaxed);
                        return;
                    }
                }
                if (i < nent) t.val[i] = v;
            } else {
                long v;
                if (!fast_long(buf, len, v)) {
                    v = strtol(buf, &q, 10);
                    if (q == buf) { bad_tok[k] = g; return; }
                    if (q != buf + len) {
                        fallback.store(1, std::memory_order_relaxed);
                        return;
                    }
                }
                if (i < nent) {
                    if (f < 2) {
                        const long lim = f == 0 ? nr : nc;
                        const bool ok = v >= 1 && v <= lim;
                        (f == 0 ? t.row : t.col)[i] = ok ? (int32_t)(v - 1) : -1;
                        if (!ok && oor[k].size() < 2) oor[k].emplace_back(g, v);
                    } else {
                        t.val[i] = (float)(int)v;
                    }
                }
            }
            ++g;
        }
    });
    if (fallback.load()) return PARSE_FALLBACK;
    int64_t first_bad = INT64_MAX;
    for (int k = 0; k < T; ++k) first_bad = std::min(first_bad, bad_tok[k]);
    // text ended inside an entry: its next field fails to parse
    if (first_bad == INT64_MAX && total < need && total % W != 0) first_bad = total;
    std::vector<long> bad_ent(T, LONG_MAX);
    parallel_for(T, [&](int k) {
        const long i0 = (long)((int64_t)nent * k / T), i1 = (long)((int64_t)nent * (k + 1) / T);
        for (long i = i0; i < i1; ++i) {
            if (field == 2) t.val[i] = 1.0f;
            if (t.row[i] < 0 || t.col[i] < 0) {
                bad_ent[k] = i;
                return;
            }
        }
    });
    long range_bad = LONG_MAX;
    for (int k = 0; k < T; ++k) range_bad = std::min(range_bad, bad_ent[k]);
    // sequential order: entry i's range check follows its last token
    const int64_t range_pos = range_bad == LONG_MAX ? INT64_MAX : (int64_t)range_bad * W + (W - 1);
    if (first_bad != INT64_MAX && first_bad <= range_pos) {
        const long i = (long)(first_bad / W);
        if (first_bad % W == 2)
            return set_error(VNM_E_FORMAT, "%s: bad value at entry %ld", path, i + 1);
        return set_error(VNM_E_FORMAT, "%s: bad entry %ld", path, i + 1);
    }
    if (range_bad != LONG_MAX) {
        auto raw = [&](int64_t g, int32_t stored) {
            if (stored >= 0) return (long)stored + 1;
            for (auto &o : oor)
                for (auto &x : o)
                    if (x.first == g) return x.second;
            return 0L;
        };
        const int64_t g0 = (int64_t)range_bad * W;
        return set_error(VNM_E_FORMAT, "%s: entry %ld (%ld, %ld) out of range", path, range_bad + 1,
                         raw(g0, t.row[range_bad]), raw(g0 + 1, t.col[range_bad]));
    }
    return VNM_OK;
}

}  // namespace

int read_mtx_tuples(const char *path, Tuples &t, int threads) {
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
    const int T = host_threads(threads);
    if (T > 1 && e - p >= PARALLEL_BYTES) {
        Tuples u;
        u.nrows = t.nrows;
        u.ncols = t.ncols;
        u.symmetric = t.symmetric;
        const int rc = parse_parallel(path, p, e, field, nr, nc, nv, T, u);
        if (rc != PARSE_FALLBACK) {
            if (rc == VNM_OK) t = std::move(u);
            return rc;
        }
    }
    t.row.resize(nv);
    t.col.resize(nv);
    t.val.resize(nv);
    return parse_serial(path, p, e, field, nr, nc, nv, t);
}

namespace {

struct Ent { int32_t r, c; float v; };

constexpr size_t PARALLEL_SORT = 1 << 16;       // fewer entries are sorted by one thread
constexpr size_t WRITE_PER_THREAD = 1 << 17;    // entries a writer thread formats per block

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

namespace {

inline bool ent_less(const Ent &a, const Ent &b) {
    if (a.r == b.r) return a.c < b.c;
    return a.r < b.r;
}

// (row, col) sort on T threads: parts sorted, then merged pairwise.  Used
// only where positions are unique, so the order is the sequential sort's.
void sort_rc_parallel(std::vector<Ent> &e, int T) {
    const size_t n = e.size();
    std::vector<size_t> b(T + 1);
    for (int k = 0; k <= T; ++k) b[k] = n * (size_t)k / (size_t)T;
    parallel_for(T, [&](int k) { std::sort(e.begin() + b[k], e.begin() + b[k + 1], ent_less); });
    std::vector<Ent> tmp(n);
    std::vector<Ent> *src = &e, *dst = &tmp;
    for (int w = 1; w < T; w *= 2) {
        const int npair = (T + 2 * w - 1) / (2 * w);
        parallel_for(npair, [&](int j) {
            const int k0 = j * 2 * w, k1 = std::min(k0 + w, T), k2 = std::min(k0 + 2 * w, T);
            std::merge(src->begin() + b[k0], src->begin() + b[k1], src->begin() + b[k1], src->begin() + b[k2],
                       dst->begin() + b[k0], ent_less);
        });
        std::swap(src, dst);
    }
    if (src != &e) e.swap(*src);
}

}  // namespace

namespace {

// Entries of the loader in input order: the tuples, then the mirrored
// off-diagonal tuples of a symmetric file (mtx/readMtx.hpp:161-169).
void loader_entries(const Tuples &t, std::vector<Ent> &e, int T) {
    const size_t n0 = t.row.size();
    if (T > 1 && n0 < PARALLEL_SORT) T = 1;
    std::vector<size_t> b(T + 1), off(T + 1, 0);
    for (int k = 0; k <= T; ++k) b[k] = n0 * (size_t)k / (size_t)T;
    if (t.symmetric)
        parallel_for(T, [&](int k) {
            size_t m = 0;
            for (size_t i = b[k]; i < b[k + 1]; ++i) m += t.col[i] != t.row[i];
            off[k + 1] = m;
        });
    for (int k = 0; k < T; ++k) off[k + 1] += off[k];
    e.resize(n0 + off[T]);
    parallel_for(T, [&](int k) {
        Ent *mir = e.data() + n0 + off[k];
        for (size_t i = b[k]; i < b[k + 1]; ++i) {
            e[i] = Ent{t.row[i], t.col[i], t.val[i]};
            if (t.symmetric && t.col[i] != t.row[i]) *mir++ = Ent{t.col[i], t.row[i], t.val[i]};
        }
    });
}

}  // namespace

void loader_order(Tuples &t, int M, HostCoo &out, int threads) {
    const int T = host_threads(threads);
    std::vector<Ent> e;
    loader_entries(t, e, T);
    out.nrows = t.nrows;
    out.ncols = t.ncols;
    out.row.clear();
    out.col.clear();
    out.val.clear();
    if (!e.empty()) {
        const size_t n = e.size();
        if (T > 1 && n >= PARALLEL_SORT) {
            // with unique positions every correct sort gives one order; a
            // repeated position sends the entries to the sequential sort,
            // whose order of the repeats decides the values (below)
            sort_rc_parallel(e, T);
            std::atomic<int> dup{0};
            parallel_for(T, [&](int k) {
                const size_t i0 = std::max<size_t>(1, n * (size_t)k / (size_t)T);
                const size_t i1 = n * (size_t)(k + 1) / (size_t)T;
                for (size_t i = i0; i < i1; ++i)
                    if (e[i].r == e[i - 1].r && e[i].c == e[i - 1].c) {
                        dup.store(1, std::memory_order_relaxed);
                        return;
                    }
            });
            if (!dup.load()) {
                out.row.resize(n);
                out.col.resize(n);
                out.val.resize(n);
                parallel_for(T, [&](int k) {
                    const size_t i0 = n * (size_t)k / (size_t)T, i1 = n * (size_t)(k + 1) / (size_t)T;
                    for (size_t i = i0; i < i1; ++i) {
                        out.row[i] = e[i].r;
                        out.col[i] = e[i].c;
                        out.val[i] = e[i].v;
                    }
                });
            } else {
                loader_entries(t, e, T);
            }
        }
        if (out.row.empty()) {
            sort_rc(e);
            // Duplicates: positions are compacted, the value array is not, so
            // entry k keeps the k-th sorted value (mtx/readMtx.hpp:197-218).
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
        }
        // readMtx sorts once more (mtx/readMtx.hpp:419); positions are
        // unique and already ordered, so that sort only keeps the order.
    }
    if (M > 0 && out.nrows % M != 0) {  // mtx/Mtx.h:106-109
        out.nrows = ((out.nrows + M - 1) / M) * M;
        out.ncols = ((out.ncols + M - 1) / M) * M;
    }
}

namespace {

inline char *put_uint(char *o, uint64_t x) {
    char tmp[24];
    int n = 0;
    do {
        tmp[n++] = (char)('0' + x % 10);
        x /= 10;
    } while (x);
    while (n) *o++ = tmp[--n];
    return o;
}

inline char *put_int(char *o, int x) {
    if (x < 0) {
        *o++ = '-';
        return put_uint(o, (uint64_t)(-(int64_t)x));
    }
    return put_uint(o, (uint64_t)x);
}

// "%.2f" of (double)v for a finite float below 2^31 in magnitude: v =
// man * 2^e2 exactly, so v * 100 = man * 100 * 2^e2 is split into its
// integer part and remainder without rounding and rounded to nearest, ties
// to even, as printf rounds.  nullptr for infinities, NaNs and larger values.
inline char *put_fixed2(char *o, float v) {
    uint32_t bits;
    memcpy(&bits, &v, 4);
    const int ex = (int)((bits >> 23) & 0xFFu);
    if (ex == 0xFF) return nullptr;
    uint64_t man = bits & 0x7FFFFFu;
    int e2;
    if (ex == 0) {
        e2 = -149;
    } else {
        man |= 0x800000u;
        e2 = ex - 150;
    }
    uint64_t q;
    if (e2 >= 0) {
        if (e2 > 7) return nullptr;
        q = (man * 100) << e2;
    } else {
        const int sh = -e2;
        const uint64_t x = man * 100;  // < 2^31
        if (sh >= 40) {
            q = 0;  // x * 2^-sh < 2^-9: rounds to 0
        } else {
            q = x >> sh;
            const uint64_t r = x & ((1ull << sh) - 1), half = 1ull << (sh - 1);
            if (r > half || (r == half && (q & 1))) ++q;
        }
    }
    if (bits >> 31) *o++ = '-';
    o = put_uint(o, q / 100);
    const unsigned f = (unsigned)(q % 100);
    *o++ = '.';
    *o++ = (char)('0' + f / 10);
    *o++ = (char)('0' + f % 10);
    return o;
}

}  // namespace

// Mtx::write_to_file (mtx/Mtx.h:412-435).  One thread: the fprintf loop.
// More: each formats a block of entries into its own buffer (integers and
// the exact %.2f above, snprintf for what that does not cover) and the
// blocks are written in order.
int write_mtx(const char *path, const HostCoo &c, int threads) {
    FILE *f = fopen(path, "w");
    if (!f) return set_error(VNM_E_IO, "cannot open %s for writing", path);
    fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n");
    fprintf(f, "%d %d %d\n", c.nrows, c.ncols, (int)c.row.size());
    const size_t n = c.row.size();
    const int T = host_threads(threads);
    if (T == 1) {
        std::vector<char> vb(1 << 20);
        setvbuf(f, vb.data(), _IOFBF, vb.size());
        for (size_t i = 0; i < n; ++i)
            fprintf(f, "%d %d %.2f\n", c.row[i] + 1, c.col[i] + 1, (double)c.val[i]);
        if (fclose(f) != 0) return set_error(VNM_E_IO, "write to %s failed", path);
        return VNM_OK;
    }
    // no more threads than blocks of entries
    const int TW = (int)std::min<size_t>((size_t)T, std::max<size_t>(1, (n + WRITE_PER_THREAD - 1) / WRITE_PER_THREAD));
    std::vector<std::string> buf(TW);
    bool ok = true;
    for (size_t b0 = 0; b0 < n && ok; b0 += (size_t)TW * WRITE_PER_THREAD) {
        const size_t b1 = std::min(n, b0 + (size_t)TW * WRITE_PER_THREAD);
        parallel_for(TW, [&](int k) {
            const size_t i0 = std::min(b1, b0 + (size_t)k * WRITE_PER_THREAD);
            const size_t i1 = std::min(b1, i0 + WRITE_PER_THREAD);
            std::string &o = buf[k];
            o.resize((i1 - i0) * 40);
            char *w = &o[0];
            for (size_t i = i0; i < i1; ++i) {
                char line[64];
                char *x = put_int(line, c.row[i] + 1);
                *x++ = ' ';
                x = put_int(x, c.col[i] + 1);
                *x++ = ' ';
                char *y = put_fixed2(x, c.val[i]);
                if (y) {
                    *y++ = '\n';
                    const size_t len = (size_t)(y - line);
                    const size_t at = (size_t)(w - &o[0]);
                    if (at + len > o.size()) {
                        o.resize(2 * o.size() + len);
                        w = &o[0] + at;
                    }
                    memcpy(w, line, len);
                    w += len;
                } else {  // infinities, NaNs, |v| >= 2^31
                    const int len = snprintf(nullptr, 0, "%d %d %.2f\n", c.row[i] + 1, c.col[i] + 1,
                                             (double)c.val[i]);
                    std::vector<char> big((size_t)len + 1);
                    snprintf(big.data(), big.size(), "%d %d %.2f\n", c.row[i] + 1, c.col[i] + 1,
                             (double)c.val[i]);
                    const size_t at = (size_t)(w - &o[0]);
                    if (at + (size_t)len > o.size()) {
                        o.resize(2 * o.size() + (size_t)len);
                        w = &o[0] + at;
                    }
                    memcpy(w, big.data(), (size_t)len);
                    w += len;
                }
            }
            o.resize((size_t)(w - &o[0]));
        });
        for (int k = 0; k < TW && ok; ++k)
            if (!buf[k].empty() && fwrite(buf[k].data(), 1, buf[k].size(), f) != buf[k].size()) ok = false;
    }
    if (fclose(f) != 0 || !ok) return set_error(VNM_E_IO, "write to %s failed", path);
    return VNM_OK;
}

// ---- binary cache of loaded matrices
//
// The reference names a binary cache next to the matrix,
// "<dir>/.<file>.<ud|d>.<nosl|sl>.bin" (mtx/readMtx.hpp:230-247,388-405: ud
// for symmetric input, nosl unless GRB_UTIL_REMOVE_SELFLOOP=0), and never
// writes or reads it.  Here the file holds the loader's COO before the
// M-padding (sorted, repeats removed), valid while the source file keeps its
// size and modification time: a 64-byte header, then row[], col[], val[].

namespace {

struct CacheHeader {
    char magic[8];  // "VNMCOO1"
    uint64_t src_size;
    int64_t src_mtime_ns;
    int64_t nnz;
    int32_t nrows, ncols;
    int32_t symmetric;
    char pad[20];
};
static_assert(sizeof(CacheHeader) == 64, "cache header layout");

int64_t mtime_ns(const struct stat &st) {
    return (int64_t)st.st_mtim.tv_sec * 1000000000ll + (int64_t)st.st_mtim.tv_nsec;
}

}  // namespace

// The banner's storage field (first line), for the cache name.
int banner_symmetric(const char *path, bool *sym) {
    FILE *f = fopen(path, "r");
    if (!f) return set_error(VNM_E_IO, "cannot open %s", path);
    char line[1100] = {0};
    const bool ok = fgets(line, sizeof line, f) != nullptr;
    fclose(f);
    if (!ok) return set_error(VNM_E_FORMAT, "%s: empty file", path);
    for (char *q = line; *q; ++q) *q = (char)tolower(*q);
    *sym = strstr(line, "symmetric") != nullptr;
    return VNM_OK;
}

std::string cache_name(const char *path, bool symmetric) {
    const std::string p(path);
    const size_t sl = p.find_last_of('/');
    const std::string dir = sl == std::string::npos ? "." : (sl == 0 ? "/" : p.substr(0, sl));
    const std::string base = sl == std::string::npos ? p : p.substr(sl + 1);
    const char *e = getenv("GRB_UTIL_REMOVE_SELFLOOP");
    const bool nosl = e == nullptr || atoi(e) != 0;
    return dir + (dir == "/" ? "" : "/") + "." + base + "." + (symmetric ? "ud" : "d") + "." +
           (nosl ? "nosl" : "sl") + ".bin";
}

bool cache_load(const std::string &name, const struct stat &src, HostCoo &c) {
    FILE *f = fopen(name.c_str(), "rb");
    if (!f) return false;
    CacheHeader h;
    bool ok = fread(&h, sizeof h, 1, f) == 1 && memcmp(h.magic, "VNMCOO1", 8) == 0 &&
              h.src_size == (uint64_t)src.st_size && h.src_mtime_ns == mtime_ns(src) && h.nnz >= 0 &&
              h.nnz <= ((int64_t)1 << 40) && h.nrows >= 0 && h.ncols >= 0;
    if (ok) {
        struct stat cs;
        ok = fstat(fileno(f), &cs) == 0 && (uint64_t)cs.st_size == sizeof h + (uint64_t)h.nnz * 12u;
    }
    if (ok) {
        c.nrows = h.nrows;
        c.ncols = h.ncols;
        c.row.resize(h.nnz);
        c.col.resize(h.nnz);
        c.val.resize(h.nnz);
        const size_t n = (size_t)h.nnz;
        ok = fread(c.row.data(), 4, n, f) == n && fread(c.col.data(), 4, n, f) == n &&
             fread(c.val.data(), 4, n, f) == n;
        // in range, and strictly increasing in (row, col) as the loader
        // leaves them: a damaged file of the right size is refused too
        for (size_t i = 0; ok && i < n; ++i)
            ok = c.row[i] >= 0 && c.row[i] < c.nrows && c.col[i] >= 0 && c.col[i] < c.ncols &&
                 (i == 0 || c.row[i] > c.row[i - 1] || (c.row[i] == c.row[i - 1] && c.col[i] > c.col[i - 1]));
    }
    fclose(f);
    return ok;
}

// Best effort: a directory that cannot be written leaves no cache.
void cache_store(const std::string &name, const struct stat &src, bool symmetric, const HostCoo &c) {
    const std::string tmp = name + ".tmp" + std::to_string((long)getpid());
    FILE *f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    CacheHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, "VNMCOO1", 8);
    h.src_size = (uint64_t)src.st_size;
    h.src_mtime_ns = mtime_ns(src);
    h.nnz = (int64_t)c.row.size();
    h.nrows = c.nrows;
    h.ncols = c.ncols;
    h.symmetric = symmetric;
    const size_t n = c.row.size();
    bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(c.row.data(), 4, n, f) == n &&
              fwrite(c.col.data(), 4, n, f) == n && fwrite(c.val.data(), 4, n, f) == n;
    ok = (fclose(f) == 0) && ok;
    if (!ok || rename(tmp.c_str(), name.c_str()) != 0) unlink(tmp.c_str());
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

int vnm_mtx_read_ex(const char *path, int M, int threads, vnm_coo *out) {
    if (!path || !out || M < 1 || threads < 0) return set_error(VNM_E_ARG, "vnm_mtx_read: bad arguments");
    Tuples t;
    int rc = read_mtx_tuples(path, t, threads);
    if (rc != VNM_OK) return rc;
    HostCoo c;
    loader_order(t, M, c, threads);
    return coo_export(c, out);
}

int vnm_mtx_read(const char *path, int M, vnm_coo *out) { return vnm_mtx_read_ex(path, M, 0, out); }

int vnm_mtx_read_cached(const char *path, int M, int threads, int *from_cache, vnm_coo *out) {
    if (!path || !out || M < 1 || threads < 0) return set_error(VNM_E_ARG, "vnm_mtx_read_cached: bad arguments");
    if (from_cache) *from_cache = 0;
    struct stat st;
    if (stat(path, &st) != 0) return set_error(VNM_E_IO, "cannot open %s", path);
    bool sym = false;
    int rc = banner_symmetric(path, &sym);
    if (rc != VNM_OK) return rc;
    const std::string cname = cache_name(path, sym);
    HostCoo c;
    if (cache_load(cname, st, c)) {
        if (from_cache) *from_cache = 1;
    } else {
        Tuples t;
        if ((rc = read_mtx_tuples(path, t, threads)) != VNM_OK) return rc;
        loader_order(t, 0, c, threads);  // unpadded: the cache serves any M
        cache_store(cname, st, sym, c);
    }
    if (c.nrows % M != 0) {  // mtx/Mtx.h:106-109, as loader_order pads
        c.nrows = ((c.nrows + M - 1) / M) * M;
        c.ncols = ((c.ncols + M - 1) / M) * M;
    }
    return coo_export(c, out);
}

int vnm_mtx_write_ex(const char *path, const vnm_coo *coo, int threads) {
    if (!path || !coo || threads < 0) return set_error(VNM_E_ARG, "vnm_mtx_write: bad arguments");
    HostCoo c;
    coo_import(coo, c);
    return write_mtx(path, c, threads);
}

int vnm_mtx_write(const char *path, const vnm_coo *coo) { return vnm_mtx_write_ex(path, coo, 0); }

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
    loader_order(t, M, c, 0);
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
    loader_order(t, M, c, 0);
    return coo_export(c, out);
}

int vnm_gen_rmat(int scale, int64_t und, double a, double b, double c, uint64_t seed, int M,
                 vnm_coo *out) {
    if (!out || M < 1) return set_error(VNM_E_ARG, "vnm_gen_rmat: bad arguments");
    Tuples t;
    int rc = gen_rmat(scale, und, a, b, c, seed, t);
    if (rc != VNM_OK) return set_error(rc, "vnm_gen_rmat: bad sizes");
    HostCoo h;
    loader_order(t, M, h, 0);
    return coo_export(h, out);
}

int vnm_gen_chung_lu(int64_t n, double gamma, double mean_degree, uint64_t seed, int M,
                     vnm_coo *out) {
    if (!out || M < 1) return set_error(VNM_E_ARG, "vnm_gen_chung_lu: bad arguments");
    Tuples t;
    int rc = gen_chung_lu(n, gamma, mean_degree, seed, t);
    if (rc != VNM_OK) return set_error(rc, "vnm_gen_chung_lu: bad sizes");
    HostCoo h;
    loader_order(t, M, h, 0);
    return coo_export(h, out);
}

void vnm_gen_normal(uint64_t seed, int64_t count, float *out) { gen_normal(seed, count, out); }

}  // extern "C"
