// Device-wide sort/scan/select primitives (rocPRIM through hipCUB), wrapped so
// the rest of the library compiles without the template-heavy headers.
#include "sortlib.h"

#include <hipcub/hipcub.hpp>

namespace vnm {

namespace {
struct TempBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t need) {
        if (need <= bytes && p) return VNM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, need ? need : 1) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(VNM_E_OOM, "scratch allocation of %zu bytes failed", need);
        }
        bytes = need;
        return VNM_OK;
    }
    ~TempBuf() {
        if (p) (void)hipFree(p);
    }
};
thread_local TempBuf g_tmp;
}  // namespace

#define SORT_TRY(expr)                                                                     \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            return set_error(VNM_E_HIP, "%s failed: %s", #expr, hipGetErrorString(_e));   \
    } while (0)

int sort_keys_u64(const uint64_t *in, uint64_t *out, int64_t n, int end_bit, hipStream_t st) {
    if (n <= 0) return VNM_OK;
    size_t bytes = 0;
    SORT_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, in, out, (int)n, 0, end_bit, st));
    int rc = g_tmp.ensure(bytes);
    if (rc) return rc;
    SORT_TRY(hipcub::DeviceRadixSort::SortKeys(g_tmp.p, bytes, in, out, (int)n, 0, end_bit, st));
    return VNM_OK;
}

int sort_pairs_u64_i32(const uint64_t *kin, uint64_t *kout, const int32_t *vin, int32_t *vout,
                       int64_t n, int end_bit, hipStream_t st) {
    if (n <= 0) return VNM_OK;
    size_t bytes = 0;
    SORT_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0,
                                                end_bit, st));
    int rc = g_tmp.ensure(bytes);
    if (rc) return rc;
    SORT_TRY(hipcub::DeviceRadixSort::SortPairs(g_tmp.p, bytes, kin, kout, vin, vout, (int)n, 0,
                                                end_bit, st));
    return VNM_OK;
}

int scan_excl_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t st) {
    if (n <= 0) return VNM_OK;
    size_t bytes = 0;
    SORT_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, st));
    int rc = g_tmp.ensure(bytes);
    if (rc) return rc;
    SORT_TRY(hipcub::DeviceScan::ExclusiveSum(g_tmp.p, bytes, in, out, (int)n, st));
    return VNM_OK;
}

int scan_incl_max_i32(const int32_t *in, int32_t *out, int64_t n, hipStream_t st) {
    if (n <= 0) return VNM_OK;
    size_t bytes = 0;
    SORT_TRY(hipcub::DeviceScan::InclusiveScan(nullptr, bytes, in, out, hipcub::Max(), (int)n, st));
    int rc = g_tmp.ensure(bytes);
    if (rc) return rc;
    SORT_TRY(hipcub::DeviceScan::InclusiveScan(g_tmp.p, bytes, in, out, hipcub::Max(), (int)n, st));
    return VNM_OK;
}

int select_flagged_i32(const int32_t *in, const uint8_t *flags, int32_t *out, int32_t *d_count,
                       int64_t n, hipStream_t st) {
    if (n <= 0) {
        SORT_TRY(hipMemsetAsync(d_count, 0, sizeof(int32_t), st));
        return VNM_OK;
    }
    size_t bytes = 0;
    SORT_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, in, flags, out, d_count, (int)n, st));
    int rc = g_tmp.ensure(bytes);
    if (rc) return rc;
    SORT_TRY(hipcub::DeviceSelect::Flagged(g_tmp.p, bytes, in, flags, out, d_count, (int)n, st));
    return VNM_OK;
}

}  // namespace vnm
