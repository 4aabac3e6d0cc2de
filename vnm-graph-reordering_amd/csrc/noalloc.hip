// vnm_spmm_noalloc: the SpMM without any device allocation, for callers that
// cannot allocate (code under HIP graph capture, serving loops, harnesses
// that forbid hipMalloc).  The packed matrix brings its format and work-item
// tables; the only scratch the SpMM needs -- the partial-sum slots of rows
// cut into work-item chunks -- comes from a static device array instead of
// the packed matrix's own buffer.
#include "graph.h"
#include "hip_common.h"
#include "packed.h"

#include <map>
#include <mutex>

namespace vnm {

// The partial-sum slots of vnm_spmm_noalloc: a static device array of
// NOALLOC_FLOATS floats (32 MiB) that each device holds until the process
// ends and that every vnm_spmm_noalloc call on the device shares.  The SpMM
// kernel writes every slot it uses in full before the fix-up pass reads it,
// so nothing needs clearing between calls.
constexpr size_t NOALLOC_FLOATS = (size_t)8 << 20;
static __device__ __align__(256) float scratch[NOALLOC_FLOATS];

static std::mutex noalloc_mu;
static std::map<int, float *> noalloc_addr;  // device -> address of `scratch` there

// Called by vnm_ctx_create with `device` current: the array's device
// address, looked up once per device (the host-side name `scratch` is not a
// device address, and each device holds its own copy of the array).
int noalloc_setup(int device, float **out) {
    std::lock_guard<std::mutex> lk(noalloc_mu);
    auto it = noalloc_addr.find(device);
    if (it != noalloc_addr.end()) {
        *out = it->second;
        return VNM_OK;
    }
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return set_error(VNM_E_HIP, "hipGetSymbolAddress of the SpMM slot array failed on device %d: %s", device,
                         hipGetErrorString(e));
    }
    noalloc_addr[device] = (float *)ptr;
    *out = (float *)ptr;
    return VNM_OK;
}

// Values per lane when no shape was chosen for (bt, n): 8 bytes per lane,
// or the widest narrower setting that fits n (no timing, which would create
// events).
static int default_vec(int bt, int n) {
    for (int v = 8 / bsize_of(bt); v >= 1; v /= 2)
        if (vec_ok(bt, n, v)) return v;
    return 0;
}

static int spmm_noalloc(vnm_packed *p, const void *B, int ldb, int n, int bt, const float *bs, float *C, int ldc) {
    int rc = spmm_check(p, B, ldb, n, bt, bs, C, ldc);
    if (rc) return rc;
    if (!p->ctx->noalloc_slots) return set_error(VNM_E_HIP, "spmm_noalloc: no slot array on this device");
    if (scratch_rows(p) * (size_t)n > NOALLOC_FLOATS)
        return set_error(VNM_E_ARG,
                         "spmm_noalloc: %zu slot rows x %d columns = %zu floats exceed the static slot array "
                         "(%zu floats); use vnm_spmm",
                         scratch_rows(p), n, scratch_rows(p) * (size_t)n, NOALLOC_FLOATS);
    auto it = p->shapes.find(shape_key(bt, n));
    const int v = it != p->shapes.end() ? it->second : default_vec(bt, n);
    if (!v) return set_error(VNM_E_ARG, "spmm_noalloc: no launch shape fits n = %d", n);
    return spmm_run(p, v, B, ldb, n, bt, bs, C, ldc, nullptr, p->ctx->noalloc_slots, false);
}

}  // namespace vnm

extern "C" int vnm_spmm_noalloc(vnm_packed *p, const void *B, int ldb, int n, int b_dtype, const float *b_scales,
                                float *C, int ldc) {
    return vnm::spmm_noalloc(p, B, ldb, n, b_dtype, b_scales, C, ldc);
}
