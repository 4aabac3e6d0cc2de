// Context, graph lifetime and host<->device transfer of the COO.
//
// Replaces the ownership model of Mtx<float> (mtx/Mtx.h:21-96: raw device
// buffers per matrix, freed in the destructor) with a context that owns one
// HIP stream and graphs that own their device buffers.  No HIP error is
// swallowed (the reference's CUDA_SAFE_CALL prints and continues,
// util/utility.h:72-83).
#include "graph.h"
#include "hip_common.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace vnm {


__global__ void k_iota(int32_t *p, int32_t n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

static int upload(vnm_graph *g, const vnm_coo *coo) {
    hipStream_t st = g->ctx->stream;
    int rc;
    const size_t n = (size_t)coo->nnz;
    if ((rc = g->row.alloc(n)) || (rc = g->col.alloc(n)) || (rc = g->val.alloc(n))) return rc;
    if (n) {
        VNM_HIP_TRY(hipMemcpyAsync(g->row.p, coo->row, n * 4, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipMemcpyAsync(g->col.p, coo->col, n * 4, hipMemcpyHostToDevice, st));
        VNM_HIP_TRY(hipMemcpyAsync(g->val.p, coo->val, n * 4, hipMemcpyHostToDevice, st));
    }
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int graph_clone(vnm_graph *src, vnm_graph **out) {
    vnm_graph *g = new vnm_graph();
    g->ctx = src->ctx;
    g->nrows = src->nrows;
    g->ncols = src->ncols;
    g->M = src->M;
    g->nnz = src->nnz;
    g->col_overflow = src->col_overflow;
    hipStream_t st = src->ctx->stream;
    int rc;
    const size_t n = (size_t)src->nnz;
    if ((rc = g->row.alloc(n)) || (rc = g->col.alloc(n)) || (rc = g->val.alloc(n)) ||
        (rc = g->perm_total.alloc(src->perm_total.n))) {
        delete g;
        return rc;
    }
    if (hipMemcpyAsync(g->row.p, src->row.p, n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(g->col.p, src->col.p, n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(g->val.p, src->val.p, n * 4, hipMemcpyDeviceToDevice, st) != hipSuccess ||
        hipMemcpyAsync(g->perm_total.p, src->perm_total.p, src->perm_total.n * 4,
                       hipMemcpyDeviceToDevice, st) != hipSuccess) {
        delete g;
        return set_error(VNM_E_HIP, "graph clone copy failed");
    }
    *out = g;
    return VNM_OK;
}

KernelTimer::KernelTimer(vnm_ctx *c, const char *n) : ctx(c), name(n) {
    if (!ctx || !ctx->timing) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        a = b = nullptr;
        (void)hipGetLastError();
        return;
    }
    (void)hipEventRecord(a, ctx->stream);
}

void KernelTimer::stop() {
    if (!a) return;
    (void)hipEventRecord(b, ctx->stream);
    ctx->pending.push_back(vnm_timer_pending{name, a, b});
    a = b = nullptr;
}

static void resolve_timers(vnm_ctx *c) {
    for (auto &p : c->pending) {
        float ms = 0.f;
        if (hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto &t = c->timers[p.name];
            t.first += ms;
            t.second += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    c->pending.clear();
}

}  // namespace vnm

using namespace vnm;

extern "C" {

int vnm_ctx_timing(vnm_ctx *c, int enable) {
    if (!c) return set_error(VNM_E_ARG, "null context");
    resolve_timers(c);
    c->timing = enable != 0;
    c->timers.clear();
    return VNM_OK;
}

int vnm_ctx_timer_read(vnm_ctx *c, const char *kernel, double *total_ms, int64_t *launches) {
    if (!c || !kernel || !total_ms || !launches) return set_error(VNM_E_ARG, "null argument");
    resolve_timers(c);
    auto it = c->timers.find(kernel);
    *total_ms = it == c->timers.end() ? 0.0 : it->second.first;
    *launches = it == c->timers.end() ? 0 : it->second.second;
    return VNM_OK;
}

int vnm_ctx_create(int device, vnm_ctx **out) {
    if (!out) return set_error(VNM_E_ARG, "null output");
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return set_error(VNM_E_NODEV, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return set_error(VNM_E_ARG, "device %d out of range", device);
    VNM_HIP_TRY(hipSetDevice(device));
    vnm_ctx *c = new vnm_ctx();
    c->device = device;
    int rc = noalloc_setup(device, &c->noalloc_slots);
    if (rc) {
        delete c;
        return rc;
    }
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return set_error(VNM_E_HIP, "hipStreamCreate failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return VNM_OK;
}

void vnm_ctx_destroy(vnm_ctx *c) {
    if (!c) return;
    resolve_timers(c);
    if (c->stream) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

void *vnm_ctx_stream(vnm_ctx *c) { return c ? (void *)c->stream : nullptr; }

int64_t vnm_ctx_launches(vnm_ctx *c) { return c ? c->launches : 0; }

int vnm_graph_create(vnm_ctx *ctx, const vnm_coo *coo, int M, vnm_graph **out) {
    if (!ctx || !coo || !out) return set_error(VNM_E_ARG, "null argument");
    if (M != 4 && M != 8 && M != 16) return set_error(VNM_E_ARG, "M must be 4, 8 or 16 (got %d)", M);
    if (coo->nrows <= 0 || coo->ncols <= 0 || coo->nnz < 0)
        return set_error(VNM_E_ARG, "bad matrix dimensions");
    if (coo->nrows % 4 != 0)
        return set_error(VNM_E_ARG, "nrows must be padded to a multiple of M (got %d)", coo->nrows);
    for (int64_t i = 0; i < coo->nnz; ++i)
        if (coo->row[i] < 0 || coo->row[i] >= coo->nrows || coo->col[i] < 0 ||
            coo->col[i] >= coo->ncols)
            return set_error(VNM_E_ARG, "entry %lld out of range", (long long)i);
    VNM_HIP_TRY(hipSetDevice(ctx->device));
    vnm_graph *g = new vnm_graph();
    g->ctx = ctx;
    g->nrows = coo->nrows;
    g->ncols = coo->ncols;
    g->M = M;
    g->nnz = coo->nnz;
    int rc = upload(g, coo);
    if (!rc) rc = g->perm_total.alloc(g->nrows);
    if (rc) {
        delete g;
        return rc;
    }
    hipLaunchKernelGGL(k_iota, dim3((g->nrows + 255) / 256), dim3(256), 0, ctx->stream,
                       g->perm_total.p, g->nrows);
    ctx->launches++;
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        delete g;
        return set_error(VNM_E_HIP, "k_iota launch failed: %s", hipGetErrorString(e));
    }
    *out = g;
    return VNM_OK;
}

void vnm_graph_destroy(vnm_graph *g) {
    if (!g) return;
    if (g->ctx && g->ctx->stream) (void)hipStreamSynchronize(g->ctx->stream);
    delete g;
}

int vnm_graph_download(vnm_graph *g, vnm_coo *out) {
    if (!g || !out) return set_error(VNM_E_ARG, "null argument");
    HostCoo c;
    c.nrows = g->nrows;
    c.ncols = g->ncols;
    c.row.resize(g->nnz);
    c.col.resize(g->nnz);
    c.val.resize(g->nnz);
    hipStream_t st = g->ctx->stream;
    if (g->nnz) {
        VNM_HIP_TRY(hipMemcpyAsync(c.row.data(), g->row.p, g->nnz * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(c.col.data(), g->col.p, g->nnz * 4, hipMemcpyDeviceToHost, st));
        VNM_HIP_TRY(hipMemcpyAsync(c.val.data(), g->val.p, g->nnz * 4, hipMemcpyDeviceToHost, st));
    }
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return coo_export(c, out);
}

int vnm_graph_perm(vnm_graph *g, int32_t *perm) {
    if (!g || !perm) return set_error(VNM_E_ARG, "null argument");
    hipStream_t st = g->ctx->stream;
    VNM_HIP_TRY(hipMemcpyAsync(perm, g->perm_total.p, g->perm_total.n * 4, hipMemcpyDeviceToHost, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    return VNM_OK;
}

int vnm_graph_shape(vnm_graph *g, int32_t *nrows, int32_t *ncols, int64_t *nnz) {
    if (!g || !nrows || !ncols || !nnz) return set_error(VNM_E_ARG, "null argument");
    *nrows = g->nrows;
    *ncols = g->ncols;
    *nnz = g->nnz;
    return VNM_OK;
}

}  // extern "C"
