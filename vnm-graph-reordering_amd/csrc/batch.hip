// Batched V:N:M SpMM and the GCN aggregation layer (gfx950).
//
// Batch: many packed matrices of uneven size multiplied in one persistent
// launch.  The work items (64 values x one column tile) of all matrices
// form one list (prefix offsets per matrix); a fixed grid of 8 workgroups per
// CU walks it, each XCD taking a contiguous range so neighbouring items of
// one matrix share that XCD's L2.  A device table holds every matrix's
// pointers (SpmmArgs).  The batch is tuned once (values per lane, timed on the
// whole batch) and then captured as a HIP graph of two kernels -- the SpMM
// and the fix-up pass (rows cut into chunks, whose partial sums are added in
// item order) -- replayed by vnm_batch_run.
//
// GCN layer (the consumer of gnn/single/pyg_gcn.py:41-74, aggr='add'):
// Y = A_vnm (X W) + bias.  X W runs on MFMA (BF16 inputs, FP32 accumulation,
// BF16 result into a workspace kept with the packed matrix), then the SpMM
// with the bias folded into its row stores and fix-up pass.
#include "graph.h"
#include "hip_common.h"
#include "packed.h"

#include <algorithm>
#include <chrono>
#include <vector>

struct vnm_batch {
    vnm_ctx *ctx = nullptr;
    int count = 0, n = 0, bt = 0, at = 0, vec = 0, T = 1;
    int total_units = 0, nfix = 0, grid = 0, per_xcd = 0, waves_per_xcd = 0;
    vnm::DBuf<vnm::SpmmArgs> descs;
    vnm::DBuf<int32_t> unit_off;
    vnm::DBuf<int4> fix;          // {matrix, row, first slot, slots}
    vnm::DBuf<float> part;        // partial-sum slots of every matrix
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~vnm_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

namespace vnm {

template <int BT, int VEC>
__global__ __launch_bounds__(256) void k_spmm_batch(const SpmmArgs *__restrict__ descs,
                                                    const int32_t *__restrict__ unit_off, int count, int T,
                                                    int total, int per_xcd, int waves_per_xcd) {
    // wave-uniform; readfirstlane lets the unit's reads be scalar loads
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    for (int u = slot * 4 + wv; u < per_xcd; u += waves_per_xcd) {
        const int unit = xcd * per_xcd + u;
        if (unit >= total) break;  // wave-uniform: every wave reaches the end
        // matrix of this unit: the last g with unit_off[g] <= unit
        int lo = 0, hi = count - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (unit_off[mid] <= unit) lo = mid;
            else hi = mid - 1;
        }
        const int local = unit - unit_off[lo];
        spmm_unit<BT, VEC>(descs[lo], local / T, local % T, T);
    }
}

// Fix-up of the listed (matrix, row) pairs: cut rows get the sum of their
// slots in item order (rows of a matrix without work items, the bias or zero).
__global__ __launch_bounds__(256) void k_fix_batch(const int4 *__restrict__ list, int nf, int n,
                                                   const SpmmArgs *__restrict__ descs) {
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (w >= nf) return;
    const int4 e = list[w];
    const SpmmArgs &a = descs[e.x];
    fix_row(a.C + (size_t)e.y * a.ldc, a.part, a.pn, e.z, e.w, n, a.bias, false);
}

__global__ void k_tag_fix(const int4 *__restrict__ fix, int nf, int g, int4 *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nf) out[i] = make_int4(g, fix[i].x, fix[i].y, fix[i].z);
}

template <int BT, int VEC>
static void launch_batch_v(vnm_batch *b, hipStream_t st) {
    hipLaunchKernelGGL((k_spmm_batch<BT, VEC>), dim3(b->grid), dim3(256), 0, st, b->descs.p, b->unit_off.p,
                       b->count, b->T, b->total_units, b->per_xcd, b->waves_per_xcd);
}

static void launch_batch_b(vnm_batch *b, hipStream_t st) {
    const int v = b->vec;
    if (b->bt == VNM_DT_F32) {
        if (v == 4) launch_batch_v<VNM_DT_F32, 4>(b, st);
        else if (v == 2) launch_batch_v<VNM_DT_F32, 2>(b, st);
        else launch_batch_v<VNM_DT_F32, 1>(b, st);
    } else if (b->bt == VNM_DT_BF16) {
        if (v == 8) launch_batch_v<VNM_DT_BF16, 8>(b, st);
        else if (v == 4) launch_batch_v<VNM_DT_BF16, 4>(b, st);
        else if (v == 2) launch_batch_v<VNM_DT_BF16, 2>(b, st);
        else launch_batch_v<VNM_DT_BF16, 1>(b, st);
    } else {
        if (v == 16) launch_batch_v<VNM_DT_FP8, 16>(b, st);
        else if (v == 8) launch_batch_v<VNM_DT_FP8, 8>(b, st);
        else if (v == 4) launch_batch_v<VNM_DT_FP8, 4>(b, st);
        else launch_batch_v<VNM_DT_FP8, 2>(b, st);
    }
}

// Set the launch shape for `vec` values per lane: tile count, units, grid.
static void batch_shape(vnm_batch *b, const std::vector<int> &items, int vec, std::vector<int32_t> &off) {
    b->vec = vec;
    b->T = b->n / (64 * vec);
    off.assign(b->count + 1, 0);
    for (int g = 0; g < b->count; ++g) off[g + 1] = off[g] + items[g] * b->T;
    b->total_units = off[b->count];
    int ncu = 256;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->ctx->device) != hipSuccess || ncu < 8)
        ncu = 256;
    b->grid = std::max(8, (ncu * 8) / 8 * 8);  // 8 workgroups of 4 waves per CU
    b->per_xcd = (b->total_units + 7) / 8;
    b->waves_per_xcd = (b->grid / 8) * 4;
}

static int batch_enqueue(vnm_batch *b, hipStream_t st) {
    if (b->total_units > 0) {
        launch_batch_b(b, st);
    }
    if (b->nfix > 0)
        hipLaunchKernelGGL(k_fix_batch, dim3((b->nfix + 3) / 4), dim3(256), 0, st, b->fix.p, b->nfix, b->n,
                           b->descs.p);
    VNM_LAUNCH_CHECK("k_spmm_batch");
    return VNM_OK;
}

static int batch_create(vnm_ctx *ctx, vnm_packed *const *ps, int count, const void *const *B, const int32_t *ldb,
                        int n, int bt, const float *const *bs, float *const *C, const int32_t *ldc,
                        const float *const *bias, int vec, vnm_batch **out) {
    if (!ctx || !ps || count < 1 || !B || !ldb || !C || !ldc || !out) return set_error(VNM_E_ARG, "bad arguments");
    if (n <= 0 || n % 64) return set_error(VNM_E_ARG, "batch: n must be a positive multiple of 64");
    if (bt != VNM_DT_F32 && bt != VNM_DT_BF16 && bt != VNM_DT_FP8)
        return set_error(VNM_E_ARG, "batch: B dtype must be F32, BF16 or FP8");
    const int at = ps[0] ? ps[0]->dtype : 0;
    for (int g = 0; g < count; ++g) {
        const vnm_packed *p = ps[g];
        if (!p || p->ctx != ctx) return set_error(VNM_E_ARG, "batch: matrix %d missing or on another context", g);
        if (p->dtype != at) return set_error(VNM_E_ARG, "batch: every matrix needs the same value dtype");
        if (p->ngroups > 0)
            return set_error(VNM_E_ARG, "batch: matrix %d has a sparse-MFMA part (pack it with vnm_pack)", g);
        if (!B[g] || !C[g] || ldb[g] < n || ldc[g] < n || (bt == VNM_DT_FP8 && (!bs || !bs[g])))
            return set_error(VNM_E_ARG, "batch: operands of matrix %d", g);
        if ((ldb[g] * bsize_of(bt)) % 16 || ldc[g] % 4 || ((uintptr_t)B[g] & 15) || ((uintptr_t)C[g] & 15) ||
            (bias && bias[g] && ((uintptr_t)bias[g] & 15)))
            return set_error(VNM_E_ARG, "batch: matrix %d operands must be 16-byte aligned", g);
        if ((uint64_t)p->ncols * (uint64_t)ldb[g] * (uint64_t)bsize_of(bt) > 0xFFFFFFFFull)
            return set_error(VNM_E_ARG, "batch: B of matrix %d must be smaller than 4 GiB", g);
    }
    if (vec && !vec_ok(bt, n, vec)) return set_error(VNM_E_ARG, "batch: %d values per lane do not fit n", vec);
    hipStream_t st = ctx->stream;
    auto b = std::make_unique<vnm_batch>();
    b->ctx = ctx;
    b->count = count;
    b->n = n;
    b->bt = bt;
    b->at = at;
    std::vector<SpmmArgs> d(count);
    std::vector<int> items(count);
    std::vector<size_t> part_off(count + 1, 0);
    std::vector<int> fix_off(count + 1, 0);
    int rc;
    for (int g = 0; g < count; ++g) {
        items[g] = ps[g]->nitems;
        part_off[g + 1] = part_off[g] + (size_t)ps[g]->nslots * n;
        fix_off[g + 1] = fix_off[g] + ps[g]->nfix;
    }
    b->nfix = fix_off[count];
    if ((rc = b->descs.alloc(count)) || (rc = b->unit_off.alloc(count + 1)) ||
        (rc = b->fix.alloc(std::max(b->nfix, 1))) || (rc = b->part.alloc(std::max<size_t>(part_off[count], 1))))
        return rc;
    for (int g = 0; g < count; ++g)
        d[g] = packed_args(ps[g], B[g], ldb[g], bs ? bs[g] : nullptr, C[g], ldc[g], bias ? bias[g] : nullptr,
                           b->part.p + part_off[g], n);
    VNM_HIP_TRY(hipMemcpyAsync(b->descs.p, d.data(), count * sizeof(SpmmArgs), hipMemcpyHostToDevice, st));
    for (int g = 0; g < count; ++g)
        if (ps[g]->nfix)
            hipLaunchKernelGGL(k_tag_fix, dim3((ps[g]->nfix + 255) / 256), dim3(256), 0, st, ps[g]->fix.p,
                               ps[g]->nfix, g, b->fix.p + fix_off[g]);
    VNM_LAUNCH_CHECK("k_tag_fix");
    std::vector<int32_t> off;
    if (vec == 0) {
        // tuner: each usable shape, one warm-up and three timed batches
        hipEvent_t e0, e1;
        VNM_HIP_TRY(hipEventCreate(&e0));
        VNM_HIP_TRY(hipEventCreate(&e1));
        float best = 1e30f;
        for (int v = 16; v >= 1; v /= 2) {
            if (!vec_ok(bt, n, v)) continue;
            batch_shape(b.get(), items, v, off);
            VNM_HIP_TRY(hipMemcpyAsync(b->unit_off.p, off.data(), (count + 1) * 4, hipMemcpyHostToDevice, st));
            if ((rc = batch_enqueue(b.get(), st))) break;
            (void)hipEventRecord(e0, st);
            for (int i = 0; i < 3 && rc == VNM_OK; ++i) rc = batch_enqueue(b.get(), st);
            (void)hipEventRecord(e1, st);
            float ms = 0.f;
            if (rc || hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                if (!rc) rc = set_error(VNM_E_HIP, "batch tuning timing failed");
                break;
            }
            if (getenv("VNM_DEBUG")) fprintf(stderr, "[vnm] batch tune bt=%d n=%d vec=%d %.4f ms\n", bt, n, v, ms / 3);
            if (ms < best) { best = ms; vec = v; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (rc) return rc;
    }
    batch_shape(b.get(), items, vec, off);
    VNM_HIP_TRY(hipMemcpyAsync(b->unit_off.p, off.data(), (count + 1) * 4, hipMemcpyHostToDevice, st));
    VNM_HIP_TRY(hipStreamSynchronize(st));
    // capture the two kernels once; vnm_batch_run replays them
    VNM_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    rc = batch_enqueue(b.get(), st);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_error(VNM_E_HIP, "batch capture failed: %s", hipGetErrorString(e));
    b->graph = graph;
    e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return set_error(VNM_E_HIP, "batch graph instantiation failed: %s", hipGetErrorString(e));
    ctx->launches += 2;
    *out = b.release();
    return VNM_OK;
}

// ------------------------------------------------------------ GCN layer

// Wt[N x K] = W[K x N]^T (BF16), so that the GEMM's LDS stage is a plain
// 16-byte copy.  Tiles of 64 x 64 through LDS, both sides coalesced.
__global__ __launch_bounds__(256) void k_transpose_w(int K, int N, const uint16_t *__restrict__ W,
                                                     uint16_t *__restrict__ Wt) {
    __shared__ uint16_t t[64][65];
    const int k0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
    for (int i = threadIdx.x; i < 64 * 64; i += 256) {
        const int r = i / 64, c = i % 64;
        if (k0 + r < K && n0 + c < N) t[r][c] = W[(size_t)(k0 + r) * N + n0 + c];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * 64; i += 256) {
        const int r = i / 64, c = i % 64;  // r: column of W, c: k
        if (n0 + r < N && k0 + c < K) Wt[(size_t)(n0 + r) * K + k0 + c] = t[c][r];
    }
}

// H[R x N] (BF16) = X[R x K] (BF16, leading dimension ldx) W[K x N], with
// W given transposed (Wt, N x K); K a multiple of 16, N of 64.  A workgroup
// takes RT row tiles of 128 rows of one XCD for one 64-column panel; a wave
// computes 32 rows x 64 columns of each (two row fragments share every W
// fragment), GK k at a time.  Wt's 64-row panel is staged in LDS (rows
// padded by 16 bytes, so the 16 lanes reading one k range hit different
// banks), once per workgroup when K <= GK.  gfx950's v_mfma_f32_16x16x32_bf16
// takes the W fragment as its first operand (lane l: column n = l%16 of the
// panel, k 8(l/16)..+7) and the X fragment as its second (row l%16, the same
// k), so it produces H transposed: lane l ends with 4 consecutive columns of
// one row, stored as 8 bytes.  In step s lane group g = l/16 takes k
// 32s + 8g .. +7, so one load covers 64 contiguous bytes of each of 16 rows;
// a tile's loads are all issued before its first MFMA, and the X loads of
// the first tile are in flight while the panel is staged (the barriers wait
// for LDS only).  Groups past the stage's k count (a K that is an odd
// multiple of 16) hold zeros in both operands.  The N/64 column panels of
// a row tile are consecutive workgroups of one XCD, so the X tile is fetched
// from HBM once and hit in that XCD's L2 by the other panels.
constexpr int GK = 256;
constexpr int GKP = GK + 8;
constexpr int GS = GK / 32;  // MFMA steps of a full stage
constexpr int GCP = 64 * (GK / 8) / 256;  // 16-byte panel copies per thread and stage

// A barrier for LDS data only: pending global loads stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// f32_to_bf16 without branches (round to nearest even, NaN kept quiet).
__device__ __forceinline__ uint32_t bf16_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    const uint32_t r = (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
    return (u & 0x7FFFFFFFu) > 0x7F800000u ? ((u >> 16) | 0x40u) : r;
}

__global__ __launch_bounds__(256) void k_gemm_xw(int R, int K, int N, const uint16_t *__restrict__ X, int ldx,
                                                 const uint16_t *__restrict__ Wt, uint16_t *__restrict__ H,
                                                 int RT) {
    typedef short bf16x4_t __attribute__((ext_vector_type(4)));
    typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
    typedef float f32x4_t __attribute__((ext_vector_type(4)));
    __shared__ __align__(16) uint16_t w[64 * GKP];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    // workgroups are dealt round-robin over the 8 XCDs: b -> XCD b % 8
    const int P = N / 64;
    const int xcd = blockIdx.x % 8, jx = blockIdx.x / 8;
    const int panel = jx % P, tgroup = jx / P;
    const int n0 = panel * 64;
    const int m = lane & 15, g = lane >> 4;
    const bool once = K <= GK;  // the whole panel fits one stage: staged once
    if (once) {
        if (K == GK) {  // every copy's load issued first
            uint4 cp[GCP];
#pragma unroll
            for (int j = 0; j < GCP; ++j) {
                const int i = threadIdx.x + 256 * j;
                cp[j] = *(const uint4 *)(Wt + (size_t)(n0 + i / (GK / 8)) * K + (i % (GK / 8)) * 8);
            }
#pragma unroll
            for (int j = 0; j < GCP; ++j) {
                const int i = threadIdx.x + 256 * j;
                *(uint4 *)&w[(i / (GK / 8)) * GKP + (i % (GK / 8)) * 8] = cp[j];
            }
        } else {
            const int per_row = K / 8;
            for (int i = threadIdx.x; i < 64 * per_row; i += 256)
                *(uint4 *)&w[(i / per_row) * GKP + (i % per_row) * 8] =
                    *(const uint4 *)(Wt + (size_t)(n0 + i / per_row) * K + (i % per_row) * 8);
        }
        lds_barrier();
    }
    for (int t = 0; t < RT; ++t) {
        const int tile = (tgroup * RT + t) * 8 + xcd;
        const int r0 = (tile * 4 + wv) * 32;
        if (once && (tile * 4) * 32 >= R) break;  // workgroup-uniform: no later tile has rows
        f32x4_t acc[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[h][nt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
        for (int kc = 0; kc < K; kc += GK) {
            const int kn = min(GK, K - kc);  // a multiple of 16
            const int S = (kn + 31) / 32;    // steps of 32 k
            bf16x8_t a[2][GS];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int row = r0 + h * 16 + m;
#pragma unroll
                for (int st = 0; st < GS; ++st) {
                    bf16x4_t lo = bf16x4_t{0, 0, 0, 0}, hi = bf16x4_t{0, 0, 0, 0};
                    const int k0 = 32 * st + 8 * g;
                    if (st < S && k0 < kn && row < R) {
                        const uint16_t *xp = X + (size_t)row * ldx + kc + k0;
                        lo = *(const bf16x4_t *)xp;  // ldx % 4 == 0: 8-byte aligned
                        hi = *(const bf16x4_t *)(xp + 4);
                    }
                    a[h][st] = __builtin_bit_cast(bf16x8_t, __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7));
                }
            }
            if (!once) {  // this chunk of the panel, after its earlier readers
                const int per_row = kn / 8;
                lds_barrier();
                for (int i = threadIdx.x; i < 64 * per_row; i += 256)
                    *(uint4 *)&w[(i / per_row) * GKP + (i % per_row) * 8] =
                        *(const uint4 *)(Wt + (size_t)(n0 + i / per_row) * K + kc + (i % per_row) * 8);
                lds_barrier();
            }
            if (r0 < R) {
#pragma unroll
                for (int st = 0; st < GS; ++st) {
                    if (st >= S) break;
                    const int k0 = 32 * st + 8 * g;
                    const bool live = k0 < kn;
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt) {
                        bf16x8_t bb = bf16x8_t{};
                        if (live) bb = *(const bf16x8_t *)&w[(nt * 16 + m) * GKP + k0];
#pragma unroll
                        for (int h = 0; h < 2; ++h)
                            acc[h][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bb, a[h][st], acc[h][nt], 0, 0, 0);
                    }
                }
            }
        }
        // lane (m, g): row r0 + 16h + m, columns n0 + 16nt + 4g .. +3
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int rr = r0 + h * 16 + m;
            if (rr < R) {
#pragma unroll
                for (int nt = 0; nt < 4; ++nt) {
                    const uint2 v = make_uint2(bf16_bits(acc[h][nt][0]) | (bf16_bits(acc[h][nt][1]) << 16),
                                               bf16_bits(acc[h][nt][2]) | (bf16_bits(acc[h][nt][3]) << 16));
                    *(uint2 *)(H + (size_t)rr * N + n0 + nt * 16 + g * 4) = v;
                }
            }
        }
    }
}

static int gcn_layer(vnm_packed *p, const uint16_t *X, int ldx, int K, const uint16_t *W, int N,
                     const float *bias, float *Y, int ldy) {
    if (!p || !X || !W || !Y) return set_error(VNM_E_ARG, "gcn: null argument");
    if (K <= 0 || K % 16 || N <= 0 || N % 64 || ldx < K || ldx % 4 || ldy < N || ldy % 4)
        return set_error(VNM_E_ARG, "gcn: K must be a multiple of 16, N of 64, ldx >= K and ldy >= N (multiples of 4)");
    if (((uintptr_t)X & 7) || ((uintptr_t)Y & 15) || (bias && ((uintptr_t)bias & 15)))
        return set_error(VNM_E_ARG, "gcn: X must be 8-byte and Y, bias 16-byte aligned");
    const int R = p->ncols;  // X W has one row per column of A
    hipStream_t st = p->ctx->stream;
    int rc;
    if (p->gcn_h.n < (size_t)R * N + (size_t)K * N) {  // X W, then W transposed
        p->gcn_h.release();
        if ((rc = p->gcn_h.alloc((size_t)R * N + (size_t)K * N))) return rc;
    }
    uint16_t *Wt = p->gcn_h.p + (size_t)R * N;  // 16-byte aligned: R * N is a multiple of 64
    hipLaunchKernelGGL(k_transpose_w, dim3((K + 63) / 64, N / 64), dim3(256), 0, st, K, N, W, Wt);
    p->ctx->launches++;
    VNM_LAUNCH_CHECK("k_transpose_w");
    KernelTimer kt(p->ctx, "k_gemm_xw");
    // row tiles of 128 per XCD, RT of them per workgroup and column panel:
    // about 4 workgroups per CU, so each stages its W panel once for RT tiles
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, p->ctx->device) != hipSuccess || ncu < 8)
        ncu = 8;
    const int P = N / 64;
    const int tiles_x = ((R + 127) / 128 + 7) / 8;  // row tiles per XCD
    const int RT = K <= GK ? std::max(1, std::min(tiles_x, tiles_x * P * 8 / (4 * ncu))) : 1;
    const int groups = (tiles_x + RT - 1) / RT;
    hipLaunchKernelGGL(k_gemm_xw, dim3(8 * P * groups), dim3(256), 0, st, R, K, N, X, ldx, Wt, p->gcn_h.p, RT);
    kt.stop();
    p->ctx->launches++;
    VNM_LAUNCH_CHECK("k_gemm_xw");
    return spmm(p, p->gcn_h.p, N, N, VNM_DT_BF16, nullptr, Y, ldy, bias);
}

}  // namespace vnm

using namespace vnm;

extern "C" {

int vnm_batch_create(vnm_ctx *ctx, vnm_packed *const *ps, int count, const void *const *B, const int32_t *ldb,
                     int n, int b_dtype, const float *const *b_scales, float *const *C, const int32_t *ldc,
                     const float *const *bias, int cols_per_lane, vnm_batch **out) {
    return batch_create(ctx, ps, count, B, ldb, n, b_dtype, b_scales, C, ldc, bias, cols_per_lane, out);
}

int vnm_batch_run(vnm_batch *b) {
    if (!b) return set_error(VNM_E_ARG, "null batch");
    KernelTimer kt(b->ctx, "k_spmm_batch");
    hipError_t e = hipGraphLaunch(b->exec, b->ctx->stream);
    kt.stop();
    if (e != hipSuccess) return set_error(VNM_E_HIP, "batch graph launch failed: %s", hipGetErrorString(e));
    b->ctx->launches += 2;
    return VNM_OK;
}

int vnm_batch_info(const vnm_batch *b, int32_t *cols_per_lane, int32_t *units, int32_t *zero_rows) {
    if (!b) return set_error(VNM_E_ARG, "null batch");
    if (cols_per_lane) *cols_per_lane = b->vec;
    if (units) *units = b->total_units;
    if (zero_rows) *zero_rows = b->nfix;
    return VNM_OK;
}

void vnm_batch_destroy(vnm_batch *b) {
    if (!b) return;
    if (b->ctx && b->ctx->stream) (void)hipStreamSynchronize(b->ctx->stream);
    delete b;
}

int vnm_gcn_layer(vnm_packed *p, const uint16_t *X, int ldx, int K, const uint16_t *W, int N, const float *bias,
                  float *Y, int ldy) {
    return gcn_layer(p, X, ldx, K, W, N, bias, Y, ldy);
}

}  // extern "C"
