// Batched ensemble: many independent systems of uneven sizes stepped together.
//
// Every system is its own Between(system, system) brute force (all bodies massive within a
// system), integrated with kick-drift-kick leapfrog in FP64 like the single-system loop. The force
// pass of a step is ONE persistent launch: a few workgroups per CU whose wavefronts pull work
// items (system, tile of 64*R affected bodies) from a queue until it is empty. Items are ordered by
// decreasing cost (system size x tile size), so large systems start first and small ones fill the
// tail. R (affected bodies per lane) is chosen per system-size bucket by the autotuner. Affecting
// bodies are read with wave-uniform addresses (scalar loads), so wavefronts need no barriers and
// run independently. The per-lane arithmetic and summation order are those of the brute-force
// kernel (FP32 tiles of 256 bodies, Kahan across tiles). A step is enqueue-only and can be
// captured into a hipGraph (the queue counter is cleared by a kernel, not a memset).
#include "pnb_internal.hpp"

#include <algorithm>
#include <vector>

namespace pnb {

constexpr int BT_TILE = 256;

struct BatchItem {
    uint32_t off;  // first particle of the system
    uint32_t n;    // system size
    uint32_t i0;   // first affected body of the tile, relative to the system
    uint32_t r;    // affected bodies per lane (1, 2, 4 or 8)
};

struct BatchState {
    size_t total = 0, systems = 0;
    float softening = 0.f;
    int checked = 1;
    std::vector<int64_t> offsets;
    std::vector<std::pair<size_t, int>> tuning;  // (n_min, per_lane)
    bool items_dirty = true;
    size_t n_items = 0;
    unsigned long long interactions = 0;
    DevBuf x, v, shadow, acc, items, queue, stage;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    size_t graph_steps = 0;
    double graph_dt = 0.0;
    hipStream_t graph_stream = nullptr;
    void drop_graph() {
        if (exec) hipGraphExecDestroy(exec);
        if (graph) hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_steps = 0;
    }
};

void batch_free(BatchState* s) {
    if (!s) return;
    s->drop_graph();
    for (DevBuf* b : {&s->x, &s->v, &s->shadow, &s->acc, &s->items, &s->queue, &s->stage}) b->release();
    delete s;
}

__device__ __forceinline__ void bt_kahan(float& sum, float& comp, float v) {
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

template <int R, bool MASK>
__device__ __forceinline__ void bt_tile(const float4* __restrict__ sh, const BatchItem w, float eps2,
                                        float4* __restrict__ acc) {
    const uint32_t lane = threadIdx.x & 63;
    const float4* sys = sh + w.off;
    float xi[R], yi[R], zi[R], ax[R], ay[R], az[R], cx[R], cy[R], cz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = w.i0 + r * 64 + lane;
        const float4 p = sys[i < w.n ? i : w.n - 1];
        xi[r] = p.x; yi[r] = p.y; zi[r] = p.z;
        ax[r] = ay[r] = az[r] = cx[r] = cy[r] = cz[r] = 0.f;
    }
    for (uint32_t j0 = 0; j0 < w.n; j0 += BT_TILE) {
        const uint32_t cnt = min(uint32_t(BT_TILE), w.n - j0);
        float px[R], py[R], pz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) px[r] = py[r] = pz[r] = 0.f;
        for (uint32_t j = 0; j < cnt; ++j) {
            const float4 q = sys[j0 + j];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const float dx = q.x - xi[r], dy = q.y - yi[r], dz = q.z - zi[r];
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
                float inv = __builtin_amdgcn_rsqf(r2);
                if (MASK) inv = (r2 > 0.0f) ? inv : 0.0f;
                const float s = (q.w * inv) * (inv * inv);
                px[r] = fmaf(dx, s, px[r]);
                py[r] = fmaf(dy, s, py[r]);
                pz[r] = fmaf(dz, s, pz[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            bt_kahan(ax[r], cx[r], px[r]);
            bt_kahan(ay[r], cy[r], py[r]);
            bt_kahan(az[r], cz[r], pz[r]);
        }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = w.i0 + r * 64 + lane;
        if (i < w.n) acc[w.off + i] = make_float4(ax[r], ay[r], az[r], 0.f);
    }
}

template <bool MASK>
__global__ __launch_bounds__(256) void bt_force_kernel(const BatchItem* __restrict__ items,
                                                       uint32_t n_items,
                                                       uint32_t* __restrict__ queue,
                                                       const float4* __restrict__ sh, float eps2,
                                                       float4* __restrict__ acc) {
    const uint32_t lane = threadIdx.x & 63;
    while (true) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(queue, 1u);
        it = __builtin_amdgcn_readfirstlane(it);
        if (it >= n_items) break;
        const BatchItem w = items[it];
        switch (w.r) {
            case 8: bt_tile<8, MASK>(sh, w, eps2, acc); break;
            case 4: bt_tile<4, MASK>(sh, w, eps2, acc); break;
            case 2: bt_tile<2, MASK>(sh, w, eps2, acc); break;
            default: bt_tile<1, MASK>(sh, w, eps2, acc); break;
        }
    }
}

__global__ void bt_queue_reset_kernel(uint32_t* q) { *q = 0u; }

__global__ void bt_pack_kernel(const double* __restrict__ x3, const double* __restrict__ v3,
                               const float* __restrict__ mu, size_t n, double4* __restrict__ x,
                               double4* __restrict__ v, float4* __restrict__ sh) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = make_double4(x3[3 * i], x3[3 * i + 1], x3[3 * i + 2], 0.0);
    v[i] = make_double4(v3[3 * i], v3[3 * i + 1], v3[3 * i + 2], 0.0);
    sh[i] = make_float4(float(x3[3 * i]), float(x3[3 * i + 1]), float(x3[3 * i + 2]), mu[i]);
}

__global__ void bt_kick_drift_kernel(double4* __restrict__ x, double4* __restrict__ v,
                                     const float4* __restrict__ acc, float4* __restrict__ sh,
                                     size_t n, double h, double dt) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = acc[i];
    double4 vi = v[i];
    vi.x += double(a.x) * h;
    vi.y += double(a.y) * h;
    vi.z += double(a.z) * h;
    double4 xi = x[i];
    xi.x += vi.x * dt;
    xi.y += vi.y * dt;
    xi.z += vi.z * dt;
    v[i] = vi;
    x[i] = xi;
    sh[i] = make_float4(float(xi.x), float(xi.y), float(xi.z), sh[i].w);
}

__global__ void bt_kick_kernel(double4* __restrict__ v, const float4* __restrict__ acc, size_t n,
                               double h) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = acc[i];
    double4 vi = v[i];
    vi.x += double(a.x) * h;
    vi.y += double(a.y) * h;
    vi.z += double(a.z) * h;
    v[i] = vi;
}

__global__ void bt_unpack_kernel(const double4* __restrict__ x, const double4* __restrict__ v,
                                 size_t n, double* __restrict__ out) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 a = x[i], b = v[i];
    out[3 * i] = a.x; out[3 * i + 1] = a.y; out[3 * i + 2] = a.z;
    out[3 * (n + i)] = b.x; out[3 * (n + i) + 1] = b.y; out[3 * (n + i) + 2] = b.z;
}

static int per_lane_for(const BatchState* s, size_t n) {
    int r = 0;
    size_t best = 0;
    for (const auto& t : s->tuning)
        if (t.first <= n && (r == 0 || t.first >= best)) {
            best = t.first;
            r = t.second;
        }
    if (r) return r;
    return n >= 256 ? 4 : 1;  // untuned default (what autotuning picks on MI355X)
}

static int build_items(pnb_ctx* c, BatchState* s) {
    std::vector<BatchItem> items;
    std::vector<uint64_t> cost;
    s->interactions = 0;
    for (size_t k = 0; k < s->systems; ++k) {
        const uint32_t off = uint32_t(s->offsets[k]);
        const uint32_t n = uint32_t(s->offsets[k + 1] - s->offsets[k]);
        if (n == 0) continue;
        s->interactions += (unsigned long long)n * n;
        const uint32_t r = uint32_t(per_lane_for(s, n));
        const uint32_t tile = 64 * r;
        for (uint32_t i0 = 0; i0 < n; i0 += tile) items.push_back({off, n, i0, r});
    }
    std::stable_sort(items.begin(), items.end(), [](const BatchItem& a, const BatchItem& b) {
        return uint64_t(a.n) * a.r > uint64_t(b.n) * b.r;
    });
    s->n_items = items.size();
    PNB_HIP(c, s->items.ensure(std::max<size_t>(1, items.size()) * sizeof(BatchItem)));
    if (!items.empty())
        PNB_HIP(c, hipMemcpyAsync(s->items.ptr, items.data(), items.size() * sizeof(BatchItem),
                                  hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    s->items_dirty = false;
    return PNB_OK;
}

static int force(pnb_ctx* c, BatchState* s) {
    const bool mask = s->checked && s->softening == 0.0f;
    const float eps2 = s->softening * s->softening;
    int occ = 0;
    hipError_t e = mask ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bt_force_kernel<true>, 256, 0)
                        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, bt_force_kernel<false>, 256, 0);
    if (e != hipSuccess || occ < 1) occ = 2;
    const unsigned grid = unsigned(std::max(1, c->cu_count) * occ);
    bt_queue_reset_kernel<<<1, 1, 0, c->stream>>>(s->queue.as<uint32_t>());
    PNB_CHECK_LAUNCH(c);
    return timed(c, PNB_PROF_BATCH, [&]() -> int {
        if (mask)
            bt_force_kernel<true><<<grid, 256, 0, c->stream>>>(s->items.as<BatchItem>(), uint32_t(s->n_items),
                                                              s->queue.as<uint32_t>(), s->shadow.as<float4>(),
                                                              eps2, s->acc.as<float4>());
        else
            bt_force_kernel<false><<<grid, 256, 0, c->stream>>>(s->items.as<BatchItem>(), uint32_t(s->n_items),
                                                               s->queue.as<uint32_t>(), s->shadow.as<float4>(),
                                                               eps2, s->acc.as<float4>());
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    });
}

static int enqueue(pnb_ctx* c, BatchState* s, size_t steps, double dt) {
    const unsigned g = div_up(s->total, 256);
    const double h = 0.5 * dt;
    for (size_t k = 0; k < steps; ++k) {
        bt_kick_drift_kernel<<<g, 256, 0, c->stream>>>(s->x.as<double4>(), s->v.as<double4>(),
                                                       s->acc.as<float4>(), s->shadow.as<float4>(),
                                                       s->total, h, dt);
        PNB_CHECK_LAUNCH(c);
        if (int rc = force(c, s)) return rc;
        bt_kick_kernel<<<g, 256, 0, c->stream>>>(s->v.as<double4>(), s->acc.as<float4>(), s->total, h);
        PNB_CHECK_LAUNCH(c);
    }
    return PNB_OK;
}

}  // namespace pnb

using namespace pnb;

extern "C" {

int pnb_batch_set_tuning(pnb_ctx* c, size_t n_min, int per_lane) {
    if (!c) return PNB_E_ARG;
    if (per_lane != 1 && per_lane != 2 && per_lane != 4 && per_lane != 8)
        return fail(c, PNB_E_ARG, "per_lane must be 1, 2, 4 or 8");
    if (!c->batch) c->batch = new BatchState();
    auto& t = c->batch->tuning;
    t.erase(std::remove_if(t.begin(), t.end(), [&](const std::pair<size_t, int>& e) { return e.first == n_min; }),
            t.end());
    t.emplace_back(n_min, per_lane);
    c->batch->items_dirty = true;
    c->batch->drop_graph();
    return PNB_OK;
}

int pnb_batch_init(pnb_ctx* c, const double* pos3, const double* vel3, const float* mu,
                   const int64_t* offsets, size_t num_systems, float softening, int checked) {
    if (!c) return PNB_E_ARG;
    if (!offsets || num_systems == 0) return fail(c, PNB_E_ARG, "no systems");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    for (size_t k = 0; k < num_systems; ++k)
        if (offsets[k + 1] < offsets[k] || offsets[0] != 0)
            return fail(c, PNB_E_ARG, "offsets must start at 0 and be non-decreasing");
    const size_t total = size_t(offsets[num_systems]);
    if (total == 0 || !pos3 || !vel3 || !mu) return fail(c, PNB_E_ARG, "no particles");
    if (total > 0x7fffffffu) return fail(c, PNB_E_ARG, "too many particles");
    if (int rc = activate(c)) return rc;
    if (!c->batch) c->batch = new BatchState();
    BatchState* s = c->batch;
    s->drop_graph();
    s->total = total;
    s->systems = num_systems;
    s->softening = softening;
    s->checked = checked;
    s->offsets.assign(offsets, offsets + num_systems + 1);
    PNB_HIP(c, s->x.ensure(total * sizeof(double4)));
    PNB_HIP(c, s->v.ensure(total * sizeof(double4)));
    PNB_HIP(c, s->shadow.ensure(total * sizeof(float4)));
    PNB_HIP(c, s->acc.ensure(total * sizeof(float4)));
    PNB_HIP(c, s->queue.ensure(sizeof(uint32_t)));
    PNB_HIP(c, s->stage.ensure(total * (6 * sizeof(double) + sizeof(float))));
    double* dx = s->stage.as<double>();
    double* dv = dx + 3 * total;
    float* dm = reinterpret_cast<float*>(dv + 3 * total);
    PNB_HIP(c, hipMemcpyAsync(dx, pos3, 3 * total * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(dv, vel3, 3 * total * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(dm, mu, total * sizeof(float), hipMemcpyHostToDevice, c->stream));
    bt_pack_kernel<<<div_up(total, 256), 256, 0, c->stream>>>(dx, dv, dm, total, s->x.as<double4>(),
                                                              s->v.as<double4>(), s->shadow.as<float4>());
    PNB_CHECK_LAUNCH(c);
    if (int rc = build_items(c, s)) return rc;
    if (int rc = force(c, s)) return rc;
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_batch_run(pnb_ctx* c, size_t steps, double dt, int use_graph) {
    if (!c) return PNB_E_ARG;
    BatchState* s = c->batch;
    if (!s || s->total == 0) return fail(c, PNB_E_STATE, "pnb_batch_init has not been called");
    if (steps == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    if (s->items_dirty)
        if (int rc = build_items(c, s)) return rc;
    if (!use_graph) return enqueue(c, s, steps, dt);
    if (!c->stream) return fail(c, PNB_E_STATE, "graph capture needs a non-null stream");
    if (!s->exec || s->graph_steps != steps || s->graph_dt != dt || s->graph_stream != c->stream) {
        s->drop_graph();
        PNB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        int rc = enqueue(c, s, steps, dt);
        c->capturing = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) {
            if (g) hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture", __FILE__, __LINE__);
        s->graph = g;
        PNB_HIP(c, hipGraphInstantiate(&s->exec, g, nullptr, nullptr, 0));
        s->graph_steps = steps;
        s->graph_dt = dt;
        s->graph_stream = c->stream;
    }
    PNB_HIP(c, hipGraphLaunch(s->exec, c->stream));
    return PNB_OK;
}

int pnb_batch_read(pnb_ctx* c, double* pos3, double* vel3, float* acc4) {
    if (!c) return PNB_E_ARG;
    BatchState* s = c->batch;
    if (!s || s->total == 0) return fail(c, PNB_E_STATE, "pnb_batch_init has not been called");
    if (int rc = activate(c)) return rc;
    const size_t n = s->total;
    double* st = s->stage.as<double>();
    bt_unpack_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(s->x.as<double4>(), s->v.as<double4>(), n, st);
    PNB_CHECK_LAUNCH(c);
    if (pos3) PNB_HIP(c, hipMemcpyAsync(pos3, st, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (vel3) PNB_HIP(c, hipMemcpyAsync(vel3, st + 3 * n, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (acc4) PNB_HIP(c, hipMemcpyAsync(acc4, s->acc.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_batch_info(pnb_ctx* c, size_t* particles, size_t* items,
                   unsigned long long* interactions_per_step) {
    if (!c) return PNB_E_ARG;
    BatchState* s = c->batch;
    if (!s || s->total == 0) return fail(c, PNB_E_STATE, "pnb_batch_init has not been called");
    if (particles) *particles = s->total;
    if (items) *items = s->n_items;
    if (interactions_per_step) *interactions_per_step = s->interactions;
    return PNB_OK;
}

}  // extern "C"
