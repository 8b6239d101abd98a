// Context, errors, profiling, device-wide scan and the massive/massless split.
#include "pnb_internal.hpp"

#include <new>

namespace pnb {

hipError_t DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && ptr) return hipSuccess;
    release();
    size_t want = bytes < 256 ? 256 : bytes;
    want = (want + 4095) & ~size_t(4095);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        cap = 0;
        (void)hipGetLastError();
        return e;
    }
    cap = want;
    return hipSuccess;
}

void DevBuf::release() {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

int fail(pnb_ctx* c, int code, const char* fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        c->last_error = buf;
    }
    return code;
}

int hip_fail(pnb_ctx* c, hipError_t e, const char* expr, const char* file, int line) {
    (void)hipGetLastError();
    int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PNB_E_OOM : PNB_E_HIP;
    return fail(c, code, "%s failed: %s (%s:%d)", expr, hipGetErrorString(e), file, line);
}

int prof_resolve(pnb_ctx* c) {
    if (c->ev_used == 0) return PNB_OK;
    PNB_HIP(c, hipEventSynchronize(c->ev_stop[c->ev_used - 1]));
    for (int k = 0; k < c->ev_used; ++k) {
        float ms = 0.f;
        PNB_HIP(c, hipEventElapsedTime(&ms, c->ev_start[k], c->ev_stop[k]));
        c->prof_ms[c->ev_kernel[k]] += ms;
        c->prof_count[c->ev_kernel[k]] += 1;
    }
    c->ev_used = 0;
    return PNB_OK;
}

int activate(pnb_ctx* c) {
    PNB_HIP(c, hipSetDevice(c->device));
    return PNB_OK;
}

// ---- exclusive scan (reduce, scan block sums recursively, add) --------------------------------

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* total) {
    __shared__ uint32_t wsum[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < SCAN_BLOCK / 64; ++w) {
        uint32_t s = wsum[w];
        if (w < wid) before += s;
        all += s;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_reduce_kernel(const uint32_t* __restrict__ in,
                                                                 size_t n,
                                                                 uint32_t* __restrict__ sums) {
    size_t base = size_t(blockIdx.x) * SCAN_TILE + size_t(threadIdx.x) * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < n) s += in[base + k];
    uint32_t total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(SCAN_BLOCK) void scan_apply_kernel(const uint32_t* in, size_t n,
                                                                const uint32_t* __restrict__ offs,
                                                                uint32_t* out, int with_total) {
    size_t base = size_t(blockIdx.x) * SCAN_TILE + size_t(threadIdx.x) * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        v[k] = (base + k < n) ? in[base + k] : 0u;
        s += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, &total) + (offs ? offs[blockIdx.x] : 0u);
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
    if (with_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == SCAN_BLOCK - 1) out[n] = run;
}

static int scan_rec(pnb_ctx* c, const uint32_t* in, uint32_t* out, size_t n, bool with_total,
                    uint32_t* tmp, size_t tmp_cap) {
    unsigned blocks = div_up(n, SCAN_TILE);
    if (blocks <= 1) {
        scan_apply_kernel<<<1, SCAN_BLOCK, 0, c->stream>>>(in, n, nullptr, out, with_total ? 1 : 0);
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    }
    if (tmp_cap < blocks) return fail(c, PNB_E_HIP, "scan scratch too small");
    uint32_t* sums = tmp;
    scan_reduce_kernel<<<blocks, SCAN_BLOCK, 0, c->stream>>>(in, n, sums);
    PNB_CHECK_LAUNCH(c);
    int rc = scan_rec(c, sums, sums, blocks, false, tmp + blocks, tmp_cap - blocks);
    if (rc) return rc;
    scan_apply_kernel<<<blocks, SCAN_BLOCK, 0, c->stream>>>(in, n, sums, out, with_total ? 1 : 0);
    PNB_CHECK_LAUNCH(c);
    return PNB_OK;
}

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0u; }

int scan_u32(pnb_ctx* c, const uint32_t* in, uint32_t* out, size_t n, bool with_total) {
    if (n == 0) {
        if (with_total) {
            zero_u32_kernel<<<1, 1, 0, c->stream>>>(out);
            PNB_CHECK_LAUNCH(c);
        }
        return PNB_OK;
    }
    size_t need = 0, level = n;
    while (level > SCAN_TILE) {
        level = div_up(level, SCAN_TILE);
        need += level;
    }
    need += 16;
    PNB_HIP(c, c->scan_tmp.ensure(need * sizeof(uint32_t)));
    return scan_rec(c, in, out, n, with_total, c->scan_tmp.as<uint32_t>(), need);
}

// ---- massive/massless split ------------------------------------------------------------------

__global__ void affecting_flags_kernel(const float4* __restrict__ p, size_t n,
                                       uint32_t* __restrict__ flags) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (p[i].w != 0.0f) ? 1u : 0u;  // GravitationalField::is_affecting
}

__global__ void compact_kernel(const float4* __restrict__ p, size_t n,
                               const uint32_t* __restrict__ rank, float4* __restrict__ dst) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && p[i].w != 0.0f) dst[rank[i]] = p[i];
}

int split_affecting(pnb_ctx* c, const float4* particles, size_t n, float4* dst,
                    uint32_t* m_host) {
    PNB_HIP(c, c->split_flags.ensure((n + 1) * sizeof(uint32_t)));
    uint32_t* flags = c->split_flags.as<uint32_t>();
    unsigned g = div_up(n, 256);
    affecting_flags_kernel<<<g, 256, 0, c->stream>>>(particles, n, flags);
    PNB_CHECK_LAUNCH(c);
    int rc = scan_u32(c, flags, flags, n, true);
    if (rc) return rc;
    compact_kernel<<<g, 256, 0, c->stream>>>(particles, n, flags, dst);
    PNB_CHECK_LAUNCH(c);
    uint32_t m = 0;
    PNB_HIP(c, hipMemcpyAsync(&m, flags + n, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    *m_host = m;
    return PNB_OK;
}

}  // namespace pnb

// ---- C ABI -----------------------------------------------------------------------------------

extern "C" {

int pnb_device_count(int* count) {
    if (!count) return PNB_E_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0) {
        (void)hipGetLastError();
        *count = 0;
        return PNB_E_NODEV;
    }
    *count = n;
    return PNB_OK;
}

int pnb_create(int device, pnb_ctx** out) {
    if (!out) return PNB_E_ARG;
    *out = nullptr;
    int n = 0;
    if (pnb_device_count(&n) != PNB_OK) return PNB_E_NODEV;
    if (device < 0 || device >= n) return PNB_E_ARG;
    pnb_ctx* c = new (std::nothrow) pnb_ctx();
    if (!c) return PNB_E_OOM;
    c->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking);
    for (int k = 0; k < pnb_ctx::EV_POOL && e == hipSuccess; ++k) {
        e = hipEventCreate(&c->ev_start[k]);
        if (e == hipSuccess) e = hipEventCreate(&c->ev_stop[k]);
    }
    if (e == hipSuccess) e = hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        pnb_destroy(c);
        return PNB_E_NODEV;
    }
    c->stream = c->own_stream;
    if (pnb::bf_noalloc_init(c) != PNB_OK) {
        (void)hipGetLastError();
        pnb_destroy(c);
        return PNB_E_HIP;
    }
    *out = c;
    return PNB_OK;
}

void pnb_destroy(pnb_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->own_stream) hipStreamSynchronize(c->own_stream);
    if (c->stream && c->stream != c->own_stream) hipStreamSynchronize(c->stream);
    pnb::bf_tune_free(c->bf_tune);
    pnb::bh_free(c->bh);
    pnb::step_free(c->step);
    pnb::batch_free(c->batch);
    pnb::mlp_free(c->mlp);
    for (pnb::DevBuf* b : {&c->h_affected, &c->h_affecting, &c->h_out, &c->split_tmp,
                           &c->split_flags, &c->bf_partial, &c->scan_tmp, &c->pairs_part, &c->d2_raw_a,
                           &c->d2_raw_b, &c->d2_a, &c->d2_b, &c->d2_out4, &c->d2_out2})
        b->release();
    for (int k = 0; k < pnb_ctx::EV_POOL; ++k) {
        if (c->ev_start[k]) hipEventDestroy(c->ev_start[k]);
        if (c->ev_stop[k]) hipEventDestroy(c->ev_stop[k]);
    }
    if (c->own_stream) hipStreamDestroy(c->own_stream);
    delete c;
}

const char* pnb_last_error(const pnb_ctx* c) { return c ? c->last_error.c_str() : "null context"; }

const char* pnb_status_string(int s) {
    switch (s) {
        case PNB_OK: return "ok";
        case PNB_E_ARG: return "invalid argument";
        case PNB_E_NODEV: return "no usable GPU";
        case PNB_E_OOM: return "device out of memory";
        case PNB_E_HIP: return "HIP runtime error";
        case PNB_E_DEPTH: return "octree depth limit exceeded";
        case PNB_E_STATE: return "invalid call order";
        default: return "unknown status";
    }
}

int pnb_set_stream(pnb_ctx* c, void* s) {
    if (!c) return PNB_E_ARG;
    c->stream = static_cast<hipStream_t>(s);
    return PNB_OK;
}

int pnb_reset_stream(pnb_ctx* c) {
    if (!c) return PNB_E_ARG;
    c->stream = c->own_stream;
    return PNB_OK;
}

void* pnb_get_stream(pnb_ctx* c) { return c ? static_cast<void*>(c->stream) : nullptr; }

int pnb_synchronize(pnb_ctx* c) {
    if (!c) return PNB_E_ARG;
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_profile_enable(pnb_ctx* c, int enable) {
    if (!c) return PNB_E_ARG;
    if (int rc = pnb::prof_resolve(c)) return rc;
    c->profile = enable != 0;
    for (int k = 0; k < PNB_PROF_NUM_KERNELS; ++k) {
        c->prof_ms[k] = 0.0;
        c->prof_count[k] = 0;
    }
    return PNB_OK;
}

int pnb_profile_read(pnb_ctx* c, int k, double* total_ms, uint64_t* launches) {
    if (!c || k < 0 || k >= PNB_PROF_NUM_KERNELS) return PNB_E_ARG;
    if (int rc = pnb::prof_resolve(c)) return rc;
    if (total_ms) *total_ms = c->prof_ms[k];
    if (launches) *launches = c->prof_count[k];
    return PNB_OK;
}

}  // extern "C"
