// Brute-force gravitational acceleration on gfx950, and the FP64 direct sum used as the accuracy reference.
//
// Replaces the reference's wgpu brute force (particular/src/gpu/bruteforce_shared.wgsl:8-31 and
// particular/src/gpu/mod.rs:179-208) with its own design for CDNA4:
//  * each lane owns R affected particles (register blocking): one affecting body fetched per
//    wave feeds R independent FMA chains;
//  * affecting bodies are shared by the whole workgroup, either staged tile by tile in LDS and
//    read back as broadcasts (variant 0) or read straight from memory with wave-uniform addresses,
//    which the compiler turns into scalar loads into SGPRs (variant 1);
//  * the pair term uses the hardware reciprocal square root (v_rsq_f32),
//    a_i += d * (mu * rsq(r2)^3), r2 = |d|^2 + eps^2;
//  * sums are plain FP32 within a tile of BF_TILE affecting bodies and Kahan-compensated across
//    tiles, so the error does not grow with the affecting count;
//  * small affected counts split the affecting range over several workgroups (j_split) whose
//    partial sums are reduced in a fixed order, so results do not depend on scheduling.
// Pair semantics follow particular/src/gravity/impls/mod.rs:151-166: with `checked` a pair at
// zero unsoftened distance contributes nothing; with softening > 0 that pair contributes
// d * finite = 0 anyway, so only checked && softening == 0 needs the explicit mask.
#include "pnb_internal.hpp"

#include <algorithm>
#include <mutex>
#include <vector>

namespace pnb {

constexpr int BF_BLOCK = 256;
constexpr int BF_TILE = 256;

struct BfTuneEntry {
    size_t n_min;
    int variant, per_lane, j_split;
};
struct BfTune {
    std::vector<BfTuneEntry> entries;
};

void bf_tune_free(BfTune* t) { delete t; }

template <int R, bool MASK>
__device__ __forceinline__ void bf_body(const float4 q, const float (&xi)[R], const float (&yi)[R],
                                        const float (&zi)[R], float eps2, float (&px)[R],
                                        float (&py)[R], float (&pz)[R]) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const float dx = q.x - xi[r];
        const float dy = q.y - yi[r];
        const float dz = q.z - zi[r];
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
        float inv = __builtin_amdgcn_rsqf(r2);
        if (MASK) inv = (r2 > 0.0f) ? inv : 0.0f;
        const float s = (q.w * inv) * (inv * inv);
        px[r] = fmaf(dx, s, px[r]);
        py[r] = fmaf(dy, s, py[r]);
        pz[r] = fmaf(dz, s, pz[r]);
    }
}

__device__ __forceinline__ void kahan_add(float& sum, float& comp, float v) {
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// grid.x: affected tiles of BF_BLOCK * R particles; grid.y: affecting ranges of j_chunk bodies.
// With gridDim.y == 1 the result goes to `out`, else to partial[blockIdx.y * n + i].
// VAR 0: LDS tile; 1: wave-uniform global reads (scalar loads).
template <int R, bool MASK, int VAR>
__global__ __launch_bounds__(BF_BLOCK) void bf_kernel(const float4* __restrict__ aff, uint32_t n,
                                                      const float4* __restrict__ afg, uint32_t m,
                                                      uint32_t j_chunk, float eps2,
                                                      float4* __restrict__ out) {
    constexpr bool LDS = VAR == 0;
    __shared__ float4 tile[LDS ? BF_TILE : 1];
    const uint32_t jb = blockIdx.y * j_chunk;
    const uint32_t je = min(m, jb + j_chunk);
    const uint32_t i0 = blockIdx.x * (BF_BLOCK * R) + threadIdx.x;

    float xi[R], yi[R], zi[R], ax[R], ay[R], az[R], cx[R], cy[R], cz[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = i0 + r * BF_BLOCK;
        const float4 p = aff[i < n ? i : n - 1];
        xi[r] = p.x; yi[r] = p.y; zi[r] = p.z;
        ax[r] = ay[r] = az[r] = 0.0f;
        cx[r] = cy[r] = cz[r] = 0.0f;
    }

    for (uint32_t j0 = jb; j0 < je; j0 += BF_TILE) {
        const uint32_t cnt = min(uint32_t(BF_TILE), je - j0);
        if (LDS) {
            __syncthreads();
            if (threadIdx.x < cnt) tile[threadIdx.x] = afg[j0 + threadIdx.x];
            __syncthreads();
        }
        float px[R], py[R], pz[R];
#pragma unroll
        for (int r = 0; r < R; ++r) px[r] = py[r] = pz[r] = 0.0f;
        if (cnt == BF_TILE) {
#pragma unroll 8
            for (int j = 0; j < BF_TILE; ++j)
                bf_body<R, MASK>(LDS ? tile[j] : afg[j0 + j], xi, yi, zi, eps2, px, py, pz);
        } else {
            for (uint32_t j = 0; j < cnt; ++j)
                bf_body<R, MASK>(LDS ? tile[j] : afg[j0 + j], xi, yi, zi, eps2, px, py, pz);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            kahan_add(ax[r], cx[r], px[r]);
            kahan_add(ay[r], cy[r], py[r]);
            kahan_add(az[r], cz[r], pz[r]);
        }
    }

    float4* dst = out + size_t(blockIdx.y) * n;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const uint32_t i = i0 + r * BF_BLOCK;
        if (i < n) dst[i] = make_float4(ax[r], ay[r], az[r], 0.0f);
    }
}

// Fixed-order compensated reduction of j_split partial sums.
__global__ void bf_reduce_kernel(const float4* __restrict__ partial, uint32_t n, uint32_t splits,
                                 float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float ax = 0, ay = 0, az = 0, cx = 0, cy = 0, cz = 0;
    for (uint32_t s = 0; s < splits; ++s) {
        const float4 p = partial[size_t(s) * n + i];
        kahan_add(ax, cx, p.x);
        kahan_add(ay, cy, p.y);
        kahan_add(az, cz, p.z);
    }
    out[i] = make_float4(ax, ay, az, 0.0f);
}

__global__ void fill_zero_kernel(float4* __restrict__ out, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

using BfKernel = void (*)(const float4*, uint32_t, const float4*, uint32_t, uint32_t, float, float4*);

template <bool MASK, int VAR>
static BfKernel pick_r(int r) {
    switch (r) {
        case 1: return bf_kernel<1, MASK, VAR>;
        case 2: return bf_kernel<2, MASK, VAR>;
        case 4: return bf_kernel<4, MASK, VAR>;
        default: return bf_kernel<8, MASK, VAR>;
    }
}

template <int VAR>
static BfKernel pick_mask(int r, bool mask) {
    return mask ? pick_r<true, VAR>(r) : pick_r<false, VAR>(r);
}

static BfKernel pick_kernel(int variant, int r, bool mask) {
    return variant == 0 ? pick_mask<0>(r, mask) : pick_mask<1>(r, mask);
}

struct BfShape {
    int variant, per_lane, j_split;
};

static BfShape choose_shape(pnb_ctx* c, size_t n, size_t m) {
    if (c->bf_tune) {
        const BfTuneEntry* best = nullptr;
        for (const auto& e : c->bf_tune->entries)
            if (e.n_min <= n && (!best || e.n_min > best->n_min)) best = &e;
        if (best) return {best->variant, best->per_lane, best->j_split};
    }
    // Default (from autotuning on MI355X): 4 affected bodies per lane with scalar-loaded
    // affecting bodies once there are at least 64 affected tiles, else 1 per lane from LDS; the
    // affecting range is split into up to 64 workgroups per CU (small workgroups balance the
    // load across CUs and XCDs), keeping at least 4 tiles of affecting bodies per split.
    const size_t cus = c->cu_count > 0 ? size_t(c->cu_count) : 256;
    const int r = n >= size_t(64) * BF_BLOCK * 4 ? 4 : 1;
    const int variant = r == 4 ? 1 : 0;
    const size_t blocks = div_up(n, size_t(BF_BLOCK) * r);
    int s = 1;
    while (s < 64 && blocks * s * 2 <= 64 * cus && m / (size_t(s) * 2) >= 4 * size_t(BF_TILE)) s <<= 1;
    return {variant, r, s};
}

struct BfPlan {
    BfShape shape;
    uint32_t splits, chunk;
};

static BfPlan plan(pnb_ctx* c, size_t n, size_t m) {
    const BfShape sh = choose_shape(c, n, m);
    uint32_t splits = uint32_t(sh.j_split);
    uint32_t chunk = uint32_t(div_up(m, splits));
    chunk = (chunk + BF_TILE - 1) / BF_TILE * BF_TILE;
    splits = uint32_t(div_up(m, chunk));
    return {sh, splits, chunk};
}

// Allocates what bf_launch(n, m) will need, so that a later launch inside a graph capture
// allocates nothing.
int bf_prepare(pnb_ctx* c, size_t n, size_t m) {
    if (n == 0 || m == 0) return PNB_OK;
    const BfPlan p = plan(c, n, m);
    if (p.splits > 1) PNB_HIP(c, c->bf_partial.ensure(size_t(p.splits) * n * sizeof(float4)));
    return PNB_OK;
}

// Scratch of the allocation-free entry point (pnb_bf_accel_f32_dev_noalloc): the partial sums of
// the affecting split, splits * n records. 256 MiB: the default shapes need at most
// 64 * CUs * BF_BLOCK * 4 records (16,777,216 on 256 CUs). Every GPU the library runs on holds its
// own copy until the process ends; every call on that GPU shares it.
constexpr size_t BF_SCRATCH_RECORDS = size_t(1) << 24;
static __device__ __align__(256) float4 scratch[BF_SCRATCH_RECORDS];
constexpr int BF_SCRATCH_DEVICES = 64;
static float4* g_scratch_dev[BF_SCRATCH_DEVICES];  // device address of `scratch` per device
static std::mutex g_scratch_mu;

// Once per device, with that device current (called by pnb_create): the device address of
// `scratch`. In host code the array's name is a host-side stand-in, not a device address.
int bf_noalloc_init(pnb_ctx* c) {
    if (c->device < 0 || c->device >= BF_SCRATCH_DEVICES) return PNB_OK;  // entry point refuses
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[c->device]) {
        void* ptr = nullptr;
        PNB_HIP(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_dev[c->device] = static_cast<float4*>(ptr);
    }
    c->noalloc_scratch = g_scratch_dev[c->device];
    return PNB_OK;
}

// The launch itself, with the split's partial sums in `partial` (at least splits * n records;
// unused when the plan has one split). No allocation.
static int bf_launch_with(pnb_ctx* c, const BfPlan& p, const float4* affected, size_t n,
                          const float4* affecting, size_t m, float softening, int checked,
                          float4* out, float4* partial) {
    const BfShape sh = p.shape;
    const bool mask = checked && softening == 0.0f;
    BfKernel k = pick_kernel(sh.variant, sh.per_lane, mask);
    const float eps2 = softening * softening;
    const uint32_t splits = p.splits, chunk = p.chunk;
    dim3 grid(div_up(n, size_t(BF_BLOCK) * sh.per_lane), splits);
    float4* dst = splits > 1 ? partial : out;
    return timed(c, PNB_PROF_BF, [&]() -> int {
        k<<<grid, BF_BLOCK, 0, c->stream>>>(affected, uint32_t(n), affecting, uint32_t(m), chunk,
                                            eps2, dst);
        PNB_CHECK_LAUNCH(c);
        if (splits > 1) {
            bf_reduce_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(dst, uint32_t(n), splits, out);
            PNB_CHECK_LAUNCH(c);
        }
        return PNB_OK;
    });
}

int bf_launch(pnb_ctx* c, const float4* affected, size_t n, const float4* affecting, size_t m,
              float softening, int checked, float4* out) {
    if (n == 0) return PNB_OK;
    if (n > 0x7fffffffu || m > 0x7fffffffu)
        return fail(c, PNB_E_ARG, "brute force supports at most 2^31-1 particles per side");
    if (m == 0) {
        fill_zero_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(out, n);
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    }
    const BfPlan p = plan(c, n, m);
    if (p.splits > 1) {
        const size_t need = size_t(p.splits) * n * sizeof(float4);
        if (c->capturing && c->bf_partial.cap < need)
            return fail(c, PNB_E_STATE, "brute-force scratch not prepared before graph capture");
        PNB_HIP(c, c->bf_partial.ensure(need));
    }
    return bf_launch_with(c, p, affected, n, affecting, m, softening, checked, out,
                          c->bf_partial.as<float4>());
}

// ---- FP64 direct sum: one workgroup per affected particle, fixed-order tree reduction --------

constexpr int DF_BLOCK = 256;

__global__ __launch_bounds__(DF_BLOCK) void direct_f64_kernel(const float4* __restrict__ aff,
                                                               const float4* __restrict__ afg,
                                                               uint32_t m, double eps2,
                                                               double* __restrict__ out3) {
    const uint32_t i = blockIdx.x;
    const float4 p = aff[i];
    const double x = p.x, y = p.y, z = p.z;
    double sx = 0, sy = 0, sz = 0;
    for (uint32_t j = threadIdx.x; j < m; j += DF_BLOCK) {
        const float4 q = afg[j];
        const double dx = double(q.x) - x, dy = double(q.y) - y, dz = double(q.z) - z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 == 0.0) continue;
        const double rs = r2 + eps2;
        const double s = double(q.w) / (rs * sqrt(rs));
        sx += dx * s; sy += dy * s; sz += dz * s;
    }
    __shared__ double red[3][DF_BLOCK];
    red[0][threadIdx.x] = sx;
    red[1][threadIdx.x] = sy;
    red[2][threadIdx.x] = sz;
    __syncthreads();
    for (int w = DF_BLOCK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out3[3 * size_t(i)] = red[0][0];
        out3[3 * size_t(i) + 1] = red[1][0];
        out3[3 * size_t(i) + 2] = red[2][0];
    }
}

// ---- FP64 brute force (glam::DVec3 / f64: particular/src/gravity/impls/glam.rs:234) ------------
// The reference's scalar kernel in FP64 (impls/mod.rs:151-166): IEEE sqrt and division, pairs
// folded in affecting order per tile; tiles of BF_TILE bodies staged in LDS, one affected body per
// lane. checked: zero unsoftened distance contributes nothing.

__global__ __launch_bounds__(BF_BLOCK) void bf64_kernel(const double4* __restrict__ aff, uint32_t n,
                                                        const double4* __restrict__ afg, uint32_t m,
                                                        double eps2, int checked,
                                                        double4* __restrict__ out) {
    __shared__ double4 tile[BF_TILE];
    const uint32_t i = blockIdx.x * BF_BLOCK + threadIdx.x;
    const double4 p = aff[i < n ? i : n - 1];
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (uint32_t j0 = 0; j0 < m; j0 += BF_TILE) {
        const uint32_t cnt = min(uint32_t(BF_TILE), m - j0);
        __syncthreads();
        if (threadIdx.x < cnt) tile[threadIdx.x] = afg[j0 + threadIdx.x];
        __syncthreads();
        for (uint32_t j = 0; j < cnt; ++j) {
            const double4 q = tile[j];
            const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
            const double r2 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
            if (checked && r2 == 0.0) continue;
            const double rs = r2 + eps2;
            const double s = q.w / (rs * sqrt(rs));
            ax += dx * s;
            ay += dy * s;
            az += dz * s;
        }
    }
    if (i < n) out[i] = make_double4(ax, ay, az, 0.0);
}

static int copy_in(pnb_ctx* c, DevBuf& b, const float* src, size_t count) {
    PNB_HIP(c, b.ensure(count * sizeof(float4)));
    if (count)
        PNB_HIP(c, hipMemcpyAsync(b.ptr, src, count * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    return PNB_OK;
}

static int copy_out(pnb_ctx* c, float* dst, const DevBuf& b, size_t count) {
    if (count)
        PNB_HIP(c, hipMemcpyAsync(dst, b.ptr, count * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

}  // namespace pnb

using namespace pnb;

extern "C" {

int pnb_bf_accel_f32_dev(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                         size_t m, float softening, int checked, float* out) {
    if (!c) return PNB_E_ARG;
    if ((n && (!affected || !out)) || (m && !affecting))
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (int rc = activate(c)) return rc;
    return bf_launch(c, reinterpret_cast<const float4*>(affected), n,
                     reinterpret_cast<const float4*>(affecting), m, softening, checked,
                     reinterpret_cast<float4*>(out));
}

int pnb_bf_accel_f32_dev_noalloc(pnb_ctx* c, const float* affected, size_t n,
                                 const float* affecting, size_t m, float softening, int checked,
                                 float* out) {
    if (!c) return PNB_E_ARG;
    if ((n && (!affected || !out)) || (m && !affecting))
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (n > 0x7fffffffu || m > 0x7fffffffu)
        return fail(c, PNB_E_ARG, "brute force supports at most 2^31-1 particles per side");
    if (!c->noalloc_scratch)
        return fail(c, PNB_E_STATE, "no allocation-free scratch on device %d", c->device);
    if (int rc = activate(c)) return rc;
    float4* o = reinterpret_cast<float4*>(out);
    if (m == 0) {
        fill_zero_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(o, n);
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    }
    const BfPlan p = plan(c, n, m);
    if (p.splits > 1 && size_t(p.splits) * n > BF_SCRATCH_RECORDS)
        return fail(c, PNB_E_ARG,
                    "the allocation-free brute force holds %zu partial sums; %u affecting ranges "
                    "x %zu affected need %zu",
                    BF_SCRATCH_RECORDS, p.splits, n, size_t(p.splits) * n);
    return bf_launch_with(c, p, reinterpret_cast<const float4*>(affected), n,
                          reinterpret_cast<const float4*>(affecting), m, softening, checked, o,
                          c->noalloc_scratch);
}

int pnb_bf_accel_f32(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                     size_t m, float softening, int checked, float* out) {
    if (!c) return PNB_E_ARG;
    if ((n && (!affected || !out)) || (m && !affecting))
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    if (int rc = copy_in(c, c->h_affected, affected, n)) return rc;
    if (int rc = copy_in(c, c->h_affecting, affecting, m)) return rc;
    PNB_HIP(c, c->h_out.ensure(n * sizeof(float4)));
    if (int rc = bf_launch(c, c->h_affected.as<float4>(), n, c->h_affecting.as<float4>(), m,
                           softening, checked, c->h_out.as<float4>()))
        return rc;
    return copy_out(c, out, c->h_out, n);
}

int pnb_reordered_bf_accel_f32_dev(pnb_ctx* c, const float* particles, size_t n, float softening,
                                   int checked, float* out) {
    if (!c) return PNB_E_ARG;
    if (n && (!particles || !out)) return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    PNB_HIP(c, c->split_tmp.ensure(n * sizeof(float4)));
    uint32_t m = 0;
    const float4* p = reinterpret_cast<const float4*>(particles);
    if (int rc = split_affecting(c, p, n, c->split_tmp.as<float4>(), &m)) return rc;
    return bf_launch(c, p, n, c->split_tmp.as<float4>(), m, softening, checked,
                     reinterpret_cast<float4*>(out));
}

int pnb_reordered_bf_accel_f32(pnb_ctx* c, const float* particles, size_t n, float softening,
                               int checked, float* out) {
    if (!c) return PNB_E_ARG;
    if (n && (!particles || !out)) return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    if (int rc = copy_in(c, c->h_affected, particles, n)) return rc;
    PNB_HIP(c, c->h_out.ensure(n * sizeof(float4)));
    if (int rc = pnb_reordered_bf_accel_f32_dev(c, c->h_affected.as<float>(), n, softening,
                                                checked, c->h_out.as<float>()))
        return rc;
    return copy_out(c, out, c->h_out, n);
}

int pnb_bf_set_tuning(pnb_ctx* c, size_t n_min, int variant, int per_lane, int j_split) {
    if (!c) return PNB_E_ARG;
    if (variant < 0 || variant > 1 ||
        (per_lane != 1 && per_lane != 2 && per_lane != 4 && per_lane != 8) || j_split < 1 ||
        j_split > 64)
        return fail(c, PNB_E_ARG, "invalid brute-force shape");
    if (!c->bf_tune) c->bf_tune = new BfTune();
    auto& v = c->bf_tune->entries;
    v.erase(std::remove_if(v.begin(), v.end(), [&](const BfTuneEntry& e) { return e.n_min == n_min; }),
            v.end());
    v.push_back({n_min, variant, per_lane, j_split});
    return PNB_OK;
}

int pnb_bf_clear_tuning(pnb_ctx* c) {
    if (!c) return PNB_E_ARG;
    if (c->bf_tune) c->bf_tune->entries.clear();
    return PNB_OK;
}

int pnb_direct_f64_dev(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                       size_t m, double softening, double* out3) {
    if (!c) return PNB_E_ARG;
    if ((n && (!affected || !out3)) || (m && !affecting))
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (n == 0) return PNB_OK;
    if (n > 0x7fffffffu || m > 0x7fffffffu) return fail(c, PNB_E_ARG, "too many particles");
    if (int rc = activate(c)) return rc;
    direct_f64_kernel<<<unsigned(n), DF_BLOCK, 0, c->stream>>>(
        reinterpret_cast<const float4*>(affected), reinterpret_cast<const float4*>(affecting),
        uint32_t(m), softening * softening, out3);
    PNB_CHECK_LAUNCH(c);
    return PNB_OK;
}

int pnb_bf_accel_f64_dev(pnb_ctx* c, const double* affected, size_t n, const double* affecting,
                         size_t m, double softening, int checked, double* out) {
    if (!c) return PNB_E_ARG;
    if ((n && (!affected || !out)) || (m && !affecting))
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (n > 0x7fffffffu || m > 0x7fffffffu) return fail(c, PNB_E_ARG, "too many particles");
    if (int rc = activate(c)) return rc;
    bf64_kernel<<<div_up(n, BF_BLOCK), BF_BLOCK, 0, c->stream>>>(
        reinterpret_cast<const double4*>(affected), uint32_t(n),
        reinterpret_cast<const double4*>(affecting), uint32_t(m), softening * softening, checked,
        reinterpret_cast<double4*>(out));
    PNB_CHECK_LAUNCH(c);
    return PNB_OK;
}

int pnb_bf_accel_f64(pnb_ctx* c, const double* affected, size_t n, const double* affecting,
                     size_t m, double softening, int checked, double* out) {
    if (!c) return PNB_E_ARG;
    if ((n && (!affected || !out)) || (m && !affecting))
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    PNB_HIP(c, c->h_affected.ensure(n * 32));
    PNB_HIP(c, c->h_affecting.ensure(std::max<size_t>(m, 1) * 32));
    PNB_HIP(c, c->h_out.ensure(n * 32));
    PNB_HIP(c, hipMemcpyAsync(c->h_affected.ptr, affected, n * 32, hipMemcpyHostToDevice, c->stream));
    if (m) PNB_HIP(c, hipMemcpyAsync(c->h_affecting.ptr, affecting, m * 32, hipMemcpyHostToDevice, c->stream));
    if (int rc = pnb_bf_accel_f64_dev(c, c->h_affected.as<double>(), n, c->h_affecting.as<double>(), m,
                                      softening, checked, c->h_out.as<double>()))
        return rc;
    PNB_HIP(c, hipMemcpyAsync(out, c->h_out.ptr, n * 32, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

}  // extern "C"
