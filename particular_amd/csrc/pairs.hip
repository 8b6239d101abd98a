// Brute force over pairs (Newton's third law) on gfx950: every unordered pair of one particle set
// is evaluated once and both accelerations come from it.
//
// Reference: sequential::BruteForcePairs / RestrictedBruteForce (particular/src/sequential.rs:
// 215-305) with Acceleration's PairInteraction (particular/src/gravity/newtonian/acceleration.rs:
// 95-109, particular/src/gravity/impls/mod.rs:185-205): for i < j, t = d / (|d|^2 + eps^2)^1.5
// with d = x_j - x_i, a_i += t mu_j, a_j -= t mu_i. The reference's pair form is unchecked (a pair
// at zero distance gives a non-finite result); `checked` masks such pairs as the one-sided brute
// force does (particular/src/gravity/impls/mod.rs:160-161), unchecked keeps the reference's form.
//
// Design for CDNA4 (one pair term feeds two accelerations, so a pair costs about 17 VALU against
// 2 x 13 for the one-sided kernel):
//  * tiles of T = 64 R bodies; one wavefront holds an i-tile in registers (R bodies per lane,
//    a plain f32 partial sum per step of T terms, the K step partials added) and streams a
//    j-tile through wave-uniform scalar loads, one body ahead;
//  * per j body the lanes fold the reaction of their R i-bodies, the wavefront sums the 64 lanes
//    with DPP (row butterflies, then row broadcasts 15 and 31) and lane 63 adds the total to the
//    j body's accumulator in LDS;
//  * a workgroup of K wavefronts takes a pair of super-tiles (P, Q), P <= Q, of K tiles each:
//    wavefront w keeps i-tile w of P and at step s takes j-tile (w + s) mod K of Q, so within a
//    step no two wavefronts touch the same LDS accumulators and after K steps every tile pair of
//    (P, Q) is done (for P == Q each unordered tile pair once, the diagonal tiles as triangles);
//  * the i-side sums of P go to partial slot Q and the j-side sums of Q to slot P (for P == Q
//    both, added, to slot P); every body receives exactly one partial per super-tile, and a last
//    pass adds a body's partials in slot order with Kahan compensation. No atomics: the result does
//    not depend on scheduling.
#include "pnb_internal.hpp"

#include <algorithm>
#include <cmath>

namespace pnb {

constexpr int PR_R = 8;               // i-bodies per lane
constexpr int PR_T = 64 * PR_R;       // bodies per tile
constexpr int PR_K = 8;               // wavefronts per workgroup = tiles per super-tile
constexpr int PR_S = PR_K * PR_T;     // bodies per super-tile
// Bodies past n: massless, far away (r^2 stays finite in f32), so that every term they take part
// in is zero without a mask.
constexpr float PR_FAR = 3.0e18f;

// Sum over the wavefront's 64 lanes, complete in lane 63 (partial sums elsewhere).
__device__ __forceinline__ float wave_sum_to_63(float v) {
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0xb1, 0xf, 0xf, false);  // quad_perm [1,0,3,2]
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x4e, 0xf, 0xf, false);  // quad_perm [2,3,0,1]
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x141, 0xf, 0xf, false); // row_half_mirror
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x140, 0xf, 0xf, false); // row_mirror
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x142, 0xa, 0xf, false); // row_bcast:15, rows 1, 3
    v += __builtin_amdgcn_update_dpp(0.0f, v, 0x143, 0xc, 0xf, false); // row_bcast:31, rows 2, 3
    return v;
}

__device__ __forceinline__ void pr_kahan(float& sum, float& comp, float v) {
    const float y = v - comp;
    const float t = sum + y;
    comp = (t - sum) - y;
    sum = t;
}

// One tile pair: the wavefront's i-bodies (registers) against the T bodies of j-tile `jt`
// (global, from body jbase). TRI: the same tile, pairs with j > i only.
template <bool MASK, bool TRI>
__device__ __forceinline__ void pr_tile(const float4* __restrict__ p, uint32_t n, uint32_t jbase,
                                        const float (&xi)[PR_R], const float (&yi)[PR_R],
                                        const float (&zi)[PR_R], const float (&mi)[PR_R],
                                        float eps2, int lane, float4* __restrict__ jacc,
                                        float (&px)[PR_R], float (&py)[PR_R], float (&pz)[PR_R]) {
    // j bodies past n stand in as massless and far away; their loads stay in bounds
    const bool full = jbase + uint32_t(PR_T) <= n;
    float4 q_next = p[min(jbase, n - 1)];  // wave-uniform: scalar loads, one body ahead
    for (uint32_t j = 0; j < uint32_t(PR_T); ++j) {
        const uint32_t jg = jbase + j;
        float4 q = q_next;
        if (!full && jg >= n) q = make_float4(PR_FAR, PR_FAR, PR_FAR, 0.0f);
        q_next = p[min(jg + 1, n - 1)];
        float rx = 0.0f, ry = 0.0f, rz = 0.0f;
#pragma unroll
        for (int r = 0; r < PR_R; ++r) {
            const float dx = q.x - xi[r];
            const float dy = q.y - yi[r];
            const float dz = q.z - zi[r];
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, fmaf(dz, dz, eps2)));
            float inv = __builtin_amdgcn_rsqf(r2);
            if (MASK) inv = (r2 > 0.0f) ? inv : 0.0f;
            if (TRI) inv = (j > uint32_t(lane + 64 * r)) ? inv : 0.0f;
            const float t = (inv * inv) * inv;
            const float si = q.w * t, sj = mi[r] * t;
            px[r] = fmaf(dx, si, px[r]);
            py[r] = fmaf(dy, si, py[r]);
            pz[r] = fmaf(dz, si, pz[r]);
            rx = fmaf(dx, sj, rx);
            ry = fmaf(dy, sj, ry);
            rz = fmaf(dz, sj, rz);
        }
        rx = wave_sum_to_63(rx);
        ry = wave_sum_to_63(ry);
        rz = wave_sum_to_63(rz);
        if (lane == 63) {  // a_j -= sum_i d mu_i t
            float4 a = jacc[j];
            a.x -= rx;
            a.y -= ry;
            a.z -= rz;
            jacc[j] = a;
        }
    }
}

// blockIdx.x -> (P, Q), P <= Q: the lower-triangle index b = Q (Q + 1) / 2 + P.
__device__ __forceinline__ void pr_decode(uint32_t b, uint32_t& P, uint32_t& Q) {
    uint32_t q = uint32_t((sqrtf(8.0f * float(b) + 1.0f) - 1.0f) * 0.5f);
    while (q * (q + 1) / 2 > b) --q;
    while ((q + 1) * (q + 2) / 2 <= b) ++q;
    Q = q;
    P = b - q * (q + 1) / 2;
}

template <bool MASK>
__global__ __launch_bounds__(64 * PR_K) __attribute__((amdgpu_waves_per_eu(4, 8))) void pairs_kernel(const float4* __restrict__ p, uint32_t n,
                                                          uint32_t n_pad, float eps2,
                                                          float4* __restrict__ part) {
    __shared__ float4 jacc[PR_K * PR_T];
    uint32_t P, Q;
    pr_decode(blockIdx.x, P, Q);
    // the wavefront index is wave-uniform: keep it (and the j-tile addresses) in SGPRs
    const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    for (uint32_t t = threadIdx.x; t < uint32_t(PR_K * PR_T); t += blockDim.x)
        jacc[t] = make_float4(0.f, 0.f, 0.f, 0.f);

    const uint32_t ibase = P * PR_S + uint32_t(w) * PR_T;
    float xi[PR_R], yi[PR_R], zi[PR_R], mi[PR_R];
    float ax[PR_R], ay[PR_R], az[PR_R];
#pragma unroll
    for (int r = 0; r < PR_R; ++r) {
        const uint32_t i = ibase + uint32_t(lane + 64 * r);
        float4 b = p[i < n ? i : n - 1];
        if (i >= n) b = make_float4(PR_FAR, PR_FAR, PR_FAR, 0.0f);
        xi[r] = b.x; yi[r] = b.y; zi[r] = b.z; mi[r] = b.w;
        ax[r] = ay[r] = az[r] = 0.0f;
    }
    __syncthreads();

    const bool diag = P == Q;
    // P == Q: the unordered tile pair {a, b} is taken by wavefront a at step (b - a) mod K when
    // that is below K / 2, by b otherwise; at step K / 2 by the lower of the two; step 0 is the
    // diagonal tile's triangle.
    const int steps = diag ? PR_K / 2 + 1 : PR_K;
    for (int s = 0; s < steps; ++s) {
        const int jt = (w + s) % PR_K;
        const bool take = !diag || s < PR_K / 2 || w < PR_K / 2;  // diag: s <= K / 2 here
        if (take) {
            float px[PR_R], py[PR_R], pz[PR_R];
#pragma unroll
            for (int r = 0; r < PR_R; ++r) px[r] = py[r] = pz[r] = 0.0f;
            const uint32_t jbase = Q * PR_S + uint32_t(jt) * PR_T;
            if (diag && s == 0)
                pr_tile<MASK, true>(p, n, jbase, xi, yi, zi, mi, eps2, lane, jacc + jt * PR_T,
                                    px, py, pz);
            else
                pr_tile<MASK, false>(p, n, jbase, xi, yi, zi, mi, eps2, lane, jacc + jt * PR_T,
                                     px, py, pz);
#pragma unroll
            for (int r = 0; r < PR_R; ++r) {  // at most K step partials: plain sums
                ax[r] += px[r];
                ay[r] += py[r];
                az[r] += pz[r];
            }
        }
        __syncthreads();  // the next step hands each j-tile's accumulators to another wavefront
    }

    // i-side of P into slot Q (P == Q: plus the j-side of the same bodies, all in slot P)
    float4* dst_i = part + size_t(Q) * n_pad;
#pragma unroll
    for (int r = 0; r < PR_R; ++r) {
        const uint32_t il = uint32_t(w) * PR_T + uint32_t(lane + 64 * r);
        if (diag) {
            const float4 a = jacc[il];
            ax[r] += a.x;
            ay[r] += a.y;
            az[r] += a.z;
        }
        dst_i[P * PR_S + il] = make_float4(ax[r], ay[r], az[r], 0.0f);
    }
    if (!diag) {  // j-side of Q into slot P
        float4* dst_j = part + size_t(P) * n_pad + size_t(Q) * PR_S;
        for (uint32_t t = threadIdx.x; t < uint32_t(PR_K * PR_T); t += blockDim.x) dst_j[t] = jacc[t];
    }
}

// Every body's partials in slot order, compensated.
__global__ void pairs_reduce_kernel(const float4* __restrict__ part, uint32_t n, uint32_t n_pad,
                                    uint32_t slots, float4* __restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float ax = 0, ay = 0, az = 0, cx = 0, cy = 0, cz = 0;
    for (uint32_t s = 0; s < slots; ++s) {
        const float4 a = part[size_t(s) * n_pad + i];
        pr_kahan(ax, cx, a.x);
        pr_kahan(ay, cy, a.y);
        pr_kahan(az, cz, a.z);
    }
    out[i] = make_float4(ax, ay, az, 0.0f);
}

__global__ void pairs_zero_kernel(float4* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// Largest set this path takes: the partials hold ceil(n / S)^2 * S records of 16 bytes (68 GB
// at this size).
constexpr size_t PR_MAX_N = size_t(1) << 22;

int pairs_launch(pnb_ctx* c, const float4* p, size_t n, float softening, int checked,
                 float4* out) {
    if (n == 0) return PNB_OK;
    if (n > PR_MAX_N)
        return fail(c, PNB_E_ARG, "pair brute force takes at most %zu particles", PR_MAX_N);
    if (n == 1) {  // no pair
        pairs_zero_kernel<<<1, 64, 0, c->stream>>>(out, 1);
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    }
    const uint32_t ns = uint32_t(div_up(n, PR_S));
    const uint32_t n_pad = ns * uint32_t(PR_S);
    const size_t need = size_t(ns) * n_pad * sizeof(float4);
    if (c->capturing && c->pairs_part.cap < need)
        return fail(c, PNB_E_STATE, "pair scratch not prepared before graph capture");
    PNB_HIP(c, c->pairs_part.ensure(need));
    float4* part = c->pairs_part.as<float4>();
    const unsigned blocks = unsigned(size_t(ns) * (ns + 1) / 2);
    const bool mask = checked && softening == 0.0f;
    const float eps2 = softening * softening;
    return timed(c, PNB_PROF_BF, [&]() -> int {
        if (mask)
            pairs_kernel<true><<<blocks, 64 * PR_K, 0, c->stream>>>(p, uint32_t(n), n_pad, eps2, part);
        else
            pairs_kernel<false><<<blocks, 64 * PR_K, 0, c->stream>>>(p, uint32_t(n), n_pad, eps2, part);
        PNB_CHECK_LAUNCH(c);
        pairs_reduce_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(part, uint32_t(n), n_pad, ns, out);
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    });
}

}  // namespace pnb

using namespace pnb;

extern "C" {

int pnb_bf_pairs_accel_f32_dev(pnb_ctx* c, const float* particles, size_t n, float softening,
                               int checked, float* out) {
    if (!c) return PNB_E_ARG;
    if (n && (!particles || !out)) return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    return pairs_launch(c, reinterpret_cast<const float4*>(particles), n, softening, checked,
                        reinterpret_cast<float4*>(out));
}

int pnb_bf_pairs_accel_f32(pnb_ctx* c, const float* particles, size_t n, float softening,
                           int checked, float* out) {
    if (!c) return PNB_E_ARG;
    if (n && (!particles || !out)) return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (n > PR_MAX_N)
        return fail(c, PNB_E_ARG, "pair brute force takes at most %zu particles", PR_MAX_N);
    if (int rc = activate(c)) return rc;
    PNB_HIP(c, c->h_affected.ensure(n * sizeof(float4)));
    PNB_HIP(c, c->h_out.ensure(n * sizeof(float4)));
    PNB_HIP(c, hipMemcpyAsync(c->h_affected.ptr, particles, n * sizeof(float4),
                              hipMemcpyHostToDevice, c->stream));
    if (int rc = pairs_launch(c, c->h_affected.as<float4>(), n, softening, checked,
                              c->h_out.as<float4>()))
        return rc;
    PNB_HIP(c, hipMemcpyAsync(out, c->h_out.ptr, n * sizeof(float4), hipMemcpyDeviceToHost,
                              c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

}  // extern "C"
