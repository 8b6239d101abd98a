// Two-dimensional particles: glam::Vec2 positions with f32 mu (the reference's 2D GPU layout,
// particular/src/gravity/impls/mod.rs:381-418, and its CPU record GravitationalField<Vec2, f32>,
// particular/src/gravity/mod.rs:12-18). Records are {x, y, mu} (12 bytes); accelerations {ax, ay}.
//
// They run on the three-dimensional path with z = 0, which reproduces the 2D reference exactly:
//  * pair term: Vec2::length_squared is x*x + y*y; the 3D form (x*x + y*y) + z*z adds an exact
//    +0, and every other operation of impls/mod.rs:151-166 is the same per component;
//  * Barnes-Hut: the 3D root box (square_with over x, y and the zero extent of z) has the 2D
//    box's x/y sides; at every level all bodies take the same z half, so the octree groups the
//    bodies exactly as the quadtree does, in the same order (the z bit is one constant per level),
//    with the same widths (x extent), centres of mass and opening decisions.
#include "pnb_internal.hpp"

#include <algorithm>

namespace pnb {

__global__ void d2_widen_kernel(const float* __restrict__ in, size_t n, float4* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float4(in[3 * i], in[3 * i + 1], 0.f, in[3 * i + 2]);
}

__global__ void d2_narrow_kernel(const float4* __restrict__ in, size_t n, float* __restrict__ out) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        const float4 a = in[i];
        out[2 * i] = a.x;
        out[2 * i + 1] = a.y;
    }
}

static int widen(pnb_ctx* c, const float* in, size_t n, DevBuf& dst) {
    PNB_HIP(c, dst.ensure(std::max<size_t>(n, 1) * sizeof(float4)));
    if (n) {
        d2_widen_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(in, n, dst.as<float4>());
        PNB_CHECK_LAUNCH(c);
    }
    return PNB_OK;
}

static int narrow(pnb_ctx* c, size_t n, float* out) {
    if (n) {
        d2_narrow_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(c->d2_out4.as<float4>(), n, out);
        PNB_CHECK_LAUNCH(c);
    }
    return PNB_OK;
}

// Widened device copies of (affected, affecting); one copy when they are the same array.
static int widen_pair(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                      size_t m, const float4** a4, const float4** b4) {
    if (int rc = widen(c, affected, n, c->d2_a)) return rc;
    *a4 = c->d2_a.as<float4>();
    if (affected == affecting && n == m) {
        *b4 = *a4;
        return PNB_OK;
    }
    if (int rc = widen(c, affecting, m, c->d2_b)) return rc;
    *b4 = c->d2_b.as<float4>();
    return PNB_OK;
}

static int stage(pnb_ctx* c, DevBuf& b, const float* src, size_t count) {
    PNB_HIP(c, b.ensure(std::max<size_t>(count, 1) * 12));
    if (count) PNB_HIP(c, hipMemcpyAsync(b.ptr, src, count * 12, hipMemcpyHostToDevice, c->stream));
    return PNB_OK;
}

}  // namespace pnb

using namespace pnb;

#define PNB_D2_ARGS_CHECK()                                                                 \
    if (!c) return PNB_E_ARG;                                                               \
    if ((n && (!affected || !out)) || (m && !affecting))                                    \
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");

extern "C" {

int pnb_bf_accel_f32_2d_dev(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                            size_t m, float softening, int checked, float* out) {
    PNB_D2_ARGS_CHECK();
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    const float4 *a4, *b4;
    if (int rc = widen_pair(c, affected, n, affecting, m, &a4, &b4)) return rc;
    PNB_HIP(c, c->d2_out4.ensure(n * sizeof(float4)));
    if (int rc = pnb_bf_accel_f32_dev(c, reinterpret_cast<const float*>(a4), n,
                                      reinterpret_cast<const float*>(b4), m, softening, checked,
                                      c->d2_out4.as<float>()))
        return rc;
    return narrow(c, n, out);
}

int pnb_bf_accel_f32_2d(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                        size_t m, float softening, int checked, float* out) {
    PNB_D2_ARGS_CHECK();
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    const bool same = affected == affecting && n == m;
    if (int rc = stage(c, c->d2_raw_a, affected, n)) return rc;
    if (!same)
        if (int rc = stage(c, c->d2_raw_b, affecting, m)) return rc;
    const float* da = c->d2_raw_a.as<float>();
    const float* db = same ? da : c->d2_raw_b.as<float>();
    PNB_HIP(c, c->d2_out2.ensure(n * 8));
    if (int rc = pnb_bf_accel_f32_2d_dev(c, da, n, db, m, softening, checked, c->d2_out2.as<float>()))
        return rc;
    PNB_HIP(c, hipMemcpyAsync(out, c->d2_out2.ptr, n * 8, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_bh_accel_f32_2d_dev(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                            size_t m, float theta, float softening, int checked, unsigned flags,
                            float* out) {
    PNB_D2_ARGS_CHECK();
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    const float4 *a4, *b4;
    if (int rc = widen_pair(c, affected, n, affecting, m, &a4, &b4)) return rc;
    PNB_HIP(c, c->d2_out4.ensure(n * sizeof(float4)));
    if (int rc = pnb_bh_accel_f32_dev(c, reinterpret_cast<const float*>(a4), n,
                                      reinterpret_cast<const float*>(b4), m, theta, softening,
                                      checked, flags, c->d2_out4.as<float>()))
        return rc;
    return narrow(c, n, out);
}

int pnb_bh_accel_f32_2d(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                        size_t m, float theta, float softening, int checked, unsigned flags,
                        float* out) {
    PNB_D2_ARGS_CHECK();
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    const bool same = affected == affecting && n == m;
    if (int rc = stage(c, c->d2_raw_a, affected, n)) return rc;
    if (!same)
        if (int rc = stage(c, c->d2_raw_b, affecting, m)) return rc;
    const float* da = c->d2_raw_a.as<float>();
    const float* db = same ? da : c->d2_raw_b.as<float>();
    PNB_HIP(c, c->d2_out2.ensure(n * 8));
    if (int rc = pnb_bh_accel_f32_2d_dev(c, da, n, db, m, theta, softening, checked, flags,
                                         c->d2_out2.as<float>()))
        return rc;
    PNB_HIP(c, hipMemcpyAsync(out, c->d2_out2.ptr, n * 8, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

}  // extern "C"
