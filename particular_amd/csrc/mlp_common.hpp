// Far-field correction MLP: shared device definitions (bh.hip build and walk, mlp.hip training).
//
// Network: 16 inputs -> 64 -> 64 -> 13, ReLU. BF16 weights and activations feed MFMA
// (v_mfma_f32_16x16x16_bf16, FP32 accumulate); biases and everything outside the products are
// FP32. The network maps a description of an internal octree node (its normalised quadrupole,
// mass fraction, size in bodies, depth and normalised octupole) to the coefficients of two
// correction fields: y[0..5] is added to the node's normalised quadrupole (a residual: a zero
// output layer leaves the quadrupole term) to give an effective second moment Q in units of
// M w^2 (quad_field), and y[6..12] are an effective traceless third moment in units of M w^3
// (oct_field), learned from zero. Every lane that accepts the node in the Barnes-Hut walk adds
// both fields at its own position to the node's monopole term. Leaves are never corrected (their
// monopole is exact).
//
// MFMA 16x16x16 bf16 operand layout (lane l of 64): A[m][k] with m = l % 16, k = 4 (l / 16) + i;
// B[k][n] with n = l % 16, k = 4 (l / 16) + i; D[m][n] with n = l % 16, m = 4 (l / 16) + i,
// i = 0..3. Operands are read from LDS arrays that keep k contiguous, so each fragment is one
// 8-byte read.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace pnb {

constexpr int MLP_IN = 16;
constexpr int MLP_H = 64;
constexpr int MLP_OUT = 13;
constexpr int MLP_QOUT = 6;  // outputs 0..5: quadrupole; 6..12: octupole
constexpr int MLP_KIN = 16;   // inputs padded to one K step
constexpr int MLP_NOUT = 16;  // outputs padded to one N tile
constexpr int MLP_OFF_W1 = 0;
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_IN * MLP_H;
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_H;
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_H * MLP_H;
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_H;
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_H * MLP_OUT;
constexpr int MLP_NPARAM = MLP_OFF_B3 + MLP_OUT;  // 6093
constexpr int MLP_FEAT = 16;                      // stored feature stride (all used)
constexpr int MLP_YS = 16;                        // stored output stride (13 used)
constexpr int MLP_GEOM = 8;  // per training row: centre of mass, M w^2, M w^3, 3 zeros

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// BF16 operand copies of the weights (as raw bf16 bits) and FP32 biases. The forward part
// (MlpForward) is a prefix, so inference kernels stage only that much in LDS.
struct MlpForward {
    uint16_t w1t[MLP_H][MLP_KIN];   // [n][k] = W1[k][n], k >= MLP_IN zero
    uint16_t w2t[MLP_H][MLP_H];     // [n][k] = W2[k][n]
    uint16_t w3t[MLP_NOUT][MLP_H];  // [n][k] = W3[k][n], n >= MLP_OUT zero
    float b1[MLP_H], b2[MLP_H], b3[MLP_NOUT];
};
struct MlpWeights : MlpForward {
    uint16_t w2[MLP_H][MLP_H];      // [n][k] = W2[n][k]   (B operand of g W2^T)
    uint16_t w3[MLP_H][MLP_NOUT];   // [n][k] = W3[n][k]   (B operand of g W3^T)
};

// FP32 -> BF16 bits, round to nearest even (gfx950 converts in hardware: v_cvt_pk_bf16_f32)
__device__ __forceinline__ uint16_t f2bf(float f) {
    return __builtin_bit_cast(uint16_t, static_cast<__bf16>(f));
}

__device__ __forceinline__ bf16x4_t frag(const uint16_t* p) {
    return *reinterpret_cast<const bf16x4_t*>(p);
}

__device__ __forceinline__ f32x4_t mfma16(bf16x4_t a, bf16x4_t b, f32x4_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Makes this wavefront's LDS writes visible to its own later reads by other lanes. Only
// wave-scope ordering is needed: the MLP tiles are private to one wavefront, and callers run them
// inside per-wave loops where a workgroup barrier could deadlock.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ bf16x4_t pack_relu(const f32x4_t& z) {
    bf16x4_t r;
#pragma unroll
    for (int i = 0; i < 4; ++i) r[i] = short(f2bf(fmaxf(z[i], 0.f)));
    return r;
}

// Forward of 16 rows computed transposed (Z^T = W^T X^T), entirely in registers: every MFMA
// result D[m][n] (m = 4 (l/16) + i, n = l % 16) is the B fragment (k = 4 (l/16) + i, n = l % 16)
// of the next layer's product. `bx` is the B fragment of X^T: features 4 (l/16) .. +3 of row l % 16.
// Returns D of Y^T: lanes 0..15 hold y[row l][0..3] (component 3 is padding).
__device__ __forceinline__ f32x4_t mlp_forward16_regs(const MlpForward& w, bf16x4_t bx) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, g = lane >> 4;
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    bf16x4_t h1[4], h2[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4_t z = mfma16(frag(&w.w1t[16 * mt + r16][4 * g]), bx, zero);
#pragma unroll
        for (int i = 0; i < 4; ++i) z[i] += w.b1[16 * mt + 4 * g + i];
        h1[mt] = pack_relu(z);
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        f32x4_t z = zero;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) z = mfma16(frag(&w.w2t[16 * mt + r16][16 * ks + 4 * g]), h1[ks], z);
#pragma unroll
        for (int i = 0; i < 4; ++i) z[i] += w.b2[16 * mt + 4 * g + i];
        h2[mt] = pack_relu(z);
    }
    f32x4_t y = zero;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) y = mfma16(frag(&w.w3t[r16][16 * ks + 4 * g]), h2[ks], y);
#pragma unroll
    for (int i = 0; i < 4; ++i) y[i] += (4 * g + i < MLP_OUT) ? w.b3[4 * g + i] : 0.f;
    return y;
}

// Per-wave scratch of a 16-row forward pass.
struct MlpTile {
    uint16_t x[16][MLP_KIN];   // inputs, row-major
    uint16_t h1[16][MLP_H];    // relu(z1), row-major
    uint16_t h2[16][MLP_H];    // relu(z2), row-major
};

// Forward of 16 rows held in t.x; writes y (16 x 3, row-major with stride 4) to `y` when y is not
// null. z1 / z2 accumulators stay in registers (D layout) for the backward pass; h1t / h2t, when
// given, receive the activations transposed ([hidden][row]) for weight gradients.
__device__ __forceinline__ void mlp_forward16(const MlpForward& w, MlpTile& t, float* y,
                                              f32x4_t (&z1)[4], f32x4_t (&z2)[4],
                                              uint16_t (*h1t)[16] = nullptr,
                                              uint16_t (*h2t)[16] = nullptr) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, g = lane >> 4;
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    const bf16x4_t ax = frag(&t.x[r16][4 * g]);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        z1[nt] = mfma16(ax, frag(&w.w1t[16 * nt + r16][4 * g]), zero);
        const float b = w.b1[16 * nt + r16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            z1[nt][i] += b;
            const uint16_t h = f2bf(fmaxf(z1[nt][i], 0.f));
            t.h1[4 * g + i][16 * nt + r16] = h;
            if (h1t) h1t[16 * nt + r16][4 * g + i] = h;
        }
    }
    wave_lds_sync();
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4_t acc = zero;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
            acc = mfma16(frag(&t.h1[r16][16 * ks + 4 * g]), frag(&w.w2t[16 * nt + r16][16 * ks + 4 * g]), acc);
        const float b = w.b2[16 * nt + r16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            acc[i] += b;
            const uint16_t h = f2bf(fmaxf(acc[i], 0.f));
            t.h2[4 * g + i][16 * nt + r16] = h;
            if (h2t) h2t[16 * nt + r16][4 * g + i] = h;
        }
        z2[nt] = acc;
    }
    wave_lds_sync();
    f32x4_t acc = zero;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks)
        acc = mfma16(frag(&t.h2[r16][16 * ks + 4 * g]), frag(&w.w3t[r16][16 * ks + 4 * g]), acc);
    if (y && r16 < MLP_OUT) {
        const float b = w.b3[r16];
#pragma unroll
        for (int i = 0; i < 4; ++i) y[(4 * g + i) * 4 + r16] = acc[i] + b;
    }
    wave_lds_sync();
}

// Field of a second moment Q = scale * q at a point: q = {xx, yy, zz, xy, xz, yz}, d = centre of
// mass - point, inv = 1 / |d| (softened like the monopole term). The quadrupole term of the
// multipole expansion, a = -Q d |d|^-5 + 2.5 (d.Q.d) d |d|^-7, is added to (ax, ay, az).
__device__ __forceinline__ void quad_field(const float (&q)[6], float dx, float dy, float dz,
                                           float inv, float scale, float& ax, float& ay, float& az) {
    const float qx = q[0] * dx + q[3] * dy + q[4] * dz;
    const float qy = q[3] * dx + q[1] * dy + q[5] * dz;
    const float qz = q[4] * dx + q[5] * dy + q[2] * dz;
    const float dqd = dx * qx + dy * qy + dz * qz;
    const float inv2 = inv * inv;
    const float s5 = scale * (inv2 * inv2 * inv);
    const float t = 2.5f * dqd * inv2;
    ax += s5 * (t * dx - qx);
    ay += s5 * (t * dy - qy);
    az += s5 * (t * dz - qz);
}

// g . (d a / d q_j) for the six components of quad_field's q, added to g6.
__device__ __forceinline__ void quad_field_grad(float dx, float dy, float dz, float inv, float gx,
                                                float gy, float gz, float (&g6)[6]) {
    const float inv2 = inv * inv;
    const float inv5 = inv2 * inv2 * inv;
    const float gd = 2.5f * (gx * dx + gy * dy + gz * dz) * inv2;  // 2.5 (g.d) / |d|^2
    g6[0] += inv5 * (gd * dx * dx - dx * gx);
    g6[1] += inv5 * (gd * dy * dy - dy * gy);
    g6[2] += inv5 * (gd * dz * dz - dz * gz);
    g6[3] += inv5 * (2.f * gd * dx * dy - (dy * gx + dx * gy));
    g6[4] += inv5 * (2.f * gd * dx * dz - (dz * gx + dx * gz));
    g6[5] += inv5 * (2.f * gd * dy * dz - (dz * gy + dy * gz));
}

// The seven coefficients of a traceless symmetric third-order tensor O, {xxx, yyy, zzz, xxy, xxz,
// xyy, xyz} (the rest follow from O_iik = 0: xzz = -xxx - xyy, yzz = -xxy - yyy,
// yyz = -xxz - zzz). With d = centre of mass - point, F_i = O_ijk d_j d_k and G = d.F:
// oct_basis gives F's derivative with respect to each coefficient (B_j, a 3-vector).
__device__ __forceinline__ void oct_basis(float dx, float dy, float dz, float (&bx)[7],
                                          float (&by)[7], float (&bz)[7]) {
    const float xx = dx * dx, yy = dy * dy, zz = dz * dz;
    const float xy = 2.f * dx * dy, xz = 2.f * dx * dz, yz = 2.f * dy * dz;
    bx[0] = xx - zz; by[0] = 0.f;     bz[0] = -xz;
    bx[1] = 0.f;     by[1] = yy - zz; bz[1] = -yz;
    bx[2] = 0.f;     by[2] = -yz;     bz[2] = zz - yy;
    bx[3] = xy;      by[3] = xx - zz; bz[3] = -yz;
    bx[4] = xz;      by[4] = -yz;     bz[4] = xx - yy;
    bx[5] = yy - zz; by[5] = xy;      bz[5] = -xz;
    bx[6] = yz;      by[6] = xz;      bz[6] = xy;
}

// Field at a point of the octupole term of a third moment O = scale * o (o as in oct_basis):
// the gradient of O_ijk r_i r_j r_k / |r|^7 (r = point - centre of mass), in terms of d = -r,
// a = scale (3 F |d|^-7 - 7 G d |d|^-9), added to (ax, ay, az). inv = 1 / |d| (softened like the
// monopole term). With o = (5/2) U~ / (M w^3), U~ the traceless part of the node's third moments
// about its centre of mass and scale = M w^3, this is the multipole expansion's octupole term.
__device__ __forceinline__ void oct_field(const float (&o)[7], float dx, float dy, float dz,
                                          float inv, float scale, float& ax, float& ay, float& az) {
    float bx[7], by[7], bz[7];
    oct_basis(dx, dy, dz, bx, by, bz);
    float fx = 0.f, fy = 0.f, fz = 0.f;
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        fx += o[j] * bx[j];
        fy += o[j] * by[j];
        fz += o[j] * bz[j];
    }
    const float g = dx * fx + dy * fy + dz * fz;
    const float inv2 = inv * inv;
    const float s7 = scale * (inv2 * inv2) * (inv2 * inv);
    const float t = 7.f * g * inv2;
    ax += s7 * (3.f * fx - t * dx);
    ay += s7 * (3.f * fy - t * dy);
    az += s7 * (3.f * fz - t * dz);
}

// g . (d a / d o_j) for the seven coefficients of oct_field's o (scale 1), added to g7.
__device__ __forceinline__ void oct_field_grad(float dx, float dy, float dz, float inv, float gx,
                                               float gy, float gz, float (&g7)[7]) {
    float bx[7], by[7], bz[7];
    oct_basis(dx, dy, dz, bx, by, bz);
    const float inv2 = inv * inv;
    const float inv7 = (inv2 * inv2) * (inv2 * inv);
    const float gd = 7.f * (gx * dx + gy * dy + gz * dz) * inv2;  // 7 (g.d) / |d|^2
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        const float c = dx * bx[j] + dy * by[j] + dz * bz[j];  // dG / do_j
        g7[j] += inv7 * (3.f * (gx * bx[j] + gy * by[j] + gz * bz[j]) - gd * c);
    }
}

}  // namespace pnb
