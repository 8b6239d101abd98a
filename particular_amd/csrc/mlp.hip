// Training and standalone inference of the far-field correction MLP (no reference counterpart).
//
// Parameters live as an FP32 master vector (Adam state beside it) and are packed into BF16 MFMA
// operand copies (MlpWeights) after every update. A training set holds rows (node features, the
// node's centre of mass, M w^2 and M w^3, accepting-lane mask, wavefront id), per wavefront the lanes'
// positions (w = softening^2), and per-lane targets (brute force minus Barnes-Hut acceleration)
// with weights 1 / |a_bf|^2. One full-batch step:
//   mlp_fwd_kernel    y = MLP(x) plus x[0..5] on y[0..5], for every row (MFMA, 16-row tiles
//                     per wavefront)
//   mlp_pred_kernel   per (128-row chunk of a wavefront, lane l): the partial sum over the chunk's
//                     rows holding l of the fields of Q = M w^2 y[0..5] (quad_field) and of
//                     O = M w^3 y[6..12] (oct_field) at the lane
//   mlp_resid_kernel  per (wavefront, lane): pred_l from the partials, loss, g_pred
//   mlp_rowgrad_kernel per row: g_y = sum over its lanes of the scaled (d field / d y)^T g_pred
//                     (quad_field_grad, oct_field_grad)
//   mlp_bwd_kernel    recomputes the forward per 16-row tile and back-propagates with MFMA;
//                     weight gradients accumulate in registers, one partial per wavefront
//   mlp_reduce_kernel fixed-order sum of the partials (deterministic)
//   mlp_adam_kernel   Adam on the FP32 master weights, then mlp_pack_kernel
// The loss and its gradients are those of oracle/mlp_oracle.py with the same BF16 rounding points.
#include "pnb_internal.hpp"
#include "mlp_common.hpp"

#include <cmath>
#include <vector>

namespace pnb {

constexpr int MLP_WAVES_PER_BLOCK = 4;
constexpr int MLP_BWD_BLOCKS = 256;

struct MlpState {
    bool ready = false;
    bool octupole = false;  // the output layer may give non-zero octupole outputs
    int adam_t = 0;
    DevBuf params, m, v, grads, weights, partial, loss_hist;
    // training set
    size_t rows = 0, waves = 0, cap_rows = 0, cap_waves = 0;
    unsigned long long dropped = 0;
    DevBuf feats, geom, mask, rwave, lanes, target, wgt, wave_rows;
    DevBuf y, gy, loss_part;
    // the loss in three parallel passes over fixed chunks of each wavefront's rows (derived from
    // wave_rows when the training set changes)
    bool chunks_ok = false;
    size_t chunks = 0;
    float inv_count = 0.f;  // 1 / lanes with a positive weight
    DevBuf chunk_wave, chunk_rb, wave_chunk, part, lane_g;
};

void mlp_free(MlpState* s) {
    if (!s) return;
    for (DevBuf* b : {&s->params, &s->m, &s->v, &s->grads, &s->weights, &s->partial, &s->loss_hist,
                      &s->feats, &s->geom, &s->mask, &s->rwave, &s->lanes, &s->target, &s->wgt,
                      &s->wave_rows, &s->y, &s->gy, &s->loss_part, &s->chunk_wave, &s->chunk_rb,
                      &s->wave_chunk, &s->part, &s->lane_g})
        b->release();
    delete s;
}

// ---- parameter packing -------------------------------------------------------------------------

__global__ void mlp_pack_kernel(const float* __restrict__ p, MlpWeights* __restrict__ w) {
    for (int t = threadIdx.x; t < MLP_H * MLP_H; t += blockDim.x) {
        const int n = t / MLP_H, k = t % MLP_H;
        w->w2t[n][k] = f2bf(p[MLP_OFF_W2 + k * MLP_H + n]);
        w->w2[n][k] = f2bf(p[MLP_OFF_W2 + n * MLP_H + k]);
    }
    for (int t = threadIdx.x; t < MLP_H * MLP_KIN; t += blockDim.x) {
        const int n = t / MLP_KIN, k = t % MLP_KIN;
        w->w1t[n][k] = k < MLP_IN ? f2bf(p[MLP_OFF_W1 + k * MLP_H + n]) : uint16_t(0);
        w->w3[n][k] = k < MLP_OUT ? f2bf(p[MLP_OFF_W3 + n * MLP_OUT + k]) : uint16_t(0);
    }
    for (int t = threadIdx.x; t < MLP_NOUT * MLP_H; t += blockDim.x) {
        const int n = t / MLP_H, k = t % MLP_H;
        w->w3t[n][k] = n < MLP_OUT ? f2bf(p[MLP_OFF_W3 + k * MLP_OUT + n]) : uint16_t(0);
    }
    for (int t = threadIdx.x; t < MLP_H; t += blockDim.x) {
        w->b1[t] = p[MLP_OFF_B1 + t];
        w->b2[t] = p[MLP_OFF_B2 + t];
    }
    for (int t = threadIdx.x; t < MLP_NOUT; t += blockDim.x) w->b3[t] = t < MLP_OUT ? p[MLP_OFF_B3 + t] : 0.f;
}

template <class W>
__device__ void load_weights(W& dst, const MlpWeights* __restrict__ src) {
    const uint32_t* s = reinterpret_cast<const uint32_t*>(src);
    uint32_t* d = reinterpret_cast<uint32_t*>(&dst);
    for (int t = threadIdx.x; t < int(sizeof(W) / 4); t += blockDim.x) d[t] = s[t];
    __syncthreads();
}

// ---- forward over rows --------------------------------------------------------------------------

__device__ __forceinline__ void load_x16(MlpTile& t, const float* __restrict__ feats, size_t row0,
                                         size_t rows, uint16_t (*xt)[16] = nullptr) {
    const int lane = threadIdx.x & 63;
    if (lane < 16) {
        const size_t r = row0 + lane;
        const bool valid = r < rows;
#pragma unroll
        for (int k = 0; k < MLP_KIN; ++k) {
            const uint16_t v = (valid && k < MLP_IN) ? f2bf(feats[r * MLP_FEAT + k]) : uint16_t(0);
            t.x[lane][k] = v;
            if (xt) xt[k][lane] = v;
        }
    }
    wave_lds_sync();
}

// y (rows x MLP_YS, the first MLP_OUT used) = MLP(feats), plus feats[0..5] on y[0..5],
// register-resident forward: the quadrupole outputs are a residual on the node's normalised
// quadrupole (features 0..5).
__global__ __launch_bounds__(256) void mlp_fwd_kernel(const float* __restrict__ feats, size_t rows,
                                                      const MlpWeights* __restrict__ wg,
                                                      float* __restrict__ y) {
    __shared__ MlpForward w;
    load_weights(w, wg);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, g = lane >> 4;
    const size_t gw = size_t(blockIdx.x) * MLP_WAVES_PER_BLOCK + wv;
    const size_t nw = size_t(gridDim.x) * MLP_WAVES_PER_BLOCK;
    // features 4g..4g+3 of row r16 of a tile, read one tile ahead (16-byte loads)
    const auto fetch = [&](size_t tile) {
        const size_t r = tile * 16 + r16;
        return r < rows ? *reinterpret_cast<const float4*>(feats + r * MLP_FEAT + 4 * g)
                        : make_float4(0.f, 0.f, 0.f, 0.f);
    };
    float4 nxv = fetch(gw);
    for (size_t tile = gw; tile * 16 < rows; tile += nw) {
        const size_t r = tile * 16 + r16;
        const float4 xv = nxv;
        nxv = fetch(tile + nw);  // past the last tile: zeros, nothing read
        const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
        bf16x4_t bx;
#pragma unroll
        for (int i = 0; i < 4; ++i) bx[i] = short(f2bf(4 * g + i < MLP_IN ? xa[i] : 0.f));
        const f32x4_t yv = mlp_forward16_regs(w, bx);
        if (g < MLP_YS / 4 && r < rows) {  // lanes of group g hold outputs 4 g .. 4 g + 3
            // plus the residual on 0..5: the row's normalised quadrupole, features 0..5 in FP32
            // (features 4g..4g+3 of this row are the lane's own)
            float o4[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int o = 4 * g + i;
                o4[i] = o < MLP_OUT ? yv[i] + (o < MLP_QOUT ? xa[i] : 0.f) : 0.f;
            }
            *reinterpret_cast<float4*>(y + r * MLP_YS + 4 * g) =
                make_float4(o4[0], o4[1], o4[2], o4[3]);
        }
    }
}

// ---- per-lane aggregation, loss, gradient of y ---------------------------------------------------

// The loss over the lanes, L = sum_lanes w |pred - target|^2 / count with
// pred = sum over the lane's accepted rows of quad_field(y_row[0..5]) + oct_field(y_row[6..12]),
// and its gradient with respect to each row's thirteen outputs, in three passes that each keep every lane of the GPU busy:
//  1. mlp_pred_kernel: per (chunk of MLP_CHUNK rows of one wavefront, lane), the partial pred;
//  2. mlp_resid_kernel: per (wavefront, lane), the chunks' partials summed in chunk order, the
//     residual, the loss share and g = 2 w (pred - target) / count;
//  3. mlp_rowgrad_kernel: per row, sum over its masked lanes (in lane order) of
//     g . d quad_field / d y, times the row's scale.
// All sums run in a fixed order, so the loss and the gradients do not depend on scheduling.
constexpr uint32_t MLP_CHUNK = 128;

__global__ __launch_bounds__(64) void mlp_pred_kernel(const float* __restrict__ y,
                                                      const uint32_t* __restrict__ wave_rows,
                                                      const uint32_t* __restrict__ chunk_wave,
                                                      const uint32_t* __restrict__ chunk_rb,
                                                      const unsigned long long* __restrict__ mask,
                                                      const float4* __restrict__ geom,
                                                      const float4* __restrict__ lanes,
                                                      float4* __restrict__ part) {
    const uint32_t c = blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t w = chunk_wave[c], rb = chunk_rb[c];
    const uint32_t re = min(rb + MLP_CHUNK, wave_rows[w + 1]);
    const float4 p = lanes[size_t(w) * 64 + lane];  // w = softening^2
    float px = 0.f, py = 0.f, pz = 0.f;
    for (uint32_t r = rb; r < re; ++r) {
        if ((mask[r] >> lane) & 1ull) {
            const float4 gm = geom[2 * size_t(r)];
            const float mw3 = geom[2 * size_t(r) + 1].x;
            const float dx = gm.x - p.x, dy = gm.y - p.y, dz = gm.z - p.z;
            const float inv = rsqrtf(dx * dx + dy * dy + dz * dz + p.w);
            float q[6], o[7];
#pragma unroll
            for (int j = 0; j < 6; ++j) q[j] = y[size_t(r) * MLP_YS + j];
#pragma unroll
            for (int j = 0; j < 7; ++j) o[j] = y[size_t(r) * MLP_YS + MLP_QOUT + j];
            quad_field(q, dx, dy, dz, inv, gm.w, px, py, pz);
            oct_field(o, dx, dy, dz, inv, mw3, px, py, pz);
        }
    }
    part[size_t(c) * 64 + lane] = make_float4(px, py, pz, 0.f);
}

__global__ __launch_bounds__(64) void mlp_resid_kernel(const uint32_t* __restrict__ wave_chunk,
                                                       const float4* __restrict__ part,
                                                       const float* __restrict__ target,
                                                       const float* __restrict__ wgt,
                                                       float inv_count,
                                                       float4* __restrict__ lane_g,
                                                       float* __restrict__ loss_part) {
    const size_t w = blockIdx.x;
    const int lane = threadIdx.x;
    float px = 0.f, py = 0.f, pz = 0.f;
    for (uint32_t c = wave_chunk[w]; c < wave_chunk[w + 1]; ++c) {
        const float4 v = part[size_t(c) * 64 + lane];
        px += v.x;
        py += v.y;
        pz += v.z;
    }
    const size_t tl = w * 64 + lane;
    const float dx = px - target[4 * tl], dy = py - target[4 * tl + 1], dz = pz - target[4 * tl + 2];
    const float wl = wgt[tl];
    float loss = wl * (dx * dx + dy * dy + dz * dz) * inv_count;
    const float k = 2.f * wl * inv_count;
    lane_g[tl] = make_float4(k * dx, k * dy, k * dz, 0.f);
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off, 64);
    if (lane == 0) loss_part[w] = loss;
}

__global__ __launch_bounds__(256) void mlp_rowgrad_kernel(size_t rows,
                                                          const uint32_t* __restrict__ rwave,
                                                          const unsigned long long* __restrict__ mask,
                                                          const float4* __restrict__ geom,
                                                          const float4* __restrict__ lanes,
                                                          const float4* __restrict__ lane_g,
                                                          float* __restrict__ gy) {
    const size_t r = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const size_t w = rwave[r];
    unsigned long long m = mask[r];
    const float4 gm = geom[2 * r];
    const float mw3 = geom[2 * r + 1].x;
    float g6[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float g7[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    while (m) {  // the row's lanes in ascending order
        const int l = __builtin_ctzll(m);
        m &= m - 1;
        const float4 p = lanes[w * 64 + l];
        const float4 g = lane_g[w * 64 + l];
        const float ex = gm.x - p.x, ey = gm.y - p.y, ez = gm.z - p.z;
        const float inv = rsqrtf(ex * ex + ey * ey + ez * ez + p.w);
        quad_field_grad(ex, ey, ez, inv, g.x, g.y, g.z, g6);
        oct_field_grad(ex, ey, ez, inv, g.x, g.y, g.z, g7);
    }
#pragma unroll
    for (int j = 0; j < MLP_YS; ++j)
        gy[r * MLP_YS + j] = j < MLP_QOUT ? gm.w * g6[j]
                             : j < MLP_OUT ? mw3 * g7[j - MLP_QOUT] : 0.f;
}

// ---- backward ------------------------------------------------------------------------------------

// Per wavefront and 16-row tile, every product in registers: each activation and gradient the
// backward pass needs is produced in two layouts by computing its product in both orientations
// (MFMA time is cheap here, LDS transposes were not). With lane l = 16 g + r (r = l % 16):
//   chain layout: D of a transposed product Z^T = W^T H^T, lane holds {hidden 4g+i, row r}, which
//     is also the A fragment of H (m = row r, k = hidden 4g+i) and the B fragment of H^T for the
//     next transposed product;
//   grad layout: D of Z = H W, lane holds {row 4g+i, hidden r}: the A fragment of H^T (m = hidden,
//     k = rows) and the B fragment of H (k = rows, n = hidden) of the weight-gradient products,
//     whose reduction runs over the tile's rows.
// Forward: Z1, Z2 in both layouts (biases: b[hidden]); backward: G_h2 in both (from g_y, read
// from memory in both layouts), G_h1 in the grad layout (from G_z2 in the chain layout);
// gradients: gW3 += H2^T g_y, gW2 += H1^T G_z2, gW1 += X^T G_z1, bias gradients summed from the
// grad layouts (rows 4g..4g+3 per lane, then across the four row groups).
__global__ __launch_bounds__(256) void mlp_bwd_kernel(const float* __restrict__ feats,
                                                      const float* __restrict__ gyin, size_t rows,
                                                      const MlpWeights* __restrict__ wg,
                                                      float* __restrict__ partial) {
    __shared__ MlpWeights w;
    load_weights(w, wg);
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, g = lane >> 4;
    const f32x4_t zero = {0.f, 0.f, 0.f, 0.f};
    f32x4_t gw1[4], gw2[4][4], gw3[4];
    float gb1[4] = {0, 0, 0, 0}, gb2[4] = {0, 0, 0, 0}, gb3 = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        gw1[a] = zero;
        gw3[a] = zero;
#pragma unroll
        for (int b = 0; b < 4; ++b) gw2[a][b] = zero;
    }
    const size_t gwave = size_t(blockIdx.x) * MLP_WAVES_PER_BLOCK + wv;
    const size_t nw = size_t(gridDim.x) * MLP_WAVES_PER_BLOCK;
    // A tile's inputs, read one tile ahead (one wavefront per SIMD here: the loads' latency is
    // hidden behind the previous tile's products): x and g_y of row r, columns 4g..4g+3 (chain
    // layout / A fragment of the row-major matrix) and of rows 4g..4g+3, column r (grad layout).
    float4 nxv, ngv;
    float nxs[4], ngs[4];
    const auto fetch = [&](size_t tile) {
        const size_t row0 = tile * 16, r = row0 + r16;
        nxv = make_float4(0.f, 0.f, 0.f, 0.f);
        ngv = nxv;
        if (r < rows) {
            nxv = *reinterpret_cast<const float4*>(feats + r * MLP_FEAT + 4 * g);
            ngv = *reinterpret_cast<const float4*>(gyin + r * MLP_YS + 4 * g);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const size_t rr = row0 + 4 * g + i;
            const bool ok = rr < rows;
            nxs[i] = (ok && r16 < MLP_IN) ? feats[rr * MLP_FEAT + r16] : 0.f;
            ngs[i] = (ok && r16 < MLP_OUT) ? gyin[rr * MLP_YS + r16] : 0.f;
        }
    };
    fetch(gwave);
    for (size_t tile = gwave; tile * 16 < rows; tile += nw) {
        const float4 xv = nxv, gv = ngv;
        float xs[4], gs[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            xs[i] = nxs[i];
            gs[i] = ngs[i];
        }
        fetch(tile + nw);  // past the last tile: zeros, nothing read
        bf16x4_t bx, gyc;
        {
            const float xa[4] = {xv.x, xv.y, xv.z, xv.w}, ga[4] = {gv.x, gv.y, gv.z, gv.w};
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int col = 4 * g + i;
                bx[i] = short(f2bf(col < MLP_IN ? xa[i] : 0.f));
                gyc[i] = short(f2bf(col < MLP_OUT ? ga[i] : 0.f));
            }
        }
        // forward: H1 in the chain layout; the signs of Z2 (chain) and Z1, Z2 (grad) as bit masks
        // (bit 4 t + i), H1 and H2 in the grad layout
        bf16x4_t h1c[4], h1g[4], h2g[4];
        uint32_t pos2c = 0, pos1g = 0, pos2g = 0;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4_t z = mfma16(frag(&w.w1t[16 * mt + r16][4 * g]), bx, zero);
#pragma unroll
            for (int i = 0; i < 4; ++i) z[i] += w.b1[16 * mt + 4 * g + i];
            h1c[mt] = pack_relu(z);
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            f32x4_t z = zero;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                z = mfma16(frag(&w.w2t[16 * mt + r16][16 * ks + 4 * g]), h1c[ks], z);
#pragma unroll
            for (int i = 0; i < 4; ++i)
                pos2c |= (z[i] + w.b2[16 * mt + 4 * g + i] > 0.f ? 1u : 0u) << (4 * mt + i);
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4_t z = mfma16(bx, frag(&w.w1t[16 * nt + r16][4 * g]), zero);
            const float b = w.b1[16 * nt + r16];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                z[i] += b;
                pos1g |= (z[i] > 0.f ? 1u : 0u) << (4 * nt + i);
            }
            h1g[nt] = pack_relu(z);
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4_t z = zero;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                z = mfma16(h1c[ks], frag(&w.w2t[16 * nt + r16][16 * ks + 4 * g]), z);
            const float b = w.b2[16 * nt + r16];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                z[i] += b;
                pos2g |= (z[i] > 0.f ? 1u : 0u) << (4 * nt + i);
            }
            h2g[nt] = pack_relu(z);
        }
        // g_z2 = (g_y W3^T) * relu'(z2): chain layout (the A operand of G_h1) and grad layout
        bf16x4_t g2c[4], g2g[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x4_t w3f = frag(&w.w3[16 * t + r16][4 * g]);
            const f32x4_t ghc = mfma16(w3f, gyc, zero);
            const f32x4_t ghg = mfma16(gyc, w3f, zero);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                g2c[t][i] = short(f2bf((pos2c >> (4 * t + i)) & 1u ? ghc[i] : 0.f));
                const float gz = (pos2g >> (4 * t + i)) & 1u ? ghg[i] : 0.f;
                gb2[t] += gz;
                g2g[t][i] = short(f2bf(gz));
            }
        }
        // g_z1 = (g_z2 W2^T) * relu'(z1), grad layout
        bf16x4_t g1g[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4_t gh = zero;
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                gh = mfma16(g2c[ks], frag(&w.w2[16 * nt + r16][16 * ks + 4 * g]), gh);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const float gz = (pos1g >> (4 * nt + i)) & 1u ? gh[i] : 0.f;
                gb1[nt] += gz;
                g1g[nt][i] = short(f2bf(gz));
            }
        }
        // weight gradients: gW3 += H2^T g_y, gW2 += H1^T g_z2, gW1 += X^T g_z1, with x and g_y
        // of rows 4g..4g+3, column r
        bf16x4_t xg, gyg;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            xg[i] = short(f2bf(xs[i]));
            gyg[i] = short(f2bf(gs[i]));
            gb3 += gs[i];
        }
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            gw3[mt] = mfma16(h2g[mt], gyg, gw3[mt]);
            gw1[mt] = mfma16(xg, g1g[mt], gw1[mt]);
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) gw2[mt][nt] = mfma16(h1g[mt], g2g[nt], gw2[mt][nt]);
        }
    }
    float* out = partial + gwave * MLP_NPARAM;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int k = 4 * g + i, n = 16 * nt + r16;
            if (k < MLP_IN) out[MLP_OFF_W1 + k * MLP_H + n] = gw1[nt][i];
        }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                out[MLP_OFF_W2 + (16 * mt + 4 * g + i) * MLP_H + 16 * nt + r16] = gw2[mt][nt][i];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i)
            if (r16 < MLP_OUT) out[MLP_OFF_W3 + (16 * mt + 4 * g + i) * MLP_OUT + r16] = gw3[mt][i];
    // bias gradients: each lane's sums over its rows 4g..4g+3, then over the four row groups
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        float a = gb1[nt], b = gb2[nt];
        a += __shfl_xor(a, 16, 64);
        a += __shfl_xor(a, 32, 64);
        b += __shfl_xor(b, 16, 64);
        b += __shfl_xor(b, 32, 64);
        if (g == 0) {
            out[MLP_OFF_B1 + 16 * nt + r16] = a;
            out[MLP_OFF_B2 + 16 * nt + r16] = b;
        }
    }
    gb3 += __shfl_xor(gb3, 16, 64);
    gb3 += __shfl_xor(gb3, 32, 64);
    if (g == 0 && r16 < MLP_OUT) out[MLP_OFF_B3 + r16] = gb3;
}

// grads[k] = sum over the partials in a fixed order: 8 groups of 32 threads per workgroup, group
// g summing its contiguous eighth of the partials for 32 parameters, then the eight group sums
// added in group order.
constexpr int MLP_RED_PARAMS = 32;
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* __restrict__ partial,
                                                         int parts, float* __restrict__ grads) {
    __shared__ float acc[8][MLP_RED_PARAMS];
    const int g = threadIdx.x / MLP_RED_PARAMS, t = threadIdx.x % MLP_RED_PARAMS;
    const int k = blockIdx.x * MLP_RED_PARAMS + t;
    const int per = (parts + 7) / 8, p0 = g * per, p1 = min(parts, p0 + per);
    float s = 0.f;
    if (k < MLP_NPARAM)
        for (int p = p0; p < p1; ++p) s += partial[size_t(p) * MLP_NPARAM + k];
    acc[g][t] = s;
    __syncthreads();
    if (g == 0 && k < MLP_NPARAM) {
        float total = acc[0][t];
#pragma unroll
        for (int q = 1; q < 8; ++q) total += acc[q][t];
        grads[k] = total;
    }
}

__global__ void mlp_adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                float* __restrict__ m, float* __restrict__ v, float lr, float c1,
                                float c2) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= MLP_NPARAM) return;
    const float gk = g[k];
    const float mk = 0.9f * m[k] + 0.1f * gk;
    const float vk = 0.999f * v[k] + 0.001f * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + 1e-8f);
}

__global__ void mlp_loss_sum_kernel(const float* __restrict__ part, size_t n, float* __restrict__ out) {
    __shared__ float red[256];
    float s = 0.f;
    for (size_t i = threadIdx.x; i < n; i += 256) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ---- host helpers ------------------------------------------------------------------------------

static uint64_t splitmix(uint64_t& s) {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

int mlp_ensure(pnb_ctx* c) {
    if (!c->mlp) c->mlp = new MlpState();
    MlpState* s = c->mlp;
    PNB_HIP(c, s->params.ensure(MLP_NPARAM * 4));
    PNB_HIP(c, s->m.ensure(MLP_NPARAM * 4));
    PNB_HIP(c, s->v.ensure(MLP_NPARAM * 4));
    PNB_HIP(c, s->grads.ensure(MLP_NPARAM * 4));
    PNB_HIP(c, s->weights.ensure(sizeof(MlpWeights)));
    return PNB_OK;
}

static int pack(pnb_ctx* c) {
    mlp_pack_kernel<<<1, 256, 0, c->stream>>>(c->mlp->params.as<float>(), c->mlp->weights.as<MlpWeights>());
    PNB_CHECK_LAUNCH(c);
    return PNB_OK;
}

const MlpWeights* mlp_weights(pnb_ctx* c) {
    return (c->mlp && c->mlp->ready) ? c->mlp->weights.as<MlpWeights>() : nullptr;
}

// Whether the network may output a non-zero octupole (a non-zero weight or bias of outputs
// 6..12): the corrected walk then fetches every accepted node's octupole with its quadrupole.
bool mlp_octupole(pnb_ctx* c) { return c->mlp && c->mlp->ready && c->mlp->octupole; }

// Appends `rows` rows and `waves` wavefronts (already on the device) to the training set.
int mlp_append(pnb_ctx* c, const float* feats, const float4* geom, const unsigned long long* mask,
               const uint32_t* row_wave_local, const uint32_t* wave_rows_local, size_t rows,
               size_t waves, const float4* lanes, const float* target, const float* wgt,
               unsigned long long dropped);

// Grows a device array, keeping its first `used` bytes.
static int grow(pnb_ctx* c, DevBuf& b, size_t used, size_t want) {
    if (want <= b.cap) return PNB_OK;
    DevBuf nb;
    PNB_HIP(c, nb.ensure(want));
    if (used && b.ptr) PNB_HIP(c, hipMemcpyAsync(nb.ptr, b.ptr, used, hipMemcpyDeviceToDevice, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    b.release();
    b = nb;
    return PNB_OK;
}

__global__ void mlp_offset_kernel(const uint32_t* __restrict__ src, size_t n, uint32_t add,
                                  uint32_t* __restrict__ dst) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + add;
}

int mlp_append(pnb_ctx* c, const float* feats, const float4* geom, const unsigned long long* mask,
               const uint32_t* row_wave_local, const uint32_t* wave_rows_local, size_t rows,
               size_t waves, const float4* lanes, const float* target, const float* wgt,
               unsigned long long dropped) {
    if (int rc = mlp_ensure(c)) return rc;
    MlpState* s = c->mlp;
    const size_t nr = s->rows + rows, nwv = s->waves + waves;
    if (nr > 0xffffffffu) return fail(c, PNB_E_ARG, "training set too large");
    auto cap = [](size_t need, size_t have) { return need <= have ? have : std::max(need, 2 * have); };
    const size_t cr = cap(nr, s->cap_rows), cw = cap(nwv, s->cap_waves);
    if (int rc = grow(c, s->feats, s->rows * MLP_FEAT * 4, cr * MLP_FEAT * 4)) return rc;
    if (int rc = grow(c, s->geom, s->rows * MLP_GEOM * 4, cr * MLP_GEOM * 4)) return rc;
    if (int rc = grow(c, s->mask, s->rows * 8, cr * 8)) return rc;
    if (int rc = grow(c, s->rwave, s->rows * 4, cr * 4)) return rc;
    if (int rc = grow(c, s->lanes, s->waves * 64 * 16, cw * 64 * 16)) return rc;
    if (int rc = grow(c, s->target, s->waves * 64 * 16, cw * 64 * 16)) return rc;
    if (int rc = grow(c, s->wgt, s->waves * 64 * 4, cw * 64 * 4)) return rc;
    if (int rc = grow(c, s->wave_rows, s->waves ? (s->waves + 1) * 4 : 0, (cw + 1) * 4)) return rc;
    s->cap_rows = cr;
    s->cap_waves = cw;
    hipStream_t st = c->stream;
    if (rows) {
        PNB_HIP(c, hipMemcpyAsync(s->feats.as<float>() + s->rows * MLP_FEAT, feats, rows * MLP_FEAT * 4, hipMemcpyDeviceToDevice, st));
        PNB_HIP(c, hipMemcpyAsync(s->geom.as<float>() + s->rows * MLP_GEOM, geom, rows * MLP_GEOM * 4, hipMemcpyDeviceToDevice, st));
        PNB_HIP(c, hipMemcpyAsync(s->mask.as<unsigned long long>() + s->rows, mask, rows * 8, hipMemcpyDeviceToDevice, st));
        mlp_offset_kernel<<<div_up(rows, 256), 256, 0, st>>>(row_wave_local, rows, uint32_t(s->waves),
                                                             s->rwave.as<uint32_t>() + s->rows);
        PNB_CHECK_LAUNCH(c);
    }
    PNB_HIP(c, hipMemcpyAsync(s->lanes.as<float4>() + s->waves * 64, lanes, waves * 64 * 16, hipMemcpyDeviceToDevice, st));
    PNB_HIP(c, hipMemcpyAsync(s->target.as<float>() + s->waves * 64 * 4, target, waves * 64 * 16, hipMemcpyDeviceToDevice, st));
    PNB_HIP(c, hipMemcpyAsync(s->wgt.as<float>() + s->waves * 64, wgt, waves * 64 * 4, hipMemcpyDeviceToDevice, st));
    mlp_offset_kernel<<<div_up(waves + 1, 256), 256, 0, st>>>(wave_rows_local, waves + 1, uint32_t(s->rows),
                                                               s->wave_rows.as<uint32_t>() + s->waves);
    PNB_CHECK_LAUNCH(c);
    PNB_HIP(c, hipStreamSynchronize(st));
    s->rows = nr;
    s->waves = nwv;
    s->dropped += dropped;
    s->chunks_ok = false;
    return PNB_OK;
}

static int step_grads(pnb_ctx* c, float* loss_dev) {
    MlpState* s = c->mlp;
    if (s->waves == 0) return fail(c, PNB_E_STATE, "the training set is empty");
    PNB_HIP(c, s->y.ensure(std::max<size_t>(1, s->rows) * MLP_YS * 4));
    PNB_HIP(c, s->gy.ensure(std::max<size_t>(1, s->rows) * MLP_YS * 4));
    PNB_HIP(c, s->loss_part.ensure(s->waves * 4));
    PNB_HIP(c, s->partial.ensure(size_t(MLP_BWD_BLOCKS) * MLP_WAVES_PER_BLOCK * MLP_NPARAM * 4));
    if (!s->chunks_ok) {
        // once per training set: lanes with a positive weight, and the row chunks
        std::vector<float> hw(s->waves * 64);
        std::vector<uint32_t> wr(s->waves + 1);
        PNB_HIP(c, hipMemcpyAsync(hw.data(), s->wgt.ptr, hw.size() * 4, hipMemcpyDeviceToHost, c->stream));
        PNB_HIP(c, hipMemcpyAsync(wr.data(), s->wave_rows.ptr, wr.size() * 4, hipMemcpyDeviceToHost, c->stream));
        PNB_HIP(c, hipStreamSynchronize(c->stream));
        size_t cnt = 0;
        for (float x : hw) cnt += x > 0.f;
        s->inv_count = 1.f / float(std::max<size_t>(cnt, 1));
        std::vector<uint32_t> cw, crb, wc(s->waves + 1, 0);
        for (size_t w = 0; w < s->waves; ++w) {
            wc[w] = uint32_t(cw.size());
            // at least one chunk per wavefront, so that every lane gets a prediction (0 if no row)
            uint32_t r = wr[w];
            do {
                cw.push_back(uint32_t(w));
                crb.push_back(r);
                r += MLP_CHUNK;
            } while (r < wr[w + 1]);
        }
        wc[s->waves] = uint32_t(cw.size());
        s->chunks = cw.size();
        PNB_HIP(c, s->chunk_wave.ensure(s->chunks * 4));
        PNB_HIP(c, s->chunk_rb.ensure(s->chunks * 4));
        PNB_HIP(c, s->wave_chunk.ensure((s->waves + 1) * 4));
        PNB_HIP(c, s->part.ensure(s->chunks * 64 * sizeof(float4)));
        PNB_HIP(c, s->lane_g.ensure(s->waves * 64 * sizeof(float4)));
        PNB_HIP(c, hipMemcpyAsync(s->chunk_wave.ptr, cw.data(), s->chunks * 4, hipMemcpyHostToDevice, c->stream));
        PNB_HIP(c, hipMemcpyAsync(s->chunk_rb.ptr, crb.data(), s->chunks * 4, hipMemcpyHostToDevice, c->stream));
        PNB_HIP(c, hipMemcpyAsync(s->wave_chunk.ptr, wc.data(), wc.size() * 4, hipMemcpyHostToDevice, c->stream));
        PNB_HIP(c, hipStreamSynchronize(c->stream));
        s->chunks_ok = true;
    }
    const float inv_count = s->inv_count;
    const MlpWeights* wg = s->weights.as<MlpWeights>();
    const unsigned tiles = div_up(std::max<size_t>(s->rows, 1), 16);
    return timed(c, PNB_PROF_MLP, [&]() -> int {
        if (s->rows) {
            mlp_fwd_kernel<<<std::min<unsigned>(1024, div_up(tiles, MLP_WAVES_PER_BLOCK)), 256, 0, c->stream>>>(
                s->feats.as<float>(), s->rows, wg, s->y.as<float>());
            PNB_CHECK_LAUNCH(c);
        }
        mlp_pred_kernel<<<unsigned(s->chunks), 64, 0, c->stream>>>(
            s->y.as<float>(), s->wave_rows.as<uint32_t>(), s->chunk_wave.as<uint32_t>(),
            s->chunk_rb.as<uint32_t>(), s->mask.as<unsigned long long>(), s->geom.as<float4>(),
            s->lanes.as<float4>(), s->part.as<float4>());
        PNB_CHECK_LAUNCH(c);
        mlp_resid_kernel<<<unsigned(s->waves), 64, 0, c->stream>>>(
            s->wave_chunk.as<uint32_t>(), s->part.as<float4>(), s->target.as<float>(),
            s->wgt.as<float>(), inv_count, s->lane_g.as<float4>(), s->loss_part.as<float>());
        PNB_CHECK_LAUNCH(c);
        if (s->rows) {
            mlp_rowgrad_kernel<<<div_up(s->rows, 256), 256, 0, c->stream>>>(
                s->rows, s->rwave.as<uint32_t>(), s->mask.as<unsigned long long>(),
                s->geom.as<float4>(), s->lanes.as<float4>(), s->lane_g.as<float4>(),
                s->gy.as<float>());
            PNB_CHECK_LAUNCH(c);
        }
        mlp_loss_sum_kernel<<<1, 256, 0, c->stream>>>(s->loss_part.as<float>(), s->waves, loss_dev);
        PNB_CHECK_LAUNCH(c);
        mlp_bwd_kernel<<<MLP_BWD_BLOCKS, 256, 0, c->stream>>>(s->feats.as<float>(), s->gy.as<float>(),
                                                             s->rows, wg, s->partial.as<float>());
        PNB_CHECK_LAUNCH(c);
        mlp_reduce_kernel<<<div_up(MLP_NPARAM, MLP_RED_PARAMS), 256, 0, c->stream>>>(
            s->partial.as<float>(), MLP_BWD_BLOCKS * MLP_WAVES_PER_BLOCK, s->grads.as<float>());
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    });
}

}  // namespace pnb

using namespace pnb;

extern "C" {

int pnb_mlp_init(pnb_ctx* c, uint64_t seed) {
    if (!c) return PNB_E_ARG;
    if (int rc = activate(c)) return rc;
    if (int rc = mlp_ensure(c)) return rc;
    MlpState* s = c->mlp;
    // He-normal hidden layers (Box-Muller over SplitMix64), zero biases, and a zero output layer:
    // a fresh network adds nothing to its residual, i.e. starts exactly at the quadrupole term
    std::vector<float> p(MLP_NPARAM, 0.f);
    uint64_t st = seed;
    auto normal = [&]() {
        const double u1 = (double(splitmix(st) >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double u2 = double(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int k = 0; k < MLP_IN * MLP_H; ++k) p[MLP_OFF_W1 + k] = float(normal() * std::sqrt(2.0 / MLP_IN));
    for (int k = 0; k < MLP_H * MLP_H; ++k) p[MLP_OFF_W2 + k] = float(normal() * std::sqrt(2.0 / MLP_H));
    std::vector<float> zeros(MLP_NPARAM, 0.f);
    PNB_HIP(c, hipMemcpyAsync(s->params.ptr, p.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(s->m.ptr, zeros.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(s->v.ptr, zeros.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    if (int rc = pack(c)) return rc;
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    s->adam_t = 0;
    s->ready = true;
    s->octupole = false;  // zero output layer
    return PNB_OK;
}

int pnb_mlp_set_params(pnb_ctx* c, const float* params, size_t n) {
    if (!c || !params || n != MLP_NPARAM) return c ? fail(c, PNB_E_ARG, "expected %d parameters", MLP_NPARAM) : PNB_E_ARG;
    if (int rc = activate(c)) return rc;
    if (int rc = mlp_ensure(c)) return rc;
    MlpState* s = c->mlp;
    std::vector<float> zeros(MLP_NPARAM, 0.f);
    PNB_HIP(c, hipMemcpyAsync(s->params.ptr, params, n * 4, hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(s->m.ptr, zeros.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(s->v.ptr, zeros.data(), MLP_NPARAM * 4, hipMemcpyHostToDevice, c->stream));
    if (int rc = pack(c)) return rc;
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    s->adam_t = 0;
    s->ready = true;
    bool oct = false;
    for (int k = 0; k < MLP_H; ++k)
        for (int o = MLP_QOUT; o < MLP_OUT; ++o) oct |= params[MLP_OFF_W3 + k * MLP_OUT + o] != 0.f;
    for (int o = MLP_QOUT; o < MLP_OUT; ++o) oct |= params[MLP_OFF_B3 + o] != 0.f;
    s->octupole = oct;
    return PNB_OK;
}

int pnb_mlp_get_params(pnb_ctx* c, float* params, size_t n) {
    if (!c || !params || n != MLP_NPARAM) return c ? fail(c, PNB_E_ARG, "expected %d parameters", MLP_NPARAM) : PNB_E_ARG;
    if (!c->mlp || !c->mlp->ready) return fail(c, PNB_E_STATE, "the MLP has not been initialised");
    if (int rc = activate(c)) return rc;
    PNB_HIP(c, hipMemcpyAsync(params, c->mlp->params.ptr, n * 4, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_mlp_num_params(void) { return MLP_NPARAM; }

int pnb_mlp_clear_dataset(pnb_ctx* c) {
    if (!c) return PNB_E_ARG;
    if (c->mlp) {
        c->mlp->rows = 0;
        c->mlp->waves = 0;
        c->mlp->dropped = 0;
        c->mlp->chunks_ok = false;
    }
    return PNB_OK;
}

int pnb_mlp_dataset_info(pnb_ctx* c, size_t* rows, size_t* waves, uint64_t* dropped) {
    if (!c) return PNB_E_ARG;
    MlpState* s = c->mlp;
    if (rows) *rows = s ? s->rows : 0;
    if (waves) *waves = s ? s->waves : 0;
    if (dropped) *dropped = s ? s->dropped : 0;
    return PNB_OK;
}

int pnb_mlp_dataset_read(pnb_ctx* c, float* feats, float* geom, uint64_t* mask, uint32_t* row_wave,
                         uint32_t* wave_rows, float* lanes, float* target, float* wgt) {
    if (!c) return PNB_E_ARG;
    MlpState* s = c->mlp;
    if (!s || s->waves == 0) return fail(c, PNB_E_STATE, "the training set is empty");
    if (int rc = activate(c)) return rc;
    hipStream_t st = c->stream;
    if (feats && s->rows) PNB_HIP(c, hipMemcpyAsync(feats, s->feats.ptr, s->rows * MLP_FEAT * 4, hipMemcpyDeviceToHost, st));
    if (geom && s->rows) PNB_HIP(c, hipMemcpyAsync(geom, s->geom.ptr, s->rows * MLP_GEOM * 4, hipMemcpyDeviceToHost, st));
    if (mask && s->rows) PNB_HIP(c, hipMemcpyAsync(mask, s->mask.ptr, s->rows * 8, hipMemcpyDeviceToHost, st));
    if (row_wave && s->rows) PNB_HIP(c, hipMemcpyAsync(row_wave, s->rwave.ptr, s->rows * 4, hipMemcpyDeviceToHost, st));
    if (wave_rows) PNB_HIP(c, hipMemcpyAsync(wave_rows, s->wave_rows.ptr, (s->waves + 1) * 4, hipMemcpyDeviceToHost, st));
    if (lanes) PNB_HIP(c, hipMemcpyAsync(lanes, s->lanes.ptr, s->waves * 64 * 16, hipMemcpyDeviceToHost, st));
    if (target) PNB_HIP(c, hipMemcpyAsync(target, s->target.ptr, s->waves * 64 * 16, hipMemcpyDeviceToHost, st));
    if (wgt) PNB_HIP(c, hipMemcpyAsync(wgt, s->wgt.ptr, s->waves * 64 * 4, hipMemcpyDeviceToHost, st));
    PNB_HIP(c, hipStreamSynchronize(st));
    return PNB_OK;
}

int pnb_mlp_dataset_load(pnb_ctx* c, const float* feats, const float* geom, const uint64_t* mask,
                         const uint32_t* row_wave, const uint32_t* wave_rows, size_t rows,
                         size_t waves, const float* lanes, const float* target, const float* wgt) {
    if (!c) return PNB_E_ARG;
    if (waves == 0 || !wave_rows || !lanes || !target || !wgt ||
        (rows && (!feats || !geom || !mask || !row_wave)))
        return fail(c, PNB_E_ARG, "incomplete training set");
    if (int rc = activate(c)) return rc;
    DevBuf tmp;
    const size_t bytes = rows * (MLP_FEAT * 4 + MLP_GEOM * 4 + 8 + 4) + (waves + 1) * 4 + waves * 64 * 36;
    PNB_HIP(c, tmp.ensure(bytes));
    char* p = tmp.as<char>();
    float4* dg = reinterpret_cast<float4*>(p); p += rows * MLP_GEOM * 4;  // 16-byte records first
    float4* dl = reinterpret_cast<float4*>(p); p += waves * 64 * 16;
    float* dt = reinterpret_cast<float*>(p); p += waves * 64 * 16;
    float* df = reinterpret_cast<float*>(p); p += rows * MLP_FEAT * 4;
    unsigned long long* dm = reinterpret_cast<unsigned long long*>(p); p += rows * 8;
    uint32_t* drw = reinterpret_cast<uint32_t*>(p); p += rows * 4;
    uint32_t* dwr = reinterpret_cast<uint32_t*>(p); p += (waves + 1) * 4;
    float* dwg = reinterpret_cast<float*>(p);
    hipStream_t st = c->stream;
    hipError_t e = hipSuccess;
    if (rows) {
        if (e == hipSuccess) e = hipMemcpyAsync(df, feats, rows * MLP_FEAT * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dg, geom, rows * MLP_GEOM * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(dm, mask, rows * 8, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(drw, row_wave, rows * 4, hipMemcpyHostToDevice, st);
    }
    if (e == hipSuccess) e = hipMemcpyAsync(dwr, wave_rows, (waves + 1) * 4, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dl, lanes, waves * 64 * 16, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dt, target, waves * 64 * 16, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dwg, wgt, waves * 64 * 4, hipMemcpyHostToDevice, st);
    if (e != hipSuccess) {
        tmp.release();
        return hip_fail(c, e, "training set upload", __FILE__, __LINE__);
    }
    int rc = mlp_append(c, df, dg, dm, drw, dwr, rows, waves, dl, dt, dwg, 0);
    PNB_HIP(c, hipStreamSynchronize(st));
    tmp.release();
    return rc;
}

// One forward/backward over the whole training set without an update: loss and gradients.
int pnb_mlp_grad(pnb_ctx* c, float* loss, float* grads, size_t n) {
    if (!c) return PNB_E_ARG;
    if (grads && n != MLP_NPARAM) return fail(c, PNB_E_ARG, "expected %d gradients", MLP_NPARAM);
    if (!c->mlp || !c->mlp->ready) return fail(c, PNB_E_STATE, "the MLP has not been initialised");
    if (int rc = activate(c)) return rc;
    MlpState* s = c->mlp;
    PNB_HIP(c, s->loss_hist.ensure(16));
    if (int rc = step_grads(c, s->loss_hist.as<float>())) return rc;
    if (loss) PNB_HIP(c, hipMemcpyAsync(loss, s->loss_hist.ptr, 4, hipMemcpyDeviceToHost, c->stream));
    if (grads) PNB_HIP(c, hipMemcpyAsync(grads, s->grads.ptr, n * 4, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

// `iterations` full-batch Adam steps; loss_history (optional) receives the loss before each step.
int pnb_mlp_train(pnb_ctx* c, size_t iterations, float lr, float* loss_history) {
    if (!c) return PNB_E_ARG;
    if (!c->mlp || !c->mlp->ready) return fail(c, PNB_E_STATE, "the MLP has not been initialised");
    if (!(lr > 0.f)) return fail(c, PNB_E_ARG, "learning rate must be > 0");
    if (int rc = activate(c)) return rc;
    MlpState* s = c->mlp;
    PNB_HIP(c, s->loss_hist.ensure(std::max<size_t>(iterations, 1) * 4));
    for (size_t it = 0; it < iterations; ++it) {
        if (int rc = step_grads(c, s->loss_hist.as<float>() + it)) return rc;
        s->adam_t += 1;
        const float c1 = 1.f - std::pow(0.9f, float(s->adam_t)), c2 = 1.f - std::pow(0.999f, float(s->adam_t));
        mlp_adam_kernel<<<div_up(MLP_NPARAM, 256), 256, 0, c->stream>>>(
            s->params.as<float>(), s->grads.as<float>(), s->m.as<float>(), s->v.as<float>(), lr, c1, c2);
        PNB_CHECK_LAUNCH(c);
        if (int rc = pack(c)) return rc;
    }
    if (iterations) s->octupole = true;  // the output layer has been updated
    if (loss_history && iterations)
        PNB_HIP(c, hipMemcpyAsync(loss_history, s->loss_hist.ptr, iterations * 4, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

// Standalone inference: y (rows x 16, 13 used) = MLP(feats (rows x 16)), plus feats[0..5] on
// y[0..5]; device pointers.
int pnb_mlp_forward_dev(pnb_ctx* c, const float* feats, size_t rows, float* y) {
    if (!c) return PNB_E_ARG;
    if (rows && (!feats || !y)) return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!c->mlp || !c->mlp->ready) return fail(c, PNB_E_STATE, "the MLP has not been initialised");
    if (rows == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    const unsigned tiles = div_up(rows, 16);
    mlp_fwd_kernel<<<std::min<unsigned>(1024, div_up(tiles, MLP_WAVES_PER_BLOCK)), 256, 0, c->stream>>>(
        feats, rows, c->mlp->weights.as<MlpWeights>(), y);
    PNB_CHECK_LAUNCH(c);
    return PNB_OK;
}

}  // extern "C"
