// Barnes-Hut on gfx950: the reference's octree rebuilt in parallel, node for node, plus a
// wavefront-coherent stackless traversal with the reference's per-particle acceptance rule.
//
// Reference behaviour (particular/src/tree/mod.rs:73-139, partition.rs:136-220,
// sequential.rs:456-506):
//  * root box = square_with(min/max of the coordinates), in f32;
//  * a node subdivides when its particles do not all share one position; child index bit a is
//    (pos[a] < centre[a]) with centre = (min + max) / 2 of the node box, in f32; bit 0 takes
//    [centre, max], bit 1 takes [min, centre]; children are built depth-first in orthant order,
//    so node ids are a pre-order and the particles of a node are contiguous in "tree order";
//  * node data = centre of mass of its particles; traversal skips a node at zero distance,
//    opens an internal node when (theta^2 d^2) < width^2 (width = x extent of its box), and
//    otherwise adds the node's pair term.
//
// Build (all on the GPU, no recursion):
//  1. bbox reduction and the root box (same f32 operations as square_with);
//  2. per particle, its path through the tree as a 63-bit key: 21 levels of the same f32 midpoint
//     descent, 3 bits per level (the child index). Sorting by key (stable) gives tree order;
//  3. for adjacent particles p-1, p: lcp[p] = number of leading levels they share (or "same
//     position"). Runs that share all 21 levels but hold distinct positions are refined with
//     further 21-level keys until separated;
//  4. with runs of identical positions [r0, r1): the deepest internal level around r0 is
//     I = max(lcp[r0], lcp[r1]); the nodes starting at r0 are levels lcp[r0]+1 .. I+1 (the last a
//     leaf). Counting them per particle and scanning gives every node its pre-order index; a
//     node's subtree ends at the first particle with lcp < level, and skip = index of the first
//     node starting there;
//  5. FP64 mass / mass-weighted position sums: leaves directly, internal nodes from their
//     children, one launch per level from the deepest up over that level's list of internal
//     nodes (built from the same per-particle level ranges); node records get the f32 centre of
//     mass.
// Traversal: one wavefront walks the pre-order node array with one shared cursor. A lane that
// opens node k lets the wave step to k+1; a lane that accepts or skips k waits until the cursor
// reaches skip[k]. Each lane therefore applies exactly the reference's decisions; node records
// are wave-uniform, so they arrive through scalar loads.
// Optional relative criterion (PNB_BH_RELATIVE, no reference counterpart): a node also opens when
// its first neglected term exceeds alpha times the particle's estimated acceleration; an
// estimate pass over one centroid per wavefront gives that acceleration, walk records carrying
// M w^p replace the plain ones, and small opened nodes are summed body by body (buckets).
#include "pnb_internal.hpp"
#include <hip/hip_fp16.h>
#include "mlp_common.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

namespace pnb {

int radix_sort_pairs(pnb_ctx* c, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                     uint32_t* vals_alt, size_t n, int key_bits, DevBuf& hist_buf,
                     DevBuf& bits_buf, uint64_t** keys_out, uint32_t** vals_out,
                     bool all_digits = false);
int split_affecting(pnb_ctx* c, const float4* particles, size_t n, float4* dst, uint32_t* m_host);

constexpr int KEY_LEVELS = 21;
constexpr int LCP_SAME = 0x3fffffff;  // adjacent particles at the same position
constexpr int MAX_LEVEL = 252;        // 12 rounds of 21 levels
constexpr int DEEP_STACK = 24;

constexpr uint32_t FLAG_NONFINITE = 1u;
constexpr uint32_t FLAG_DEPTH = 2u;
constexpr uint32_t FLAG_CAPACITY = 4u;  // capturable build: more nodes or levels than sized for

struct __align__(32) NodeRec {
    float cx, cy, cz, mass;  // centre of mass (f32) and total mu
    float w2;                // width^2 of an internal node; -inf for a leaf (never opened)
    uint32_t skip;           // pre-order index of the first node after this subtree
    uint32_t begin, end;     // particle range in tree order
};
static_assert(sizeof(NodeRec) == 32, "node record is 32 bytes");


// Mass-weighted second moments of a node about its FP64 centre of mass,
// T = sum m (x - c)(x - c)^T, as {xx, yy, zz, xy, xz, yz}; built for the learned correction only.
// FP64 node sums in one 64-byte record: m = (mass, mass-weighted x, y, z) and p = (x, y, z,
// count) of the positions, for the centre of a massless node. One record per child keeps the
// bottom-up pass at one cache line per child for the sums.
struct __align__(64) NodeSums {
    double4 m, p;
};

struct NodeMoments {
    double xx = 0, yy = 0, zz = 0, xy = 0, xz = 0, yz = 0;
    // third moments U = sum m (x - c)(x - c)(x - c) about the same centre:
    // {xxx, yyy, zzz, xxy, xxz, xyy, yyz, xzz, yzz, xyz}
    double u[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
};
// Traceless quadrupole normalised by mass and width: Q = (3 T - tr(T) I) / (M w^2), f32,
// a = {xx, yy, zz, xy}, b = {xz, yz, 0, 0}.
struct __align__(32) NodeQuad {
    float4 a, b;
};

// Per internal node (built with the moments): its normalised octupole (5/2) U~ / (M w^3), U~ the
// traceless part of its third moments about its centre of mass, as the seven coefficients
// {xxx, yyy, zzz, xxy, xxz, xyy, xyz} (oct_basis); the octupole term of its multipole expansion
// is oct_field of these with scale M w^3.
struct __align__(32) NodeDesc {
    float oct[7];
    float pad;
};
constexpr float OCT_FEATURE_SCALE = 32.f;  // network input f9..15 = 32 x NodeDesc::oct

struct BhInfo {
    float mn[3], mx[3];  // root box
    uint32_t flags;
    int max_level;
    uint32_t total;     // node count (S[M])
    uint32_t max_rows;  // most rows one wavefront listed in the last row-listing walk
    unsigned long long steps, accepted;  // traversal counters of the last walk
    uint32_t lvl_count[MAX_LEVEL + 2];   // internal nodes per level
    uint32_t lvl_off[MAX_LEVEL + 2];     // their first slot in the level-grouped list
};

// Per-level counters that many workgroups update: one 128-byte line each, so that atomics on
// different levels do not queue behind each other in one L2 line.
constexpr int LVL_STRIDE = 32;  // uint32 per counter
constexpr int LVL_SLOTS = MAX_LEVEL + 2;

struct BhState {
    size_t m = 0;
    size_t nodes = 0;
    size_t internal = 0;  // internal nodes (listed by level in ilist)
    int max_level = -1;
    bool built = false;
    const uint64_t* skeys = nullptr;  // sorted keys (points into keys or keys_alt)
    const uint32_t* sidx = nullptr;   // tree order -> original index
    DevBuf keys, keys_alt, idx, idx_alt, sp, lcp, cnt, ip, key2, rec, sums, level, width;
    DevBuf skipv;            // per node: the skip pointer alone (the bottom-up pass's child walk)
    DevBuf ilist;            // internal nodes grouped by level (bottom-up pass)
    DevBuf lvl;              // padded per-level counts and list cursors (2 x LVL_SLOTS lines)
    DevBuf mom, quad, desc;  // second moments (FP64), normalised quadrupoles, child shares
    DevBuf qh, oh;           // per node: MLP(node) as halves, the learned correction's moments
    bool has_moments = false;
    // relative opening criterion (PNB_BH_RELATIVE): walk records carrying M w^3, per affected
    // particle alpha |a| from the estimate pass, and the criterion's parameters
    DevBuf rrec, est, cen, cen_spread, est_spread;
    DevBuf wkey, wkey_alt, wval, wval_alt;  // wavefronts ordered by their estimate
    bool rrec_ok = false;
    int rrec_power = 0, rrec_bucket = 0;
    int rel_bucket = 24;  // nodes of at most this many bodies are summed directly when opened
    float rel_alpha = 1e-4f, rel_theta_est = 1.2f;
    bool rel_per_particle = false;  // else one estimate per wavefront, at its centroid
    DevBuf info, bbox_part, hist, bits, afg_tmp;
    // capturable builds (pnb_bh_set_capturable): sized once by an ordinary build, then no host
    // round trip; dd_last: the last build's node count and status are only on the device
    bool dd_mode = false, dd_last = false, dd_mlp = false;
    size_t dd_m = 0, dd_n = 0, dd_cap = 0;
    int dd_levels = 0;
    std::vector<uint32_t> last_lvl;  // internal nodes per level of the last ordinary build
    std::vector<uint32_t> dd_plan;   // the same, of the build that sized the capturable one
    // reference-exact centres of mass (PNB_BH_REF_COM): per body its node at the current level,
    // (group, input index) keys and the bodies regrouped in that order
    DevBuf rc_cur, rc_keys, rc_keys_alt, rc_vals, rc_vals_alt, rc_ord;
    DevBuf akeys, akeys_alt, aidx, aidx_alt, apts, h_aff, h_afg, h_out;
    DevBuf c_starts, c_pts, c_bh, c_bf, c_list, c_count, c_target, c_wgt, c_rows, c_rwave,
        c_cfeats, c_cgeom, c_cmask, c_lanes;
    DevBuf feat_out;  // pnb_bh_node_features
    float last_theta = 0.f;
    unsigned long long last_steps = 0, last_accepted = 0;
};

void bh_free(BhState* s) {
    if (!s) return;
    for (DevBuf* b : {&s->keys, &s->keys_alt, &s->idx, &s->idx_alt, &s->sp, &s->lcp, &s->cnt,
                      &s->ip, &s->key2, &s->rec, &s->sums, &s->level, &s->width, &s->skipv,
                      &s->ilist, &s->lvl, &s->mom, &s->quad, &s->desc, &s->qh, &s->oh, &s->rrec, &s->est,
                      &s->cen, &s->cen_spread, &s->est_spread, &s->wkey, &s->wkey_alt, &s->wval, &s->wval_alt,
                      &s->info, &s->bbox_part, &s->hist, &s->bits, &s->afg_tmp, &s->akeys,
                      &s->akeys_alt, &s->aidx, &s->aidx_alt, &s->apts, &s->h_aff, &s->h_afg,
                      &s->h_out, &s->c_starts, &s->c_pts, &s->c_bh, &s->c_bf, &s->c_list,
                      &s->c_count, &s->c_target, &s->c_wgt, &s->c_rows, &s->c_rwave, &s->c_cfeats,
                      &s->c_cgeom, &s->c_cmask, &s->c_lanes, &s->feat_out})
        b->release();
    delete s;
}

// ---- f32 helpers with the reference's rounding (no contraction) ------------------------------
// hipcc contracts a*b + c into one FMA by default, also through __fadd_rn / __fmul_rn (plain
// operators in HIP); the reference rounds every product and sum, so these switch contraction off.

__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
    return a + b;
}
__device__ __forceinline__ float sub_rn(float a, float b) {
#pragma clang fp contract(off)
    return a - b;
}
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
    return a * b;
}

__device__ __forceinline__ float midpoint(float a, float b) { return mul_rn(add_rn(a, b), 0.5f); }

// glam Vec3::length_squared: (x*x + y*y) + z*z
__device__ __forceinline__ float norm2_ref(float x, float y, float z) {
    return add_rn(add_rn(mul_rn(x, x), mul_rn(y, y)), mul_rn(z, z));
}

__device__ __forceinline__ uint64_t descend_key(float x, float y, float z, float* mn, float* mx,
                                                int levels) {
    uint64_t key = 0;
    for (int l = 0; l < levels; ++l) {
        const float cx = midpoint(mn[0], mx[0]), cy = midpoint(mn[1], mx[1]),
                    cz = midpoint(mn[2], mx[2]);
        const uint32_t bx = x < cx, by = y < cy, bz = z < cz;
        if (bx) mx[0] = cx; else mn[0] = cx;
        if (by) mx[1] = cy; else mn[1] = cy;
        if (bz) mx[2] = cz; else mn[2] = cz;
        key = (key << 3) | (bx | (by << 1) | (bz << 2));
    }
    return key;
}

__device__ __forceinline__ bool same_pos(const float4& a, const float4& b) {
    return a.x == b.x && a.y == b.y && a.z == b.z;
}

// ---- 1. bounding box -------------------------------------------------------------------------

constexpr int BB_BLOCKS = 512;

__global__ __launch_bounds__(256) void bh_bbox_kernel(const float4* __restrict__ p, size_t m,
                                                      float* __restrict__ part,
                                                      BhInfo* __restrict__ info) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    bool bad = false;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < m;
         i += size_t(gridDim.x) * blockDim.x) {
        const float4 q = p[i];
        bad |= !(isfinite(q.x) && isfinite(q.y) && isfinite(q.z));
        mn[0] = fminf(mn[0], q.x); mx[0] = fmaxf(mx[0], q.x);
        mn[1] = fminf(mn[1], q.y); mx[1] = fmaxf(mx[1], q.y);
        mn[2] = fminf(mn[2], q.z); mx[2] = fmaxf(mx[2], q.z);
    }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
        }
    __shared__ float red[4][6];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int a = 0; a < 3; ++a) { red[w][a] = mn[a]; red[w][3 + a] = mx[a]; }
    if (bad) atomicOr(&info->flags, FLAG_NONFINITE);
    __syncthreads();
    if (threadIdx.x < 6) {
        float v = red[0][threadIdx.x];
        for (int ww = 1; ww < 4; ++ww)
            v = threadIdx.x < 3 ? fminf(v, red[ww][threadIdx.x]) : fmaxf(v, red[ww][threadIdx.x]);
        part[blockIdx.x * 6 + threadIdx.x] = v;
    }
}

// square_with (partition.rs:136-153): centre = (min+max)/2, half = max(0, sizes...)/2,
// box = centre -+ half, all in f32.
// One wavefront: the partial boxes strided over the lanes, then a lane reduction (min and max
// are exact, so the order does not matter).
__global__ void bh_root_kernel(const float* __restrict__ part, int blocks, BhInfo* __restrict__ info) {
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = threadIdx.x; b < blocks; b += 64)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], part[b * 6 + a]);
            mx[a] = fmaxf(mx[a], part[b * 6 + 3 + a]);
        }
    for (int off = 32; off > 0; off >>= 1)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], __shfl_xor(mn[a], off, 64));
            mx[a] = fmaxf(mx[a], __shfl_xor(mx[a], off, 64));
        }
    if (threadIdx.x != 0) return;
    float centre[3], big = 0.0f;
    for (int a = 0; a < 3; ++a) {
        centre[a] = midpoint(mn[a], mx[a]);
        big = fmaxf(big, sub_rn(mx[a], mn[a]));
    }
    const float half = midpoint(big, 0.0f);
    for (int a = 0; a < 3; ++a) {
        info->mn[a] = sub_rn(centre[a], half);
        info->mx[a] = add_rn(centre[a], half);
    }
}

__global__ void bh_info_init_kernel(BhInfo* info, uint32_t* __restrict__ lvl) {
    for (int l = threadIdx.x; l < 2 * LVL_SLOTS; l += blockDim.x) lvl[l * LVL_STRIDE] = 0;
    if (threadIdx.x == 0) {
        info->flags = 0;
        info->max_level = -1;
        info->total = 0;
        info->max_rows = 0;
        info->steps = 0;
        info->accepted = 0;
    }
    for (int l = threadIdx.x; l < MAX_LEVEL + 2; l += blockDim.x) info->lvl_count[l] = 0;
}

// ---- 2. keys ------------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void bh_keys_kernel(const float4* __restrict__ p, size_t m,
                                                      const BhInfo* __restrict__ info,
                                                      uint64_t* __restrict__ keys,
                                                      uint32_t* __restrict__ idx) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= m) return;
    const float4 q = p[i];
    float mn[3] = {info->mn[0], info->mn[1], info->mn[2]};
    float mx[3] = {info->mx[0], info->mx[1], info->mx[2]};
    keys[i] = descend_key(q.x, q.y, q.z, mn, mx, KEY_LEVELS);
    idx[i] = uint32_t(i);
}

__global__ __launch_bounds__(256) void bh_gather_kernel(const float4* __restrict__ src, size_t m,
                                                        const uint32_t* __restrict__ idx,
                                                        float4* __restrict__ dst) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < m) dst[i] = src[idx[i]];
}

// ---- 3. common prefixes and deep runs -----------------------------------------------------------

__global__ __launch_bounds__(256) void bh_lcp_kernel(const uint64_t* __restrict__ keys,
                                                     const float4* __restrict__ sp, size_t m,
                                                     int* __restrict__ lcp) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= m) return;
    if (p == 0) {
        lcp[0] = -1;
        return;
    }
    const uint64_t a = keys[p - 1], b = keys[p];
    int v;
    if (a != b) v = (__clzll(a ^ b) - 1) / 3;
    else v = same_pos(sp[p - 1], sp[p]) ? LCP_SAME : KEY_LEVELS;
    lcp[p] = v;
}

// One thread per run of equal 21-level keys that holds distinct positions: extend the keys by
// further 21-level rounds from the run's box and insertion-sort the run (stable), until every
// adjacent pair is either separated or at the same position.
__global__ __launch_bounds__(256) void bh_deep_kernel(const uint64_t* __restrict__ keys, size_t m,
                                                      float4* __restrict__ sp,
                                                      uint32_t* __restrict__ idx,
                                                      int* __restrict__ lcp,
                                                      uint64_t* __restrict__ key2,
                                                      BhInfo* __restrict__ info) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p + 1 >= m) return;
    const uint64_t k = keys[p];
    if (keys[p + 1] != k || (p > 0 && keys[p - 1] == k)) return;
    size_t r1 = p + 1;
    bool deep = false;
    while (r1 < m && keys[r1] == k) {
        deep |= lcp[r1] == KEY_LEVELS;
        ++r1;
    }
    if (!deep) return;

    struct Work {
        uint32_t b, e;
        int base;
        float mn[3], mx[3];
    };
    Work st[DEEP_STACK];
    int top = 0;
    {
        Work w;
        w.b = uint32_t(p);
        w.e = uint32_t(r1);
        w.base = KEY_LEVELS;
        for (int a = 0; a < 3; ++a) { w.mn[a] = info->mn[a]; w.mx[a] = info->mx[a]; }
        const float4 q = sp[p];
        descend_key(q.x, q.y, q.z, w.mn, w.mx, KEY_LEVELS);
        st[top++] = w;
    }
    while (top > 0) {
        Work w = st[--top];
        if (w.base + KEY_LEVELS > MAX_LEVEL) {
            atomicOr(&info->flags, FLAG_DEPTH);
            return;
        }
        for (uint32_t q = w.b; q < w.e; ++q) {
            const float4 v = sp[q];
            float mn[3] = {w.mn[0], w.mn[1], w.mn[2]}, mx[3] = {w.mx[0], w.mx[1], w.mx[2]};
            key2[q] = descend_key(v.x, v.y, v.z, mn, mx, KEY_LEVELS);
        }
        for (uint32_t q = w.b + 1; q < w.e; ++q) {  // stable insertion sort by key2
            const uint64_t kk = key2[q];
            const float4 pv = sp[q];
            const uint32_t iv = idx[q];
            uint32_t r = q;
            while (r > w.b && key2[r - 1] > kk) {
                key2[r] = key2[r - 1];
                sp[r] = sp[r - 1];
                idx[r] = idx[r - 1];
                --r;
            }
            key2[r] = kk;
            sp[r] = pv;
            idx[r] = iv;
        }
        uint32_t q = w.b;
        while (q < w.e) {
            const uint32_t s = q++;
            bool sub_deep = false;
            while (q < w.e && key2[q] == key2[s]) {
                const bool same = same_pos(sp[q - 1], sp[q]);
                lcp[q] = same ? LCP_SAME : w.base + KEY_LEVELS;
                sub_deep |= !same;
                ++q;
            }
            if (q < w.e) lcp[q] = w.base + (__clzll(key2[q - 1] ^ key2[q]) - 1) / 3;
            if (sub_deep) {
                if (top == DEEP_STACK) {
                    atomicOr(&info->flags, FLAG_DEPTH);
                    return;
                }
                Work c;
                c.b = s;
                c.e = q;
                c.base = w.base + KEY_LEVELS;
                for (int a = 0; a < 3; ++a) { c.mn[a] = w.mn[a]; c.mx[a] = w.mx[a]; }
                const float4 v = sp[s];
                descend_key(v.x, v.y, v.z, c.mn, c.mx, KEY_LEVELS);
                st[top++] = c;
            }
        }
    }
}

// ---- 4. node counts, pre-order indices, node records --------------------------------------------

// Workgroups of the level kernels below take contiguous runs of BH_CHUNK-particle chunks, so
// that each issues one global atomic per level for its whole range: per-chunk atomics on the
// few level counters serialised in L2 (0.2 ms per kernel at 4M particles).
constexpr int BH_CHUNK = 256;

__device__ __forceinline__ void chunk_range(size_t m, size_t& c0, size_t& c1) {
    const size_t chunks = (m + BH_CHUNK - 1) / BH_CHUNK;
    const size_t per = (chunks + gridDim.x - 1) / gridDim.x;
    c0 = min(chunks, size_t(blockIdx.x) * per);
    c1 = min(chunks, c0 + per);
}

// Node counts per particle: a run of equal positions starting at p (lcp[p] != LCP_SAME) starts
// the nodes of levels lcp[p] + 1 .. I + 1 (I + 1 is its leaf); also the deepest level and the
// number of internal nodes per level.
__global__ __launch_bounds__(256) void bh_count_kernel(const int* __restrict__ lcp, size_t m,
                                                       uint32_t* __restrict__ cnt,
                                                       int* __restrict__ ip,
                                                       BhInfo* __restrict__ info,
                                                       uint32_t* __restrict__ lvl_count) {
    __shared__ uint32_t hist[LVL_SLOTS];
    for (int l = threadIdx.x; l < LVL_SLOTS; l += blockDim.x) hist[l] = 0;
    __syncthreads();
    const bool lead = (threadIdx.x & 63) == 0;
    size_t c0, c1;
    chunk_range(m, c0, c1);
    int deepest = -1;
    for (size_t ch = c0; ch < c1; ++ch) {
        const size_t p = ch * BH_CHUNK + threadIdx.x;
        int top = -1, lo = MAX_LEVEL + 1;
        if (p < m) {
            const int lp = lcp[p];  // -1 at p = 0
            if (lp == LCP_SAME) {
                cnt[p] = 0;
            } else {
                size_t r1 = p + 1;
                while (r1 < m && lcp[r1] == LCP_SAME) ++r1;
                const int nxt = r1 < m ? lcp[r1] : -1;
                const int I = max(lp, nxt);
                cnt[p] = uint32_t(I + 1 - lp);
                ip[p] = I;
                top = I + 1;
                lo = lp + 1;
            }
        }
        const int hi = top - 1;  // internal levels lo .. hi
        int wlo = lo, wtop = top;
        for (int off = 32; off > 0; off >>= 1) {
            wtop = max(wtop, __shfl_xor(wtop, off, 64));
            wlo = min(wlo, __shfl_xor(wlo, off, 64));
        }
        deepest = max(deepest, wtop);
        // levels above MAX_LEVEL are not counted: the build fails on such a tree
        const int whi = min(wtop - 1, MAX_LEVEL);
        for (int l = wlo; l <= whi; ++l) {
            const uint64_t in = __ballot(lo <= l && l <= hi);
            if (lead && in) atomicAdd(&hist[l], uint32_t(__popcll(in)));
        }
    }
    if (lead && deepest >= 0) {  // only when it raises the maximum seen so far
        if (deepest > __atomic_load_n(&info->max_level, __ATOMIC_RELAXED))
            atomicMax(&info->max_level, deepest);
        if (deepest > MAX_LEVEL) atomicOr(&info->flags, FLAG_DEPTH);
    }
    __syncthreads();
    for (int l = threadIdx.x; l < LVL_SLOTS; l += blockDim.x)
        if (hist[l]) atomicAdd(&lvl_count[l * LVL_STRIDE], hist[l]);
}

// Copies the per-level counts into info (for the host) and sets the list cursors to their
// exclusive scan.
__global__ __launch_bounds__(256) void bh_level_offsets_kernel(BhInfo* __restrict__ info,
                                                               const uint32_t* __restrict__ lvl_count,
                                                               uint32_t* __restrict__ lvl_cur) {
    __shared__ uint32_t v[LVL_SLOTS];
    for (int l = threadIdx.x; l < LVL_SLOTS; l += blockDim.x) v[l] = lvl_count[l * LVL_STRIDE];
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int l = 0; l < LVL_SLOTS; ++l) {
            const uint32_t n_l = v[l];
            v[l] = acc;
            info->lvl_count[l] = n_l;
            info->lvl_off[l] = acc;
            acc += n_l;
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < LVL_SLOTS; l += blockDim.x) lvl_cur[l * LVL_STRIDE] = v[l];
}

// ilist: the internal nodes of each level together (level l from the exclusive scan of the
// counts). The run starting at particle p holds internal nodes S[p] + (l - lcp[p] - 1) for the
// levels lcp[p] < l <= ip[p]. A workgroup counts its range per level, reserves one range per
// level with a global atomic, then hands out slots from LDS. The order inside a level is
// arbitrary: each node's sums depend only on its own children.
__global__ __launch_bounds__(256) void bh_level_list_kernel(const int* __restrict__ lcp,
                                                            const int* __restrict__ ip,
                                                            const uint32_t* __restrict__ S,
                                                            size_t m,
                                                            uint32_t* __restrict__ lvl_cur,
                                                            uint32_t* __restrict__ ilist,
                                                            const BhInfo* __restrict__ info) {
    if (info->flags) return;  // a capturable build that did not fit: nothing to list
    __shared__ uint32_t cnt[LVL_SLOTS];
    for (int l = threadIdx.x; l < LVL_SLOTS; l += blockDim.x) cnt[l] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    size_t c0, c1;
    chunk_range(m, c0, c1);
    auto run = [&](size_t p, int& lo, int& hi, uint32_t& base) {
        lo = MAX_LEVEL + 1;
        hi = -1;
        base = 0;
        if (p < m) {
            const int lp = lcp[p];
            if (lp != LCP_SAME) {
                lo = lp + 1;
                hi = min(ip[p], MAX_LEVEL);
                base = S[p];
            }
        }
    };
    auto wave_range = [](int lo, int hi, int& wlo, int& whi) {
        wlo = lo;
        whi = hi;
        for (int off = 32; off > 0; off >>= 1) {
            wlo = min(wlo, __shfl_xor(wlo, off, 64));
            whi = max(whi, __shfl_xor(whi, off, 64));
        }
    };
    for (size_t ch = c0; ch < c1; ++ch) {  // pass 1: counts
        int lo, hi, wlo, whi;
        uint32_t base;
        run(ch * BH_CHUNK + threadIdx.x, lo, hi, base);
        wave_range(lo, hi, wlo, whi);
        for (int l = wlo; l <= whi; ++l) {
            const uint64_t in = __ballot(lo <= l && l <= hi);
            if (lane == 0 && in) atomicAdd(&cnt[l], uint32_t(__popcll(in)));
        }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < LVL_SLOTS; l += blockDim.x)
        if (cnt[l]) cnt[l] = atomicAdd(&lvl_cur[l * LVL_STRIDE], cnt[l]);  // now: next free slot
    __syncthreads();
    const uint64_t below = (uint64_t(1) << lane) - 1;
    for (size_t ch = c0; ch < c1; ++ch) {  // pass 2: entries
        int lo, hi, wlo, whi;
        uint32_t base;
        run(ch * BH_CHUNK + threadIdx.x, lo, hi, base);
        wave_range(lo, hi, wlo, whi);
        for (int l = wlo; l <= whi; ++l) {
            const bool mine = lo <= l && l <= hi;
            const uint64_t in = __ballot(mine);
            if (!in) continue;
            uint32_t slot = 0;
            if (lane == 0) slot = atomicAdd(&cnt[l], uint32_t(__popcll(in)));
            slot = __shfl(slot, 0, 64);
            if (mine) ilist[slot + uint32_t(__popcll(in & below))] = base + uint32_t(l - lo);
        }
    }
}

__device__ __forceinline__ bool same_prefix(uint64_t a, uint64_t b, int level) {
    return level == 0 || (a >> (63 - 3 * level)) == (b >> (63 - 3 * level));
}

// One thread per particle; run starts (lcp != LCP_SAME) write the nodes that start there.
// S is the exclusive scan of the per-particle node counts (m + 1 entries).
__global__ __launch_bounds__(256) void bh_node_kernel(
    const uint64_t* __restrict__ keys, const float4* __restrict__ sp, const int* __restrict__ lcp,
    const int* __restrict__ ip, const uint32_t* __restrict__ S, size_t m,
    const BhInfo* __restrict__ info, NodeRec* __restrict__ rec, NodeSums* __restrict__ sums,
    uint8_t* __restrict__ level, float* __restrict__ width,
    NodeMoments* __restrict__ mom, NodeQuad* __restrict__ quad, uint32_t* __restrict__ skipv) {
    const size_t p = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= m || info->flags) return;  // flags: a capturable build that did not fit
    const int lp = lcp[p];
    if (lp == LCP_SAME) return;
    const int I = ip[p];
    const uint32_t base = S[p];
    const float4 q = sp[p];

    // widths along the x-axis descent (levels 0 .. I+1), written for the levels starting here
    float mn = info->mn[0], mx = info->mx[0];
    for (int l = 0; l <= I + 1; ++l) {
        if (l > 0) {
            const float c = midpoint(mn, mx);
            if (q.x < c) mx = c; else mn = c;
        }
        if (l > lp) {
            const uint32_t k = base + uint32_t(l - lp - 1);
            const float wd = sub_rn(mx, mn);
            width[k] = wd;
            level[k] = uint8_t(l);
            rec[k].w2 = (l <= I) ? mul_rn(wd, wd) : -INFINITY;
            rec[k].begin = uint32_t(p);
        }
    }
    // subtree ends, deepest level first (ranges only grow towards the root)
    const uint64_t kp = keys[p];
    size_t e = p + 1;
    for (int l = I + 1; l > lp; --l) {
        if (l > KEY_LEVELS) {
            while (e < m && lcp[e] >= l) ++e;
        } else if (e < m && same_prefix(keys[e], kp, l)) {
            // gallop to a bracket [lo, hi) with prefix(lo) equal and prefix(hi) different or m
            size_t lo = e, step = 1, hi = e + 1;
            while (hi < m && same_prefix(keys[hi], kp, l)) {
                lo = hi;
                step <<= 1;
                hi = lo + step;
            }
            if (hi > m) hi = m;
            while (hi - lo > 1) {
                const size_t mid = lo + (hi - lo) / 2;
                if (same_prefix(keys[mid], kp, l)) lo = mid; else hi = mid;
            }
            e = hi;
        }
        const uint32_t k = base + uint32_t(l - lp - 1);
        rec[k].end = uint32_t(e);
        rec[k].skip = S[e];
        skipv[k] = S[e];
        if (l == I + 1) {  // the leaf: a run of identical positions
            double tot = 0.0;
            for (size_t r = p; r < e; ++r) tot += double(sp[r].w);
            const double cnt_d = double(e - p);
            sums[k].m = make_double4(tot, tot * q.x, tot * q.y, tot * q.z);
            sums[k].p = make_double4(cnt_d * q.x, cnt_d * q.y, cnt_d * q.z, cnt_d);
            if (mom)  // one position: no spread about the centre (its moments are never read)
                quad[k] = NodeQuad{make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
            rec[k].cx = q.x;
            rec[k].cy = q.y;
            rec[k].cz = q.z;
            rec[k].mass = float(tot);
        }
    }
}

// FP64 sums of the internal nodes of one level from their children (pre-order: first child at
// k+1, next sibling at the previous child's skip), then the f32 centre of mass
// (impls/mod.rs:126-131: mass-weighted mean, or the plain mean when the total mass is 0).
// With `mom`, also the second moments by the parallel-axis rule
// T = sum over children of (T_c + M_c (c_c - c)(c_c - c)^T), and the normalised quadrupole.
// One internal node k: its sums from its children's (and, with moments, its second and third
// moments, normalised quadrupole and octupole), its record's centre of mass and mass.
__device__ __forceinline__ void up_node(uint32_t k, NodeRec* __restrict__ rec,
                                        NodeSums* __restrict__ sums,
                                        NodeMoments* __restrict__ mom,
                                        NodeQuad* __restrict__ quad,
                                        NodeDesc* __restrict__ desc,
                                        const float* __restrict__ width,
                                        const uint32_t* __restrict__ skipv,
                                        const float4* __restrict__ sp) {
    {
        const NodeRec r = rec[k];
        double4 a = make_double4(0, 0, 0, 0);
        // with moments, in the same pass over the children: the second and third moments about
        // o (the first massive child's centre of mass) by the parallel-axis rule, with
        // u_c = c_c - o, S = sum T_c + M_c u_c u_c and U = sum U_c + sym(u_c, T_c) + M_c u_c u_c u_c,
        // and the first moments D = sum M_c u_c; all shifted to the centre of mass below. A leaf
        // child (skip = c + 1: one position) has no moments of its own and none is read.
        NodeMoments so;
        double ox = 0, oy = 0, oz = 0, dx_ = 0, dy_ = 0, dz_ = 0, mpos = 0;
        bool have_o = false;
        size_t c = k + 1;
        while (c < r.skip) {
            const double4 s1 = sums[c].m;
            a.x += s1.x; a.y += s1.y; a.z += s1.z; a.w += s1.w;
            const uint32_t nx = skipv[c];
            if (mom) {
                if (s1.x > 0.0) {
                    const NodeMoments tc = nx == c + 1 ? NodeMoments{} : mom[c];
                    const double im = 1.0 / s1.x;
                    const double cx = s1.y * im, cy = s1.z * im, cz = s1.w * im;
                    if (!have_o) {
                        ox = cx; oy = cy; oz = cz;
                        have_o = true;
                    }
                    const double ux = cx - ox, uy = cy - oy, uz = cz - oz;
                    const double M = s1.x;
                    so.xx += tc.xx + M * ux * ux;
                    so.yy += tc.yy + M * uy * uy;
                    so.zz += tc.zz + M * uz * uz;
                    so.xy += tc.xy + M * ux * uy;
                    so.xz += tc.xz + M * ux * uz;
                    so.yz += tc.yz + M * uy * uz;
                    so.u[0] += tc.u[0] + 3.0 * ux * tc.xx + M * ux * ux * ux;
                    so.u[1] += tc.u[1] + 3.0 * uy * tc.yy + M * uy * uy * uy;
                    so.u[2] += tc.u[2] + 3.0 * uz * tc.zz + M * uz * uz * uz;
                    so.u[3] += tc.u[3] + 2.0 * ux * tc.xy + uy * tc.xx + M * ux * ux * uy;
                    so.u[4] += tc.u[4] + 2.0 * ux * tc.xz + uz * tc.xx + M * ux * ux * uz;
                    so.u[5] += tc.u[5] + ux * tc.yy + 2.0 * uy * tc.xy + M * ux * uy * uy;
                    so.u[6] += tc.u[6] + 2.0 * uy * tc.yz + uz * tc.yy + M * uy * uy * uz;
                    so.u[7] += tc.u[7] + ux * tc.zz + 2.0 * uz * tc.xz + M * ux * uz * uz;
                    so.u[8] += tc.u[8] + uy * tc.zz + 2.0 * uz * tc.yz + M * uy * uz * uz;
                    so.u[9] += tc.u[9] + ux * tc.yz + uy * tc.xz + uz * tc.xy + M * ux * uy * uz;
                    dx_ += M * ux; dy_ += M * uy; dz_ += M * uz;
                    mpos += M;
                }
            }
            c = nx > c ? nx : c + 1;
        }
        sums[k].m = a;
        float cx, cy, cz;
        if (a.x != 0.0) {
            cx = float(a.y / a.x); cy = float(a.z / a.x); cz = float(a.w / a.x);
        } else {
            // Massless: the plain mean of the particles' positions (impls/mod.rs:127-128). Only
            // nodes of zero mass keep count-weighted position sums in `p` (leaves always do), so
            // with every child massless they add up; otherwise (masses of both signs summing to
            // zero) the node's own particles are summed.
            double4 b = make_double4(0, 0, 0, 0);
            bool massless = true;
            for (size_t q = k + 1; q < r.skip;) {
                const NodeSums sc = sums[q];
                if (sc.m.x != 0.0) {
                    massless = false;
                    break;
                }
                b.x += sc.p.x; b.y += sc.p.y; b.z += sc.p.z; b.w += sc.p.w;
                const uint32_t nx = skipv[q];
                q = nx > q ? nx : q + 1;
            }
            if (!massless) {
                b = make_double4(0, 0, 0, 0);
                for (uint32_t q = r.begin; q < r.end; ++q) {
                    const float4 v = sp[q];
                    b.x += v.x; b.y += v.y; b.z += v.z; b.w += 1.0;
                }
            }
            sums[k].p = b;
            cx = float(b.x / b.w); cy = float(b.y / b.w); cz = float(b.z / b.w);
        }
        rec[k].cx = cx;
        rec[k].cy = cy;
        rec[k].cz = cz;
        rec[k].mass = float(a.x);
        if (mom) {
            NodeMoments t;
            NodeQuad qd{make_float4(0.f, 0.f, 0.f, 0.f), make_float4(0.f, 0.f, 0.f, 0.f)};
            NodeDesc ds{{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, 0.f};
            if (a.x > 0.0 && width[k] > 0.0f) {
                // T = sum over massive children of T_c + M_c (c_c - p)(c_c - p)^T with p the
                // centre of mass: with u_c = c_c - o and v = p - o,
                // T = S_o - v D^T - D v^T + M v v^T (M, D over the same children)
                const double vx = a.y / a.x - ox, vy = a.z / a.x - oy, vz = a.w / a.x - oz;
                t.xx = so.xx - 2.0 * vx * dx_ + mpos * vx * vx;
                t.yy = so.yy - 2.0 * vy * dy_ + mpos * vy * vy;
                t.zz = so.zz - 2.0 * vz * dz_ + mpos * vz * vz;
                t.xy = so.xy - vx * dy_ - dx_ * vy + mpos * vx * vy;
                t.xz = so.xz - vx * dz_ - dx_ * vz + mpos * vx * vz;
                t.yz = so.yz - vy * dz_ - dy_ * vz + mpos * vy * vz;
                const double w = width[k];
                const double f = 1.0 / (a.x * w * w), tr = t.xx + t.yy + t.zz;
                qd.a = make_float4(float((3.0 * t.xx - tr) * f), float((3.0 * t.yy - tr) * f),
                                   float((3.0 * t.zz - tr) * f), float(3.0 * t.xy * f));
                qd.b = make_float4(float(3.0 * t.xz * f), float(3.0 * t.yz * f), 0.f, 0.f);
            }
            if (a.x > 0.0 && width[k] > 0.0f) {
                // third moments about the centre of mass p = o + v:
                // U_ijk - (v_i S_jk + v_j S_ik + v_k S_ij) + (v_i v_j D_k + v_i v_k D_j + v_j v_k D_i)
                //   - M v_i v_j v_k, then the normalised octupole from their traceless part
                const double v[3] = {a.y / a.x - ox, a.z / a.x - oy, a.w / a.x - oz};
                const double S[3][3] = {{so.xx, so.xy, so.xz}, {so.xy, so.yy, so.yz},
                                        {so.xz, so.yz, so.zz}};
                const double D[3] = {dx_, dy_, dz_};
                // index triples of {xxx, yyy, zzz, xxy, xxz, xyy, yyz, xzz, yzz, xyz}
                constexpr int I3[10][3] = {{0, 0, 0}, {1, 1, 1}, {2, 2, 2}, {0, 0, 1}, {0, 0, 2},
                                           {0, 1, 1}, {1, 1, 2}, {0, 2, 2}, {1, 2, 2}, {0, 1, 2}};
#pragma unroll
                for (int q = 0; q < 10; ++q) {
                    const int i0 = I3[q][0], i1 = I3[q][1], i2 = I3[q][2];
                    t.u[q] = so.u[q] - (v[i0] * S[i1][i2] + v[i1] * S[i0][i2] + v[i2] * S[i0][i1]) +
                             (v[i0] * v[i1] * D[i2] + v[i0] * v[i2] * D[i1] +
                              v[i1] * v[i2] * D[i0]) -
                             mpos * v[i0] * v[i1] * v[i2];
                }
                // traceless part: U~_ijk = U_ijk - (d_ij V_k + d_ik V_j + d_jk V_i) / 5, V_k = U_llk
                const double* u = t.u;
                const double vx = u[0] + u[5] + u[7], vy = u[3] + u[1] + u[8],
                             vz = u[4] + u[6] + u[2];
                const double w = width[k];
                const double f = 2.5 / (a.x * w * w * w);
                ds.oct[0] = float((u[0] - 0.6 * vx) * f);
                ds.oct[1] = float((u[1] - 0.6 * vy) * f);
                ds.oct[2] = float((u[2] - 0.6 * vz) * f);
                ds.oct[3] = float((u[3] - 0.2 * vy) * f);
                ds.oct[4] = float((u[4] - 0.2 * vz) * f);
                ds.oct[5] = float((u[5] - 0.2 * vx) * f);
                ds.oct[6] = float(u[9] * f);
            }
            mom[k] = t;
            quad[k] = qd;
            desc[k] = ds;
        }
    }
}

template <bool MOM>
__global__ __launch_bounds__(256) void bh_up_kernel(NodeRec* __restrict__ rec,
                                                    NodeSums* __restrict__ sums,
                                                    const uint32_t* __restrict__ nodes_at_level,
                                                    uint32_t count, NodeMoments* __restrict__ mom,
                                                    NodeQuad* __restrict__ quad,
                                                    NodeDesc* __restrict__ desc,
                                                    const float* __restrict__ width,
                                                    const uint32_t* __restrict__ skipv,
                                                    const float4* __restrict__ sp) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += gridDim.x * blockDim.x)
        up_node(nodes_at_level[t], rec, sums, MOM ? mom : nullptr, quad, desc, width, skipv, sp);
}

// Consecutive small levels (deepest first) in one workgroup: a barrier between levels makes one
// level's sums visible to the next (the waves of a workgroup share one CU and its cache), so a
// run of small levels costs one launch instead of one per level.
constexpr int UP_RUN_MAX = 32;
constexpr uint32_t UP_RUN_NODES = 1024;  // a level joins a run when it has at most this many
struct UpRun {
    uint32_t off[UP_RUN_MAX], cnt[UP_RUN_MAX];
    int levels;
};

template <bool MOM>
__global__ __launch_bounds__(1024) void bh_up_run_kernel(NodeRec* __restrict__ rec,
                                                         NodeSums* __restrict__ sums,
                                                         const uint32_t* __restrict__ ilist,
                                                         UpRun run, NodeMoments* __restrict__ mom,
                                                         NodeQuad* __restrict__ quad,
                                                         NodeDesc* __restrict__ desc,
                                                         const float* __restrict__ width,
                                                         const uint32_t* __restrict__ skipv,
                                                         const float4* __restrict__ sp) {
    for (int j = 0; j < run.levels; ++j) {
        for (uint32_t t = threadIdx.x; t < run.cnt[j]; t += blockDim.x)
            up_node(ilist[run.off[j] + t], rec, sums, MOM ? mom : nullptr, quad, desc, width, skipv,
                    sp);
        __syncthreads();
    }
}

// ---- capturable builds: node count and status stay on the device ----------------------------
// The scan's total against the capacities sized by an earlier ordinary build; a tree that does
// not fit flags the build (every later kernel of it then returns at once and the walk writes NaN).
__global__ void bh_dd_check_kernel(const uint32_t* __restrict__ S, size_t m, uint32_t cap,
                                   int levels, BhInfo* __restrict__ info) {
    const uint32_t total = S[m];
    info->total = total;
    if (total > cap || info->max_level > levels) atomicOr(&info->flags, FLAG_CAPACITY);
}

// The bottom-up sums of one level, its count and place in the level-grouped list read from the
// device (the host launches every level up to the sized depth; levels the tree lacks are empty).
// With `mom` (the learned correction), also the node moments, as in the ordinary build.
template <bool MOM>
__global__ __launch_bounds__(256) void bh_up_level_dd_kernel(NodeRec* __restrict__ rec,
                                                             NodeSums* __restrict__ sums,
                                                             const uint32_t* __restrict__ ilist,
                                                             const BhInfo* __restrict__ info,
                                                             int level,
                                                             const uint32_t* __restrict__ skipv,
                                                             const float4* __restrict__ sp,
                                                             NodeMoments* __restrict__ mom,
                                                             NodeQuad* __restrict__ quad,
                                                             NodeDesc* __restrict__ desc,
                                                             const float* __restrict__ width) {
    if (info->flags) return;
    const uint32_t cnt = info->lvl_count[level], off = info->lvl_off[level];
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < cnt; t += gridDim.x * blockDim.x)
        up_node(ilist[off + t], rec, sums, MOM ? mom : nullptr, quad, desc, width, skipv, sp);
}

// A run of levels hi, hi - 1, ..., lo in one workgroup, a barrier between levels (levels the tree
// lacks are empty).
template <bool MOM>
__global__ __launch_bounds__(1024) void bh_up_levels_dd_kernel(NodeRec* __restrict__ rec,
                                                               NodeSums* __restrict__ sums,
                                                               const uint32_t* __restrict__ ilist,
                                                               const BhInfo* __restrict__ info,
                                                               int hi, int lo,
                                                               const uint32_t* __restrict__ skipv,
                                                               const float4* __restrict__ sp,
                                                               NodeMoments* __restrict__ mom,
                                                               NodeQuad* __restrict__ quad,
                                                               NodeDesc* __restrict__ desc,
                                                               const float* __restrict__ width) {
    if (info->flags) return;  // uniform: every thread returns, or none
    for (int L = hi; L >= lo; --L) {
        const uint32_t cnt = info->lvl_count[L], off = info->lvl_off[L];
        for (uint32_t t = threadIdx.x; t < cnt; t += blockDim.x)
            up_node(ilist[off + t], rec, sums, MOM ? mom : nullptr, quad, desc, width, skipv,
                    sp);
        __syncthreads();
    }
}

// Reference-exact node data (PNB_BH_REF_COM): the f32 folds of impls/mod.rs:126-131 over each
// node's slice in the order the reference holds it. The reference fills a child's slice by pushing
// the parent's bodies in order (tree/mod.rs:116-125), so every node folds its bodies in input
// order. Tree order agrees with that inside a leaf (its bodies share one key, and the sort keeps
// equal keys in input order) but not across the children of an internal node. Leaves therefore
// fold tree order; the internal nodes of level L fold `ord`, the bodies regrouped by the level-L
// node that holds them and sorted by input index inside each group. One thread per node: the
// folds are sequential by definition, so a node of k bodies costs k dependent steps on one lane.
__device__ __forceinline__ void ref_com_fold(const float4* __restrict__ src, uint32_t b,
                                             uint32_t e, NodeRec* __restrict__ r) {
    float tot = 0.0f;
    for (uint32_t q = b; q < e; ++q) tot = add_rn(tot, src[q].w);
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    if (tot == 0.0f) {
        for (uint32_t q = b; q < e; ++q) {
            const float4 v = src[q];
            cx = add_rn(cx, v.x); cy = add_rn(cy, v.y); cz = add_rn(cz, v.z);
        }
        const float len = float(e - b);
        cx = __fdiv_rn(cx, len); cy = __fdiv_rn(cy, len); cz = __fdiv_rn(cz, len);
    } else {
        for (uint32_t q = b; q < e; ++q) {
            const float4 v = src[q];
            const float f = __fdiv_rn(v.w, tot);
            cx = add_rn(cx, mul_rn(v.x, f));
            cy = add_rn(cy, mul_rn(v.y, f));
            cz = add_rn(cz, mul_rn(v.z, f));
        }
    }
    r->cx = cx;
    r->cy = cy;
    r->cz = cz;
    r->mass = tot;
}

__global__ __launch_bounds__(256) void bh_ref_com_leaf_kernel(NodeRec* __restrict__ rec,
                                                              const float4* __restrict__ sp,
                                                              size_t total) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= total || __float_as_int(rec[k].w2) >= 0) return;  // internal nodes: by level
    ref_com_fold(sp, rec[k].begin, rec[k].end, &rec[k]);
}

// Per body (tree position q): cur[q] moves to the node of level L that holds q (the child of
// cur[q] whose range covers q, when cur[q] is internal; a leaf above level L keeps it), and the
// sort key is (that node's first tree position, input index). Every body of an internal node of
// level L then shares its group; groups tile [0, m) in tree order, so after the sort each such
// node's bodies occupy its own range [begin, end), in input order.
__global__ __launch_bounds__(256) void bh_ref_com_group_kernel(const NodeRec* __restrict__ rec,
                                                               const uint32_t* __restrict__ skipv,
                                                               const uint32_t* __restrict__ sidx,
                                                               uint32_t m, int descend,
                                                               uint32_t* __restrict__ cur,
                                                               uint64_t* __restrict__ keys,
                                                               uint32_t* __restrict__ vals) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= m) return;
    uint32_t k = descend ? cur[q] : 0u;
    if (descend && __float_as_int(rec[k].w2) >= 0) {
        const uint32_t last = skipv[k];  // children are k + 1, skip[k + 1], ... before skip[k]
        uint32_t ch = k + 1;
        while (rec[ch].end <= q) {
            const uint32_t nx = skipv[ch];
            if (nx >= last) break;
            ch = nx;
        }
        k = ch;
    }
    cur[q] = k;
    keys[q] = (uint64_t(rec[k].begin) << 32) | sidx[q];
    vals[q] = q;
}

__global__ __launch_bounds__(256) void bh_gather_keyed_kernel(const float4* __restrict__ afg,
                                                              const uint64_t* __restrict__ keys,
                                                              uint32_t m,
                                                              float4* __restrict__ ord) {
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < m) ord[q] = afg[uint32_t(keys[q])];
}

__global__ __launch_bounds__(256) void bh_ref_com_internal_kernel(NodeRec* __restrict__ rec,
                                                                  const uint32_t* __restrict__ list,
                                                                  uint32_t n,
                                                                  const float4* __restrict__ ord) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = list[i];
    ref_com_fold(ord, rec[k].begin, rec[k].end, &rec[k]);
}

// ---- learned correction: node features and the node network ------------------------------------
// Features of an internal node (the network's input; zero for leaves):
//   f0..5  normalised quadrupole Q~ = (3T - tr(T) I) / (M w^2), {xx, yy, zz, xy, xz, yz}
//   f6     1 + log2(node mass / root mass) / 32
//   f7     log2(bodies in the node) / 32
//   f8     depth / 64
//   f9..15 normalised octupole (5/2) U~ / (M w^3) (NodeDesc::oct) times 32 (its entries are
//          about 0.01, the other features' about 0.1 to 1)
// The network's outputs are two correction fields: y[0..5] = f0..5 + MLP(f)[0..5] is an
// effective second moment in units of M w^2 (a residual on the quadrupole) and y[6..12] =
// MLP(f)[6..12] an effective traceless third moment in units of M w^3, learned from zero; every
// lane that accepts the node adds quad_field(M w^2 y[0..5]) + oct_field(M w^3 y[6..12]) at its
// own position. With a zero output layer (a fresh network) this is the quadrupole term of the
// node's multipole expansion; the octupole term (y[6..12] = f9..15 / 32) and whatever else lowers
// the brute-force residual is what training adds.

__device__ __forceinline__ void node_features(const NodeRec& r, const NodeQuad& q, uint8_t level,
                                              const NodeDesc& dsc, float root_mass,
                                              float (&f)[MLP_KIN]) {
#pragma unroll
    for (int i = 0; i < MLP_KIN; ++i) f[i] = 0.f;
    if (!(r.w2 >= 0.0f) || !(r.mass > 0.0f)) return;  // leaf or massless: no correction
    f[0] = q.a.x;
    f[1] = q.a.y;
    f[2] = q.a.z;
    f[3] = q.a.w;
    f[4] = q.b.x;
    f[5] = q.b.y;
    f[6] = 1.f + __builtin_amdgcn_logf(fmaxf(r.mass * __builtin_amdgcn_rcpf(root_mass), 1e-30f)) *
                     (1.f / 32.f);
    f[7] = __builtin_amdgcn_logf(float(r.end - r.begin)) * (1.f / 32.f);
    f[8] = float(level) * (1.f / 64.f);
#pragma unroll
    for (int i = 0; i < 7; ++i) f[9 + i] = OCT_FEATURE_SCALE * dsc.oct[i];
}

// The learned moments of a node, read with scalar loads when a lane accepts it: a QuadH entry
// when the network has only quadrupole outputs, an OctH entry (which holds the quadrupole too)
// when it has octupole outputs.
// QuadH (16 bytes): the quadrupole y[0..5] (units of M w^2) as halves. y is O(1) and the
// correction it feeds is itself a small fraction of the force, so 11-bit mantissas cost nothing
// measurable, while the walk's scalar loads take half of what floats would.
// OctH (64 bytes, one scalar load): the octupole y[6..12] (units of M w^3) times the node's width
// w, in FP32, laid out for F_i = O_ijk d_j d_k written with the trace removed,
// O_izz = -(O_ixx + O_iyy):
//   F_x = O_xxx a + O_xyy b + O_xxy (2xy) + 2 O_xxz xz + 2 O_xyz yz,
//   F_y = O_xxy a + O_yyy b + O_xyy (2xy) + 2 O_xyz xz + 2 O_yyz yz,
//   F_z = O_xxz a + O_yyz b + 2 O_xyz xy + 2 O_xzz xz + 2 O_yzz yz,  a = xx - zz, b = yy - zz,
// o = {xxx, xyy, xxy, yyy, xxz, yyz, 2xxz, 2xyz, 2yyz, 2xzz, 2yzz}: one FMA per coefficient and
// product (a vector instruction reads one scalar operand here), fifteen for F, and no square
// root for w^3; then the quadrupole as in QuadH and a flag, non-zero when the octupole is worth
// adding: its largest coefficient is at least OCT_MIN, which leaves out nodes whose octupole
// vanishes (two equal masses) or nearly does, about a third of the accepted ones on Plummer
// spheres, at no measurable cost in accuracy.
struct __align__(16) QuadH {
    __half q[8];  // y[0..5], 0, 0
};
struct __align__(64) OctH {
    float o[11];  // w * octupole, the layout above
    __half q[6];  // y[0..5]
    uint32_t flag, pad;
};
static_assert(sizeof(OctH) == 64, "octupole entry is 64 bytes");
constexpr float OCT_MIN = 3e-3f;

// qh[k] = MLP(features of node k) for the internal nodes listed in inodes (zero for massless
// ones; the walk reads qh only for internal nodes): 16 nodes per MFMA tile, one tile per
// wavefront and iteration, weights staged in LDS. dd_info (a capturable build): the count is the
// build's internal-node total, read on the device, and a build that did not fit writes nothing.
__global__ __launch_bounds__(256) void mlp_node_kernel(const NodeRec* __restrict__ rec,
                                                       const NodeQuad* __restrict__ quad,
                                                       const uint8_t* __restrict__ level,
                                                       const NodeDesc* __restrict__ desc,
                                                       const uint32_t* __restrict__ inodes,
                                                       uint32_t count,
                                                       const MlpWeights* __restrict__ wg,
                                                       QuadH* __restrict__ qh,
                                                       OctH* __restrict__ oh,
                                                       const BhInfo* __restrict__ dd_info) {
    if (dd_info) {
        if (dd_info->flags) return;  // uniform
        count = dd_info->lvl_off[LVL_SLOTS - 1] + dd_info->lvl_count[LVL_SLOTS - 1];
    }
    __shared__ MlpForward w;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(wg);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&w);
        for (int t = threadIdx.x; t < int(sizeof(MlpForward) / 4); t += blockDim.x) dst[t] = src[t];
        __syncthreads();
    }
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, g = lane >> 4;
    const float root_mass = rec[0].mass;
    for (uint32_t t0 = (blockIdx.x * 4 + wv) * 16; t0 < count; t0 += gridDim.x * 4 * 16) {
        const bool listed = t0 + r16 < count;
        const uint32_t k = listed ? inodes[t0 + r16] : 0u;
        float f[MLP_KIN];
        NodeRec r{};
        r.w2 = -INFINITY;
        if (listed) {
            r = rec[k];
            node_features(r, quad[k], level[k], desc[k], root_mass, f);
        } else {
#pragma unroll
            for (int q = 0; q < MLP_KIN; ++q) f[q] = 0.f;
        }
        bf16x4_t bx;  // features 4 g .. 4 g + 3 of node k (static indices, selected by g)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float v = g == 0 ? f[q] : g == 1 ? f[4 + q] : g == 2 ? f[8 + q] : f[12 + q];
            bx[q] = short(f2bf(v));
        }
        f32x4_t yv = mlp_forward16_regs(w, bx);  // lanes of group g: outputs 4 g .. 4 g + 3
        // the residual: the node's own normalised quadrupole, f0..5 in FP32
#pragma unroll
        for (int q = 0; q < 4; ++q) yv[q] += g == 0 ? f[q] : (g == 1 && q < 2) ? f[4 + q] : 0.f;
        const float keep = (r.w2 >= 0.0f && r.mass > 0.0f) ? 1.f : 0.f;
        // the octupole outputs 6..12 of node r16 gathered in its group-0 lane: 6, 7 from lane
        // r16 + 16, 8..11 from r16 + 32, 12 from r16 + 48
        float oc[7];
        oc[0] = __shfl(yv[2], r16 + 16, 64);
        oc[1] = __shfl(yv[3], r16 + 16, 64);
        oc[2] = __shfl(yv[0], r16 + 32, 64);
        oc[3] = __shfl(yv[1], r16 + 32, 64);
        oc[4] = __shfl(yv[2], r16 + 32, 64);
        oc[5] = __shfl(yv[3], r16 + 32, 64);
        oc[6] = __shfl(yv[0], r16 + 48, 64);
        const float q4 = __shfl(yv[0], r16 + 16, 64), q5 = __shfl(yv[1], r16 + 16, 64);
        if (listed && g == 0) {
            float big = 0.f;
#pragma unroll
            for (int j = 0; j < 7; ++j) {
                oc[j] *= keep;
                big = fmaxf(big, fabsf(oc[j]));
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) qh[k].q[q] = __float2half(keep * yv[q]);
            qh[k].q[6] = __float2half(0.f);
            qh[k].q[7] = __float2half(0.f);
            // w * O in the walk's layout (OctH), the dependent components from O_iik = 0
            const float wd = keep > 0.f ? sqrtf(r.w2) : 0.f;
            const float xxx = wd * oc[0], yyy = wd * oc[1], zzz = wd * oc[2], xxy = wd * oc[3],
                        xxz = wd * oc[4], xyy = wd * oc[5], xyz = wd * oc[6];
            const float xzz = -(xxx + xyy), yzz = -(xxy + yyy), yyz = -(xxz + zzz);
            OctH e;
            e.o[0] = xxx;
            e.o[1] = xyy;
            e.o[2] = xxy;
            e.o[3] = yyy;
            e.o[4] = xxz;
            e.o[5] = yyz;
            e.o[6] = 2.f * xxz;
            e.o[7] = 2.f * xyz;
            e.o[8] = 2.f * yyz;
            e.o[9] = 2.f * xzz;
            e.o[10] = 2.f * yzz;
#pragma unroll
            for (int q = 0; q < 4; ++q) e.q[q] = __float2half(keep * yv[q]);
            e.q[4] = __float2half(keep * q4);
            e.q[5] = __float2half(keep * q5);
            e.flag = big >= OCT_MIN ? 1u : 0u;
            e.pad = 0u;
            oh[k] = e;
        }
        if (listed && g == 1) {
            qh[k].q[4] = __float2half(keep * yv[0]);
            qh[k].q[5] = __float2half(keep * yv[1]);
        }
    }
}

// Every node's features (pnb_bh_node_features): MLP_FEAT floats per node, pre-order.
__global__ __launch_bounds__(256) void bh_features_kernel(const NodeRec* __restrict__ rec,
                                                          const NodeQuad* __restrict__ quad,
                                                          const uint8_t* __restrict__ level,
                                                          const NodeDesc* __restrict__ desc,
                                                          size_t total, float* __restrict__ out) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= total) return;
    float f[MLP_KIN];
    node_features(rec[k], quad[k], level[k], desc[k], rec[0].mass, f);
#pragma unroll
    for (int i = 0; i < MLP_FEAT; ++i) out[k * MLP_FEAT + i] = f[i];
}

// ---- traversal -----------------------------------------------------------------------------------

constexpr int WALK_PLAIN = 0;  // plain Barnes-Hut
constexpr int WALK_ROWS = 1;   // also list the internal nodes each wavefront accepted (rows)
constexpr int WALK_QUAD = 2;   // add quad_field(q) of every accepted internal node
constexpr int WALK_OCT = 3;    // WALK_QUAD with the network's octupole outputs (oct_field)

// Per wavefront, the internal nodes that at least one of its lanes accepted, in walk order:
// {node index, accepting-lane count, accepting-lane mask (low, high 32 bits)}.
struct WalkRows {
    uint4* rows;      // [wave][cap]
    uint32_t* count;  // [wave] rows found; only the first cap are stored
    uint32_t cap;
};

// Advances the wavefront's cursor: (k, r) = (open != 0) ? (kn, rn) : (ks, rs), with the
// operands and results in SGPRs (one scalar compare and four scalar selects). Left to itself the
// compiler copies both candidate records into VGPRs and selects them with v_cndmask.
__device__ __forceinline__ uint64_t pack2(uint32_t lo, uint32_t hi) {
    return uint64_t(lo) | (uint64_t(hi) << 32);
}

template <typename Off, bool SEL3>
__device__ __forceinline__ void advance(uint64_t open, Off kn, Off ks, Off& k, NodeRec& r,
                                        const NodeRec& rn, const NodeRec& rs) {
    const uint64_t n0 = pack2(__float_as_uint(rn.cx), __float_as_uint(rn.cy));
    const uint64_t s0 = pack2(__float_as_uint(rs.cx), __float_as_uint(rs.cy));
    const uint64_t n1 = pack2(__float_as_uint(rn.cz), __float_as_uint(rn.mass));
    const uint64_t s1 = pack2(__float_as_uint(rs.cz), __float_as_uint(rs.mass));
    const uint64_t n2 = pack2(__float_as_uint(rn.w2), rn.skip);
    const uint64_t s2 = pack2(__float_as_uint(rs.w2), rs.skip);
    // begin/end: selected only when SEL3 (the relative criterion keeps M w^3 in begin); otherwise
    // naming them makes each record one 32-byte scalar load
    const uint64_t n3 = pack2(rn.begin, rn.end), s3 = pack2(rs.begin, rs.end);
    uint64_t o0, o1, o2, o3 = 0;
    Off ok;
    if constexpr (SEL3) {
        if constexpr (sizeof(Off) == 4) {
            asm volatile(
                "s_cmp_lg_u64 %5, 0\n\t"
                "s_cselect_b64 %0, %6, %7\n\t"
                "s_cselect_b64 %1, %8, %9\n\t"
                "s_cselect_b64 %2, %10, %11\n\t"
                "s_cselect_b64 %3, %14, %15\n\t"
                "s_cselect_b32 %4, %12, %13"
                : "=&s"(o0), "=&s"(o1), "=&s"(o2), "=&s"(o3), "=&s"(ok)
                : "s"(open), "s"(n0), "s"(s0), "s"(n1), "s"(s1), "s"(n2), "s"(s2), "s"(kn),
                  "s"(ks), "s"(n3), "s"(s3)
                : "scc");
        } else {
            asm volatile(
                "s_cmp_lg_u64 %5, 0\n\t"
                "s_cselect_b64 %0, %6, %7\n\t"
                "s_cselect_b64 %1, %8, %9\n\t"
                "s_cselect_b64 %2, %10, %11\n\t"
                "s_cselect_b64 %3, %14, %15\n\t"
                "s_cselect_b64 %4, %12, %13"
                : "=&s"(o0), "=&s"(o1), "=&s"(o2), "=&s"(o3), "=&s"(ok)
                : "s"(open), "s"(n0), "s"(s0), "s"(n1), "s"(s1), "s"(n2), "s"(s2), "s"(kn),
                  "s"(ks), "s"(n3), "s"(s3)
                : "scc");
        }
    } else if constexpr (sizeof(Off) == 4) {
        asm volatile(
            "s_cmp_lg_u64 %4, 0\n\t"
            "s_cselect_b64 %0, %5, %6\n\t"
            "s_cselect_b64 %1, %7, %8\n\t"
            "s_cselect_b64 %2, %9, %10\n\t"
            "s_cselect_b32 %3, %11, %12"
            : "=&s"(o0), "=&s"(o1), "=&s"(o2), "=&s"(ok)
            : "s"(open), "s"(n0), "s"(s0), "s"(n1), "s"(s1), "s"(n2), "s"(s2), "s"(kn), "s"(ks),
              "s"(n3), "s"(s3)
            : "scc");
    } else {
        asm volatile(
            "s_cmp_lg_u64 %4, 0\n\t"
            "s_cselect_b64 %0, %5, %6\n\t"
            "s_cselect_b64 %1, %7, %8\n\t"
            "s_cselect_b64 %2, %9, %10\n\t"
            "s_cselect_b64 %3, %11, %12"
            : "=&s"(o0), "=&s"(o1), "=&s"(o2), "=&s"(ok)
            : "s"(open), "s"(n0), "s"(s0), "s"(n1), "s"(s1), "s"(n2), "s"(s2), "s"(kn), "s"(ks),
              "s"(n3), "s"(s3)
            : "scc");
    }
    r.cx = __uint_as_float(uint32_t(o0));
    r.cy = __uint_as_float(uint32_t(o0 >> 32));
    r.cz = __uint_as_float(uint32_t(o1));
    r.mass = __uint_as_float(uint32_t(o1 >> 32));
    r.w2 = __uint_as_float(uint32_t(o2));
    r.skip = uint32_t(o2 >> 32);
    if constexpr (SEL3) {
        r.begin = uint32_t(o3);
        r.end = uint32_t(o3 >> 32);
    }
    k = ok;
}

// The corrected walk's partial sums: the terms of PART_STEPS wavefront steps (at most one per
// lane and step) summed in plain FP32, then added to the compensated sum.
constexpr uint32_t PART_STEPS = 16;
__device__ __forceinline__ void fold_partial(float& ax, float& ay, float& az, float& cx, float& cy,
                                             float& cz, float& px, float& py, float& pz) {
    float y = px - cx, t = ax + y;
    cx = (t - ax) - y; ax = t;
    y = py - cy; t = ay + y;
    cy = (t - ay) - y; ay = t;
    y = pz - cz; t = az + y;
    cz = (t - az) - y; az = t;
    px = py = pz = 0.f;
}

__device__ __forceinline__ float half_lo(uint32_t v) {
    return __half2float(__ushort_as_half(uint16_t(v)));
}
__device__ __forceinline__ float half_hi(uint32_t v) {
    return __half2float(__ushort_as_half(uint16_t(v >> 16)));
}


// One wavefront walks the pre-order node records with one cursor (k, in SGPRs); lane decisions
// are the reference's (see the file comment). The walk is bound by its instruction stream as a
// whole (about 33 VALU, 15 SALU and 2 scalar loads per step for the plain walk, and the latency
// of those loads), so the step keeps everything wave-uniform in SGPRs: the records of both
// possible next nodes (k + 1, skip) are requested at the top of the step and picked with scalar
// selects at its end, and the corrected walks fetch the learned entry of node k only when some
// lane accepts it as an internal node (prefetching it for both candidates costs more
// scalar-cache misses than it hides).
// WIDE: 64-bit byte offsets, for trees whose records span 4 GiB or more (2^27 nodes).
// REL: the relative criterion. Each lane holds A = alpha |a~| (est_in[i >> est_shift], from an
// estimate pass per particle, est_shift = 0, or per wavefront, est_shift = 6) and
// the records carry K = M w^p in `begin` (0 for leaves); an internal node is opened when the
// geometric rule opens it or when the size of its first neglected term, M w^p / r^(p+2)
// (r^2 = d^2 + eps^2), exceeds A: p = 2 for the plain walk (K > A r^4, the quadrupole is
// neglected), p = 3 for WALK_QUAD (K / r > A r^4, the octupole is the first term it lacks) and
// p = 4 for WALK_OCT (K / r^2 > A r^4: the network adds octupoles, the hexadecapole
// is the first term it lacks).
// REL buckets: a record whose `end` is non-zero packs the first body (27 bits) and the count
// (5 bits) of a small internal node; lanes that open it add its bodies (`bodies`, the affecting
// particles in tree order) directly and skip its subtree.
// est_out (estimate pass): writes alpha |a| per particle, in walk order, instead of out.
// wave_order (REL, optional): the launch's k-th wavefront (k < ceil(n / 64)) walks the particles
// of wavefront wave_order[k], so that the heaviest wavefronts start first.
// lanes: only the first `lanes` lanes of each wavefront walk (64 except in the estimate pass).
// packed (not with REL): the wavefront's particles are consecutive, `lanes` of them from
// (launch wavefront index) * lanes, instead of 64 from (launch wavefront index) * 64. Narrow
// wavefronts walk smaller unions of their particles' traversals, so that a small problem runs as
// many short walks instead of a few long ones; every lane adds exactly the nodes of its own
// pre-order traversal, in that order, either way, so the plain walk's results are bit-identical;
// the corrected walk's partial sums are grouped by wavefront steps, so its last bits can differ.
// dd: a capturable build's tree (`total` is then its capacity; the count is read from `info`).
// WALK_OCT: the network has octupole outputs; every accepted internal node's octupole entry is
// fetched together with its quadrupole entry, so that the two loads overlap. A mode of its own
// rather than a run-time flag: behind a flag the compiler copied the ten octupole components
// from SGPRs into VGPRs on every accepting step (ten VALU moves).
template <int MODE, bool WIDE, bool REL>
__global__ __launch_bounds__(256) void bh_walk_kernel(const NodeRec* __restrict__ nodes,
                                                      uint32_t total,
                                                      const float4* __restrict__ pts,
                                                      const uint32_t* __restrict__ perm,
                                                      uint32_t n, float theta2, float eps2,
                                                      float4* __restrict__ out,
                                                      BhInfo* __restrict__ info, WalkRows rows,
                                                      const QuadH* __restrict__ qh,
                                                      const OctH* __restrict__ oh,
                                                      const float* __restrict__ est_in,
                                                      uint32_t est_shift,
                                                      float* __restrict__ est_out, float alpha,
                                                      const uint32_t* __restrict__ wave_order,
                                                      const float4* __restrict__ bodies,
                                                      uint32_t lanes, int dd, int packed) {
    constexpr bool QUAD = MODE == WALK_QUAD || MODE == WALK_OCT;
    constexpr bool OCT = MODE == WALK_OCT;
    // the geometric corrected walk is bound by its arithmetic: its terms go into plain partial
    // sums (no per-term compensation); the relative one keeps the per-term form (with partial
    // sums the compiler moved its cursor mask out of the scalar registers)
    constexpr bool PARTIAL = QUAD && !REL;
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (packed) i = (i >> 6) * lanes + (threadIdx.x & 63);
    if (dd) {
        // capturable build: the node count and the build's status are on the device; a build
        // that did not fit gives NaN (pnb_bh_check reports why)
        if (info->flags) {
            if (i < n && uint32_t(threadIdx.x & 63) < lanes)
                out[perm ? perm[i] : i] = make_float4(NAN, NAN, NAN, 0.f);
            return;
        }
        total = info->total;
    }
    if (REL && wave_order) {
        const uint32_t k = i >> 6;  // this wavefront's place in the launch
        if (k < ((n + 63) >> 6)) i = (wave_order[k] << 6) | (i & 63);
    }
    const int lane = threadIdx.x & 63;
    const bool valid = i < n && uint32_t(lane) < lanes;
    const float4 p = pts[valid ? i : n - 1];
    const float A = REL ? est_in[(valid ? i : n - 1) >> est_shift] : 0.f;
    float ax = 0.f, ay = 0.f, az = 0.f, cx = 0.f, cy = 0.f, cz = 0.f;
    float px = 0.f, py = 0.f, pz = 0.f;  // PARTIAL: the partial sums
    uint32_t acc_wave = 0;               // QUAD: accepted terms of the wavefront
    using Off = std::conditional_t<WIDE, uint64_t, uint32_t>;
    Off resume = valid ? Off(0) : ~Off(0);
    uint32_t accepted = 0, steps = 0, nrows = 0;
    const uint32_t wave_slot = i / 64;

    // The cursor is a byte offset, ko = 32 k, so each record address is the array base plus one
    // scalar offset (32 bits unless WIDE: the host takes WIDE when 32 (total + 1) >= 2^32).
    const char* base = reinterpret_cast<const char*>(nodes);
    const Off end = Off(total) * Off(sizeof(NodeRec));
    Off ko = 0;
    NodeRec r{};
    if (total > 0) r = nodes[0];  // no tree (m = 0): nodes may not even be allocated
    while (ko < end) {
        // in bounds: k + 1 <= total and skip <= total, and the array holds total + 1 records
        const Off so = Off(r.skip) * Off(sizeof(NodeRec));
        const NodeRec r_next = *reinterpret_cast<const NodeRec*>(base + (ko + Off(32)));
        const NodeRec r_skip = *reinterpret_cast<const NodeRec*>(base + so);
        const bool active = ko >= resume;
        const float dx = sub_rn(r.cx, p.x), dy = sub_rn(r.cy, p.y), dz = sub_rn(r.cz, p.z);
        const float d2 = norm2_ref(dx, dy, dz);
        const bool nz = d2 != 0.0f;
        // REL with QUAD: the reciprocal distance is needed by every lane's criterion, not
        // only by the accepting lanes' pair terms
        constexpr bool REL_INV = REL && QUAD;
        const float r2_rel = REL ? add_rn(d2, eps2) : 0.f;
        const float inv_rel = REL_INV ? __builtin_amdgcn_rsqf(r2_rel) : 0.f;
        const float K = __uint_as_float(r.begin);
        // QUAD: K / r (p = 3), or K / r^2 (p = 4, when the network adds octupoles)
        const float lhs = REL ? (REL_INV ? mul_rn(OCT ? mul_rn(K, inv_rel) : K, inv_rel) : K)
                              : 0.f;
        const float rhs = REL ? mul_rn(mul_rn(A, r2_rel), r2_rel) : 0.f;
        const bool crit = mul_rn(theta2, d2) < r.w2 || (REL && rhs < lhs);
        bool open = active && nz && crit;
        const bool acc = active && nz && !crit;
        // the wavefront's open mask from the compares' own lane masks (a ballot of the combined
        // bool would first copy it into a VGPR)
        const uint64_t live_mask = __builtin_amdgcn_ballot_w64(active) &
                                   __builtin_amdgcn_ballot_w64(nz);
        uint64_t crit_mask = __builtin_amdgcn_fcmpf(mul_rn(theta2, d2), r.w2, 4 /*olt*/);
        if constexpr (REL) crit_mask |= __builtin_amdgcn_fcmpf(rhs, lhs, 4 /*olt*/);
        uint64_t open_mask = live_mask & crit_mask;
        const uint64_t acc_mask = live_mask & ~crit_mask;
        if constexpr (REL) {
            const uint32_t bucket = r.end;
            if (bucket != 0 && open_mask != 0) {  // a small node that some lanes open
                if (open) {
                    const uint32_t b0 = bucket & 0x7ffffffu, b1 = b0 + (bucket >> 27);
                    for (uint32_t q = b0; q < b1; ++q) {
                        const float4 b = bodies[q];  // wave-uniform: a scalar load
                        const float ex = sub_rn(b.x, p.x), ey = sub_rn(b.y, p.y),
                                    ez = sub_rn(b.z, p.z);
                        const float e2 = norm2_ref(ex, ey, ez);
                        if (e2 != 0.0f) {
                            const float iv = __builtin_amdgcn_rsqf(add_rn(e2, eps2));
                            const float sb = (b.w * iv) * (iv * iv);
                            const float bx = fmaf(ex, sb, -cx), by = fmaf(ey, sb, -cy),
                                        bz = fmaf(ez, sb, -cz);
                            float t = ax + bx;
                            cx = (t - ax) - bx; ax = t;
                            t = ay + by;
                            cy = (t - ay) - by; ay = t;
                            t = az + bz;
                            cz = (t - az) - bz; az = t;
                            ++accepted;
                        }
                    }
                }
                open = false;  // nobody descends: every lane continues after the subtree
                open_mask = 0;
            }
        }
        if (acc) {
            const float r2 = REL ? r2_rel : add_rn(d2, eps2);
            const float inv = REL_INV ? inv_rel : __builtin_amdgcn_rsqf(r2);
            const float inv2 = inv * inv;
            const float s = (r.mass * inv) * inv2;
            if constexpr (QUAD) {
                // the terms are chained onto v by FMAs: the vector terms first, in place, and the
                // monopole's direction term, shared by leaves and internal nodes, last. v is the
                // partial sum (PARTIAL) or the new term less the compensation, added at once.
                float vx = PARTIAL ? px : -cx, vy = PARTIAL ? py : -cy, vz = PARTIAL ? pz : -cz;
                float st = s;
                if (__float_as_int(r.w2) >= 0) {  // internal (leaves: -inf)
                    // internal node: the monopole plus the field of Q = M w^2 y[0..5]
                    // (quad_field's formula, w2 = width^2) and, when the node has one, of
                    // O = M w^3 y[6..12] (oct_field's): with u = |d|^-2 and s5 = M w^2 |d|^-5,
                    // d (s + s5 u (5/2 dQd - 7 u G)) + s5 (3 u F - Q d), G = d.F, F = O d d
                    // the quadrupole's halves: QuadH, or (OCT) inside the node's one OctH entry
                    uint32_t h0, h1, h2;
                    OctH oe;
                    if constexpr (OCT) {
                        oe = *reinterpret_cast<const OctH*>(reinterpret_cast<const char*>(oh) +
                                                            (ko << 1));  // 64-byte entries
                        const uint32_t* hq = reinterpret_cast<const uint32_t*>(oe.q);
                        h0 = hq[0];
                        h1 = hq[1];
                        h2 = hq[2];
                        // an empty asm that reads the whole entry here: one load and one wait
                        // (otherwise the compiler fetches the octupole behind the flag's test)
                        asm volatile("" ::"s"(oe.o[0]), "s"(oe.o[1]), "s"(oe.o[2]),
                                     "s"(oe.o[3]), "s"(oe.o[4]), "s"(oe.o[5]), "s"(oe.o[6]),
                                     "s"(oe.o[7]), "s"(oe.o[8]), "s"(oe.o[9]), "s"(oe.o[10]),
                                     "s"(h0), "s"(h1), "s"(h2), "s"(oe.flag));
                    } else {
                        const uint4 h = *reinterpret_cast<const uint4*>(
                            reinterpret_cast<const char*>(qh) + (ko >> 1));  // 16-byte entries
                        h0 = h.x;
                        h1 = h.y;
                        h2 = h.z;
                    }
                    const float q[6] = {half_lo(h0), half_hi(h0), half_lo(h1),
                                        half_hi(h1), half_lo(h2), half_hi(h2)};
                    // fma chains from +0 so that every half-precision product is one v_fma_mix
                    const float qx = fmaf(q[4], dz, fmaf(q[3], dy, fmaf(q[0], dx, 0.f)));
                    const float qy = fmaf(q[5], dz, fmaf(q[1], dy, fmaf(q[3], dx, 0.f)));
                    const float qz = fmaf(q[2], dz, fmaf(q[5], dy, fmaf(q[4], dx, 0.f)));
                    const float dqd = dx * qx + dy * qy + dz * qz;
                    const float s5 = s * (r.w2 * inv2);  // M w^2 |d|^-5
                    const float s5u = s5 * inv2;
                    // the quadrupole, then the octupole added to the same values in place (an
                    // if without an else: two exclusive branches cost register copies at their
                    // join)
                    st = fmaf(s5u, 2.5f * dqd, s);
                    vx = fmaf(-s5, qx, vx);
                    vy = fmaf(-s5, qy, vy);
                    vz = fmaf(-s5, qz, vz);
                    if (OCT && oe.flag) {  // this node's octupole counts
                        // F_i = O_ijk d_j d_k from the traceless form (OctH)
                        const float* o = oe.o;
                        const float ta = dx * dx - dz * dz, tb = dy * dy - dz * dz;
                        const float pxy = dx * dy, pxz = dx * dz, pyz = dy * dz;
                        const float p2xy = pxy + pxy;
                        const float fx = fmaf(o[0], ta, fmaf(o[1], tb, fmaf(o[2], p2xy,
                                         fmaf(o[6], pxz, o[7] * pyz))));
                        const float fy = fmaf(o[2], ta, fmaf(o[3], tb, fmaf(o[1], p2xy,
                                         fmaf(o[7], pxz, o[8] * pyz))));
                        const float fz = fmaf(o[4], ta, fmaf(o[5], tb, fmaf(o[7], pxy,
                                         fmaf(o[9], pxz, o[10] * pyz))));
                        const float gdd = dx * fx + dy * fy + dz * fz;
                        // M w^3 |d|^-7 = s5 u w, the w being in oe.o
                        st = fmaf(-7.f * s5u * inv2, gdd, st);
                        const float s7x3 = 3.f * s5u;
                        vx = fmaf(s7x3, fx, vx);
                        vy = fmaf(s7x3, fy, vy);
                        vz = fmaf(s7x3, fz, vz);
                    }
                }
                vx = fmaf(dx, st, vx);
                vy = fmaf(dy, st, vy);
                vz = fmaf(dz, st, vz);
                if constexpr (PARTIAL) {
                    px = vx;
                    py = vy;
                    pz = vz;
                } else {
                    float t = ax + vx;
                    cx = (t - ax) - vx; ax = t;
                    t = ay + vy;
                    cy = (t - ay) - vy; ay = t;
                    t = az + vz;
                    cz = (t - az) - vz; az = t;
                }
            } else {
                const float yx = fmaf(dx, s, -cx), yy = fmaf(dy, s, -cy), yz = fmaf(dz, s, -cz);
                float t = ax + yx;
                cx = (t - ax) - yx; ax = t;
                t = ay + yy;
                cy = (t - ay) - yy; ay = t;
                t = az + yz;
                cz = (t - az) - yz; az = t;
                ++accepted;  // per lane (a wave-wide popcount costs scalar issue slots)
            }
        }
        if constexpr (QUAD) {
            // the corrected walk's accepted terms, counted per wavefront from the accepting mask
            // with scalar instructions; PARTIAL: the partial sums added with compensation every
            // PART_STEPS wavefront steps (and once at the end)
            acc_wave += uint32_t(__popcll(acc_mask));
            if constexpr (PARTIAL) {
                if ((steps & (PART_STEPS - 1)) == PART_STEPS - 1)
                    fold_partial(ax, ay, az, cx, cy, cz, px, py, pz);
            }
        }
        if (MODE == WALK_ROWS && __float_as_int(r.w2) >= 0) {  // internal: a row if accepted
            const unsigned long long m = acc_mask;
            if (m) {
                if (lane == 0 && nrows < rows.cap)
                    rows.rows[size_t(wave_slot) * rows.cap + nrows] =
                        make_uint4(uint32_t(ko >> 5), uint32_t(__popcll(m)), uint32_t(m),
                                   uint32_t(m >> 32));
                ++nrows;
            }
        }
        // active && !open: resume at skip. An inactive lane sits inside a skipped subtree that
        // contains node k, so its resume point is already >= skip(k) and the max keeps it.
        if (!open) resume = max(resume, so);
        ++steps;
        // skip > k for every node, and skip = k + 1 for a leaf, so the two candidates agree there
        advance<Off, REL>(open_mask, ko + Off(32), so, ko, r, r_next, r_skip);
    }
    if constexpr (PARTIAL) fold_partial(ax, ay, az, cx, cy, cz, px, py, pz);
    if (MODE == WALK_ROWS && lane == 0 && wave_slot * 64 < n) {  // wavefronts holding particles
        rows.count[wave_slot] = nrows;
        atomicMax(&info->max_rows, nrows);
    }
    if (valid) {
        if (est_out)
            est_out[i] = alpha * sqrtf(ax * ax + ay * ay + az * az);
        else
            out[perm ? perm[i] : i] = make_float4(ax, ay, az, 0.f);
    }
    unsigned long long acc64 = accepted;
    for (int off = 32; off > 0; off >>= 1) acc64 += __shfl_xor(acc64, off, 64);
    acc64 += acc_wave;
    if (lane == 0) {
        atomicAdd(&info->steps, (unsigned long long)steps);
        atomicAdd(&info->accepted, acc64);
    }
}

// ---- host side --------------------------------------------------------------------------------

static int build_tree(pnb_ctx* c, const float4* afg, size_t m, bool ref_com = false,
                      bool moments = false) {
    if (!c->bh) c->bh = new BhState();
    BhState* s = c->bh;
    s->built = false;
    s->has_moments = false;
    s->rrec_ok = false;
    s->dd_last = false;
    s->m = m;
    s->nodes = 0;
    s->internal = 0;
    s->max_level = -1;
    s->last_lvl.clear();
    PNB_HIP(c, s->info.ensure(sizeof(BhInfo)));
    BhInfo* info = s->info.as<BhInfo>();
    PNB_HIP(c, s->lvl.ensure(size_t(2) * LVL_SLOTS * LVL_STRIDE * sizeof(uint32_t)));
    uint32_t* lvl_count = s->lvl.as<uint32_t>();
    uint32_t* lvl_cur = lvl_count + LVL_SLOTS * LVL_STRIDE;
    bh_info_init_kernel<<<1, 256, 0, c->stream>>>(info, lvl_count);
    PNB_CHECK_LAUNCH(c);
    if (m == 0) {
        s->built = true;
        return PNB_OK;
    }
    if (m > 0x7fffffffu) return fail(c, PNB_E_ARG, "Barnes-Hut supports at most 2^31-1 bodies");
    const unsigned g = div_up(m, 256);
    PNB_HIP(c, s->bbox_part.ensure(BB_BLOCKS * 6 * sizeof(float)));
    PNB_HIP(c, s->keys.ensure(m * 8));
    PNB_HIP(c, s->keys_alt.ensure(m * 8));
    PNB_HIP(c, s->idx.ensure(m * 4));
    PNB_HIP(c, s->idx_alt.ensure(m * 4));
    PNB_HIP(c, s->sp.ensure(m * sizeof(float4)));
    PNB_HIP(c, s->lcp.ensure(m * sizeof(int)));
    PNB_HIP(c, s->cnt.ensure((m + 1) * 4));
    PNB_HIP(c, s->ip.ensure(m * sizeof(int)));
    PNB_HIP(c, s->key2.ensure(m * 8));

    return timed(c, PNB_PROF_BH_BUILD, [&]() -> int {
        const int bb = int(std::min<size_t>(BB_BLOCKS, div_up(m, 256)));
        bh_bbox_kernel<<<bb, 256, 0, c->stream>>>(afg, m, s->bbox_part.as<float>(), info);
        PNB_CHECK_LAUNCH(c);
        bh_root_kernel<<<1, 64, 0, c->stream>>>(s->bbox_part.as<float>(), bb, info);
        PNB_CHECK_LAUNCH(c);
        bh_keys_kernel<<<g, 256, 0, c->stream>>>(afg, m, info, s->keys.as<uint64_t>(), s->idx.as<uint32_t>());
        PNB_CHECK_LAUNCH(c);
        uint64_t* sk;
        uint32_t* si;
        int rc = timed(c, PNB_PROF_SORT, [&]() -> int {
            return radix_sort_pairs(c, s->keys.as<uint64_t>(), s->idx.as<uint32_t>(),
                                    s->keys_alt.as<uint64_t>(), s->idx_alt.as<uint32_t>(), m, 63,
                                    s->hist, s->bits, &sk, &si);
        });
        if (rc) return rc;
        s->skeys = sk;
        s->sidx = si;
        bh_gather_kernel<<<g, 256, 0, c->stream>>>(afg, m, si, s->sp.as<float4>());
        PNB_CHECK_LAUNCH(c);
        int* lcp = s->lcp.as<int>();
        bh_lcp_kernel<<<g, 256, 0, c->stream>>>(sk, s->sp.as<float4>(), m, lcp);
        PNB_CHECK_LAUNCH(c);
        bh_deep_kernel<<<g, 256, 0, c->stream>>>(sk, m, s->sp.as<float4>(), si, lcp,
                                                 s->key2.as<uint64_t>(), info);
        PNB_CHECK_LAUNCH(c);
        uint32_t* cnt = s->cnt.as<uint32_t>();
        const unsigned lg = std::min<unsigned>(g, 4 * std::max(c->cu_count, 1));
        bh_count_kernel<<<lg, 256, 0, c->stream>>>(lcp, m, cnt, s->ip.as<int>(), info, lvl_count);
        PNB_CHECK_LAUNCH(c);
        bh_level_offsets_kernel<<<1, 256, 0, c->stream>>>(info, lvl_count, lvl_cur);
        PNB_CHECK_LAUNCH(c);
        if ((rc = scan_u32(c, cnt, cnt, m, true))) return rc;
        BhInfo h;
        PNB_HIP(c, hipMemcpyAsync(&h, info, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        uint32_t total = 0;
        PNB_HIP(c, hipMemcpyAsync(&total, cnt + m, 4, hipMemcpyDeviceToHost, c->stream));
        PNB_HIP(c, hipStreamSynchronize(c->stream));
        if (h.flags & FLAG_NONFINITE) return fail(c, PNB_E_ARG, "non-finite affecting position");
        if (h.flags & FLAG_DEPTH)
            return fail(c, PNB_E_DEPTH, "octree deeper than %d levels", MAX_LEVEL);
        s->nodes = total;
        s->max_level = h.max_level;
        // one record past the last node: the walk prefetches records k + 1 and skip <= total
        PNB_HIP(c, s->rec.ensure(size_t(total + 1) * sizeof(NodeRec)));
        PNB_HIP(c, s->sums.ensure(size_t(total) * sizeof(NodeSums)));
        PNB_HIP(c, s->level.ensure(total));
        PNB_HIP(c, s->width.ensure(size_t(total) * sizeof(float)));
        NodeMoments* mom = nullptr;
        NodeQuad* quad = nullptr;
        NodeDesc* desc = nullptr;
        if (moments) {
            PNB_HIP(c, s->mom.ensure(size_t(total) * sizeof(NodeMoments)));
            PNB_HIP(c, s->quad.ensure(size_t(total) * sizeof(NodeQuad)));
            PNB_HIP(c, s->desc.ensure(size_t(total) * sizeof(NodeDesc)));
            mom = s->mom.as<NodeMoments>();
            quad = s->quad.as<NodeQuad>();
            desc = s->desc.as<NodeDesc>();
        }
        NodeRec* rec = s->rec.as<NodeRec>();
        PNB_HIP(c, s->skipv.ensure(size_t(total) * sizeof(uint32_t)));
        uint32_t* skipv = s->skipv.as<uint32_t>();
        bh_node_kernel<<<g, 256, 0, c->stream>>>(sk, s->sp.as<float4>(), lcp, s->ip.as<int>(),
                                                 cnt, m, info, rec, s->sums.as<NodeSums>(),
                                                 s->level.as<uint8_t>(),
                                                 s->width.as<float>(), mom, quad, skipv);
        PNB_CHECK_LAUNCH(c);
        // bottom-up sums, one launch per level over that level's internal nodes
        const int levels = std::max(h.max_level, 0);  // internal levels 0 .. max_level - 1
        std::vector<uint32_t> off(size_t(levels) + 1, 0);
        for (int L = 0; L < levels; ++L) off[L + 1] = off[L] + h.lvl_count[L];
        PNB_HIP(c, s->ilist.ensure(std::max<size_t>(off[levels], 1) * sizeof(uint32_t)));
        s->internal = off[levels];
        s->last_lvl.assign(h.lvl_count, h.lvl_count + levels);
        uint32_t* ilist = s->ilist.as<uint32_t>();
        bh_level_list_kernel<<<lg, 256, 0, c->stream>>>(lcp, s->ip.as<int>(), cnt, m, lvl_cur,
                                                        ilist, info);
        PNB_CHECK_LAUNCH(c);
        for (int L = levels - 1; L >= 0;) {
            if (h.lvl_count[L] > UP_RUN_NODES) {
                const uint32_t n_l = h.lvl_count[L];
                (mom ? bh_up_kernel<true> : bh_up_kernel<false>)<<<div_up(n_l, 256), 256, 0,
                                                                   c->stream>>>(
                    rec, s->sums.as<NodeSums>(), ilist + off[L], n_l, mom, quad,
                    desc, s->width.as<float>(), skipv, s->sp.as<float4>());
                PNB_CHECK_LAUNCH(c);
                --L;
                continue;
            }
            UpRun run{};
            for (; L >= 0 && h.lvl_count[L] <= UP_RUN_NODES && run.levels < UP_RUN_MAX; --L) {
                run.off[run.levels] = off[L];
                run.cnt[run.levels] = h.lvl_count[L];
                ++run.levels;
            }
            (mom ? bh_up_run_kernel<true> : bh_up_run_kernel<false>)<<<1, 1024, 0, c->stream>>>(
                rec, s->sums.as<NodeSums>(), ilist, run, mom, quad, desc, s->width.as<float>(),
                skipv, s->sp.as<float4>());
            PNB_CHECK_LAUNCH(c);
        }
        s->has_moments = moments;
        if (ref_com) {
            bh_ref_com_leaf_kernel<<<div_up(total, 256), 256, 0, c->stream>>>(
                rec, s->sp.as<float4>(), total);
            PNB_CHECK_LAUNCH(c);
            PNB_HIP(c, s->rc_cur.ensure(m * 4));
            PNB_HIP(c, s->rc_keys.ensure(m * 8));
            PNB_HIP(c, s->rc_keys_alt.ensure(m * 8));
            PNB_HIP(c, s->rc_vals.ensure(m * 4));
            PNB_HIP(c, s->rc_vals_alt.ensure(m * 4));
            PNB_HIP(c, s->rc_ord.ensure(m * sizeof(float4)));
            for (int L = 0; L < levels; ++L) {
                bh_ref_com_group_kernel<<<g, 256, 0, c->stream>>>(
                    rec, skipv, si, uint32_t(m), L > 0, s->rc_cur.as<uint32_t>(),
                    s->rc_keys.as<uint64_t>(), s->rc_vals.as<uint32_t>());
                PNB_CHECK_LAUNCH(c);
                uint64_t* ko;
                uint32_t* vo;
                if ((rc = radix_sort_pairs(c, s->rc_keys.as<uint64_t>(), s->rc_vals.as<uint32_t>(),
                                           s->rc_keys_alt.as<uint64_t>(),
                                           s->rc_vals_alt.as<uint32_t>(), m, 63, s->hist, s->bits,
                                           &ko, &vo)))
                    return rc;
                bh_gather_keyed_kernel<<<g, 256, 0, c->stream>>>(afg, ko, uint32_t(m),
                                                                 s->rc_ord.as<float4>());
                PNB_CHECK_LAUNCH(c);
                if (h.lvl_count[L] == 0) continue;
                bh_ref_com_internal_kernel<<<div_up(h.lvl_count[L], 256), 256, 0, c->stream>>>(
                    rec, ilist + off[L], h.lvl_count[L], s->rc_ord.as<float4>());
                PNB_CHECK_LAUNCH(c);
            }
        }

        s->built = true;
        return PNB_OK;
    });
}

// Whether work is being captured on the context's stream into a graph.
static bool stream_capturing(pnb_ctx* c) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &st) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return st != hipStreamCaptureStatusNone;
}

// Capacities of a capturable build from an ordinary build of the same size (its node count and
// depth known on the host): a quarter more nodes and 8 more levels, so that later trees of
// similar inputs fit.
static int dd_size(pnb_ctx* c, size_t m, bool mlp) {
    BhState* s = c->bh;
    const size_t cap = s->nodes + s->nodes / 4 + 1024;
    if (cap + 1 >= (size_t(1) << 27))
        return fail(c, PNB_E_ARG, "capturable Barnes-Hut trees are limited to 2^27 nodes");
    PNB_HIP(c, s->rec.ensure((cap + 1) * sizeof(NodeRec)));
    PNB_HIP(c, s->sums.ensure(cap * sizeof(NodeSums)));
    PNB_HIP(c, s->level.ensure(cap));
    PNB_HIP(c, s->width.ensure(cap * sizeof(float)));
    PNB_HIP(c, s->skipv.ensure(cap * sizeof(uint32_t)));
    PNB_HIP(c, s->ilist.ensure(cap * sizeof(uint32_t)));
    if (mlp) {  // the learned correction's node data and walk entries
        PNB_HIP(c, s->mom.ensure(cap * sizeof(NodeMoments)));
        PNB_HIP(c, s->quad.ensure(cap * sizeof(NodeQuad)));
        PNB_HIP(c, s->desc.ensure(cap * sizeof(NodeDesc)));
        PNB_HIP(c, s->qh.ensure((cap + 1) * sizeof(QuadH)));
        PNB_HIP(c, s->oh.ensure((cap + 1) * sizeof(OctH)));
    }
    s->dd_mlp = mlp;
    s->dd_cap = cap;
    s->dd_levels = std::min(std::max(s->max_level, 0) + 8, MAX_LEVEL + 1);
    s->dd_plan = s->last_lvl;
    s->dd_m = m;
    return PNB_OK;
}

// The octree of a capturable build: the ordinary build's kernels with no host round trip (every
// digit of the sort, the node count checked and used on the device, the bottom-up sums launched
// for every level up to the sized depth). Buffers were sized by dd_size for these m.
static int build_tree_dd(pnb_ctx* c, const float4* afg, size_t m, bool moments) {
    BhState* s = c->bh;
    s->built = false;
    s->has_moments = false;
    s->rrec_ok = false;
    s->m = m;
    BhInfo* info = s->info.as<BhInfo>();
    uint32_t* lvl_count = s->lvl.as<uint32_t>();
    uint32_t* lvl_cur = lvl_count + LVL_SLOTS * LVL_STRIDE;
    bh_info_init_kernel<<<1, 256, 0, c->stream>>>(info, lvl_count);
    PNB_CHECK_LAUNCH(c);
    s->nodes = s->dd_cap;  // an upper bound; the count itself is info->total
    s->max_level = s->dd_levels;
    s->internal = 0;
    s->dd_last = true;
    if (m == 0) {
        s->built = true;
        return PNB_OK;
    }
    const unsigned g = div_up(m, 256);
    return timed(c, PNB_PROF_BH_BUILD, [&]() -> int {
        const int bb = int(std::min<size_t>(BB_BLOCKS, div_up(m, 256)));
        bh_bbox_kernel<<<bb, 256, 0, c->stream>>>(afg, m, s->bbox_part.as<float>(), info);
        PNB_CHECK_LAUNCH(c);
        bh_root_kernel<<<1, 64, 0, c->stream>>>(s->bbox_part.as<float>(), bb, info);
        PNB_CHECK_LAUNCH(c);
        bh_keys_kernel<<<g, 256, 0, c->stream>>>(afg, m, info, s->keys.as<uint64_t>(), s->idx.as<uint32_t>());
        PNB_CHECK_LAUNCH(c);
        uint64_t* sk;
        uint32_t* si;
        int rc = timed(c, PNB_PROF_SORT, [&]() -> int {
            return radix_sort_pairs(c, s->keys.as<uint64_t>(), s->idx.as<uint32_t>(),
                                    s->keys_alt.as<uint64_t>(), s->idx_alt.as<uint32_t>(), m, 63,
                                    s->hist, s->bits, &sk, &si, true);
        });
        if (rc) return rc;
        s->skeys = sk;
        s->sidx = si;
        float4* sp = s->sp.as<float4>();
        bh_gather_kernel<<<g, 256, 0, c->stream>>>(afg, m, si, sp);
        PNB_CHECK_LAUNCH(c);
        int* lcp = s->lcp.as<int>();
        bh_lcp_kernel<<<g, 256, 0, c->stream>>>(sk, sp, m, lcp);
        PNB_CHECK_LAUNCH(c);
        bh_deep_kernel<<<g, 256, 0, c->stream>>>(sk, m, sp, si, lcp, s->key2.as<uint64_t>(), info);
        PNB_CHECK_LAUNCH(c);
        uint32_t* cnt = s->cnt.as<uint32_t>();
        const unsigned lg = std::min<unsigned>(g, 4 * std::max(c->cu_count, 1));
        bh_count_kernel<<<lg, 256, 0, c->stream>>>(lcp, m, cnt, s->ip.as<int>(), info, lvl_count);
        PNB_CHECK_LAUNCH(c);
        bh_level_offsets_kernel<<<1, 256, 0, c->stream>>>(info, lvl_count, lvl_cur);
        PNB_CHECK_LAUNCH(c);
        if ((rc = scan_u32(c, cnt, cnt, m, true))) return rc;
        bh_dd_check_kernel<<<1, 1, 0, c->stream>>>(cnt, m, uint32_t(s->dd_cap), s->dd_levels, info);
        PNB_CHECK_LAUNCH(c);
        NodeRec* rec = s->rec.as<NodeRec>();
        uint32_t* skipv = s->skipv.as<uint32_t>();
        NodeMoments* mom = moments ? s->mom.as<NodeMoments>() : nullptr;
        NodeQuad* quad = moments ? s->quad.as<NodeQuad>() : nullptr;
        NodeDesc* desc = moments ? s->desc.as<NodeDesc>() : nullptr;
        const float* width = s->width.as<float>();
        bh_node_kernel<<<g, 256, 0, c->stream>>>(sk, sp, lcp, s->ip.as<int>(), cnt, m, info, rec,
                                                 s->sums.as<NodeSums>(), s->level.as<uint8_t>(),
                                                 s->width.as<float>(), mom, quad, skipv);
        PNB_CHECK_LAUNCH(c);
        uint32_t* ilist = s->ilist.as<uint32_t>();
        bh_level_list_kernel<<<lg, 256, 0, c->stream>>>(lcp, s->ip.as<int>(), cnt, m, lvl_cur,
                                                        ilist, info);
        PNB_CHECK_LAUNCH(c);
        // Bottom-up sums on the sizing build's plan: a level that held more than half a run's
        // nodes then has a launch of its own, runs of the others one single-workgroup launch.
        // The counts are read on the device, so a level that outgrew its plan is still summed
        // (a run only takes longer, a launch strides over more nodes).
        const auto planned = [&](int L) -> uint32_t {
            return size_t(L) < s->dd_plan.size() ? s->dd_plan[size_t(L)] : 0u;
        };
        const unsigned max_grid = 4 * unsigned(std::max(c->cu_count, 1));
        for (int L = s->dd_levels - 1; L >= 0;) {
            if (planned(L) > UP_RUN_NODES / 2) {
                const uint32_t n_l = planned(L);
                const unsigned ug = std::min<unsigned>(div_up(n_l + n_l / 4, 256u), max_grid);
                (mom ? bh_up_level_dd_kernel<true> : bh_up_level_dd_kernel<false>)<<<
                    ug, 256, 0, c->stream>>>(rec, s->sums.as<NodeSums>(), ilist, info, L, skipv,
                                             sp, mom, quad, desc, width);
                PNB_CHECK_LAUNCH(c);
                --L;
                continue;
            }
            const int hi = L;
            while (L >= 0 && planned(L) <= UP_RUN_NODES / 2) --L;
            (mom ? bh_up_levels_dd_kernel<true> : bh_up_levels_dd_kernel<false>)<<<
                1, 1024, 0, c->stream>>>(rec, s->sums.as<NodeSums>(), ilist, info, hi, L + 1, skipv,
                                         sp, mom, quad, desc, width);
            PNB_CHECK_LAUNCH(c);
        }
        s->has_moments = moments;
        s->built = true;
        return PNB_OK;
    });
}

// After a capturable build: waits for it and takes its node count, depth and status to the host.
static int dd_resolve(pnb_ctx* c) {
    BhState* s = c->bh;
    if (!s || !s->dd_last) return PNB_OK;
    BhInfo h;
    PNB_HIP(c, hipMemcpyAsync(&h, s->info.ptr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    s->dd_last = false;
    if (h.flags) {
        s->built = false;
        if (h.flags & FLAG_NONFINITE) return fail(c, PNB_E_ARG, "non-finite affecting position");
        if (h.flags & FLAG_DEPTH) return fail(c, PNB_E_DEPTH, "octree deeper than %d levels", MAX_LEVEL);
        return fail(c, PNB_E_STATE,
                    "the capturable Barnes-Hut build was sized for %zu nodes and %d levels; this "
                    "tree needs more (the accelerations are NaN): call again outside a capture to "
                    "size it anew", s->dd_cap, s->dd_levels);
    }
    s->nodes = s->m ? h.total : 0;
    s->max_level = s->m ? h.max_level : -1;
    size_t internal = 0;
    for (int L = 0; L < std::max(s->max_level, 0); ++L) internal += h.lvl_count[L];
    s->internal = internal;
    return PNB_OK;
}

const MlpWeights* mlp_weights(pnb_ctx* c);  // mlp.hip: nullptr until initialised
bool mlp_octupole(pnb_ctx* c);               // mlp.hip: the network has octupole outputs

// Walk records for the relative criterion: node k's record with begin = the bits of K = M w^p,
// p = 2, 3 or 4 (f32, each operation rounded; 0 for leaves, which are never opened) and end =
// first body | count << 27 for an internal node of at most `bucket` (<= 31) bodies whose first
// body is below 2^27, else 0; for the total + 1 records (the last is the sentinel the walk
// prefetches).
__global__ void bh_rel_rec_kernel(const NodeRec* __restrict__ rec, size_t total, int power,
                                  uint32_t bucket, NodeRec* __restrict__ rrec) {
    const size_t k = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k > total) return;
    NodeRec r = rec[k];
    float K = 0.f;
    uint32_t packed = 0;
    if (k < total && __float_as_int(r.w2) >= 0) {
        K = mul_rn(r.mass, r.w2);
        if (power == 3) K = mul_rn(K, __fsqrt_rn(r.w2));
        if (power == 4) K = mul_rn(K, r.w2);
        const uint32_t count = r.end - r.begin;
        if (count <= bucket && r.begin < (1u << 27)) packed = r.begin | (count << 27);
    }
    r.begin = __float_as_uint(K);
    r.end = packed;
    rrec[k] = r;
}

// Centroid of each wavefront's particles (64 consecutive particles in walk order; the last one
// may hold fewer): a fixed-order butterfly sum per wavefront, {x, y, z, 0}.
__global__ __launch_bounds__(256) void bh_wave_centroid_kernel(const float4* __restrict__ pts,
                                                               uint32_t n,
                                                               float4* __restrict__ cen) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < n;
    const float4 p = valid ? pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    float x = p.x, y = p.y, z = p.z, c = valid ? 1.f : 0.f;
    for (int off = 32; off > 0; off >>= 1) {
        x += __shfl_xor(x, off, 64);
        y += __shfl_xor(y, off, 64);
        z += __shfl_xor(z, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    if ((threadIdx.x & 63) == 0 && i < n) cen[i >> 6] = make_float4(x / c, y / c, z / c, 0.f);
}

// Sort keys for the wavefronts of the relative walk: the binary exponent of alpha |a~|
// (non-negative floats order as their bits), so that a stable ascending sort starts the
// wavefronts with the smallest estimates, whose particles open the most nodes, first, and keeps
// the spatial (tree) order within each factor-of-two band for the caches.
__global__ void bh_wave_keys_kernel(const float* __restrict__ est, uint32_t waves,
                                    uint64_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < waves) {
        keys[w] = __float_as_uint(est[w]) >> 23;
        vals[w] = w;
    }
}

// The estimate pass walks EST_GROUP centroids per wavefront (its other lanes idle), so that more,
// shorter wavefronts hide the walk's latency; every centroid's decisions and sums are its own,
// so the estimates do not depend on the grouping.
constexpr uint32_t EST_GROUP = 8;

__global__ void bh_cen_spread_kernel(const float4* __restrict__ cen, uint32_t waves,
                                     float4* __restrict__ spread, uint32_t groups) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < groups * 64) {
        const uint32_t w = (t >> 6) * EST_GROUP + (t & 63) % EST_GROUP;  // idle lanes: a copy
        spread[t] = cen[w < waves ? w : waves - 1];
    }
}

__global__ void bh_est_gather_kernel(const float* __restrict__ spread, uint32_t waves,
                                     float* __restrict__ est) {
    const uint32_t w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w < waves) est[w] = spread[(w / EST_GROUP) * 64 + w % EST_GROUP];
}

template <int MODE, bool REL>
static void launch_walk(bool wide, unsigned grid, hipStream_t st, const NodeRec* nr, uint32_t total,
                        const float4* pts, const uint32_t* perm, uint32_t n, float theta2,
                        float eps2, float4* out, BhInfo* info, const WalkRows& rows,
                        const QuadH* qh, const float* est_in = nullptr, uint32_t est_shift = 0,
                        float* est_out = nullptr, float alpha = 0.f,
                        const uint32_t* wave_order = nullptr, const float4* bodies = nullptr,
                        uint32_t lanes = 64, int dd = 0, int packed = 0,
                        const OctH* oh = nullptr, int oct = 0) {
    if constexpr (MODE == WALK_QUAD) {
        if (oct) {
            launch_walk<WALK_OCT, REL>(wide, grid, st, nr, total, pts, perm, n, theta2, eps2, out,
                                       info, rows, qh, est_in, est_shift, est_out, alpha,
                                       wave_order, bodies, lanes, dd, packed, oh, 0);
            return;
        }
    }
    if (wide)
        bh_walk_kernel<MODE, true, REL><<<grid, 256, 0, st>>>(
            nr, total, pts, perm, n, theta2, eps2, out, info, rows, qh, oh, est_in, est_shift,
            est_out, alpha, wave_order, bodies, lanes, dd, packed);
    else
        bh_walk_kernel<MODE, false, REL><<<grid, 256, 0, st>>>(
            nr, total, pts, perm, n, theta2, eps2, out, info, rows, qh, oh, est_in, est_shift,
            est_out, alpha, wave_order, bodies, lanes, dd, packed);
}

// Particles per wavefront of the plain and corrected walks: the largest power of two (at most
// 64, at least 1) that still gives WALK_WAVES_PER_CU wavefronts per CU; 64 for a problem that
// fits one wavefront. Measured at theta = 0.5 on Plummer spheres (walk time, 256 CUs): 1,024
// bodies 173 us at 64 lanes, 97 us at 1; 4,096: 352 us at 64, 225-238 us at 2-8; 16,384:
// 577 us at 64, 382 us at 8; 65,536: 695 us at 64, 630 us at 32; from 262,144 on, 64 is best.
// PNB_WALK_LANES (1..64) overrides it.
constexpr size_t WALK_WAVES_PER_CU = 8;
static uint32_t walk_lanes(const pnb_ctx* c, size_t n) {
    if (const char* e = std::getenv("PNB_WALK_LANES")) {
        const long v = std::strtol(e, nullptr, 10);
        if (v >= 1 && v <= 64) return uint32_t(v);
    }
    if (n <= 64) return 64;
    const size_t target = WALK_WAVES_PER_CU * size_t(std::max(c->cu_count, 1));
    uint32_t lanes = 64;
    while (lanes > 1 && div_up(n, size_t(lanes)) < target) lanes >>= 1;
    return lanes;
}

// quad: add the learned correction (the per-node moments in s->qh). rel: the relative criterion
// in addition to the geometric one: an estimate pass (plain walk at s->rel_theta_est) gives
// alpha |a~| per particle or, by default, per wavefront at the centroid of its 64 particles; then
// the walk opens a node also when M w^p / r^(p+2) > alpha |a~| (p = 2, or 3 with quad). Both
// passes count in the walk's time and in the step/accepted counters.
static int walk(pnb_ctx* c, const float4* pts, const uint32_t* perm, size_t n, float theta,
                float softening, float4* out, const WalkRows* rows = nullptr, bool quad = false,
                bool rel = false, bool dd = false) {
    BhState* s = c->bh;
    const float theta2 = theta * theta;  // (self.theta * self.theta), sequential.rs:491
    const float eps2 = softening * softening;
    const uint32_t total = uint32_t(s->nodes);
    rel = rel && total > 0 && !rows;
    const size_t waves = div_up(n, 64);
    if (rel) {
        PNB_HIP(c, s->rrec.ensure((size_t(total) + 1) * sizeof(NodeRec)));
        PNB_HIP(c, s->est.ensure((s->rel_per_particle ? n : waves) * sizeof(float)));
        if (!s->rel_per_particle) {
            const size_t groups = div_up(waves, size_t(EST_GROUP));
            PNB_HIP(c, s->cen.ensure(waves * sizeof(float4)));
            PNB_HIP(c, s->cen_spread.ensure(groups * 64 * sizeof(float4)));
            PNB_HIP(c, s->est_spread.ensure(groups * 64 * sizeof(float)));
            PNB_HIP(c, s->wkey.ensure(waves * sizeof(uint64_t)));
            PNB_HIP(c, s->wkey_alt.ensure(waves * sizeof(uint64_t)));
            PNB_HIP(c, s->wval.ensure(waves * sizeof(uint32_t)));
            PNB_HIP(c, s->wval_alt.ensure(waves * sizeof(uint32_t)));
        }
    }
    return timed(c, PNB_PROF_BH_WALK, [&]() -> int {
        const unsigned grid = div_up(n, 256);
        BhInfo* info = s->info.as<BhInfo>();
        const NodeRec* nr = s->rec.as<NodeRec>();
        // 64-bit record offsets once the records (and the one past the last) reach 4 GiB;
        // PNB_WIDE_WALK=1 takes them for any tree (tests compare both forms)
        const char* force = std::getenv("PNB_WIDE_WALK");
        const bool wide = (uint64_t(total) + 1) * sizeof(NodeRec) >= (uint64_t(1) << 32) ||
                          (force && force[0] == '1');
        if (rel) {
            // the power of the first neglected term: the quadrupole (plain walk), the octupole
            // (with the correction's quadrupole) or the hexadecapole (and its octupole)
            const int power = quad ? (mlp_octupole(c) ? 4 : 3) : 2;
            if (!s->rrec_ok || s->rrec_power != power || s->rrec_bucket != s->rel_bucket) {
                bh_rel_rec_kernel<<<div_up(size_t(total) + 1, 256), 256, 0, c->stream>>>(
                    nr, total, power, uint32_t(s->rel_bucket), s->rrec.as<NodeRec>());
                PNB_CHECK_LAUNCH(c);
                s->rrec_ok = true;
                s->rrec_power = power;
                s->rrec_bucket = s->rel_bucket;
            }
            const float te = s->rel_theta_est;
            float* est = s->est.as<float>();
            uint32_t shift = 0;
            const uint32_t* order = nullptr;
            if (s->rel_per_particle) {
                launch_walk<WALK_PLAIN, false>(wide, grid, c->stream, nr, total, pts, perm,
                                               uint32_t(n), te * te, eps2, out, info, WalkRows{},
                                               nullptr, nullptr, 0, est, s->rel_alpha);
                PNB_CHECK_LAUNCH(c);
            } else {
                float4* cen = s->cen.as<float4>();
                bh_wave_centroid_kernel<<<grid, 256, 0, c->stream>>>(pts, uint32_t(n), cen);
                PNB_CHECK_LAUNCH(c);
                const uint32_t groups = uint32_t(div_up(waves, size_t(EST_GROUP)));
                float4* spread = s->cen_spread.as<float4>();
                float* est_spread = s->est_spread.as<float>();
                bh_cen_spread_kernel<<<div_up(size_t(groups) * 64, 256), 256, 0, c->stream>>>(
                    cen, uint32_t(waves), spread, groups);
                PNB_CHECK_LAUNCH(c);
                // lanes 0 .. EST_GROUP-1 of each wavefront walk; in the last one only those with a
                // centroid (i < n excludes the rest)
                const uint32_t n_est = (groups - 1) * 64 + (uint32_t(waves) - (groups - 1) * EST_GROUP);
                launch_walk<WALK_PLAIN, false>(wide, div_up(size_t(groups) * 64, 256), c->stream,
                                               nr, total, spread, nullptr, n_est, te * te,
                                               eps2, out, info, WalkRows{}, nullptr, nullptr, 0,
                                               est_spread, s->rel_alpha, nullptr, nullptr,
                                               EST_GROUP);
                PNB_CHECK_LAUNCH(c);
                bh_est_gather_kernel<<<div_up(waves, 256), 256, 0, c->stream>>>(
                    est_spread, uint32_t(waves), est);
                PNB_CHECK_LAUNCH(c);
                shift = 6;
            }
            // Heaviest first: the wavefronts with the smallest estimates open the most nodes, and
            // with few wavefronts per CU their tail sets the walk's time (two Plummer spheres at
            // 1e-4 rms: the corrected walk went from 11.8 to 9.2 ms at 1M bodies, 64 wavefronts
            // per CU); with many (256 per CU at 4M bodies) the tail is absorbed and reordering
            // cost 1.5 %. The order changes no result: every wavefront's sums are its own.
            if (!s->rel_per_particle && waves < size_t(128) * size_t(std::max(c->cu_count, 1))) {
                bh_wave_keys_kernel<<<div_up(waves, 256), 256, 0, c->stream>>>(
                    est, uint32_t(waves), s->wkey.as<uint64_t>(), s->wval.as<uint32_t>());
                PNB_CHECK_LAUNCH(c);
                uint64_t* sk;
                uint32_t* sv;
                if (int rc = radix_sort_pairs(c, s->wkey.as<uint64_t>(), s->wval.as<uint32_t>(),
                                              s->wkey_alt.as<uint64_t>(),
                                              s->wval_alt.as<uint32_t>(), waves, 8, s->hist,
                                              s->bits, &sk, &sv))
                    return rc;
                order = sv;
            }
            const NodeRec* rr = s->rrec.as<NodeRec>();
            const float4* bodies = s->sp.as<float4>();  // the affecting particles, tree order
            if (quad)
                launch_walk<WALK_QUAD, true>(wide, grid, c->stream, rr, total, pts, perm,
                                             uint32_t(n), theta2, eps2, out, info, WalkRows{},
                                             s->qh.as<QuadH>(), est, shift, nullptr, 0.f, order,
                                             bodies, 64, 0, 0, s->oh.as<OctH>(),
                                             mlp_octupole(c) ? 1 : 0);
            else
                launch_walk<WALK_PLAIN, true>(wide, grid, c->stream, rr, total, pts, perm,
                                              uint32_t(n), theta2, eps2, out, info, WalkRows{},
                                              nullptr, est, shift, nullptr, 0.f, order, bodies);
        } else if (rows) {
            launch_walk<WALK_ROWS, false>(wide, grid, c->stream, nr, total, pts, perm, uint32_t(n),
                                          theta2, eps2, out, info, *rows, nullptr);
        } else {
            const uint32_t lanes = walk_lanes(c, n);
            const unsigned pgrid = unsigned(div_up(div_up(n, size_t(lanes)) * 64, size_t(256)));
            if (quad)
                launch_walk<WALK_QUAD, false>(wide, pgrid, c->stream, nr, total, pts, perm,
                                              uint32_t(n), theta2, eps2, out, info, WalkRows{},
                                              s->qh.as<QuadH>(), nullptr, 0, nullptr, 0.f, nullptr,
                                              nullptr, lanes, dd ? 1 : 0, 1,
                                              s->oh.as<OctH>(), mlp_octupole(c) ? 1 : 0);
            else
                launch_walk<WALK_PLAIN, false>(wide, pgrid, c->stream, nr, total, pts, perm,
                                               uint32_t(n), theta2, eps2, out, info, WalkRows{},
                                               nullptr, nullptr, 0, nullptr, 0.f, nullptr, nullptr,
                                               lanes, dd ? 1 : 0, 1);
        }
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    });
}

// The learned correction's per-node moments: qh[k] = MLP(node k) for the current tree.
static int node_network(pnb_ctx* c) {
    BhState* s = c->bh;
    const MlpWeights* wg = mlp_weights(c);
    if (!wg) return fail(c, PNB_E_STATE, "PNB_BH_MLP needs an initialised MLP (pnb_mlp_init)");
    if (!s->has_moments) return fail(c, PNB_E_STATE, "octree built without node moments");
    const size_t total = s->nodes;  // a capturable build: its capacity
    // one entry past the last node, as for the records (the walk prefetches k + 1 and skip); a
    // capturable build's entries were sized with its buffers, so nothing grows here
    PNB_HIP(c, s->qh.ensure((size_t(total) + 1) * sizeof(QuadH)));
    PNB_HIP(c, s->oh.ensure((size_t(total) + 1) * sizeof(OctH)));
    if (total == 0) return PNB_OK;
    const bool dd = s->dd_last;  // the internal-node count is on the device
    if (!dd && s->internal == 0) return PNB_OK;
    return timed(c, PNB_PROF_MLP, [&]() -> int {
        const size_t tiles = div_up(dd ? total : s->internal, 16);
        const unsigned grid = unsigned(std::min<size_t>(div_up(tiles, 4), 16 * size_t(std::max(c->cu_count, 1))));
        mlp_node_kernel<<<grid, 256, 0, c->stream>>>(s->rec.as<NodeRec>(), s->quad.as<NodeQuad>(),
                                                     s->level.as<uint8_t>(), s->desc.as<NodeDesc>(),
                                                     s->ilist.as<uint32_t>(), uint32_t(s->internal),
                                                     wg, s->qh.as<QuadH>(), s->oh.as<OctH>(),
                                                     dd ? s->info.as<BhInfo>() : nullptr);
        PNB_CHECK_LAUNCH(c);
        return PNB_OK;
    });
}

// Orders the affected particles along the tree's key order so that each wavefront holds
// neighbours, walks, and scatters results back to the caller's order.
static int accel_dev(pnb_ctx* c, const float4* affected, size_t n, const float4* afg, size_t m,
                     float theta, float softening, float4* out, bool mlp = false,
                     bool ref_com = false, bool rel = false) {
    if (n == 0) return PNB_OK;
    if (n > 0x7fffffffu) return fail(c, PNB_E_ARG, "Barnes-Hut supports at most 2^31-1 particles");
    if (mlp && !mlp_weights(c))
        return fail(c, PNB_E_STATE, "PNB_BH_MLP needs an initialised MLP (pnb_mlp_init)");
    if (!c->bh) c->bh = new BhState();
    bool dd = false;
    if (c->bh->dd_mode) {
        if (ref_com || rel)
            return fail(c, PNB_E_ARG, "capturable Barnes-Hut takes neither PNB_BH_REF_COM nor "
                                      "PNB_BH_RELATIVE");
        BhState* s = c->bh;
        const bool sized = s->dd_cap > 0 && s->dd_m == m && s->dd_mlp == mlp &&
                           ((affected == afg && n == m) || s->dd_n == n);
        if (sized) {
            dd = true;
        } else if (stream_capturing(c)) {
            return fail(c, PNB_E_STATE, "capturable Barnes-Hut: call once outside the capture "
                                        "with these sizes and flags first (it sizes the tree's "
                                        "buffers)");
        }
    }
    // inside an outside capture: no events on the stream
    const bool was_capturing = c->capturing;
    if (dd) c->capturing = c->capturing || stream_capturing(c);
    struct Restore {
        pnb_ctx* c;
        bool v;
        ~Restore() { c->capturing = v; }
    } restore{c, was_capturing};
    if (dd) {
        if (int rc = build_tree_dd(c, afg, m, mlp)) return rc;
    } else {
        if (int rc = build_tree(c, afg, m, ref_com, mlp)) return rc;
        if (c->bh->dd_mode) {
            // The sizing call of a capturable build: capacities from this tree, then the tree
            // again in the sized buffers (growing a buffer discards its contents, so the tree
            // just built may be gone), built the capturable way.
            if (int rc = dd_size(c, m, mlp)) return rc;
            c->bh->dd_n = n;
            dd = true;
            if (int rc = build_tree_dd(c, afg, m, mlp)) return rc;
        }
    }
    BhState* s = c->bh;
    const bool quad = mlp && m;
    if (quad)
        if (int rc = node_network(c)) return rc;
    if (affected == afg && n == m)
        return walk(c, s->sp.as<float4>(), s->sidx, n, theta, softening, out, nullptr, quad, rel,
                    dd);
    const unsigned g = div_up(n, 256);
    PNB_HIP(c, s->akeys.ensure(n * 8));
    PNB_HIP(c, s->akeys_alt.ensure(n * 8));
    PNB_HIP(c, s->aidx.ensure(n * 4));
    PNB_HIP(c, s->aidx_alt.ensure(n * 4));
    PNB_HIP(c, s->apts.ensure(n * sizeof(float4)));
    if (m == 0) {
        // no tree: every particle gets zero (the reference's root is None)
        return walk(c, affected, nullptr, n, theta, softening, out, nullptr, false, false, dd);
    }
    bh_keys_kernel<<<g, 256, 0, c->stream>>>(affected, n, s->info.as<BhInfo>(),
                                             s->akeys.as<uint64_t>(), s->aidx.as<uint32_t>());
    PNB_CHECK_LAUNCH(c);
    uint64_t* sk;
    uint32_t* si;
    if (int rc = radix_sort_pairs(c, s->akeys.as<uint64_t>(), s->aidx.as<uint32_t>(),
                                  s->akeys_alt.as<uint64_t>(), s->aidx_alt.as<uint32_t>(), n, 63,
                                  s->hist, s->bits, &sk, &si, dd))
        return rc;
    bh_gather_kernel<<<g, 256, 0, c->stream>>>(affected, n, si, s->apts.as<float4>());
    PNB_CHECK_LAUNCH(c);
    return walk(c, s->apts.as<float4>(), si, n, theta, softening, out, nullptr, quad, rel, dd);
}

static int copy_h2d(pnb_ctx* c, DevBuf& b, const float* src, size_t count) {
    PNB_HIP(c, b.ensure(std::max<size_t>(count, 1) * sizeof(float4)));
    if (count)
        PNB_HIP(c, hipMemcpyAsync(b.ptr, src, count * sizeof(float4), hipMemcpyHostToDevice, c->stream));
    return PNB_OK;
}

// ---- training data for the learned correction -------------------------------------------------

__global__ void mlp_gather_waves_kernel(const float4* __restrict__ sp, const uint32_t* __restrict__ starts,
                                        uint32_t waves, float eps2, float4* __restrict__ dst,
                                        float4* __restrict__ lanes) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < waves * 64) {
        const float4 p = sp[starts[i / 64] + (i % 64)];
        dst[i] = p;
        lanes[i] = make_float4(p.x, p.y, p.z, eps2);
    }
}

// target = brute force - Barnes-Hut, weight = 1 / |brute force|^2 (0 when it vanishes)
__global__ void mlp_target_kernel(const float4* __restrict__ bf, const float4* __restrict__ bh,
                                  uint32_t n, float* __restrict__ target, float* __restrict__ wgt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 a = bf[i], b = bh[i];
    target[4 * i] = a.x - b.x;
    target[4 * i + 1] = a.y - b.y;
    target[4 * i + 2] = a.z - b.z;
    target[4 * i + 3] = 0.f;
    const float n2 = a.x * a.x + a.y * a.y + a.z * a.z;
    wgt[i] = n2 > 0.f ? 1.f / n2 : 0.f;
}

// One thread per stored row: the node's features (node_features, as mlp_node_kernel computes
// them) and geometry {centre of mass, M w^2, M w^3, 0, 0, 0} (two float4 per row), written
// compactly at wave_rows[w] + r.
__global__ void mlp_row_feats_kernel(const NodeRec* __restrict__ nodes,
                                     const float* __restrict__ widths,
                                     const NodeQuad* __restrict__ quad,
                                     const uint8_t* __restrict__ level,
                                     const NodeDesc* __restrict__ desc,
                                     const uint4* __restrict__ rows, uint32_t cap,
                                     const uint32_t* __restrict__ wave_rows, uint32_t waves,
                                     float* __restrict__ feats, float4* __restrict__ geom,
                                     unsigned long long* __restrict__ mask,
                                     uint32_t* __restrict__ rwave) {
    const uint32_t w = blockIdx.y;
    if (w >= waves) return;
    const uint32_t b = wave_rows[w], cnt = wave_rows[w + 1] - b;
    const float root_mass = nodes[0].mass;
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < cnt; r += gridDim.x * blockDim.x) {
        const uint4 row = rows[size_t(w) * cap + r];
        const NodeRec nr = nodes[row.x];
        float f[MLP_KIN];
        node_features(nr, quad[row.x], level[row.x], desc[row.x], root_mass, f);
        const size_t dst = size_t(b) + r;
#pragma unroll
        for (int q = 0; q < MLP_FEAT; ++q) feats[dst * MLP_FEAT + q] = f[q];
        const float wd = widths[row.x];
        geom[2 * dst] = make_float4(nr.cx, nr.cy, nr.cz, nr.mass * wd * wd);
        geom[2 * dst + 1] = make_float4(nr.mass * wd * wd * wd, 0.f, 0.f, 0.f);
        mask[dst] = (unsigned long long)row.z | ((unsigned long long)row.w << 32);
        rwave[dst] = w;
    }
}

int mlp_append(pnb_ctx* c, const float* feats, const float4* geom, const unsigned long long* mask,
               const uint32_t* row_wave_local, const uint32_t* wave_rows_local, size_t rows,
               size_t waves, const float4* lanes, const float* target, const float* wgt,
               unsigned long long dropped);

static int collect(pnb_ctx* c, const float4* p, size_t n, float theta, float softening,
                   size_t waves, uint64_t seed, uint32_t cap) {
    if (int rc = build_tree(c, p, n, false, true)) return rc;
    BhState* s = c->bh;
    const size_t full = n / 64;
    const size_t W = std::min(waves, full);
    if (W == 0) return fail(c, PNB_E_ARG, "need at least 64 particles");
    std::vector<uint32_t> starts(W);
    const size_t stride = full / W;
    const size_t offset = size_t(seed % stride);
    for (size_t w = 0; w < W; ++w) starts[w] = uint32_t((offset + w * stride) * 64);
    const size_t nr = W * 64;
    PNB_HIP(c, s->c_starts.ensure(W * 4));
    PNB_HIP(c, s->c_pts.ensure(nr * sizeof(float4)));
    PNB_HIP(c, s->c_lanes.ensure(nr * sizeof(float4)));
    PNB_HIP(c, s->c_bh.ensure(nr * sizeof(float4)));
    PNB_HIP(c, s->c_bf.ensure(nr * sizeof(float4)));
    PNB_HIP(c, s->c_list.ensure(W * cap * sizeof(uint4)));
    PNB_HIP(c, s->c_count.ensure(W * 4));
    PNB_HIP(c, s->c_target.ensure(nr * 16));
    PNB_HIP(c, s->c_wgt.ensure(nr * 4));
    PNB_HIP(c, s->c_rows.ensure((W + 1) * 4));
    PNB_HIP(c, hipMemcpyAsync(s->c_starts.ptr, starts.data(), W * 4, hipMemcpyHostToDevice, c->stream));
    mlp_gather_waves_kernel<<<div_up(nr, 256), 256, 0, c->stream>>>(
        s->sp.as<float4>(), s->c_starts.as<uint32_t>(), uint32_t(W), softening * softening,
        s->c_pts.as<float4>(), s->c_lanes.as<float4>());
    PNB_CHECK_LAUNCH(c);
    WalkRows rows{s->c_list.as<uint4>(), s->c_count.as<uint32_t>(), cap};
    if (int rc = walk(c, s->c_pts.as<float4>(), nullptr, nr, theta, softening, s->c_bh.as<float4>(),
                      &rows))
        return rc;
    if (int rc = bf_launch(c, s->c_pts.as<float4>(), nr, s->sp.as<float4>(), n, softening, 1,
                           s->c_bf.as<float4>()))
        return rc;
    mlp_target_kernel<<<div_up(nr, 256), 256, 0, c->stream>>>(s->c_bf.as<float4>(), s->c_bh.as<float4>(),
                                                            uint32_t(nr), s->c_target.as<float>(),
                                                            s->c_wgt.as<float>());
    PNB_CHECK_LAUNCH(c);
    std::vector<uint32_t> counts(W), offs(W + 1, 0);
    PNB_HIP(c, hipMemcpyAsync(counts.data(), s->c_count.ptr, W * 4, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    unsigned long long dropped = 0;
    uint32_t most = 0;
    for (size_t w = 0; w < W; ++w) {
        const uint32_t kept = std::min(counts[w], cap);
        dropped += counts[w] - kept;
        most = std::max(most, kept);
        offs[w + 1] = offs[w] + kept;
    }
    const size_t R = offs[W];
    PNB_HIP(c, s->c_cfeats.ensure(std::max<size_t>(R, 1) * MLP_FEAT * 4));
    PNB_HIP(c, s->c_cgeom.ensure(std::max<size_t>(R, 1) * MLP_GEOM * sizeof(float)));
    PNB_HIP(c, s->c_cmask.ensure(std::max<size_t>(R, 1) * 8));
    PNB_HIP(c, s->c_rwave.ensure(std::max<size_t>(R, 1) * 4));
    PNB_HIP(c, hipMemcpyAsync(s->c_rows.ptr, offs.data(), (W + 1) * 4, hipMemcpyHostToDevice, c->stream));
    if (R) {
        const dim3 grid(unsigned(std::min<uint32_t>(div_up(most, 256), 64)), unsigned(W));
        mlp_row_feats_kernel<<<grid, 256, 0, c->stream>>>(
            s->rec.as<NodeRec>(), s->width.as<float>(), s->quad.as<NodeQuad>(),
            s->level.as<uint8_t>(), s->desc.as<NodeDesc>(), s->c_list.as<uint4>(), cap,
            s->c_rows.as<uint32_t>(), uint32_t(W), s->c_cfeats.as<float>(), s->c_cgeom.as<float4>(),
            s->c_cmask.as<unsigned long long>(), s->c_rwave.as<uint32_t>());
        PNB_CHECK_LAUNCH(c);
    }
    return mlp_append(c, s->c_cfeats.as<float>(), s->c_cgeom.as<float4>(),
                      s->c_cmask.as<unsigned long long>(), s->c_rwave.as<uint32_t>(),
                      s->c_rows.as<uint32_t>(), R, W, s->c_lanes.as<float4>(),
                      s->c_target.as<float>(), s->c_wgt.as<float>(), dropped);
}

}  // namespace pnb

using namespace pnb;

#define PNB_BH_ARGS_CHECK()                                                                 \
    if (!c) return PNB_E_ARG;                                                               \
    if ((n && (!affected || !out)) || (m && !affecting))                                    \
        return fail(c, PNB_E_ARG, "null pointer with non-zero length");                     \
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");

extern "C" {

int pnb_bh_accel_f32_dev(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                         size_t m, float theta, float softening, int checked, unsigned flags,
                         float* out) {
    PNB_BH_ARGS_CHECK();
    (void)checked;  // the traversal skips zero-distance nodes for both (sequential.rs:485-487)
    if (flags & ~PNB_BH_ALL_FLAGS) return fail(c, PNB_E_ARG, "unknown Barnes-Hut flags");
    if (int rc = activate(c)) return rc;
    return accel_dev(c, reinterpret_cast<const float4*>(affected), n,
                     reinterpret_cast<const float4*>(affecting), m, theta, softening,
                     reinterpret_cast<float4*>(out), (flags & PNB_BH_MLP) != 0,
                     (flags & PNB_BH_REF_COM) != 0, (flags & PNB_BH_RELATIVE) != 0);
}

int pnb_bh_accel_f32(pnb_ctx* c, const float* affected, size_t n, const float* affecting,
                     size_t m, float theta, float softening, int checked, unsigned flags,
                     float* out) {
    PNB_BH_ARGS_CHECK();
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    if (!c->bh) c->bh = new BhState();
    BhState* s = c->bh;
    const bool same = affected == affecting && n == m;
    if (int rc = copy_h2d(c, s->h_afg, affecting, m)) return rc;
    if (!same)
        if (int rc = copy_h2d(c, s->h_aff, affected, n)) return rc;
    PNB_HIP(c, s->h_out.ensure(n * sizeof(float4)));
    const float* daff = same ? s->h_afg.as<float>() : s->h_aff.as<float>();
    if (int rc = pnb_bh_accel_f32_dev(c, daff, n, s->h_afg.as<float>(), m, theta, softening,
                                      checked, flags, s->h_out.as<float>()))
        return rc;
    PNB_HIP(c, hipMemcpyAsync(out, s->h_out.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_reordered_bh_accel_f32_dev(pnb_ctx* c, const float* particles, size_t n, float theta,
                                   float softening, int checked, unsigned flags, float* out) {
    if (!c) return PNB_E_ARG;
    if (n && (!particles || !out)) return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    if (!c->bh) c->bh = new BhState();
    BhState* s = c->bh;
    if (s->dd_mode && stream_capturing(c))
        return fail(c, PNB_E_STATE, "the massive/massless split reads its count on the host: "
                                    "capture pnb_bh_accel_f32_dev instead");
    PNB_HIP(c, s->afg_tmp.ensure(n * sizeof(float4)));
    uint32_t m = 0;
    const float4* p = reinterpret_cast<const float4*>(particles);
    if (int rc = split_affecting(c, p, n, s->afg_tmp.as<float4>(), &m)) return rc;
    const float4* afg = s->afg_tmp.as<float4>();
    if (m == n) afg = p;  // every particle is affecting: same storage, the tree order is reused
    (void)checked;
    if (flags & ~PNB_BH_ALL_FLAGS) return fail(c, PNB_E_ARG, "unknown Barnes-Hut flags");
    return accel_dev(c, p, n, afg, m, theta, softening, reinterpret_cast<float4*>(out),
                     (flags & PNB_BH_MLP) != 0, (flags & PNB_BH_REF_COM) != 0,
                     (flags & PNB_BH_RELATIVE) != 0);
}

int pnb_reordered_bh_accel_f32(pnb_ctx* c, const float* particles, size_t n, float theta,
                               float softening, int checked, unsigned flags, float* out) {
    if (!c) return PNB_E_ARG;
    if (n && (!particles || !out)) return fail(c, PNB_E_ARG, "null pointer with non-zero length");
    if (n == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    if (!c->bh) c->bh = new BhState();
    BhState* s = c->bh;
    if (int rc = copy_h2d(c, s->h_aff, particles, n)) return rc;
    PNB_HIP(c, s->h_out.ensure(n * sizeof(float4)));
    if (int rc = pnb_reordered_bh_accel_f32_dev(c, s->h_aff.as<float>(), n, theta, softening,
                                                checked, flags, s->h_out.as<float>()))
        return rc;
    PNB_HIP(c, hipMemcpyAsync(out, s->h_out.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_bh_set_relative(pnb_ctx* c, float alpha, float estimate_theta, int per_particle,
                        int bucket) {
    if (!c) return PNB_E_ARG;
    if (!(alpha >= 0.0f) || !std::isfinite(alpha))
        return fail(c, PNB_E_ARG, "alpha must be finite and >= 0");
    if (!(estimate_theta > 0.0f) || !std::isfinite(estimate_theta))
        return fail(c, PNB_E_ARG, "estimate_theta must be finite and > 0");
    if (bucket < 0 || bucket > 31) return fail(c, PNB_E_ARG, "bucket must be in 0 .. 31");
    if (!c->bh) c->bh = new BhState();
    c->bh->rel_alpha = alpha;
    c->bh->rel_theta_est = estimate_theta;
    c->bh->rel_per_particle = per_particle != 0;
    c->bh->rel_bucket = bucket;
    return PNB_OK;
}

int pnb_bh_set_capturable(pnb_ctx* c, int enable) {
    if (!c) return PNB_E_ARG;
    if (!c->bh) c->bh = new BhState();
    BhState* s = c->bh;
    if (int rc = activate(c)) return rc;
    (void)dd_resolve(c);  // takes the last build's status off the device; its failure is not this call's
    s->dd_mode = enable != 0;
    s->dd_cap = 0;  // the next call sizes anew
    s->dd_m = s->dd_n = 0;
    return PNB_OK;
}

int pnb_bh_check(pnb_ctx* c) {
    if (!c) return PNB_E_ARG;
    if (!c->bh) return PNB_OK;
    if (int rc = activate(c)) return rc;
    return dd_resolve(c);
}

int pnb_bh_tree_size(pnb_ctx* c, size_t* nodes, int* max_level) {
    if (!c) return PNB_E_ARG;
    if (c->bh && c->bh->dd_last) {
        if (int rc = activate(c)) return rc;
        if (int rc = dd_resolve(c)) return rc;
    }
    if (!c->bh || !c->bh->built) return fail(c, PNB_E_STATE, "no octree has been built");
    if (nodes) *nodes = c->bh->nodes;
    if (max_level) *max_level = c->bh->max_level;
    return PNB_OK;
}

int pnb_bh_node_features(pnb_ctx* c, float* out, size_t cap) {
    if (!c) return PNB_E_ARG;
    if (c->bh && c->bh->dd_last) {
        if (int rc = activate(c)) return rc;
        if (int rc = dd_resolve(c)) return rc;
    }
    BhState* s = c->bh;
    if (!s || !s->built) return fail(c, PNB_E_STATE, "no octree has been built");
    if (!s->has_moments) return fail(c, PNB_E_STATE, "the last octree was built without moments");
    if (cap < s->nodes || (s->nodes && !out)) return fail(c, PNB_E_ARG, "output too small");
    if (int rc = activate(c)) return rc;
    if (s->nodes == 0) return PNB_OK;
    const size_t bytes = s->nodes * MLP_FEAT * sizeof(float);
    PNB_HIP(c, s->feat_out.ensure(bytes));
    bh_features_kernel<<<div_up(s->nodes, size_t(256)), 256, 0, c->stream>>>(
        s->rec.as<NodeRec>(), s->quad.as<NodeQuad>(), s->level.as<uint8_t>(),
        s->desc.as<NodeDesc>(), s->nodes, s->feat_out.as<float>());
    PNB_CHECK_LAUNCH(c);
    PNB_HIP(c, hipMemcpyAsync(out, s->feat_out.ptr, bytes, hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_bh_tree_export(pnb_ctx* c, pnb_tree_node* out, size_t cap, uint32_t* order, size_t m) {
    if (!c) return PNB_E_ARG;
    if (c->bh && c->bh->dd_last) {
        if (int rc = activate(c)) return rc;
        if (int rc = dd_resolve(c)) return rc;
    }
    BhState* s = c->bh;
    if (!s || !s->built) return fail(c, PNB_E_STATE, "no octree has been built");
    if (cap < s->nodes || (order && m < s->m)) return fail(c, PNB_E_ARG, "output too small");
    if (int rc = activate(c)) return rc;
    const size_t T = s->nodes;
    if (T) {
        NodeRec* rec = new NodeRec[T];
        uint8_t* lev = new uint8_t[T];
        float* wid = new float[T];
        NodeSums* sm = new NodeSums[T];
        NodeQuad* qd = s->has_moments ? new NodeQuad[T] : nullptr;
        hipError_t e = hipMemcpyAsync(rec, s->rec.ptr, T * sizeof(NodeRec), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(lev, s->level.ptr, T, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(wid, s->width.ptr, T * 4, hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(sm, s->sums.ptr, T * sizeof(NodeSums), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess && qd)
            e = hipMemcpyAsync(qd, s->quad.ptr, T * sizeof(NodeQuad), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess)
            for (size_t k = 0; k < T; ++k) {
                pnb_tree_node& o = out[k];
                o.com[0] = rec[k].cx; o.com[1] = rec[k].cy; o.com[2] = rec[k].cz;
                o.mass = rec[k].mass;
                o.mass64 = sm[k].m.x;
                o.width = wid[k];
                o.skip = rec[k].skip;
                o.begin = rec[k].begin;
                o.count = rec[k].end - rec[k].begin;
                o.level = lev[k];
                o.internal = rec[k].w2 >= 0.0f ? 1u : 0u;
                const float q6[6] = {qd ? qd[k].a.x : 0.f, qd ? qd[k].a.y : 0.f,
                                     qd ? qd[k].a.z : 0.f, qd ? qd[k].a.w : 0.f,
                                     qd ? qd[k].b.x : 0.f, qd ? qd[k].b.y : 0.f};
                for (int q = 0; q < 6; ++q) o.quad[q] = q6[q];
            }
        delete[] rec;
        delete[] lev;
        delete[] wid;
        delete[] sm;
        delete[] qd;
        if (e != hipSuccess) return hip_fail(c, e, "tree export copy", __FILE__, __LINE__);
    }
    if (order && s->m)
        PNB_HIP(c, hipMemcpy(order, s->sidx, s->m * 4, hipMemcpyDeviceToHost));
    return PNB_OK;
}

int pnb_bh_stats(pnb_ctx* c, uint64_t* wave_steps, uint64_t* accepted) {
    if (!c) return PNB_E_ARG;
    if (!c->bh || !c->bh->built) return fail(c, PNB_E_STATE, "no octree has been built");
    if (int rc = activate(c)) return rc;
    BhInfo h;
    PNB_HIP(c, hipMemcpyAsync(&h, c->bh->info.ptr, sizeof(h), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    if (wave_steps) *wave_steps = h.steps;
    if (accepted) *accepted = h.accepted;
    return PNB_OK;
}

int pnb_mlp_collect(pnb_ctx* c, const float* particles, size_t n, float theta, float softening,
                    size_t waves, uint64_t seed, size_t row_cap) {
    if (!c) return PNB_E_ARG;
    if (!particles || n == 0 || waves == 0) return fail(c, PNB_E_ARG, "no particles or waves");
    if (row_cap == 0 || row_cap > 65536) return fail(c, PNB_E_ARG, "row_cap must be in 1..65536");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (int rc = activate(c)) return rc;
    return collect(c, reinterpret_cast<const float4*>(particles), n, theta, softening, waves, seed,
                   uint32_t(row_cap));
}

}  // extern "C"
