// Stable LSD radix sort of (uint64 key, uint32 value) pairs, 8-bit digits, for gfx950.
//
// One pass per digit, three launches: per-tile digit histograms (digit-major, so one device-wide
// exclusive scan turns them into scatter offsets), the scan, and a scatter in which every
// wavefront ranks its keys with ballots (for each lane, the set of lanes holding the same digit is
// the AND of eight ballot masks; the rank is the popcount of that set below the lane). A tile of
// 2048 keys (the fastest of 1024, 2048 and 4096 at 4M keys) is split into four contiguous wave
// segments processed in order, so the ranking is stable. Digits on which all keys agree are
// skipped (found from one OR/AND reduction).
#include "pnb_internal.hpp"

namespace pnb {

constexpr int RS_BLOCK = 256;
constexpr int RS_WAVES = RS_BLOCK / 64;
constexpr int RS_ROUNDS = 8;                       // 64-key rounds per wave
constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;      // 2048 keys per workgroup
constexpr int RS_WAVE_SPAN = 64 * RS_ROUNDS;       // 512 keys per wave segment

__global__ __launch_bounds__(RS_BLOCK) void rs_hist_kernel(const uint64_t* __restrict__ keys,
                                                           size_t n, int shift,
                                                           uint32_t* __restrict__ hist,
                                                           uint32_t tiles) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = size_t(blockIdx.x) * RS_TILE;
#pragma unroll 4
    for (int k = 0; k < RS_ROUNDS; ++k) {
        const size_t i = base + size_t(k) * RS_BLOCK + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xffu], 1u);
    }
    __syncthreads();
    hist[size_t(threadIdx.x) * tiles + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_BLOCK) void rs_scatter_kernel(
    const uint64_t* __restrict__ keys, const uint32_t* __restrict__ vals, size_t n, int shift,
    const uint32_t* __restrict__ offsets, uint32_t tiles, uint64_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out) {
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint32_t goff[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int d = lane; d < 256; d += 64) wcnt[w][d] = 0;
    goff[threadIdx.x] = offsets[size_t(threadIdx.x) * tiles + blockIdx.x];
    __syncthreads();

    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t seg = size_t(blockIdx.x) * RS_TILE + size_t(w) * RS_WAVE_SPAN;
    uint64_t key[RS_ROUNDS];
    uint32_t val[RS_ROUNDS];
    uint32_t rank[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t i = seg + size_t(r) * 64 + lane;
        const bool valid = i < n;
        key[r] = valid ? keys[i] : 0ull;
        val[r] = valid ? vals[i] : 0u;
        const uint32_t d = uint32_t(key[r] >> shift) & 0xffu;
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1u);
            peers &= ((d >> b) & 1u) ? bal : ~bal;
        }
        const uint32_t before = wcnt[w][d];
        rank[r] = before + uint32_t(__popcll(peers & lt_mask));
        // the lowest lane of each digit group advances the wave's running count for that digit
        if (valid && (peers & lt_mask) == 0) wcnt[w][d] = before + uint32_t(__popcll(peers));
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive prefix over the wave segments, plus the tile's global offset
    {
        const int d = threadIdx.x;
        uint32_t run = goff[d];
#pragma unroll
        for (int ww = 0; ww < RS_WAVES; ++ww) {
            const uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = run;
            run += c;
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t i = seg + size_t(r) * 64 + lane;
        if (i < n) {
            const uint32_t d = uint32_t(key[r] >> shift) & 0xffu;
            const uint32_t pos = wcnt[w][d] + rank[r];
            keys_out[pos] = key[r];
            vals_out[pos] = val[r];
        }
    }
}

__global__ __launch_bounds__(256) void rs_bits_kernel(const uint64_t* __restrict__ keys, size_t n,
                                                      unsigned long long* __restrict__ or_and) {
    unsigned long long o = 0ull, a = ~0ull;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += size_t(gridDim.x) * blockDim.x) {
        o |= keys[i];
        a &= keys[i];
    }
    for (int off = 32; off > 0; off >>= 1) {
        o |= __shfl_xor(o, off, 64);
        a &= __shfl_xor(a, off, 64);
    }
    // one pair of atomics per workgroup (per wavefront, they queued on the two words: 0.1 ms)
    __shared__ unsigned long long wo[4], wa[4];
    if ((threadIdx.x & 63) == 0) {
        wo[threadIdx.x >> 6] = o;
        wa[threadIdx.x >> 6] = a;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < int(blockDim.x >> 6); ++w) {
            o |= wo[w];
            a &= wa[w];
        }
        atomicOr(&or_and[0], o);
        atomicAnd(&or_and[1], a);
    }
}

__global__ void rs_bits_init_kernel(unsigned long long* or_and) {
    or_and[0] = 0ull;
    or_and[1] = ~0ull;
}

// Small inputs: every key's place is the number of keys below it plus the number of equal keys
// before it in the input, which is exactly the stable order of the radix path. One thread per
// key; a wavefront reads the keys it counts against 64 at a time (one per lane, coalesced) and
// broadcasts them lane by lane with v_readlane into scalar registers, so each compare is a
// vector-scalar compare with no memory access. A workgroup counts against at most RANK_SPAN keys:
// inputs above it split the count over blockIdx.y and add the parts with atomics before a
// scatter launch.
constexpr uint32_t SMALL_SORT = 16384;
constexpr uint32_t RANK_SPAN = 1024;

__device__ __forceinline__ uint64_t lane_key(uint64_t v, int k) {
    const uint32_t lo = __builtin_amdgcn_readlane(uint32_t(v), k);
    const uint32_t hi = __builtin_amdgcn_readlane(uint32_t(v >> 32), k);
    return (uint64_t(hi) << 32) | lo;
}

// Keys of at most 63 bits: "kj <= ki" is "kj < ki + 1", so keys before position i count below
// ki + 1 and the others below ki; positions past the span read as ~0, which counts for neither.
__global__ __launch_bounds__(256) void rank_count_kernel(const uint64_t* __restrict__ keys,
                                                         const uint32_t* __restrict__ vals,
                                                         uint32_t n, uint64_t mask,
                                                         uint32_t* __restrict__ rank,
                                                         uint64_t* __restrict__ keys_out,
                                                         uint32_t* __restrict__ vals_out) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wbase = i - lane;  // the wavefront's first position
    const uint64_t ki = i < n ? (keys[i] & mask) : 0ull;
    const uint32_t j0 = blockIdx.y * RANK_SPAN, j1 = min(n, j0 + RANK_SPAN);
    uint32_t cc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // independent counts: no carry chain
    uint64_t next = j0 + lane < j1 ? (keys[j0 + lane] & mask) : ~0ull;
    for (uint32_t t0 = j0; t0 < j1; t0 += 64) {
        const uint64_t kj = next;
        const uint32_t tn = t0 + 64 + lane;  // the next 64 keys load while these are counted
        next = tn < j1 ? (keys[tn] & mask) : ~0ull;
        if (t0 == wbase) {  // the wavefront's own keys: position decides per lane
#pragma unroll
            for (int k = 0; k < 64; ++k)
                cc[k & 7] += lane_key(kj, k) < (uint32_t(k) < lane ? ki + 1 : ki) ? 1u : 0u;
        } else {
            const uint64_t thr = t0 < wbase ? ki + 1 : ki;
#pragma unroll
            for (int k = 0; k < 64; k += 8) {  // eight broadcasts, then eight compares
                uint64_t b[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) b[u] = lane_key(kj, k + u);
#pragma unroll
                for (int u = 0; u < 8; ++u) cc[u] += b[u] < thr ? 1u : 0u;
            }
        }
    }
    const uint32_t c = ((cc[0] + cc[1]) + (cc[2] + cc[3])) + ((cc[4] + cc[5]) + (cc[6] + cc[7]));
    if (i >= n) return;
    if (gridDim.y > 1) {
        atomicAdd(&rank[i], c);
        return;
    }
    keys_out[c] = keys[i];
    vals_out[c] = vals[i];
}

__global__ __launch_bounds__(256) void rank_zero_kernel(uint32_t* __restrict__ rank, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rank[i] = 0u;
}

__global__ __launch_bounds__(256) void rank_scatter_kernel(const uint64_t* __restrict__ keys,
                                                           const uint32_t* __restrict__ vals,
                                                           uint32_t n,
                                                           const uint32_t* __restrict__ rank,
                                                           uint64_t* __restrict__ keys_out,
                                                           uint32_t* __restrict__ vals_out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys_out[rank[i]] = keys[i];
    vals_out[rank[i]] = vals[i];
}

// Sorts keys[0..n) with their values, stably; the result ends in (*keys_out, *vals_out), which is
// either the input pair or the alternate pair. Only bits [0, key_bits) are considered. Up to
// SMALL_SORT pairs (keys of at most 63 bits) are ranked by counting, into the alternate pair;
// either way equal keys keep their input order.
// Digits no two keys differ in are skipped after one read-back to the host, unless all_digits.
int radix_sort_pairs(pnb_ctx* c, uint64_t* keys, uint32_t* vals, uint64_t* keys_alt,
                     uint32_t* vals_alt, size_t n, int key_bits, DevBuf& hist_buf,
                     DevBuf& bits_buf, uint64_t** keys_out, uint32_t** vals_out,
                     bool all_digits) {
    *keys_out = keys;
    *vals_out = vals;
    if (n <= 1) return PNB_OK;
    if (n <= SMALL_SORT && key_bits <= 63) {
        const uint64_t mask = key_bits >= 64 ? ~0ull : ((1ull << key_bits) - 1);
        const uint32_t nn = uint32_t(n), g = div_up(nn, 256u), parts = div_up(nn, RANK_SPAN);
        uint32_t* rank = nullptr;
        if (parts > 1) {
            PNB_HIP(c, hist_buf.ensure(size_t(nn) * sizeof(uint32_t)));
            rank = hist_buf.as<uint32_t>();
            rank_zero_kernel<<<g, 256, 0, c->stream>>>(rank, nn);
            PNB_CHECK_LAUNCH(c);
        }
        rank_count_kernel<<<dim3(g, parts), 256, 0, c->stream>>>(keys, vals, nn, mask, rank,
                                                                keys_alt, vals_alt);
        PNB_CHECK_LAUNCH(c);
        if (parts > 1) {
            rank_scatter_kernel<<<g, 256, 0, c->stream>>>(keys, vals, nn, rank, keys_alt, vals_alt);
            PNB_CHECK_LAUNCH(c);
        }
        *keys_out = keys_alt;
        *vals_out = vals_alt;
        return PNB_OK;
    }
    if (n > 0xffffffffu) return fail(c, PNB_E_ARG, "radix sort supports at most 2^32-1 keys");
    const uint32_t tiles = div_up(n, RS_TILE);
    PNB_HIP(c, hist_buf.ensure((size_t(256) * tiles + 1) * sizeof(uint32_t)));
    // one digit: sorted in one pass, no need to look for it first; all_digits: every pass runs
    // (no read-back, so the sort can be captured in a graph)
    uint64_t varying = all_digits ? ~0ull : 0xffull;
    if (key_bits > 8 && !all_digits) {
        // skip the digits no two keys differ in (found on the device, read back once)
        PNB_HIP(c, bits_buf.ensure(2 * sizeof(unsigned long long)));
        unsigned long long* oa = bits_buf.as<unsigned long long>();
        rs_bits_init_kernel<<<1, 1, 0, c->stream>>>(oa);
        PNB_CHECK_LAUNCH(c);
        rs_bits_kernel<<<std::min<unsigned>(1024, div_up(n, 256)), 256, 0, c->stream>>>(keys, n, oa);
        PNB_CHECK_LAUNCH(c);
        unsigned long long h_oa[2];
        PNB_HIP(c, hipMemcpyAsync(h_oa, oa, sizeof(h_oa), hipMemcpyDeviceToHost, c->stream));
        PNB_HIP(c, hipStreamSynchronize(c->stream));
        varying = h_oa[0] ^ h_oa[1];
    }
    uint32_t* hist = hist_buf.as<uint32_t>();
    uint64_t *kin = keys, *kout = keys_alt;
    uint32_t *vin = vals, *vout = vals_alt;
    for (int shift = 0; shift < key_bits; shift += 8) {
        if (((varying >> shift) & 0xffull) == 0) continue;
        rs_hist_kernel<<<tiles, RS_BLOCK, 0, c->stream>>>(kin, n, shift, hist, tiles);
        PNB_CHECK_LAUNCH(c);
        if (int rc = scan_u32(c, hist, hist, size_t(256) * tiles, false)) return rc;
        rs_scatter_kernel<<<tiles, RS_BLOCK, 0, c->stream>>>(kin, vin, n, shift, hist, tiles, kout, vout);
        PNB_CHECK_LAUNCH(c);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    *keys_out = kin;
    *vals_out = vin;
    return PNB_OK;
}

}  // namespace pnb
