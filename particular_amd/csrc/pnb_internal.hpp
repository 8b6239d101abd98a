// Internal declarations shared by the particular_amd HIP sources (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/particular_amd.h"

namespace pnb {

// Grow-only device allocation owned by a context (the reference re-allocates per call,
// particular/src/gpu/resources.rs:27-34; here buffers only grow).
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes);
    void release();
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

struct BfTune;   // brute-force kernel shape table (bf.hip)
struct BhState;  // Barnes-Hut build state (bh.hip)
struct StepState;   // leapfrog loop state (step.hip)
struct BatchState;  // batched ensemble state (batch.hip)
struct MlpState;    // learned far-field correction (mlp.hip)

}  // namespace pnb

// The context: the GpuResources analogue (particular/src/gpu/mod.rs:84-143). One device, one
// stream, grow-only buffers. One thread at a time per context.
struct pnb_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;  // own_stream unless pnb_set_stream installed another
    std::string last_error;
    int cu_count = 0;

    // scratch for host-pointer entry points and the massive/massless split
    pnb::DevBuf h_affected, h_affecting, h_out, split_tmp, split_flags, bf_partial;
    pnb::DevBuf scan_tmp;
    pnb::DevBuf pairs_part;  // pair brute force: one partial per body and super-tile
    // device address of the allocation-free entry point's static scratch on this device
    float4* noalloc_scratch = nullptr;
    // two-dimensional entry points: {x, y, mu} records staged, widened to float4, results narrowed
    pnb::DevBuf d2_raw_a, d2_raw_b, d2_a, d2_b, d2_out4, d2_out2;

    // per-kernel timing (enabled by pnb_profile_enable): event pairs recorded on the stream around
    // each launch, resolved lazily so that profiling does not serialise host and device
    bool profile = false;
    bool capturing = false;  // inside hipStreamBeginCapture: no events, no allocation
    double prof_ms[PNB_PROF_NUM_KERNELS] = {};
    uint64_t prof_count[PNB_PROF_NUM_KERNELS] = {};
    static constexpr int EV_POOL = 1024;
    hipEvent_t ev_start[EV_POOL] = {}, ev_stop[EV_POOL] = {};
    int ev_kernel[EV_POOL] = {};
    int ev_used = 0;
    int ev_open = 0;  // slots whose stop event is not recorded yet (nested timed launches)

    pnb::BfTune* bf_tune = nullptr;
    pnb::BhState* bh = nullptr;
    pnb::StepState* step = nullptr;
    pnb::BatchState* batch = nullptr;
    pnb::MlpState* mlp = nullptr;
};

namespace pnb {

int fail(pnb_ctx* c, int code, const char* fmt, ...);
int hip_fail(pnb_ctx* c, hipError_t e, const char* expr, const char* file, int line);
int activate(pnb_ctx* c);  // hipSetDevice(c->device)

// Device-wide exclusive scan of n uint32 values (in place allowed). out[n] receives the total
// when out has room for n + 1 entries and with_total is set.
int scan_u32(pnb_ctx* c, const uint32_t* in, uint32_t* out, size_t n, bool with_total);

// Stable compaction of the affecting particles (mu != 0) of `particles` into `dst`; writes the
// count to *m_dev (device) and returns it in *m_host. Ordered::new (storage.rs:85-95).
int split_affecting(pnb_ctx* c, const float4* particles, size_t n, float4* dst, uint32_t* m_host);

// Brute force on device pointers, asynchronous on c->stream.
int bf_launch(pnb_ctx* c, const float4* affected, size_t n, const float4* affecting, size_t m,
              float softening, int checked, float4* out);

// Folds all recorded event pairs into the per-kernel totals (waits for the last one).
int prof_resolve(pnb_ctx* c);

// Timed launch helper: brackets `launch` with events when profiling is on (not while capturing).
template <class F>
int timed(pnb_ctx* c, int kernel_id, F&& launch) {
    if (!c->profile || c->capturing) return launch();
    if (c->ev_used == pnb_ctx::EV_POOL && c->ev_open == 0)
        if (int rc = prof_resolve(c)) return rc;
    if (c->ev_used == pnb_ctx::EV_POOL) return launch();  // pool full inside a nested launch
    const int k = c->ev_used++;
    c->ev_kernel[k] = kernel_id;
    c->ev_open++;
    hipError_t e = hipEventRecord(c->ev_start[k], c->stream);
    int rc = (e == hipSuccess) ? launch() : hip_fail(c, e, "hipEventRecord", __FILE__, __LINE__);
    if (rc == PNB_OK) {
        e = hipEventRecord(c->ev_stop[k], c->stream);
        if (e != hipSuccess) rc = hip_fail(c, e, "hipEventRecord", __FILE__, __LINE__);
    }
    c->ev_open--;
    if (rc != PNB_OK) c->ev_used = k;  // drop this slot and any nested ones
    return rc;
}

// Looks up the allocation-free brute force's static scratch on c->device (once per device).
int bf_noalloc_init(pnb_ctx* c);

void bf_tune_free(BfTune*);
void bh_free(BhState*);
void step_free(StepState*);
void batch_free(BatchState*);
void mlp_free(MlpState*);

inline unsigned div_up(size_t a, size_t b) { return static_cast<unsigned>((a + b - 1) / b); }

}  // namespace pnb

#define PNB_HIP(c, expr)                                                         \
    do {                                                                         \
        hipError_t e_ = (expr);                                                  \
        if (e_ != hipSuccess) return pnb::hip_fail((c), e_, #expr, __FILE__, __LINE__); \
    } while (0)

#define PNB_CHECK_LAUNCH(c) PNB_HIP((c), hipGetLastError())
