// Kick-drift-kick leapfrog over the brute-force acceleration, with FP64 positions and velocities.
//
// particular has no integrator (particular/README.md:22-25); its tests and examples step with
// symplectic Euler around Interaction::compute (newtonian/mod.rs:318-330,
// examples/pocket-solar-system/src/orbit_prediction.rs:158-170). This loop keeps the state on
// the GPU: x and v in FP64 (the drift accumulates many small increments), an f32 shadow
// {x, y, z, mu} of every particle for the force kernel, and the affecting (mu != 0) particles
// compacted once, in their original order, by a precomputed rank (Ordered::new,
// particular/src/storage.rs:85-95).
//
// One step: v += a dt/2; x += v dt; refresh the shadows; a = brute force; v += a dt/2.
// The steps can be captured once into a hipGraph and replayed (no memset nodes, no allocation
// inside the capture).
#include "pnb_internal.hpp"

namespace pnb {

int bf_prepare(pnb_ctx* c, size_t n, size_t m);

struct StepState {
    size_t n = 0, m = 0;
    float softening = 0.0f;
    int checked = 1;
    DevBuf x, v, acc, shadow, afg, rank, stage;
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

void step_free(StepState* s) {
    if (!s) return;
    s->drop_graph();
    for (DevBuf* b : {&s->x, &s->v, &s->acc, &s->shadow, &s->afg, &s->rank, &s->stage}) b->release();
    delete s;
}

__global__ void lf_pack_kernel(const double* __restrict__ x3, const double* __restrict__ v3,
                               const float* __restrict__ mu, size_t n, double4* __restrict__ x,
                               double4* __restrict__ v, float4* __restrict__ shadow,
                               uint32_t* __restrict__ flags) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x3[3 * i], py = x3[3 * i + 1], pz = x3[3 * i + 2];
    x[i] = make_double4(px, py, pz, 0.0);
    v[i] = make_double4(v3[3 * i], v3[3 * i + 1], v3[3 * i + 2], 0.0);
    shadow[i] = make_float4(float(px), float(py), float(pz), mu[i]);
    flags[i] = mu[i] != 0.0f ? 1u : 0u;
}

__global__ void lf_compact_kernel(const float4* __restrict__ shadow, const uint32_t* __restrict__ rank,
                                  size_t n, float4* __restrict__ afg) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n && shadow[i].w != 0.0f) afg[rank[i]] = shadow[i];
}

__global__ void lf_kick_drift_kernel(double4* __restrict__ x, double4* __restrict__ v,
                                     const float4* __restrict__ acc, float4* __restrict__ shadow,
                                     const uint32_t* __restrict__ rank, float4* __restrict__ afg,
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
    const float mu = shadow[i].w;
    const float4 s = make_float4(float(xi.x), float(xi.y), float(xi.z), mu);
    shadow[i] = s;
    if (mu != 0.0f) afg[rank[i]] = s;
}

__global__ void lf_kick_kernel(double4* __restrict__ v, const float4* __restrict__ acc, size_t n,
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

__global__ void lf_unpack_kernel(const double4* __restrict__ x, const double4* __restrict__ v,
                                 size_t n, double* __restrict__ out) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 a = x[i], b = v[i];
    out[3 * i] = a.x; out[3 * i + 1] = a.y; out[3 * i + 2] = a.z;
    out[3 * (n + i)] = b.x; out[3 * (n + i) + 1] = b.y; out[3 * (n + i) + 2] = b.z;
}

static int enqueue_steps(pnb_ctx* c, StepState* s, size_t steps, double dt) {
    const unsigned g = div_up(s->n, 256);
    const double h = 0.5 * dt;
    for (size_t k = 0; k < steps; ++k) {
        int rc = timed(c, PNB_PROF_DRIFT, [&]() -> int {
            lf_kick_drift_kernel<<<g, 256, 0, c->stream>>>(
                s->x.as<double4>(), s->v.as<double4>(), s->acc.as<float4>(), s->shadow.as<float4>(),
                s->rank.as<uint32_t>(), s->afg.as<float4>(), s->n, h, dt);
            PNB_CHECK_LAUNCH(c);
            return PNB_OK;
        });
        if (rc) return rc;
        rc = bf_launch(c, s->shadow.as<float4>(), s->n, s->afg.as<float4>(), s->m, s->softening,
                       s->checked, s->acc.as<float4>());
        if (rc) return rc;
        lf_kick_kernel<<<g, 256, 0, c->stream>>>(s->v.as<double4>(), s->acc.as<float4>(), s->n, h);
        PNB_CHECK_LAUNCH(c);
    }
    return PNB_OK;
}

}  // namespace pnb

using namespace pnb;

extern "C" {

int pnb_leapfrog_init(pnb_ctx* c, const double* pos3, const double* vel3, const float* mu,
                      size_t n, float softening, int checked) {
    if (!c) return PNB_E_ARG;
    if (n == 0 || !pos3 || !vel3 || !mu) return fail(c, PNB_E_ARG, "leapfrog needs n > 0 particles");
    if (n > 0x7fffffffu) return fail(c, PNB_E_ARG, "too many particles");
    if (!(softening >= 0.0f)) return fail(c, PNB_E_ARG, "softening must be >= 0");
    if (int rc = activate(c)) return rc;
    if (!c->step) c->step = new StepState();
    StepState* s = c->step;
    s->drop_graph();
    s->n = n;
    s->softening = softening;
    s->checked = checked;
    PNB_HIP(c, s->x.ensure(n * sizeof(double4)));
    PNB_HIP(c, s->v.ensure(n * sizeof(double4)));
    PNB_HIP(c, s->acc.ensure(n * sizeof(float4)));
    PNB_HIP(c, s->shadow.ensure(n * sizeof(float4)));
    PNB_HIP(c, s->afg.ensure(n * sizeof(float4)));
    PNB_HIP(c, s->rank.ensure((n + 1) * sizeof(uint32_t)));
    PNB_HIP(c, s->stage.ensure(n * (6 * sizeof(double) + sizeof(float))));
    double* dx = s->stage.as<double>();
    double* dv = dx + 3 * n;
    float* dm = reinterpret_cast<float*>(dv + 3 * n);
    PNB_HIP(c, hipMemcpyAsync(dx, pos3, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(dv, vel3, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PNB_HIP(c, hipMemcpyAsync(dm, mu, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    const unsigned g = div_up(n, 256);
    uint32_t* rank = s->rank.as<uint32_t>();
    lf_pack_kernel<<<g, 256, 0, c->stream>>>(dx, dv, dm, n, s->x.as<double4>(), s->v.as<double4>(),
                                             s->shadow.as<float4>(), rank);
    PNB_CHECK_LAUNCH(c);
    if (int rc = scan_u32(c, rank, rank, n, true)) return rc;
    uint32_t m = 0;
    PNB_HIP(c, hipMemcpyAsync(&m, rank + n, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    s->m = m;
    lf_compact_kernel<<<g, 256, 0, c->stream>>>(s->shadow.as<float4>(), rank, n, s->afg.as<float4>());
    PNB_CHECK_LAUNCH(c);
    if (int rc = bf_prepare(c, n, m)) return rc;
    if (int rc = bf_launch(c, s->shadow.as<float4>(), n, s->afg.as<float4>(), m, softening, checked,
                           s->acc.as<float4>()))
        return rc;
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_leapfrog_run(pnb_ctx* c, size_t steps, double dt, int use_graph) {
    if (!c) return PNB_E_ARG;
    StepState* s = c->step;
    if (!s || s->n == 0) return fail(c, PNB_E_STATE, "pnb_leapfrog_init has not been called");
    if (steps == 0) return PNB_OK;
    if (int rc = activate(c)) return rc;
    if (!use_graph) return enqueue_steps(c, s, steps, dt);
    if (!c->stream) return fail(c, PNB_E_STATE, "graph capture needs a non-null stream");
    if (!s->exec || s->graph_steps != steps || s->graph_dt != dt || s->graph_stream != c->stream) {
        s->drop_graph();
        PNB_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        c->capturing = true;
        int rc = enqueue_steps(c, s, steps, dt);
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

int pnb_leapfrog_read(pnb_ctx* c, double* pos3, double* vel3, float* acc4) {
    if (!c) return PNB_E_ARG;
    StepState* s = c->step;
    if (!s || s->n == 0) return fail(c, PNB_E_STATE, "pnb_leapfrog_init has not been called");
    if (int rc = activate(c)) return rc;
    const size_t n = s->n;
    double* st = s->stage.as<double>();
    lf_unpack_kernel<<<div_up(n, 256), 256, 0, c->stream>>>(s->x.as<double4>(), s->v.as<double4>(), n, st);
    PNB_CHECK_LAUNCH(c);
    if (pos3) PNB_HIP(c, hipMemcpyAsync(pos3, st, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (vel3) PNB_HIP(c, hipMemcpyAsync(vel3, st + 3 * n, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (acc4) PNB_HIP(c, hipMemcpyAsync(acc4, s->acc.ptr, n * sizeof(float4), hipMemcpyDeviceToHost, c->stream));
    PNB_HIP(c, hipStreamSynchronize(c->stream));
    return PNB_OK;
}

int pnb_leapfrog_views(pnb_ctx* c, const float** particles_dev, const float** affecting_dev,
                       const float** acc_dev, size_t* n, size_t* m) {
    if (!c) return PNB_E_ARG;
    StepState* s = c->step;
    if (!s || s->n == 0) return fail(c, PNB_E_STATE, "pnb_leapfrog_init has not been called");
    if (particles_dev) *particles_dev = s->shadow.as<float>();
    if (affecting_dev) *affecting_dev = s->afg.as<float>();
    if (acc_dev) *acc_dev = s->acc.as<float>();
    if (n) *n = s->n;
    if (m) *m = s->m;
    return PNB_OK;
}

}  // extern "C"
