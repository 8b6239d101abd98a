// particular_amd C++ host interface: the reference's Interaction/storage API over the C ABI.
//
// Mirrors particular 0.8 (Rust) for 3D f32 particles:
//   Between<S1, S2>                     particular/src/lib.rs:299-300
//   Interaction<Storage>::compute       particular/src/lib.rs:364-370
//   Acceleration<CHECKED>               particular/src/gravity/newtonian/acceleration.rs:17-44
//   AccelerationSoftened<CHECKED>       particular/src/gravity/newtonian/acceleration_softened.rs:17-49
//   Ordered / Reordered                 particular/src/storage.rs:51-205
//   GpuResources -> hip::HipResources   particular/src/gpu/mod.rs:84-143
//   gpu::BruteForce -> hip::BruteForce  particular/src/gpu/mod.rs:149-208
//   parallel::BarnesHut -> hip::BarnesHut  particular/src/parallel.rs:298-367
//   GpuCompute sugar -> hip_brute_force / hip_barnes_hut  particular/src/gpu/mod.rs:13-37
// Storages lower to Between as the reference's blanket impls do (storage.rs:207-241). Results
// are returned eagerly as std::vector<Vec3>, like gpu::BruteForce's Vec::IntoIter. Failures throw
// particular_amd::Error (the reference panics).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "particular_amd.h"

namespace particular_amd {

struct Vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
};

// GravitationalField<Vec3, f32> (particular/src/gravity/mod.rs:12-18): 16 bytes, C layout.
struct GravitationalField {
    Vec3 position;
    float m = 0.f;
    bool is_affecting() const { return m != 0.f; }  // gravity/mod.rs:27-34
};
static_assert(sizeof(GravitationalField) == 16, "particle record is 16 bytes");

template <class S1, class S2>
struct Between {
    S1 affected;
    S2 affecting;
};

template <class T>
struct Slice {
    const T* data = nullptr;
    size_t len = 0;
    Slice() = default;
    Slice(const T* d, size_t n) : data(d), len(n) {}
    Slice(const std::vector<T>& v) : data(v.data()), len(v.size()) {}  // NOLINT
};

class Error : public std::runtime_error {
   public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

   private:
    int status_;
};

template <bool CHECKED>
struct Acceleration {
    static constexpr bool checked_v = CHECKED;
    float softening() const { return 0.f; }
    static Acceleration<true> checked() { return {}; }
    static Acceleration<false> unchecked() { return {}; }
};

template <bool CHECKED>
struct AccelerationSoftened {
    static constexpr bool checked_v = CHECKED;
    float eps = 0.f;
    float softening() const { return eps; }
    static AccelerationSoftened<true> checked(float s) { return {s}; }
    static AccelerationSoftened<false> unchecked(float s) { return {s}; }
};

// Affecting particles first, each group in original order (storage.rs:51-138).
class Ordered {
   public:
    template <class F>
    static Ordered make(Slice<GravitationalField> unordered, F is_affecting) {
        Ordered o;
        for (size_t i = 0; i < unordered.len; ++i)
            if (is_affecting(unordered.data[i])) o.particles_.push_back(unordered.data[i]);
        o.affecting_len_ = o.particles_.size();
        for (size_t i = 0; i < unordered.len; ++i)
            if (!is_affecting(unordered.data[i])) o.particles_.push_back(unordered.data[i]);
        return o;
    }
    static Ordered make(Slice<GravitationalField> unordered) {
        return make(unordered, [](const GravitationalField& p) { return p.is_affecting(); });
    }
    size_t affecting_len() const { return affecting_len_; }
    Slice<GravitationalField> affecting() const { return {particles_.data(), affecting_len_}; }
    Slice<GravitationalField> non_affecting() const {
        return {particles_.data() + affecting_len_, particles_.size() - affecting_len_};
    }
    Slice<GravitationalField> particles() const { return {particles_.data(), particles_.size()}; }

   private:
    std::vector<GravitationalField> particles_;
    size_t affecting_len_ = 0;
};

// The original particles plus an Ordered copy (storage.rs:140-205). With the default affecting
// test the split is done on the GPU and the copy is never made.
class Reordered {
   public:
    explicit Reordered(Slice<GravitationalField> unordered) : unordered(unordered) {}
    Slice<GravitationalField> unordered;
    const Ordered& ordered() const {
        if (!ordered_built_) {
            ordered_ = Ordered::make(unordered);
            ordered_built_ = true;
        }
        return ordered_;
    }
    Slice<GravitationalField> affecting() const { return ordered().affecting(); }

   private:
    mutable Ordered ordered_;
    mutable bool ordered_built_ = false;
};

namespace hip {

// GpuResources analogue: one context (device, stream, grow-only buffers). Reuse it across calls.
class HipResources {
   public:
    explicit HipResources(int device = 0) {
        int rc = pnb_create(device, &ctx_);
        if (rc != PNB_OK) throw Error(rc, pnb_status_string(rc));
    }
    ~HipResources() { pnb_destroy(ctx_); }
    HipResources(const HipResources&) = delete;
    HipResources& operator=(const HipResources&) = delete;
    pnb_ctx* get() const { return ctx_; }
    void check(int rc) const {
        if (rc != PNB_OK) throw Error(rc, pnb_last_error(ctx_));
    }
    // Barnes-Hut without host round trips for device-pointer calls (pnb_bh_set_capturable), and
    // the status of the last such build (pnb_bh_check).
    void set_bh_capturable(bool enable) { check(pnb_bh_set_capturable(ctx_, enable ? 1 : 0)); }
    void bh_check() { check(pnb_bh_check(ctx_)); }

   private:
    pnb_ctx* ctx_ = nullptr;
};

namespace detail {
inline std::vector<Vec3> unpad(const std::vector<float>& out) {
    std::vector<Vec3> v(out.size() / 4);
    for (size_t i = 0; i < v.size(); ++i) v[i] = {out[4 * i], out[4 * i + 1], out[4 * i + 2]};
    return v;
}
inline const float* f(const Slice<GravitationalField>& s) {
    return reinterpret_cast<const float*>(s.data);
}
}  // namespace detail

// gpu::BruteForce replacement (particular/src/gpu/mod.rs:149-208).
template <class T>
struct BruteForce {
    HipResources& resources;
    T interaction;

    std::vector<Vec3> compute(Between<Slice<GravitationalField>, Slice<GravitationalField>> b) {
        std::vector<float> out(4 * b.affected.len);
        resources.check(pnb_bf_accel_f32(resources.get(), detail::f(b.affected), b.affected.len,
                                         detail::f(b.affecting), b.affecting.len,
                                         interaction.softening(), T::checked_v, out.data()));
        return detail::unpad(out);
    }
    std::vector<Vec3> compute(Slice<GravitationalField> s) { return compute({s, s}); }
    std::vector<Vec3> compute(const Ordered& o) { return compute({o.particles(), o.affecting()}); }
    std::vector<Vec3> compute(const Reordered& r) {
        std::vector<float> out(4 * r.unordered.len);
        resources.check(pnb_reordered_bf_accel_f32(resources.get(), detail::f(r.unordered),
                                                   r.unordered.len, interaction.softening(),
                                                   T::checked_v, out.data()));
        return detail::unpad(out);
    }
};

// parallel::BarnesHut replacement (particular/src/parallel.rs:298-367). relative_alpha >= 0 adds
// the relative opening criterion (PNB_BH_RELATIVE, no reference counterpart); the default keeps
// the reference's criterion alone.
template <class T>
struct BarnesHut {
    HipResources& resources;
    float theta;
    T interaction;
    float relative_alpha = -1.0f;
    float estimate_theta = 1.2f;
    bool estimate_per_particle = false;
    int bucket = 24;

    unsigned flags() {
        if (relative_alpha < 0.0f) return 0u;
        resources.check(pnb_bh_set_relative(resources.get(), relative_alpha, estimate_theta,
                                            estimate_per_particle ? 1 : 0, bucket));
        return PNB_BH_RELATIVE;
    }

    std::vector<Vec3> compute(Between<Slice<GravitationalField>, Slice<GravitationalField>> b) {
        std::vector<float> out(4 * b.affected.len);
        const unsigned fl = flags();
        resources.check(pnb_bh_accel_f32(resources.get(), detail::f(b.affected), b.affected.len,
                                         detail::f(b.affecting), b.affecting.len, theta,
                                         interaction.softening(), T::checked_v, fl, out.data()));
        return detail::unpad(out);
    }
    std::vector<Vec3> compute(Slice<GravitationalField> s) { return compute({s, s}); }
    std::vector<Vec3> compute(const Ordered& o) { return compute({o.particles(), o.affecting()}); }
    std::vector<Vec3> compute(const Reordered& r) {
        std::vector<float> out(4 * r.unordered.len);
        const unsigned fl = flags();
        resources.check(pnb_reordered_bh_accel_f32(resources.get(), detail::f(r.unordered),
                                                   r.unordered.len, theta,
                                                   interaction.softening(), T::checked_v, fl,
                                                   out.data()));
        return detail::unpad(out);
    }
};

}  // namespace hip

// GpuCompute sugar (particular/src/gpu/mod.rs:19-30).
template <class S, class T>
std::vector<Vec3> hip_brute_force(const S& storage, hip::HipResources& res, T interaction) {
    return hip::BruteForce<T>{res, interaction}.compute(storage);
}

template <class S, class T>
std::vector<Vec3> hip_barnes_hut(const S& storage, hip::HipResources& res, float theta,
                                 T interaction) {
    return hip::BarnesHut<T>{res, theta, interaction}.compute(storage);
}

}  // namespace particular_amd
