// The reference's own algorithm tests, ported to the C++ host interface over the HIP backend:
//   acceleration_error!  particular/src/gravity/newtonian/mod.rs:228-277
//   circular_orbit!      particular/src/gravity/newtonian/mod.rs:281-347
//   instantiations       particular/src/gravity/newtonian/mod.rs:382-423 (tolerances), :511-518
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "particular_amd.hpp"

using namespace particular_amd;

static int failures = 0;

static void expect(bool ok, const char* what, double value, double bound) {
    std::printf("%-48s %.3e (bound %.1e) %s\n", what, value, bound, ok ? "ok" : "FAIL");
    if (!ok) ++failures;
}

template <class Alg>
static void acceleration_error(Alg& cm, float epsilon, const char* name) {
    const std::vector<GravitationalField> massive = {
        {{0, 0, 0}, 20}, {{1, 1, 1}, 30}, {{-3, -3, -3}, 40}};
    const std::vector<GravitationalField> particles = {
        {{10, 10, 10}, 0}, massive[0], massive[1], massive[2], {{30, 30, 30}, 0},
        {{-45, -45, -45}, 0}};
    Reordered reordered{Slice<GravitationalField>(particles)};
    const std::vector<Vec3> computed = cm.compute(reordered);
    double worst = 0;
    for (size_t i = 0; i < particles.size(); ++i) {
        // the test's own expected value, in the same precision (mod.rs:253-262)
        Vec3 acc;
        for (const auto& p2 : massive) {
            const Vec3 p1 = particles[i].position;
            const Vec3 dir{p2.position.x - p1.x, p2.position.y - p1.y, p2.position.z - p1.z};
            const float mag2 = dir.x * dir.x + dir.y * dir.y + dir.z * dir.z;
            if (mag2 != 0.0f) {
                const float f = p2.m * std::sqrt(1.0f / mag2) / mag2;
                acc.x += dir.x * f; acc.y += dir.y * f; acc.z += dir.z * f;
            }
        }
        // error = |splat(1) - computed / acceleration| (mod.rs:268-271)
        const float ex = 1.0f - computed[i].x / acc.x, ey = 1.0f - computed[i].y / acc.y,
                    ez = 1.0f - computed[i].z / acc.z;
        const double err = std::sqrt(double(ex * ex + ey * ey + ez * ez));
        if (err > worst) worst = err;
    }
    char what[96];
    std::snprintf(what, sizeof what, "acceleration_error %s", name);
    expect(worst <= epsilon, what, worst, epsilon);
}

template <class Alg>
static void circular_orbit(Alg& cm, int orbit_count, float epsilon, const char* name) {
    const float dt = 1.0f / 60.0f;
    std::vector<GravitationalField> particles = {{{0, 0, 0}, 1e6f}, {{100, 0, 0}, 0}};
    std::vector<Vec3> velocities = {{0, 0, 0}, {0, 100, 0}};
    auto distance = [&]() {
        const float dx = particles[0].position.x - particles[1].position.x;
        const float dy = particles[0].position.y - particles[1].position.y;
        const float dz = particles[0].position.z - particles[1].position.z;
        return std::sqrt(dx * dx + dy * dy + dz * dz);
    };
    auto energy = [&](float r) { return -(particles[0].m + particles[1].m) / (r + r); };
    const float d0 = distance(), e0 = energy(d0);
    const float period = 6.2831855f * std::sqrt((d0 * d0 * d0) / particles[0].m);
    const long steps = std::lround(period / dt) * orbit_count;
    for (long s = 0; s < steps; ++s) {
        const std::vector<Vec3> a = cm.compute(Slice<GravitationalField>(particles));
        for (size_t i = 0; i < 2; ++i) {
            velocities[i].x += a[i].x * dt; velocities[i].y += a[i].y * dt; velocities[i].z += a[i].z * dt;
        }
        for (size_t i = 0; i < 2; ++i) {
            particles[i].position.x += velocities[i].x * dt;
            particles[i].position.y += velocities[i].y * dt;
            particles[i].position.z += velocities[i].z * dt;
        }
    }
    const float d1 = distance(), e1 = energy(d1);
    char what[96];
    std::snprintf(what, sizeof what, "circular_orbit distance %s", name);
    expect(std::fabs(1.0f - d0 / d1) < epsilon, what, std::fabs(1.0f - d0 / d1), epsilon);
    std::snprintf(what, sizeof what, "circular_orbit energy %s", name);
    expect(std::fabs(1.0f - e0 / e1) < epsilon, what, std::fabs(1.0f - e0 / e1), epsilon);
}

int main(int argc, char** argv) {
    const int orbits = argc > 1 ? std::atoi(argv[1]) : 100;
    try {
        hip::HipResources res(0);
        hip::BruteForce<Acceleration<true>> bf{res, Acceleration<true>::checked()};
        acceleration_error(bf, 1e-2f, "hip::BruteForce");
        circular_orbit(bf, orbits, 1e-2f, "hip::BruteForce");
        hip::BarnesHut<Acceleration<true>> bh0{res, 0.0f, Acceleration<true>::checked()};
        acceleration_error(bh0, 1e-2f, "hip::BarnesHut(0)");
        circular_orbit(bh0, orbits, 1e-2f, "hip::BarnesHut(0)");
        hip::BarnesHut<Acceleration<true>> bh5{res, 0.5f, Acceleration<true>::checked()};
        acceleration_error(bh5, 5e-1f, "hip::BarnesHut(0.5)");
        circular_orbit(bh5, orbits, 1e-1f, "hip::BarnesHut(0.5)");
        // the relative opening criterion (no reference counterpart), theta = 1 as the floor
        hip::BarnesHut<Acceleration<true>> bhr{res, 1.0f, Acceleration<true>::checked(), 1e-3f};
        acceleration_error(bhr, 5e-1f, "hip::BarnesHut(relative 1e-3)");
        circular_orbit(bhr, orbits, 1e-1f, "hip::BarnesHut(relative 1e-3)");
        // GpuCompute-style sugar over a Reordered storage
        const std::vector<GravitationalField> ps = {{{0, 0, 0}, 2}, {{1, 0, 0}, 0}};
        Reordered r{Slice<GravitationalField>(ps)};
        const auto a = hip_brute_force(r, res, AccelerationSoftened<true>::checked(0.0f));
        expect(std::fabs(a[1].x + 2.0f) < 1e-6f, "hip_brute_force(Reordered) two-body", a[1].x, -2);
    } catch (const Error& e) {
        std::printf("error %d: %s\n", e.status(), e.what());
        return 2;
    }
    return failures ? 1 : 0;
}
