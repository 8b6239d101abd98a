// Host-side logic of the C++ interface and the C ABI's argument checks, without a GPU: the
// storage lowering (Ordered / Reordered, particular/src/storage.rs:51-205), the no-device error
// path of HipResources, and the null-context / null-pointer checks of the entry points. Built with
// -fsanitize=address,undefined on the host (`make -C tests/cpp san`) and run by
// tests/test_sanitizers.py; exit status 0 when every expectation holds.
#include "particular_amd.hpp"

#include <cstdio>
#include <vector>

using namespace particular_amd;

static int failures = 0;
static void expect(bool ok, const char* what) {
    if (!ok) {
        std::printf("FAIL %s\n", what);
        ++failures;
    }
}

int main() {
    // Ordered: affecting first, each group in its original order
    std::vector<GravitationalField> ps;
    for (int i = 0; i < 97; ++i)
        ps.push_back({{float(i), float(-i), 0.5f * i}, (i % 3 == 0) ? 0.0f : float(i)});
    const Ordered o = Ordered::make(Slice<GravitationalField>(ps));
    expect(o.affecting_len() == 64, "affecting count");
    expect(o.particles().len == ps.size(), "all particles kept");
    float last = -1;
    for (size_t i = 0; i < o.affecting().len; ++i) {
        expect(o.affecting().data[i].m != 0.0f, "affecting group");
        expect(o.affecting().data[i].position.x > last, "affecting group order");
        last = o.affecting().data[i].position.x;
    }
    for (size_t i = 0; i < o.non_affecting().len; ++i)
        expect(o.non_affecting().data[i].m == 0.0f, "non-affecting group");
    const Ordered custom = Ordered::make(Slice<GravitationalField>(ps),
                                         [](const GravitationalField& p) { return p.position.x > 50; });
    expect(custom.affecting_len() == 46, "custom predicate");
    Reordered r{Slice<GravitationalField>(ps)};
    expect(r.affecting().len == 64 && r.unordered.len == ps.size(), "Reordered lowering");
    const Ordered empty = Ordered::make(Slice<GravitationalField>());
    expect(empty.particles().len == 0 && empty.affecting_len() == 0, "empty storage");

    // without a GPU: context creation reports the missing device; every entry point refuses a
    // null context
    int count = -1;
    const int dc = pnb_device_count(&count);
    if (dc == PNB_E_NODEV) {
        expect(count == 0, "no device count");
        try {
            hip::HipResources res(0);
            expect(false, "HipResources without a device");
        } catch (const Error& e) {
            expect(e.status() == PNB_E_NODEV, "HipResources reports PNB_E_NODEV");
        }
    }
    float out[8] = {};
    expect(pnb_bf_accel_f32(nullptr, nullptr, 0, nullptr, 0, 0.f, 1, nullptr) == PNB_E_ARG, "bf null ctx");
    expect(pnb_bf_pairs_accel_f32(nullptr, nullptr, 1, 0.f, 1, out) == PNB_E_ARG, "pairs null ctx");
    expect(pnb_bh_accel_f32(nullptr, nullptr, 0, nullptr, 0, 0.5f, 0.f, 1, 0, nullptr) == PNB_E_ARG,
           "bh null ctx");
    expect(pnb_bf_accel_f32_dev_noalloc(nullptr, nullptr, 0, nullptr, 0, 0.f, 1, nullptr) == PNB_E_ARG,
           "noalloc null ctx");
    expect(pnb_synchronize(nullptr) == PNB_E_ARG, "synchronize null ctx");
    expect(pnb_status_string(PNB_E_DEPTH) != nullptr, "status string");
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
