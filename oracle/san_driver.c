/*
 * Exercises every oracle entry point on small and edge-case inputs, for a build with
 * -fsanitize=address,undefined (`make -C oracle san`; run by tests/test_sanitizers.py). Test
 * infrastructure only. Exit status 0 when every call returned its expected status; the sanitizers
 * abort the process on the first report.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static unsigned long long rng_state = 0x9e3779b97f4a7c15ull;
static float frand(void) { /* splitmix64 to [-1, 1) */
    unsigned long long z = (rng_state += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    z ^= z >> 31;
    return (float)((double)(z >> 11) / 9007199254740992.0 * 2.0 - 1.0);
}

static int failures = 0;
static void expect(int ok, const char *what) {
    if (!ok) {
        printf("FAIL %s\n", what);
        ++failures;
    }
}

static float *particles(size_t n, int massless_every, int dup_every) {
    float *p = calloc(4 * (n ? n : 1), sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        if (dup_every && i > 0 && i % (size_t)dup_every == 0) {
            memcpy(p + 4 * i, p + 4 * (i - 1), 4 * sizeof(float));
            continue;
        }
        p[4 * i] = frand(); p[4 * i + 1] = frand(); p[4 * i + 2] = frand();
        p[4 * i + 3] = (massless_every && i % (size_t)massless_every == 0) ? 0.0f : 0.5f + 0.5f * frand();
    }
    return p;
}

static void tree_case(const float *p, size_t m, int max_depth, int expect_tree) {
    orc_tree *t = orc_tree_build(p, m, max_depth);
    expect((t != NULL) == expect_tree, "tree build status");
    if (!t) return;
    const size_t nodes = orc_tree_num_nodes(t);
    const orc_node *nd = orc_tree_nodes(t);
    const uint32_t *order = orc_tree_order(t);
    uint32_t *skips = calloc(nodes ? nodes : 1, sizeof(uint32_t));
    orc_tree_skips(t, skips);
    for (size_t k = 0; k < nodes; ++k) expect(skips[k] > k && skips[k] <= nodes, "skip range");
    for (size_t i = 0; i < m; ++i) expect(order[i] < m, "order range");
    if (nodes) expect(nd[0].count == (int32_t)m, "root holds every body");
    float lo[3], hi[3];
    orc_tree_root_box(t, lo, hi);
    (void)orc_tree_max_level(t);
    const size_t n = m < 300 ? m : 300;
    float *out = calloc(4 * (n ? n : 1), sizeof(float));
    uint32_t *visits = calloc(n ? n : 1, sizeof(uint32_t));
    float *est = calloc(n ? n : 1, sizeof(float));
    for (size_t i = 0; i < n; ++i) est[i] = 1e-3f;
    const float thetas[] = {0.0f, 0.5f, 1.0f};
    for (int k = 0; k < 3; ++k) {
        expect(orc_bh_f32(t, p, n, thetas[k], 0.0f, 1, 0, 0, out, visits, 2) == 0, "bh f32");
        expect(orc_bh_f32(t, p, n, thetas[k], 1e-3f, 1, 1, 1, out, NULL, 1) == 0, "bh com64");
        expect(orc_bh_rel_f32(t, p, n, thetas[k], 0.0f, est, 2, 24, p, 1, 0, out, visits, 2) == 0,
               "bh rel p2");
        expect(orc_bh_rel_f32(t, p, n, thetas[k], 1e-3f, est, 3, 0, p, 0, 1, out, NULL, 1) == 0,
               "bh rel p3");
    }
    free(est); free(visits); free(out); free(skips);
    orc_tree_free(t);
}

int main(void) {
    const size_t sizes[] = {0, 1, 2, 3, 17, 257, 1000};
    for (size_t s = 0; s < sizeof sizes / sizeof sizes[0]; ++s) {
        const size_t n = sizes[s];
        float *p = particles(n, 5, 7);
        float *out = calloc(4 * (n ? n : 1), sizeof(float));
        double *out64 = calloc(3 * (n ? n : 1), sizeof(double));
        expect(orc_bf_f32(p, n, p, n, 0.0f, 1, out, 2) == 0, "bf checked");
        expect(orc_bf_f32(p, n, p, n, 1e-2f, 0, out, 1) == 0, "bf softened unchecked");
        expect(orc_direct_f64(p, n, p, n, 0.0, out64, 2) == 0, "direct f64");
        expect(orc_bf_pairs_f32(p, n, n, 0.0f, 1, out) == 0, "pairs checked");
        expect(orc_bf_pairs_f32(p, n, n / 2, 1e-3f, 1, out) == 0, "pairs restricted");
        expect(orc_bf_pairs_f32(p, n, n + 1, 0.0f, 1, out) == 1, "pairs restrict_len > n refused");
        if (n) {
            float a[3];
            orc_accel_pair(p, p + 4 * (n - 1), 0.0f, 1, a);
        }
        tree_case(p, n, 300, 1);
        free(out64); free(out); free(p);
    }
    /* all bodies at one point (a single leaf), all massless, and a deep cluster */
    float *same = particles(40, 0, 1);
    tree_case(same, 40, 300, 1);
    free(same);
    float *massless = particles(100, 1, 0);
    tree_case(massless, 100, 300, 1);
    free(massless);
    float *deep = particles(50, 0, 0);
    for (int i = 0; i < 10; ++i) {
        deep[4 * i] = (float)i * 1e-30f; deep[4 * i + 1] = 0.0f; deep[4 * i + 2] = 0.0f;
    }
    tree_case(deep, 50, 300, 1);
    tree_case(deep, 50, 8, 0); /* deeper than max_depth: refused */
    free(deep);
    printf("%s (%d failures)\n", failures ? "FAILED" : "ok", failures);
    return failures ? 1 : 0;
}
