/*
 * CPU oracle for particular's acceleration path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * A restatement, in C, of the reference algorithms for glam::Vec3 / f32:
 *   pair kernel  particular/src/gravity/impls/mod.rs:151-166, impls/glam.rs:232
 *   norm         glam Vec3::length_squared = (x*x + y*y) + z*z (impls/glam.rs:180)
 *   BruteForce   particular/src/sequential.rs:181-194
 *   BruteForcePairs particular/src/sequential.rs:215-305, impls/mod.rs:185-205
 *   COM          particular/src/gravity/impls/mod.rs:120-134
 *   root box     particular/src/tree/partition.rs:121-188
 *   tree build   particular/src/tree/mod.rs:73-139, partition.rs:193-220
 *   BarnesHut    particular/src/sequential.rs:456-506, newtonian/acceleration.rs:121-126
 * Build flags: -O2 -ffp-contract=off, no fast-math (IEEE sqrtf and division, no reassociation).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* glam Vec3::length_squared: self.dot(self) = (x*x + y*y) + z*z, each op rounded. */
static inline float norm2f(float x, float y, float z) {
    float r = x * x;
    r = r + y * y;
    r = r + z * z;
    return r;
}

/* impls/mod.rs:151-166: d = p2 - at; n = |d|^2; if CHECKED && n == 0 { d } else
 * { ns = n + eps*eps; d * (m / (ns * sqrt(ns))) } */
static inline void pair(float px, float py, float pz, float pm, float ax, float ay, float az,
                        float eps, int checked, float *ox, float *oy, float *oz) {
    float dx = px - ax, dy = py - ay, dz = pz - az;
    float n = norm2f(dx, dy, dz);
    if (checked && n == 0.0f) {
        *ox = dx; *oy = dy; *oz = dz;
        return;
    }
    float ns = n + eps * eps;
    float s = pm / (ns * sqrtf(ns));
    *ox = dx * s; *oy = dy * s; *oz = dz * s;
}

void orc_accel_pair(const float *p2, const float *at, float softening, int checked, float *out3) {
    pair(p2[0], p2[1], p2[2], p2[3], at[0], at[1], at[2], softening, checked, &out3[0], &out3[1],
         &out3[2]);
}

static void set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

/* sequential.rs:189-193: affecting.iter().fold(U::default(), |acc, p2| acc + compute(p1, p2)) */
int orc_bf_f32(const float *affected, size_t n, const float *affecting, size_t m, float softening,
               int checked, float *out, int threads) {
    if ((n && !affected) || (m && !affecting) || (n && !out)) return 1;
    set_threads(threads);
    long long nn = (long long)n;
#pragma omp parallel for schedule(dynamic, 64)
    for (long long i = 0; i < nn; ++i) {
        const float *a = affected + 4 * i;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (size_t j = 0; j < m; ++j) {
            const float *p = affecting + 4 * j;
            float ox, oy, oz;
            pair(p[0], p[1], p[2], p[3], a[0], a[1], a[2], softening, checked, &ox, &oy, &oz);
            sx = sx + ox; sy = sy + oy; sz = sz + oz;
        }
        out[4 * i + 0] = sx; out[4 * i + 1] = sy; out[4 * i + 2] = sz; out[4 * i + 3] = 0.0f;
    }
    return 0;
}

/* sequential.rs:232-263 (RestrictedBruteForce::compute) with Acceleration's pair term
 * (newtonian/acceleration.rs:95-109 -> impls/mod.rs:185-205): for i < restrict_len and
 * j = i+1 .. n-1, dir = acceleration_at::<false>(GravitationalField(x_j, 1), x_i, eps), i.e.
 * d * (1 / (r2s * sqrt(r2s))) with no zero-distance check; output_i += dir * mu_j (a separate
 * fold, added to out[i] after the j loop) and out[j] += -dir * mu_i. checked = 1 (not the
 * reference's behaviour, which has no check on this path) drops pairs at zero unsoftened
 * distance instead. Sequential by definition. */
int orc_bf_pairs_f32(const float *p, size_t n, size_t restrict_len, float softening, int checked,
                     float *out) {
    if ((n && (!p || !out)) || restrict_len > n) return 1;
    for (size_t k = 0; k < 4 * n; ++k) out[k] = 0.0f;
    for (size_t i = 0; i < restrict_len; ++i) {
        const float *a = p + 4 * i;
        float sx = 0.0f, sy = 0.0f, sz = 0.0f;
        for (size_t j = i + 1; j < n; ++j) {
            const float *b = p + 4 * j;
            float dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
            float r2 = norm2f(dx, dy, dz);
            if (checked && r2 == 0.0f) continue;
            float ns = r2 + softening * softening;
            float f = 1.0f / (ns * sqrtf(ns));
            float ux = dx * f, uy = dy * f, uz = dz * f;
            sx = sx + ux * b[3]; sy = sy + uy * b[3]; sz = sz + uz * b[3];
            float *o = out + 4 * j;
            o[0] = o[0] + (-ux) * a[3]; o[1] = o[1] + (-uy) * a[3]; o[2] = o[2] + (-uz) * a[3];
        }
        out[4 * i] = out[4 * i] + sx;
        out[4 * i + 1] = out[4 * i + 1] + sy;
        out[4 * i + 2] = out[4 * i + 2] + sz;
    }
    return 0;
}

int orc_direct_f64(const float *affected, size_t n, const float *affecting, size_t m,
                   double softening, double *out3, int threads) {
    if ((n && !affected) || (m && !affecting) || (n && !out3)) return 1;
    set_threads(threads);
    double e2 = softening * softening;
    long long nn = (long long)n;
#pragma omp parallel for schedule(dynamic, 16)
    for (long long i = 0; i < nn; ++i) {
        double ax = affected[4 * i], ay = affected[4 * i + 1], az = affected[4 * i + 2];
        double sx = 0, sy = 0, sz = 0;
        for (size_t j = 0; j < m; ++j) {
            const float *p = affecting + 4 * j;
            double dx = (double)p[0] - ax, dy = (double)p[1] - ay, dz = (double)p[2] - az;
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            double rs = r2 + e2;
            double s = (double)p[3] / (rs * sqrt(rs));
            sx += dx * s; sy += dy * s; sz += dz * s;
        }
        out3[3 * i] = sx; out3[3 * i + 1] = sy; out3[3 * i + 2] = sz;
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Octree                                                                                     */
/* ------------------------------------------------------------------------------------------ */

struct orc_tree {
    orc_node *nodes;
    size_t len, cap;
    uint32_t *order; /* tree particle order (original indices) */
    size_t m;
    const float *p;  /* borrowed during build only */
    int max_level;
    int max_depth;
    int failed;
    float rmin[3], rmax[3];
};

/* MidPoint for floats (partition.rs:55-59): (a + b) / 2 */
static inline float midpoint(float a, float b) { return (a + b) / 2.0f; }

/* impls/mod.rs:120-134, in f32 with the reference's fold orders, plus the FP64 reduction. */
static void centre_of_mass(const orc_tree *t, const uint32_t *idx, size_t cnt, orc_node *nd) {
    const float *p = t->p;
    float tot = 0.0f;
    double tot64 = 0.0;
    for (size_t i = 0; i < cnt; ++i) {
        tot = tot + p[4 * idx[i] + 3];
        tot64 += (double)p[4 * idx[i] + 3];
    }
    float cx = 0.0f, cy = 0.0f, cz = 0.0f;
    double dx = 0, dy = 0, dz = 0;
    if (tot == 0.0f) {
        for (size_t i = 0; i < cnt; ++i) {
            const float *q = p + 4 * idx[i];
            cx = cx + q[0]; cy = cy + q[1]; cz = cz + q[2];
        }
        float len = (float)cnt;
        cx = cx / len; cy = cy / len; cz = cz / len;
    } else {
        for (size_t i = 0; i < cnt; ++i) {
            const float *q = p + 4 * idx[i];
            float f = q[3] / tot;
            cx = cx + q[0] * f; cy = cy + q[1] * f; cz = cz + q[2] * f;
        }
    }
    if (tot64 == 0.0) {
        for (size_t i = 0; i < cnt; ++i) {
            const float *q = p + 4 * idx[i];
            dx += q[0]; dy += q[1]; dz += q[2];
        }
        dx /= (double)cnt; dy /= (double)cnt; dz /= (double)cnt;
    } else {
        for (size_t i = 0; i < cnt; ++i) {
            const float *q = p + 4 * idx[i];
            dx += (double)q[0] * (double)q[3];
            dy += (double)q[1] * (double)q[3];
            dz += (double)q[2] * (double)q[3];
        }
        dx /= tot64; dy /= tot64; dz /= tot64;
    }
    nd->com[0] = cx; nd->com[1] = cy; nd->com[2] = cz; nd->mass = tot;
    nd->com64[0] = dx; nd->com64[1] = dy; nd->com64[2] = dz; nd->mass64 = tot64;
}

static int push_node(orc_tree *t) {
    if (t->len == t->cap) {
        size_t nc = t->cap ? t->cap * 2 : 1024;
        orc_node *nn = (orc_node *)realloc(t->nodes, nc * sizeof(orc_node));
        if (!nn) { t->failed = 1; return -1; }
        t->nodes = nn;
        t->cap = nc;
    }
    return (int)t->len++;
}

/* tree/mod.rs:91-139 (build_node_with). idx is this node's slice, in the order the reference
 * holds it (stable partition of the parent's slice); tmp is scratch of the same length. */
static int build(orc_tree *t, uint32_t *idx, uint32_t *tmp, size_t cnt, const float bmin[3],
                 const float bmax[3], int level, int32_t begin) {
    if (cnt == 0) return -1;
    if (level > t->max_depth) { t->failed = 2; return -1; }
    int id = push_node(t);
    if (id < 0) return -1;
    orc_node *nd = &t->nodes[id];
    memset(nd, 0, sizeof(*nd));
    for (int c = 0; c < 8; ++c) nd->child[c] = -1;
    nd->level = level;
    nd->begin = begin;
    nd->count = (int32_t)cnt;
    for (int a = 0; a < 3; ++a) { nd->bmin[a] = bmin[a]; nd->bmax[a] = bmax[a]; }
    centre_of_mass(t, idx, cnt, nd);
    if (level > t->max_level) t->max_level = level;

    /* input.windows(2).any(|d| coordinates(&d[0]) != coordinates(&d[1])) */
    const float *p = t->p;
    int differ = 0;
    for (size_t i = 1; i < cnt && !differ; ++i) {
        const float *a = p + 4 * idx[i - 1], *b = p + 4 * idx[i];
        differ = (a[0] != b[0]) || (a[1] != b[1]) || (a[2] != b[2]);
    }
    if (!differ) return id;

    float center[3];
    for (int a = 0; a < 3; ++a) center[a] = midpoint(bmin[a], bmax[a]);
    /* index = sum_i (position[i] < center[i]) << i  (tree/mod.rs:121-123), stable partition */
    size_t counts[8] = {0}, off[8];
    for (size_t i = 0; i < cnt; ++i) {
        const float *q = p + 4 * idx[i];
        int c = (q[0] < center[0]) | ((q[1] < center[1]) << 1) | ((q[2] < center[2]) << 2);
        counts[c]++;
    }
    off[0] = 0;
    for (int c = 1; c < 8; ++c) off[c] = off[c - 1] + counts[c - 1];
    size_t w[8];
    memcpy(w, off, sizeof(w));
    for (size_t i = 0; i < cnt; ++i) {
        const float *q = p + 4 * idx[i];
        int c = (q[0] < center[0]) | ((q[1] < center[1]) << 1) | ((q[2] < center[2]) << 2);
        tmp[w[c]++] = idx[i];
    }
    memcpy(idx, tmp, cnt * sizeof(uint32_t));

    int32_t child[8];
    for (int c = 0; c < 8; ++c) {
        /* subdivide (partition.rs:193-220): bit j == 0 -> [center, max], else [min, center] */
        float cmin[3], cmax[3];
        for (int a = 0; a < 3; ++a) {
            if ((c & (1 << a)) == 0) { cmin[a] = center[a]; cmax[a] = bmax[a]; }
            else { cmin[a] = bmin[a]; cmax[a] = center[a]; }
        }
        child[c] = build(t, idx + off[c], tmp + off[c], counts[c], cmin, cmax, level + 1,
                         begin + (int32_t)off[c]);
        if (t->failed) return -1;
    }
    nd = &t->nodes[id]; /* the node array may have moved */
    nd->internal = 1;
    for (int c = 0; c < 8; ++c) nd->child[c] = child[c];
    return id;
}

orc_tree *orc_tree_build(const float *affecting, size_t m, int max_depth) {
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    if (!t) return NULL;
    t->m = m;
    t->p = affecting;
    t->max_depth = max_depth > 0 ? max_depth : 4096;
    t->max_level = -1;
    t->order = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
    uint32_t *tmp = (uint32_t *)malloc((m ? m : 1) * sizeof(uint32_t));
    if (!t->order || !tmp) { free(tmp); orc_tree_free(t); return NULL; }
    for (size_t i = 0; i < m; ++i) t->order[i] = (uint32_t)i;

    /* BoundingBox::square_with (partition.rs:136-153) via with/extend (:109-132) */
    float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < m; ++i)
        for (int a = 0; a < 3; ++a) {
            mn[a] = fminf(mn[a], affecting[4 * i + a]);
            mx[a] = fmaxf(mx[a], affecting[4 * i + a]);
        }
    float center[3], size[3];
    for (int a = 0; a < 3; ++a) { center[a] = midpoint(mn[a], mx[a]); size[a] = mx[a] - mn[a]; }
    float big = 0.0f; /* size.into_iter().fold(S::default(), S::max) */
    for (int a = 0; a < 3; ++a) big = fmaxf(big, size[a]);
    float half = midpoint(big, 0.0f); /* MidPoint::half */
    for (int a = 0; a < 3; ++a) { t->rmin[a] = center[a] - half; t->rmax[a] = center[a] + half; }

    build(t, t->order, tmp, m, t->rmin, t->rmax, 0, 0);
    free(tmp);
    t->p = NULL;
    if (t->failed) { orc_tree_free(t); return NULL; }
    return t;
}

void orc_tree_free(orc_tree *t) {
    if (!t) return;
    free(t->nodes);
    free(t->order);
    free(t);
}

size_t orc_tree_num_nodes(const orc_tree *t) { return t ? t->len : 0; }
int orc_tree_max_level(const orc_tree *t) { return t ? t->max_level : -1; }
const orc_node *orc_tree_nodes(const orc_tree *t) { return t ? t->nodes : NULL; }
const uint32_t *orc_tree_order(const orc_tree *t) { return t ? t->order : NULL; }
void orc_tree_root_box(const orc_tree *t, float *bmin3, float *bmax3) {
    for (int a = 0; a < 3; ++a) { bmin3[a] = t->rmin[a]; bmax3[a] = t->rmax[a]; }
}

/* End of every node's subtree in the pre-order node array (the index of the first node after
 * it), from the child links: a leaf ends at k + 1, an internal node where its last non-empty child
 * ends. Children always follow their parent, so one backward pass settles every node. */
void orc_tree_skips(const orc_tree *t, uint32_t *out) {
    for (size_t k = t->len; k-- > 0;) {
        uint32_t end = (uint32_t)k + 1;
        for (int c = 0; c < 8; ++c) {
            const int32_t ch = t->nodes[k].child[c];
            if (ch >= 0 && out[ch] > end) end = out[ch];
        }
        out[k] = end;
    }
}

/* sequential.rs:465-505. LIFO stack from the root; None children skipped; d2 = |com - x|^2
 * (acceleration.rs:121-126); skip when d2 == 0; open an Internal node when
 * (theta*theta)*d2 < width*width with width = bbox.max[0] - bbox.min[0] (partition.rs:183-188);
 * otherwise add the pair kernel with the node's {com, mass}. */
int orc_bh_f32(const orc_tree *t, const float *affected, size_t n, float theta, float softening,
               int checked, int com64, int accum64, float *out, uint32_t *visits, int threads) {
    if (!t || (n && (!affected || !out))) return 1;
    set_threads(threads);
    const orc_node *nodes = t->nodes;
    size_t stack_cap = (size_t)(t->max_level + 2) * 8 + 8;
    float th2 = theta * theta;
    long long nn = (long long)n;
    int rc = 0;
#pragma omp parallel
    {
        int32_t *stack = (int32_t *)malloc(stack_cap * sizeof(int32_t));
        if (!stack) {
#pragma omp atomic write
            rc = 3;
        }
#pragma omp for schedule(dynamic, 64)
        for (long long i = 0; i < nn; ++i) {
            if (!stack) continue;
            const float *a = affected + 4 * i;
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            double dx64 = 0.0, dy64 = 0.0, dz64 = 0.0;
            uint32_t acc = 0;
            size_t sp = 0;
            if (t->len) stack[sp++] = 0;
            while (sp) {
                int32_t id = stack[--sp];
                if (id < 0) continue;
                const orc_node *nd = &nodes[id];
                float cx, cy, cz;
                if (com64) { cx = (float)nd->com64[0]; cy = (float)nd->com64[1]; cz = (float)nd->com64[2]; }
                else { cx = nd->com[0]; cy = nd->com[1]; cz = nd->com[2]; }
                float ms = com64 ? (float)nd->mass64 : nd->mass;
                float d2 = norm2f(cx - a[0], cy - a[1], cz - a[2]);
                if (d2 == 0.0f) continue;
                float w = nd->bmax[0] - nd->bmin[0];
                if (nd->internal && (th2 * d2) < (w * w)) {
                    for (int c = 0; c < 8; ++c) stack[sp++] = nd->child[c];
                } else {
                    if (accum64) {
                        /* the same decisions, the pair term and the sum in FP64 */
                        double ex = (double)cx - a[0], ey = (double)cy - a[1], ez = (double)cz - a[2];
                        double r2 = ex * ex + ey * ey + ez * ez + (double)softening * softening;
                        double s = (double)ms / (r2 * sqrt(r2));
                        dx64 += ex * s; dy64 += ey * s; dz64 += ez * s;
                    } else {
                        float ox, oy, oz;
                        pair(cx, cy, cz, ms, a[0], a[1], a[2], softening, checked, &ox, &oy, &oz);
                        sx = sx + ox; sy = sy + oy; sz = sz + oz;
                    }
                    acc++;
                }
            }
            if (accum64) { sx = (float)dx64; sy = (float)dy64; sz = (float)dz64; }
            out[4 * i] = sx; out[4 * i + 1] = sy; out[4 * i + 2] = sz; out[4 * i + 3] = 0.0f;
            if (visits) visits[i] = acc;
        }
        free(stack);
    }
    return rc;
}

/* The relative opening criterion of the GPU backend (no reference counterpart): the traversal of
 * orc_bh_f32 (sequential.rs:456-506) in which an internal node is also opened when
 * lhs > (A * r2) * r2, with r2 = d2 + softening^2, w2 = width * width, A = est[i] (alpha |a~| of
 * affected particle i) and lhs = mass * w2 (order 2) or ((mass * w2) * sqrt(w2)) * (1 / sqrt(r2))
 * (order 3), every f32 operation rounded (IEEE sqrt and division). est = +inf gives orc_bh_f32's
 * decisions; est = 0 opens every internal node of positive mass. bucket > 0: an opened internal
 * node of at most `bucket` particles whose first tree-order particle is below 2^27 is summed
 * particle by particle in tree order (pairs at zero distance skipped, each counted in visits)
 * instead of being descended; `affecting` is the array the tree was built from. */
int orc_bh_rel_f32(const orc_tree *t, const float *affected, size_t n, float theta,
                   float softening, const float *est, int order, int bucket,
                   const float *affecting, int com64, int accum64, float *out, uint32_t *visits,
                   int threads) {
    if (!t || (n && (!affected || !out || !est)) || (order != 2 && order != 3)) return 1;
    if (bucket < 0 || (bucket > 0 && !affecting)) return 1;
    set_threads(threads);
    const orc_node *nodes = t->nodes;
    size_t stack_cap = (size_t)(t->max_level + 2) * 8 + 8;
    float th2 = theta * theta;
    float eps2 = softening * softening;
    long long nn = (long long)n;
    int rc = 0;
#pragma omp parallel
    {
        int32_t *stack = (int32_t *)malloc(stack_cap * sizeof(int32_t));
        if (!stack) {
#pragma omp atomic write
            rc = 3;
        }
#pragma omp for schedule(dynamic, 64)
        for (long long i = 0; i < nn; ++i) {
            if (!stack) continue;
            const float *a = affected + 4 * i;
            const float A = est[i];
            float sx = 0.0f, sy = 0.0f, sz = 0.0f;
            double dx64 = 0.0, dy64 = 0.0, dz64 = 0.0;
            uint32_t acc = 0;
            size_t sp = 0;
            if (t->len) stack[sp++] = 0;
            while (sp) {
                int32_t id = stack[--sp];
                if (id < 0) continue;
                const orc_node *nd = &nodes[id];
                float cx, cy, cz;
                if (com64) { cx = (float)nd->com64[0]; cy = (float)nd->com64[1]; cz = (float)nd->com64[2]; }
                else { cx = nd->com[0]; cy = nd->com[1]; cz = nd->com[2]; }
                float ms = com64 ? (float)nd->mass64 : nd->mass;
                float d2 = norm2f(cx - a[0], cy - a[1], cz - a[2]);
                if (d2 == 0.0f) continue;
                float w = nd->bmax[0] - nd->bmin[0];
                float w2 = w * w;
                int open = 0;
                if (nd->internal) {
                    float r2 = d2 + eps2;
                    float lhs = ms * w2;
                    if (order == 3) lhs = (lhs * sqrtf(w2)) * (1.0f / sqrtf(r2));
                    float rhs = (A * r2) * r2;
                    open = (th2 * d2) < w2 || rhs < lhs;
                }
                if (open && nd->count <= bucket && nd->begin < (1 << 27)) {
                    for (int32_t q = nd->begin; q < nd->begin + nd->count; ++q) {
                        const float *b = affecting + 4 * (size_t)t->order[q];
                        float ex = b[0] - a[0], ey = b[1] - a[1], ez = b[2] - a[2];
                        if (norm2f(ex, ey, ez) == 0.0f) continue;
                        if (accum64) {
                            double fx = (double)b[0] - a[0], fy = (double)b[1] - a[1],
                                   fz = (double)b[2] - a[2];
                            double q2 = fx * fx + fy * fy + fz * fz + (double)softening * softening;
                            double s = (double)b[3] / (q2 * sqrt(q2));
                            dx64 += fx * s; dy64 += fy * s; dz64 += fz * s;
                        } else {
                            float ox, oy, oz;
                            pair(b[0], b[1], b[2], b[3], a[0], a[1], a[2], softening, 1, &ox, &oy, &oz);
                            sx = sx + ox; sy = sy + oy; sz = sz + oz;
                        }
                        acc++;
                    }
                } else if (open) {
                    for (int c = 0; c < 8; ++c) stack[sp++] = nd->child[c];
                } else {
                    if (accum64) {
                        double ex = (double)cx - a[0], ey = (double)cy - a[1], ez = (double)cz - a[2];
                        double q2 = ex * ex + ey * ey + ez * ez + (double)softening * softening;
                        double s = (double)ms / (q2 * sqrt(q2));
                        dx64 += ex * s; dy64 += ey * s; dz64 += ez * s;
                    } else {
                        float ox, oy, oz;
                        pair(cx, cy, cz, ms, a[0], a[1], a[2], softening, 1, &ox, &oy, &oz);
                        sx = sx + ox; sy = sy + oy; sz = sz + oz;
                    }
                    acc++;
                }
            }
            if (accum64) { sx = (float)dx64; sy = (float)dy64; sz = (float)dz64; }
            out[4 * i] = sx; out[4 * i + 1] = sy; out[4 * i + 2] = sz; out[4 * i + 3] = 0.0f;
            if (visits) visits[i] = acc;
        }
        free(stack);
    }
    return rc;
}
