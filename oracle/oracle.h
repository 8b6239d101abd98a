/*
 * CPU oracle for particular's acceleration path (f32, 3D) -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and
 * only as the checker or the timed CPU baseline. The product path (particular_amd) never calls it.
 *
 * Particle layout everywhere: 4 floats per particle, {x, y, z, w}; for affecting particles w is
 * the standard gravitational parameter mu (particular/src/gravity/mod.rs:12-18), for affected
 * particles w is ignored. Outputs are 4 floats per particle {ax, ay, az, 0}.
 *
 * Every function states the reference lines it restates. All f32 arithmetic is compiled with
 * -ffp-contract=off so each operation rounds exactly as the Rust reference does.
 */
#ifndef PARTICULAR_ORACLE_H
#define PARTICULAR_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One pair interaction, the scalar kernel of particular/src/gravity/impls/mod.rs:151-166 for
 * glam::Vec3/f32 (instantiated at impls/glam.rs:232). p2 = {x,y,z,mu}, at = {x,y,z}. */
void orc_accel_pair(const float *p2, const float *at, float softening, int checked, float *out3);

/* sequential::BruteForce (particular/src/sequential.rs:181-194): per affected particle, a fold
 * over the affecting slice in index order. parallel::BruteForce (parallel.rs:216-232) runs the
 * same fold per particle on many threads, so results are identical for any `threads`. */
int orc_bf_f32(const float *affected, size_t n, const float *affecting, size_t m, float softening,
               int checked, float *out, int threads);

/* sequential::BruteForcePairs over a slice (restrict_len = n) or an Ordered storage (restrict_len =
 * affecting count, massive particles first): RestrictedBruteForce (particular/src/sequential.rs:
 * 215-263) with the pair term of particular/src/gravity/impls/mod.rs:185-205, in the reference's
 * order. checked = 0 is the reference (no zero-distance check on this path); checked = 1 drops
 * pairs at zero unsoftened distance. Single-threaded. */
int orc_bf_pairs_f32(const float *p, size_t n, size_t restrict_len, float softening, int checked,
                     float *out);

/* FP64 direct sum, the accuracy reference: a_i = sum_j mu_j d_ij / (|d_ij|^2 + eps^2)^{3/2}, pairs
 * with |d_ij| == 0 skipped. Output 3 doubles per particle. */
int orc_direct_f64(const float *affected, size_t n, const float *affecting, size_t m,
                   double softening, double *out3, int threads);

/* Node of the restated octree, in the reference's node order (depth-first pre-order, children
 * in orthant order 0..7; particular/src/tree/mod.rs:108-134). */
typedef struct {
    int32_t child[8];  /* -1 for None */
    int32_t internal;  /* Node::Internal vs Node::External */
    int32_t level;     /* depth below the root */
    int32_t begin;     /* first particle of the node in tree particle order */
    int32_t count;     /* number of particles in the node */
    float bmin[3], bmax[3]; /* bounding box the node was built with */
    float com[3], mass;     /* f32 centre of mass exactly as the reference computes it */
    double com64[3], mass64;/* the same reduction in FP64 */
} orc_node;

typedef struct orc_tree orc_tree;

/* RootedOrthtree::new (particular/src/storage.rs:20-33) over `affecting`: square root box
 * (partition.rs:136-153), recursive subdivision (tree/mod.rs:91-139), centre of mass per node
 * (impls/mod.rs:120-134). Returns NULL on allocation failure or when the recursion exceeds
 * `max_depth` (the reference would recurse without bound). */
orc_tree *orc_tree_build(const float *affecting, size_t m, int max_depth);
void orc_tree_free(orc_tree *t);
size_t orc_tree_num_nodes(const orc_tree *t);
int orc_tree_max_level(const orc_tree *t);
const orc_node *orc_tree_nodes(const orc_tree *t);
/* Original indices of the affecting particles in tree order (the order of the leaves). */
const uint32_t *orc_tree_order(const orc_tree *t);
/* The cubic root box. */
void orc_tree_root_box(const orc_tree *t, float *bmin3, float *bmax3);
/* Subtree ends in pre-order: out[k] = index of the first node after node k's subtree. */
void orc_tree_skips(const orc_tree *t, uint32_t *out);

/* sequential::BarnesHut traversal (particular/src/sequential.rs:456-506) for every affected
 * particle. com64 = 0 uses the reference's f32 node data, 1 uses the FP64 reduction rounded to
 * f32. accum64 = 1 keeps every decision but evaluates and sums the accepted pair terms in FP64
 * (isolates the summation error of an implementation). `visits` (optional, n entries) receives
 * the number of accepted node interactions. */
int orc_bh_f32(const orc_tree *t, const float *affected, size_t n, float theta, float softening,
               int checked, int com64, int accum64, float *out, uint32_t *visits, int threads);

/* orc_bh_f32 with the GPU backend's relative opening criterion (PNB_BH_RELATIVE, no reference
 * counterpart): an internal node is also opened when M w^p / r^(p+2) > est[i], p = order (2 or
 * 3), r^2 = d2 + eps^2, est[i] = alpha |a~_i| supplied by the caller; an opened node of at most
 * `bucket` particles is summed particle by particle (`affecting`: the array the tree was built
 * from); see oracle.c for the f32 operation order. */
int orc_bh_rel_f32(const orc_tree *t, const float *affected, size_t n, float theta,
                   float softening, const float *est, int order, int bucket,
                   const float *affecting, int com64, int accum64, float *out, uint32_t *visits,
                   int threads);

#ifdef __cplusplus
}
#endif

#endif
