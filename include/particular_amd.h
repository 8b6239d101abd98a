/*
 * particular_amd -- MI355X (gfx950) backend for particular's gravitational acceleration path.
 *
 * C ABI. Plain pointers and sizes only. Every entry point returns an int status (PNB_OK = 0);
 * pnb_last_error() gives the message of the last failure on a context.
 *
 * Particle records are 16 bytes, {x, y, z, w} as 4 floats:
 *   affecting particles: w = mu (standard gravitational parameter), the
 *     GravitationalField<Vec3, f32> record of particular/src/gravity/mod.rs:12-18 and the
 *     `Affecting` vec4 of the reference GPU layout (particular/src/gravity/impls/mod.rs:430-434);
 *   affected particles: w ignored (the padded `Affected` vec3 of the same layout);
 *   outputs: {ax, ay, az, 0} (the padded `Interaction` vec3).
 *
 * The `_dev` entry points take device pointers and are asynchronous on the context's stream;
 * the others take host pointers and are synchronous on return, like the reference's
 * gpu::BruteForce::compute (particular/src/gpu/mod.rs:186-207, which blocks on the result).
 *
 * Interaction parameters mirror particular's interaction types:
 *   Acceleration::checked()                softening = 0, checked = 1
 *     (particular/src/gravity/newtonian/acceleration.rs:27-33)
 *   Acceleration::unchecked()              softening = 0, checked = 0
 *   AccelerationSoftened::checked(eps)     softening = eps, checked = 1
 *     (particular/src/gravity/newtonian/acceleration_softened.rs:30-37)
 *   AccelerationSoftened::unchecked(eps)   softening = eps, checked = 0
 * With checked = 1 a pair whose unsoftened squared distance is 0 contributes nothing
 * (particular/src/gravity/impls/mod.rs:160-161).
 *
 * A context is used by one host thread at a time (the reference takes &mut GpuResources,
 * particular/src/gpu/mod.rs:152).
 */
#ifndef PARTICULAR_AMD_H
#define PARTICULAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PNB_OK 0
#define PNB_E_ARG 1   /* invalid argument: null pointer with a non-zero length, bad theta, ... */
#define PNB_E_NODEV 2 /* no usable GPU */
#define PNB_E_OOM 3   /* device allocation failed */
#define PNB_E_HIP 4   /* a HIP runtime call failed */
#define PNB_E_DEPTH 5 /* octree deeper than the build supports (pathological input) */
#define PNB_E_STATE 6 /* call out of order, e.g. a step loop that was never initialised */

/* Kernel ids for pnb_profile_read. */
#define PNB_PROF_BF 0          /* brute-force acceleration kernel */
#define PNB_PROF_BH_BUILD 1    /* whole octree build (keys, sort, nodes, centres of mass) */
#define PNB_PROF_BH_WALK 2     /* Barnes-Hut traversal kernel */
#define PNB_PROF_DRIFT 3       /* leapfrog kick + drift kernel */
#define PNB_PROF_BATCH 4       /* batched-ensemble force kernel */
#define PNB_PROF_MLP 5         /* MLP training / inference kernels */
#define PNB_PROF_SORT 6        /* radix sort inside the octree build */
#define PNB_PROF_NUM_KERNELS 8

typedef struct pnb_ctx pnb_ctx;

/* ---- context: the GpuResources analogue (particular/src/gpu/mod.rs:84-143) ---------------- */

/* Number of visible GPUs; returns PNB_E_NODEV (and *count = 0) when there are none. */
int pnb_device_count(int *count);
/* Creates a context on `device` with its own stream (default priority). Replaces the lazily
 * initialised GpuResources::new / get_or_init (particular/src/gpu/mod.rs:96-142). */
int pnb_create(int device, pnb_ctx **out);
void pnb_destroy(pnb_ctx *ctx);
const char *pnb_last_error(const pnb_ctx *ctx);
const char *pnb_status_string(int status);
/* Makes later calls launch on `hip_stream` (a hipStream_t; NULL is HIP's null stream, which is
 * what a framework's default stream usually is). pnb_reset_stream returns to the context's own
 * stream. Graph capture (use_graph) needs a non-null stream. */
int pnb_set_stream(pnb_ctx *ctx, void *hip_stream);
int pnb_reset_stream(pnb_ctx *ctx);
/* The stream the context currently launches on, as a hipStream_t. */
void *pnb_get_stream(pnb_ctx *ctx);
int pnb_synchronize(pnb_ctx *ctx);

/* Per-kernel device time, measured with HIP events on the context stream around each launch of
 * the kernels named by PNB_PROF_*. Enabling resets the counters. */
int pnb_profile_enable(pnb_ctx *ctx, int enable);
int pnb_profile_read(pnb_ctx *ctx, int kernel_id, double *total_ms, uint64_t *launches);

/* ---- brute force ----------------------------------------------------------------------- */

/* Between(affected[0..n), affecting[0..m)) brute force. Replaces gpu::BruteForce::compute
 * (particular/src/gpu/mod.rs:179-208) and matches sequential::BruteForce
 * (particular/src/sequential.rs:181-194): out_i = sum_j pair(affected_i, affecting_j).
 * m = 0 gives zeros, n = 0 is a no-op (particular/src/gpu/resources.rs:272-280). */
int pnb_bf_accel_f32(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                     size_t m, float softening, int checked, float *out);
int pnb_bf_accel_f32_dev(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                         size_t m, float softening, int checked, float *out);

/* pnb_bf_accel_f32_dev without any device allocation. No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The same kernels and shapes as pnb_bf_accel_f32_dev, so the same results; the
 * partial sums of the affecting split go to a static 256 MiB array per GPU (splits * n records of
 * 16 bytes, at most 16,777,216; the default shapes stay within it on 256 CUs) instead of the
 * context's buffers. A shape that needs more (a pnb_bf_set_tuning entry with a large split) is
 * rejected with PNB_E_ARG. Every call on one GPU shares that array, whatever the context: calls
 * that overlap on one GPU (several streams or threads) race. */
int pnb_bf_accel_f32_dev_noalloc(pnb_ctx *ctx, const float *affected, size_t n,
                                 const float *affecting, size_t m, float softening, int checked,
                                 float *out);

/* Reordered lowering (particular/src/storage.rs:219-229): the affecting set is the particles
 * with mu != 0, in their original relative order (Ordered::new, storage.rs:85-95); every particle
 * is affected and out follows the original order. */
int pnb_reordered_bf_accel_f32(pnb_ctx *ctx, const float *particles, size_t n, float softening,
                               int checked, float *out);
int pnb_reordered_bf_accel_f32_dev(pnb_ctx *ctx, const float *particles, size_t n,
                                   float softening, int checked, float *out);

/* Brute force over pairs for one particle set (every particle affects and is affected): each
 * unordered pair is evaluated once and gives both accelerations (Newton's third law). Replaces
 * sequential::BruteForcePairs (particular/src/sequential.rs:275-305, RestrictedBruteForce
 * :215-263) with Acceleration's pair term (particular/src/gravity/impls/mod.rs:185-205):
 * out_i = sum_{j != i} mu_j d_ij / (|d_ij|^2 + eps^2)^1.5. Its Ordered and Reordered forms
 * (sequential.rs:288-305, 368-385) give the same sums over the whole set (pairs of two massless
 * particles add zero). checked = 0 keeps the reference's unchecked pair form (a pair at zero
 * distance is non-finite); checked = 1 drops such pairs as pnb_bf_accel_f32 does. Sums are
 * compensated, in a fixed order. n up to 2^22; device scratch of 16 * ceil(n / 4096)^2 * 4096
 * bytes (one partial sum per body and block of 4,096 bodies: 4.3 GB at n = 2^20). */
int pnb_bf_pairs_accel_f32(pnb_ctx *ctx, const float *particles, size_t n, float softening,
                           int checked, float *out);
int pnb_bf_pairs_accel_f32_dev(pnb_ctx *ctx, const float *particles, size_t n, float softening,
                               int checked, float *out);

/* Brute-force kernel shape. variant: 0 = affecting tile staged in LDS, 1 = affecting particles
 * read as wave-uniform scalar loads. per_lane: affected particles per lane (1, 2, 4, 8).
 * j_split: affecting ranges per affected block (1..64, partial sums reduced in fixed order).
 * Applies to problems with n >= n_min until replaced; pnb_bf_clear_tuning drops all entries. */
int pnb_bf_set_tuning(pnb_ctx *ctx, size_t n_min, int variant, int per_lane, int j_split);
int pnb_bf_clear_tuning(pnb_ctx *ctx);

/* ---- Barnes-Hut -----------------------------------------------------------------------------
 * Replaces parallel::BarnesHut (particular/src/parallel.rs:342-367): the octree is built from
 * the affecting particles exactly as RootedOrthtree::new builds it (particular/src/storage.rs:20-33,
 * particular/src/tree/mod.rs:73-139, same f32 box arithmetic, same cells, same pre-order), with
 * the centre of mass of every node reduced in FP64; every affected particle then applies the
 * traversal rule of sequential::BarnesHut (particular/src/sequential.rs:465-505): skip a node at
 * zero distance, open an internal node when (theta*theta)*d^2 < width^2, else add its pair term.
 * theta = 0 degenerates to brute force over the leaves (sequential.rs:441-442). `checked` is
 * accepted for symmetry with the reference signature; the traversal's zero-distance skip makes
 * both forms identical. flags: a combination of PNB_BH_MLP and PNB_BH_REF_COM. Sizes: n up to
 * 2^31 - 1; any octree the node counts (uint32) allow, trees of 2^27 nodes or more being walked
 * with 64-bit record offsets (the environment variable PNB_WIDE_WALK=1 forces that form, for
 * tests). */
#define PNB_BH_MLP 1u /* add the learned correction of each accepted internal node (see below) */
/* Node data exactly as the reference computes it: f32 sequential folds of mass and mass-weighted
 * position per node (particular/src/gravity/impls/mod.rs:126-131) over the node's bodies in input
 * order, the order of the reference's per-node slices (particular/src/tree/mod.rs:116-125),
 * instead of the FP64 reduction; bit-identical node data. The folds are sequential, so a node of
 * k bodies costs k dependent steps (the root: all of them); meant for reference parity checks,
 * not for speed. The default FP64 centres of mass differ from these by the reference's own f32
 * rounding, which at 4M bodies moves accelerations by up to 4.4e-5 rms (theta = 0.8). */
#define PNB_BH_REF_COM 2u
/* Relative opening criterion (no reference counterpart; the criterion of tree codes that bound
 * each node's force error by a fraction of the particle's own acceleration, e.g. GADGET-2,
 * Springel 2005, section 3.1). An internal node is opened when the geometric rule above opens it
 * OR when M w^p / r^(p+2) > alpha |a~|: M the node's mass, w its width, r^2 = d^2 + eps^2, |a~|
 * the acceleration from an estimate pass (a plain walk at estimate_theta over the same tree; its
 * time is part of the walk's) at the particle itself or, by default, at the centroid of the 64
 * consecutive particles of its wavefront in the walk's spatial order (1/64 of the cost, the same
 * rms error), and p the order of the first term the walk leaves out:
 * 2 for the plain walk (the quadrupole), 3 with PNB_BH_MLP (whose learned moment stands in for
 * the quadrupole) and 4 with PNB_BH_MLP when the network has octupole outputs (a non-zero weight
 * or bias of outputs 6..12, or any training). Each node's error is then bounded relative to the
 * particle's own acceleration;
 * `theta` stays a geometric floor (1 opens every node whose width exceeds the distance). alpha
 * and estimate_theta come from pnb_bh_set_relative (defaults 1e-4 and 1.2). Combines with the
 * other flags. */
#define PNB_BH_RELATIVE 4u
#define PNB_BH_ALL_FLAGS (PNB_BH_MLP | PNB_BH_REF_COM | PNB_BH_RELATIVE)
/* Parameters of PNB_BH_RELATIVE for the context's later calls: alpha >= 0 (0 opens every internal
 * node with mass), estimate_theta > 0, per_particle != 0 for one estimate per particle instead of
 * one per wavefront, and bucket (0 .. 31, default 24): an internal node of at most `bucket`
 * bodies that a particle opens is summed body by body (in tree order) instead of being walked
 * (0: never; nodes whose first body lies at or beyond 2^27 are always walked). */
int pnb_bh_set_relative(pnb_ctx *ctx, float alpha, float estimate_theta, int per_particle,
                        int bucket);

int pnb_bh_accel_f32(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                     size_t m, float theta, float softening, int checked, unsigned flags,
                     float *out);
int pnb_bh_accel_f32_dev(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                         size_t m, float theta, float softening, int checked, unsigned flags,
                         float *out);
/* Reordered lowering (particular/src/storage.rs:219-229): tree from the particles with mu != 0. */
int pnb_reordered_bh_accel_f32(pnb_ctx *ctx, const float *particles, size_t n, float theta,
                               float softening, int checked, unsigned flags, float *out);
int pnb_reordered_bh_accel_f32_dev(pnb_ctx *ctx, const float *particles, size_t n, float theta,
                                   float softening, int checked, unsigned flags, float *out);

/* Capturable Barnes-Hut (no reference counterpart; for step loops captured in a hipGraph and
 * for callers that cannot wait on the host). With enable != 0, pnb_bh_accel_f32_dev (flags 0,
 * Between on device pointers) builds and walks the same tree, with the same results, without
 * any host round trip or allocation once sized: the first call at given particle counts is an
 * ordinary call that sizes the buffers (a quarter more nodes and 8 more levels than its tree
 * needs) and must not be captured; later calls at the same counts can be captured on the
 * context's stream (pnb_set_stream) and replayed with new particle data. A later tree that
 * needs more nodes or levels than were sized is flagged on the device: its accelerations are NaN
 * and pnb_bh_check then returns PNB_E_STATE (call again outside a capture to size anew). The other
 * learned correction (PNB_BH_MLP) runs in this mode too: the node moments and the network pass
 * over the internal nodes use the device's counts, and the sizing call's flag is part of the
 * sizes (a call with the other flag sizes anew, so inside a capture it is refused). The other
 * flags (PNB_BH_REF_COM, PNB_BH_RELATIVE) are refused in this mode; the Reordered entry points
 * run in it but read the massive count on the host, so they are refused inside a capture.
 * Changing the mode forgets the sizes. */
int pnb_bh_set_capturable(pnb_ctx *ctx, int enable);
/* Waits for the context's stream and reports the status of the last capturable build: PNB_OK,
 * PNB_E_STATE (did not fit its sizes), PNB_E_DEPTH or PNB_E_ARG (non-finite positions). */
int pnb_bh_check(pnb_ctx *ctx);

/* The last octree built on a context, for inspection (the reference exposes its tree through
 * RootedOrthtree::get, particular/src/storage.rs:41-45). Nodes are in the reference's order
 * (depth-first pre-order, children in orthant order). */
typedef struct {
    float com[3];     /* centre of mass, f32 (rounded from the FP64 reduction) */
    float mass;       /* total mu, f32 */
    double mass64;    /* total mu, FP64 */
    float width;      /* x extent of the node's box (partition.rs:183-188) */
    uint32_t skip;    /* index of the first node after this node's subtree */
    uint32_t begin;   /* first particle of the node, in tree order */
    uint32_t count;   /* particles in the node */
    uint32_t level;   /* depth below the root */
    uint32_t internal;/* 1 for Node::Internal, 0 for Node::External */
    /* Traceless quadrupole about the centre of mass, normalised by mass and width:
     * (3 T - tr(T) I) / (mass w^2) with T = sum mu (x - com)(x - com)^T, as {xx, yy, zz, xy, xz,
     * yz}. Filled when the tree was built for the learned correction, zero otherwise (and for
     * leaves). */
    float quad[6];
} pnb_tree_node;

int pnb_bh_tree_size(pnb_ctx *ctx, size_t *nodes, int *max_level);
/* Copies the nodes (cap >= nodes) and, if `order` is not NULL, the original index of every
 * affecting particle in tree order (m >= affecting count). */
int pnb_bh_tree_export(pnb_ctx *ctx, pnb_tree_node *nodes, size_t cap, uint32_t *order, size_t m);
/* Counters of the walks since the last build: wavefront node steps and accepted pair terms (with
 * PNB_BH_RELATIVE, the estimate pass's included). */
int pnb_bh_stats(pnb_ctx *ctx, uint64_t *wave_steps, uint64_t *accepted);

/* Two-dimensional particles (glam::Vec2 / f32; the reference's 2D GPU layout,
 * particular/src/gravity/impls/mod.rs:381-418): records {x, y, mu} (12 bytes, the repr(C)
 * GravitationalField<Vec2, f32>), outputs {ax, ay} (8 bytes). Same semantics as the 3D entry
 * points (brute force: sequential.rs:181-194; Barnes-Hut on the reference's quadtree:
 * sequential.rs:456-506, flags as pnb_bh_accel_f32). Computed on the 3D path with z = 0, which
 * reproduces the 2D arithmetic and the quadtree exactly. */
int pnb_bf_accel_f32_2d(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                        size_t m, float softening, int checked, float *out);
int pnb_bf_accel_f32_2d_dev(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                            size_t m, float softening, int checked, float *out);
int pnb_bh_accel_f32_2d(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                        size_t m, float theta, float softening, int checked, unsigned flags,
                        float *out);
int pnb_bh_accel_f32_2d_dev(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                            size_t m, float theta, float softening, int checked, unsigned flags,
                            float *out);

/* FP64 brute force for glam::DVec3 / f64 particles (CPU-only in the reference:
 * particular/src/gravity/impls/glam.rs:234): records are 4 doubles {x, y, z, mu} (32 B); the
 * reference's scalar kernel in FP64 with IEEE sqrt and division (impls/mod.rs:151-166). */
int pnb_bf_accel_f64(pnb_ctx *ctx, const double *affected, size_t n, const double *affecting,
                     size_t m, double softening, int checked, double *out);
int pnb_bf_accel_f64_dev(pnb_ctx *ctx, const double *affected, size_t n, const double *affecting,
                         size_t m, double softening, int checked, double *out);

/* ---- accuracy reference ---------------------------------------------------------------- */

/* FP64 direct sum on the GPU (device pointers): out3[3 i + k] = sum_j mu_j d / (|d|^2+eps^2)^1.5
 * with pairs at zero distance skipped. */
int pnb_direct_f64_dev(pnb_ctx *ctx, const float *affected, size_t n, const float *affecting,
                       size_t m, double softening, double *out3);


/* ---- leapfrog step loop ---------------------------------------------------------------------
 * particular has no integrator (particular/README.md:22-25); its tests step with symplectic Euler
 * around compute (particular/src/gravity/newtonian/mod.rs:318-330). This loop keeps the state
 * on the GPU: kick-drift-kick leapfrog, FP64 positions and velocities, brute-force accelerations
 * from an f32 shadow, affecting set = particles with mu != 0 (compacted once, original order). */

/* Uploads pos3/vel3 (n x 3 doubles) and mu (n floats) and computes the initial accelerations. */
int pnb_leapfrog_init(pnb_ctx *ctx, const double *pos3, const double *vel3, const float *mu,
                      size_t n, float softening, int checked);
/* Enqueues `steps` KDK steps (v += a dt/2; x += v dt; a = F(x); v += a dt/2) on the context
 * stream; use_graph captures them once into a hipGraph and replays it. Asynchronous. */
int pnb_leapfrog_run(pnb_ctx *ctx, size_t steps, double dt, int use_graph);
/* Copies the state back (any pointer may be NULL): pos3/vel3 n x 3 doubles, acc4 n x 4 floats. */
int pnb_leapfrog_read(pnb_ctx *ctx, double *pos3, double *vel3, float *acc4);
/* Device views of the f32 particle shadow (n x 4), the compacted affecting set (m x 4) and the
 * accelerations (n x 4); valid until the next init. */
int pnb_leapfrog_views(pnb_ctx *ctx, const float **particles_dev, const float **affecting_dev,
                       const float **acc_dev, size_t *n, size_t *m);

/* ---- batched ensemble -----------------------------------------------------------------------
 * Many independent systems of uneven sizes stepped together; no reference counterpart (the
 * reference computes one storage per call). System k owns particles [offsets[k], offsets[k+1])
 * and is its own Between(system, system) brute force; all systems advance with the KDK leapfrog
 * of the step loop above. Each step's force pass is one persistent launch over a queue of
 * (system, tile) items; tile shapes per system-size bucket come from pnb_batch_set_tuning. */

/* Affected bodies per lane (1, 2, 4, 8) for systems with at least n_min bodies. */
int pnb_batch_set_tuning(pnb_ctx *ctx, size_t n_min, int per_lane);
/* Uploads the ensemble (pos3/vel3: total x 3 doubles, mu: total floats, offsets: num_systems+1
 * int64 starting at 0) and computes the initial accelerations. */
int pnb_batch_init(pnb_ctx *ctx, const double *pos3, const double *vel3, const float *mu,
                   const int64_t *offsets, size_t num_systems, float softening, int checked);
/* Enqueues `steps` KDK steps for every system; use_graph captures them once into a hipGraph. */
int pnb_batch_run(pnb_ctx *ctx, size_t steps, double dt, int use_graph);
int pnb_batch_read(pnb_ctx *ctx, double *pos3, double *vel3, float *acc4);
/* Particle count, work items per force pass and pair interactions per step (sum of N_k^2). */
int pnb_batch_info(pnb_ctx *ctx, size_t *particles, size_t *items,
                   unsigned long long *interactions_per_step);

/* ---- learned far-field correction (Barnes-Hut + MLP) -----------------------------------------
 * No reference counterpart. A 16 -> 64 -> 64 -> 13 ReLU MLP (BF16 weights on MFMA, FP32
 * accumulate) corrects the monopole term of every internal node a lane accepts in the Barnes-Hut
 * walk (pnb_bh_accel_f32* with PNB_BH_MLP). Its input describes the node (pnb_bh_node_features):
 * f0..5 the normalised quadrupole Q = (3T - tr(T) I) / (M w^2) (see pnb_tree_node),
 * f6 1 + log2(mass / root mass) / 32, f7 log2(bodies) / 32, f8 depth / 64, f9..15 32 times the
 * normalised octupole (5/2) U~ / (M w^3), U~ the traceless part of the third moments about the
 * centre of mass, as {xxx, yyy, zzz, xxy, xxz, xyy, xyz}. Its outputs are two fields' coefficients: y[0..5] =
 * f0..5 + MLP(f)[0..5] (a residual on the quadrupole) is a second moment Q in units of M w^2 and
 * y[6..12] = MLP(f)[6..12] (learned from zero) a traceless third moment O in units of M w^3 (same
 * seven components). Every lane that accepts the node adds, with d = centre of mass - lane,
 * a = -Q d |d|^-5 + 2.5 (d.Q.d) d |d|^-7 + 3 F |d|^-7 - 7 (d.F) d |d|^-9 with F_i = O_ijk d_j d_k;
 * with y[6..12] = f9..15 / 32 the last two terms are the octupole term of the node's multipole
 * expansion; they are left out for a node whose largest y[6..12] coefficient is below 3e-3. The
 * network runs once per node after the octree build. Training data are generated
 * by the method itself: per-lane targets are brute force minus Barnes-Hut. */

int pnb_mlp_num_params(void);
/* He-normal hidden layers from a seed (SplitMix64 + Box-Muller), zero biases and a zero output
 * layer (a fresh network adds nothing to its residual: the correction is the quadrupole term);
 * resets the Adam state. */
int pnb_mlp_init(pnb_ctx *ctx, uint64_t seed);
/* Parameters as one FP32 vector: W1[16][64], b1[64], W2[64][64], b2[64], W3[64][13], b3[13]. */
int pnb_mlp_set_params(pnb_ctx *ctx, const float *params, size_t n);
int pnb_mlp_get_params(pnb_ctx *ctx, float *params, size_t n);
/* Builds the octree of `particles` (device, n x 4, all affecting), walks `waves` seeded
 * wavefronts of 64 consecutive particles in tree order at `theta` recording their rows (one row =
 * one internal node some lane of the wavefront accepted, with the accepting-lane mask; at most
 * row_cap per wavefront, the rest counted as dropped), computes their brute-force accelerations,
 * and appends rows and targets to the training set. */
int pnb_mlp_collect(pnb_ctx *ctx, const float *particles, size_t n, float theta, float softening,
                    size_t waves, uint64_t seed, size_t row_cap);
int pnb_mlp_clear_dataset(pnb_ctx *ctx);
int pnb_mlp_dataset_info(pnb_ctx *ctx, size_t *rows, size_t *waves, uint64_t *dropped);
/* Training set on the host: feats rows x 16, geom rows x 8 (node centre of mass, M w^2, M w^3,
 * three zeros), mask rows (64-bit lane masks), row_wave rows, wave_rows waves + 1 (row offsets), lanes
 * waves x 64 x 4 (lane positions, w = softening^2), target waves x 64 x 4, wgt waves x 64. */
int pnb_mlp_dataset_read(pnb_ctx *ctx, float *feats, float *geom, uint64_t *mask,
                         uint32_t *row_wave, uint32_t *wave_rows, float *lanes, float *target,
                         float *wgt);
int pnb_mlp_dataset_load(pnb_ctx *ctx, const float *feats, const float *geom,
                         const uint64_t *mask, const uint32_t *row_wave,
                         const uint32_t *wave_rows, size_t rows, size_t waves,
                         const float *lanes, const float *target, const float *wgt);
/* Loss and gradients over the whole training set, without an update. */
int pnb_mlp_grad(pnb_ctx *ctx, float *loss, float *grads, size_t n);
/* Full-batch Adam steps; loss_history (optional, `iterations` floats) gets the loss before each. */
int pnb_mlp_train(pnb_ctx *ctx, size_t iterations, float lr, float *loss_history);
/* Inference on device rows: y (rows x 16, 13 used) = MLP(feats (rows x 16)), plus feats[0..5]
 * on y[0..5]. */
int pnb_mlp_forward_dev(pnb_ctx *ctx, const float *feats, size_t rows, float *y);
/* The network's input of every node of the last octree built with PNB_BH_MLP, for inspection:
 * out (cap >= nodes, host) gets 16 floats per node in pre-order, as the node network reads them
 * (zeros for leaves and massless nodes). PNB_E_STATE when the last build had no moments. */
int pnb_bh_node_features(pnb_ctx *ctx, float *out, size_t cap);

#ifdef __cplusplus
}
#endif

#endif /* PARTICULAR_AMD_H */
