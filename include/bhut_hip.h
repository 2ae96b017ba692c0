/*
 * bhut_hip.h -- C ABI of libbhut_hip.so, the MI355X (gfx950) implementation
 * of bhut's Barnes-Hut force path.
 *
 * Plain pointers and sizes only.  Functions without a suffix take HOST
 * pointers and are synchronous; `_dev` variants take DEVICE pointers (for
 * example torch.Tensor.data_ptr() of tensors on the same device), enqueue on
 * the context stream and return without synchronising, except where noted.
 *
 * Status: 0 on success, a negative BH_E* code on failure; bh_last_error()
 * returns the message of the calling thread's last failure.
 *
 * Reference interfaces are cited as path:line in maroba/bhut.  The reference
 * has no C ABI of its own; every entry point below replaces a Python function
 * on the NumPy path, and bhut_amd/_native.py binds them with ctypes.
 */
#ifndef BHUT_HIP_H
#define BHUT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BH_OK 0
#define BH_EINVAL (-1)    /* invalid argument (maps to ValueError)          */
#define BH_EHIP (-2)      /* HIP runtime error                              */
#define BH_ENOMEM (-3)    /* device allocation failed                       */
#define BH_EEMPTY (-4)    /* zero particles (build.py:81-82)                */
#define BH_ENODEV (-5)    /* no HIP device                                  */
#define BH_EOVERFLOW (-6) /* an internal capacity was exceeded              */

typedef struct bh_ctx bh_ctx;   /* device, stream, cached allocations */
typedef struct bh_tree bh_tree; /* a tree resident in device memory   */

/* Options of the FP32 relative-coordinate, compensated, grouped form.
 * Zero fields take the defaults in brackets. */
typedef struct {
  int32_t group_size; /* targets per group: 64 (one per lane) or 128 [64]   */
  int32_t kahan;      /* 1: Kahan-compensated per-target sums, 0: plain     */
  int32_t stack_cap;  /* initial LDS stack entries of the group walk, <= 8192;
                         doubled (up to 8192) when a walk overflows [1024] */
  int32_t multipole;  /* accepted cells: 1 monopole, 2 quadrupole [2]       */
} bh_f32_opts;

/* ---------------------------------------------------------------- context */
const char *bh_last_error(void);
int bh_version(void);
int bh_device_count(void);
int bh_ctx_create(int device, bh_ctx **out);
void bh_ctx_destroy(bh_ctx *ctx);
int bh_ctx_synchronize(bh_ctx *ctx);
void *bh_ctx_stream(bh_ctx *ctx);             /* the hipStream_t used        */
/* Enqueue all further work on `stream` (a hipStream_t, e.g. torch's current
 * stream); NULL returns to the context's own stream.  Synchronises first. */
int bh_ctx_set_stream(bh_ctx *ctx, void *stream);
int bh_ctx_trim(bh_ctx *ctx);                 /* release cached allocations  */
/* Upper bound, in bytes, on the FP32 form's interaction lists held at once;
 * larger problems are walked and evaluated in batches of groups.  0 selects
 * the default (a quarter of the device memory, at most 64 GiB).  Results do
 * not depend on it. */
int bh_ctx_set_list_budget(bh_ctx *ctx, int64_t bytes);
/* Per-kernel timing with HIP events recorded on the context stream around
 * the hot kernels (walks, list builds, evaluations).  enable=1 clears the
 * records and starts recording; the summary is one "name launches total_ms"
 * line per kernel (synchronises). */
int bh_ctx_profile(bh_ctx *ctx, int32_t enable);
int bh_ctx_profile_summary(bh_ctx *ctx, char *buf, int64_t cap);

/* ------------------------------------------------------------------ build */
/* build_tree (bhut/tree/build.py:19-107, via bhut/api.py:155-191): bounding
 * box, Morton codes, stable sort, median-split topology, tight node boxes,
 * mass and centre of mass.  pos is (n, dim) row-major float64. */
int bh_tree_build(bh_ctx *ctx, const double *pos, const double *mass, int64_t n, int32_t dim,
                  int32_t leaf_size, bh_tree **out);
int bh_tree_build_dev(bh_ctx *ctx, const double *d_pos, const double *d_mass, int64_t n,
                      int32_t dim, int32_t leaf_size, bh_tree **out); /* enqueued, returns at once */
/* TreeData (bhut/tree/node.py:15-77) in host memory plus the source
 * positions/masses the traversal will use (bh.py:428 `positions, masses`). */
int bh_tree_import(bh_ctx *ctx, int64_t n, int64_t num_nodes, int32_t dim, int32_t leaf_size,
                   const double *center, const double *half_size, const double *mass,
                   const double *com, const int64_t *first_child, const int64_t *child_count,
                   const int64_t *start, const int64_t *count, const int64_t *perm,
                   const double *src_pos, const double *src_mass, bh_tree **out);
int bh_tree_clone(const bh_tree *t, bh_tree **out);
/* Batched form: a forest of nsys independent systems whose particles are
 * consecutive runs of counts[s] rows of pos/mass (counts is host memory).
 * Every system gets exactly the tree build_tree would give it alone
 * (build.py:19-107); one segmented sort on (system, Morton code), one
 * topology launch and one upward pass serve all systems.  Self accelerations
 * (bh_accel_self*, bh_accel_f32_self*), refit, lists and counters accept a
 * forest; each system's particles feel only their own system. */
int bh_forest_build(bh_ctx *ctx, const double *pos, const double *mass, const int64_t *counts, int64_t nsys,
                    int32_t dim, int32_t leaf_size, bh_tree **out);
int bh_forest_build_dev(bh_ctx *ctx, const double *d_pos, const double *d_mass, const int64_t *counts,
                        int64_t nsys, int32_t dim, int32_t leaf_size, bh_tree **out); /* enqueued */
/* Number of systems and their particle / node offsets (nsys + 1 entries
 * each; either array may be NULL). */
int bh_tree_systems(const bh_tree *t, int64_t *nsys, int64_t *particle_offsets, int64_t *node_offsets);
void bh_tree_free(bh_tree *t);
int bh_tree_shape(const bh_tree *t, int64_t *n, int64_t *num_nodes, int32_t *dim,
                  int32_t *leaf_size, int32_t *depth);
/* LDS stack entries the FP32 group walk on this tree has grown to, for self
 * (out[0]) and external (out[1]) targets; 0 where it never had to grow.
 * Later walks on the tree start there.  No reference counterpart: the
 * autotuner reads it to report the stack a candidate actually ran with. */
int bh_tree_f32_stack(const bh_tree *t, int32_t *out);
/* TreeData arrays (node.py:59-77) copied to host buffers; any may be NULL. */
int bh_tree_export(const bh_tree *t, double *center, double *half_size, double *mass, double *com,
                   int64_t *first_child, int64_t *child_count, int64_t *start, int64_t *count,
                   uint8_t *is_leaf, int64_t *perm);

/* ------------------------------------------------------------------ refit */
/* should_refit_vs_rebuild (bhut/tree/refit.py:195-251): *refit_ok = 1 when
 * every particle lies within 1.1 root half-sizes of the stored root centre. */
int bh_tree_should_refit(bh_tree *t, const double *pos, int64_t n, int32_t dim, int32_t *refit_ok);
int bh_tree_should_refit_dev(bh_tree *t, const double *d_pos, int64_t n, int32_t dim,
                             int32_t *refit_ok); /* synchronises */
/* refit_tree (refit.py:17-109): new mass and centre of mass from positions in
 * ORIGINAL order; centres and half sizes stay as built.  mass may be NULL to
 * keep the current masses. */
int bh_tree_refit(bh_tree *t, const double *pos, const double *mass);
int bh_tree_refit_dev(bh_tree *t, const double *d_pos, const double *d_mass);

/* ---------------------------------------------------- accelerations (FP64) */
/* barnes_hut_accelerations (bhut/traverse/bh.py:428-482; evaluate path of
 * api.py:367-429 and Tree.evaluate_accelerations api.py:733-786): out is
 * (n, dim) in the original particle order; source i skips target i. */
int bh_accel_self(bh_tree *t, double theta, double softening, double G, double *out);
int bh_accel_self_dev(bh_tree *t, double theta, double softening, double G, double *d_out);
/* barnes_hut_accelerations_targets (bh.py:30-80; Tree.accelerations with
 * targets, api.py:608-688): no self exclusion. */
int bh_accel_targets(bh_tree *t, const double *targets, int64_t m, double theta, double softening,
                     double G, double *out);
int bh_accel_targets_dev(bh_tree *t, const double *d_targets, int64_t m, double theta,
                         double softening, double G, double *d_out);

/* bh_accel_targets_dev without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * targets' Morton ordering lives in a fixed workspace the library holds on
 * every device it runs on (24.25 MiB, set up by bh_ctx_create), so m is at
 * most 1048576 per call (BH_EINVAL above that: split larger queries).  Same
 * results as bh_accel_targets_dev bit for bit.  Calls on one device share
 * the workspace: calls that overlap on one device (other threads, other
 * contexts' streams) race and must be serialised by the caller.  Enqueued
 * on the context stream; returns at once. */
int bh_accel_targets_dev_noalloc(bh_tree *t, const double *d_targets, int64_t m, double theta,
                                 double softening, double G, double *d_out);

/* ------------------------------------------- accelerations (FP32, chosen) */
/* Accelerations at m HOST-resident targets streamed through the device in
 * chunks (the Dask target chunks of bhut/api.py:28-78 and api.py:690-731,
 * map_blocks in bhut/backends/dask_.py:173-228): two lane streams, each
 * driven by a host thread of its own, take alternate chunks through pinned
 * staging memory (gather, copy in, walk, copy out, scatter), so one lane's
 * copies and host work overlap the other's kernels; device memory holds two
 * chunks whatever m is.  precision 64: the plain walk (bh_accel_targets);
 * 32: the FP32 form (bh_accel_f32_targets, opts as there).  chunk <= 0
 * selects 1048576; chunks are rounded up to whole groups of 64.  The targets
 * are put in the tree's Morton order on the host first, so every result is
 * bit for bit what one call on all targets returns.  Synchronous. */
int bh_accel_targets_streamed(bh_tree *t, const double *targets, int64_t m, int64_t chunk, double theta,
                              double softening, double G, int32_t precision, const bh_f32_opts *opts,
                              double *out);

/* Same contract as bh_accel_self / bh_accel_targets, computed in FP32 on
 * coordinates relative to leaf and group centres, with a conservative group
 * opening test; opts may be NULL for defaults.  stats (may be NULL) receives
 * 4 int64: monopole interactions, particle pairs, groups, list entries. */
int bh_accel_f32_self(bh_tree *t, double theta, double softening, double G,
                      const bh_f32_opts *opts, double *out, int64_t *stats);
int bh_accel_f32_self_dev(bh_tree *t, double theta, double softening, double G,
                          const bh_f32_opts *opts, double *d_out, int64_t *stats);
int bh_accel_f32_targets(bh_tree *t, const double *targets, int64_t m, double theta,
                         double softening, double G, const bh_f32_opts *opts, double *out,
                         int64_t *stats);
int bh_accel_f32_targets_dev(bh_tree *t, const double *d_targets, int64_t m, double theta,
                             double softening, double G, const bh_f32_opts *opts, double *d_out,
                             int64_t *stats);

/* ------------------------------------------------------------ potential */
/* Softened gravitational potential phi_i = -G sum m_j / sqrt(r^2 + eps^2)
 * through the plain FP64 walk (the same visits and opening decisions as
 * bh_accel_self / bh_accel_targets), out has m entries in the original
 * target order.  bhut documents `bhut.potential(positions, masses, theta)`
 * (docs/getting-started.md:36) but does not implement it; there is no
 * reference function to follow.  targets == NULL selects self mode
 * (m == n, self pair excluded). */
int bh_potential(bh_tree *t, const double *targets, int64_t m, double theta, double softening, double G,
                 double *out);
int bh_potential_dev(bh_tree *t, const double *d_targets, int64_t m, double theta, double softening,
                     double G, double *d_out);

/* ------------------------------------------------- lists and statistics */
/* Per-target interaction lists of the walk in _barnes_hut_single_target
 * (bh.py:485-548): node ids in pop order, leaves and accepted monopoles.
 * targets == NULL selects self mode (m must equal n).  ptr has m+1 entries
 * (original target order); ids receives min(total, cap) entries. */
int bh_interaction_lists(bh_tree *t, const double *targets, int64_t m, double theta,
                         int64_t *ptr, int32_t *ids, int64_t cap, int64_t *total);
/* Exact per-target counters replacing compute_tree_statistics (bh.py:604-700):
 * out is (m, 5) int64: nodes visited, opening tests, monopoles, leaves,
 * particle pairs (self excluded).  targets == NULL selects self mode. */
int bh_counters(bh_tree *t, const double *targets, int64_t m, double theta, int64_t *out);

/* --------------------------------------------------------- direct sums */
/* FP64 direct summation, the theta = 0 limit of the walk (bh.py:600-601):
 * acc (m, dim) and/or pot (m) (Plummer potential -G m / sqrt(r^2 + eps^2)),
 * either may be NULL.  self_mode skips j == i (m must equal n). */
int bh_direct(bh_ctx *ctx, const double *src, const double *mass, int64_t n, const double *targets,
              int64_t m, int32_t dim, int32_t self_mode, double softening, double G, double *acc,
              double *pot);
int bh_direct_dev(bh_ctx *ctx, const double *d_src, const double *d_mass, int64_t n,
                  const double *d_targets, int64_t m, int32_t dim, int32_t self_mode,
                  double softening, double G, double *d_acc, double *d_pot);

/* ------------------------------------------------------ leapfrog driver */
/* Kick-drift-kick leapfrog over positions / velocities resident in device
 * memory (original particle order, (n, dim) float64; the library updates
 * them in place).  bhut has no integrator: its galaxy example rebuilds a
 * Tree per right-hand-side call (examples/galaxy/galaxy_simulation.py:
 * 277-284).  Each step refits the tree (refit_tree, refit.py:17-109), or
 * rebuilds it every rebuild_every steps (Tree.rebuild, api.py:564-606), and
 * evaluates forces with the plain FP64 walk (precision 64) or the FP32
 * grouped form (32).  With use_graph, each kind of step (refit / rebuild)
 * runs eagerly once, is then captured into a HIP graph and replayed; the
 * simulation runs on a stream of its own and synchronises after every step.
 * Calls on one simulation must not overlap. */
typedef struct bh_sim bh_sim;
int bh_sim_create(bh_ctx *ctx, double *d_pos, double *d_vel, const double *d_mass, int64_t n, int32_t dim,
                  int32_t leaf_size, double theta, double softening, double G, int32_t precision,
                  const bh_f32_opts *opts, bh_sim **out); /* also evaluates the initial forces */
int bh_sim_run(bh_sim *sim, double dt, int64_t steps, int32_t rebuild_every, int32_t use_graph);
/* current accelerations (n, dim) into device memory */
int bh_sim_accelerations(bh_sim *sim, double *d_acc);
/* capacities of the FP32 interaction lists used by captured steps: entries
 * per group of each list region, and work items (both >= 1); too small a
 * capacity is detected, regrown, and the step's force redone eagerly.  Drops
 * the captured graphs. */
int bh_sim_set_list_capacity(bh_sim *sim, int64_t region, int64_t items);
/* 8 int64: steps, rebuilds, refits, graph launches, graph captures, list
 * regrows, list region (entries per group), item capacity */
int bh_sim_info(const bh_sim *sim, int64_t *out);
void bh_sim_destroy(bh_sim *sim);

#ifdef __cplusplus
}
#endif
#endif /* BHUT_HIP_H */
