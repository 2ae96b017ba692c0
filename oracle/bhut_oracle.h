/*
 * bhut_oracle.h -- CPU restatement of bhut's Barnes-Hut force path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * implementation in bhut_amd/csrc and the CPU baseline timed by bench.py.
 * It is never linked into, loaded by, or called from the product path.
 *
 * Every function restates the reference algorithm (maroba/bhut) step by
 * step, in the same floating-point order, and cites the file:line it follows.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#ifndef BHUT_ORACLE_H
#define BHUT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirror of bhut/tree/node.py:15-77 (TreeData) with int64 indices. */
typedef struct {
  int64_t M;          /* number of nodes */
  int64_t N;          /* number of particles */
  int64_t dim;        /* 2 or 3 */
  int64_t leaf_size;
  double *center;     /* M x dim */
  double *half_size;  /* M */
  double *mass;       /* M */
  double *com;        /* M x dim */
  int64_t *first_child; /* M, -1 for leaves */
  int64_t *child_count; /* M */
  int64_t *start;       /* M */
  int64_t *count;       /* M */
  uint8_t *is_leaf;     /* M */
  int64_t *perm;        /* N */
} orc_tree;

/* Leaf formula selector: the Numba loop (bh.py:226-280, CI default with
 * the [accel] extra) or the pure-Python loop (bh.py:384-425). */
enum { ORC_LEAF_NUMBA = 0, ORC_LEAF_PYTHON = 1 };

/* numpy float64 sum of a contiguous 1-D slice (pairwise order). */
double orc_np_sum(const double *a, int64_t n);

void orc_aabb(const double *pos, int64_t n, int dim, double *pmin, double *pmax);
void orc_normalize(const double *pos, int64_t n, int dim, const double *pmin,
                   const double *pmax, double *out01);
void orc_morton(const double *p01, int64_t n, int dim, uint64_t *codes);
void orc_argsort_stable(const uint64_t *keys, int64_t n, int64_t *perm);

int64_t orc_max_nodes(int64_t n, int64_t leaf_size);
/* Fills the topology arrays of t (allocated with orc_max_nodes entries) and
 * sets t->M.  Returns M, or -1 on node-index overflow (build.py:148-149). */
int64_t orc_topology(orc_tree *t);
/* Full build_tree (build.py:19-107).  t arrays must hold orc_max_nodes(n)
 * nodes and n perm entries.  Returns M (>0) or a negative error. */
int64_t orc_build(const double *pos, const double *mass, int64_t n, int dim,
                  int64_t leaf_size, orc_tree *t);
/* refit_tree (refit.py:17-109): writes mass/com of t in place from new
 * original-order positions and masses; geometry is left untouched. */
void orc_refit(orc_tree *t, const double *pos, const double *mass);
/* should_refit_vs_rebuild (refit.py:195-251): returns 1 for refit. */
int orc_should_refit(const orc_tree *t, const double *pos, int64_t n, int dim);

/* Accelerations.  self_mode=1: targets are the sources (bh.py:428-548) and
 * target i skips source i; self_mode=0: external targets (bh.py:30-223).
 * nthreads <= 0 uses the OpenMP default. */
void orc_accel(const orc_tree *t, const double *src_pos, const double *src_mass,
               const double *tgt, int64_t m, int self_mode, double eps,
               double theta, double G, int leaf_formula, int nthreads,
               double *out);
void orc_accel_subset(const orc_tree *t, const double *src_pos, const double *src_mass,
                      const int64_t *idx, int64_t k, double eps, double theta, double G,
                      int leaf_formula, int nthreads, double *out);

/* Softened potential -G sum m / sqrt(r2 + eps2) over the same walk
 * (accepted nodes and leaf partial sums in pop order, self skipped in
 * self_mode).  No bhut counterpart: the definition the library follows. */
void orc_potential(const orc_tree *t, const double *src_pos, const double *src_mass,
                   const double *tgt, int64_t m, int self_mode, double eps, double theta,
                   double G, int nthreads, double *out);

/* Interaction list of one target in pop order: node ids of leaves summed
 * directly and internal nodes used as monopoles.  Returns the list length
 * (entries beyond cap are counted but not written). */
int64_t orc_target_list(const orc_tree *t, const double *tpos, double theta,
                        int64_t *out, int64_t cap);

/* Exact per-target counters: pops, opening tests, monopoles, leaves,
 * particle-particle pairs (self excluded).  out has 5*m entries. */
void orc_counters(const orc_tree *t, const double *tgt, int64_t m,
                  int self_mode, double theta, int nthreads, int64_t *out);

/* FP64 direct sum in source index order with the Numba formula; self_mode
 * skips j == i.  Pairs with r2+eps2 == 0 contribute zero. */
void orc_direct(const double *src_pos, const double *src_mass, int64_t n,
                const double *tgt, int64_t m, int dim, int self_mode,
                double eps, double G, int nthreads, double *out);
/* FP64 direct-sum Plummer-softened potential -G m / sqrt(r2 + eps2). */
void orc_direct_potential(const double *src_pos, const double *src_mass,
                          int64_t n, const double *tgt, int64_t m, int dim,
                          int self_mode, double eps, double G, int nthreads,
                          double *out);

#ifdef __cplusplus
}
#endif
#endif
