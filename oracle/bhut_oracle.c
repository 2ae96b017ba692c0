/*
 * bhut_oracle.c -- CPU restatement of bhut's Barnes-Hut force path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path and the CPU
 * baseline of bench.py.  Nothing in the product links or calls this file.
 *
 * Compiled with -ffp-contract=off so every a*b+c below rounds twice, exactly
 * as NumPy/Numba do; the one place the reference fuses (OpenBLAS ddot inside
 * np.linalg.norm) uses an explicit fma().
 *
 * Reference files cited as path:line relative to the bhut repository.
 */
#include "bhut_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* NumPy arithmetic restated                                           */
/* ------------------------------------------------------------------ */

/* numpy's pairwise summation of a contiguous float64 run (the inner loop
 * np.add.reduce uses for a 1-D slice): sequential from -0.0 below 8
 * elements, eight interleaved accumulators up to 128, recursive halving
 * (on a multiple of 8) above.  Used by np.sum at build.py:263,284 and
 * refit.py:140. */
static double np_pairwise(const double *a, int64_t n) {
  if (n < 8) {
    double r = -0.0;
    for (int64_t i = 0; i < n; i++) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int64_t i;
    for (i = 8; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return np_pairwise(a, n2) + np_pairwise(a + n2, n - n2);
}

/* np.sum over a 1-D float64 array: the reduction starts from the additive
 * identity +0.0 and runs its inner loop over buffers of at most 8192
 * elements, adding the pairwise sum of each buffer in turn (pinned against
 * np.sum at lengths up to 250 001). */
double orc_np_sum(const double *a, int64_t n) {
  double acc = 0.0;
  for (int64_t s = 0; s < n; s += 8192) acc += np_pairwise(a + s, n - s < 8192 ? n - s : 8192);
  return acc;
}

/* np.min / np.max along axis 0: NaN propagates (bbox.py:43-44). */
static double np_min2(double acc, double v) {
  if (isnan(acc)) return acc;
  if (isnan(v) || v < acc) return v;
  return acc;
}
static double np_max2(double acc, double v) {
  if (isnan(acc)) return acc;
  if (isnan(v) || v > acc) return v;
  return acc;
}

/* np.linalg.norm of a small float64 vector (bh.py:593): sqrt(x.dot(x)),
 * where OpenBLAS' ddot accumulates with fused multiply-adds from x0*x0. */
static double np_norm(const double *v, int dim) {
  double s = v[0] * v[0];
  for (int d = 1; d < dim; d++) s = fma(v[d], v[d], s);
  return sqrt(s);
}

/* ------------------------------------------------------------------ */
/* bbox.py / morton.py                                                 */
/* ------------------------------------------------------------------ */

/* aabb (bbox.py:14-46). */
void orc_aabb(const double *pos, int64_t n, int dim, double *pmin, double *pmax) {
  for (int d = 0; d < dim; d++) {
    double lo = pos[d], hi = pos[d];
    for (int64_t i = 1; i < n; i++) {
      lo = np_min2(lo, pos[i * dim + d]);
      hi = np_max2(hi, pos[i * dim + d]);
    }
    pmin[d] = lo;
    pmax[d] = hi;
  }
}

/* normalize_to_unit (bbox.py:49-96): (p - pmin) / extent + offset, then
 * np.minimum(., 1 - 1e-12); degenerate axes use extent 1 and offset 0.5. */
void orc_normalize(const double *pos, int64_t n, int dim, const double *pmin,
                   const double *pmax, double *out01) {
  const double cap = 1.0 - 1e-12;
  for (int d = 0; d < dim; d++) {
    double extent = pmax[d] - pmin[d];
    double ext = (extent <= 0) ? 1.0 : extent;
    double off = (extent <= 0) ? 0.5 : 0.0;
    for (int64_t i = 0; i < n; i++) {
      double v = (pos[i * dim + d] - pmin[d]) / ext + off;
      /* np.minimum propagates NaN */
      out01[i * dim + d] = (isnan(v) || v < cap) ? v : cap;
    }
  }
}

/* Quantisation of morton.py:87-89: truncate p*2^21 to uint64 and clamp to
 * 2^21-1.  On the x86-64 hosts NumPy runs on, a NaN converts to 2^63, which
 * the clamp maps to 2^21-1. */
static uint64_t quantize21(double p01) {
  const uint64_t maxc = (1ull << 21) - 1;
  if (isnan(p01)) return maxc;
  double s = p01 * 2097152.0;
  uint64_t q = (s <= 0.0) ? 0 : (uint64_t)s;
  return q > maxc ? maxc : q;
}

/* morton_codes (morton.py:48-102), written as the definition of the code:
 * bit b of axis d lands at bit b*dim + d. */
void orc_morton(const double *p01, int64_t n, int dim, uint64_t *codes) {
  for (int64_t i = 0; i < n; i++) {
    uint64_t c = 0;
    for (int d = 0; d < dim; d++) {
      uint64_t q = quantize21(p01[i * dim + d]);
      for (int b = 0; b < 21; b++) c |= ((q >> b) & 1ull) << (b * dim + d);
    }
    codes[i] = c;
  }
}

/* np.argsort(kind='stable') (morton.py:135): bottom-up merge sort of indices
 * by key, ties kept in index order. */
void orc_argsort_stable(const uint64_t *keys, int64_t n, int64_t *perm) {
  int64_t *tmp = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
  for (int64_t i = 0; i < n; i++) perm[i] = i;
  int64_t *src = perm, *dst = tmp;
  for (int64_t w = 1; w < n; w *= 2) {
    for (int64_t lo = 0; lo < n; lo += 2 * w) {
      int64_t mid = lo + w < n ? lo + w : n;
      int64_t hi = lo + 2 * w < n ? lo + 2 * w : n;
      int64_t a = lo, b = mid, k = lo;
      while (a < mid && b < hi) dst[k++] = (keys[src[b]] < keys[src[a]]) ? src[b++] : src[a++];
      while (a < mid) dst[k++] = src[a++];
      while (b < hi) dst[k++] = src[b++];
    }
    int64_t *t = src; src = dst; dst = t;
  }
  if (src != perm) memcpy(perm, src, (size_t)n * sizeof(int64_t));
  free(tmp);
}

/* ------------------------------------------------------------------ */
/* build.py                                                            */
/* ------------------------------------------------------------------ */

/* Node budget of _build_tree_iterative (build.py:124). */
int64_t orc_max_nodes(int64_t n, int64_t leaf_size) {
  int64_t m = 4 * n / leaf_size;
  return m > 10 ? m : 10;
}

/* _build_tree_iterative (build.py:110-212): LIFO stack of (node, start,
 * end); a node with more than leaf_size particles gets children
 * (node_count, node_count+1) at the count median, left pushed first so the
 * right child is popped (and numbered) first. */
int64_t orc_topology(orc_tree *t) {
  const int64_t N = t->N, L = t->leaf_size;
  const int64_t cap = orc_max_nodes(N, L);
  for (int64_t i = 0; i < cap; i++) {
    t->first_child[i] = -1;
    t->child_count[i] = 0;
    t->start[i] = 0;
    t->count[i] = 0;
    t->is_leaf[i] = 0;
  }
  /* stack depth is bounded by the tree depth + 1 <= 2 + log2(N) */
  int64_t sn[130], ss[130], se[130];
  int sp = 0;
  sn[sp] = 0; ss[sp] = 0; se[sp] = N; sp++;
  int64_t node_count = 1;
  while (sp > 0) {
    sp--;
    int64_t node = sn[sp], s = ss[sp], e = se[sp];
    int64_t c = e - s;
    if (node >= cap) return -1;
    t->start[node] = s;
    t->count[node] = c;
    if (c <= L) {
      t->is_leaf[node] = 1;
      continue;
    }
    t->is_leaf[node] = 0;
    int64_t mid = s + c / 2;
    if (node_count + 2 > cap) { /* build.py:166-169 */
      t->is_leaf[node] = 1;
      continue;
    }
    t->first_child[node] = node_count;
    t->child_count[node] = 2;
    if (s < mid) { sn[sp] = node_count; ss[sp] = s; se[sp] = mid; sp++; }
    if (mid < e) { sn[sp] = node_count + 1; ss[sp] = mid; se[sp] = e; sp++; }
    node_count += 2;
  }
  t->M = node_count;
  return node_count;
}

/* _compute_node_geometry (build.py:215-242): root from the global box,
 * every other node from the tight box of its sorted particles. */
static void geometry(orc_tree *t, const double *spos, const double *pmin, const double *pmax) {
  const int dim = (int)t->dim;
  double ext_max = NAN;
  for (int d = 0; d < dim; d++) {
    t->center[d] = (pmin[d] + pmax[d]) / 2;
    double e = pmax[d] - pmin[d];
    ext_max = d == 0 ? e : np_max2(ext_max, e);
  }
  t->half_size[0] = ext_max / 2;
  for (int64_t k = 1; k < t->M; k++) {
    int64_t s = t->start[k], c = t->count[k];
    if (c <= 0) continue;
    double lo[3], hi[3];
    orc_aabb(spos + s * dim, c, dim, lo, hi);
    double em = NAN;
    for (int d = 0; d < dim; d++) {
      t->center[k * dim + d] = (lo[d] + hi[d]) / 2;
      double e = hi[d] - lo[d];
      em = d == 0 ? e : np_max2(em, e);
    }
    t->half_size[k] = em / 2;
  }
}

/* Leaf mass and centre of mass in NumPy order (build.py:258-272,
 * refit.py:139-149): mass = np.sum(m[s:e]); com = np.sum(m[:,None]*x,
 * axis=0) / mass, the axis-0 sum running sequentially from +0.0 down the
 * rows; mass <= 0 (or NaN) falls back to the geometric centre. */
static void leaf_mass_com(const double *m, const double *x, int64_t c, int dim,
                          const double *center, double *out_mass, double *out_com) {
  double total = orc_np_sum(m, c);
  *out_mass = total;
  if (total > 0) {
    for (int d = 0; d < dim; d++) {
      double s = 0.0;
      for (int64_t i = 0; i < c; i++) {
        double w = m[i] * x[i * dim + d];
        s = s + w;
      }
      out_com[d] = s / total;
    }
  } else {
    for (int d = 0; d < dim; d++) out_com[d] = center[d];
  }
}

/* _compute_mass_and_com (build.py:245-295): nodes from M-1 down to 0;
 * internal nodes sum their two children with np.sum semantics. */
static void mass_com(orc_tree *t, const double *spos, const double *smass) {
  const int dim = (int)t->dim;
  for (int64_t k = t->M - 1; k >= 0; k--) {
    double *com = t->com + k * dim;
    const double *ctr = t->center + k * dim;
    if (t->is_leaf[k]) {
      int64_t s = t->start[k], c = t->count[k];
      if (c > 0) {
        leaf_mass_com(smass + s, spos + s * dim, c, dim, ctr, &t->mass[k], com);
      } else {
        t->mass[k] = 0.0;
        for (int d = 0; d < dim; d++) com[d] = ctr[d];
      }
    } else {
      int64_t fc = t->first_child[k], nc = t->child_count[k];
      if (nc > 0) {
        double cm[2];
        for (int64_t j = 0; j < nc; j++) cm[j] = t->mass[fc + j];
        double total = orc_np_sum(cm, nc);
        if (total > 0) {
          for (int d = 0; d < dim; d++) {
            double s = 0.0;
            for (int64_t j = 0; j < nc; j++) {
              double w = cm[j] * t->com[(fc + j) * dim + d];
              s = s + w;
            }
            com[d] = s / total;
          }
        } else {
          for (int d = 0; d < dim; d++) com[d] = ctr[d];
        }
        t->mass[k] = total;
      } else {
        t->mass[k] = 0.0;
        for (int d = 0; d < dim; d++) com[d] = ctr[d];
      }
    }
  }
}

/* build_tree (build.py:19-107). */
int64_t orc_build(const double *pos, const double *mass, int64_t n, int dim,
                  int64_t leaf_size, orc_tree *t) {
  if (n <= 0) return -2;           /* "Cannot build tree with zero particles" */
  t->N = n;
  t->dim = dim;
  t->leaf_size = leaf_size;
  double pmin[3], pmax[3];
  orc_aabb(pos, n, dim, pmin, pmax);
  double *p01 = (double *)malloc((size_t)n * dim * sizeof(double));
  uint64_t *codes = (uint64_t *)malloc((size_t)n * sizeof(uint64_t));
  orc_normalize(pos, n, dim, pmin, pmax, p01);
  orc_morton(p01, n, dim, codes);
  orc_argsort_stable(codes, n, t->perm);
  free(p01);
  free(codes);
  double *spos = (double *)malloc((size_t)n * dim * sizeof(double));
  double *smass = (double *)malloc((size_t)n * sizeof(double));
  for (int64_t i = 0; i < n; i++) {
    int64_t p = t->perm[i];
    for (int d = 0; d < dim; d++) spos[i * dim + d] = pos[p * dim + d];
    smass[i] = mass[p];
  }
  int64_t M = orc_topology(t);
  if (M < 0) { free(spos); free(smass); return -3; }
  for (int64_t k = 0; k < M; k++) {
    t->half_size[k] = 0.0;
    t->mass[k] = 0.0;
    for (int d = 0; d < dim; d++) { t->center[k * dim + d] = 0.0; t->com[k * dim + d] = 0.0; }
  }
  geometry(t, spos, pmin, pmax);
  mass_com(t, spos, smass);
  free(spos);
  free(smass);
  return M;
}

/* ------------------------------------------------------------------ */
/* refit.py                                                            */
/* ------------------------------------------------------------------ */

/* refit_tree (refit.py:17-109): leaves gather new_positions[perm[s:e]]
 * (refit.py:112-149); internal nodes add only children with mass > 0,
 * starting from 0.0 (refit.py:152-192).  center/half_size stay stale. */
void orc_refit(orc_tree *t, const double *pos, const double *mass) {
  const int dim = (int)t->dim;
  int64_t L = t->leaf_size;
  double *lm = (double *)malloc((size_t)(L > 0 ? L : 1) * sizeof(double));
  double *lx = (double *)malloc((size_t)(L > 0 ? L : 1) * dim * sizeof(double));
  int64_t lcap = L;
  for (int64_t k = 0; k < t->M; k++) {
    if (!t->is_leaf[k]) continue;
    int64_t s = t->start[k], c = t->count[k];
    double *com = t->com + k * dim;
    if (c == 0) {
      t->mass[k] = 0.0;
      for (int d = 0; d < dim; d++) com[d] = 0.0;
      continue;
    }
    if (c > lcap) {
      lm = (double *)realloc(lm, (size_t)c * sizeof(double));
      lx = (double *)realloc(lx, (size_t)c * dim * sizeof(double));
      lcap = c;
    }
    for (int64_t i = 0; i < c; i++) {
      int64_t p = t->perm[s + i];
      lm[i] = mass[p];
      for (int d = 0; d < dim; d++) lx[i * dim + d] = pos[p * dim + d];
    }
    leaf_mass_com(lm, lx, c, dim, t->center + k * dim, &t->mass[k], com);
  }
  free(lm);
  free(lx);
  for (int64_t k = t->M - 1; k >= 0; k--) {
    if (t->is_leaf[k]) continue;
    int64_t fc = t->first_child[k], nc = t->child_count[k];
    double *com = t->com + k * dim;
    if (fc == -1 || nc == 0) {
      t->mass[k] = 0.0;
      for (int d = 0; d < dim; d++) com[d] = t->center[k * dim + d];
      continue;
    }
    double total = 0.0, wc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < nc; j++) {
      int64_t ch = fc + j;
      if (ch >= t->M) break;
      double cmass = t->mass[ch];
      if (cmass > 0) {
        total += cmass;
        for (int d = 0; d < dim; d++) {
          double w = cmass * t->com[ch * dim + d];
          wc[d] = wc[d] + w;
        }
      }
    }
    t->mass[k] = total;
    for (int d = 0; d < dim; d++) com[d] = total > 0 ? wc[d] / total : t->center[k * dim + d];
  }
}

/* should_refit_vs_rebuild (refit.py:195-251). */
int orc_should_refit(const orc_tree *t, const double *pos, int64_t n, int dim) {
  if (n != t->N) return 0;
  if (dim != t->dim) return 0;
  const double lim = 1.1 * t->half_size[0];
  for (int64_t i = 0; i < n; i++) {
    double md = NAN;
    for (int d = 0; d < dim; d++) {
      double a = fabs(pos[i * dim + d] - t->center[d]);
      md = d == 0 ? a : np_max2(md, a);
    }
    if (md > lim) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------ */
/* traverse/bh.py + traverse/kernels.py                                */
/* ------------------------------------------------------------------ */

/* acc_monopole, single-vector branch (kernels.py:57-69):
 * r2 = np.sum(dx**2); r2s = r2 + eps**2; zero if r2s == 0;
 * f = G*m / r2s**1.5 (libm pow); a = f*dx. */
static void acc_monopole(const double *dx, double m, double eps, double G, int dim, double *a) {
  double r2 = -0.0;
  for (int d = 0; d < dim; d++) { double q = dx[d] * dx[d]; r2 += q; }
  r2 = 0.0 + r2;
  double r2s = r2 + eps * eps;
  if (r2s == 0) {
    for (int d = 0; d < dim; d++) a[d] = 0.0;
    return;
  }
  double f = (G * m) / pow(r2s, 1.5);
  for (int d = 0; d < dim; d++) a[d] = f * dx[d];
}

/* _leaf_acceleration_pure_python (bh.py:384-425) and its external twin
 * (bh.py:283-311). */
static void leaf_python(const orc_tree *t, int64_t leaf, const double *tp, int64_t tidx,
                        const double *pos, const double *mass, double eps, double G,
                        double *acc) {
  const int dim = (int)t->dim;
  for (int d = 0; d < dim; d++) acc[d] = 0.0;
  int64_t s = t->start[leaf], c = t->count[leaf];
  if (s == -1 || c <= 0) return;
  for (int64_t i = s; i < s + c; i++) {
    int64_t p = t->perm[i];
    if (p == tidx) continue;
    double dx[3], a[3];
    for (int d = 0; d < dim; d++) dx[d] = pos[p * dim + d] - tp[d];
    acc_monopole(dx, mass[p], eps, G, dim, a);
    for (int d = 0; d < dim; d++) acc[d] = acc[d] + a[d];
  }
}

/* _numba_leaf_particle_loop (bh.py:226-280) and its external twin
 * (bh.py:174-223).  Returns 0 if Numba would raise ZeroDivisionError
 * (r == 0 for a counted pair); the caller then recomputes the whole leaf
 * with the pure-Python loop (bh.py:307-311, 339-343). */
static int leaf_numba(const orc_tree *t, int64_t leaf, const double *tp, int64_t tidx,
                      const double *pos, const double *mass, double eps, double G,
                      double *acc) {
  const int dim = (int)t->dim;
  for (int d = 0; d < dim; d++) acc[d] = 0.0;
  int64_t s = t->start[leaf], c = t->count[leaf];
  if (s == -1 || c <= 0) return 1;
  const double eps2 = eps * eps;
  for (int64_t i = s; i < s + c; i++) {
    int64_t p = t->perm[i];
    if (p == tidx) continue;
    double dx[3];
    for (int d = 0; d < dim; d++) dx[d] = pos[p * dim + d] - tp[d];
    double rsq = 0.0;
    for (int d = 0; d < dim; d++) { double q = dx[d] * dx[d]; rsq += q; }
    rsq += eps2;
    double r = sqrt(rsq);
    if (r == 0.0) return 0;
    double rinv = 1.0 / r;
    double ric = rinv * rinv * rinv;
    double mag = G * mass[p] * ric;
    for (int d = 0; d < dim; d++) { double q = mag * dx[d]; acc[d] += q; }
  }
  return 1;
}

/* Softened potential of one source, G m / sqrt(r2 + eps^2) with r2 summed
 * x, y, z in order (0 when the softened distance is 0).  bhut has no
 * potential; this is the definition the library's walk evaluates. */
static double pair_potential(const double *dx, double m, double eps, double G, int dim) {
  double r2 = dx[0] * dx[0];
  for (int d = 1; d < dim; d++) { double q = dx[d] * dx[d]; r2 = r2 + q; }
  double r2s = r2 + eps * eps;
  if (r2s == 0.0) return 0.0;
  return (G * m) / sqrt(r2s);
}

/* Potential of one leaf: partial sum from 0.0 in leaf order, self skipped. */
static double leaf_potential(const orc_tree *t, int64_t leaf, const double *tp, int64_t tidx,
                             const double *pos, const double *mass, double eps, double G) {
  const int dim = (int)t->dim;
  double pp = 0.0;
  int64_t s = t->start[leaf], c = t->count[leaf];
  if (s == -1 || c <= 0) return 0.0;
  for (int64_t i = s; i < s + c; i++) {
    int64_t p = t->perm[i];
    if (p == tidx) continue;
    double dx[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < dim; d++) dx[d] = pos[p * dim + d] - tp[d];
    pp = pp - pair_potential(dx, mass[p], eps, G, dim);
  }
  return pp;
}

/* _should_open_node (bh.py:576-601): d = np.linalg.norm(target - com);
 * d == 0 always opens; otherwise open iff (2*half_size)/d >= theta. */
static int should_open_theta(const orc_tree *t, int64_t k, const double *tp, double theta) {
  const int dim = (int)t->dim;
  double v[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < dim; d++) v[d] = tp[d] - t->com[k * dim + d];
  double dist = np_norm(v, dim);
  if (dist == 0) return 1;
  double size = 2 * t->half_size[k];
  double ratio = size / dist;
  return ratio >= theta;
}

/* Per-target walk shared by the acceleration, list and counter drivers:
 * _barnes_hut_single_target (bh.py:485-548) / _external (bh.py:83-146).
 * The stack holds at most depth+1 <= 130 entries. */
typedef struct {
  double *acc;        /* accumulate here (NULL: do not compute) */
  double *phi;        /* softened potential (NULL: do not compute) */
  int64_t *list;      /* interaction list (NULL: not recorded) */
  int64_t list_cap, list_len;
  int64_t cnt[5];     /* pops, tests, monopoles, leaves, pairs */
} walk_out;

static void walk(const orc_tree *t, const double *tp, int64_t tidx, const double *pos,
                 const double *mass, double eps, double theta, double G, int leaf_formula,
                 walk_out *w) {
  const int dim = (int)t->dim;
  int64_t stack[130];
  int sp = 0;
  stack[sp++] = 0;
  if (w->acc) for (int d = 0; d < dim; d++) w->acc[d] = 0.0;
  while (sp > 0) {
    int64_t k = stack[--sp];
    w->cnt[0]++;
    if (k < 0 || k >= t->M) continue;
    if (t->mass[k] == 0) continue;
    if (t->first_child[k] == -1) {
      w->cnt[3]++;
      int64_t pairs = t->count[k];
      if (tidx >= 0)
        for (int64_t i = t->start[k]; i < t->start[k] + t->count[k]; i++)
          if (t->perm[i] == tidx) pairs--;
      w->cnt[4] += pairs;
      if (w->list) { if (w->list_len < w->list_cap) w->list[w->list_len] = k; w->list_len++; }
      if (w->phi) *w->phi = *w->phi + leaf_potential(t, k, tp, tidx, pos, mass, eps, G);
      if (w->acc) {
        double a[3];
        if (leaf_formula == ORC_LEAF_PYTHON || !leaf_numba(t, k, tp, tidx, pos, mass, eps, G, a))
          leaf_python(t, k, tp, tidx, pos, mass, eps, G, a);
        for (int d = 0; d < dim; d++) w->acc[d] = w->acc[d] + a[d];
      }
    } else {
      w->cnt[1]++;
      if (should_open_theta(t, k, tp, theta)) {
        int64_t ch = t->first_child[k];
        for (int64_t i = 0; i < t->child_count[k]; i++) {
          if (ch != -1) { stack[sp++] = ch; ch += 1; }
        }
      } else {
        w->cnt[2]++;
        if (w->list) { if (w->list_len < w->list_cap) w->list[w->list_len] = k; w->list_len++; }
        if (w->phi) {
          double dx[3] = {0.0, 0.0, 0.0};
          for (int d = 0; d < dim; d++) dx[d] = t->com[k * dim + d] - tp[d];
          *w->phi = *w->phi - pair_potential(dx, t->mass[k], eps, G, dim);
        }
        if (w->acc) {
          double dx[3], a[3];
          for (int d = 0; d < dim; d++) dx[d] = t->com[k * dim + d] - tp[d];
          acc_monopole(dx, t->mass[k], eps, G, dim, a);
          for (int d = 0; d < dim; d++) w->acc[d] = w->acc[d] + a[d];
        }
      }
    }
  }
}

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* barnes_hut_accelerations (bh.py:428-482) / _targets (bh.py:30-80):
 * targets in original order, each walked independently. */
void orc_accel(const orc_tree *t, const double *src_pos, const double *src_mass,
               const double *tgt, int64_t m, int self_mode, double eps, double theta,
               double G, int leaf_formula, int nthreads, double *out) {
  const int dim = (int)t->dim;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; i++) {
    walk_out w;
    memset(&w, 0, sizeof(w));
    w.acc = out + i * dim;
    walk(t, tgt + i * dim, self_mode ? i : -1, src_pos, src_mass, eps, theta, G, leaf_formula, &w);
  }
}

/* orc_accel in self mode for the particles idx[0..k) only: particle idx[q]
 * is the target and skips itself (bh.py:475-480 for i = idx[q]); out is
 * (k, dim). */
void orc_accel_subset(const orc_tree *t, const double *src_pos, const double *src_mass,
                      const int64_t *idx, int64_t k, double eps, double theta, double G,
                      int leaf_formula, int nthreads, double *out) {
  const int dim = (int)t->dim;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < k; q++) {
    walk_out w;
    memset(&w, 0, sizeof(w));
    w.acc = out + q * dim;
    walk(t, src_pos + idx[q] * dim, idx[q], src_pos, src_mass, eps, theta, G, leaf_formula, &w);
  }
}

/* Potential over the same walk as orc_accel: accepted nodes and leaf partial
 * sums added in pop order. */
void orc_potential(const orc_tree *t, const double *src_pos, const double *src_mass,
                   const double *tgt, int64_t m, int self_mode, double eps, double theta,
                   double G, int nthreads, double *out) {
  const int dim = (int)t->dim;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; i++) {
    walk_out w;
    memset(&w, 0, sizeof(w));
    out[i] = 0.0;
    w.phi = out + i;
    walk(t, tgt + i * dim, self_mode ? i : -1, src_pos, src_mass, eps, theta, G, ORC_LEAF_NUMBA, &w);
  }
}

int64_t orc_target_list(const orc_tree *t, const double *tpos, double theta, int64_t *out,
                        int64_t cap) {
  walk_out w;
  memset(&w, 0, sizeof(w));
  w.list = out;
  w.list_cap = cap;
  walk(t, tpos, -1, NULL, NULL, 0.0, theta, 1.0, ORC_LEAF_NUMBA, &w);
  return w.list_len;
}

/* Exact counters; the reference's compute_tree_statistics (bh.py:604-700)
 * counts count-1 pairs per leaf whether or not the target is in it, so it
 * is not used here. */
void orc_counters(const orc_tree *t, const double *tgt, int64_t m, int self_mode,
                  double theta, int nthreads, int64_t *out) {
  const int dim = (int)t->dim;
  set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < m; i++) {
    walk_out w;
    memset(&w, 0, sizeof(w));
    walk(t, tgt + i * dim, self_mode ? i : -1, NULL, NULL, 0.0, theta, 1.0, ORC_LEAF_NUMBA, &w);
    for (int j = 0; j < 5; j++) out[i * 5 + j] = w.cnt[j];
  }
}

void orc_direct(const double *src_pos, const double *src_mass, int64_t n, const double *tgt,
                int64_t m, int dim, int self_mode, double eps, double G, int nthreads,
                double *out) {
  set_threads(nthreads);
  const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < m; i++) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 0; j < n; j++) {
      if (self_mode && j == i) continue;
      double dx[3], rsq = 0.0;
      for (int d = 0; d < dim; d++) { dx[d] = src_pos[j * dim + d] - tgt[i * dim + d]; double q = dx[d] * dx[d]; rsq += q; }
      rsq += eps2;
      if (rsq == 0.0) continue;
      double rinv = 1.0 / sqrt(rsq);
      double mag = G * src_mass[j] * (rinv * rinv * rinv);
      for (int d = 0; d < dim; d++) { double q = mag * dx[d]; acc[d] += q; }
    }
    for (int d = 0; d < dim; d++) out[i * dim + d] = acc[d];
  }
}

void orc_direct_potential(const double *src_pos, const double *src_mass, int64_t n,
                          const double *tgt, int64_t m, int dim, int self_mode, double eps,
                          double G, int nthreads, double *out) {
  set_threads(nthreads);
  const double eps2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = 0; i < m; i++) {
    double phi = 0.0;
    for (int64_t j = 0; j < n; j++) {
      if (self_mode && j == i) continue;
      double rsq = 0.0;
      for (int d = 0; d < dim; d++) { double dx = src_pos[j * dim + d] - tgt[i * dim + d]; double q = dx * dx; rsq += q; }
      rsq += eps2;
      if (rsq == 0.0) continue;
      double q = (G * src_mass[j]) / sqrt(rsq);
      phi -= q;
    }
    out[i] = phi;
  }
}
