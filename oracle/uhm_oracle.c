/*
 * uhm_oracle.c -- CPU restatement of the reference's multifrontal
 * factorize / solve path.  TEST INFRASTRUCTURE ONLY: it is the checker for
 * the GPU library and the timed CPU baseline; nothing in the product links,
 * loads or calls it.
 *
 * It follows the reference algorithm step by step (file:line per function):
 *   analysis   Mesh_::lock             UHM/src/mesh/mesh.cxx:438-456
 *              Element_ connectivity   UHM/src/mesh/element.cxx:272-392
 *              Scheduler_::load        UHM/src/operation/scheduler.cxx:157-178
 *   mapper     Helper_::set_mapper     UHM/include/uhm/matrix/uhm/helper.hxx:98-158
 *   extend-add Helper_::_merge_A       helper.hxx:191-222, fla/matrix.cxx:284-303
 *   LU front   lu_piv_hier             UHM/src/matrix/uhm/fla/lu_piv/decompose.cxx:124-153
 *              linal::dense::lu_piv    LINAL/src/dense/lu_piv.cxx:47-111
 *              apply_pivots            LINAL/src/dense/pivot.cxx:46-91
 *   Chol front chol_hier               UHM/src/matrix/uhm/fla/chol/decompose.cxx:85-98
 *              linal::dense::chol      LINAL/src/dense/chol.cxx:46-92
 *   solves     solve_lu_piv_{1,2}_hier lu_piv/solve.cxx:161-231
 *              solve_chol_{1,2}_hier   chol/solve.cxx:152-202
 *   flops      LAWN-41                 LINAL/src/util.cxx:56-101, UHM/src/mesh/element.cxx:128-195
 *
 * libflame's FLA_LU_piv / FLA_Chol (not in this container) are restated as
 * the unblocked GEPP of LAPACK dgetf2 (first maximal |a|, LAPACK i?amax;
 * multipliers by the reciprocal of a normal pivot) and unblocked Cholesky
 * on each b-wide panel.  The right-hand side is merged
 * once at solve time (x_f = b_f + sum of the children's xb); the reference
 * merges the b part during factorization and the update during the solve,
 * which is the same sum in another order.
 *
 * This file is test infrastructure: only the tests, the smoke check and the
 * CPU-baseline leg of the benchmark load it.  Parity status, each pinned by
 * tests/test_oracle.py:
 *   front LU        LAPACK getrf golden factors and pivots (scipy, committed)
 *   front Cholesky  numpy.linalg.cholesky golden factors (committed)
 *   lock / levels   independent lowest-common-ancestor formulation, hand-
 *                   derived maps of the two-element example mesh
 *   mapper          run-partition invariants, hand-derived example maps
 *   solve           dense numpy.linalg.solve of the assembled matrix (committed)
 *   flops           LAWN-41 closed forms
 * Not pinnable here: libflame's blocked-panel rounding (the reference's
 * tests store no values); the restatement follows LAPACK's unblocked GEPP.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { SEP_NONE = 0, SEP_FACTOR = 1, SEP_SCHUR = 2 };

typedef struct {
  int n;       /* number of nodes in the element's node set */
  int* node;   /* sorted ascending node index */
  int* flag;   /* separation flag */
  int cap;
} nset;

typedef struct {
  /* mesh */
  int n_nodes, n_elem;
  int *ndof, *kind;
  int *parent, *cptr, *cidx;
  /* analysis */
  nset* nodes;
  int** owner;
  int *n_owner, *cap_owner;
  int* gen;
  int* offset;
  int *fs, *ss;
  int **fac, **facoff, **sch, **schoff;
  int *nfac, *nsch;
  int n_levels;
  int *lvl_ptr, *lvl_elem; /* elements per generation, leaves first, id order */
  /* numeric */
  int nrhs;
  double** F;   /* n x n fronts, column major (working copy) */
  double** K;   /* n x n assembly input of each front (copy-in), zero if none */
  int** piv;
  double** X;   /* n x nrhs */
  double** B;   /* n x nrhs (leaf right-hand sides) */
  double** marg; /* per front, fs relative pivot margins (LU): (|best| - |second|) / |best| */
  int info;
} orc;

/* ------------------------------------------------------------------ sets */
static void ns_insert(nset* s, int node, int flag, int* is_new) {
  int lo = 0, hi = s->n;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    if (s->node[mid] < node) lo = mid + 1;
    else hi = mid;
  }
  if (lo < s->n && s->node[lo] == node) {
    s->flag[lo] = flag;
    *is_new = 0;
    return;
  }
  if (s->n == s->cap) {
    s->cap = s->cap ? 2 * s->cap : 16;
    s->node = (int*)realloc(s->node, s->cap * sizeof(int));
    s->flag = (int*)realloc(s->flag, s->cap * sizeof(int));
  }
  memmove(s->node + lo + 1, s->node + lo, (s->n - lo) * sizeof(int));
  memmove(s->flag + lo + 1, s->flag + lo, (s->n - lo) * sizeof(int));
  s->node[lo] = node;
  s->flag[lo] = flag;
  s->n++;
  *is_new = 1;
}

static void own_add(orc* o, int node, int e) {
  for (int k = 0; k < o->n_owner[node]; ++k)
    if (o->owner[node][k] == e) return;
  if (o->n_owner[node] == o->cap_owner[node]) {
    o->cap_owner[node] = o->cap_owner[node] ? 2 * o->cap_owner[node] : 4;
    o->owner[node] = (int*)realloc(o->owner[node], o->cap_owner[node] * sizeof(int));
  }
  o->owner[node][o->n_owner[node]++] = e;
}

static void own_remove(orc* o, int node, int e) {
  for (int k = 0; k < o->n_owner[node]; ++k)
    if (o->owner[node][k] == e) {
      o->owner[node][k] = o->owner[node][--o->n_owner[node]];
      return;
    }
}

static int is_leaf(const orc* o, int e) { return o->cptr[e + 1] == o->cptr[e]; }

/* ------------------------------------------------------------ analysis */
/* Element_::update_generation (element.cxx:63-72), post order. */
static int gen_rec(orc* o, int e) {
  if (is_leaf(o, e)) return o->gen[e] = 0;
  int g = 0;
  for (int k = o->cptr[e]; k < o->cptr[e + 1]; ++k) {
    int cg = gen_rec(o, o->cidx[k]);
    if (cg < g) g = cg;
  }
  return o->gen[e] = g - 1;
}

static long long key_of(const orc* o, int node) {
  return o->offset[node] >= 0 ? (long long)o->offset[node] : 2147483647LL + node;
}

static const orc* g_sort_orc;
static int cmp_by_offset(const void* a, const void* b) {
  long long x = key_of(g_sort_orc, *(const int*)a), y = key_of(g_sort_orc, *(const int*)b);
  return (x > y) - (x < y);
}

void* orc_analyze(int n_nodes, const int* ndof, const int* kind, int n_elem, const int* eptr, const int* enode,
                  const int* parent, const int* cptr, const int* cidx) {
  orc* o = (orc*)calloc(1, sizeof(orc));
  o->n_nodes = n_nodes;
  o->n_elem = n_elem;
  o->ndof = (int*)malloc(n_nodes * sizeof(int));
  o->kind = (int*)malloc(n_nodes * sizeof(int));
  memcpy(o->ndof, ndof, n_nodes * sizeof(int));
  memcpy(o->kind, kind, n_nodes * sizeof(int));
  o->parent = (int*)malloc(n_elem * sizeof(int));
  memcpy(o->parent, parent, n_elem * sizeof(int));
  o->cptr = (int*)malloc((n_elem + 1) * sizeof(int));
  memcpy(o->cptr, cptr, (n_elem + 1) * sizeof(int));
  o->cidx = (int*)malloc((cptr[n_elem] + 1) * sizeof(int));
  memcpy(o->cidx, cidx, cptr[n_elem] * sizeof(int));
  o->nodes = (nset*)calloc(n_elem, sizeof(nset));
  o->owner = (int**)calloc(n_nodes, sizeof(int*));
  o->n_owner = (int*)calloc(n_nodes, sizeof(int));
  o->cap_owner = (int*)calloc(n_nodes, sizeof(int));
  o->gen = (int*)calloc(n_elem, sizeof(int));
  o->offset = (int*)malloc(n_nodes * sizeof(int));
  for (int i = 0; i < n_nodes; ++i) o->offset[i] = -1;
  o->fs = (int*)calloc(n_elem, sizeof(int));
  o->ss = (int*)calloc(n_elem, sizeof(int));
  o->fac = (int**)calloc(n_elem, sizeof(int*));
  o->facoff = (int**)calloc(n_elem, sizeof(int*));
  o->sch = (int**)calloc(n_elem, sizeof(int*));
  o->schoff = (int**)calloc(n_elem, sizeof(int*));
  o->nfac = (int*)calloc(n_elem, sizeof(int));
  o->nsch = (int*)calloc(n_elem, sizeof(int));

  /* leaves hold the user's node lists; Element_::add_node registers owners */
  for (int e = 0; e < n_elem; ++e)
    for (int k = eptr[e]; k < eptr[e + 1]; ++k) {
      int is_new;
      ns_insert(&o->nodes[e], enode[k], SEP_NONE, &is_new);
      if (is_new) own_add(o, enode[k], e);
    }
  for (int e = 0; e < n_elem; ++e)
    if (o->parent[e] < 0) gen_rec(o, e);

  /* Scheduler_::load: generations, deepest (0) first, element id order */
  int gmin = 0;
  for (int e = 0; e < n_elem; ++e)
    if (o->gen[e] < gmin) gmin = o->gen[e];
  o->n_levels = 1 - gmin;
  o->lvl_ptr = (int*)calloc(o->n_levels + 1, sizeof(int));
  o->lvl_elem = (int*)malloc(n_elem * sizeof(int));
  for (int e = 0; e < n_elem; ++e) o->lvl_ptr[-o->gen[e] + 1]++;
  for (int L = 0; L < o->n_levels; ++L) o->lvl_ptr[L + 1] += o->lvl_ptr[L];
  {
    int* pos = (int*)malloc(o->n_levels * sizeof(int));
    for (int L = 0; L < o->n_levels; ++L) pos[L] = o->lvl_ptr[L];
    for (int e = 0; e < n_elem; ++e) o->lvl_elem[pos[-o->gen[e]]++] = e;
    free(pos);
  }

  /* restore_connectivity on leaves (element.cxx:362-381) */
  for (int e = 0; e < n_elem; ++e) {
    if (!is_leaf(o, e)) continue;
    nset* s = &o->nodes[e];
    for (int k = 0; k < s->n; ++k) {
      int nd = s->node[k];
      s->flag[k] = SEP_NONE;
      /* Node_::clean_connectivity: drop non-leaf owners */
      for (int q = 0; q < o->n_owner[nd];) {
        if (!is_leaf(o, o->owner[nd][q])) o->owner[nd][q] = o->owner[nd][--o->n_owner[nd]];
        else ++q;
      }
      own_add(o, nd, e);
    }
  }
  /* op_update_connectivity leaf -> root (operation/element.cxx:342-349) */
  for (int t = 0; t < n_elem; ++t) {
    int e = o->lvl_elem[t];
    if (!is_leaf(o, e)) {
      /* merge_nodes_from_children (element.cxx:336-360) */
      nset* s = &o->nodes[e];
      s->n = 0;
      for (int k = o->cptr[e]; k < o->cptr[e + 1]; ++k) {
        int c = o->cidx[k];
        nset* cs = &o->nodes[c];
        for (int q = 0; q < cs->n; ++q)
          if (cs->flag[q] == SEP_SCHUR) {
            int is_new;
            ns_insert(s, cs->node[q], SEP_NONE, &is_new);
            if (is_new) own_add(o, cs->node[q], e);
            own_remove(o, cs->node[q], c);
          }
      }
    }
    /* separate_nodes (element.cxx:272-286) */
    nset* s = &o->nodes[e];
    for (int q = 0; q < s->n; ++q) {
      int nd = s->node[q];
      s->flag[q] = (o->n_owner[nd] == 1 && o->kind[nd] == 0) ? SEP_FACTOR : SEP_SCHUR;
    }
  }
  /* numbering (element.cxx:383-392, util.cxx:59-63) */
  {
    int g = 0;
    for (int t = 0; t < n_elem; ++t) {
      int e = o->lvl_elem[t];
      nset* s = &o->nodes[e];
      for (int q = 0; q < s->n; ++q)
        if (s->flag[q] == SEP_FACTOR) {
          o->offset[s->node[q]] = g;
          g += o->ndof[s->node[q]];
        }
    }
  }
  /* arrange_nodes (element.cxx:288-334) */
  g_sort_orc = o;
  for (int e = 0; e < n_elem; ++e) {
    nset* s = &o->nodes[e];
    int nf = 0, nsc = 0;
    int* f = (int*)malloc((s->n + 1) * sizeof(int));
    int* sc = (int*)malloc((s->n + 1) * sizeof(int));
    for (int q = 0; q < s->n; ++q) {
      if (!o->ndof[s->node[q]]) continue;
      if (s->flag[q] == SEP_FACTOR) f[nf++] = s->node[q];
      else sc[nsc++] = s->node[q];
    }
    qsort(f, nf, sizeof(int), cmp_by_offset);
    qsort(sc, nsc, sizeof(int), cmp_by_offset);
    o->fac[e] = f;
    o->sch[e] = sc;
    o->nfac[e] = nf;
    o->nsch[e] = nsc;
    o->facoff[e] = (int*)malloc((nf + 1) * sizeof(int));
    o->schoff[e] = (int*)malloc((nsc + 1) * sizeof(int));
    int off = 0;
    for (int q = 0; q < nf; ++q) {
      o->facoff[e][q] = off;
      off += o->ndof[f[q]];
    }
    o->fs[e] = off;
    off = 0;
    for (int q = 0; q < nsc; ++q) {
      o->schoff[e][q] = off;
      off += o->ndof[sc[q]];
    }
    o->ss[e] = off;
  }
  return o;
}

void orc_free(void* h) {
  orc* o = (orc*)h;
  if (!o) return;
  for (int e = 0; e < o->n_elem; ++e) {
    free(o->nodes[e].node);
    free(o->nodes[e].flag);
    free(o->fac[e]);
    free(o->facoff[e]);
    free(o->sch[e]);
    free(o->schoff[e]);
    if (o->F) free(o->F[e]);
    if (o->K) free(o->K[e]);
    if (o->piv) free(o->piv[e]);
    if (o->X) free(o->X[e]);
    if (o->B) free(o->B[e]);
    if (o->marg) free(o->marg[e]);
  }
  for (int i = 0; i < o->n_nodes; ++i) free(o->owner[i]);
  free(o->nodes);
  free(o->owner);
  free(o->n_owner);
  free(o->cap_owner);
  free(o->gen);
  free(o->offset);
  free(o->fs);
  free(o->ss);
  free(o->fac);
  free(o->facoff);
  free(o->sch);
  free(o->schoff);
  free(o->nfac);
  free(o->nsch);
  free(o->lvl_ptr);
  free(o->lvl_elem);
  free(o->F);
  free(o->K);
  free(o->marg);
  free(o->piv);
  free(o->X);
  free(o->B);
  free(o->ndof);
  free(o->kind);
  free(o->parent);
  free(o->cptr);
  free(o->cidx);
  free(o);
}

int orc_n_levels(void* h) { return ((orc*)h)->n_levels; }
int orc_level(void* h, int L, int* elems) {
  orc* o = (orc*)h;
  int n = o->lvl_ptr[L + 1] - o->lvl_ptr[L];
  if (elems) memcpy(elems, o->lvl_elem + o->lvl_ptr[L], n * sizeof(int));
  return n;
}
void orc_element(void* h, int e, int* out) {
  orc* o = (orc*)h;
  out[0] = o->gen[e];
  out[1] = o->fs[e];
  out[2] = o->ss[e];
  out[3] = o->nfac[e];
  out[4] = o->nsch[e];
}
void orc_lists(void* h, int e, int* fac, int* facoff, int* sch, int* schoff) {
  orc* o = (orc*)h;
  memcpy(fac, o->fac[e], o->nfac[e] * sizeof(int));
  memcpy(facoff, o->facoff[e], o->nfac[e] * sizeof(int));
  memcpy(sch, o->sch[e], o->nsch[e] * sizeof(int));
  memcpy(schoff, o->schoff[e], o->nsch[e] * sizeof(int));
}
int orc_node_offset(void* h, int node) { return ((orc*)h)->offset[node]; }

/* Helper_::set_mapper (helper.hxx:98-158): runs {offs_c, offs_p, fs_p, n}. */
static int find_local(const orc* o, int e, int node, int* fs_p) {
  for (int q = 0; q < o->nfac[e]; ++q)
    if (o->fac[e][q] == node) {
      *fs_p = 0;
      return o->facoff[e][q];
    }
  for (int q = 0; q < o->nsch[e]; ++q)
    if (o->sch[e][q] == node) {
      *fs_p = 1;
      return o->schoff[e][q];
    }
  return -1;
}

static int mapper_list(const orc* o, int p, int nlist, const int* nodes, const int* offs, int* runs) {
  int nr = 0;
  int prev_fs = -1, prev_off = 0, prev_n = 0;
  for (int q = 0; q < nlist; ++q) {
    int fs_p;
    int op = find_local(o, p, nodes[q], &fs_p);
    if (op < 0) return -1;
    int nd = o->ndof[nodes[q]];
    if (q && fs_p == prev_fs && op - prev_off == prev_n) {
      runs[4 * (nr - 1) + 3] += nd;
    } else {
      runs[4 * nr + 0] = offs[q];
      runs[4 * nr + 1] = op;
      runs[4 * nr + 2] = fs_p;
      runs[4 * nr + 3] = nd;
      nr++;
    }
    prev_fs = fs_p;
    prev_off = op;
    prev_n = nd;
  }
  return nr;
}

int orc_mapper(void* h, int c, int* runs) {
  orc* o = (orc*)h;
  int p = o->parent[c];
  if (p < 0) return 0;
  return mapper_list(o, p, o->nsch[c], o->sch[c], o->schoff[c], runs);
}

/* ------------------------------------------------------------ numerics */
static void alloc_numeric(orc* o, int nrhs) {
  if (o->F && o->nrhs == nrhs) return;
  o->nrhs = nrhs;
  o->F = (double**)calloc(o->n_elem, sizeof(double*));
  o->K = (double**)calloc(o->n_elem, sizeof(double*));
  o->piv = (int**)calloc(o->n_elem, sizeof(int*));
  o->X = (double**)calloc(o->n_elem, sizeof(double*));
  o->B = (double**)calloc(o->n_elem, sizeof(double*));
  o->marg = (double**)calloc(o->n_elem, sizeof(double*));
}

/* Per-front storage is allocated on first use, so a sample of levels of a
 * large mesh fits in host memory.  The assembly input K exists only for
 * fronts that receive a copy-in (the leaves); the others start from zero. */
static void ensure(orc* o, int e) {
  if (o->F[e]) return;
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  o->piv[e] = (int*)calloc(o->fs[e] + 1, sizeof(int));
  o->marg[e] = (double*)calloc(o->fs[e] + 1, sizeof(double));
  o->X[e] = (double*)calloc(n * o->nrhs + 1, sizeof(double));
  o->B[e] = (double*)calloc(n * o->nrhs + 1, sizeof(double));
  o->F[e] = (double*)calloc(n * n + 1, sizeof(double));
}

static void ensure_input(orc* o, int e) {
  ensure(o, e);
  if (o->K[e]) return;
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  o->K[e] = (double*)calloc(n * n + 1, sizeof(double));
}

void orc_init_numeric(void* h, int nrhs) { alloc_numeric((orc*)h, nrhs); }

/* copy_in through a phony child (UHM/src/mesh/interf.cxx:108-204):
 * F[map i, map j] += K[i, j]; user node order nods[0..nn). */
static int user_map(const orc* o, int e, int nn, const int* nods, int* map) {
  int m = 0;
  for (int q = 0; q < nn; ++q) {
    int fs_p, op = find_local(o, e, nods[q], &fs_p);
    int nd = o->ndof[nods[q]];
    if (!nd) continue;
    if (op < 0) return -1;
    for (int d = 0; d < nd; ++d) map[m + d] = (fs_p ? o->fs[e] : 0) + op + d;
    m += nd;
  }
  return m;
}

int orc_copy_in(void* h, int e, int nn, const int* nods, int side, int m, const double* buf) {
  orc* o = (orc*)h;
  int* map = (int*)malloc((m + 1) * sizeof(int));
  int mm = user_map(o, e, nn, nods, map);
  if (mm != m) {
    free(map);
    return -1;
  }
  ensure(o, e);
  int n = o->fs[e] + o->ss[e];
  if (side == 1) {
    ensure_input(o, e);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) o->K[e][map[i] + (size_t)map[j] * n] += buf[i + (size_t)j * m];
  } else {
    for (int r = 0; r < o->nrhs; ++r)
      for (int i = 0; i < m; ++i) o->B[e][map[i] + (size_t)r * n] += buf[i + (size_t)r * m];
  }
  free(map);
  return 0;
}

/* front-order input (the assembled leaf stage) */
void orc_set_front(void* h, int e, const double* F) {
  orc* o = (orc*)h;
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  ensure_input(o, e);
  memcpy(o->K[e], F, n * n * sizeof(double));
}
void orc_set_rhs_front(void* h, int e, const double* b) {
  orc* o = (orc*)h;
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  ensure(o, e);
  memcpy(o->B[e], b, n * o->nrhs * sizeof(double));
}
/* replace the working values of front e (n x n), e.g. by factored values
 * computed elsewhere, as the children's input of orc_assemble */
void orc_set_factored(void* h, int e, const double* F) {
  orc* o = (orc*)h;
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  ensure(o, e);
  memcpy(o->F[e], F, n * n * sizeof(double));
}
/* the assembly input of front e (n x n, front order; zero without copy-in) */
void orc_get_input(void* h, int e, double* out) {
  orc* o = (orc*)h;
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  ensure(o, e);
  if (o->K[e]) memcpy(out, o->K[e], n * n * sizeof(double));
  else memset(out, 0, n * n * sizeof(double));
}

/* ---- dense kernels on column-major views (ld = front size)
 * Every entry accumulates its updates in the order of the textbook loops
 * (ascending inner index p), whatever the blocking and the thread count, so
 * a result does not depend on either.  par: OpenMP inside the kernel (one
 * front at a time); fronts processed in parallel pass 0. */
#define AT(A, ld, i, j) ((A)[(size_t)(i) + (size_t)(j) * (ld)])
#define KB_ 64   /* inner-index block */
#define IB_ 256  /* row block */
static inline int imin(int a, int b) { return a < b ? a : b; }

/* c[i0..ie) -= sum over p in [p0, pe) of a_p[i] * b[p], p ascending */
static inline void axpy_block(double* c, int i0, int ie, const double* A, int lda, int p0, int pe, const double* b,
                              int bstride) {
  int p = p0;
  for (; p + 4 <= pe; p += 4) {
    const double b0 = b[(size_t)p * bstride], b1 = b[(size_t)(p + 1) * bstride];
    const double b2 = b[(size_t)(p + 2) * bstride], b3 = b[(size_t)(p + 3) * bstride];
    const double* a0 = A + (size_t)p * lda;
    const double* a1 = a0 + lda;
    const double* a2 = a1 + lda;
    const double* a3 = a2 + lda;
    for (int i = i0; i < ie; ++i) {
      double t = c[i];
      t -= a0[i] * b0;
      t -= a1[i] * b1;
      t -= a2[i] * b2;
      t -= a3[i] * b3;
      c[i] = t;
    }
  }
  for (; p < pe; ++p) {
    const double bp = b[(size_t)p * bstride];
    const double* a = A + (size_t)p * lda;
    for (int i = i0; i < ie; ++i) c[i] -= a[i] * bp;
  }
}

/* C[m x n] -= A[m x k] * B[k x n] */
static void gemm_sub(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* Cm, int ldc,
                     int par) {
  for (int p0 = 0; p0 < k; p0 += KB_) {
    const int pe = imin(p0 + KB_, k);
#pragma omp parallel for schedule(static) if (par)
    for (int j = 0; j < n; ++j)
      for (int i0 = 0; i0 < m; i0 += IB_)
        axpy_block(Cm + (size_t)j * ldc, i0, imin(i0 + IB_, m), A, lda, p0, pe, B + (size_t)j * ldb, 1);
  }
}

/* C[m x n] -= A[m x k] * A2[n x k]^T, lower part only when lower */
static void gemm_nt_sub(int m, int n, int k, const double* A, int lda, const double* A2, int ld2, double* Cm, int ldc,
                        int lower, int par) {
  for (int p0 = 0; p0 < k; p0 += KB_) {
    const int pe = imin(p0 + KB_, k);
#pragma omp parallel for schedule(dynamic, 8) if (par)
    for (int j = 0; j < n; ++j) {
      const int ib = lower ? j : 0;
      for (int i0 = ib; i0 < m; i0 += IB_)
        axpy_block(Cm + (size_t)j * ldc, i0, imin(i0 + IB_, m), A, lda, p0, pe, A2 + j, ld2);
    }
  }
}

/* B = L^-1 B, L unit lower k x k: column j runs p = 0 .. k-1,
 * b[i] -= L[i, p] b[p] for i > p; blocked over p (diagonal block, then the
 * rows below it), which keeps that order for every entry. */
static void trsm_llnu(int k, int n, const double* L, int ldl, double* B, int ldb, int par) {
  for (int p0 = 0; p0 < k; p0 += KB_) {
    const int pe = imin(p0 + KB_, k);
#pragma omp parallel for schedule(static) if (par)
    for (int j = 0; j < n; ++j) {
      double* b = B + (size_t)j * ldb;
      for (int p = p0; p < pe; ++p) {
        const double bp = b[p];
        for (int i = p + 1; i < pe; ++i) b[i] -= AT(L, ldl, i, p) * bp;
      }
      for (int i0 = pe; i0 < k; i0 += IB_) axpy_block(b, i0, imin(i0 + IB_, k), L, ldl, p0, pe, b, 1);
    }
  }
}

/* B = B * U^-1, U upper non-unit k x k, B m x k: row by row, column j gets
 * b[i, j] -= b[i, p] U[p, j] for p < j ascending, then / U[j, j]. */
static void trsm_runn(int m, int k, const double* U, int ldu, double* B, int ldb, int par) {
#pragma omp parallel for schedule(static) if (par)
  for (int i0 = 0; i0 < m; i0 += 64) {
    const int ie = imin(i0 + 64, m);
    for (int j = 0; j < k; ++j) {
      double* bj = B + (size_t)j * ldb;
      axpy_block(bj, i0, ie, B, ldb, 0, j, U + (size_t)j * ldu, 1);
      const double d = AT(U, ldu, j, j);
      for (int i = i0; i < ie; ++i) bj[i] /= d;
    }
  }
}

/* B = B * L^-T, L lower non-unit k x k, B m x k (as trsm_runn with
 * U[p, j] = L[j, p]) */
static void trsm_rltn(int m, int k, const double* L, int ldl, double* B, int ldb, int par) {
#pragma omp parallel for schedule(static) if (par)
  for (int i0 = 0; i0 < m; i0 += 64) {
    const int ie = imin(i0 + 64, m);
    for (int j = 0; j < k; ++j) {
      double* bj = B + (size_t)j * ldb;
      axpy_block(bj, i0, ie, B, ldb, 0, j, L + j, ldl);
      const double d = AT(L, ldl, j, j);
      for (int i = i0; i < ie; ++i) bj[i] /= d;
    }
  }
}

/* swap rows r1, r2 of columns [c0, c1) */
static void swap_rows(double* A, int ld, int r1, int r2, int c0, int c1) {
  if (r1 == r2) return;
  for (int j = c0; j < c1; ++j) {
    double t = AT(A, ld, r1, j);
    AT(A, ld, r1, j) = AT(A, ld, r2, j);
    AT(A, ld, r2, j) = t;
  }
}

/* Unblocked GEPP on the panel rows [k0, m) x cols [k0, k0+kb) of A (leading
 * dimension ld), swaps applied inside the panel columns only; relative
 * pivots p[k] = pivot row - k; marg[k] (if not null) = relative margin of
 * the chosen pivot over the next largest candidate.  Returns 1 + first
 * zero-pivot row or 0. */
static int getf2_panel(double* A, int ld, int k0, int kb, int m, int* p, double* marg, int par) {
  int info = 0;
  for (int j = k0; j < k0 + kb; ++j) {
    int pr = j;
    double best = -1.0, second = 0.0;
    for (int i = j; i < m; ++i) {
      double a = fabs(AT(A, ld, i, j));
      if (a > best) {
        if (best > second) second = best;
        best = a;
        pr = i;
      } else if (a > second) {
        second = a;
      }
    }
    p[j] = pr - j;
    if (marg) marg[j] = best > 0.0 ? (best - second) / best : 0.0;
    swap_rows(A, ld, j, pr, k0, k0 + kb);
    /* multipliers as LAPACK dgetf2 forms them: times the reciprocal of a
     * normal pivot, by division for a subnormal one */
    const double d = AT(A, ld, j, j);
    if (fabs(d) >= DBL_MIN) {
      const double r = 1.0 / d;
      for (int i = j + 1; i < m; ++i) AT(A, ld, i, j) *= r;
    } else if (d != 0.0) {
      for (int i = j + 1; i < m; ++i) AT(A, ld, i, j) /= d;
    } else if (!info) {
      info = j + 1;
    }
#pragma omp parallel for schedule(static) if (par && (m - j) * (k0 + kb - j) > 32768)
    for (int c = j + 1; c < k0 + kb; ++c) {
      const double u = AT(A, ld, j, c);
      if (u == 0.0) continue;
      for (int i = j + 1; i < m; ++i) AT(A, ld, i, c) -= AT(A, ld, i, j) * u;
    }
  }
  return info;
}

/* lu_piv_hier: blocked LU of ATL with tile b (pivots inside ATL), then
 * ATR <- L^-1 P ATR, ABL <- ABL U^-1, ABR -= ABL ATR. */
static int front_lu(double* F, int n, int fs, int b, int* p, double* marg, int par) {
  int info = 0;
  for (int k0 = 0; k0 < fs; k0 += b) {
    int kb = fs - k0 < b ? fs - k0 : b;
    int r = getf2_panel(F, n, k0, kb, fs, p, marg, par);
    if (r && !info) info = r;
    for (int j = k0; j < k0 + kb; ++j) {
      int pr = j + p[j];
      swap_rows(F, n, j, pr, 0, k0);        /* apply_pivots on [A10; A20] */
      swap_rows(F, n, j, pr, k0 + kb, fs);  /* apply_pivots on [A12; A22] */
    }
    trsm_llnu(kb, fs - k0 - kb, &AT(F, n, k0, k0), n, &AT(F, n, k0, k0 + kb), n, par);
    gemm_sub(fs - k0 - kb, fs - k0 - kb, kb, &AT(F, n, k0 + kb, k0), n, &AT(F, n, k0, k0 + kb), n,
             &AT(F, n, k0 + kb, k0 + kb), n, par);
  }
  const int ss = n - fs;
  if (fs && ss) {
    for (int j = 0; j < fs; ++j) swap_rows(F, n, j, j + p[j], fs, n); /* apply_pivots(p, ATR) */
    trsm_llnu(fs, ss, F, n, &AT(F, n, 0, fs), n, par);
    trsm_runn(ss, fs, F, n, &AT(F, n, fs, 0), n, par);
    gemm_sub(ss, ss, fs, &AT(F, n, fs, 0), n, &AT(F, n, 0, fs), n, &AT(F, n, fs, fs), n, par);
  }
  return info;
}

/* chol_hier: blocked Cholesky of ATL (lower) with tile b, then
 * ABL <- ABL L^-T, ABR -= ABL ABL^T (lower). */
static int front_chol(double* F, int n, int fs, int b, int par) {
  int info = 0;
  for (int k0 = 0; k0 < fs; k0 += b) {
    int kb = fs - k0 < b ? fs - k0 : b;
    for (int j = k0; j < k0 + kb; ++j) { /* FLA_Chol on A11 */
      double d = AT(F, n, j, j);
      if (!(d > 0.0) && !info) info = j + 1;
      d = sqrt(d);
      AT(F, n, j, j) = d;
      for (int i = j + 1; i < k0 + kb; ++i) AT(F, n, i, j) /= d;
      for (int c = j + 1; c < k0 + kb; ++c) {
        const double l = AT(F, n, c, j);
        for (int i = c; i < k0 + kb; ++i) AT(F, n, i, c) -= AT(F, n, i, j) * l;
      }
    }
    int m2 = fs - k0 - kb;
    trsm_rltn(m2, kb, &AT(F, n, k0, k0), n, &AT(F, n, k0 + kb, k0), n, par);
    gemm_nt_sub(m2, m2, kb, &AT(F, n, k0 + kb, k0), n, &AT(F, n, k0 + kb, k0), n, &AT(F, n, k0 + kb, k0 + kb), n, 1,
                par);
  }
  const int ss = n - fs;
  if (fs && ss) {
    trsm_rltn(ss, fs, F, n, &AT(F, n, fs, 0), n, par);
    gemm_nt_sub(ss, ss, fs, &AT(F, n, fs, 0), n, &AT(F, n, fs, 0), n, &AT(F, n, fs, fs), n, 1, par);
  }
  return info;
}

/* op_merge(A): parent (zeroed, or its own copy-in) += children's ABR in
 * children order, through the mapper runs (helper.hxx:191-222). */
static void merge_A(orc* o, int e, int par) {
  const int n = o->fs[e] + o->ss[e];
  if (o->K[e]) memcpy(o->F[e], o->K[e], (size_t)n * n * sizeof(double));
  else memset(o->F[e], 0, (size_t)n * n * sizeof(double));
  for (int k = o->cptr[e]; k < o->cptr[e + 1]; ++k) {
    int c = o->cidx[k];
    int nc = o->fs[c] + o->ss[c], fc = o->fs[c];
    int* runs = (int*)malloc((4 * o->nsch[c] + 4) * sizeof(int));
    int nr = orc_mapper(o, c, runs);
    /* column runs are independent: each parent entry still receives the
     * children in children order */
#pragma omp parallel for schedule(dynamic, 1) if (par)
    for (int jr = 0; jr < nr; ++jr)
      for (int ir = 0; ir < nr; ++ir) {
        int ioc = runs[4 * ir], jo = runs[4 * jr];
        int iop = runs[4 * ir + 1] + (runs[4 * ir + 2] ? o->fs[e] : 0);
        int jop = runs[4 * jr + 1] + (runs[4 * jr + 2] ? o->fs[e] : 0);
        for (int jj = 0; jj < runs[4 * jr + 3]; ++jj)
          for (int ii = 0; ii < runs[4 * ir + 3]; ++ii)
            AT(o->F[e], n, iop + ii, jop + jj) += AT(o->F[c], nc, fc + ioc + ii, fc + jo + jj);
      }
    free(runs);
  }
}

/* Merged input of front e (0 + own copy-in + children's ABR, merge order)
 * into out, without factoring; its children must have been factored. */
void orc_assemble(void* h, int e, double* out) {
  orc* o = (orc*)h;
  ensure(o, e);
  merge_A(o, e, 1);
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  memcpy(out, o->F[e], n * n * sizeof(double));
}

/* Levels [L0, L1] (leaves are level 0); every child of those fronts must
 * have been factored.  A level with at least as many fronts as threads
 * runs its fronts in parallel; a smaller one runs them one after another,
 * each with OpenMP inside (the results are the same bits either way). */
int orc_factor_levels(void* h, int method, int b, int L0, int L1) {
  orc* o = (orc*)h;
  o->info = 0;
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  for (int L = L0; L <= L1 && L < o->n_levels; ++L) {
    const int lo = o->lvl_ptr[L], hi = o->lvl_ptr[L + 1];
    for (int t = lo; t < hi; ++t) ensure(o, o->lvl_elem[t]);
    const int outer = (hi - lo) >= nth;
    int first = hi;
#pragma omp parallel for schedule(dynamic, 1) reduction(min : first) if (outer)
    for (int t = lo; t < hi; ++t) {
      int e = o->lvl_elem[t];
      merge_A(o, e, !outer);
      int n = o->fs[e] + o->ss[e];
      int r = (method == 1) ? front_chol(o->F[e], n, o->fs[e], b, !outer)
                            : front_lu(o->F[e], n, o->fs[e], b, o->piv[e], o->marg[e], !outer);
      if (r && t < first) first = t;
    }
    if (first < hi && !o->info) o->info = o->lvl_elem[first] + 1;
  }
  return o->info;
}

int orc_factor(void* h, int method, int b) { return orc_factor_levels(h, method, b, 0, ((orc*)h)->n_levels - 1); }

/* Forward (leaf -> root) then backward (root -> leaf) solves. */
void orc_solve(void* h, int method) {
  orc* o = (orc*)h;
  int nth = 1;
#ifdef _OPENMP
  nth = omp_get_max_threads();
#endif
  for (int e = 0; e < o->n_elem; ++e) ensure(o, e);
  for (int L = 0; L < o->n_levels; ++L) {
    const int outer = (o->lvl_ptr[L + 1] - o->lvl_ptr[L]) >= nth;
#pragma omp parallel for schedule(dynamic, 1) if (outer)
    for (int t = o->lvl_ptr[L]; t < o->lvl_ptr[L + 1]; ++t) {
      int e = o->lvl_elem[t];
      int n = o->fs[e] + o->ss[e], fs = o->fs[e];
      double* x = o->X[e];
      memcpy(x, o->B[e], (size_t)n * o->nrhs * sizeof(double));
      for (int k = o->cptr[e]; k < o->cptr[e + 1]; ++k) { /* merge_rhs_x */
        int c = o->cidx[k], nc = o->fs[c] + o->ss[c];
        int* runs = (int*)malloc((4 * o->nsch[c] + 4) * sizeof(int));
        int nr = orc_mapper(o, c, runs);
        for (int r = 0; r < o->nrhs; ++r)
          for (int q = 0; q < nr; ++q) {
            int op = runs[4 * q + 1] + (runs[4 * q + 2] ? fs : 0);
            for (int d = 0; d < runs[4 * q + 3]; ++d)
              x[op + d + (size_t)r * n] += o->X[c][o->fs[c] + runs[4 * q] + d + (size_t)r * nc];
          }
        free(runs);
      }
      const double* F = o->F[e];
#pragma omp parallel for schedule(static) if (!outer)
      for (int r = 0; r < o->nrhs; ++r) {
        double* xr = x + (size_t)r * n;
        if (method != 1)
          for (int k = 0; k < fs; ++k) {
            int pr = k + o->piv[e][k];
            double t2 = xr[k];
            xr[k] = xr[pr];
            xr[pr] = t2;
          }
        for (int k = 0; k < fs; ++k) {
          if (method == 1) xr[k] /= AT(F, n, k, k);
          const double xk = xr[k];
          for (int i = k + 1; i < n; ++i) xr[i] -= AT(F, n, i, k) * xk;
        }
      }
    }
  }
  for (int L = o->n_levels - 1; L >= 0; --L) {
    const int outer = (o->lvl_ptr[L + 1] - o->lvl_ptr[L]) >= nth;
#pragma omp parallel for schedule(dynamic, 1) if (outer)
    for (int t = o->lvl_ptr[L]; t < o->lvl_ptr[L + 1]; ++t) {
      int e = o->lvl_elem[t];
      int n = o->fs[e] + o->ss[e], fs = o->fs[e];
      double* x = o->X[e];
      int p = o->parent[e];
      if (p >= 0) { /* branch_rhs_x */
        int np = o->fs[p] + o->ss[p];
        int* runs = (int*)malloc((4 * o->nsch[e] + 4) * sizeof(int));
        int nr = orc_mapper(o, e, runs);
        for (int r = 0; r < o->nrhs; ++r)
          for (int q = 0; q < nr; ++q) {
            int op = runs[4 * q + 1] + (runs[4 * q + 2] ? o->fs[p] : 0);
            for (int d = 0; d < runs[4 * q + 3]; ++d)
              x[fs + runs[4 * q] + d + (size_t)r * n] = o->X[p][op + d + (size_t)r * np];
          }
        free(runs);
      }
      const double* F = o->F[e];
#pragma omp parallel for schedule(static) if (!outer)
      for (int r = 0; r < o->nrhs; ++r) {
        double* xr = x + (size_t)r * n;
        if (method == 1) {
          for (int i = 0; i < fs; ++i) {
            double s = 0.0;
            for (int k = fs; k < n; ++k) s += AT(F, n, k, i) * xr[k];
            xr[i] -= s;
          }
          for (int k = fs - 1; k >= 0; --k) {
            double s = xr[k];
            for (int i = k + 1; i < fs; ++i) s -= AT(F, n, i, k) * xr[i];
            xr[k] = s / AT(F, n, k, k);
          }
        } else {
          for (int k = fs; k < n; ++k) {
            const double xk = xr[k];
            for (int i = 0; i < fs; ++i) xr[i] -= AT(F, n, i, k) * xk;
          }
          for (int k = fs - 1; k >= 0; --k) {
            xr[k] /= AT(F, n, k, k);
            const double xk = xr[k];
            for (int i = 0; i < k; ++i) xr[i] -= AT(F, n, i, k) * xk;
          }
        }
      }
    }
  }
}

void orc_get_front(void* h, int e, double* out) {
  orc* o = (orc*)h;
  ensure(o, e);
  size_t n = (size_t)(o->fs[e] + o->ss[e]);
  memcpy(out, o->F[e], n * n * sizeof(double));
}
void orc_get_pivots(void* h, int e, int* out) {
  orc* o = (orc*)h;
  ensure(o, e);
  memcpy(out, o->piv[e], o->fs[e] * sizeof(int));
}
/* relative pivot margins of the last LU factorization of front e */
void orc_get_margins(void* h, int e, double* out) {
  orc* o = (orc*)h;
  ensure(o, e);
  memcpy(out, o->marg[e], o->fs[e] * sizeof(double));
}
/* solution in the global numbering: x[offset(node) + d] */
void orc_get_x_global(void* h, double* out, int ndof) {
  orc* o = (orc*)h;
  for (int e = 0; e < o->n_elem; ++e) ensure(o, e);
  for (int e = 0; e < o->n_elem; ++e) {
    int n = o->fs[e] + o->ss[e];
    for (int q = 0; q < o->nfac[e]; ++q) {
      int nd = o->fac[e][q];
      for (int d = 0; d < o->ndof[nd]; ++d)
        for (int r = 0; r < o->nrhs; ++r)
          out[o->offset[nd] + d + (size_t)r * ndof] = o->X[e][o->facoff[e][q] + d + (size_t)r * n];
    }
  }
}

/* LAWN-41 counts (LINAL/src/util.cxx:60-101) summed like Mesh_::estimate_cost. */
double orc_flops(void* h, int method) {
  orc* o = (orc*)h;
  double s = 0.0;
  for (int e = 0; e < o->n_elem; ++e) {
    double f = o->fs[e], q = o->ss[e];
    if (method == 1) {
      s += ((1.0 / 6.0) * f * f * f + 0.5 * f * f + (1.0 / 3.0) * f) + ((1.0 / 6.0) * f * f * f - (1.0 / 6.0) * f);
      s += 0.5 * q * f * (f + 1.0) + 0.5 * q * f * (f - 1.0);
      s += f * q * (q + 1.0);
    } else {
      s += (0.5 * f * f * f - (1.0 / 6.0) * f * f * f + 0.5 * f * f - 0.5 * f * f + (2.0 / 3.0) * f) +
           (0.5 * f * f * f - (1.0 / 6.0) * f * f * f - 0.5 * f * f + (1.0 / 6.0) * f);
      s += 0.5 * q * f * (f + 1.0) + 0.5 * q * f * (f - 1.0);
      s += 0.5 * q * f * (f + 1.0) + 0.5 * q * f * (f - 1.0);
      s += 2.0 * q * q * f;
    }
  }
  return s;
}

/* algorithmic flops of the fronts in levels [L0, L1] */
double orc_level_flops(void* h, int method, int L0, int L1) {
  orc* o = (orc*)h;
  double s = 0.0;
  for (int L = L0; L <= L1 && L < o->n_levels; ++L)
    for (int t = o->lvl_ptr[L]; t < o->lvl_ptr[L + 1]; ++t) {
      const int e = o->lvl_elem[t];
      const double f = o->fs[e], q = o->ss[e];
      if (method == 1)
        s += (f * f * f / 3.0 + f * f / 2.0 + f / 6.0) + q * f * f + f * q * (q + 1.0);
      else
        s += (2.0 * f * f * f / 3.0 - f * f / 2.0 + 5.0 * f / 6.0) + 2.0 * q * f * f + 2.0 * q * q * f;
    }
  return s;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Dense single-front entry points for the golden-vector tests. */
int orc_dense_lu(double* F, int n, int fs, int b, int* p) { return front_lu(F, n, fs, b, p, NULL, 1); }
/* One unblocked panel (getf2_panel with k0 = 0) of an m x kb column-major
 * block with leading dimension ld: relative pivots and their margins. */
int orc_panel_lu(double* A, int ld, int m, int kb, int* p, double* marg) {
  return getf2_panel(A, ld, 0, kb, m, p, marg, 1);
}
int orc_dense_chol(double* F, int n, int fs, int b) { return front_chol(F, n, fs, b, 1); }
