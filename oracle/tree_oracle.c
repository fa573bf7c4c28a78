/*
 * CPU oracle for the multigrid Poisson-Helmholtz path on ADAPTIVE
 * quadtrees/octrees.  TEST INFRASTRUCTURE ONLY: the checker that the GPU
 * library's tree path is compared against.  Only tests/, the smoke test and
 * the CPU-baseline leg of the benchmark may load it; the product library never
 * links or calls it.
 *
 * It restates, in plain C99 with the reference's floating-point evaluation
 * order, the reference algorithm on a `tree` grid (DIM = 2 or 3):
 *
 *   cells/flags    grid/tree.h:24-52 (Cell: flags active/leaf, neighbors, pid)
 *   storage        one dense array per level of (2^l+4)^dim cells, GHOSTS=2;
 *                  only cells the reference allocates are marked allocated
 *                  (grid/tree.h:929-972 alloc_children, 1488-1584 init_grid)
 *   refinement     grid/tree-common.h:21-92 (refine_cell with the TWO_ONE
 *                  parent-neighbour refinement), grid/tree.h:1032-1041
 *                  (increment_neighbors: the 3^dim ring gets children)
 *   caches         grid/tree.h:556-684 (update_cache_f with FBOUNDARY):
 *                  pre-order traversal of grid/foreach_cell.h:66-135,
 *                  active[l], prolongation[l], boundary[l], leaves, faces
 *   traversals     grid/tree.h:721-740 (foreach_level, foreach_coarse_level,
 *                  foreach_level_or_leaf: level l, then leaves of l-1 .. 0)
 *   box boundary   grid/tree.h:1139-1363 (normal_neighbor, diagonal_neighbor
 *                  _2D/_3D, tangential_neighbor_x for face vectors, BGHOSTS=1)
 *   tree boundary  grid/tree-common.h:587-716 (tree_boundary_level: restrict
 *                  below l, then per level halo prolongation + box BC)
 *   halo_face      grid/tree-common.h:373-409
 *   refine_face_x  grid/tree-common.h:464-511 (alpha halos)
 *   restriction    grid/multigrid-common.h:26-63, tree-common.h:795-799
 *   prolongation   grid/multigrid-common.h:124-146 (bilinear, refine_bilinear)
 *   relax          poisson.h:271-338 (GS in place over foreach_level_or_leaf,
 *                  or weighted Jacobi 2/3)
 *   residual       poisson.h:355-372 (the TREE form: face fluxes g, halo_face,
 *                  then the leaf divergence)
 *   mg_cycle       poisson.h:32-90, mg_solve 139-235, poisson 402-441
 *   automatic BC   grid/stencils.h:228-470: the residual's face loop reads `a`
 *                  with a one-cell stencil, so `a` gets the full tree boundary
 *                  (tree_boundary_level(list, -1)) before every residual; `g`
 *                  gets halo_face; `res` gets the full tree boundary at the
 *                  restriction(res) of mg_cycle (tree_restriction).
 *
 *   adaptation     grid/tree-common.h:98-131, 165-303 (coarsen_cell,
 *                  adapt_wavelet's refinement and coarsening passes),
 *                  grid/tree.h:990-1059 (free_children, decrement_neighbors)
 *
 * A tree whose leaves are all on one level ("full") takes the reference's
 * tree_is_full() shortcuts: box BCs only at the given level, and
 * multigrid_restriction for restriction().
 *
 * Red-black Gauss-Seidel (an extension, as in the uniform-grid oracle): the
 * cells of foreach_level_or_leaf(l) are coloured by the parity of i+j+k at
 * their own level; red first, then the box BC of levels 0..l (ghost cells
 * only, halos are left as they are), then black.
 *
 * Boundary conditions are the reference's dirichlet/neumann/symmetry forms
 * (common.h:957-960) with the value evaluated at the centre of the boundary
 * face between the interior cell and the ghost; the values are given per
 * level and side on every boundary face of that level.
 *
 * Planted mistakes: built with -DORC_PLANT=k (k > 0), the oracle makes one
 * plausible mistake on purpose, so that the tests can show that their checks
 * catch it (oracle/tree.py builds these variants into a temporary directory;
 * the normal build has ORC_PLANT 0):
 *  11  adapt_wavelet clears the children's flags before each estimated
 *      scalar (only the last scalar decides)
 *  12  adapt_wavelet estimates only the first scalar of the list
 *
 * Build with -ffp-contract=off (the reference is compiled without FMA).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifndef DIM
#error "compile with -DDIM=2 or 3"
#endif
#if DIM < 2 || DIM > 3
#error "the tree oracle supports DIM=2 and DIM=3"
#endif

#define GH 2
#define NCH (1 << DIM)

/* cell flags; the last four are adapt_wavelet's user flags
   (grid/tree-common.h:197-198, 201, 218: refined, too_fine, too_coarse,
   just_fine) */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT(k) (ORC_PLANT == (k))

enum { C_ACTIVE = 1, C_LEAF = 2, C_FB = 4, C_REFINED = 8, C_TOO_FINE = 16, C_TOO_COARSE = 32, C_JUST_FINE = 64 };
/* fields, per level, structure of arrays; T_S is a user scalar (the one
   adapt_wavelet estimates) */
enum { T_A, T_B, T_RES, T_DA, T_LAM, T_C, T_G };
#define T_AL (T_G + DIM)
#define T_S (T_AL + DIM)
/* T_U + d: the face velocity of project(); T_DB, T_RB: the adjoints of
   da and res in the transposed V-cycle */
#define T_U (T_S + 1)
#define T_DB (T_U + DIM)
#define T_RB (T_DB + 1)
/* user scalars 1..NUS-1 (user scalar 0 is T_S): T_SX + k - 1 */
#define NUS 4
#define T_SX (T_RB + 1)
#define NVT (T_SX + NUS - 1)

enum { BC_SYMMETRY = 0, BC_DIRICHLET = 1, BC_NEUMANN = 2 };
enum { SM_GS = 0, SM_JACOBI = 1, SM_RBGS = 2 };
/* which boundary condition a scalar uses */
enum { BK_A = 0, BK_HOM = 1, BK_SYM = 2 };

typedef struct {
  int n, len;
  size_t ncell;
  unsigned char *alloc, *flags;
  unsigned short *nb;
  signed char *pid;
  double *f[NVT];
  long nc; /* allocated groups of children on this level (Layer.nc) */
} otl;

typedef struct {
  int *p;
  int n, m;
} olist;

typedef struct {
  int level, idx, flags;
} oface;

typedef struct {
  int i;
  double resb, resa, sum;
  int nrelax, minlevel;
} ot_mgstats;

typedef struct {
  int depth, cap;
  otl *L;
  double L0, X0, Y0, Z0;
  /* caches */
  olist *active, *prolong, *bnd;
  olist leaves_l, leaves_i; /* leaf cache: level and index */
  oface *faces;
  int nfaces, mfaces;
  int dirty;
  long nleaves;
  /* problem */
  int bc[2 * DIM];
  double **bcv[2 * DIM]; /* [side][level] -> n^(DIM-1) values */
  int bcv_n[2 * DIM];    /* levels given; deeper levels read zero */
  int lambda_field;
  double lambda_val;
  int alpha_field;
  double alpha_val;
  int smoother;
  double TOLERANCE;
  int NITERMAX, NITERMIN;
} ot_tree;

/* ----------------------------------------------------------- indexing */

static inline size_t lidx(const otl *L, int i, int j, int k) {
#if DIM == 2
  (void)k;
  return (size_t)i * L->len + j;
#else
  return ((size_t)i * L->len + j) * L->len + k;
#endif
}

static inline ptrdiff_t lstride(const otl *L, int d) {
#if DIM == 2
  return d == 0 ? L->len : 1;
#else
  return d == 0 ? (ptrdiff_t)L->len * L->len : d == 1 ? L->len : 1;
#endif
}

static inline void lijk(const otl *L, size_t idx, int c[3]) {
#if DIM == 2
  c[0] = (int)(idx / L->len);
  c[1] = (int)(idx % L->len);
  c[2] = 0;
#else
  size_t m = L->len;
  c[2] = (int)(idx % m);
  c[1] = (int)((idx / m) % m);
  c[0] = (int)(idx / (m * m));
#endif
}

static inline double delta_l(const ot_tree *t, int l) { return t->L0 * (1. / (1 << l)); }

static inline int is_boundary(const otl *L, size_t c) { return L->pid[c] < 0; }
static inline int is_leaf(const otl *L, size_t c) { return L->flags[c] & C_LEAF; }
static inline int is_refined(const otl *L, size_t c) {
  return !(L->flags[c] & C_LEAF) && L->nb[c] && L->pid[c] >= 0;
}
static inline int is_prolongation(const otl *L, size_t c) {
  return !(L->flags[c] & C_LEAF) && !L->nb[c] && L->pid[c] >= 0;
}
/* allocated and not a boundary cell (tree.h:1136-1137 is_neighbor) */
static inline int is_nbr(const otl *L, size_t c) { return L->alloc[c] && L->pid[c] >= 0; }

static void list_add(olist *l, int v) {
  if (l->n == l->m) {
    l->m = l->m ? 2 * l->m : 64;
    l->p = realloc(l->p, sizeof(int) * l->m);
  }
  l->p[l->n++] = v;
}

/* ---------------------------------------------------------- levels */

static void level_init(otl *L, int l) {
  L->n = 1 << l;
  L->len = L->n + 2 * GH;
  L->ncell = 1;
  for (int d = 0; d < DIM; d++) L->ncell *= L->len;
  L->alloc = calloc(L->ncell, 1);
  L->flags = calloc(L->ncell, 1);
  L->nb = calloc(L->ncell, sizeof(unsigned short));
  L->pid = calloc(L->ncell, 1);
  for (int f = 0; f < NVT; f++) L->f[f] = calloc(L->ncell, sizeof(double));
  L->nc = 0;
}

static void level_free(otl *L) {
  free(L->alloc); free(L->flags); free(L->nb); free(L->pid);
  for (int f = 0; f < NVT; f++) free(L->f[f]);
}

static void add_level(ot_tree *t) {
  int l = t->depth + 1;
  if (l >= t->cap) {
    int cap = t->cap * 2;
    t->L = realloc(t->L, sizeof(otl) * cap);
    t->active = realloc(t->active, sizeof(olist) * cap);
    t->prolong = realloc(t->prolong, sizeof(olist) * cap);
    t->bnd = realloc(t->bnd, sizeof(olist) * cap);
    for (int i = t->cap; i < cap; i++) {
      memset(&t->active[i], 0, sizeof(olist));
      memset(&t->prolong[i], 0, sizeof(olist));
      memset(&t->bnd[i], 0, sizeof(olist));
    }
    t->cap = cap;
  }
  level_init(&t->L[l], l);
  t->depth = l;
}

/* ---------------------------------------------------------- setup */

/* init_grid(1), grid/tree.h:1488-1584: the root and its two ghost rings,
   pids right/left before top/bottom before front/back (common.h:950-951:
   right=0, left=1, top=2, bottom=3, front=4, back=5) */
ot_tree *ot_new(double L0, double X0, double Y0, double Z0) {
  ot_tree *t = calloc(1, sizeof(ot_tree));
  t->cap = 4;
  t->L = calloc(t->cap, sizeof(otl));
  t->active = calloc(t->cap, sizeof(olist));
  t->prolong = calloc(t->cap, sizeof(olist));
  t->bnd = calloc(t->cap, sizeof(olist));
  t->L0 = L0; t->X0 = X0; t->Y0 = Y0; t->Z0 = Z0;
  t->depth = 0;
  level_init(&t->L[0], 0);
  otl *L = &t->L[0];
  for (size_t c = 0; c < L->ncell; c++) {
    int x[3];
    lijk(L, c, x);
    L->alloc[c] = 1;
    int k = x[0] - GH, m = x[1] - GH, n = DIM == 3 ? x[2] - GH : 0;
    L->pid[c] = k > 0 ? -1 - 0 : k < 0 ? -1 - 1 : m > 0 ? -1 - 2 : m < 0 ? -1 - 3 : n > 0 ? -1 - 4 : n < 0 ? -1 - 5 : 0;
  }
  size_t root = lidx(L, GH, GH, GH);
  L->flags[root] = C_LEAF | C_ACTIVE;
  L->pid[root] = 0;
  t->dirty = 1;
  for (int b = 0; b < 2 * DIM; b++) t->bc[b] = BC_SYMMETRY;
  t->lambda_val = 0.;
  t->alpha_val = 1.;
  t->smoother = SM_GS;
  t->TOLERANCE = 1e-3;
  t->NITERMAX = 100;
  t->NITERMIN = 1;
  return t;
}

static void free_bcv(ot_tree *t, int b) {
  if (!t->bcv[b]) return;
  for (int l = 0; l < t->bcv_n[b]; l++) free(t->bcv[b][l]);
  free(t->bcv[b]);
  t->bcv[b] = NULL;
  t->bcv_n[b] = 0;
}

void ot_free(ot_tree *t) {
  for (int l = 0; l <= t->depth; l++) level_free(&t->L[l]);
  for (int l = 0; l < t->cap; l++) {
    free(t->active[l].p); free(t->prolong[l].p); free(t->bnd[l].p);
  }
  for (int b = 0; b < 2 * DIM; b++) free_bcv(t, b);
  free(t->L); free(t->active); free(t->prolong); free(t->bnd);
  free(t->leaves_l.p); free(t->leaves_i.p); free(t->faces);
  free(t);
}

int ot_dimension(void) { return DIM; }
int ot_depth(ot_tree *t) { return t->depth; }
void ot_set_niter(ot_tree *t, int nmax, int nmin) { t->NITERMAX = nmax; t->NITERMIN = nmin; }
void ot_set_tolerance(ot_tree *t, double tol) { t->TOLERANCE = tol; }
void ot_set_smoother(ot_tree *t, int s) { t->smoother = s; }
void ot_set_lambda_const(ot_tree *t, double v) { t->lambda_field = 0; t->lambda_val = v; }
void ot_set_alpha_const(ot_tree *t, double v) { t->alpha_field = 0; t->alpha_val = v; }

/* ---------------------------------------------------------- refinement */

/* alloc_children, grid/tree.h:929-972 */
static void alloc_children(ot_tree *t, int l, const int c[3]) {
  if (l == t->depth) add_level(t);
  otl *F = &t->L[l + 1];
  int fi = 2 * c[0] - GH, fj = 2 * c[1] - GH, fk = 2 * c[2] - GH;
  if (F->alloc[lidx(F, fi, fj, DIM == 3 ? fk : 0)]) return;
  signed char pid = t->L[l].pid[lidx(&t->L[l], c[0], c[1], c[2])];
  F->nc++;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
        size_t q = lidx(F, fi + a, fj + b, DIM == 3 ? fk + e : 0);
        F->alloc[q] = 1;
        F->flags[q] = 0;
        F->nb[q] = 0;
        F->pid[q] = pid;
        /* mempool_alloc0: a new cell's fields start at zero */
        for (int f = 0; f < NVT; f++) F->f[f][q] = 0.;
      }
}

/* increment_neighbors, grid/tree.h:1032-1041: every cell of the 3^dim ring
   (the cell itself included) counts one more refined neighbour and gets
   children when its count leaves zero */
static void increment_neighbors(ot_tree *t, int l, const int c[3]) {
  otl *L = &t->L[l];
  size_t q0 = lidx(L, c[0], c[1], c[2]);
  t->dirty = 1;
  if (L->nb[q0]++ == 0) alloc_children(t, l, c);
  for (int a = -1; a <= 1; a++)
    for (int b = -1; b <= 1; b++)
      for (int e = (DIM == 3 ? -1 : 0); e <= (DIM == 3 ? 1 : 0); e++) {
        int n[3] = {c[0] + a, c[1] + b, c[2] + e};
        L = &t->L[l];
        size_t q = lidx(L, n[0], n[1], n[2]);
        if (L->nb[q]++ == 0) alloc_children(t, l, n);
      }
  L = &t->L[l];
  L->nb[q0]--;
}

static void halo_bilinear(ot_tree *t, int f, int l, size_t q);

/* refine_cell, grid/tree-common.h:21-67 (no MPI, no refined cache): the
   coarser neighbours the 2:1 rule needs are refined first and get `flag`;
   the fields of `list` are prolongated into the new children with
   refine_bilinear (multigrid-common.h:142-146) */
static int refine_cell_f(ot_tree *t, int l, const int c[3], int flag, const int *list, int nlist) {
  int nr = 0;
  if (l > 0) {
    int ch[3], pc[3];
    for (int d = 0; d < 3; d++) {
      ch[d] = 2 * ((c[d] + GH) % 2) - 1;
      pc[d] = (c[d] + GH) / 2;
    }
    if (DIM == 2) { ch[2] = 1; pc[2] = 0; }
    for (int a = 0; a != 2 * ch[0]; a += ch[0])
      for (int b = 0; b != 2 * ch[1]; b += ch[1])
        for (int e = 0; e != (DIM == 3 ? 2 * ch[2] : 1); e += (DIM == 3 ? ch[2] : 1)) {
          int p[3] = {pc[0] + a, pc[1] + b, DIM == 3 ? pc[2] + e : 0};
          otl *P = &t->L[l - 1];
          size_t q = lidx(P, p[0], p[1], p[2]);
          if (P->pid[q] >= 0 && (P->flags[q] & C_LEAF)) {
            nr += refine_cell_f(t, l - 1, p, flag, list, nlist);
            t->L[l - 1].flags[q] |= flag;
          }
        }
  }
  increment_neighbors(t, l, c);
  otl *L = &t->L[l];
  size_t q = lidx(L, c[0], c[1], c[2]);
  unsigned char cflag = (L->flags[q] & C_ACTIVE) ? (C_ACTIVE | C_LEAF) : C_LEAF;
  otl *F = &t->L[l + 1];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < (DIM == 3 ? 2 : 1); e++)
        F->flags[lidx(F, 2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0)] |= cflag;
  for (int m = 0; m < nlist; m++) halo_bilinear(t, list[m], l, q);
  L->flags[q] &= ~C_LEAF;
  return nr + 1;
}

static int refine_cell(ot_tree *t, int l, const int c[3]) { return refine_cell_f(t, l, c, 0, NULL, 0); }

/* free_children, grid/tree.h:990-1004: the layer goes when its last group
   of children does (update_depth (-1)) */
static void free_children(ot_tree *t, int l, const int c[3]) {
  otl *F = &t->L[l + 1];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
        size_t q = lidx(F, 2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0);
        F->alloc[q] = 0;
        F->flags[q] = 0;
        F->nb[q] = 0;
      }
  if (--F->nc == 0) {
    if (l + 1 != t->depth) {
      fprintf(stderr, "tree oracle: layer %d emptied below the depth %d\n", l + 1, t->depth);
      abort();
    }
    level_free(F);
    memset(F, 0, sizeof(otl));
    t->depth--;
    t->dirty = 1;
  }
}

/* decrement_neighbors, grid/tree.h:1043-1059 */
static void decrement_neighbors(ot_tree *t, int l, const int c[3]) {
  t->dirty = 1;
  for (int a = -1; a <= 1; a++)
    for (int b = -1; b <= 1; b++)
      for (int e = (DIM == 3 ? -1 : 0); e <= (DIM == 3 ? 1 : 0); e++) {
        int n[3] = {c[0] + a, c[1] + b, c[2] + e};
        otl *L = &t->L[l];
        size_t q = lidx(L, n[0], n[1], n[2]);
        if (L->alloc[q]) {
          L->nb[q]--;
          if (L->nb[q] == 0) free_children(t, l, n);
        }
      }
  otl *L = &t->L[l];
  size_t q0 = lidx(L, c[0], c[1], c[2]);
  if (L->nb[q0]) {
    otl *F = &t->L[l + 1];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
          size_t q = lidx(F, 2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0);
          F->flags[q] = 0;
          F->pid[q] = L->pid[q0];
        }
  }
}

static double child_avg(ot_tree *t, int f, int l, size_t q);

/* coarsen_cell, grid/tree-common.h:98-131 (no MPI): refused when a child
   has a refined neighbour; the fields of `list` take the children's average
   (restriction_average) */
static int coarsen_cell(ot_tree *t, int l, const int c[3], const int *list, int nlist) {
  otl *L = &t->L[l], *F = &t->L[l + 1];
  size_t q = lidx(L, c[0], c[1], c[2]);
  int pid = L->pid[q];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
        size_t p = lidx(F, 2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0);
        if (F->nb[p] || (F->pid[p] < 0 && F->pid[p] != pid)) return 0;
      }
  for (int m = 0; m < nlist; m++) L->f[list[m]][q] = child_avg(t, list[m], l, q);
  L->flags[q] |= C_LEAF;
  decrement_neighbors(t, l, c);
  return 1;
}

/* ---------------------------------------------------------- caches */

static void visit(ot_tree *t, int l, int i, int j, int k) {
  otl *L = &t->L[l];
  size_t q = lidx(L, i, j, k);
  if (!L->alloc[q]) return;
  int c[3] = {i, j, k};
  if (L->pid[q] == 0 && (L->flags[q] & C_ACTIVE)) list_add(&t->active[l], (int)q);
  if (!is_boundary(L, q)) {
    /* boundary cells in the 3^dim ring, once per level (FBOUNDARY) */
    for (int a = -1; a <= 1; a++)
      for (int b = -1; b <= 1; b++)
        for (int e = (DIM == 3 ? -1 : 0); e <= (DIM == 3 ? 1 : 0); e++) {
          size_t n = lidx(L, i + a, j + b, k + e);
          if (L->alloc[n] && is_boundary(L, n) && !(L->flags[n] & C_FB)) {
            list_add(&t->bnd[l], (int)n);
            L->flags[n] |= C_FB;
          }
        }
  }
  if (is_leaf(L, q)) {
    if (L->pid[q] == 0) {
      list_add(&t->leaves_l, l);
      list_add(&t->leaves_i, (int)q);
      int flags = 0;
      for (int d = 0; d < DIM; d++) {
        size_t m = q - lstride(L, d);
        if (is_boundary(L, m) || is_prolongation(L, m) || is_leaf(L, m)) flags |= 1 << d;
      }
      if (flags) {
        if (t->nfaces == t->mfaces) {
          t->mfaces = t->mfaces ? 2 * t->mfaces : 256;
          t->faces = realloc(t->faces, sizeof(oface) * t->mfaces);
        }
        t->faces[t->nfaces++] = (oface){l, (int)q, flags};
      }
      for (int d = 0; d < DIM; d++) {
        size_t p = q + lstride(L, d);
        if (is_boundary(L, p) || is_prolongation(L, p) || (L->pid[p] != 0 && is_leaf(L, p))) {
          if (t->nfaces == t->mfaces) {
            t->mfaces = t->mfaces ? 2 * t->mfaces : 256;
            t->faces = realloc(t->faces, sizeof(oface) * t->mfaces);
          }
          t->faces[t->nfaces++] = (oface){l, (int)p, 1 << d};
        }
      }
      if (L->nb[q] > 0) list_add(&t->prolong[l], (int)q);
    }
    return;
  }
  (void)c;
  if (l < t->depth)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++)
          visit(t, l + 1, 2 * i - GH + a, 2 * j - GH + b, DIM == 3 ? 2 * k - GH + e : 0);
}

/* update_cache_f, grid/tree.h:556-684 */
void ot_update(ot_tree *t) {
  if (!t->dirty) return;
  for (int l = 0; l <= t->depth; l++) t->active[l].n = t->prolong[l].n = t->bnd[l].n = 0;
  t->leaves_l.n = t->leaves_i.n = 0;
  t->nfaces = 0;
  visit(t, 0, GH, GH, DIM == 3 ? GH : 0);
  for (int l = 0; l <= t->depth; l++)
    for (int n = 0; n < t->bnd[l].n; n++) t->L[l].flags[t->bnd[l].p[n]] &= ~C_FB;
  t->nleaves = t->leaves_l.n;
  t->dirty = 0;
}

/* tree_is_full, grid/tree.h:1633-1638 */
static int is_full(ot_tree *t) {
  ot_update(t);
  long full = 1L << (t->depth * DIM);
  return t->nleaves == full;
}

long ot_nleaves(ot_tree *t) { ot_update(t); return t->nleaves; }

/* the leaves in cache order: level and cell centre */
void ot_leaves(ot_tree *t, int *level, double *xyz) {
  ot_update(t);
  for (long n = 0; n < t->nleaves; n++) {
    int l = t->leaves_l.p[n], c[3];
    otl *L = &t->L[l];
    lijk(L, t->leaves_i.p[n], c);
    double D = delta_l(t, l);
    if (level) level[n] = l;
    if (xyz) {
      xyz[3 * n] = (c[0] - GH + 0.5) * D + t->X0;
      xyz[3 * n + 1] = (c[1] - GH + 0.5) * D + t->Y0;
      xyz[3 * n + 2] = DIM == 3 ? (c[2] - GH + 0.5) * D + t->Z0 : 0.;
    }
  }
}

/* one pass of the refine() macro (tree-common.h:308-326) over the leaves
   marked in cache order; the fixed point reached by repeated passes does not
   depend on the visiting order, since refine_cell only refines what 2:1
   balance requires */
static void mesh_fields_boundary(ot_tree *t);
static void refine_cell_fields(ot_tree *t, int l, const int c[3]);

/* one pass of refine (cond) (grid/tree-common.h:308-326): boundary (all),
   then refine_cell (point, all) on every marked leaf in foreach_leaf()
   order; the user scalars, a and b get refine_bilinear in the new cells */
int ot_refine_marked(ot_tree *t, const unsigned char *mark) {
  ot_update(t);
  mesh_fields_boundary(t);
  long n = t->nleaves;
  int *ll = malloc(sizeof(int) * n), *li = malloc(sizeof(int) * n);
  memcpy(ll, t->leaves_l.p, sizeof(int) * n);
  memcpy(li, t->leaves_i.p, sizeof(int) * n);
  int refined = 0;
  for (long q = 0; q < n; q++) {
    if (!mark[q]) continue;
    int l = ll[q], c[3];
    otl *L = &t->L[l];
    if (!(L->flags[li[q]] & C_LEAF)) continue; /* refined by 2:1 meanwhile */
    lijk(L, li[q], c);
    refine_cell_fields(t, l, c);
    refined++;
  }
  free(ll); free(li);
  t->dirty = 1;
  return refined;
}

/* exported caches for index parity: per level lists of (i,j,k) */
int ot_cache_count(ot_tree *t, int which, int l) {
  ot_update(t);
  if (l > t->depth) return 0;
  olist *c = which == 0 ? &t->active[l] : which == 1 ? &t->prolong[l] : &t->bnd[l];
  return c->n;
}
void ot_cache_get(ot_tree *t, int which, int l, int *ijk) {
  ot_update(t);
  olist *c = which == 0 ? &t->active[l] : which == 1 ? &t->prolong[l] : &t->bnd[l];
  for (int n = 0; n < c->n; n++) lijk(&t->L[l], c->p[n], &ijk[3 * n]);
}
int ot_face_count(ot_tree *t) { ot_update(t); return t->nfaces; }
/* faces: level, i, j, k, flags per entry; xyz of each face entry's low-face
   centres are computed by the caller from these */
void ot_face_get(ot_tree *t, int *out) {
  ot_update(t);
  for (int n = 0; n < t->nfaces; n++) {
    int c[3];
    lijk(&t->L[t->faces[n].level], t->faces[n].idx, c);
    out[5 * n] = t->faces[n].level;
    out[5 * n + 1] = c[0]; out[5 * n + 2] = c[1]; out[5 * n + 3] = c[2];
    out[5 * n + 4] = t->faces[n].flags;
  }
}
/* cell flags and neighbour counts of a level, for index parity */
void ot_level_meta(ot_tree *t, int l, unsigned char *alloc, unsigned char *flags,
                   unsigned short *nb, signed char *pid) {
  otl *L = &t->L[l];
  if (alloc) memcpy(alloc, L->alloc, L->ncell);
  if (flags) memcpy(flags, L->flags, L->ncell);
  if (nb) memcpy(nb, L->nb, L->ncell * sizeof(unsigned short));
  if (pid) memcpy(pid, L->pid, L->ncell);
}

/* ---------------------------------------------------------- fields */

/* leaf values in cache order */
void ot_set_leaf(ot_tree *t, int f, const double *v) {
  ot_update(t);
  for (long n = 0; n < t->nleaves; n++) t->L[t->leaves_l.p[n]].f[f][t->leaves_i.p[n]] = v[n];
}
void ot_get_leaf(ot_tree *t, int f, double *v) {
  ot_update(t);
  for (long n = 0; n < t->nleaves; n++) v[n] = t->L[t->leaves_l.p[n]].f[f][t->leaves_i.p[n]];
}
/* face values (alpha) of the face cache, one value per set flag bit in
   entry order then dimension order (a foreach_face() assignment) */
void ot_set_faces(ot_tree *t, const double *v) {
  ot_update(t);
  long q = 0;
  for (int n = 0; n < t->nfaces; n++)
    for (int d = 0; d < DIM; d++)
      if (t->faces[n].flags & (1 << d)) t->L[t->faces[n].level].f[T_AL + d][t->faces[n].idx] = v[q++];
  t->alpha_field = 1;
}
void ot_use_lambda_field(ot_tree *t) { t->lambda_field = 1; }
double *ot_level_field(ot_tree *t, int l, int f) { return t->L[l].f[f]; }

/* boundary values per side for every level 0..depth, each level n^(DIM-1)
   values in C order of the tangential indices */
void ot_set_bc(ot_tree *t, int b, int kind, const double *vals) {
  free_bcv(t, b);
  t->bc[b] = kind;
  if (!vals) return;
  t->bcv[b] = calloc(t->depth + 1, sizeof(double *));
  t->bcv_n[b] = t->depth + 1;
  size_t off = 0;
  for (int l = 0; l <= t->depth; l++) {
    size_t m = 1;
    for (int d = 1; d < DIM; d++) m *= (size_t)1 << l;
    t->bcv[b][l] = malloc(sizeof(double) * m);
    memcpy(t->bcv[b][l], vals + off, sizeof(double) * m);
    off += m;
  }
}

/* ---------------------------------------------------------- box boundary */

/* value of the BC of side b at the face between interior cell P (coords c,
   level l) and its ghost, for the field's BC kind */
static inline double bc_value(const ot_tree *t, int bk, int b, int l, const int c[3], double s) {
  int kind = bk == BK_SYM ? BC_SYMMETRY : t->bc[b];
  if (kind == BC_SYMMETRY) return s;
  if (bk == BK_HOM) return kind == BC_DIRICHLET ? -s : s;
  int d = b / 2, n = 1 << l;
  size_t ti = 0;
  for (int e = 0; e < DIM; e++)
    if (e != d) ti = ti * n + (size_t)(c[e] - GH);
  double v = t->bcv[b] && l < t->bcv_n[b] ? t->bcv[b][l][ti] : 0.;
  if (kind == BC_DIRICHLET) return 2. * v - s;
  return delta_l(t, l) * v + s;
}

static inline int bid_of(const otl *L, size_t q) { return -L->pid[q] - 1; }

/* box_boundary_level for one scalar field f at level l (tree.h:1286-1312
   with normal_neighbor 1139-1166 and diagonal_neighbor 1168-1258) */
static void box_scalar(ot_tree *t, int f, int bk, int l) {
  otl *L = &t->L[l];
  double *s = L->f[f];
  for (int n = 0; n < t->bnd[l].n; n++) {
    size_t q = t->bnd[l].p[n];
    int g[3];
    lijk(L, q, g);
    int done = 0;
    /* normal neighbour */
    for (int d = 0; d < DIM && !done; d++)
      for (int i = -1; i <= 1 && !done; i += 2) {
        size_t p = q + i * lstride(L, d);
        if (is_nbr(L, p)) {
          int c[3] = {g[0], g[1], g[2]};
          c[d] += i;
          s[q] = bc_value(t, bk, bid_of(L, q), l, c, s[p]);
          done = 1;
        }
      }
    /* 2D diagonal (in each plane for 3D, foreach_dimension rotation) */
    for (int pl = 0; pl < (DIM == 3 ? 3 : 1) && !done; pl++) {
      int da = pl, db = (pl + 1) % DIM;
      for (int i = -1; i <= 1 && !done; i += 2)
        for (int j = -1; j <= 1 && !done; j += 2) {
          size_t nn = q + i * lstride(L, da) + j * lstride(L, db);
          size_t n1 = q + i * lstride(L, da), n2 = q + j * lstride(L, db);
          if (L->alloc[nn] && is_nbr(L, nn) && L->alloc[n1] && is_boundary(L, n1) && L->alloc[n2] &&
              is_boundary(L, n2)) {
            int c[3] = {g[0], g[1], g[2]};
            c[da] += i;
            c[db] += j;
            double sn = s[nn];
            s[q] = (bc_value(t, bk, bid_of(L, n1), l, c, sn) + bc_value(t, bk, bid_of(L, n2), l, c, sn) - sn);
            done = 1;
          }
        }
    }
#if DIM == 3
    for (int i = -1; i <= 1 && !done; i += 2)
      for (int j = -1; j <= 1 && !done; j += 2)
        for (int k = -1; k <= 1 && !done; k += 2) {
          size_t n0 = q + i * lstride(L, 0) + j * lstride(L, 1) + k * lstride(L, 2);
          size_t n1 = q + i * lstride(L, 0) + j * lstride(L, 1);
          size_t n2 = q + i * lstride(L, 0) + k * lstride(L, 2);
          size_t n3 = q + j * lstride(L, 1) + k * lstride(L, 2);
          if (is_nbr(L, n0) && is_boundary(L, n1) && is_boundary(L, n2) && is_boundary(L, n3)) {
            int c[3] = {g[0] + i, g[1] + j, g[2] + k};
            double sn = s[n0];
            s[q] = (bc_value(t, bk, bid_of(L, n1), l, c, sn) + bc_value(t, bk, bid_of(L, n2), l, c, sn) +
                    bc_value(t, bk, bid_of(L, n3), l, c, sn) - 2. * sn);
            done = 1;
          }
        }
#endif
    if (!done) s[q] = NAN; /* `undefined`: never read by the solver */
  }
}

/* the face-vector part of box_boundary_level (tree.h:1313-1353) for alpha
   with the default face-vector BCs: the normal components have none
   (NULL), the tangential ones are symmetry (cartesian-common.h:546-576,
   the VN/VT slots), so a ghost's tangential face copies the face of the
   tangential neighbour found by tangential_neighbor_x (tree.h:1260-1280),
   or is zero when there is none */
static void box_faces(ot_tree *t, int l) {
  otl *L = &t->L[l];
  for (int n = 0; n < t->bnd[l].n; n++) {
    size_t q = t->bnd[l].p[n];
    for (int d = 0; d < DIM; d++) {
      double *v = L->f[T_AL + d];
      /* i = -1: no normal neighbour on the low side */
      size_t pm = q - lstride(L, d), pp = q + lstride(L, d);
      (void)pp;
      if (is_nbr(L, pm)) continue; /* normal: NULL BC, nothing */
      /* i = +1 normal case also sets nothing; only i = -1 reaches here */
      int t1 = (d + 1) % DIM, t2 = (d + 2) % DIM;
      long found = -1;
      for (int j = -1; j <= 1 && found < 0; j += 2) {
        size_t a = q + j * lstride(L, t1), b = a - lstride(L, d);
        if (is_nbr(L, a) || is_nbr(L, b)) { found = (long)a; break; }
#if DIM == 3
        a = q + j * lstride(L, t2); b = a - lstride(L, d);
        if (is_nbr(L, a) || is_nbr(L, b)) { found = (long)a; break; }
#else
        (void)t2;
#endif
      }
      v[q] = found >= 0 ? v[found] : 0.;
    }
  }
}

/* ---------------------------------------------------------- restriction */

static double child_avg(ot_tree *t, int f, int l, size_t q) {
  otl *L = &t->L[l], *F = &t->L[l + 1];
  int c[3];
  lijk(L, q, c);
  double sum = 0.;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < (DIM == 3 ? 2 : 1); e++)
        sum += F->f[f][lidx(F, 2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0)];
  return sum / NCH;
}

/* the tangential child offsets of face_average / halo_face per component
   (foreach_dimension rotation of fine(v.x,0,j,k)) */
static const int ftan[3][4][3] = {
#if DIM == 3
  {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}},
  {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 1}},
  {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}},
#else
  {{0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 0}},
  {{0, 0, 0}, {1, 0, 0}, {0, 0, 0}, {0, 0, 0}},
  {{0}},
#endif
};

/* mean of the fine faces of component d on the low (hi=0) or high (hi=1)
   side of cell q at level l */
static double fine_face_avg(ot_tree *t, int f, int d, int l, size_t q, int hi) {
  otl *L = &t->L[l], *F = &t->L[l + 1];
  int c[3];
  lijk(L, q, c);
  int base[3] = {2 * c[0] - GH, 2 * c[1] - GH, 2 * c[2] - GH};
  base[d] += 2 * hi;
  double s = 0.;
#if DIM == 2
  for (int m = 0; m < 2; m++) s += F->f[f][lidx(F, base[0] + ftan[d][m][0], base[1] + ftan[d][m][1], 0)];
  return s / 2.;
#else
  for (int m = 0; m < 4; m++)
    s += F->f[f][lidx(F, base[0] + ftan[d][m][0], base[1] + ftan[d][m][1], base[2] + ftan[d][m][2])];
  return s / 4.;
#endif
}

/* restriction of one level (foreach_coarse_level): scalar f (average) or
   the face vector alpha (face_average, both faces of each coarse cell) */
static void restrict_level(ot_tree *t, int f, int l) {
  otl *L = &t->L[l];
  olist *A = &t->active[l];
  for (int n = 0; n < A->n; n++) {
    size_t q = A->p[n];
    if (is_leaf(L, q)) continue;
    if (f == T_AL) {
      for (int d = 0; d < DIM; d++) {
        L->f[T_AL + d][q] = fine_face_avg(t, T_AL + d, d, l, q, 0);
        L->f[T_AL + d][q + lstride(L, d)] = fine_face_avg(t, T_AL + d, d, l, q, 1);
      }
    } else
      L->f[f][q] = child_avg(t, f, l, q);
  }
}

/* ---------------------------------------------------------- prolongation */

/* bilinear (multigrid-common.h:124-140) for the cell c at level l */
static double bilinear(ot_tree *t, int f, int l, const int c[3]) {
  otl *P = &t->L[l - 1];
  const double *s = P->f[f];
  int pc[3], ch[3];
  for (int d = 0; d < 3; d++) {
    pc[d] = (c[d] + GH) / 2;
    ch[d] = 2 * ((c[d] + GH) % 2) - 1;
  }
#if DIM == 2
  double c0 = s[lidx(P, pc[0], pc[1], 0)], c1 = s[lidx(P, pc[0] + ch[0], pc[1], 0)];
  double c2 = s[lidx(P, pc[0], pc[1] + ch[1], 0)], c3 = s[lidx(P, pc[0] + ch[0], pc[1] + ch[1], 0)];
  return (9. * c0 + 3. * (c1 + c2) + c3) / 16.;
#else
  double c000 = s[lidx(P, pc[0], pc[1], pc[2])];
  double c100 = s[lidx(P, pc[0] + ch[0], pc[1], pc[2])];
  double c010 = s[lidx(P, pc[0], pc[1] + ch[1], pc[2])];
  double c001 = s[lidx(P, pc[0], pc[1], pc[2] + ch[2])];
  double c110 = s[lidx(P, pc[0] + ch[0], pc[1] + ch[1], pc[2])];
  double c101 = s[lidx(P, pc[0] + ch[0], pc[1], pc[2] + ch[2])];
  double c011 = s[lidx(P, pc[0], pc[1] + ch[1], pc[2] + ch[2])];
  double c111 = s[lidx(P, pc[0] + ch[0], pc[1] + ch[1], pc[2] + ch[2])];
  return (27. * c000 + 9. * (c100 + c010 + c001) + 3. * (c110 + c101 + c011) + c111) / 64.;
#endif
}

/* refine_bilinear of a halo cell (children of a prolongation leaf) */
static void halo_bilinear(ot_tree *t, int f, int l, size_t q) {
  otl *L = &t->L[l], *F = &t->L[l + 1];
  int c[3];
  lijk(L, q, c);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
        int fc[3] = {2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0};
        F->f[f][lidx(F, fc[0], fc[1], fc[2])] = bilinear(t, f, l + 1, fc);
      }
}

/* refine_face_x (tree-common.h:464-511) for alpha, component d, halo of the
   prolongation leaf q at level l */
static void halo_face_prolong(ot_tree *t, int d, int l, size_t q) {
  otl *L = &t->L[l], *F = &t->L[l + 1];
  double *v = L->f[T_AL + d], *fv = F->f[T_AL + d];
  int t1 = (d + 1) % DIM, t2 = (d + 2) % DIM;
  ptrdiff_t sd = lstride(L, d), s1 = lstride(L, t1), s2 = lstride(L, t2);
  (void)s2;
  int c[3];
  lijk(L, q, c);
  int base[3] = {2 * c[0] - GH, 2 * c[1] - GH, 2 * c[2] - GH};
  size_t m = q - sd, p = q + sd;
  /* child (o, j[, k]) of the prolongation leaf along d, tangential j, k */
#define FCH(o, jj, kk)                                                                             \
  ({                                                                                               \
    int _x[3] = {base[0], base[1], base[2]};                                                       \
    _x[d] += (o);                                                                                  \
    _x[t1] += (jj);                                                                                \
    if (DIM == 3) _x[t2] += (kk);                                                                  \
    lidx(F, _x[0], _x[1], DIM == 3 ? _x[2] : 0);                                                   \
  })
  if (!is_refined(L, m)) {
    double g1 = (v[q + s1] - v[q - s1]) / 8.;
#if DIM == 3
    double g2 = (v[q + s2] - v[q - s2]) / 8.;
    for (int j = 0; j <= 1; j++)
      for (int k = 0; k <= 1; k++) fv[FCH(0, j, k)] = v[q] + (2 * j - 1) * g1 + (2 * k - 1) * g2;
#else
    for (int j = 0; j <= 1; j++) fv[FCH(0, j, 0)] = v[q] + (2 * j - 1) * g1;
#endif
  }
  if (!is_refined(L, p) && L->nb[p]) {
    double g1 = (v[p + s1] - v[p - s1]) / 8.;
#if DIM == 3
    double g2 = (v[p + s2] - v[p - s2]) / 8.;
    for (int j = 0; j <= 1; j++)
      for (int k = 0; k <= 1; k++) fv[FCH(2, j, k)] = v[p] + (2 * j - 1) * g1 + (2 * k - 1) * g2;
#else
    for (int j = 0; j <= 1; j++) fv[FCH(2, j, 0)] = v[p] + (2 * j - 1) * g1;
#endif
  }
  {
    double g1 = (v[q + s1] - v[q - s1] + v[p + s1] - v[p - s1]) / 16.;
#if DIM == 3
    double g2 = (v[q + s2] - v[q - s2] + v[p + s2] - v[p - s2]) / 16.;
    for (int j = 0; j <= 1; j++)
      for (int k = 0; k <= 1; k++) fv[FCH(1, j, k)] = (v[q] + v[p]) / 2. + (2 * j - 1) * g1 + (2 * k - 1) * g2;
#else
    for (int j = 0; j <= 1; j++) fv[FCH(1, j, 0)] = (v[q] + v[p]) / 2. + (2 * j - 1) * g1;
#endif
  }
#undef FCH
}

/* ---------------------------------------------------------- tree boundary */

/* tree_boundary_level ({f}, l) for one scalar with BC kind bk
   (tree-common.h:587-716); l < 0 means depth() */
static void tree_boundary_scalar(ot_tree *t, int f, int bk, int l) {
  int depth = l < 0 ? t->depth : l;
  if (is_full(t)) {
    box_scalar(t, f, bk, depth);
    return;
  }
  for (int m = depth - 1; m >= 0; m--) restrict_level(t, f, m);
  box_scalar(t, f, bk, 0);
  for (int i = 0; i < depth; i++) {
    olist *P = &t->prolong[i];
    for (int n = 0; n < P->n; n++) halo_bilinear(t, f, i, P->p[n]);
    box_scalar(t, f, bk, i + 1);
  }
}

/* halo_face for a face field f (T_G or T_AL), tree-common.h:373-409 */
static void halo_face(ot_tree *t, int f) {
  for (int l = t->depth - 1; l >= 0; l--) {
    otl *L = &t->L[l];
    olist *P = &t->prolong[l];
    for (int n = 0; n < P->n; n++) {
      size_t q = P->p[n];
      for (int d = 0; d < DIM; d++) {
        ptrdiff_t sd = lstride(L, d);
        if (is_refined(L, q - sd)) L->f[f + d][q] = fine_face_avg(t, f + d, d, l, q, 0);
        if (is_refined(L, q + sd)) L->f[f + d][q + sd] = fine_face_avg(t, f + d, d, l, q, 1);
      }
    }
  }
}

/* restriction ({alpha, lambda}) of poisson() (poisson.h:418-420) through
   tree_restriction: boundary() = halo_face (alpha, flux part) then
   tree_boundary_level of both; a full tree then also runs
   multigrid_restriction (lambda before alpha on each level) */
void ot_restriction_alpha_lambda(ot_tree *t) {
  if (!t->alpha_field && !t->lambda_field) return;
  ot_update(t);
  if (t->alpha_field) halo_face(t, T_AL);
  int full = is_full(t), depth = t->depth;
  if (full) {
    if (t->lambda_field) box_scalar(t, T_LAM, BK_SYM, depth);
    if (t->alpha_field) box_faces(t, depth);
    for (int l = depth - 1; l >= 0; l--) {
      if (t->lambda_field) restrict_level(t, T_LAM, l);
      if (t->alpha_field) restrict_level(t, T_AL, l);
      if (t->lambda_field) box_scalar(t, T_LAM, BK_SYM, l);
      if (t->alpha_field) box_faces(t, l);
    }
    return;
  }
  for (int l = depth - 1; l >= 0; l--) {
    if (t->lambda_field) restrict_level(t, T_LAM, l);
    if (t->alpha_field) restrict_level(t, T_AL, l);
  }
  if (t->lambda_field) box_scalar(t, T_LAM, BK_SYM, 0);
  if (t->alpha_field) box_faces(t, 0);
  for (int i = 0; i < depth; i++) {
    olist *P = &t->prolong[i];
    for (int n = 0; n < P->n; n++) {
      if (t->lambda_field) halo_bilinear(t, T_LAM, i, P->p[n]);
      if (t->alpha_field)
        for (int d = 0; d < DIM; d++) halo_face_prolong(t, d, i, P->p[n]);
    }
    if (t->lambda_field) box_scalar(t, T_LAM, BK_SYM, i + 1);
    if (t->alpha_field) box_faces(t, i + 1);
  }
}

/* restriction (res) of mg_cycle (poisson.h:42): tree_restriction */
static void restriction_res(ot_tree *t) {
  if (is_full(t)) {
    box_scalar(t, T_RES, BK_SYM, t->depth);
    for (int l = t->depth - 1; l >= 0; l--) {
      restrict_level(t, T_RES, l);
      box_scalar(t, T_RES, BK_SYM, l);
    }
  } else
    tree_boundary_scalar(t, T_RES, BK_SYM, -1);
}

/* ---------------------------------------------------------- relax/residual */

static inline double alpha_at(const ot_tree *t, const otl *L, int d, size_t q) {
  return t->alpha_field ? L->f[T_AL + d][q] : t->alpha_val;
}
static inline double lambda_at(const ot_tree *t, const otl *L, size_t q) {
  return t->lambda_field ? L->f[T_LAM][q] : t->lambda_val;
}

static inline double relax_value(const ot_tree *t, const otl *L, size_t q, double D2) {
  const double *a = L->f[T_DA];
  double n = -D2 * L->f[T_RES][q], d = -lambda_at(t, L, q) * D2;
  for (int c = 0; c < DIM; c++) {
    ptrdiff_t sd = lstride(L, c);
    double ah = alpha_at(t, L, c, q + sd), al = alpha_at(t, L, c, q);
    n += ah * a[q + sd] + al * a[q - sd];
    d += ah + al;
  }
  return n / d;
}

/* the cells of foreach_level_or_leaf(l), in order, passed to fn */
#define FOREACH_LEVEL_OR_LEAF(t, l, ...)                                                           \
  for (int _l1 = (l); _l1 >= 0; _l1--) {                                                           \
    otl *L = &(t)->L[_l1];                                                                         \
    double D = delta_l((t), _l1), D2 = D * D;                                                      \
    (void)D2;                                                                                      \
    olist *A = &(t)->active[_l1];                                                                  \
    for (int _n = 0; _n < A->n; _n++) {                                                            \
      size_t q = A->p[_n];                                                                         \
      if (_l1 == (l) || is_leaf(L, q)) { __VA_ARGS__ }                                             \
    }                                                                                              \
  }

void ot_relax(ot_tree *t, int l) {
  ot_update(t);
  if (t->smoother == SM_GS) {
    FOREACH_LEVEL_OR_LEAF(t, l, { L->f[T_DA][q] = relax_value(t, L, q, D2); })
  } else if (t->smoother == SM_JACOBI) {
    FOREACH_LEVEL_OR_LEAF(t, l, { L->f[T_C][q] = relax_value(t, L, q, D2); })
    FOREACH_LEVEL_OR_LEAF(t, l, { L->f[T_DA][q] = (L->f[T_DA][q] + 2. * L->f[T_C][q]) / 3.; })
  } else {
    for (int colour = 0; colour < 2; colour++) {
      if (colour == 1)
        for (int m = 0; m <= l; m++) box_scalar(t, T_DA, BK_HOM, m);
      FOREACH_LEVEL_OR_LEAF(t, l, {
        int c[3];
        lijk(L, q, c);
        if (((c[0] + c[1] + c[2]) & 1) == colour) L->f[T_DA][q] = relax_value(t, L, q, D2);
      })
    }
  }
}

/* residual, TREE form (poisson.h:355-372) */
double ot_residual(ot_tree *t) {
  ot_update(t);
  tree_boundary_scalar(t, T_A, BK_A, -1);
  for (int n = 0; n < t->nfaces; n++) {
    otl *L = &t->L[t->faces[n].level];
    double D = delta_l(t, t->faces[n].level);
    size_t q = t->faces[n].idx;
    for (int d = 0; d < DIM; d++)
      if (t->faces[n].flags & (1 << d)) {
        const double *a = L->f[T_A];
        L->f[T_G + d][q] = alpha_at(t, L, d, q) * ((a[q] - a[q - lstride(L, d)]) / D);
      }
  }
  halo_face(t, T_G);
  double maxres = 0.;
  for (long n = 0; n < t->nleaves; n++) {
    int l = t->leaves_l.p[n];
    otl *L = &t->L[l];
    size_t q = t->leaves_i.p[n];
    double D = delta_l(t, l);
    double res = L->f[T_B][q] - lambda_at(t, L, q) * L->f[T_A][q];
    for (int d = 0; d < DIM; d++) {
      const double *g = L->f[T_G + d];
      res -= (g[q + lstride(L, d)] - g[q]) / D;
    }
    L->f[T_RES][q] = res;
    if (fabs(res) > maxres) maxres = fabs(res);
  }
  return maxres;
}

/* boundary_level (da, l) of mg_cycle */
static void boundary_da(ot_tree *t, int l) { tree_boundary_scalar(t, T_DA, BK_HOM, l); }

void ot_mg_cycle(ot_tree *t, int nrelax, int minlevel, int maxlevel) {
  ot_update(t);
  restriction_res(t);
  minlevel = minlevel < maxlevel ? minlevel : maxlevel;
  for (int l = minlevel; l <= maxlevel; l++) {
    if (l == minlevel) {
      FOREACH_LEVEL_OR_LEAF(t, l, { L->f[T_DA][q] = 0.; })
    } else {
      otl *L = &t->L[l];
      olist *A = &t->active[l];
      for (int n = 0; n < A->n; n++) {
        int c[3];
        lijk(L, A->p[n], c);
        L->f[T_DA][A->p[n]] = bilinear(t, T_DA, l, c);
      }
    }
    boundary_da(t, l);
    for (int it = 0; it < nrelax; it++) {
      ot_relax(t, l);
      boundary_da(t, l);
    }
  }
  for (long n = 0; n < t->nleaves; n++) {
    otl *L = &t->L[t->leaves_l.p[n]];
    size_t q = t->leaves_i.p[n];
    L->f[T_A][q] += L->f[T_DA][q];
  }
}

double ot_sum_b(ot_tree *t) {
  ot_update(t);
  double sum = 0.;
  for (long n = 0; n < t->nleaves; n++) sum += t->L[t->leaves_l.p[n]].f[T_B][t->leaves_i.p[n]];
  return sum;
}

int ot_mg_solve(ot_tree *t, double tolerance, int nrelax, int minlevel, ot_mgstats *out) {
  ot_mgstats s;
  memset(&s, 0, sizeof s);
  s.sum = ot_sum_b(t);
  s.nrelax = nrelax > 0 ? nrelax : 4;
  double resb;
  resb = s.resb = s.resa = ot_residual(t);
  if (tolerance == 0.) tolerance = t->TOLERANCE;
  for (s.i = 0; s.i < t->NITERMAX && (s.i < t->NITERMIN || s.resa > tolerance); s.i++) {
    ot_mg_cycle(t, s.nrelax, minlevel, t->depth);
    s.resa = ot_residual(t);
    if (s.resa > tolerance) {
      if (resb / s.resa < 1.2 && s.nrelax < 100) s.nrelax++;
      else if (resb / s.resa > 10 && s.nrelax > 2) s.nrelax--;
    }
    resb = s.resa;
  }
  s.minlevel = minlevel;
  if (out) *out = s;
  return 0;
}

int ot_poisson(ot_tree *t, double tolerance, int nrelax, int minlevel, ot_mgstats *out) {
  ot_restriction_alpha_lambda(t);
  double defaultol = t->TOLERANCE;
  if (tolerance) t->TOLERANCE = tolerance;
  int r = ot_mg_solve(t, 0., nrelax, minlevel > 1 ? minlevel : 1, out);
  if (tolerance) t->TOLERANCE = defaultol;
  return r;
}

/* project(), poisson.h:469-509, on the tree.  uf: the face velocity on the
   face cache (one value per set flag bit, entry order then dimension, as
   ot_set_faces), updated in place; p is T_A.  The divergence's stencil
   reads uf.x[1], so uf first gets its automatic boundary, halo_face (the
   fine fluxes averaged onto the coarse faces of resolution jumps); the
   face update's stencil reads p[-1], so p first gets the full tree
   boundary with its own conditions. */
int ot_project(ot_tree *t, double *uf, double dt, int nrelax, ot_mgstats *out) {
  ot_update(t);
  if (dt == 0.) dt = 1.;
  if (nrelax == 0) nrelax = 4;
  long v = 0;
  for (int n = 0; n < t->nfaces; n++)
    for (int d = 0; d < DIM; d++)
      if (t->faces[n].flags & (1 << d)) t->L[t->faces[n].level].f[T_U + d][t->faces[n].idx] = uf[v++];
  halo_face(t, T_U);
  for (long n = 0; n < t->nleaves; n++) {
    int l = t->leaves_l.p[n];
    otl *L = &t->L[l];
    size_t q = t->leaves_i.p[n];
    double div = 0.;
    for (int d = 0; d < DIM; d++) div += L->f[T_U + d][q + lstride(L, d)] - L->f[T_U + d][q];
    div /= dt * delta_l(t, l);
    L->f[T_B][q] = div;
  }
  int r = ot_poisson(t, t->TOLERANCE / (dt * dt), nrelax, 0, out);
  tree_boundary_scalar(t, T_A, BK_A, -1);
  v = 0;
  for (int n = 0; n < t->nfaces; n++) {
    otl *L = &t->L[t->faces[n].level];
    double D = delta_l(t, t->faces[n].level);
    size_t q = t->faces[n].idx;
    for (int d = 0; d < DIM; d++)
      if (t->faces[n].flags & (1 << d)) {
        const double *p = L->f[T_A];
        L->f[T_U + d][q] -= dt * alpha_at(t, L, d, q) * ((p[q] - p[q - lstride(L, d)]) / D);
        uf[v++] = L->f[T_U + d][q];
      }
  }
  return r;
}

/* ------------------------------------------------ the transposed operator

   The adjoint of a tree solve needs the transpose of the leaf operator
   r(x) = -lambda x - sum_d (g_d[hi] - g_d[])/Delta, g = halo_face(alpha
   (x[] - x[-1])/Delta) on the face cache, x after its tree boundary with the
   homogeneous conditions (the residual of poisson.h:355-372 with b = 0 and
   zero boundary values).  The tree operator is not symmetric (the halo
   values are bilinear prolongations of restrictions), so its transpose is
   built stage by stage, each stage of the forward chain transposed in
   reverse order:

     residual_T    y (leaves) -> xbar[q] = -lambda y, gbar[q] += y/Delta,
                   gbar[hi] -= y/Delta
     halo_face_T   each coarse target face's adjoint goes to its fine faces
                   (divided by 2^(dim-1)) and the target's is cleared
     faces_T       c = (alpha gbar)/Delta of each face-cache face in cache
                   order: xbar[] += c, xbar[-1] -= c
     boundary_T    tree_boundary_level's stages in reverse: box ghosts to
                   their sources (the homogeneous rule's factor), halo
                   children to the coarse ring (the bilinear weights, over
                   16 in 2D and 64 in 3D), coarse cells to their children
                   (divided by 2^dim), each written cell's adjoint cleared.

   xbar lives in T_C, gbar in T_G + d.  Every accumulation runs in the
   order written here (the order the GPU's gathers reproduce). */

/* the homogeneous (or, sym = 1, symmetry) box rule of ghost q:
   ghost = k * s[src]; src < 0: none */
#define HOMF(b) (sym ? 1. : (t->bc[b] == BC_DIRICHLET ? -1. : 1.))
static long box_rule_hom(ot_tree *t, int l, size_t q, double *k, int sym) {
  otl *L = &t->L[l];
  int g[3];
  lijk(L, q, g);
  for (int d = 0; d < DIM; d++)
    for (int i = -1; i <= 1; i += 2) {
      size_t p = q + i * lstride(L, d);
      if (is_nbr(L, p)) {
        *k = HOMF(bid_of(L, q));
        return (long)p;
      }
    }
  for (int pl = 0; pl < (DIM == 3 ? 3 : 1); pl++) {
    int da = pl, db = (pl + 1) % DIM;
    for (int i = -1; i <= 1; i += 2)
      for (int j = -1; j <= 1; j += 2) {
        size_t nn = q + i * lstride(L, da) + j * lstride(L, db);
        size_t n1 = q + i * lstride(L, da), n2 = q + j * lstride(L, db);
        if (L->alloc[nn] && is_nbr(L, nn) && L->alloc[n1] && is_boundary(L, n1) && L->alloc[n2] &&
            is_boundary(L, n2)) {
          double h1 = HOMF(bid_of(L, n1));
          double h2 = HOMF(bid_of(L, n2));
          *k = h1 + h2 - 1.;
          return (long)nn;
        }
      }
  }
#if DIM == 3
  for (int i = -1; i <= 1; i += 2)
    for (int j = -1; j <= 1; j += 2)
      for (int kk = -1; kk <= 1; kk += 2) {
        size_t n0 = q + i * lstride(L, 0) + j * lstride(L, 1) + kk * lstride(L, 2);
        size_t n1 = q + i * lstride(L, 0) + j * lstride(L, 1);
        size_t n2 = q + i * lstride(L, 0) + kk * lstride(L, 2);
        size_t n3 = q + j * lstride(L, 1) + kk * lstride(L, 2);
        if (is_nbr(L, n0) && is_boundary(L, n1) && is_boundary(L, n2) && is_boundary(L, n3)) {
          double h1 = HOMF(bid_of(L, n1));
          double h2 = HOMF(bid_of(L, n2));
          double h3 = HOMF(bid_of(L, n3));
          *k = h1 + h2 + h3 - 2.;
          return (long)n0;
        }
      }
#endif
  return -1;
}

#undef HOMF

static void box_T(ot_tree *t, int f, int l, int sym) {
  otl *L = &t->L[l];
  double *x = L->f[f];
  for (int n = 0; n < t->bnd[l].n; n++) {
    size_t q = t->bnd[l].p[n];
    double k;
    long src = box_rule_hom(t, l, q, &k, sym);
    if (src >= 0) x[src] += k * x[q];
    x[q] = 0.;
  }
}

/* bilinear (multigrid-common.h:124-140) of fine cell fc at level l+1
   transposed: its adjoint onto the coarse ring at level l, then cleared */
static void bilinear_T(ot_tree *t, int f, int l, const int fc[3]) {
  otl *L = &t->L[l], *F = &t->L[l + 1];
  double *xc = L->f[f], *xf = F->f[f];
  size_t fq = lidx(F, fc[0], fc[1], fc[2]);
  double cb = xf[fq];
  int pc[3], ch[3];
  for (int d = 0; d < 3; d++) {
    pc[d] = (fc[d] + GH) / 2;
    ch[d] = 2 * ((fc[d] + GH) % 2) - 1;
  }
#if DIM == 2
  xc[lidx(L, pc[0], pc[1], 0)] += (9. * cb) / 16.;
  xc[lidx(L, pc[0] + ch[0], pc[1], 0)] += (3. * cb) / 16.;
  xc[lidx(L, pc[0], pc[1] + ch[1], 0)] += (3. * cb) / 16.;
  xc[lidx(L, pc[0] + ch[0], pc[1] + ch[1], 0)] += cb / 16.;
#else
  xc[lidx(L, pc[0], pc[1], pc[2])] += (27. * cb) / 64.;
  xc[lidx(L, pc[0] + ch[0], pc[1], pc[2])] += (9. * cb) / 64.;
  xc[lidx(L, pc[0], pc[1] + ch[1], pc[2])] += (9. * cb) / 64.;
  xc[lidx(L, pc[0], pc[1], pc[2] + ch[2])] += (9. * cb) / 64.;
  xc[lidx(L, pc[0] + ch[0], pc[1] + ch[1], pc[2])] += (3. * cb) / 64.;
  xc[lidx(L, pc[0] + ch[0], pc[1], pc[2] + ch[2])] += (3. * cb) / 64.;
  xc[lidx(L, pc[0], pc[1] + ch[1], pc[2] + ch[2])] += (3. * cb) / 64.;
  xc[lidx(L, pc[0] + ch[0], pc[1] + ch[1], pc[2] + ch[2])] += cb / 64.;
#endif
  xf[fq] = 0.;
}

/* the children of prolongation leaf q at level l back onto the coarse ring */
static void halo_bilinear_T(ot_tree *t, int f, int l, size_t q) {
  otl *L = &t->L[l];
  int c[3];
  lijk(L, q, c);
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
        int fc[3] = {2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0};
        bilinear_T(t, f, l, fc);
      }
}

/* coarse (non-leaf active) cells of level l onto their children */
static void restrict_T(ot_tree *t, int f, int l) {
  otl *L = &t->L[l], *F = &t->L[l + 1];
  olist *A = &t->active[l];
  for (int n = 0; n < A->n; n++) {
    size_t q = A->p[n];
    if (is_leaf(L, q)) continue;
    int c[3];
    lijk(L, q, c);
    double pb = L->f[f][q];
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++)
          F->f[f][lidx(F, 2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0)] += pb / NCH;
    L->f[f][q] = 0.;
  }
}

/* tree_boundary_scalar (f, BK_HOM or BK_SYM, lev) transposed (lev < 0:
   depth) */
static void tree_boundary_T(ot_tree *t, int f, int lev, int sym) {
  int depth = lev < 0 ? t->depth : lev;
  if (is_full(t)) {
    box_T(t, f, depth, sym);
    return;
  }
  for (int i = depth - 1; i >= 0; i--) {
    box_T(t, f, i + 1, sym);
    olist *P = &t->prolong[i];
    for (int n = 0; n < P->n; n++) halo_bilinear_T(t, f, i, P->p[n]);
  }
  box_T(t, f, 0, sym);
  for (int m = 0; m < depth; m++) restrict_T(t, f, m);
}

/* halo_face (T_G) transposed */
static void halo_face_T(ot_tree *t) {
  for (int l = 0; l < t->depth; l++) {
    otl *L = &t->L[l], *F = &t->L[l + 1];
    olist *P = &t->prolong[l];
    for (int n = 0; n < P->n; n++) {
      size_t q = P->p[n];
      int c[3];
      lijk(L, q, c);
      for (int d = 0; d < DIM; d++) {
        ptrdiff_t sd = lstride(L, d);
        for (int hi = 0; hi <= 1; hi++) {
          size_t tq = hi ? q + sd : q;
          if (!is_refined(L, hi ? q + sd : q - sd)) continue;
          double v = L->f[T_G + d][tq];
          int base[3] = {2 * c[0] - GH, 2 * c[1] - GH, 2 * c[2] - GH};
          base[d] += 2 * hi;
#if DIM == 2
          for (int m = 0; m < 2; m++)
            F->f[T_G + d][lidx(F, base[0] + ftan[d][m][0], base[1] + ftan[d][m][1], 0)] += v / 2.;
#else
          for (int m = 0; m < 4; m++)
            F->f[T_G + d][lidx(F, base[0] + ftan[d][m][0], base[1] + ftan[d][m][1], base[2] + ftan[d][m][2])] +=
                v / 4.;
#endif
          L->f[T_G + d][tq] = 0.;
        }
      }
    }
  }
}

/* residual_T and halo_face_T: gbar of y (leaf cache order), xbar's leaf
   values -lambda y; everything else of T_C and T_G cleared first */
static void adjoint_faces(ot_tree *t, const double *y) {
  for (int l = 0; l <= t->depth; l++) {
    otl *L = &t->L[l];
    memset(L->f[T_C], 0, L->ncell * sizeof(double));
    for (int d = 0; d < DIM; d++) memset(L->f[T_G + d], 0, L->ncell * sizeof(double));
  }
  for (long n = 0; n < t->nleaves; n++) {
    int l = t->leaves_l.p[n];
    otl *L = &t->L[l];
    size_t q = t->leaves_i.p[n];
    double D = delta_l(t, l);
    L->f[T_C][q] = -lambda_at(t, L, q) * y[n];
    for (int d = 0; d < DIM; d++) {
      double tt = y[n] / D;
      L->f[T_G + d][q] += tt;
      L->f[T_G + d][q + lstride(L, d)] -= tt;
    }
  }
  halo_face_T(t);
}

/* out = r^T y (leaf cache order) */
void ot_residual_transpose(ot_tree *t, const double *y, double *out) {
  ot_update(t);
  adjoint_faces(t, y);
  for (int n = 0; n < t->nfaces; n++) {
    otl *L = &t->L[t->faces[n].level];
    double D = delta_l(t, t->faces[n].level);
    size_t q = t->faces[n].idx;
    for (int d = 0; d < DIM; d++)
      if (t->faces[n].flags & (1 << d)) {
        double c = (alpha_at(t, L, d, q) * L->f[T_G + d][q]) / D;
        L->f[T_C][q] += c;
        L->f[T_C][q - lstride(L, d)] -= c;
      }
  }
  tree_boundary_T(t, T_C, -1, 0);
  for (long n = 0; n < t->nleaves; n++) out[n] = t->L[t->leaves_l.p[n]].f[T_C][t->leaves_i.p[n]];
}

/* one red-black pass of relax (level l, colour) transposed: each relaxed
   cell's adjoint goes to res (-Delta^2/d) and to its neighbours (alpha/d),
   then is cleared (the pass overwrote the cell without reading it) */
static void relax_pass_T(ot_tree *t, int l, int colour) {
  FOREACH_LEVEL_OR_LEAF(t, l, {
    int c[3];
    lijk(L, q, c);
    if (((c[0] + c[1] + c[2]) & 1) == colour) {
      double *db = L->f[T_DB];
      double d = -lambda_at(t, L, q) * D2;
      for (int cc = 0; cc < DIM; cc++) {
        ptrdiff_t sd = lstride(L, cc);
        d += alpha_at(t, L, cc, q + sd) + alpha_at(t, L, cc, q);
      }
      double v = db[q];
      L->f[T_RB][q] += (-D2 * v) / d;
      for (int cc = 0; cc < DIM; cc++) {
        ptrdiff_t sd = lstride(L, cc);
        db[q + sd] += (alpha_at(t, L, cc, q + sd) * v) / d;
        db[q - sd] += (alpha_at(t, L, cc, q) * v) / d;
      }
      db[q] = 0.;
    }
  })
}

/* mg_cycle (poisson.h:32-90 on the tree, red-black relax) transposed: the
   map res (leaves) -> da (leaves) of one V-cycle from da = 0, applied to
   y (leaf order) in reverse, giving out (leaf order) */
void ot_mg_cycle_T(ot_tree *t, const double *y, double *out, int nrelax, int minlevel, int maxlevel) {
  ot_update(t);
  for (int l = 0; l <= t->depth; l++) {
    otl *L = &t->L[l];
    memset(L->f[T_DB], 0, L->ncell * sizeof(double));
    memset(L->f[T_RB], 0, L->ncell * sizeof(double));
  }
  for (long n = 0; n < t->nleaves; n++) t->L[t->leaves_l.p[n]].f[T_DB][t->leaves_i.p[n]] = y[n];
  minlevel = minlevel < maxlevel ? minlevel : maxlevel;
  for (int l = maxlevel; l >= minlevel; l--) {
    for (int it = 0; it < nrelax; it++) {
      tree_boundary_T(t, T_DB, l, 0);
      relax_pass_T(t, l, 1);
      for (int m = l; m >= 0; m--) box_T(t, T_DB, m, 0);
      relax_pass_T(t, l, 0);
    }
    tree_boundary_T(t, T_DB, l, 0);
    if (l == minlevel) {
      FOREACH_LEVEL_OR_LEAF(t, l, { L->f[T_DB][q] = 0.; })
    } else {
      otl *L = &t->L[l];
      olist *A = &t->active[l];
      for (int n = 0; n < A->n; n++) {
        int c[3];
        lijk(L, A->p[n], c);
        bilinear_T(t, T_DB, l - 1, c);
      }
    }
  }
  /* restriction (res): tree_restriction with the symmetry rules */
  if (is_full(t)) {
    for (int l = 0; l < t->depth; l++) {
      box_T(t, T_RB, l, 1);
      restrict_T(t, T_RB, l);
    }
    box_T(t, T_RB, t->depth, 1);
  } else
    tree_boundary_T(t, T_RB, -1, 1);
  for (long n = 0; n < t->nleaves; n++) out[n] = t->L[t->leaves_l.p[n]].f[T_RB][t->leaves_i.p[n]];
}

/* The adjoint of poisson() on the tree, about the state a of the last
   solve: psi solves L^T psi = g for L x = -r(x), by defect correction with
   the transposed V-cycle M^T (M: the cycle mg_solve runs, red-black,
   minlevel 1, nrelax fixed), which contracts L^T's error exactly as M does
   L's (I - M^T L^T is the transpose of I - L M):
     psi = 0, rho = g;  repeat psi += M^T rho, rho = g + r^T psi
   until max|rho| <= tolerance (at least one cycle, at most nitermax).
   Then dJ/db = psi, dJ/dalpha on every face value of the face cache
   (entry order, then dimension) = gbar(psi) (a[] - a[-1])/Delta with a
   after its tree boundary, and dJ/dlambda = -psi a at the leaves.
   galpha / glambda may be NULL.  stats: i, resb (max|g|), resa. */
int ot_adjoint(ot_tree *t, const double *g, double tolerance, int nitermax, int nrelax, double *psi,
               double *galpha, double *glambda, ot_mgstats *st) {
  ot_update(t);
  long N = t->nleaves;
  double *rho = malloc(N * sizeof(double)), *tr = malloc(N * sizeof(double));
  double resb = 0., resa;
  for (long n = 0; n < N; n++) {
    rho[n] = g[n];
    psi[n] = 0.;
    if (fabs(rho[n]) > resb) resb = fabs(rho[n]);
  }
  resa = resb;
  int it;
  for (it = 0; it < nitermax && (it < 1 || resa > tolerance); it++) {
    ot_mg_cycle_T(t, rho, tr, nrelax, 1, t->depth);
    for (long n = 0; n < N; n++) psi[n] += tr[n];
    ot_residual_transpose(t, psi, tr);
    resa = 0.;
    for (long n = 0; n < N; n++) {
      rho[n] = g[n] + tr[n];
      if (fabs(rho[n]) > resa) resa = fabs(rho[n]);
    }
  }
  if (galpha) {
    tree_boundary_scalar(t, T_A, BK_A, -1);
    adjoint_faces(t, psi);
    long v = 0;
    for (int n = 0; n < t->nfaces; n++) {
      otl *L = &t->L[t->faces[n].level];
      double D = delta_l(t, t->faces[n].level);
      size_t q = t->faces[n].idx;
      const double *a = L->f[T_A];
      for (int d = 0; d < DIM; d++)
        if (t->faces[n].flags & (1 << d)) galpha[v++] = L->f[T_G + d][q] * ((a[q] - a[q - lstride(L, d)]) / D);
    }
  }
  if (glambda)
    for (long n = 0; n < N; n++) glambda[n] = -psi[n] * t->L[t->leaves_l.p[n]].f[T_A][t->leaves_i.p[n]];
  if (st) {
    memset(st, 0, sizeof *st);
    st->i = it;
    st->resb = resb;
    st->resa = resa;
    st->nrelax = nrelax;
    st->minlevel = 1;
  }
  free(rho);
  free(tr);
  return 0;
}

/* the whole tree boundary of a field (for tests): l < 0 = depth */
void ot_boundary(ot_tree *t, int f, int bk, int l) {
  ot_update(t);
  tree_boundary_scalar(t, f, bk, l);
}

/* ---------------------------------------------------------- adaptation */

/* the fields adapt_wavelet keeps up to date (its default `list`, all
   fields): the user scalars and the solver's a and b, with their boundary
   conditions */
#define NADAPT (NUS + 2)
static const int adapt_list[NADAPT] = {T_S, T_A, T_B, T_SX, T_SX + 1, T_SX + 2};
static const int adapt_bk[NADAPT] = {BK_SYM, BK_A, BK_SYM, BK_SYM, BK_SYM, BK_SYM};
static int user_scalar(int k) { return k == 0 ? T_S : T_SX + k - 1; }
/* the field index of user scalar k, or -1 */
int ot_user_scalar(int k) { return k >= 0 && k < NUS ? user_scalar(k) : -1; }

/* the estimated scalars of one adaptation and their tolerances */
typedef struct {
  int n, f[NUS];
  double max[NUS];
} adapt_slist;

/* the refinement pass of adapt_wavelet (grid/tree-common.h:196-265): a
   pre-order foreach_cell() in which `continue` skips a cell's children */
static void adapt_refine_visit(ot_tree *t, int l, const int c[3], const adapt_slist *S, int maxlevel, int minlevel,
                               int *nf) {
  otl *L = &t->L[l];
  size_t q = lidx(L, c[0], c[1], c[2]);
  if (!L->alloc[q] || !(L->flags[q] & C_ACTIVE)) return;
  if (L->flags[q] & C_LEAF) {
    if (L->flags[q] & C_TOO_COARSE) {
      L->flags[q] &= ~C_TOO_COARSE;
      refine_cell_f(t, l, c, C_REFINED, adapt_list, NADAPT);
      (*nf)++;
    }
    return;
  }
  if (L->flags[q] & C_REFINED) { /* refined earlier in this pass: skip its children */
    L->flags[q] &= ~C_TOO_COARSE;
    return;
  }
  {
    otl *F = &t->L[l + 1];
    size_t ch[NCH];
    double sc[NCH];
    int m = 0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++, m++)
          ch[m] = lidx(F, 2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0);
    const int lev = l + 1;
    /* each estimated scalar in turn updates the children's flags */
    for (int k = 0; k < (PLANT(12) ? 1 : S->n); k++) {
      const int f = S->f[k];
      const double max = S->max[k];
      if (PLANT(11))
        for (m = 0; m < NCH; m++) F->flags[ch[m]] &= ~(C_TOO_COARSE | C_TOO_FINE | C_JUST_FINE);
      for (m = 0; m < NCH; m++) sc[m] = F->f[f][ch[m]];
      halo_bilinear(t, f, l, q); /* s.prolongation (point, s) */
      for (m = 0; m < NCH; m++) {
        unsigned char *fl = &F->flags[ch[m]];
        double e = fabs(sc[m] - F->f[f][ch[m]]);
        if (e > max && lev < maxlevel) {
          *fl &= ~C_TOO_FINE;
          *fl |= C_TOO_COARSE;
        } else if ((e <= max / 1.5 || lev > maxlevel) && !(*fl & (C_TOO_COARSE | C_JUST_FINE))) {
          if (lev >= minlevel) *fl |= C_TOO_FINE;
        } else if (!(*fl & C_TOO_COARSE)) {
          *fl &= ~C_TOO_FINE;
          *fl |= C_JUST_FINE;
        }
        F->f[f][ch[m]] = sc[m];
      }
    }
    for (m = 0; m < NCH; m++) {
      unsigned char *fl = &F->flags[ch[m]];
      *fl &= ~C_JUST_FINE;
      if (!(*fl & C_LEAF)) {
        *fl &= ~C_TOO_COARSE;
        if (lev >= maxlevel) *fl |= C_TOO_FINE;
      } else if (!(*fl & C_ACTIVE))
        *fl &= ~C_TOO_COARSE;
    }
  }
  if (l < t->depth)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
          int cc[3] = {2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0};
          adapt_refine_visit(t, l + 1, cc, S, maxlevel, minlevel, nf);
        }
}

/* one level l of the coarsening passes (grid/tree-common.h:268-294) */
static void adapt_coarsen_visit(ot_tree *t, int m, const int c[3], int l, int *nc) {
  if (m > t->depth) return;
  otl *L = &t->L[m];
  size_t q = lidx(L, c[0], c[1], c[2]);
  if (!L->alloc[q]) return;
  if (!is_boundary(L, q)) {
    if (m == l) {
      if (!(L->flags[q] & C_LEAF)) {
        if (L->flags[q] & C_REFINED)
          L->flags[q] &= ~(C_REFINED | C_TOO_FINE);
        else if (L->flags[q] & C_TOO_FINE) {
          if (coarsen_cell(t, m, c, adapt_list, NADAPT)) (*nc)++;
          L = &t->L[m];
          L->flags[q] &= ~C_TOO_FINE; /* do not coarsen the parent */
        }
      }
      if (L->flags[q] & C_TOO_FINE)
        L->flags[q] &= ~C_TOO_FINE;
      else if (m > 0) {
        otl *P = &t->L[m - 1];
        size_t p = lidx(P, (c[0] + GH) / 2, (c[1] + GH) / 2, DIM == 3 ? (c[2] + GH) / 2 : 0);
        if (P->flags[p] & C_TOO_FINE) P->flags[p] &= ~C_TOO_FINE;
      }
      return;
    } else if (L->flags[q] & C_LEAF)
      return;
  }
  if (m < t->depth)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
          int cc[3] = {2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0};
          adapt_coarsen_visit(t, m + 1, cc, l, nc);
        }
}

/* boundary (all) for the fields that follow the mesh: box ghosts of the
   depth on a full tree (every leaf is there), else the tree boundary */
static void mesh_fields_boundary(ot_tree *t) {
  if (is_full(t))
    for (int m = 0; m < NADAPT; m++) box_scalar(t, adapt_list[m], adapt_bk[m], t->depth);
  else
    for (int m = 0; m < NADAPT; m++) tree_boundary_scalar(t, adapt_list[m], adapt_bk[m], -1);
}

static void refine_cell_fields(ot_tree *t, int l, const int c[3]) {
  refine_cell_f(t, l, c, 0, adapt_list, NADAPT);
}

/* adapt_wavelet (slist, max, maxlevel, minlevel) on a Cartesian tree
   (grid/tree-common.h:165-303, no MPI) for the user scalars us[0..n-1]
   (indices 0..NUS-1) with tolerances max[]; every user scalar, a and b are
   refined and coarsened with the mesh (the default list, all fields).
   Returns the number of coarsened cells (-1 for a bad scalar index); *nf
   receives the number of refined ones. */
int ot_adapt_wavelet_list(ot_tree *t, int n, const int *us, const double *max, int maxlevel, int minlevel,
                          int *nf) {
  adapt_slist S;
  if (n < 1 || n > NUS) return -1;
  S.n = n;
  for (int k = 0; k < n; k++) {
    if (us[k] < 0 || us[k] >= NUS) return -1;
    S.f[k] = user_scalar(us[k]);
    S.max[k] = max[k];
  }
  ot_update(t);
  /* boundary (list); restriction (slist) */
  if (is_full(t)) {
    for (int m = 0; m < NADAPT; m++) box_scalar(t, adapt_list[m], adapt_bk[m], t->depth);
    for (int k = 0; k < n; k++)
      for (int l = t->depth - 1; l >= 0; l--) { /* multigrid_restriction */
        restrict_level(t, S.f[k], l);
        box_scalar(t, S.f[k], BK_SYM, l);
      }
  } else
    for (int m = 0; m < NADAPT; m++) tree_boundary_scalar(t, adapt_list[m], adapt_bk[m], -1);
  if (minlevel < 1) minlevel = 1;
  int nrf = 0, ncs = 0;
  const int root[3] = {GH, GH, DIM == 3 ? GH : 0};
  adapt_refine_visit(t, 0, root, &S, maxlevel, minlevel, &nrf);
  for (int l = t->depth; l >= 0; l--) adapt_coarsen_visit(t, 0, root, l, &ncs);
  t->dirty = 1;
  if (nrf || ncs) /* mpi_boundary_update (list) */
    for (int m = 0; m < NADAPT; m++) tree_boundary_scalar(t, adapt_list[m], adapt_bk[m], -1);
  if (nf) *nf = nrf;
  return ncs;
}

/* adapt_wavelet ({s}, {max}, maxlevel, minlevel): user scalar 0 */
int ot_adapt_wavelet(ot_tree *t, double max, int maxlevel, int minlevel, int *nf) {
  const int us = 0;
  return ot_adapt_wavelet_list(t, 1, &us, &max, maxlevel, minlevel, nf);
}

/* refine one leaf (refine_cell with the 2:1 cascade, no fields): restore()
   rebuilding a dumped tree (output.h:1340-1360) */
int ot_refine_cell(ot_tree *t, int l, int i, int j, int k) {
  if (l < 0 || l > t->depth) return 0;
  otl *L = &t->L[l];
  size_t q = lidx(L, i, j, k);
  if (!L->alloc[q] || !(L->flags[q] & C_LEAF)) return 0;
  int c[3] = {i, j, k};
  refine_cell(t, l, c);
  t->dirty = 1;
  return 1;
}

/* ---------------------------------------------------------- unrefine() */

/* the non-leaf cells foreach_cell() visits (pre-order, not descending into
   leaves): level and centre; returns their number (arrays may be NULL) */
static long nonleaf_visit(ot_tree *t, int l, const int c[3], int *level, double *xyz, long n) {
  otl *L = &t->L[l];
  size_t q = lidx(L, c[0], c[1], c[2]);
  if (!L->alloc[q] || (L->flags[q] & C_LEAF)) return n;
  if (level) level[n] = l;
  if (xyz) {
    double D = delta_l(t, l);
    xyz[3 * n] = (c[0] - GH + 0.5) * D + t->X0;
    xyz[3 * n + 1] = (c[1] - GH + 0.5) * D + t->Y0;
    xyz[3 * n + 2] = DIM == 3 ? (c[2] - GH + 0.5) * D + t->Z0 : 0.;
  }
  n++;
  if (l < t->depth)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
          int cc[3] = {2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0};
          n = nonleaf_visit(t, l + 1, cc, level, xyz, n);
        }
  return n;
}

long ot_nonleaf(ot_tree *t, int *level, double *xyz) {
  const int root[3] = {GH, GH, DIM == 3 ? GH : 0};
  return nonleaf_visit(t, 0, root, level, xyz, 0);
}

static long unrefine_mark(ot_tree *t, int l, const int c[3], const unsigned char *mark, long n) {
  otl *L = &t->L[l];
  size_t q = lidx(L, c[0], c[1], c[2]);
  if (!L->alloc[q] || (L->flags[q] & C_LEAF)) return n;
  if (mark[n++]) L->flags[q] |= C_TOO_FINE;
  if (l < t->depth)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
          int cc[3] = {2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0};
          n = unrefine_mark(t, l + 1, cc, mark, n);
        }
  return n;
}

static void unrefine_pass(ot_tree *t, int m, const int c[3], int l, int *nc) {
  if (m > t->depth) return;
  otl *L = &t->L[m];
  size_t q = lidx(L, c[0], c[1], c[2]);
  if (!L->alloc[q] || (L->flags[q] & C_LEAF)) return;
  if (m == l) {
    if (L->flags[q] & C_TOO_FINE) {
      if (coarsen_cell(t, m, c, adapt_list, NADAPT)) (*nc)++;
      L = &t->L[m];
      L->flags[q] &= ~C_TOO_FINE;
    }
    return;
  }
  if (m < t->depth)
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < (DIM == 3 ? 2 : 1); e++) {
          int cc[3] = {2 * c[0] - GH + a, 2 * c[1] - GH + b, DIM == 3 ? 2 * c[2] - GH + e : 0};
          unrefine_pass(t, m + 1, cc, l, nc);
        }
}

/* unrefine (cond), grid/tree-common.h:348-371, with cond given as one mark
   per cell of ot_nonleaf's list: marked cells are coarsened from the depth
   up (coarsen_cell, which the 2:1 rule may refuse), the user scalars, a
   and b taking the children's average; then the tree boundary of those
   fields.  Returns the number of coarsened cells. */
int ot_unrefine(ot_tree *t, const unsigned char *mark) {
  ot_update(t);
  /* the fields hold their restriction where coarsen_cell reads it */
  if (is_full(t)) {
    for (int m = 0; m < NADAPT; m++)
      for (int l = t->depth - 1; l >= 0; l--) restrict_level(t, adapt_list[m], l);
  } else
    for (int m = 0; m < NADAPT; m++) tree_boundary_scalar(t, adapt_list[m], adapt_bk[m], -1);
  const int root[3] = {GH, GH, DIM == 3 ? GH : 0};
  unrefine_mark(t, 0, root, mark, 0);
  int nc = 0;
  for (int l = t->depth; l >= 0; l--) unrefine_pass(t, 0, root, l, &nc);
  t->dirty = 1;
  if (nc)
    for (int m = 0; m < NADAPT; m++) tree_boundary_scalar(t, adapt_list[m], adapt_bk[m], -1);
  return nc;
}
