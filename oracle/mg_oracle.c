/*
 * CPU oracle for the multigrid Poisson-Helmholtz path.  TEST INFRASTRUCTURE
 * ONLY: this file is the checker that the GPU library is compared against.
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it; the product library never links or calls it.
 *
 * It restates, in plain C99 and with the same floating-point evaluation
 * order, the reference algorithm on a uniform `multigrid` grid:
 *
 *   storage      grid/multigrid.h:16-73 (one AoS array per level, GHOSTS=2,
 *                (2^l+4)^dim cells, i-major, the k index fastest)
 *   traversal    grid/multigrid.h:191-241 (foreach_level / foreach order)
 *   ghost fill   grid/multigrid.h:543-594 (box_boundary_level, BGHOSTS=1,
 *                direction order right,left,top,bottom,front,back, full
 *                tangential extent) and grid/multigrid.h:638-695 (periodic
 *                copy of both ghost layers, level-0 special case)
 *   BC formulas  common.h:957-960 (dirichlet / neumann and homogeneous forms),
 *                grid/cartesian-common.h:464-467 (symmetry)
 *   restriction  grid/multigrid-common.h:26-32 (restriction_average),
 *                42-63 (face_average), 368-404 (multigrid_restriction)
 *   prolongation grid/multigrid-common.h:124-140 (bilinear)
 *   relax        poisson.h:271-338 (GS in place, or weighted Jacobi 2/3)
 *   residual     poisson.h:373-389 (the !TREE "naive" form)
 *   mg_cycle     poisson.h:32-90
 *   mg_solve     poisson.h:139-235 (nrelax adaptation 198-204)
 *   poisson      poisson.h:402-441
 *
 * One extension that the reference does not have: a red-black Gauss-Seidel
 * smoother (colour = parity of i+j+k, red first, ghost fill between the two
 * colours).  It is the order-independent smoother that the GPU path uses by
 * default, so the oracle defines its exact semantics.
 *
 * Planted mistakes: built with -DORC_PLANT=k (k > 0), the oracle makes one
 * plausible mistake on purpose, so that the tests can show that their
 * mathematical checks catch it (tests/test_oracle_pins.py builds these
 * variants into a temporary directory; the normal build has ORC_PLANT 0):
 *   1  Jacobi update (2a + c)/3 instead of (a + 2c)/3 (poisson.h:325-328)
 *   2  red-black colours swapped (black cells first)
 *   3  no ghost fill of da between the two colours
 *   4  Neumann ghost s - Delta g instead of Delta g + s (common.h:959)
 *   5  nrelax rule with its two thresholds swapped (poisson.h:198-204)
 *   6  relax pairs alpha.x[] with a[1] and alpha.x[1] with a[-1]
 *   7  residual with +lambda a instead of -lambda a (poisson.h:376)
 *   8  embedded flux with its value's sign flipped (embed.h:700)
 *   9  embedded face gradient with the weights (1 + fs) and (1 - fs)
 *      swapped (embed.h:73-74)
 *  10  dirichlet_gradient's quadratic interpolation with a1 and a3
 *      swapped (embed.h:369-370)
 *
 * Build once per dimension: -DDIM=1, 2 or 3, with -ffp-contract=off so that
 * no multiply-add is fused (the reference C is compiled without FMA).
 * Built with -fopenmp, the loops are `omp for schedule(static)` over the
 * outer index exactly like the reference's foreach; the in-place GS smoother
 * then races across thread chunks, as the reference's does.  Run with one
 * thread for the deterministic oracle.
 */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef DIM
#error "compile with -DDIM=1, 2 or 3"
#endif

#define GHOSTS 2

#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif
#define PLANT(k) (ORC_PLANT == (k))

enum { F_A, F_B, F_RES, F_DA, F_LAMBDA, F_C, F_AX };
#define NV (F_AX + DIM)

/* BC_VALUE: the ghost takes the expression itself (`a[left] = expr;`, the
   BC function returning expr; homogeneous form 0), as the reference's
   embedded-boundary test cases set their box conditions */
enum { BC_SYMMETRY = 0, BC_DIRICHLET = 1, BC_NEUMANN = 2, BC_PERIODIC = 3, BC_VALUE = 5 };
enum { SM_GS = 0, SM_JACOBI = 1, SM_RBGS = 2 };

typedef struct {
  int i;
  double resb, resa, sum;
  int nrelax, minlevel;
} orc_mgstats;

#include "embed_oracle.h"

typedef struct {
  int depth;
  double **d;
  double L0, X0, Y0, Z0;
  int bc[2 * DIM];
  double *bcv[2 * DIM];
  int periodic[DIM];
  int lambda_field;
  double lambda_val;
  int alpha_field;
  double alpha_val;
  int smoother;
  int tree; /* full quadtree/octree semantics: Morton GS, tree corner ghosts */
  double TOLERANCE;
  int NITERMAX, NITERMIN;
#if DIM == 2
  orc_embed emb; /* embedded boundary (emb.kind 0: none) */
  int emb_bad;   /* a cut cell's stencil left the level's interior */
#endif
} orc_grid;

#if DIM == 1
#define POWD(n) ((size_t)(n))
#elif DIM == 2
#define POWD(n) ((size_t)(n) * (size_t)(n))
#else
#define POWD(n) ((size_t)(n) * (size_t)(n) * (size_t)(n))
#endif

static inline int msize(int l) { return (1 << l) + 2 * GHOSTS; }

/* cell offset of (i,j,k) at level l, multigrid.h:59-72 */
static inline size_t cidx(int l, int i, int j, int k) {
  size_t m = (size_t)msize(l);
#if DIM == 1
  (void)m; (void)j; (void)k;
  return (size_t)i;
#elif DIM == 2
  (void)k;
  return (size_t)i * m + (size_t)j;
#else
  return ((size_t)i * m + (size_t)j) * m + (size_t)k;
#endif
}

/* cell stride of a unit step along dimension d at level l */
static inline ptrdiff_t dstride(int l, int d) {
  ptrdiff_t m = msize(l);
#if DIM == 1
  (void)d; (void)m;
  return 1;
#elif DIM == 2
  return d == 0 ? m : 1;
#else
  return d == 0 ? m * m : d == 1 ? m : 1;
#endif
}

#define REC(g, l, i, j, k) (&(g)->d[l][cidx(l, i, j, k) * NV])

static inline double delta_l(const orc_grid *g, int l) {
  return g->L0 * (1. / (1 << l)); /* cartesian-common.h:11, multigrid.h:14 */
}

/* ---------------------------------------------------------------- setup */

orc_grid *orc_new(int depth) {
  orc_grid *g = calloc(1, sizeof(orc_grid));
  g->depth = depth;
  g->d = calloc(depth + 1, sizeof(double *));
  for (int l = 0; l <= depth; l++) {
    size_t len = POWD(msize(l)) * NV;
    g->d[l] = calloc(len, sizeof(double)); /* reset(all, 0.), multigrid.h:758 */
  }
  g->L0 = 1.;
  g->lambda_field = 0;
  g->lambda_val = 0.;
  g->alpha_field = 0;
  g->alpha_val = 1.;
  g->smoother = SM_GS;
  g->TOLERANCE = 1e-3; /* poisson.h:102-103 */
  g->NITERMAX = 100;
  g->NITERMIN = 1;
  return g;
}

void orc_free(orc_grid *g) {
  if (!g) return;
  for (int l = 0; l <= g->depth; l++) free(g->d[l]);
  free(g->d);
  for (int b = 0; b < 2 * DIM; b++) free(g->bcv[b]);
#if DIM == 2
  if (g->emb.cs) {
    for (int l = 0; l <= g->depth; l++) {
      free(g->emb.cs[l]);
      free(g->emb.fsx[l]);
      free(g->emb.fsy[l]);
    }
    free(g->emb.cs);
    free(g->emb.fsx);
    free(g->emb.fsy);
    free(g->emb.bc);
  }
#endif
  free(g);
}

int orc_dimension(void) { return DIM; }
int orc_nvar(void) { return NV; }

void orc_set_geometry(orc_grid *g, double L0, double X0, double Y0, double Z0) {
  g->L0 = L0; g->X0 = X0; g->Y0 = Y0; g->Z0 = Z0;
}

void orc_set_niter(orc_grid *g, int nitermax, int nitermin) {
  g->NITERMAX = nitermax;
  g->NITERMIN = nitermin;
}

void orc_set_tolerance(orc_grid *g, double tol) { g->TOLERANCE = tol; }
void orc_set_smoother(orc_grid *g, int s) { g->smoother = s; }
void orc_set_tree(orc_grid *g, int tree) { g->tree = tree; }
void orc_set_lambda_const(orc_grid *g, double v) { g->lambda_field = 0; g->lambda_val = v; }
void orc_use_lambda_field(orc_grid *g) { g->lambda_field = 1; }
void orc_set_alpha_const(orc_grid *g, double v) { g->alpha_field = 0; g->alpha_val = v; }
void orc_use_alpha_field(orc_grid *g) { g->alpha_field = 1; }

/* Boundary condition of `a` on direction b (0 right, 1 left, 2 top,
   3 bottom, 4 front, 5 back).  `vals` holds the BC expression (Dirichlet
   value or Neumann gradient) at the finest level for every tangential
   index over the full extent 0..n+3 of each tangential dimension, in the
   order the tangential indices appear in (i,j,k); NULL means 0. */
void orc_set_bc(orc_grid *g, int b, int kind, const double *vals) {
  g->bc[b] = kind;
  free(g->bcv[b]);
  g->bcv[b] = NULL;
  if (kind == BC_PERIODIC) {
    g->periodic[b / 2] = 1;
    g->bc[b ^ 1] = BC_PERIODIC;
    return;
  }
  if (vals) {
    size_t len = 1;
    for (int t = 0; t < DIM - 1; t++) len *= (size_t)msize(g->depth);
    g->bcv[b] = malloc(len * sizeof(double));
    memcpy(g->bcv[b], vals, len * sizeof(double));
  }
}

double *orc_level_data(orc_grid *g, int l) { return g->d[l]; }

/* copy an interior SoA array (C order over i,j,k) in or out */
static void copy_cells(orc_grid *g, int f, int l, double *buf, int in) {
  int n = 1 << l;
  size_t p = 0;
#if DIM >= 2
  int nj = n;
#else
  int nj = 1;
#endif
#if DIM >= 3
  int nk = n;
#else
  int nk = 1;
#endif
  for (int i = 0; i < n; i++)
    for (int j = 0; j < nj; j++)
      for (int k = 0; k < nk; k++, p++) {
        double *r = REC(g, l, i + GHOSTS, j + (DIM >= 2) * GHOSTS, k + (DIM >= 3) * GHOSTS);
        if (in) r[f] = buf[p];
        else buf[p] = r[f];
      }
}

void orc_set_cell(orc_grid *g, int f, int l, const double *src) { copy_cells(g, f, l, (double *)src, 1); }
void orc_get_cell(orc_grid *g, int f, int l, double *dst) { copy_cells(g, f, l, dst, 0); }

/* face component c: n+1 faces along c, n along the others (C order) */
static void copy_faces(orc_grid *g, int c, int l, double *buf, int in) {
  int n = 1 << l;
  int ext[3] = {1, 1, 1};
  for (int d = 0; d < DIM; d++) ext[d] = n + (d == c);
  size_t p = 0;
  for (int i = 0; i < ext[0]; i++)
    for (int j = 0; j < ext[1]; j++)
      for (int k = 0; k < ext[2]; k++, p++) {
        double *r = REC(g, l, i + GHOSTS, j + (DIM >= 2) * GHOSTS, k + (DIM >= 3) * GHOSTS);
        if (in) r[F_AX + c] = buf[p];
        else buf[p] = r[F_AX + c];
      }
}

void orc_set_face(orc_grid *g, int c, int l, const double *src) { copy_faces(g, c, l, (double *)src, 1); }
void orc_get_face(orc_grid *g, int c, int l, double *dst) { copy_faces(g, c, l, dst, 0); }

/* ------------------------------------------------------------ ghost fill */

/* The value of the BC of direction b applied to the interior value s at
   tangential position (u,v) of that side (common.h:957-960). */
static double bc_apply(orc_grid *g, int b, int l, int homogeneous, int use_bc, double s, int u, int v) {
  int kind = use_bc ? g->bc[b] : BC_SYMMETRY;
  int m = msize(l);
  const double *vals = (use_bc && !homogeneous && l == g->depth) ? g->bcv[b] : NULL;
  double expr = vals ? vals[(size_t)u * (DIM == 3 ? m : 1) + v] : 0.;
  switch (kind) {
  case BC_DIRICHLET: return homogeneous ? -s : 2. * (expr) - s;
  case BC_NEUMANN: return homogeneous ? s : PLANT(4) ? s - delta_l(g, l) * (expr) : delta_l(g, l) * (expr) + s;
  case BC_VALUE: return homogeneous ? 0. : expr;
  default: return s;
  }
}

/* Edge and corner ghosts of a full tree (grid/tree.h:1168-1258, used by
   box_boundary_level 1286-1363): from the diagonal interior cell n,
     2 outside dimensions:  b1(n) + b2(n) - s[n]
     3 outside dimensions:  b1(n) + b2(n) + b3(n) - 2 s[n]
   where b_i is the BC of the side crossed by the i-th face neighbour of the
   ghost; in 2D b1 is the y side, in 3D the order is z, y, x.  (The
   multigrid fill instead reflects the cell once per direction.) */
static void tree_corners(orc_grid *g, int f, int l, int homogeneous, int use_bc) {
  if (DIM < 2) return;
  int n = 1 << l;
  for (int d = 0; d < DIM; d++)
    if (g->periodic[d]) return; /* periodic trees are not handled */
  int lo = GHOSTS - 1, hi = n + GHOSTS;
#if DIM == 2
  for (int i = lo; i <= hi; i++)
    for (int j = lo; j <= hi; j++) {
      int k = 0;
#else
  for (int i = lo; i <= hi; i++)
    for (int j = lo; j <= hi; j++)
      for (int k = lo; k <= hi; k++) {
#endif
        int c[3] = {i, j, k}, mir[3] = {i, j, k}, out[3] = {0, 0, 0}, nout = 0;
        for (int d = 0; d < DIM; d++) {
          if (c[d] == lo) { out[d] = 1 + 1; mir[d] = GHOSTS; nout++; }         /* left / bottom / back */
          else if (c[d] == hi) { out[d] = 1; mir[d] = n + GHOSTS - 1; nout++; } /* right / top / front */
        }
        if (nout < 2) continue;
        double s = REC(g, l, mir[0], mir[1], mir[2])[f];
        double bv[3] = {0., 0., 0.};
        for (int d = 0; d < DIM; d++) {
          if (!out[d]) continue;
          int b = 2 * d + (out[d] == 2);
          int t[2] = {0, 0}, q = 0;
          for (int e = 0; e < DIM; e++)
            if (e != d) t[q++] = mir[e];
          bv[d] = bc_apply(g, b, l, homogeneous, use_bc, s, t[0], DIM == 3 ? t[1] : 0);
        }
        double r;
#if DIM == 2
        r = bv[1] + bv[0] - s;
#else
        if (nout == 2) {
          int d1 = -1, d2 = -1;
          for (int d = 0; d < 3; d++)
            if (out[d]) { if (d1 < 0) d1 = d; else d2 = d; }
          r = bv[d2] + bv[d1] - s;
        } else {
          r = bv[2] + bv[1] + bv[0] - 2. * s;
        }
#endif
        REC(g, l, i, j, k)[f] = r;
      }
}

/* box_boundary_level, multigrid.h:543-594, for one cell-centred field.
   mode 0: the field's inhomogeneous BC (only `a` has one, and only values
   at the finest level); mode 1: the homogeneous version. */
static void box_boundary(orc_grid *g, int f, int l, int homogeneous, int use_bc) {
  int n = 1 << l, m = msize(l);
  double D = delta_l(g, l);
  for (int b = 0; b < 2 * DIM; b++) {
    int kind = use_bc ? g->bc[b] : BC_SYMMETRY;
    if (kind == BC_PERIODIC) continue;
    int d = b / 2, side = (b % 2 == 0) ? 1 : -1;
    int plane = side > 0 ? n + GHOSTS - 1 : GHOSTS;
    ptrdiff_t sd = dstride(l, d);
    const double *vals = (use_bc && !homogeneous && l == g->depth) ? g->bcv[b] : NULL;
    /* tangential dimensions in (i,j,k) order */
    int t1 = -1, t2 = -1;
    for (int e = 0; e < DIM; e++)
      if (e != d) { if (t1 < 0) t1 = e; else t2 = e; }
    int e1 = t1 >= 0 ? m : 1, e2 = t2 >= 0 ? m : 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int u = 0; u < e1; u++)
      for (int v = 0; v < e2; v++) {
        int idx[3] = {0, 0, 0};
        idx[d] = plane;
        if (t1 >= 0) idx[t1] = u;
        if (t2 >= 0) idx[t2] = v;
        double *in = REC(g, l, idx[0], idx[1], idx[2]);
        double *gh = in + side * sd * NV;
        double s = in[f], r;
        double expr = vals ? vals[(size_t)u * e2 + v] : 0.;
        switch (kind) {
        case BC_DIRICHLET: r = homogeneous ? -s : 2. * (expr) - s; break;
        case BC_NEUMANN: r = homogeneous ? s : PLANT(4) ? s - D * (expr) : D * (expr) + s; break;
        case BC_VALUE: r = homogeneous ? 0. : expr; break;
        default: r = s; break;
        }
        gh[f] = r;
      }
  }
  if (g->tree) tree_corners(g, f, l, homogeneous, use_bc);
}

/* periodic_boundary_level_x, multigrid.h:638-695: both ghost layers */
static void periodic_copy(orc_grid *g, int f, int l) {
  int any = 0;
  for (int d = 0; d < DIM; d++) any |= g->periodic[d];
  if (!any) return;
  int n = 1 << l, m = msize(l);
  if (l == 0) {
    /* foreach_level(0) ... foreach_neighbor(): every cell of the 5^dim
       block takes the root value (once per periodic dimension) */
    size_t len = POWD(m);
    double v = REC(g, 0, GHOSTS, (DIM >= 2) * GHOSTS, (DIM >= 3) * GHOSTS)[f];
    for (size_t c = 0; c < len; c++) g->d[0][c * NV + f] = v;
    return;
  }
  for (int d = 0; d < DIM; d++) {
    if (!g->periodic[d]) continue;
    ptrdiff_t sd = dstride(l, d);
    int t1 = -1, t2 = -1;
    for (int e = 0; e < DIM; e++)
      if (e != d) { if (t1 < 0) t1 = e; else t2 = e; }
    int e1 = t1 >= 0 ? m : 1, e2 = t2 >= 0 ? m : 1;
    for (int u = 0; u < e1; u++)
      for (int v = 0; v < e2; v++) {
        int idx[3] = {0, 0, 0};
        if (t1 >= 0) idx[t1] = u;
        if (t2 >= 0) idx[t2] = v;
        double *base = REC(g, l, idx[0], idx[1], idx[2]);
        for (int i = 0; i < GHOSTS; i++)
          base[i * sd * NV + f] = base[(i + n) * sd * NV + f];
        for (int i = n + GHOSTS; i < n + 2 * GHOSTS; i++)
          base[i * sd * NV + f] = base[(i - n) * sd * NV + f];
      }
  }
}

/* boundary_level(list, l) for one cell-centred field (box then periodic,
   the order in which init_grid registers them, multigrid.h:733-745) */
static void boundary_level(orc_grid *g, int f, int l, int homogeneous, int use_bc) {
  box_boundary(g, f, l, homogeneous, use_bc);
  periodic_copy(g, f, l);
}

void orc_boundary(orc_grid *g, int f, int l, int homogeneous) {
  boundary_level(g, f, l, homogeneous, f == F_A || f == F_DA);
}

/* Face vectors have no box BC (cartesian-common.h:566-576).  The periodic
   copy includes a component only when its "VT" companion carries the
   periodic BC (multigrid.h:636-651, cartesian-common.h:903-917): that is
   true for x in 2D and 3D and for y in 3D, not for y in 2D nor z in 3D. */
static void periodic_copy_faces(orc_grid *g, int l) {
  int n = 1 << l, m = msize(l);
  for (int d = 0; d < DIM; d++) {
    if (!g->periodic[d]) continue;
    int copies = (DIM == 1) || d == 0 || (DIM == 3 && d == 1);
    if (!copies) continue;
    if (l == 0) {
      for (int c = 0; c < DIM; c++) {
        size_t len = POWD(m);
        double v = REC(g, 0, GHOSTS, (DIM >= 2) * GHOSTS, (DIM >= 3) * GHOSTS)[F_AX + c];
        for (size_t q = 0; q < len; q++) g->d[0][q * NV + F_AX + c] = v;
      }
      continue;
    }
    ptrdiff_t sd = dstride(l, d);
    int t1 = -1, t2 = -1;
    for (int e = 0; e < DIM; e++)
      if (e != d) { if (t1 < 0) t1 = e; else t2 = e; }
    int e1 = t1 >= 0 ? m : 1, e2 = t2 >= 0 ? m : 1;
    for (int u = 0; u < e1; u++)
      for (int v = 0; v < e2; v++) {
        int idx[3] = {0, 0, 0};
        if (t1 >= 0) idx[t1] = u;
        if (t2 >= 0) idx[t2] = v;
        double *base = REC(g, l, idx[0], idx[1], idx[2]);
        for (int c = 0; c < DIM; c++) {
          for (int i = 0; i < GHOSTS; i++)
            base[i * sd * NV + F_AX + c] = base[(i + n) * sd * NV + F_AX + c];
          for (int i = n + GHOSTS; i < n + 2 * GHOSTS; i++)
            base[i * sd * NV + F_AX + c] = base[(i - n) * sd * NV + F_AX + c];
        }
      }
  }
}

/* ---------------------------------------------------------- loop helpers */

#if DIM == 1
#define LOOP_LEVEL(l, n)                                  \
  _Pragma("omp parallel for schedule(static)")            \
  for (int i = GHOSTS; i < (n) + GHOSTS; i++) {           \
    const int j = 0, k = 0; (void)j; (void)k;
#define END_LOOP }
#elif DIM == 2
#define LOOP_LEVEL(l, n)                                  \
  _Pragma("omp parallel for schedule(static)")            \
  for (int i = GHOSTS; i < (n) + GHOSTS; i++)             \
    for (int j = GHOSTS; j < (n) + GHOSTS; j++) {         \
      const int k = 0; (void)k;
#define END_LOOP }
#else
#define LOOP_LEVEL(l, n)                                  \
  _Pragma("omp parallel for schedule(static)")            \
  for (int i = GHOSTS; i < (n) + GHOSTS; i++)             \
    for (int j = GHOSTS; j < (n) + GHOSTS; j++)           \
      for (int k = GHOSTS; k < (n) + GHOSTS; k++) {
#define END_LOOP }
#endif

/* ------------------------------------------------------------ restriction */

/* restriction_average, multigrid-common.h:26-32: foreach_child order is
   (i,j,k) with k fastest, multigrid.h:357-368 */
static inline double child_sum(orc_grid *g, int f, int l, int i, int j, int k) {
  double sum = 0.;
  int fi = 2 * i - GHOSTS, fj = 2 * j - GHOSTS, fk = 2 * k - GHOSTS;
#if DIM == 1
  (void)fj; (void)fk;
  for (int a = 0; a < 2; a++) sum += REC(g, l + 1, fi + a, 0, 0)[f];
#elif DIM == 2
  (void)fk;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) sum += REC(g, l + 1, fi + a, fj + b, 0)[f];
#else
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int c = 0; c < 2; c++) sum += REC(g, l + 1, fi + a, fj + b, fk + c)[f];
#endif
  return sum;
}

static void restrict_level_average(orc_grid *g, int f, int l) {
  int n = 1 << l;
  LOOP_LEVEL(l, n)
    REC(g, l, i, j, k)[f] = child_sum(g, f, l, i, j, k) / (1 << DIM);
  END_LOOP
}

/* face_average, multigrid-common.h:42-58, with the tangential child order
   that foreach_dimension's rotation gives each component (rotate.lex:163-170) */
__attribute__((unused)) static const int face_tan[3][4][3] = {
#if DIM == 3
  {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}},
  {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 1}},
  {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}},
#elif DIM == 2
  {{0, 0, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 0}},
  {{0, 0, 0}, {1, 0, 0}, {0, 0, 0}, {0, 0, 0}},
  {{0}},
#else
  {{0}}, {{0}}, {{0}},
#endif
};

static inline double face_avg(orc_grid *g, int c, int l, int i, int j, int k, int hi) {
  int fi = 2 * i - GHOSTS, fj = 2 * j - GHOSTS, fk = 2 * k - GHOSTS;
  int off[3] = {0, 0, 0};
  off[c] = 2 * hi;
#if DIM == 1
  (void)fj; (void)fk;
  return REC(g, l + 1, fi + off[0], 0, 0)[F_AX];
#elif DIM == 2
  (void)fk;
  double s = 0.;
  for (int t = 0; t < 2; t++)
    s += REC(g, l + 1, fi + off[0] + face_tan[c][t][0], fj + off[1] + face_tan[c][t][1], 0)[F_AX + c];
  return s / 2.;
#else
  double s = 0.;
  for (int t = 0; t < 4; t++)
    s += REC(g, l + 1, fi + off[0] + face_tan[c][t][0], fj + off[1] + face_tan[c][t][1],
             fk + off[2] + face_tan[c][t][2])[F_AX + c];
  return s / 4.;
#endif
}

static void restrict_level_faces(orc_grid *g, int l) {
  int n = 1 << l;
  LOOP_LEVEL(l, n)
    double *r = REC(g, l, i, j, k);
    for (int c = 0; c < DIM; c++) {
      r[F_AX + c] = face_avg(g, c, l, i, j, k, 0);
      r[dstride(l, c) * NV + F_AX + c] = face_avg(g, c, l, i, j, k, 1);
    }
  END_LOOP
}

/* multigrid_restriction for {res}: levels depth-1..0, each followed by
   boundary_iterate on that level (res inherits b's default symmetry BC) */
void orc_restriction_res(orc_grid *g) {
  for (int l = g->depth - 1; l >= 0; l--) {
    restrict_level_average(g, F_RES, l);
    boundary_level(g, F_RES, l, 0, 0);
  }
}

/* poisson.h:418-420: restriction ({alpha, lambda}); lambda (listdef) is
   restricted before alpha (listc) on each level, multigrid-common.h:388-396 */
void orc_restriction_alpha_lambda(orc_grid *g) {
  if (!g->alpha_field && !g->lambda_field) return;
  for (int l = g->depth - 1; l >= 0; l--) {
    if (g->lambda_field) restrict_level_average(g, F_LAMBDA, l);
    if (g->alpha_field) restrict_level_faces(g, l);
    if (g->lambda_field) boundary_level(g, F_LAMBDA, l, 0, 0);
    if (g->alpha_field) periodic_copy_faces(g, l);
  }
}

/* ----------------------------------------------------------- prolongation */

/* bilinear, multigrid-common.h:124-140; child offsets multigrid.h:175-188 */
static inline double bilinear(orc_grid *g, int f, int l, int i, int j, int k) {
  int pi = (i + GHOSTS) / 2, cx = 2 * ((i + GHOSTS) % 2) - 1;
#if DIM == 1
  (void)j; (void)k;
  const double *c = REC(g, l - 1, pi, 0, 0);
  return (3. * c[f] + REC(g, l - 1, pi + cx, 0, 0)[f]) / 4.;
#elif DIM == 2
  (void)k;
  int pj = (j + GHOSTS) / 2, cy = 2 * ((j + GHOSTS) % 2) - 1;
  double c0 = REC(g, l - 1, pi, pj, 0)[f];
  double c1 = REC(g, l - 1, pi + cx, pj, 0)[f];
  double c2 = REC(g, l - 1, pi, pj + cy, 0)[f];
  double c3 = REC(g, l - 1, pi + cx, pj + cy, 0)[f];
  return (9. * c0 + 3. * (c1 + c2) + c3) / 16.;
#else
  int pj = (j + GHOSTS) / 2, cy = 2 * ((j + GHOSTS) % 2) - 1;
  int pk = (k + GHOSTS) / 2, cz = 2 * ((k + GHOSTS) % 2) - 1;
  double c000 = REC(g, l - 1, pi, pj, pk)[f];
  double c100 = REC(g, l - 1, pi + cx, pj, pk)[f];
  double c010 = REC(g, l - 1, pi, pj + cy, pk)[f];
  double c001 = REC(g, l - 1, pi, pj, pk + cz)[f];
  double c110 = REC(g, l - 1, pi + cx, pj + cy, pk)[f];
  double c101 = REC(g, l - 1, pi + cx, pj, pk + cz)[f];
  double c011 = REC(g, l - 1, pi, pj + cy, pk + cz)[f];
  double c111 = REC(g, l - 1, pi + cx, pj + cy, pk + cz)[f];
  return (27. * c000 + 9. * (c100 + c010 + c001) + 3. * (c110 + c101 + c011) + c111) / 64.;
#endif
}

/* ------------------------------------------------------- relax / residual */

static inline double lambda_at(const orc_grid *g, const double *r) {
  return g->lambda_field ? r[F_LAMBDA] : g->lambda_val;
}

/* alpha on the low face of the cell (alpha.x[]) and on the high face
   (alpha.x[1], stored in the next cell's record) */
static inline double alpha_lo(const orc_grid *g, const double *r, int c) {
  return g->alpha_field ? r[F_AX + c] : g->alpha_val;
}
static inline double alpha_hi(const orc_grid *g, const double *r, ptrdiff_t sd, int c) {
  return g->alpha_field ? r[sd * NV + F_AX + c] : g->alpha_val;
}

/* the Gauss-Seidel / Jacobi value n/d of poisson.h:303-320 */
static inline double relax_value(const orc_grid *g, int l, const double *r, int src, double D2) {
  double n = -D2 * r[F_RES], d = -lambda_at(g, r) * D2;
  for (int c = 0; c < DIM; c++) {
    ptrdiff_t sd = dstride(l, c);
    double ah = alpha_hi(g, r, sd, c), al = alpha_lo(g, r, c);
    if (PLANT(6)) n += al * r[sd * NV + src] + ah * r[-sd * NV + src];
    else n += ah * r[sd * NV + src] + al * r[-sd * NV + src];
    d += ah + al;
  }
  return n / d;
}

#if DIM == 2
/* --------------------------------------------- embedded boundaries, 2D */

/* cs / fs of level l at interior cell (i,j) or face; a read outside the
   interior marks the grid (the reference would read a ghost there) */
static inline double emb_cs(orc_grid *g, int l, int i, int j) {
  int n = 1 << l;
  if (i < 0 || i >= n || j < 0 || j >= n) {
    g->emb_bad = 1;
    return 0.;
  }
  return E_CS(&g->emb, l, i, j);
}
static inline double emb_fx(orc_grid *g, int l, int i, int j) {
  int n = 1 << l;
  if (i < 0 || i > n || j < 0 || j >= n) {
    g->emb_bad = 1;
    return 0.;
  }
  return E_FX(&g->emb, l, i, j);
}
static inline double emb_fy(orc_grid *g, int l, int i, int j) {
  int n = 1 << l;
  if (i < 0 || i >= n || j < 0 || j > n) {
    g->emb_bad = 1;
    return 0.;
  }
  return E_FY(&g->emb, l, i, j);
}
/* the face fraction normal to direction d (rot: y) and the tangential one,
   at offsets (u along the rotated x, v along the rotated y) */
static inline double emb_fn(orc_grid *g, int l, int ci, int cj, int rot, int u, int v) {
  return rot ? emb_fy(g, l, ci + v, cj + u) : emb_fx(g, l, ci + u, cj + v);
}
static inline double emb_ft(orc_grid *g, int l, int ci, int cj, int rot, int u, int v) {
  return rot ? emb_fx(g, l, ci + v, cj + u) : emb_fy(g, l, ci + u, cj + v);
}
static inline double emb_csr(orc_grid *g, int l, int ci, int cj, int rot, int u, int v) {
  return rot ? emb_cs(g, l, ci + v, cj + u) : emb_cs(g, l, ci + u, cj + v);
}
/* field f at offsets (u, v) of interior cell (ci, cj), ghosts included */
static inline double emb_s(orc_grid *g, int l, int ci, int cj, int rot, int u, int v, int f) {
  int di = rot ? v : u, dj = rot ? u : v;
  return REC(g, l, ci + di + GHOSTS, cj + dj + GHOSTS, 0)[f];
}

#define QUADRATIC(x, a1, a2, a3) \
  (((a1) * ((x) - 1.) + (a3) * ((x) + 1.)) * (x) / 2. - (a2) * ((x) - 1.) * ((x) + 1.))

/* dirichlet_gradient_x (embed.h:374-433); rot = 1: the y variant */
static double dirichlet_gradient_r(orc_grid *g, int l, int ci, int cj, int f, int rot, coord2 n, coord2 p,
                                   double bc, double *coef) {
  double D = delta_l(g, l);
  /* the rotated frame: x is the direction of the larger normal component */
  double nx = rot ? n.y : n.x, ny = rot ? n.x : n.y, px = rot ? p.y : p.x, py = rot ? p.x : p.y;
  nx = -nx;
  ny = -ny;
  double d[2], v[2] = {NODATA, NODATA};
  int defined = 1;
  if (defined && !emb_fn(g, l, ci, cj, rot, (nx > 0.), 0)) defined = 0;
  if (defined && !emb_ft(g, l, ci, cj, rot, 0, (ny > 0.))) defined = 0;
  if (defined)
    for (int m = 0; m <= 1; m++) {
      int i = (m + 1) * isign(nx);
      d[m] = (i - px) / nx;
      double y1 = py + d[m] * ny;
      int j = y1 > 0.5 ? 1 : y1 < -0.5 ? -1 : 0;
      y1 -= j;
      if (emb_fn(g, l, ci, cj, rot, i + (i < 0), j) && emb_ft(g, l, ci, cj, rot, i, j) &&
          emb_ft(g, l, ci, cj, rot, i, j + 1) && emb_csr(g, l, ci, cj, rot, i, j - 1) &&
          emb_csr(g, l, ci, cj, rot, i, j) && emb_csr(g, l, ci, cj, rot, i, j + 1))
        v[m] = PLANT(10) ? QUADRATIC(y1, (emb_s(g, l, ci, cj, rot, i, j + 1, f)), (emb_s(g, l, ci, cj, rot, i, j, f)),
                                     (emb_s(g, l, ci, cj, rot, i, j - 1, f)))
                         : QUADRATIC(y1, (emb_s(g, l, ci, cj, rot, i, j - 1, f)), (emb_s(g, l, ci, cj, rot, i, j, f)),
                                     (emb_s(g, l, ci, cj, rot, i, j + 1, f)));
      else
        break;
    }
  if (v[0] == NODATA) {
    double q = fabs(px / nx);
    d[0] = 1e-3 > q ? 1e-3 : q;
    *coef = -1. / (d[0] * D);
    return bc / (d[0] * D);
  }
  *coef = 0.;
  if (v[1] != NODATA) return (d[1] * (bc - v[0]) / d[0] - d[0] * (bc - v[1]) / d[1]) / ((d[1] - d[0]) * D);
  return (bc - v[0]) / (d[0] * D);
}

/* embed_flux (embed.h:654-702) of field f at interior cell (ci, cj) of
   level l: returns the coefficient of the cell's own value, *val the rest;
   bc / dirichlet: the condition's value there and its kind */
static double embed_flux2(orc_grid *g, int l, int ci, int cj, int f, double bc, int dirichlet, double *val) {
  const orc_embed *E = &g->emb;
  *val = 0.;
  double c = E_CS(E, l, ci, cj);
  if (c >= 1. || c <= 0.) return 0.;
  double grad = bc;
  if (!grad && !dirichlet) return 0.;
  coord2 n = facet_normal(E, l, ci, cj), p;
  double alpha = orc_line_alpha(c, n.x, n.y);
  double area = line_length_center(n, alpha, &p);
  double coef = 0.;
  if (dirichlet) {
    normalize2(&n);
    grad = dirichlet_gradient_r(g, l, ci, cj, f, fabs(n.x) >= fabs(n.y) ? 0 : 1, n, p, grad, &coef);
  }
  const double *r = REC(g, l, ci + GHOSTS, cj + GHOSTS, 0);
  double mua = 0., fa = 0.;
  for (int d = 0; d < 2; d++) {
    mua += alpha_lo(g, r, d) + alpha_hi(g, r, dstride(l, d), d);
    fa += (d ? E_FY(E, l, ci, cj) : E_FX(E, l, ci, cj)) + (d ? E_FY(E, l, ci, cj + 1) : E_FX(E, l, ci + 1, cj));
  }
  double D = delta_l(g, l);
  *val = PLANT(8) ? mua / (fa + SEPS) * grad * area / D : -mua / (fa + SEPS) * grad * area / D;
  return -mua / (fa + SEPS) * coef * area / D;
}

/* face_gradient_x (a, h) of the residual at interior cell (ci, cj) of the
   finest level, direction d, h = 0 (low face) or 1 (high face): the
   embedded form on partial faces when `third` is set (embed.h:61-76,
   179-195) */
static double emb_face_gradient(orc_grid *g, int ci, int cj, int d, int h, int f) {
  int l = g->depth, rot = d;
  double D = delta_l(g, l);
  double fsf = emb_fn(g, l, ci, cj, rot, h, 0);
  double am = emb_s(g, l, ci, cj, rot, h - 1, 0, f), a0 = emb_s(g, l, ci, cj, rot, h, 0, f);
  if (!(g->emb.third && fsf < 1. && fsf > 0.)) return (a0 - am) / D;
  int j = isign(emb_fn(g, l, ci, cj, rot, h, 1) - emb_fn(g, l, ci, cj, rot, h, -1));
  if (emb_fn(g, l, ci, cj, rot, h, j) > 0.5 && emb_ft(g, l, ci, cj, rot, h, j + (j < 0)) &&
      emb_ft(g, l, ci, cj, rot, h - 1, j + (j < 0)) && emb_csr(g, l, ci, cj, rot, h, j) &&
      emb_csr(g, l, ci, cj, rot, h - 1, j))
    return ((PLANT(9) ? 1. - fsf : 1. + fsf) * (a0 - am) +
            (PLANT(9) ? 1. + fsf : 1. - fsf) *
                (emb_s(g, l, ci, cj, rot, h, j, f) - emb_s(g, l, ci, cj, rot, h - 1, j, f))) /
           (2. * D);
  return (a0 - am) / D;
}

/* the relax value with the EMBED branch (poisson.h:303-320): returns n/d,
   or 0 and *zero = 1 when d is zero (then b is zeroed too) */
static double relax_value_embed(orc_grid *g, int l, int i, int j, int src, double D2, int *zero) {
  double *r = REC(g, l, i, j, 0);
  double n = -D2 * r[F_RES], d = -lambda_at(g, r) * D2;
  for (int c = 0; c < DIM; c++) {
    ptrdiff_t sd = dstride(l, c);
    double ah = alpha_hi(g, r, sd, c), al = alpha_lo(g, r, c);
    n += ah * r[sd * NV + src] + al * r[-sd * NV + src];
    d += ah + al;
  }
  /* da carries the homogeneous condition: value 0, of the same kind */
  double c, e = embed_flux2(g, l, i - GHOSTS, j - GHOSTS, src, 0., g->emb.kind == 1, &c);
  n -= c * D2;
  d += e * D2;
  *zero = !d;
  return d ? n / d : 0.;
}

/* bilinear_embed (embed.h:738-757): injection when a coarse cell that
   bilinear would use is empty; coarse cs ghosts are the symmetric mirror */
static inline double emb_cs_ghost(orc_grid *g, int l, int i, int j) {
  int n = 1 << l;
  i = i < 0 ? 0 : i >= n ? n - 1 : i;
  j = j < 0 ? 0 : j >= n ? n - 1 : j;
  return E_CS(&g->emb, l, i, j);
}
static double bilinear_embed(orc_grid *g, int f, int l, int i, int j) {
  int pi = (i + GHOSTS) / 2, cx = 2 * ((i + GHOSTS) % 2) - 1;
  int pj = (j + GHOSTS) / 2, cy = 2 * ((j + GHOSTS) % 2) - 1;
  int qi = pi - GHOSTS, qj = pj - GHOSTS;
  if (!emb_cs_ghost(g, l - 1, qi, qj) || !emb_cs_ghost(g, l - 1, qi + cx, qj) ||
      !emb_cs_ghost(g, l - 1, qi, qj + cy) || !emb_cs_ghost(g, l - 1, qi + cx, qj + cy))
    return REC(g, l - 1, pi, pj, 0)[f];
  return bilinear(g, f, l, i, j, 0);
}

/* set the embedded boundary: cs (n^2), fsx ((n+1) x n), fsy (n x (n+1)) of
   the finest level, restricted to every level (restriction_average,
   face_average); kind 1 Dirichlet / 2 Neumann with the condition's value
   per finest cell (at the fragment's centroid); third: embedded face
   gradients in the residual */
void orc_set_embed(orc_grid *g, const double *cs, const double *fsx, const double *fsy, int kind,
                   const double *bc, int third) {
  orc_embed *E = &g->emb;
  int D = g->depth;
  if (!E->cs) {
    E->cs = calloc(D + 1, sizeof(double *));
    E->fsx = calloc(D + 1, sizeof(double *));
    E->fsy = calloc(D + 1, sizeof(double *));
    for (int l = 0; l <= D; l++) {
      size_t n = (size_t)1 << l;
      E->cs[l] = calloc(n * n, sizeof(double));
      E->fsx[l] = calloc((n + 1) * n, sizeof(double));
      E->fsy[l] = calloc(n * (n + 1), sizeof(double));
    }
    E->bc = calloc(((size_t)1 << D) * ((size_t)1 << D), sizeof(double));
  }
  size_t nf = (size_t)1 << D;
  memcpy(E->cs[D], cs, nf * nf * sizeof(double));
  memcpy(E->fsx[D], fsx, (nf + 1) * nf * sizeof(double));
  memcpy(E->fsy[D], fsy, nf * (nf + 1) * sizeof(double));
  if (bc) memcpy(E->bc, bc, nf * nf * sizeof(double));
  else memset(E->bc, 0, nf * nf * sizeof(double));
  for (int l = D - 1; l >= 0; l--) {
    int n = 1 << l, m = 2 * n;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double sum = 0.;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++) sum += E->cs[l + 1][(size_t)(2 * i + a) * m + 2 * j + b];
        E->cs[l][(size_t)i * n + j] = sum / (1 << DIM);
      }
    for (int i = 0; i <= n; i++)
      for (int j = 0; j < n; j++) {
        double sum = 0.;
        for (int t = 0; t < 2; t++) sum += E->fsx[l + 1][(size_t)(2 * i) * m + 2 * j + t];
        E->fsx[l][(size_t)i * n + j] = sum / 2.;
      }
    for (int i = 0; i < n; i++)
      for (int j = 0; j <= n; j++) {
        double sum = 0.;
        for (int t = 0; t < 2; t++) sum += E->fsy[l + 1][(size_t)(2 * i + t) * (m + 1) + 2 * j];
        E->fsy[l][(size_t)i * (n + 1) + j] = sum / 2.;
      }
  }
  E->kind = kind;
  E->third = third;
  g->emb_bad = 0;
}

/* nonzero when a cut cell's stencil left a level's interior since the
   embedded boundary was set */
int orc_embed_bad(orc_grid *g) { return g->emb_bad; }
#endif /* DIM == 2 */

void orc_relax(orc_grid *g, int l) {
  int n = 1 << l;
  double D = delta_l(g, l), D2 = D * D;
#if DIM == 2
  if (g->emb.kind) {
    /* the EMBED branch: a cut cell also reads cells of its own colour
       (dirichlet_gradient), so each red-black pass computes its cells from
       the state before the pass; GS in place, Jacobi as usual */
    if (g->smoother == SM_GS) {
      LOOP_LEVEL(l, n)
        int z;
        double v = relax_value_embed(g, l, i, j, F_DA, D2, &z);
        double *r = REC(g, l, i, j, k);
        if (z) r[F_DA] = r[F_RES] = 0.;
        else r[F_DA] = v;
      END_LOOP
    } else if (g->smoother == SM_JACOBI) {
      LOOP_LEVEL(l, n)
        int z;
        double v = relax_value_embed(g, l, i, j, F_DA, D2, &z);
        double *r = REC(g, l, i, j, k);
        if (z) r[F_C] = r[F_RES] = 0.;
        else r[F_C] = v;
      END_LOOP
      LOOP_LEVEL(l, n)
        double *r = REC(g, l, i, j, k);
        r[F_DA] = (r[F_DA] + 2. * r[F_C]) / 3.;
      END_LOOP
    } else {
      for (int colour = 0; colour < 2; colour++) {
        if (colour == 1) boundary_level(g, F_DA, l, 1, 1);
        LOOP_LEVEL(l, n)
          if (((i + j + k) & 1) == colour) {
            int z;
            double v = relax_value_embed(g, l, i, j, F_DA, D2, &z);
            double *r = REC(g, l, i, j, k);
            if (z) r[F_C] = r[F_RES] = 0.;
            else r[F_C] = v;
          }
        END_LOOP
        LOOP_LEVEL(l, n)
          if (((i + j + k) & 1) == colour) {
            double *r = REC(g, l, i, j, k);
            r[F_DA] = r[F_C];
          }
        END_LOOP
      }
    }
    return;
  }
#endif
  if (g->smoother == SM_GS && g->tree) {
    /* a full tree visits level l in the pre-order of foreach_cell
       (grid/foreach_cell.h:66-123): Morton order, x the high bit */
    size_t ncell = POWD(n);
    for (size_t code = 0; code < ncell; code++) {
      int c[3] = {0, 0, 0};
      for (int b = 0; b < l; b++)
        for (int d = 0; d < DIM; d++)
          c[d] |= (int)((code >> (DIM * b + DIM - 1 - d)) & 1) << b;
      double *r = REC(g, l, c[0] + GHOSTS, (DIM >= 2) * (c[1] + GHOSTS), (DIM >= 3) * (c[2] + GHOSTS));
      r[F_DA] = relax_value(g, l, r, F_DA, D2);
    }
  } else if (g->smoother == SM_GS) {
    LOOP_LEVEL(l, n)
      double *r = REC(g, l, i, j, k);
      r[F_DA] = relax_value(g, l, r, F_DA, D2);
    END_LOOP
  } else if (g->smoother == SM_JACOBI) {
    LOOP_LEVEL(l, n)
      double *r = REC(g, l, i, j, k);
      r[F_C] = relax_value(g, l, r, F_DA, D2);
    END_LOOP
    LOOP_LEVEL(l, n)
      double *r = REC(g, l, i, j, k);
      r[F_DA] = PLANT(1) ? (2. * r[F_DA] + r[F_C]) / 3. : (r[F_DA] + 2. * r[F_C]) / 3.;
    END_LOOP
  } else {
    for (int colour = 0; colour < 2; colour++) {
      if (colour == 1 && !PLANT(3)) boundary_level(g, F_DA, l, 1, 1);
      LOOP_LEVEL(l, n)
        if (((i + j + k) & 1) == (PLANT(2) ? 1 - colour : colour)) {
          double *r = REC(g, l, i, j, k);
          r[F_DA] = relax_value(g, l, r, F_DA, D2);
        }
      END_LOOP
    }
  }
}

/* residual, poisson.h:373-389 (!TREE), finest level only; the ghost fill of
   `a` with its inhomogeneous BC comes first (stencils.h end_stencil ->
   boundary_internal -> boundary_level(list, depth)) */
double orc_residual(orc_grid *g) {
  int l = g->depth, n = 1 << l;
  double D = delta_l(g, l);
  boundary_level(g, F_A, l, 0, 1);
  double maxres = 0.;
#if DIM == 1
#pragma omp parallel for schedule(static) reduction(max : maxres)
  for (int i = GHOSTS; i < n + GHOSTS; i++) {
    const int j = 0, k = 0;
#elif DIM == 2
#pragma omp parallel for schedule(static) reduction(max : maxres)
  for (int i = GHOSTS; i < n + GHOSTS; i++)
    for (int j = GHOSTS; j < n + GHOSTS; j++) {
      const int k = 0;
#else
#pragma omp parallel for schedule(static) reduction(max : maxres)
  for (int i = GHOSTS; i < n + GHOSTS; i++)
    for (int j = GHOSTS; j < n + GHOSTS; j++)
      for (int k = GHOSTS; k < n + GHOSTS; k++) {
#endif
        double *r = REC(g, l, i, j, k);
        double res = PLANT(7) ? r[F_B] + lambda_at(g, r) * r[F_A] : r[F_B] - lambda_at(g, r) * r[F_A];
#if DIM == 2
        if (g->emb.kind) {
          /* the EMBED branch (poisson.h:373-385): embedded face gradients,
             then the embedded flux with the inhomogeneous condition */
          int ci = i - GHOSTS, cj = j - GHOSTS;
          for (int c = 0; c < DIM; c++) {
            ptrdiff_t sd = dstride(l, c);
            res += (alpha_lo(g, r, c) * emb_face_gradient(g, ci, cj, c, 0, F_A) -
                    alpha_hi(g, r, sd, c) * emb_face_gradient(g, ci, cj, c, 1, F_A)) /
                   D;
          }
          double cv, e = embed_flux2(g, l, ci, cj, F_A, g->emb.bc[(size_t)ci * n + cj], g->emb.kind == 1, &cv);
          res += cv - e * r[F_A];
          r[F_RES] = res;
          if (fabs(res) > maxres) maxres = fabs(res);
          continue;
        }
#endif
        for (int c = 0; c < DIM; c++) {
          ptrdiff_t sd = dstride(l, c);
          double a0 = r[F_A], am = r[-sd * NV + F_A], ap = r[sd * NV + F_A];
          res += (alpha_lo(g, r, c) * ((a0 - am) / D) - alpha_hi(g, r, sd, c) * ((ap - a0) / D)) / D;
        }
        r[F_RES] = res;
        if (fabs(res) > maxres) maxres = fabs(res);
      }
  return maxres;
}

/* ---------------------------------------------------------- mg_cycle/solve */

void orc_mg_cycle(orc_grid *g, int nrelax, int minlevel, int maxlevel) {
  orc_restriction_res(g);
  minlevel = minlevel < maxlevel ? minlevel : maxlevel;
  for (int l = minlevel; l <= maxlevel; l++) {
    int n = 1 << l;
    if (l == minlevel) {
      LOOP_LEVEL(l, n)
        REC(g, l, i, j, k)[F_DA] = 0.;
      END_LOOP
    } else {
#if DIM == 2
      /* bilinear_embed is the multigrid grid's (embed.h:738-757); a full
         tree keeps bilinear */
      if (g->emb.kind && !g->tree) {
        LOOP_LEVEL(l, n)
          REC(g, l, i, j, k)[F_DA] = bilinear_embed(g, F_DA, l, i, j);
        END_LOOP
      } else
#endif
      LOOP_LEVEL(l, n)
        REC(g, l, i, j, k)[F_DA] = bilinear(g, F_DA, l, i, j, k);
      END_LOOP
    }
    boundary_level(g, F_DA, l, 1, 1);
    for (int it = 0; it < nrelax; it++) {
      orc_relax(g, l);
      boundary_level(g, F_DA, l, 1, 1);
    }
  }
  int l = g->depth, n = 1 << l;
  LOOP_LEVEL(l, n)
    double *r = REC(g, l, i, j, k);
    r[F_A] += r[F_DA];
  END_LOOP
}

double orc_sum_b(orc_grid *g) {
  int l = g->depth, n = 1 << l;
  double sum = 0.;
  /* serial i-major order (the reference's reduction order with 1 thread) */
#if DIM == 1
  for (int i = GHOSTS; i < n + GHOSTS; i++) { const int j = 0, k = 0;
#elif DIM == 2
  for (int i = GHOSTS; i < n + GHOSTS; i++)
    for (int j = GHOSTS; j < n + GHOSTS; j++) { const int k = 0;
#else
  for (int i = GHOSTS; i < n + GHOSTS; i++)
    for (int j = GHOSTS; j < n + GHOSTS; j++)
      for (int k = GHOSTS; k < n + GHOSTS; k++) {
#endif
        sum += REC(g, l, i, j, k)[F_B];
      }
  return sum;
}

/* the nrelax adaptation of mg_solve after each cycle, poisson.h:198-204:
   while not converged, one more sweep when the cycle reduced the residual by
   less than 1.2 (at most 100), one fewer when by more than 10 (at least 2) */
int orc_nrelax_update(double resb, double resa, double tolerance, int nrelax) {
  double lo = PLANT(5) ? 10 : 1.2, hi = PLANT(5) ? 1.2 : 10;
  if (resa > tolerance) {
    if (resb / resa < lo && nrelax < 100) nrelax++;
    else if (resb / resa > hi && nrelax > 2) nrelax--;
  }
  return nrelax;
}

int orc_mg_solve(orc_grid *g, double tolerance, int nrelax, int minlevel, orc_mgstats *out) {
  orc_mgstats s;
  memset(&s, 0, sizeof s);
  s.sum = orc_sum_b(g);
  s.nrelax = nrelax > 0 ? nrelax : 4;
  double resb;
  resb = s.resb = s.resa = orc_residual(g);
  if (tolerance == 0.) tolerance = g->TOLERANCE;
  for (s.i = 0; s.i < g->NITERMAX && (s.i < g->NITERMIN || s.resa > tolerance); s.i++) {
    orc_mg_cycle(g, s.nrelax, minlevel, g->depth);
    s.resa = orc_residual(g);
    s.nrelax = orc_nrelax_update(resb, s.resa, tolerance, s.nrelax);
    resb = s.resa;
  }
  s.minlevel = minlevel;
  if (s.resa > tolerance)
    fprintf(stderr,
            "WARNING: convergence for a not reached after %d iterations\n"
            "  res: %g sum: %g nrelax: %d\n",
            s.i, s.resa, s.sum, s.nrelax);
  if (out) *out = s;
  return 0;
}

/* poisson(), poisson.h:402-441 (alpha/lambda defaults are the grid's
   constant settings; TOLERANCE is overridden for the call only) */
int orc_poisson(orc_grid *g, double tolerance, int nrelax, int minlevel, orc_mgstats *out) {
  orc_restriction_alpha_lambda(g);
  double defaultol = g->TOLERANCE;
  if (tolerance) g->TOLERANCE = tolerance;
  int r = orc_mg_solve(g, 0., nrelax, minlevel > 1 ? minlevel : 1, out);
  if (tolerance) g->TOLERANCE = defaultol;
  return r;
}

/* project(), poisson.h:469-509: div = sum over dimensions of
   (uf.x[1] - uf.x[]), divided by dt*Delta; poisson() for p = a with
   tolerance TOLERANCE/sq(dt); then on every face (walls included, the
   multigrid foreach_face, grid/multigrid.h:255-279, 353-355)
   uf.x[] -= dt*alpha.x[]*face_gradient_x(p, 0), with
   face_gradient_x(a,i) = (a[i] - a[i-1])/Delta (common.h:1288) and the
   ghosts of p from its (inhomogeneous) boundary conditions.  uf[c] holds
   face component c in the user layout: n+1 faces along c, C order. */
static inline size_t uface(int n, int c, int i, int j, int k) {
  int ext[3] = {1, 1, 1};
  for (int d = 0; d < DIM; d++) ext[d] = n + (d == c);
  return ((size_t)i * ext[1] + j) * ext[2] + k;
}

int orc_project(orc_grid *g, double *const *uf, double dt, int nrelax, orc_mgstats *out) {
  int l = g->depth, n = 1 << l;
  if (dt == 0.) dt = 1.;        /* poisson.h:475 */
  if (nrelax == 0) nrelax = 4;  /* poisson.h:476 */
  double D = delta_l(g, l);
  LOOP_LEVEL(l, n)
    int ci = i - GHOSTS, cj = (DIM >= 2) ? j - GHOSTS : 0, ck = (DIM >= 3) ? k - GHOSTS : 0;
    double div = 0.;
    for (int c = 0; c < DIM; c++) {
      int hi[3] = {ci, cj, ck};
      hi[c] += 1;
      div += uf[c][uface(n, c, hi[0], hi[1], hi[2])] - uf[c][uface(n, c, ci, cj, ck)];
    }
    div /= dt * D;
    REC(g, l, i, j, k)[F_B] = div;
  END_LOOP
  int r = orc_poisson(g, g->TOLERANCE / (dt * dt), nrelax, 0, out);
  /* the stencil read of p below fills its ghosts first */
  boundary_level(g, F_A, l, 0, 1);
  for (int c = 0; c < DIM; c++) {
    int ext[3] = {1, 1, 1};
    for (int d = 0; d < DIM; d++) ext[d] = n + (d == c);
    ptrdiff_t sd = dstride(l, c);
    for (int i = 0; i < ext[0]; i++)
      for (int j = 0; j < ext[1]; j++)
        for (int k = 0; k < ext[2]; k++) {
          const double *rc = REC(g, l, i + GHOSTS, j + (DIM >= 2) * GHOSTS, k + (DIM >= 3) * GHOSTS);
          double alpha = g->alpha_field ? rc[F_AX + c] : g->alpha_val;
          double grad = (rc[F_A] - rc[-sd * NV + F_A]) / D;
          double *u = &uf[c][uface(n, c, i, j, k)];
          *u -= dt * alpha * grad;
        }
  }
  return r;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n);
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
