/* A plain C99 client of include/bmg.h, the way a qcc-compiled Basilisk
   program's drop-in poisson.h uses it: it checks the struct layouts the
   drop-in relies on, the error convention, and (with the argument
   "solve", on a GPU) one poisson() through bmg_poisson_view on storage laid
   out as grid/multigrid.h:16-73 lays it out, verified by a residual
   computed here on the host. */
#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bmg.h"

#define CHECK(cond, code)                                      \
  do {                                                         \
    if (!(cond)) {                                             \
      fprintf(stderr, "check failed: %s (%s)\n", #cond,        \
              bmg_last_error());                               \
      return code;                                             \
    }                                                          \
  } while (0)

enum { NV = 4, FA = 0, FB = 1, FRES = 2, FSPARE = 3 };

static double *cell(double *lv, int n, int i, int j, int h) {
  return lv + ((size_t)(i + 2) * (size_t)(n + 4) + (size_t)(j + 2)) * NV + h;
}

static int solve(void) {
  const int depth = 6, n = 1 << depth;
  const double D = 1. / n, PI = 3.14159265358979323846;
  char *d[7];
  double *lv[7];
  for (int l = 0; l <= depth; l++) {
    size_t m = (size_t)(1 << l) + 4;
    lv[l] = calloc(m * m * NV, sizeof(double));
    CHECK(lv[l] != NULL, 20);
    d[l] = (char *)lv[l];
  }
  for (int i = -2; i < n + 2; i++)
    for (int j = -2; j < n + 2; j++) *cell(lv[depth], n, i, j, FSPARE) = 1000. + i + 0.001 * j;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double x = (i + 0.5) * D, y = (j + 0.5) * D;
      *cell(lv[depth], n, i, j, FB) = -2. * PI * PI * sin(PI * x) * sin(PI * y);
    }
  bmg_multigrid_view v;
  memset(&v, 0, sizeof v);
  v.dimension = 2;
  v.depth = depth;
  v.d = d;
  v.datasize = NV * sizeof(double);
  v.nvarmax = 64;
  v.constant = NULL;
  v.L0 = 1.;
  bmg_poisson_args p;
  memset(&p, 0, sizeof p);
  p.a = FA;
  p.b = FB;
  p.tolerance = 1e-9;
  p.res = FRES;
  bmg_view_bc bc;
  memset(&bc, 0, sizeof bc);
  for (int s = 0; s < 4; s++) {
    bc.kind[s] = BMG_BC_AFFINE;
    bc.sigma[s] = -1; /* dirichlet(0): ghost = -s */
  }
  bmg_mgstats st;
  int rc = bmg_poisson_view(&v, &p, &bc, 1e-3, 100, 1, &st);
  CHECK(rc == BMG_OK, 21);
  CHECK(st.resa <= 1e-9 && st.i > 0, 22);
  /* the residual b - div(grad a), poisson.h:373-389 with alpha = 1 */
  double maxres = 0., maxerr = 0.;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      double a0 = *cell(lv[depth], n, i, j, FA);
      double ap = i < n - 1 ? *cell(lv[depth], n, i + 1, j, FA) : -a0;
      double am = i > 0 ? *cell(lv[depth], n, i - 1, j, FA) : -a0;
      double bp = j < n - 1 ? *cell(lv[depth], n, i, j + 1, FA) : -a0;
      double bm = j > 0 ? *cell(lv[depth], n, i, j - 1, FA) : -a0;
      double r = *cell(lv[depth], n, i, j, FB) - (ap - 2. * a0 + am + bp - 2. * a0 + bm) / (D * D);
      if (fabs(r) > maxres) maxres = fabs(r);
      double x = (i + 0.5) * D, y = (j + 0.5) * D;
      double e = fabs(a0 - sin(PI * x) * sin(PI * y));
      if (e > maxerr) maxerr = e;
      CHECK(fabs(*cell(lv[depth], n, i, j, FRES) - r) <= 1e-9, 23);
    }
  for (int i = -2; i < n + 2; i++)
    for (int j = -2; j < n + 2; j++) CHECK(*cell(lv[depth], n, i, j, FSPARE) == 1000. + i + 0.001 * j, 24);
  CHECK(maxres <= 2e-9, 25);
  CHECK(maxerr < 1e-3, 26);
  printf("solve ok: %d cycles, resa %.3g, host residual %.3g, error %.3g\n", st.i, st.resa, maxres, maxerr);
  for (int l = 0; l <= depth; l++) free(lv[l]);
  CHECK(bmg_view_release() == BMG_OK, 27);
  return 0;
}

int main(int argc, char **argv) {
  /* the layouts of mgstats (poisson.h:108-114) and of the view structs */
  CHECK(sizeof(bmg_mgstats) == 40, 10);
  CHECK(offsetof(bmg_mgstats, resb) == 8 && offsetof(bmg_mgstats, sum) == 24, 11);
  CHECK(offsetof(bmg_mgstats, nrelax) == 32 && offsetof(bmg_mgstats, minlevel) == 36, 12);
  CHECK(offsetof(bmg_poisson_args, tolerance) == 24 && sizeof(bmg_poisson_args) == 48, 13);
  CHECK(offsetof(bmg_multigrid_view, d) == 8 && offsetof(bmg_multigrid_view, datasize) == 16, 14);
  printf("library %s\n", bmg_version());
  /* the error convention: a status, and a message for the last failure */
  CHECK(bmg_view_configure(7, BMG_RELAX_RBGS) == BMG_ERR_ARG, 15);
  CHECK(strlen(bmg_last_error()) > 0, 16);
  CHECK(bmg_view_configure(BMG_FP64, BMG_RELAX_RBGS) == BMG_OK, 17);
  bmg_grid *g = NULL;
  CHECK(bmg_grid_new(&g, 4, 3, 1., 0., 0., 0.) == BMG_ERR_ARG, 18);
  if (argc > 1 && strcmp(argv[1], "solve") == 0) return solve();
  printf("layouts ok\n");
  return 0;
}
