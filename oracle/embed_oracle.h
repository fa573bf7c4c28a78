/* Embedded boundaries (embed.h) in the CPU oracle, 2D: TEST INFRASTRUCTURE
   ONLY, included by mg_oracle.c.  A restatement, in the reference's
   evaluation order, of

     line_alpha                geometry.h:28-50
     line_area                 geometry.h:133-172
     line_length_center        geometry.h:352-404 (plane_area_center in 2D)
     line_center               geometry.h:482-529
     fractions (2D)            fractions.h:129-290: line fractions on the
                               edges from a vertex level set, then the cell
                               fraction from the average normal
     facet_normal (from fs)    fractions.h:414-431
     normalize                 common.h:922-930
     dirichlet_gradient (2D)   embed.h:371-457 (quadratic, degenerate case)
     embed_face_gradient (2D)  embed.h:61-76 (face_condition), used by the
                               residual for faces with 0 < fs < 1 when
                               `third` is set (embed.h:179-195)
     embed_flux                embed.h:654-702
     relax, EMBED branch       poisson.h:277-279, 309-318
     residual, EMBED branch    poisson.h:351-352, 380-385
     bilinear_embed            embed.h:738-757 (multigrid prolongation)

   cs and fs live beside the level records, interior cells and faces only;
   every read of them and of the relaxed field by a cut cell's stencil must
   stay inside the level's interior (and the box ghost layer for the field):
   emb_check() reports a level that breaks this. */

#if DIM == 2

typedef struct {
  double x, y;
} coord2;

#define SEPS 1e-30
#define NODATA 1e30
static inline int isign(double x) { return x > 0 ? 1 : -1; }
static inline double dclamp(double x, double a, double b) { return x < a ? a : x > b ? b : x; }

double orc_line_alpha(double c, double nx, double ny) {
  double alpha, n1, n2;
  n1 = fabs(nx);
  n2 = fabs(ny);
  if (n1 > n2) {
    double t = n1;
    n1 = n2;
    n2 = t;
  }
  c = dclamp(c, 0., 1.);
  double v1 = n1 / 2.;
  if (c <= v1 / n2)
    alpha = sqrt(2. * c * n1 * n2);
  else if (c <= 1. - v1 / n2)
    alpha = c * n2 + v1;
  else
    alpha = n1 + n2 - sqrt(2. * n1 * n2 * (1. - c));
  if (nx < 0.) alpha += nx;
  if (ny < 0.) alpha += ny;
  return alpha - (nx + ny) / 2.;
}

double orc_line_area(double nx, double ny, double alpha) {
  double a, v, area;
  alpha += (nx + ny) / 2.;
  if (nx < 0.) {
    alpha -= nx;
    nx = -nx;
  }
  if (ny < 0.) {
    alpha -= ny;
    ny = -ny;
  }
  if (alpha <= 0.) return 0.;
  if (alpha >= nx + ny) return 1.;
  if (nx < 1e-10)
    area = alpha / ny;
  else if (ny < 1e-10)
    area = alpha / nx;
  else {
    v = alpha * alpha;
    a = alpha - nx;
    if (a > 0.) v -= a * a;
    a = alpha - ny;
    if (a > 0.) v -= a * a;
    area = v / (2. * nx * ny);
  }
  return dclamp(area, 0., 1.);
}

/* centroid of the interface fragment (relative to the cell centre, in
   cell units) and its length */
static double line_length_center(coord2 m, double alpha, coord2 *p) {
  alpha += (m.x + m.y) / 2.;
  coord2 n = m;
  if (n.x < 0.) {
    alpha -= n.x;
    n.x = -n.x;
  }
  if (n.y < 0.) {
    alpha -= n.y;
    n.y = -n.y;
  }
  p->x = p->y = 0.;
  if (alpha <= 0. || alpha >= n.x + n.y) return 0.;
  if (n.x < 1e-4) {
    p->x = 0.;
    p->y = (m.y < 0. ? 1. - alpha : alpha) - 0.5;
    return 1.;
  }
  if (n.y < 1e-4) {
    p->y = 0.;
    p->x = (m.x < 0. ? 1. - alpha : alpha) - 0.5;
    return 1.;
  }
  if (alpha >= n.x) {
    p->x += 1.;
    p->y += (alpha - n.x) / n.y;
  } else
    p->x += alpha / n.x;
  double ax = p->x, ay = p->y;
  if (alpha >= n.y) {
    p->y += 1.;
    ay -= 1.;
    p->x += (alpha - n.y) / n.x;
    ax -= (alpha - n.y) / n.x;
  } else {
    p->y += alpha / n.y;
    ay -= alpha / n.y;
  }
  p->x /= 2.;
  p->x = dclamp(p->x, 0., 1.);
  if (m.x < 0.) p->x = 1. - p->x;
  p->x -= 0.5;
  p->y /= 2.;
  p->y = dclamp(p->y, 0., 1.);
  if (m.y < 0.) p->y = 1. - p->y;
  p->y -= 0.5;
  return sqrt(ax * ax + ay * ay);
}

/* centroid of the fraction a of the cell under the line (m, alpha) */
static void line_center(coord2 m, double alpha, double a, coord2 *p) {
  alpha += (m.x + m.y) / 2.;
  coord2 n = m;
  if (n.x < 0.) {
    alpha -= n.x;
    n.x = -n.x;
  }
  if (n.y < 0.) {
    alpha -= n.y;
    n.y = -n.y;
  }
  if (alpha <= 0.) {
    p->x = p->y = -0.5;
    return;
  }
  if (alpha >= n.x + n.y) {
    p->x = p->y = 0.;
    return;
  }
  if (n.x < 1e-4) {
    p->x = 0.;
    p->y = isign(m.y) * (a / 2. - 0.5);
    return;
  }
  if (n.y < 1e-4) {
    p->y = 0.;
    p->x = isign(m.x) * (a / 2. - 0.5);
    return;
  }
  p->x = p->y = alpha * alpha * alpha;
  {
    double b = alpha - n.x;
    if (b > 0.) {
      p->x -= b * b * (alpha + 2. * n.x);
      p->y -= b * b * b;
    }
  }
  {
    double b = alpha - n.y;
    if (b > 0.) {
      p->y -= b * b * (alpha + 2. * n.y);
      p->x -= b * b * b;
    }
  }
  p->x /= 6. * (n.x * n.x) * n.y * a;
  p->x = isign(m.x) * (p->x - 0.5);
  p->y /= 6. * (n.y * n.y) * n.x * a;
  p->y = isign(m.y) * (p->y - 0.5);
}

/* fractions() from the vertex level set phi ((n+1)^2, vertex (i,j) is the
   lower-left corner of cell (i,j)): face fractions fsx ((n+1) x n: left
   face of cell (i,j) at [i*n+j]), fsy (n x (n+1): bottom face at
   [i*(n+1)+j]) and cell fractions cs (n x n) */
void orc_fractions(int n, const double *phi, double val, double *cs, double *fsx, double *fsy) {
  const int m = n + 1;
#define PHI(i, j) phi[(size_t)(i) * m + (j)]
  /* edges along x (y faces), then along y (x faces) */
  for (int i = 0; i < n; i++)
    for (int j = 0; j <= n; j++) {
      double p0 = PHI(i, j), p1 = PHI(i + 1, j), f;
      if ((p0 - val) * (p1 - val) < 0.) {
        f = (p0 - val) / (p0 - p1);
        if (p0 < val) f = 1. - f;
      } else
        f = (p0 > val || p1 > val);
      fsy[(size_t)i * m + j] = f;
    }
  for (int i = 0; i <= n; i++)
    for (int j = 0; j < n; j++) {
      double p0 = PHI(i, j), p1 = PHI(i, j + 1), f;
      if ((p0 - val) * (p1 - val) < 0.) {
        f = (p0 - val) / (p0 - p1);
        if (p0 < val) f = 1. - f;
      } else
        f = (p0 > val || p1 > val);
      fsx[(size_t)i * n + j] = f;
    }
  /* px = fsy (edges along x), py = fsx (edges along y) */
#define PX(i, j) fsy[(size_t)(i) * m + (j)]
#define PY(i, j) fsx[(size_t)(i) * n + (j)]
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      coord2 nn;
      double nsum = 0.;
      nn.x = PY(i, j) - PY(i + 1, j);
      nsum += fabs(nn.x);
      nn.y = PX(i, j) - PX(i, j + 1);
      nsum += fabs(nn.y);
      double c;
      if (nsum == 0.)
        c = PX(i, j);
      else {
        nn.x /= nsum;
        nn.y /= nsum;
        double alpha = 0., ni = 0.;
        for (int e = 0; e <= 1; e++) {
          double pxv = PX(i, j + e);
          if (pxv > 0. && pxv < 1.) {
            double a = isign(PHI(i, j + e) - val) * (pxv - 0.5);
            alpha += nn.x * a + nn.y * (e - 0.5);
            ni++;
          }
          double pyv = PY(i + e, j);
          if (pyv > 0. && pyv < 1.) {
            double a = isign(PHI(i + e, j) - val) * (pyv - 0.5);
            alpha += nn.y * a + nn.x * (e - 0.5);
            ni++;
          }
        }
        if (ni == 0)
          c = PX(i, j) > PY(i, j) ? PX(i, j) : PY(i, j);
        else if (ni != 4)
          c = orc_line_area(nn.x, nn.y, alpha / ni);
        else
          c = 0.;
      }
      cs[(size_t)i * n + j] = c;
    }
#undef PX
#undef PY
#undef PHI
}

/* ---------------------------------------------------------- level state */

typedef struct {
  int kind;  /* 0: none; 1: Dirichlet; 2: Neumann (the a[embed] condition) */
  int third; /* embed_face_gradient on partial faces in the residual */
  double **cs, **fsx, **fsy; /* per level */
  double *bc;                /* finest level, per cell: the condition's value */
} orc_embed;

static inline double E_CS(const orc_embed *E, int l, int i, int j) {
  return E->cs[l][(size_t)i * (1 << l) + j];
}
static inline double E_FX(const orc_embed *E, int l, int i, int j) {
  return E->fsx[l][(size_t)i * (1 << l) + j];
}
static inline double E_FY(const orc_embed *E, int l, int i, int j) {
  return E->fsy[l][(size_t)i * ((1 << l) + 1) + j];
}

/* facet_normal from the face fractions (|n|_1 = 1), fractions.h:414-431 */
static coord2 facet_normal(const orc_embed *E, int l, int i, int j) {
  coord2 n;
  double nn = 0.;
  n.x = E_FX(E, l, i, j) - E_FX(E, l, i + 1, j);
  nn += fabs(n.x);
  n.y = E_FY(E, l, i, j) - E_FY(E, l, i, j + 1);
  nn += fabs(n.y);
  if (nn > 0.) {
    n.x /= nn;
    n.y /= nn;
  } else {
    n.x = 1. / 2;
    n.y = 1. / 2;
  }
  return n;
}

static void normalize2(coord2 *n) {
  double norm = 0.;
  norm += n->x * n->x;
  norm += n->y * n->y;
  norm = sqrt(norm);
  n->x /= norm;
  n->y /= norm;
}

/* geometry of the fragment in cell (i,j) (interior indices): returns the
   length, fills the centroid p and the Euclidean normal ne */
static double embed_geometry2(const orc_embed *E, int l, int i, int j, coord2 *p, coord2 *ne) {
  coord2 n = facet_normal(E, l, i, j);
  double alpha = orc_line_alpha(E_CS(E, l, i, j), n.x, n.y);
  double area = line_length_center(n, alpha, p);
  *ne = n;
  normalize2(ne);
  return area;
}

/* fragment centroids of every cut cell of level l (in cell units, relative
   to the centre; zero elsewhere), the fragment length and the Euclidean
   normal, as embed_area_center and embed_geometry give them */
void orc_embed_geometry(int n, const double *cs, const double *fsx, const double *fsy, double *px, double *py,
                        double *area, double *nx, double *ny) {
  orc_embed E;
  double *c[1] = {(double *)cs}, *fx[1] = {(double *)fsx}, *fy[1] = {(double *)fsy};
  int l = 0;
  while ((1 << l) < n) l++;
  double *cl[32], *fxl[32], *fyl[32];
  cl[l] = c[0];
  fxl[l] = fx[0];
  fyl[l] = fy[0];
  E.cs = cl;
  E.fsx = fxl;
  E.fsy = fyl;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      size_t q = (size_t)i * n + j;
      px[q] = py[q] = area[q] = nx[q] = ny[q] = 0.;
      double cc = cs[q];
      if (cc > 0. && cc < 1.) {
        coord2 p, ne;
        area[q] = embed_geometry2(&E, l, i, j, &p, &ne);
        px[q] = p.x;
        py[q] = p.y;
        nx[q] = ne.x;
        ny[q] = ne.y;
      }
    }
}

/* the centroid of the fluid part of every cut cell (line_center, as the
   reference's test cases place the right-hand side) */
void orc_cut_centroids(int n, const double *cs, const double *fsx, const double *fsy, double *px, double *py) {
  orc_embed E;
  int l = 0;
  while ((1 << l) < n) l++;
  double *cl[32], *fxl[32], *fyl[32];
  cl[l] = (double *)cs;
  fxl[l] = (double *)fsx;
  fyl[l] = (double *)fsy;
  E.cs = cl;
  E.fsx = fxl;
  E.fsy = fyl;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      size_t q = (size_t)i * n + j;
      px[q] = py[q] = 0.;
      double cc = cs[q];
      if (cc > 0. && cc < 1.) {
        coord2 nn = facet_normal(&E, l, i, j), p;
        double alpha = orc_line_alpha(cc, nn.x, nn.y);
        line_center(nn, alpha, cc, &p);
        px[q] = p.x;
        py[q] = p.y;
      }
    }
}

#endif /* DIM == 2 */
