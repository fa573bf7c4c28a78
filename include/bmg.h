/*
 * bmg.h -- C ABI of libbmg, the MI355X multigrid Poisson-Helmholtz solver.
 *
 * libbmg keeps a uniform multigrid hierarchy resident in GPU memory and runs
 * the reference's solver on it: mg_cycle, mg_solve and poisson() with their
 * relax and residual steps, restriction, prolongation and ghost handling.
 * Every entry point returns an int status (BMG_OK on success); the message
 * of the last failure is available from bmg_last_error().
 *
 * Two families of entry points:
 *
 *  - bmg_grid_* / bmg_set_* / bmg_poisson ... work on SoA arrays in C order
 *    (i slowest, k fastest, the reference's storage order) passed as plain
 *    host or device pointers.  They back the Python package and the tests.
 *
 *  - bmg_poisson_view reads and writes the reference's own storage
 *    (multigrid->d[l], datasize, field index s.i, GHOSTS = 2) so that the
 *    drop-in poisson.h can forward poisson() unchanged.
 *
 * Reference interfaces replaced (paths relative to the reference's src/):
 *   mgstats                        poisson.h:108-114   -> bmg_mgstats
 *   mg_cycle(...)                  poisson.h:32-90     -> bmg_mg_cycle
 *   mg_solve(struct MGSolve)       poisson.h:125-235   -> bmg_mg_solve with poisson()'s own
 *                                  relax/residual pair (poisson.h:271-391); bmg_viscosity for
 *                                  viscosity()'s pair (viscosity.h:176); other callbacks have
 *                                  no device counterpart
 *   relax / residual callbacks     poisson.h:271-391   -> bmg_relax, bmg_residual
 *   poisson(struct Poisson)        poisson.h:258-265, 402-441 -> bmg_poisson, bmg_poisson_view
 *   NITERMAX, NITERMIN, TOLERANCE  poisson.h:102-103   -> bmg_set_niter, bmg_set_tolerance
 *   restriction ({alpha,lambda})   grid/multigrid-common.h:368-404 -> bmg_restriction_alpha_lambda
 *   boundary_level (box, periodic) grid/multigrid.h:543-701 -> implicit in every kernel
 */
#ifndef BMG_H
#define BMG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
enum {
  BMG_OK = 0,
  BMG_ERR_ARG = 1,          /* invalid argument (shape, level, kind ...) */
  BMG_ERR_HIP = 2,          /* a HIP runtime call failed */
  BMG_ERR_UNSUPPORTED = 3,  /* valid request this build does not run */
  BMG_ERR_NOMEM = 4,        /* device allocation failed */
  BMG_ERR_NAN = 5           /* the residual held NaN (stats still filled) */
};

/* mgstats, poisson.h:108-114 (same layout) */
typedef struct {
  int i;              /* number of iterations */
  double resb, resa;  /* maximum residual before and after the iterations */
  double sum;         /* sum of r.h.s. */
  int nrelax;         /* number of relaxations */
  int minlevel;       /* minimum level of the multigrid hierarchy */
} bmg_mgstats;

/* boundary kinds of the unknown `a` on one side of the box */
enum {
  BMG_BC_SYMMETRY = 0,  /* ghost = s              (default, neumann(0)) */
  BMG_BC_DIRICHLET = 1, /* ghost = 2 v - s        common.h:957 */
  BMG_BC_NEUMANN = 2,   /* ghost = Delta g + s    common.h:959 */
  BMG_BC_PERIODIC = 3,  /* periodic in that dimension, grid/multigrid.h:638-695 */
  BMG_BC_AFFINE = 4     /* ghost = c + sigma s, c given per face (see bmg_set_bc_affine) */
};

/* smoothers: the reference's in-place Gauss-Seidel (poisson.h:296) in the
   i-major order of foreach_level (grid/multigrid.h:191-215), run by
   hyperplanes i+j+k = s with the serial loop's iterates (multigrid grids;
   refused for the full-tree kind, whose order is Morton, and by
   bmg_viscosity), its -DJACOBI weighted Jacobi (poisson.h:293-294,
   325-328), and red-black GS (the default: one pass over memory per sweep) */
enum { BMG_RELAX_GS = 0, BMG_RELAX_JACOBI = 1, BMG_RELAX_RBGS = 2 };

/* the a[embed] condition of an embedded boundary (embed.h) */
enum { BMG_EMBED_NONE = 0, BMG_EMBED_DIRICHLET = 1, BMG_EMBED_NEUMANN = 2 };

/* arithmetic of the cycle: plain FP64, or FP32 smoothing / restriction /
   prolongation inside FP64 residual, correction and norm */
enum { BMG_FP64 = 0, BMG_MIXED = 1 };

/* user scalars of a tree */
enum { BMG_USER_SCALARS = 4 };

/* field identifiers */
enum {
  BMG_F_A = 0,       /* unknown (finest level) */
  BMG_F_B = 1,       /* right-hand side (finest level) */
  BMG_F_RES = 2,     /* residual (any level) */
  BMG_F_DA = 3,      /* correction (any level) */
  BMG_F_LAMBDA = 4,  /* Helmholtz coefficient field (any level) */
  BMG_F_ALPHA = 5,   /* face coefficient: BMG_F_ALPHA + c for component c (any level) */
  BMG_F_S = 8        /* trees: user scalar k is BMG_F_S + k, k < BMG_USER_SCALARS (leaf
                        fields that adapt_wavelet may estimate and that follow the mesh) */
};

typedef struct bmg_grid bmg_grid;

/* library */
const char *bmg_version(void);
const char *bmg_last_error(void);
int bmg_device_count(int *count);

/* grid lifetime: a uniform hierarchy of levels 0..depth over the box
   [X0, X0+L0] x ... (init_grid, grid/multigrid.h:723-759) */
int bmg_grid_new(bmg_grid **g, int dimension, int depth, double L0,
                 double X0, double Y0, double Z0);
int bmg_grid_free(bmg_grid *g);
int bmg_grid_info(const bmg_grid *g, int *dimension, int *depth);

/* ghost rule of edge and corner cells: the multigrid grid reflects once per
   direction (grid/multigrid.h:543-594); a full quadtree/octree extrapolates
   from the diagonal cell (grid/tree.h:1168-1258).  Face ghosts and all
   other arithmetic are the same. */
enum { BMG_GRID_MULTIGRID = 0, BMG_GRID_TREE = 1 };

/* solver configuration */
int bmg_set_grid_kind(bmg_grid *g, int kind);
int bmg_set_precision(bmg_grid *g, int precision);
int bmg_set_smoother(bmg_grid *g, int smoother);
int bmg_set_tolerance(bmg_grid *g, double tolerance);           /* TOLERANCE */
int bmg_set_niter(bmg_grid *g, int nitermax, int nitermin);     /* NITERMAX, NITERMIN */
int bmg_set_stream(bmg_grid *g, void *hip_stream);              /* NULL: the grid's own */

/* boundary conditions of `a`; direction 0 right, 1 left, 2 top, 3 bottom,
   4 front, 5 back (common.h:947-951).  `values` (host memory, may be NULL
   for 0) holds the BC expression on each boundary face of the finest level,
   n^(dim-1) values in C order over the tangential indices.  The correction
   field uses the homogeneous version (poisson.h:155-157). */
int bmg_set_bc(bmg_grid *g, int direction, int kind, const double *values);
/* general affine BC: ghost = c + sigma * s, with sigma = +1, -1 or 0 (0:
   the ghost is the expression itself, `a[left] = expr;`, with the
   homogeneous form 0) */
int bmg_set_bc_affine(bmg_grid *g, int direction, int sigma, const double *c);

/* Embedded boundary (embed.h), 2D grids: the EMBED branch of relax and
   residual (poisson.h:277-279, 309-318, 351-352, 380-385) with embed_flux
   (embed.h:654-702), dirichlet_gradient (embed.h:371-457), the multigrid
   prolongation bilinear_embed (embed.h:738-757) and, with `third`, the
   embedded face gradients of the residual (embed.h:61-76, 179-195).
   cs: the finest level's volume fractions (n x n, C order); fs_x: the face
   fractions of the x faces ((n+1) x n, face (i, j) the left face of cell
   (i, j)), fs_y of the y faces (n x (n+1)); fs_z unused (NULL).  Coarser
   levels take restriction_average of cs and face_average of fs.  kind:
   BMG_EMBED_DIRICHLET or BMG_EMBED_NEUMANN, the a[embed] condition, with
   bc its value in every cut cell (0 < cs < 1) at the fragment's centroid
   (the value of the expression, or of the normal gradient; NULL: 0;
   bmg_embed_geometry gives the centroids); BMG_EMBED_NONE removes the
   boundary.  alpha is the caller's (usually fs times a coefficient).  The
   grid must be FP64 (any smoother: GS in the reference's order; a
   red-black pass computes its cells from the state before the pass, since
   a cut cell's stencil reaches cells of its own parity), and every cut
   cell's stencil must stay
   inside the levels the solve relaxes (minlevel and up): otherwise the
   solve returns BMG_ERR_UNSUPPORTED.  Host arrays. */
int bmg_set_embed(bmg_grid *g, const double *cs, const double *fs_x, const double *fs_y, const double *fs_z,
                  int kind, const double *bc, int third);
/* the geometry of the cut cells of an n x n level (embed_geometry,
   embed.h:221-230): per cell the fragment's centroid (px, py, relative to
   the cell centre in cell sizes: embed_area_center adds px Delta to x), its
   length and its unit normal (nx, ny); zero outside cut cells.  Any output
   may be NULL.  Host arrays. */
int bmg_embed_geometry(int n, const double *cs, const double *fs_x, const double *fs_y, double *px, double *py,
                       double *area, double *nx, double *ny);

/* fields: `ptr` is host memory (on_device = 0) or device memory (1).
   Cell fields are n^dim values; face component c is n+1 along c and n
   along the other dimensions.  alpha and lambda are constant (1 and 0, the
   reference's unityf and zeroc) until a field is set. */
int bmg_set_field(bmg_grid *g, int field, int level, const double *ptr, int on_device);
int bmg_get_field(bmg_grid *g, int field, int level, double *ptr, int on_device);
int bmg_set_alpha_const(bmg_grid *g, double alpha);
int bmg_set_lambda_const(bmg_grid *g, double lambda);

/* the solver, on the resident fields */
int bmg_restriction_alpha_lambda(bmg_grid *g);
int bmg_residual(bmg_grid *g, double *maxres);
int bmg_relax(bmg_grid *g, int level);
int bmg_mg_cycle(bmg_grid *g, int nrelax, int minlevel, int maxlevel);
int bmg_mg_solve(bmg_grid *g, double tolerance, int nrelax, int minlevel,
                 bmg_mgstats *stats);
int bmg_poisson(bmg_grid *g, double tolerance, int nrelax, int minlevel,
                bmg_mgstats *stats);
int bmg_synchronize(bmg_grid *g);

/* project(), poisson.h:469-509 (struct Project, poisson.h:460-466): the
   divergence of the face velocity uf, divided by dt*Delta, becomes b; p is
   the grid's `a` (initial guess in, pressure out) with its boundary
   conditions, alpha the grid's face coefficients (unityf unless set);
   poisson() runs with tolerance TOLERANCE/sq(dt); then every face of uf,
   walls included, gets uf -= dt*alpha*(p[] - p[-1])/Delta.  uf_x/y/z are
   the face components in the user layout (n+1 faces along their own
   direction, C order), device pointers when on_device, else host arrays
   (copied in and out).  dt = 0 means 1 and nrelax = 0 means 4, as in the
   reference.  Returns the poisson() status; the face update is done even
   when that status is BMG_ERR_NAN. */
int bmg_project(bmg_grid *g, double *uf_x, double *uf_y, double *uf_z, double dt, int nrelax, int on_device,
                bmg_mgstats *stats);

/* centered_gradient (p, g), navier-stokes/centered.h:387-406, with p the
   grid's `a` (its boundary conditions give the wall ghosts), alpha the
   grid's face coefficients, fm = unityf and SEPS = 0: on every face
   gf = a_f - alpha (p[] - p[-1])/Delta, then per cell
   g = (gf[] + gf[1])/2.  a_x/a_y/a_z: the face acceleration in the face
   layout of bmg_set_field (n+1 along its component), all NULL for zerof;
   g_x/g_y/g_z: the cell outputs (n^dim).  Device pointers when on_device,
   else host arrays. */
int bmg_centered_gradient(bmg_grid *g, const double *a_x, const double *a_y, const double *a_z,
                          double *g_x, double *g_y, double *g_z, int on_device);
/* correction (dt), navier-stokes/centered.h:322-327: u += dt g per cell
   and component (n^dim values each). */
int bmg_correction(bmg_grid *g, double *u_x, double *u_y, double *u_z, const double *g_x,
                   const double *g_y, const double *g_z, double dt, int on_device);

/* viscosity(), viscosity.h:163-178 (struct Viscosity, viscosity.h:3-10):
   the implicit viscous step u - dt/rho div(2 mu D(u)) = u_old, solved by
   mg_solve's loop (poisson.h:139-235: minlevel 0, tolerance TOLERANCE,
   NITERMAX/NITERMIN, nrelax = 0 meaning 4) around residual_viscosity (the
   multigrid grid's "naive" form, viscosity.h:134-157) and relax_viscosity
   (viscosity.h:27-96) run by the grid's smoother (red-black by default,
   the reference's JACOBI variant with BMG_RELAX_JACOBI, its default in-place
   Gauss-Seidel in the i-major order with BMG_RELAX_GS), after restriction
   ({mu, rho}).  u_x/u_y/u_z: the cell components (n^dim, C order), updated
   in place; mu_x/y/z: face fields in the layout of bmg_set_field (n+1 along
   their component), or all NULL for the constant mu_c; rho: a cell field,
   or NULL for the constant rho_c.  Walls at rest: each component's ghost is
   sigma times the mirrored cell, sigma[6*c + side] = +1 or -1 (NULL: the
   reference's default vector BCs, -1 for the normal component and +1 for the
   tangential ones; -1 everywhere is no slip).  Periodic dimensions
   (grid/multigrid.h:600-701) wrap: the ghosts copy the opposite cells, on
   level 0 every ghost is the root value, and face n of mu along a periodic
   dimension is face 0 (the value given for face n is not read).  2D and 3D.
   Device pointers when on_device, else host arrays.  FP64, whatever the
   grid's precision. */
int bmg_viscosity(bmg_grid *g, double *u_x, double *u_y, double *u_z, const double *mu_x, const double *mu_y,
                  const double *mu_z, double mu_c, const double *rho, double rho_c, double dt, const int *sigma,
                  int nrelax, int on_device, bmg_mgstats *stats);
/* viscosity_explicit(), viscosity.h:180-190: one explicit step, res =
   residual_viscosity with the right-hand side u, then u += res; stats get
   resb = max |res| (the rest zero).  Arguments as bmg_viscosity (no
   nrelax); the AXI form follows bmg_set_axi. */
int bmg_viscosity_explicit(bmg_grid *g, double *u_x, double *u_y, double *u_z, const double *mu_x,
                           const double *mu_y, const double *mu_z, double mu_c, const double *rho, double rho_c,
                           double dt, const int *sigma, int on_device, bmg_mgstats *stats);
/* the AXI form of viscosity() (axi.h, viscosity.h:12-14) on a 2D grid: x is
   the axis and y the radius, and the y component's lambda is
   1 + dt/rho (mu.x[] + mu.x[1] + mu.y[] + mu.y[0,1])/2/y^2 with y the cell
   centre, (j + 1/2) Delta + Y0.  The caller passes the metric-weighted
   fields axi.h gives the solvers (mu times fm: the cell's y on x faces, the
   face's y on y faces; rho times cm = y); poisson()'s coefficients carry the
   metric the same way and need no switch.  axi = 0 turns it off. */
int bmg_set_axi(bmg_grid *g, int axi);

/* dump() / restore() (output.h:1043-1390, non-MPI, file version 170901) on
   the multigrid grid: foreach_cell() visits every cell of every level in
   pre-order; the leaves are the finest level; a coarse cell's values are
   the restriction of the finest level (computed on the GPU).  `fields`
   (BMG_F_A, BMG_F_B) are written / read under `names`; restore needs the
   file's depth to equal the grid's, takes the origin and box size from the
   file and sets the finest level (zero for a name the file lacks). */
int bmg_grid_dump(bmg_grid *g, const char *path, const int *fields, const char *const *names, int nfields,
                  double time, int iter);
int bmg_grid_restore(bmg_grid *g, const char *path, const int *fields, const char *const *names, int nfields,
                     double *time, int *iter);

/* Gradient of an objective J(a) with respect to the coefficients (no
   counterpart in the reference: the adjoint behind the torch.autograd
   binding).  L(a) = lambda a + div(alpha grad a) is self-adjoint, so the
   adjoint state psi is one more solve of the same operator with homogeneous
   BCs and right-hand side dJ/da, and dJ/db = psi.  Given psi (finest level,
   n^dim values) and the grid's current `a` and BCs, this writes
     dJ/dalpha_f = (grad psi . grad a)_f on interior faces,
                   -psi_i (ghost(a_i) - a_i)/Delta^2 on wall faces,
     dJ/dlambda  = -psi a,
   in the layout of bmg_set_field (face component c: n+1 along c).  NULL
   outputs are skipped; periodic boundaries are refused. */
int bmg_coef_gradient(bmg_grid *g, const double *psi, double *galpha_x, double *galpha_y,
                      double *galpha_z, double *glambda, int on_device);

/* kernel timing with HIP events on the grid's stream; kind 0 = fine-level
   relax (one sweep per launch), 1 = residual, 2 = restriction,
   3 = prolongation, 4 = correction, 5 = coarse-level relax, 6 = fine-level
   relax of the two-sweep kernel (two sweeps per launch; FP32 levels of
   256^3 cells and more) */
int bmg_profile_enable(bmg_grid *g, int enable);
int bmg_profile_get(bmg_grid *g, int kind, double *total_ms, long *count);

/* ---------------------------------------------------------------- views of
   the reference's multigrid storage (grid/multigrid.h:16-73) */

typedef struct {
  int dimension;          /* 1, 2 or 3 */
  int depth;              /* grid->depth */
  char **d;               /* multigrid->d: level l holds (2^l+4)^dim records */
  size_t datasize;        /* bytes per cell record */
  int nvarmax;            /* _NVARMAX: handles >= nvarmax are constants */
  const double *constant; /* _constant: value of handle h is constant[h - nvarmax] */
  double L0, X0, Y0, Z0;
} bmg_multigrid_view;

/* struct Poisson, poisson.h:258-265, with field handles as ints (s.i) */
typedef struct {
  int a, b;
  int alpha[3];           /* face vector components; 0 or a constant handle allowed */
  int lambda;
  double tolerance;
  int nrelax, minlevel;
  int res;                /* handle of the residual output, or -1 */
} bmg_poisson_args;

/* Boundary of `a` as the drop-in reads it from the BC function pointers:
   per direction, kind (BMG_BC_AFFINE or BMG_BC_PERIODIC), sigma = the
   homogeneous BC applied to s = 1, and c = the inhomogeneous BC applied to
   s = 0 on each boundary face of the finest level (n^(dim-1) values). */
typedef struct {
  int kind[6];
  int sigma[6];
  const double *c[6];
} bmg_view_bc;

/* poisson() on the reference's storage: the finest level's record array
   ((2^depth+4)^dim records of datasize bytes) is uploaded once, the fields
   are gathered from it on the device, the whole solve runs on the GPU, and
   the records come back in one copy with `a` (and `res` when p->res >= 0)
   updated; every other byte comes back as it was.  Not re-entrant: the
   device grid and staging buffers are one per process, kept between
   calls (bmg_view_release frees them). */
int bmg_poisson_view(const bmg_multigrid_view *v, const bmg_poisson_args *p,
                     const bmg_view_bc *bc, double TOLERANCE, int NITERMAX,
                     int NITERMIN, bmg_mgstats *stats);
/* precision (BMG_FP64, BMG_MIXED) and smoother (BMG_RELAX_*) of the
   following bmg_poisson_view calls; the default is BMG_FP64 with red-black
   GS, BMG_RELAX_GS gives the reference's own relax bit for bit (no
   counterpart in the reference, whose struct Poisson has no such field) */
int bmg_view_configure(int precision, int smoother);
int bmg_view_release(void);

/* poisson() with no device allocation (no counterpart in the reference: it
   is for callers that cannot allocate -- code under HIP graph capture,
   tight serving loops, harnesses that forbid hipMalloc).  The same solve as
   bmg_grid_new + bmg_set_* + bmg_poisson + bmg_get_field(BMG_F_A), with the
   whole hierarchy (levels, coefficients, residuals, partial norms) carved
   from one static device scratch of 64 MiB that each GPU holds from the
   library's load until the process ends.  Limit: a problem whose arrays
   need more than that (about 3D 64^3 or 2D 512^2 in FP64) is refused with
   BMG_ERR_ARG before any kernel runs.  The caller passes the inputs and the
   output: a (initial guess in, solution out), b, alpha and lambda are
   device arrays in the layout of bmg_set_field; the BC values bc_c are
   host arrays as for bmg_set_bc_affine.  Every kernel and copy runs on
   `stream` (NULL: the null stream); the call returns after it is done.
   All calls on one GPU share its scratch, which keeps the last call's
   contents: calls that overlap on one GPU (two host threads, two streams)
   race and must be serialised by the caller. */
typedef struct {
  int dimension, depth;
  double L0, X0, Y0, Z0;
  int precision;            /* BMG_FP64 or BMG_MIXED */
  int smoother;             /* BMG_RELAX_* */
  int bc_kind[6];           /* BMG_BC_AFFINE or BMG_BC_PERIODIC per direction */
  int bc_sigma[6];          /* affine: ghost = c + sigma s, sigma = +1 or -1 */
  const double *bc_c[6];    /* affine: n^(dim-1) host values per side, or NULL for 0 */
  double *a;                /* device, n^dim */
  const double *b;          /* device, n^dim */
  const double *alpha[3];   /* device face components (n+1 along c), or all NULL */
  double alpha_c;           /* the constant alpha when alpha[] are NULL (1: unityf) */
  const double *lambda;     /* device, n^dim, or NULL */
  double lambda_c;          /* the constant lambda when NULL (0: zeroc) */
  double TOLERANCE;
  int NITERMAX, NITERMIN;
  void *stream;
} bmg_noalloc_problem;
int bmg_poisson_noalloc(const bmg_noalloc_problem *q, double tolerance, int nrelax,
                        int minlevel, bmg_mgstats *stats);

/* ---------------------------------------------------------------- adaptive
   quadtrees and octrees (grid/tree.h, grid/tree-common.h)

   A bmg_tree starts as the root cell of the box [X0, X0+L0]^dim
   (init_grid(1), grid/tree.h:1488-1584) and is refined on the host with the
   reference's 2:1-balanced refine_cell (grid/tree-common.h:21-67).  Its
   leaves, faces and level lists are the reference's caches in the same
   (pre-order) order (update_cache_f, grid/tree.h:556-684).  Before the first
   solver call after a refinement the tree is compiled into per-level slot
   tables in GPU memory; the solver then runs the reference's tree algorithm
   (tree_boundary_level with halo prolongation, halo_face, the TREE residual
   poisson.h:355-372) on the GPU.  Leaf fields are passed in leaf (cache)
   order.  Sides: 0 right, 1 left, 2 top, 3 bottom, 4 front, 5 back. */

typedef struct bmg_tree bmg_tree;

int bmg_tree_new(bmg_tree **t, int dimension, double L0, double X0, double Y0, double Z0);
int bmg_tree_free(bmg_tree *t);
/* depth, number of leaves and of allocated cells (ghosts and halos included) */
int bmg_tree_info(bmg_tree *t, int *depth, long *nleaves, long *ncells);
/* leaf levels and cell centres (3 doubles per leaf), leaf order */
int bmg_tree_leaves(bmg_tree *t, int *level, double *xyz);
/* one pass of refine() (grid/tree-common.h:308-326): refine every leaf
   whose mark (leaf order) is non-zero, with 2:1 balancing; repeat passes
   until no leaf is marked.  Once the tree holds fields (it has been
   compiled by a field or solver call), each pass starts with boundary (all)
   and the user scalars that are set, a and b follow the mesh
   (refine_bilinear into the new leaves, refine_cell (point, all)); alpha
   and a lambda field must be set again.  Before that there is nothing to
   carry. */
int bmg_tree_refine(bmg_tree *t, const unsigned char *mark, long *nrefined);
/* adapt_wavelet ({s}, {max}, maxlevel, minlevel) (grid/tree-common.h:165-303)
   for the leaf scalar s (BMG_F_S, set first): the children of every refined
   cell whose |value - bilinear prolongation| exceeds max are refined (below
   maxlevel), cells whose children all stay within max/1.5 are coarsened (at
   minlevel and above; minlevel < 1 counts as 1), keeping 2:1 balance.  The
   estimate and the field updates run on the GPU, the passes over the tree
   on the host.  The user scalars that are set, a and b follow the mesh
   (refine_bilinear into new leaves, the children's average into coarsened
   ones); alpha and a lambda field must
   be set again, and boundary values of `a` for levels the tree grew into
   read zero until set again.  nf and nc (may be NULL) receive the numbers of
   refined and coarsened cells (astats). */
int bmg_tree_adapt_wavelet(bmg_tree *t, double max, int maxlevel, int minlevel, int *nf, int *nc);
/* adapt_wavelet (slist, max, maxlevel, minlevel) (grid/tree-common.h:165-303,
   struct Adapt at 157-163) with slist the n user scalars fields[]
   (BMG_F_S + k, each set first) and max[] their tolerances: each scalar in
   turn updates the children's flags, so a cell is refined when any scalar
   asks for it and coarsened when all allow it.  Every user scalar that is
   set, a and b follow the mesh (the reference's default list, all fields);
   the rest as bmg_tree_adapt_wavelet, which is this call with {BMG_F_S}. */
int bmg_tree_adapt_wavelet_list(bmg_tree *t, int n, const int *fields, const double *max, int maxlevel,
                                int minlevel, int *nf, int *nc);
/* the non-leaf cells that foreach_cell() visits, in its pre-order: level
   and centre (3 doubles) each; level/xyz may be NULL to get the count */
int bmg_tree_nonleaf(bmg_tree *t, long *n, int *level, double *xyz);
/* unrefine (cond) (grid/tree-common.h:348-371) with cond given as one mark
   per cell of bmg_tree_nonleaf's list: the marked cells are coarsened from
   the depth up (coarsen_cell, which the 2:1 rule may refuse); the user
   scalars that are set, a and b take the children's average; alpha and a
   lambda field must be set again.
   ncoarsened (may be NULL) receives the number of coarsened cells. */
int bmg_tree_unrefine(bmg_tree *t, const unsigned char *mark, long *ncoarsened);
/* dump() (output.h:1043-1118, the non-MPI branch, file version 170901) of
   the leaf fields `fields` (BMG_F_A, BMG_F_B, BMG_F_S + k) under `names`: the
   header (struct DumpHeader: t, the field count with the subtree "size"
   first, iter, depth, npe = 1, version), the names, {X0, Y0, Z0, L0}, then
   every active cell in foreach_cell() order: unsigned flags (2 for a leaf),
   the subtree size and the field values (non-leaf cells: the restriction
   of their leaves).  Written to `path`~ and renamed, as the reference does. */
int bmg_tree_dump(bmg_tree *t, const char *path, const int *fields, const char *const *names, int nfields,
                  double time, int iter);
/* restore() (output.h:1221-1390) of such a file: the tree is rebuilt from
   init_grid(1) by the same traversal, refining every leaf whose record is
   not a leaf (refine_cell with 2:1 balancing); the origin and box size come
   from the file; `fields` take the leaf values of the columns named `names`
   (zero when the file has no such column, like the reference's reset).
   alpha faces and a lambda field must be set again.  time and iter (may be
   NULL) receive the header's t and i. */
int bmg_tree_restore(bmg_tree *t, const char *path, const int *fields, const char *const *names, int nfields,
                     double *time, int *iter);
/* the cache lists of a level (which: 0 active, 1 prolongation, 2 boundary)
   as (i,j,k) triples with the reference's GHOSTS = 2 offset; ijk may be
   NULL to get the count */
int bmg_tree_cache(bmg_tree *t, int which, int level, int *n, int *ijk);
/* the face cache: (level, i, j, k, flags) per entry, flags bit d = face
   normal to dimension d (foreach_face order); out may be NULL */
int bmg_tree_faces(bmg_tree *t, long *n, int *out);

int bmg_tree_set_precision(bmg_tree *t, int precision);
int bmg_tree_set_smoother(bmg_tree *t, int smoother);
int bmg_tree_set_tolerance(bmg_tree *t, double tolerance);
int bmg_tree_set_niter(bmg_tree *t, int nitermax, int nitermin);
/* boundary condition of `a` on one side (symmetry, Dirichlet or Neumann):
   the expression on every boundary face of every level 0..depth,
   concatenated, n_l^(dim-1) values per level in C order of the tangential
   indices (may be NULL for 0) */
int bmg_tree_set_bc(bmg_tree *t, int side, int kind, const double *values);
/* leaf fields (BMG_F_A, BMG_F_B, BMG_F_S, BMG_F_LAMBDA to set; also
   BMG_F_RES and BMG_F_DA to read), one value per leaf in leaf order */
int bmg_tree_set_leaf_field(bmg_tree *t, int field, const double *values, int on_device);
int bmg_tree_get_leaf_field(bmg_tree *t, int field, double *values, int on_device);
/* alpha on the faces of the face cache: one value per set flag bit, entry
   order then dimension (a foreach_face() assignment) */
int bmg_tree_set_alpha_faces(bmg_tree *t, const double *values, int on_device);
int bmg_tree_set_alpha_const(bmg_tree *t, double alpha);
int bmg_tree_set_lambda_const(bmg_tree *t, double lambda);
/* restriction ({alpha, lambda}) through tree_restriction (halo_face, face
   averages, refine_face_x into the halos) */
int bmg_tree_restriction_alpha_lambda(bmg_tree *t);
int bmg_tree_residual(bmg_tree *t, double *maxres);
int bmg_tree_relax(bmg_tree *t, int level);
int bmg_tree_mg_cycle(bmg_tree *t, int nrelax, int minlevel, int maxlevel);
int bmg_tree_mg_solve(bmg_tree *t, double tolerance, int nrelax, int minlevel,
                      bmg_mgstats *stats);
int bmg_tree_poisson(bmg_tree *t, double tolerance, int nrelax, int minlevel,
                     bmg_mgstats *stats);
/* project(), poisson.h:469-509, on the tree: uf holds the face velocity on
   the face cache (one value per set flag bit, entry then dimension order,
   as bmg_tree_set_alpha_faces) and is updated in place; p is the tree's
   `a` (initial guess in, pressure out) with its boundary conditions.  The
   divergence (per leaf, over dt*Delta, after halo_face of uf) becomes b,
   poisson() runs with TOLERANCE/dt^2, then every face of the cache gets
   uf -= dt alpha (p[] - p[-1])/Delta after the full tree boundary of p.
   dt = 0 means 1, nrelax = 0 means 4.  Device memory when on_device, else
   a host array. */
int bmg_tree_project(bmg_tree *t, double *uf, double dt, int nrelax, int on_device, bmg_mgstats *stats);
/* The adjoint of poisson() on the tree (no counterpart in the reference).
   The leaf operator r(x) = -lambda x - div(alpha grad x) (the TREE
   residual, poisson.h:355-372, with b = 0 and homogeneous boundary
   conditions) is not symmetric on adaptive trees (halo values are bilinear
   prolongations of restrictions), so its adjoint applies the transposes of
   the forward stages in reverse order.
   bmg_tree_residual_transpose: out = r^T y (leaf order).
   bmg_tree_cycle_transpose: out = M^T y with M one V-cycle from da = 0 with
   red-black relax (res -> da on the leaves; minlevel as for mg_cycle).
   bmg_tree_adjoint: psi solves L^T psi = g (L = -r, about the state `a` of
   the last solve) by defect correction with the transposed V-cycle,
   psi += M^T rho, rho = g + r^T psi, until max|rho| <= tolerance (at least
   one, at most nitermax cycles; nrelax <= 0 means 4); then dJ/db = psi,
   galpha (face values, order of bmg_tree_set_alpha_faces) = the adjoint
   flux of psi times (a[] - a[-1])/Delta, glambda (leaves) = -psi a.
   galpha and glambda may be NULL; stats: i, resb = max|g|, resa.  FP64;
   device pointers when on_device, else host arrays. */
int bmg_tree_residual_transpose(bmg_tree *t, const double *y, double *out, int on_device);
int bmg_tree_cycle_transpose(bmg_tree *t, const double *y, double *out, int nrelax, int minlevel,
                             int on_device);
int bmg_tree_adjoint(bmg_tree *t, const double *g, double tolerance, int nitermax, int nrelax,
                     double *psi, double *galpha, double *glambda, int on_device, bmg_mgstats *stats);
int bmg_tree_synchronize(bmg_tree *t);
/* timing of the relax kernel (no counterpart in the reference): while
   enabled, HIP events on the tree's stream bracket every relax launch of
   more than 2048 cells (the cycle then runs without its graph);
   bmg_tree_profile_get returns their total time, the number of launches
   and the cells they relaxed since enabling.  Enabling resets the sums. */
int bmg_tree_profile_enable(bmg_tree *t, int enable);
int bmg_tree_profile_get(bmg_tree *t, double *total_ms, long *launches, long *cells);

#ifdef __cplusplus
}
#endif

#endif /* BMG_H */
