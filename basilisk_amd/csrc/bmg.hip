// libbmg host side: the device-resident multigrid hierarchy and the
// reference's solver loop (poisson.h:32-441) driving the kernels of
// bmg_kernels.h, behind the C ABI of include/bmg.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../../include/bmg.h"
#include "bmg_internal.h"
#include "bmg_embed_kernels.h"
#include "bmg_kernels.h"
#include "bmg_visc_kernels.h"

using namespace bmg;

// the scratch of bmg_poisson_noalloc (see dalloc below)
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
thread_local std::string g_err;
}

int bmg::fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace {

size_t powd(size_t n, int dim) {
  size_t r = 1;
  for (int d = 0; d < dim; d++) r *= n;
  return r;
}

unsigned nblocks(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// one level of the hierarchy: arrays in the cycle precision (double or
// float) plus FP64 copies of the coefficients
struct Level {
  int n = 0, log2n = 0;
  size_t ncell = 0, nwall = 0;
  void* da[2] = {nullptr, nullptr};  // [1]: Jacobi output buffer
  void* res = nullptr;
  void* alo[3] = {nullptr, nullptr, nullptr};
  void* awall[3] = {nullptr, nullptr, nullptr};
  void* lam = nullptr;
  double* alo64[3] = {nullptr, nullptr, nullptr};
  double* awall64[3] = {nullptr, nullptr, nullptr};
  double* lam64 = nullptr;
};

enum {
  PROF_RELAX_FINE = 0,
  PROF_RESIDUAL,
  PROF_RESTRICT,
  PROF_PROLONG,
  PROF_CORRECT,
  PROF_RELAX_COARSE,
  PROF_RELAX_FINE2,
  PROF_N
};

}  // namespace

struct bmg_grid {
  int dim = 0, depth = 0;
  double L0 = 1., X0[3] = {0., 0., 0.};
  int precision = BMG_FP64, smoother = BMG_RELAX_RBGS;
  double TOLERANCE = 1e-3;
  int NITERMAX = 100, NITERMIN = 1;
  std::vector<Level> L;
  double* a = nullptr;
  double* a2 = nullptr;     // second buffer of a for the fused correction (allocated on first use)
  bool pending_correct = false;  // the last cycle left a += da to the next residual
  double* b = nullptr;
  double* res64 = nullptr;  // FP64 residual on the finest level
  int bc_kind[6] = {0, 0, 0, 0, 0, 0};
  int bc_sigma[6] = {1, 1, 1, 1, 1, 1};
  double* bc_c[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int periodic[3] = {0, 0, 0};
  bool alpha_field = false, lambda_field = false;
  double alpha_c = 1., lambda_c = 0.;
  bool coarse_coef_ready = false;  // FP32 copies / restricted levels valid
  bool coarse_res_fresh = false;   // the residual kernel also wrote res[depth-1]
  Part* part = nullptr;            // per-block partial max / NaN count
  hipStream_t own_stream = nullptr, stream = nullptr;
  Stat* stat = nullptr;
  double* partial = nullptr;
  double* scalar = nullptr;
  int npartial = 1024;
  bool allocated_precision[2] = {false, false};
  // profiling
  bool prof = false;
  int rb_override = -1;  // BMG_RB_KERNEL: kernel for red-black levels of >= 4096 cells
  int rb_reverse = 1;    // black passes walk the blocks backwards (BMG_RB_REVERSE=0: forwards)
  int sweep2_min = 256;  // BMG_SWEEP2_MIN: smallest level edge of the two-sweep kernel (0: off)
  int grid_kind = BMG_GRID_MULTIGRID;  // ghost rule of edges and corners
  // the cycle as a HIP graph per (nrelax, minlevel, maxlevel, deferred
  // correction, fresh coarse res); dropped whenever a kernel argument of the
  // cycle could change (BMG_GRAPHS=0: plain launches)
  bool use_graphs = true;
  std::map<std::tuple<int, int, int, int, int>, hipGraphExec_t> cycle_graphs;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
  double prof_ms[PROF_N] = {0};
  long prof_n[PROF_N] = {0};
  // viscosity() workspace, allocated on first use: per level kViscSlots
  // arrays (da, the Jacobi output and res per component, mu low and wall
  // faces per component, rho), and r = u_old on the finest level
  std::vector<double*> visc;
  double* visc_r[3] = {nullptr, nullptr, nullptr};
  // viscosity(): the AXI form (2D; x the axis, y the radius)
  bool axi = false;
  // embedded boundary (bmg_set_embed), null when none
  struct EmbState* emb = nullptr;
};

// the embedded boundary of a 2D grid: per level the cut cells' records on
// the device, and on the finest level the condition's values and the face
// codes of the residual's embedded gradients
struct EmbState {
  int kind = 0, third = 0;
  struct Lev {
    int* cell = nullptr;
    EmbRec* rec = nullptr;
    double* cs = nullptr;
    int nrec = 0;
    bool bad = false;  // a cut cell's stencil leaves the level's interior
  };
  std::vector<Lev> L;
  double* bc = nullptr;
  signed char* fcode[2] = {nullptr, nullptr};
  double* fs[2] = {nullptr, nullptr};
  bool fine_bad = false;  // an embedded face gradient leaves the interior
};

namespace {

size_t esize(const bmg_grid* g) { return g->precision == BMG_MIXED ? sizeof(float) : sizeof(double); }

// --- device memory --------------------------------------------------------
//
// Every device array of a grid comes from dalloc.  Normally that is
// hipMalloc; inside bmg_poisson_noalloc it carves 256-byte aligned pieces
// out of `scratch`, a static device array of kScratchBytes that each GPU
// holds from the library's load until the process ends, and dfree leaves
// such pieces alone.

struct Arena {
  char* base;
  size_t size, used;
};
thread_local Arena* t_arena = nullptr;

// the device address of `scratch` on each device, resolved once per device
// with that device current (the name is only a host-side stand-in)
std::mutex g_scratch_mu;
std::vector<char*> g_scratch_addr;

int scratch_address(char** out) {
  int dev = 0;
  HIPCHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_addr.size() <= dev) g_scratch_addr.resize(dev + 1, nullptr);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(BMG_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch_addr[dev] = (char*)p;
  }
  *out = g_scratch_addr[dev];
  return BMG_OK;
}

bool in_scratch(const void* p) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  for (char* b : g_scratch_addr)
    if (b && (const char*)p >= b && (const char*)p < b + kScratchBytes) return true;
  return false;
}

int dalloc(void** p, size_t bytes) {
  if (bytes == 0) bytes = 8;
  if (Arena* A = t_arena) {
    const size_t need = (bytes + 255) & ~(size_t)255;
    if (need > A->size - A->used)
      return fail(BMG_ERR_ARG, "the problem needs more than the %zu-byte scratch of bmg_poisson_noalloc", A->size);
    *p = A->base + A->used;
    A->used += need;
    return BMG_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return fail(BMG_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return BMG_OK;
}

void dfree(void*& p) {
  if (p && !in_scratch(p)) (void)hipFree(p);
  p = nullptr;
}

template <typename T>
void dfreet(T*& p) {
  if (p && !in_scratch((const void*)p)) (void)hipFree((void*)p);
  p = nullptr;
}

// --- profiling events around a kernel class
struct ProfScope {
  bmg_grid* g;
  int kind;
  hipEvent_t e1 = nullptr;
  ProfScope(bmg_grid* g_, int kind_) : g(g_), kind(kind_) {
    if (!g->prof) return;
    hipEvent_t e0;
    if (g->ev_pool.size() >= 2) {
      e0 = g->ev_pool.back(); g->ev_pool.pop_back();
      e1 = g->ev_pool.back(); g->ev_pool.pop_back();
    } else {
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
    }
    (void)hipEventRecord(e0, g->stream);
    g->ev_used.push_back({kind, {e0, e1}});
  }
  ~ProfScope() {
    if (g->prof) (void)hipEventRecord(e1, g->stream);
  }
};

void prof_collect(bmg_grid* g) {
  if (g->ev_used.empty()) return;
  (void)hipStreamSynchronize(g->stream);
  for (auto& u : g->ev_used) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, u.second.first, u.second.second);
    g->prof_ms[u.first] += ms;
    g->prof_n[u.first] += 1;
    g->ev_pool.push_back(u.second.first);
    g->ev_pool.push_back(u.second.second);
  }
  g->ev_used.clear();
}

int check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(BMG_ERR_HIP, "kernel launch failed: %s", hipGetErrorString(e));
  return BMG_OK;
}

// allocate the per-level arrays of one precision
void drop_graphs(bmg_grid* g) {
  for (auto& kv : g->cycle_graphs)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  g->cycle_graphs.clear();
}

int alloc_levels(bmg_grid* g) {
  drop_graphs(g);
  int p = g->precision;
  if (g->allocated_precision[p]) return BMG_OK;
  // arrays of the other precision are released first
  for (auto& lv : g->L) {
    for (int s = 0; s < 2; s++) dfree(lv.da[s]);
    if (lv.res != (void*)g->res64) dfree(lv.res);
    lv.res = nullptr;
    for (int d = 0; d < 3; d++) {
      if (lv.alo[d] != (void*)lv.alo64[d]) dfree(lv.alo[d]);
      if (lv.awall[d] != (void*)lv.awall64[d]) dfree(lv.awall[d]);
      lv.alo[d] = lv.awall[d] = nullptr;
    }
    if (lv.lam != (void*)lv.lam64) dfree(lv.lam);
    lv.lam = nullptr;
  }
  g->allocated_precision[0] = g->allocated_precision[1] = false;
  size_t es = esize(g);
  for (int l = 0; l <= g->depth; l++) {
    Level& lv = g->L[l];
    TRY(dalloc(&lv.da[0], lv.ncell * es));
    // second buffer for the ping-pong sweeps (Jacobi, streaming red-black)
    TRY(dalloc(&lv.da[1], lv.ncell * es));
    // in FP64 the finest residual is res64 itself
    if (p == BMG_FP64 && l == g->depth) lv.res = g->res64;
    else TRY(dalloc(&lv.res, lv.ncell * es));
    for (int d = 0; d < g->dim; d++) {
      if (p == BMG_FP64) {
        lv.alo[d] = lv.alo64[d];
        lv.awall[d] = lv.awall64[d];
      } else {
        TRY(dalloc(&lv.alo[d], lv.ncell * es));
        TRY(dalloc(&lv.awall[d], lv.nwall * es));
      }
    }
    if (p == BMG_FP64) lv.lam = lv.lam64;
    else TRY(dalloc(&lv.lam, lv.ncell * es));
  }
  g->allocated_precision[p] = true;
  g->coarse_coef_ready = false;
  g->coarse_res_fresh = false;
  return BMG_OK;
}

// the second correction buffer of the ping-pong smoothers (Jacobi and the
// hyperplane Gauss-Seidel)
int ensure_jacobi_buffers(bmg_grid* g) {
  if (g->smoother != BMG_RELAX_JACOBI && g->smoother != BMG_RELAX_GS) return BMG_OK;
  size_t es = esize(g);
  for (auto& lv : g->L)
    if (!lv.da[1]) TRY(dalloc(&lv.da[1], lv.ncell * es));
  return BMG_OK;
}

template <int DIM>
BCH level_bc(const bmg_grid* g, int l) {
  BCH bc;
  bool anyp = false;
  for (int d = 0; d < 3; d++) {
    bc.periodic[d] = d < g->dim ? g->periodic[d] : 0;
    anyp |= bc.periodic[d] != 0;
  }
  for (int b = 0; b < 6; b++) bc.sigma[b] = g->bc_sigma[b];
  bc.tree = g->grid_kind == BMG_GRID_TREE;
  // on level 0 the periodic copy sets every ghost to the root value
  // (grid/multigrid.h:655-664)
  if (l == 0 && anyp)
    for (int b = 0; b < 6; b++) bc.sigma[b] = 1;
  return bc;
}

template <typename T>
Coef<T> level_coef(const bmg_grid* g, int l) {
  const Level& lv = g->L[l];
  Coef<T> k;
  for (int d = 0; d < 3; d++) {
    k.lo[d] = g->alpha_field && d < g->dim ? (const T*)lv.alo[d] : nullptr;
    k.wall[d] = g->alpha_field && d < g->dim ? (const T*)lv.awall[d] : nullptr;
  }
  if (!g->alpha_field) k.lo[0] = nullptr;
  k.alpha = (T)g->alpha_c;
  k.lam = g->lambda_field ? (const T*)lv.lam : nullptr;
  k.lam_c = (T)g->lambda_c;
  return k;
}

Coef<double> fine_coef64(const bmg_grid* g) {
  const Level& lv = g->L[g->depth];
  Coef<double> k;
  for (int d = 0; d < 3; d++) {
    k.lo[d] = g->alpha_field && d < g->dim ? lv.alo64[d] : nullptr;
    k.wall[d] = g->alpha_field && d < g->dim ? lv.awall64[d] : nullptr;
  }
  if (!g->alpha_field) k.lo[0] = nullptr;
  k.alpha = g->alpha_c;
  k.lam = g->lambda_field ? lv.lam64 : nullptr;
  k.lam_c = g->lambda_c;
  return k;
}

double delta(const bmg_grid* g, int l) { return g->L0 * (1. / (double)(1 << l)); }

// --- embedded boundaries, host side (2D) -----------------------------------
//
// The geometry of the cut cells, from the volume and face fractions, in the
// reference's expressions and order: facet_normal from the face fractions
// (fractions.h:414-431), line_alpha (geometry.h:28-50), line_length_center
// (geometry.h:352-404), normalize (common.h:922-930), and which points
// dirichlet_gradient interpolates (embed.h:374-433).

namespace embgeo {

struct V2 {
  double x, y;
};

double line_alpha(double c, V2 n) {
  double alpha, n1 = fabs(n.x), n2 = fabs(n.y);
  if (n1 > n2) std::swap(n1, n2);
  c = c < 0. ? 0. : (c > 1. ? 1. : c);
  const double v1 = n1 / 2.;
  if (c <= v1 / n2)
    alpha = sqrt(2. * c * n1 * n2);
  else if (c <= 1. - v1 / n2)
    alpha = c * n2 + v1;
  else
    alpha = n1 + n2 - sqrt(2. * n1 * n2 * (1. - c));
  if (n.x < 0.) alpha += n.x;
  if (n.y < 0.) alpha += n.y;
  return alpha - (n.x + n.y) / 2.;
}

double clamp01(double v) { return v < 0. ? 0. : (v > 1. ? 1. : v); }

// fragment length; p: its centroid relative to the cell centre (cell units)
double line_length_center(V2 m, double alpha, V2* p) {
  alpha += (m.x + m.y) / 2.;
  V2 n = m;
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
  } else {
    p->x += alpha / n.x;
  }
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
  p->x = clamp01(p->x / 2.);
  if (m.x < 0.) p->x = 1. - p->x;
  p->x -= 0.5;
  p->y = clamp01(p->y / 2.);
  if (m.y < 0.) p->y = 1. - p->y;
  p->y -= 0.5;
  return sqrt(ax * ax + ay * ay);
}

// one level's fractions: cs (n x n), fx ((n+1) x n, left faces), fy
// (n x (n+1), bottom faces), interior only; a read outside marks `bad`
struct Frac {
  int n = 0;
  std::vector<double> cs, fx, fy;
  mutable bool bad = false;
  double c(int i, int j) const {
    if (i < 0 || i >= n || j < 0 || j >= n) return bad = true, 0.;
    return cs[(size_t)i * n + j];
  }
  double x(int i, int j) const {
    if (i < 0 || i > n || j < 0 || j >= n) return bad = true, 0.;
    return fx[(size_t)i * n + j];
  }
  double y(int i, int j) const {
    if (i < 0 || i >= n || j < 0 || j > n) return bad = true, 0.;
    return fy[(size_t)i * (n + 1) + j];
  }
  // along direction d (0: x, 1: y) at offsets (u along d, v across) from
  // cell (i, j): the face normal to d, the face across, the cell
  double fn(int d, int i, int j, int u, int v) const { return d ? y(i + v, j + u) : x(i + u, j + v); }
  double ft(int d, int i, int j, int u, int v) const { return d ? x(i + v, j + u) : y(i + u, j + v); }
  double cr(int d, int i, int j, int u, int v) const { return d ? c(i + v, j + u) : c(i + u, j + v); }
};

V2 facet_normal(const Frac& F, int i, int j) {
  V2 n;
  double nn = 0.;
  n.x = F.x(i, j) - F.x(i + 1, j);
  nn += fabs(n.x);
  n.y = F.y(i, j) - F.y(i, j + 1);
  nn += fabs(n.y);
  if (nn > 0.) {
    n.x /= nn;
    n.y /= nn;
  } else {
    n.x = n.y = 1. / 2;
  }
  return n;
}

void normalize(V2* n) {
  double norm = 0.;
  norm += n->x * n->x;
  norm += n->y * n->y;
  norm = sqrt(norm);
  n->x /= norm;
  n->y /= norm;
}

int isign(double v) { return v > 0 ? 1 : -1; }

// the record of cut cell (i, j) of a level; dirichlet: work out the
// interpolation of dirichlet_gradient too
EmbRec cut_record(const Frac& F, int i, int j, bool dirichlet, V2* centroid, V2* unit_normal) {
  EmbRec R;
  memset(&R, 0, sizeof R);
  const double c = F.c(i, j);
  V2 n = facet_normal(F, i, j), p;
  const double alpha = line_alpha(c, n);
  R.area = line_length_center(n, alpha, &p);
  double fa = 0.;
  fa += F.x(i, j) + F.x(i + 1, j);
  fa += F.y(i, j) + F.y(i, j + 1);
  R.fa = fa;
  normalize(&n);
  if (centroid) *centroid = p;
  if (unit_normal) *unit_normal = n;
  if (!dirichlet) return R;
  const int rot = fabs(n.x) >= fabs(n.y) ? 0 : 1;
  double nx = rot ? n.y : n.x, ny = rot ? n.x : n.y;
  const double px = rot ? p.y : p.x, py = rot ? p.x : p.y;
  nx = -nx;
  ny = -ny;
  bool defined = true;
  if (defined && F.fn(rot, i, j, (nx > 0.), 0) == 0.) defined = false;
  if (defined && F.ft(rot, i, j, 0, (ny > 0.)) == 0.) defined = false;
  int ok = 0;
  double d[2] = {0., 0.};
  if (defined)
    for (int m = 0; m <= 1; m++) {
      const int ii = (m + 1) * isign(nx);
      d[m] = (ii - px) / nx;
      double y1 = py + d[m] * ny;
      const int jj = y1 > 0.5 ? 1 : y1 < -0.5 ? -1 : 0;
      y1 -= jj;
      if (F.fn(rot, i, j, ii + (ii < 0), jj) && F.ft(rot, i, j, ii, jj) && F.ft(rot, i, j, ii, jj + 1) &&
          F.cr(rot, i, j, ii, jj - 1) && F.cr(rot, i, j, ii, jj) && F.cr(rot, i, j, ii, jj + 1)) {
        for (int t = 0; t < 3; t++) {
          const int u = ii, v = jj - 1 + t;  // along the direction, across it
          const int di = rot ? v : u, dj = rot ? u : v;
          R.off[m][t] = di * F.n + dj;
        }
        R.y1[m] = y1;
        ok++;
      } else {
        break;
      }
    }
  R.mode = ok;
  if (ok == 0) {
    const double q = fabs(px / nx);
    d[0] = 1e-3 > q ? 1e-3 : q;
  }
  R.d[0] = d[0];
  R.d[1] = d[1];
  return R;
}

// the code of a partial face of the finest level (embed_face_gradient's
// choice, embed.h:61-76): face (fi, ft) normal to d; 0 plain, else j
signed char face_code(const Frac& F, int d, int fi, int ft) {
  // offsets are taken from the cell (fi, ft) on the face's high side
  const int i = d ? ft : fi, j = d ? fi : ft;
  const double fsf = F.fn(d, i, j, 0, 0);
  if (!(fsf < 1. && fsf > 0.)) return 0;
  const int jj = isign(F.fn(d, i, j, 0, 1) - F.fn(d, i, j, 0, -1));
  if (F.fn(d, i, j, 0, jj) > 0.5 && F.ft(d, i, j, 0, jj + (jj < 0)) && F.ft(d, i, j, -1, jj + (jj < 0)) &&
      F.cr(d, i, j, 0, jj) && F.cr(d, i, j, -1, jj))
    return (signed char)jj;
  return 0;
}

}  // namespace embgeo

// the device view of level l's embedded boundary
EmbLevel emb_level(const bmg_grid* g, int l) {
  EmbLevel E;
  memset(&E, 0, sizeof E);
  const EmbState* S = g->emb;
  E.cell = S->L[l].cell;
  E.rec = S->L[l].rec;
  E.cs = S->L[l].cs;
  E.dirichlet = S->kind == BMG_EMBED_DIRICHLET;
  if (l == g->depth) {
    E.bc = S->bc;
    for (int d = 0; d < 2; d++) {
      E.fcode[d] = S->fcode[d];
      E.fs[d] = S->fs[d];
    }
  }
  return E;
}

void emb_free(bmg_grid* g) {
  EmbState* S = g->emb;
  if (!S) return;
  for (auto& lv : S->L) {
    dfreet(lv.cell);
    dfreet(lv.rec);
    dfreet(lv.cs);
  }
  dfreet(S->bc);
  for (int d = 0; d < 2; d++) {
    dfreet(S->fcode[d]);
    dfreet(S->fs[d]);
  }
  delete S;
  g->emb = nullptr;
}

// an embedded boundary is usable from level minlevel up
int emb_check(const bmg_grid* g, int minlevel) {
  const EmbState* S = g->emb;
  if (!S) return BMG_OK;
  if (g->precision != BMG_FP64) return fail(BMG_ERR_UNSUPPORTED, "embedded boundaries run in FP64");
  for (int l = std::max(minlevel, 0); l <= g->depth; l++)
    if (S->L[l].bad)
      return fail(BMG_ERR_UNSUPPORTED,
                  "the embedded boundary reaches outside level %d's interior (raise minlevel or move it away from "
                  "the box)",
                  l);
  if (S->fine_bad) return fail(BMG_ERR_UNSUPPORTED, "an embedded face gradient reaches outside the finest level");
  return BMG_OK;
}

// --- kernels per dimension / precision --------------------------------

// relax kernel of a level: the two-launch in-place red-black, the
// single-pass streaming red-black (3D, n >= 128) or Jacobi
// relax kernel of a level: two-launch in-place red-black (tiny levels),
// one-launch fused red-black, single-pass streaming red-black (3D,
// n >= 128) or Jacobi
enum { RK_INPLACE_RB, RK_FUSED_RB, RK_STREAM_RB, RK_STREAM2_RB, RK_STREAM3_RB, RK_PAIR_RB, RK_JACOBI, RK_GS, RK_EMBED };

// a cell-count threshold with an environment override (tuning knob)
size_t size_knob(const char* name, size_t dflt) {
  const char* e = getenv(name);
  return e && atoll(e) > 0 ? (size_t)atoll(e) : dflt;
}

// tile rows of the column-streamed sweep: 32 in FP32 (1024 threads), 16 in
// FP64 (512 threads; BMG_STREAM_TJ=16 selects them in FP32 too)
int stream_tj(size_t es) {
  const char* e = getenv("BMG_STREAM_TJ");
  return es == 4 && !(e && atoi(e) == 16) ? kSTJ32 : kSTJ;
}

// plane chunks of the column-streamed kernel (BMG_STREAM_CHUNKS overrides,
// 1..n/4): about two blocks per CU in FP32, one in FP64 (register bound),
// with chunks of at least 4 planes
int stream_chunks(int n, size_t es) {
  if (const char* e = getenv("BMG_STREAM_CHUNKS")) {
    int c = atoi(e);
    if (c >= 1 && c <= n / 4) return c;
  }
  const int tj = stream_tj(es), tiles = (n / kSTK) * (n / tj);
  const int target = tj == kSTJ32 ? 256 : (es == 4 ? 512 : 256);
  const int minplanes = (int)size_knob("BMG_STREAM_MINPLANES", 4);
  return std::max(1, std::min(n / std::max(1, minplanes), target / tiles));
}

// smallest level edge that takes the column-streamed kernel (BMG_STREAM3_MIN
// overrides, 128 or more); at 128^3 its chunks of 4 planes still beat the
// two-launch in-place kernel
int stream3_min() {
  if (const char* e = getenv("BMG_STREAM3_MIN")) {
    int v = atoi(e);
    if (v >= 128) return v;
  }
  return 128;
}

template <int DIM>
int level_kernel(const bmg_grid* g, int l) {
  if (g->emb) return RK_EMBED;
  if (g->smoother == BMG_RELAX_JACOBI) return RK_JACOBI;
  if (g->smoother == BMG_RELAX_GS) return RK_GS;
  // the column-streamed single-pass kernel: 3D levels of 128^3 cells and
  // more (fewer planes give too few blocks), walls in every direction
  const bool walls = !g->periodic[0] && !g->periodic[1] && !g->periodic[2];
  // (buffer resources cover at most 2^31 bytes: up to 512^3 cells)
  const bool stream3 = DIM == 3 && g->L[l].n >= stream3_min() && g->L[l].ncell <= (1u << 27) && walls;
  if (g->rb_override >= 0 && g->L[l].ncell >= 4096 &&
      ((g->rb_override != RK_STREAM_RB && g->rb_override != RK_STREAM2_RB && g->rb_override != RK_STREAM3_RB) ||
       (g->rb_override == RK_STREAM3_RB ? stream3 : DIM == 3 && g->L[l].n >= 128)))
    return g->rb_override;
  // measured at 256^3: the single-pass column-streamed kernel reads each
  // operand once and beats the two-launch in-place kernel, which is the
  // fastest on every other level now that the cycle replays as a graph (the
  // one-launch fused kernel, which recomputes red neighbours, only paid off
  // while launches were issued one by one; BMG_FUSED_MIN brings it back for
  // levels of at least that many cells, below BMG_INPLACE_MIN)
  if (stream3 && g->rb_override < 0) return RK_STREAM3_RB;
  const size_t nc = g->L[l].ncell;
  if (nc >= size_knob("BMG_FUSED_MIN", ~(size_t)0) && nc < size_knob("BMG_INPLACE_MIN", 1u << 20))
    return RK_FUSED_RB;
  return RK_INPLACE_RB;
}

// planes per workgroup of the streaming kernels (BMG_STREAM_CHUNK overrides,
// a power of two dividing n)
int stream_chunk(int n) {
  if (const char* e = getenv("BMG_STREAM_CHUNK")) {
    int c = atoi(e);
    if (c >= 4 && c <= n && (c & (c - 1)) == 0) return c;
  }
  return n >= 256 ? 16 : 8;
}


// largest level that the single-workgroup coarse kernel takes (at most 4096
// cells; BMG_COARSE_TOP overrides within that limit)
int coarse_top(const bmg_grid* g) {
  int top = g->dim == 3 ? 3 : g->dim == 2 ? 5 : 10;
  if (const char* e = getenv("BMG_COARSE_TOP")) {
    int v = atoi(e), lim = g->dim == 3 ? 4 : g->dim == 2 ? 6 : 12;
    if (v >= 0 && v <= lim) top = v;
  }
  return std::min(top, g->depth);
}

// one relax sweep of level l on da[src]; the ping-pong kernels write
// da[src^1], the in-place kernel da[src].  Returns the buffer holding the
// result.
template <int DIM, typename T>
int sweep(bmg_grid* g, int l, int src, int* rc) {
  Level& lv = g->L[l];
  Geo<DIM> geo{lv.n, lv.log2n};
  BCH bc = level_bc<DIM>(g, l);
  Coef<T> k = level_coef<T>(g, l);
  double D = delta(g, l);
  T D2 = (T)(D * D);
  int out = src;
  {
    ProfScope ps(g, l == g->depth ? PROF_RELAX_FINE : PROF_RELAX_COARSE);
    switch (level_kernel<DIM>(g, l)) {
      case RK_EMBED:
        // the EMBED branch (2D, FP64): Jacobi in one pass; the reference's
        // Gauss-Seidel order as hyperplanes in one workgroup; red-black as
        // two passes, each from the state before it, that end in da[src]
        if constexpr (DIM == 2 && sizeof(T) == 8) {
          const EmbLevel E = emb_level(g, l);
          Geo<2> g2{lv.n, lv.log2n};
          const Coef<double> k2 = level_coef<double>(g, l);
          double* d0 = (double*)lv.da[src];
          double* d1 = (double*)lv.da[src ^ 1];
          double* r = (double*)lv.res;
          if (g->smoother == BMG_RELAX_JACOBI) {
            hipLaunchKernelGGL(k_relax_embed, dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream, d0, d1, r, k2,
                               g2, bc, D * D, D, -1, E);
            out = src ^ 1;
          } else if (g->smoother == BMG_RELAX_GS) {
            hipLaunchKernelGGL(k_gs_embed_block, dim3(1), dim3(1024), 0, g->stream, d0, d1, r, k2, g2, bc, D * D, D,
                               E);
            out = src ^ 1;
          } else {
            hipLaunchKernelGGL(k_relax_embed, dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream, d0, d1, r, k2,
                               g2, bc, D * D, D, 0, E);
            hipLaunchKernelGGL(k_relax_embed, dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream, d1, d0, r, k2,
                               g2, bc, D * D, D, 1, E);
          }
        } else {
          *rc = fail(BMG_ERR_UNSUPPORTED, "embedded boundaries are 2D FP64");
        }
        break;
      case RK_JACOBI:
        hipLaunchKernelGGL((k_relax_jacobi<DIM, T>), dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream,
                           (const T*)lv.da[src], (T*)lv.da[src ^ 1], (const T*)lv.res, k, geo, bc, D2);
        out = src ^ 1;
        break;
      case RK_GS: {
        // hyperplanes i+j+k = s in order: one workgroup for the whole sweep
        // while a hyperplane has at most 4096 cells (BMG_GS_BLOCK_MAX
        // overrides), else one launch each
        const T* din = (const T*)lv.da[src];
        T* dout = (T*)lv.da[src ^ 1];
        const int smax = DIM * (lv.n - 1);
        if (powd(lv.n, DIM - 1) <= size_knob("BMG_GS_BLOCK_MAX", 4096)) {
          hipLaunchKernelGGL((k_gs_block<DIM, T>), dim3(1), dim3(1024), 0, g->stream, din, dout,
                             (const T*)lv.res, k, geo, bc, D2);
        } else {
          const size_t rows = DIM == 3 ? (size_t)lv.n * lv.n : (size_t)lv.n;
          for (int s = 0; s <= smax; s++)
            hipLaunchKernelGGL((k_gs_plane<DIM, T>), dim3(nblocks(rows)), dim3(kBlock), 0, g->stream, din, dout,
                               (const T*)lv.res, k, geo, bc, D2, s);
        }
        out = src ^ 1;
        break;
      }
      case RK_FUSED_RB:
        hipLaunchKernelGGL((k_relax_rb_fused<DIM, T>), dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream,
                           (const T*)lv.da[src], (T*)lv.da[src ^ 1], (const T*)lv.res, k, geo, bc, D2);
        out = src ^ 1;
        break;
      case RK_STREAM_RB:
        if constexpr (DIM == 3) {
          int chunk = stream_chunk(lv.n);
          unsigned nb = (unsigned)((lv.n / kTK) * (lv.n / kTJ) * (lv.n / chunk));
          const T* din = (const T*)lv.da[src];
          T* dout = (T*)lv.da[src ^ 1];
          const T* r = (const T*)lv.res;
          bool af = k.lo[0] != nullptr, lf = k.lam != nullptr;
          if (af && lf)
            hipLaunchKernelGGL((k_rb_stream3<T, true, true>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout, r, k,
                               geo, bc, D2, chunk);
          else if (af)
            hipLaunchKernelGGL((k_rb_stream3<T, true, false>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout, r,
                               k, geo, bc, D2, chunk);
          else if (lf)
            hipLaunchKernelGGL((k_rb_stream3<T, false, true>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout, r,
                               k, geo, bc, D2, chunk);
          else
            hipLaunchKernelGGL((k_rb_stream3<T, false, false>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout, r,
                               k, geo, bc, D2, chunk);
        }
        out = src ^ 1;
        break;
      case RK_STREAM2_RB:
        if constexpr (DIM == 3) {
          int chunk = stream_chunk(lv.n);
          unsigned nb = (unsigned)((lv.n / kTK) * (lv.n / kTJ) * (lv.n / chunk));
          const T* din = (const T*)lv.da[src];
          T* dout = (T*)lv.da[src ^ 1];
          const T* r = (const T*)lv.res;
          bool af = k.lo[0] != nullptr, lf = k.lam != nullptr;
          if (af && lf)
            hipLaunchKernelGGL((k_rb_stream3b<T, true, true>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout, r,
                               k, geo, bc, D2, chunk);
          else if (af)
            hipLaunchKernelGGL((k_rb_stream3b<T, true, false>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout, r,
                               k, geo, bc, D2, chunk);
          else if (lf)
            hipLaunchKernelGGL((k_rb_stream3b<T, false, true>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout, r,
                               k, geo, bc, D2, chunk);
          else
            hipLaunchKernelGGL((k_rb_stream3b<T, false, false>), dim3(nb), dim3(kBlock), 0, g->stream, din, dout,
                               r, k, geo, bc, D2, chunk);
        }
        out = src ^ 1;
        break;
      case RK_STREAM3_RB:
        if constexpr (DIM == 3) {
          const int nch = stream_chunks(lv.n, sizeof(T)), tj = stream_tj(sizeof(T));
          if (lv.n % kSTK || g->periodic[0] || g->periodic[1] || g->periodic[2]) {
            *rc = fail(BMG_ERR_ARG, "internal: column-streamed sweep on level %d", l);
            return src;
          }
          const unsigned nb = (unsigned)((lv.n / kSTK) * (lv.n / tj) * nch);
          const T* din = (const T*)lv.da[src];
          T* dout = (T*)lv.da[src ^ 1];
          const T* r = (const T*)lv.res;
          const bool af = k.lo[0] != nullptr, lf = k.lam != nullptr;
          auto go = [&](auto tag) {
            constexpr int TJ = decltype(tag)::value;
            const dim3 blk(TJ * kSTK / 2);
            if (af && lf)
              hipLaunchKernelGGL((k_rb_stream3c<T, true, true, TJ>), dim3(nb), blk, 0, g->stream, din, dout, r, k,
                                 geo, bc, D2, nch);
            else if (af)
              hipLaunchKernelGGL((k_rb_stream3c<T, true, false, TJ>), dim3(nb), blk, 0, g->stream, din, dout, r, k,
                                 geo, bc, D2, nch);
            else if (lf)
              hipLaunchKernelGGL((k_rb_stream3c<T, false, true, TJ>), dim3(nb), blk, 0, g->stream, din, dout, r, k,
                                 geo, bc, D2, nch);
            else
              hipLaunchKernelGGL((k_rb_stream3c<T, false, false, TJ>), dim3(nb), blk, 0, g->stream, din, dout, r, k,
                                 geo, bc, D2, nch);
          };
          if constexpr (sizeof(T) == 4) {
            if (tj == kSTJ32) go(std::integral_constant<int, kSTJ32>());
            else go(std::integral_constant<int, kSTJ>());
          } else {
            go(std::integral_constant<int, kSTJ>());
          }
        }
        out = src ^ 1;
        break;
      case RK_PAIR_RB:
        for (int colour = 0; colour < 2; colour++)
          hipLaunchKernelGGL((k_relax_rb_pair<DIM, T>), dim3(nblocks(lv.ncell / 2)), dim3(kBlock), 0, g->stream,
                             (T*)lv.da[src], (const T*)lv.res, k, geo, bc, D2, colour, colour & g->rb_reverse);
        break;
      default: {
        size_t half = lv.n > 1 ? lv.ncell / 2 : 1;
        for (int colour = 0; colour < 2; colour++)
          hipLaunchKernelGGL((k_relax_rb<DIM, T>), dim3(nblocks(half)), dim3(kBlock), 0, g->stream,
                             (T*)lv.da[src], (const T*)lv.res, k, geo, bc, D2, colour, colour & g->rb_reverse);
      }
    }
  }
  *rc = check_launch();
  return out;
}

// FP32 levels that take the two-sweep kernel: the column-streamed ones of
// at least sweep2_min cells per edge
template <int DIM, typename T>
bool use_sweep2(const bmg_grid* g, int l) {
  return DIM == 3 && sizeof(T) == 4 && g->sweep2_min > 0 && g->L[l].n >= g->sweep2_min &&
         level_kernel<DIM>(g, l) == RK_STREAM3_RB;
}

// plane chunks of the two-sweep kernel (BMG_SWEEP2_CHUNKS overrides,
// 1..n/8): about one block per CU
int sweep2_chunks(int n) {
  if (const char* e = getenv("BMG_SWEEP2_CHUNKS")) {
    const int c = atoi(e);
    if (c >= 1 && c <= n / 8) return c;
  }
  const int tiles = (n / kTTK) * (n / kTTJ);
  return std::max(1, std::min(n / 8, 256 / tiles));
}

// two red-black sweeps of level l in one pass, da[src] -> da[src^1]
template <int DIM, typename T>
int sweep2(bmg_grid* g, int l, int src, int* rc) {
  if constexpr (DIM == 3 && sizeof(T) == 4) {
    Level& lv = g->L[l];
    Geo<DIM> geo{lv.n, lv.log2n};
    const BCH bc = level_bc<DIM>(g, l);
    const Coef<T> k = level_coef<T>(g, l);
    const double D = delta(g, l);
    const T D2 = (T)(D * D);
    if (lv.n % kTTK || lv.ncell > (1u << 27)) {
      *rc = fail(BMG_ERR_ARG, "internal: two-sweep kernel on level %d", l);
      return src;
    }
    const int nch = sweep2_chunks(lv.n);
    const unsigned nb = (unsigned)((lv.n / kTTK) * (lv.n / kTTJ) * nch);
    const T* din = (const T*)lv.da[src];
    T* dout = (T*)lv.da[src ^ 1];
    const T* r = (const T*)lv.res;
    const bool af = k.lo[0] != nullptr, lf = k.lam != nullptr;
    ProfScope ps(g, l == g->depth ? PROF_RELAX_FINE2 : PROF_RELAX_COARSE);
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(nb), dim3(kTBlock), 0, g->stream, din, dout, r, k, geo, bc, D2, nch);
    };
    if (af && lf) go(k_rb_stream3t<T, true, true>);
    else if (af) go(k_rb_stream3t<T, true, false>);
    else if (lf) go(k_rb_stream3t<T, false, true>);
    else go(k_rb_stream3t<T, false, false>);
    *rc = check_launch();
    return src ^ 1;
  } else {
    *rc = fail(BMG_ERR_ARG, "internal: two-sweep kernel is 3D FP32");
    return src;
  }
}

// the public single relax (bmg_relax): result kept in da[0]
template <int DIM, typename T>
int relax_level(bmg_grid* g, int l) {
  TRY(ensure_jacobi_buffers(g));
  int rc = BMG_OK;
  int out = sweep<DIM, T>(g, l, 0, &rc);
  if (out == 1) std::swap(g->L[l].da[0], g->L[l].da[1]);
  return rc;
}

template <int DIM, typename T>
int coarse_block(bmg_grid* g, int nrelax, int minlevel, int top) {
  static_assert(sizeof(CoarseArgs<T, DIM>) <= 4096, "kernel argument block too large");
  if (top >= kCoarseLevels || minlevel < 0 || minlevel > top)
    return fail(BMG_ERR_ARG, "internal: coarse levels %d..%d out of range", minlevel, top);
  CoarseArgs<T, DIM> A;
  memset((void*)&A, 0, sizeof A);
  for (int l = minlevel; l <= top; l++) {
    Level& lv = g->L[l];
    CoarseLevel<T, DIM>& C = A.L[l];
    C.da = (T*)lv.da[0];
    C.da2 = (T*)lv.da[1];
    C.res = (T*)lv.res;
    C.k = level_coef<T>(g, l);
    C.g = Geo<DIM>{lv.n, lv.log2n};
    C.bc = level_bc<DIM>(g, l);
    double D = delta(g, l);
    C.D2 = (T)(D * D);
  }
  A.minlevel = minlevel;
  A.top = top;
  A.nrelax = nrelax;
  A.jacobi = g->smoother == BMG_RELAX_JACOBI;
  A.gs = g->smoother == BMG_RELAX_GS;
  ProfScope ps(g, PROF_RELAX_COARSE);
  // threads of the workgroup (BMG_COARSE_BLOCK: a multiple of 64 up to 1024)
  size_t nt = size_knob("BMG_COARSE_BLOCK", 1024);
  if (nt > 1024 || nt % 64) nt = 1024;
  hipLaunchKernelGGL((k_coarse_cycle<DIM, T>), dim3(1), dim3((unsigned)nt), 0, g->stream, A);
  return check_launch();
}

// the residual kernel that also applies the correction and restricts:
// 3D levels of whole 16 x 64 tiles (BMG_FUSED_RESIDUAL=0 turns it off)
bool fused_residual(const bmg_grid* g) {
  const Level& lf = g->L[g->depth];
  if (g->dim != 3 || lf.n < 128 || lf.n % kSTK || lf.ncell > (1u << 27)) return false;
  const char* e = getenv("BMG_FUSED_RESIDUAL");
  return !(e && e[0] == '0');
}

// defer: leave a += da to the next residual (fused_residual grids only)
template <int DIM, typename T>
int mg_cycle_t(bmg_grid* g, int nrelax, int minlevel, int maxlevel, bool defer) {
  // restriction (res), poisson.h:42; levels below minlevel are never read
  // by the cycle and are not computed
  minlevel = std::min(minlevel, maxlevel);
  if (minlevel < 0) minlevel = 0;
  // (no single-workgroup coarse levels with an embedded boundary)
  int top = g->emb ? -1 : std::min(coarse_top(g), maxlevel);
  bool use_block = top >= minlevel;
  int lowest = use_block ? top : minlevel;
  int lstart = g->coarse_res_fresh ? g->depth - 2 : g->depth - 1;
  g->coarse_res_fresh = false;
  {
    ProfScope ps(g, PROF_RESTRICT);
    for (int l = lstart; l >= lowest; l--) {
      Level& lc = g->L[l];
      Geo<DIM> gc{lc.n, lc.log2n};
      hipLaunchKernelGGL((k_restrict<DIM, T>), dim3(nblocks(lc.ncell)), dim3(kBlock), 0, g->stream,
                         (const T*)g->L[l + 1].res, (T*)lc.res, gc);
    }
  }
  TRY(check_launch());
  // the coarse levels in one workgroup; they end in da[0]
  if (use_block) TRY((coarse_block<DIM, T>(g, nrelax, minlevel, top)));
  for (int l = use_block ? top + 1 : minlevel; l <= maxlevel; l++) {
    Level& lv = g->L[l];
    // ping-pong sweeps start in da[flips & 1] so that they end in da[0]
    // (a two-sweep launch flips the buffers once)
    const int rk = level_kernel<DIM>(g, l);
    const bool two = use_sweep2<DIM, T>(g, l);
    const int flips = two ? nrelax - nrelax / 2 : nrelax;
    const bool inplace = rk == RK_INPLACE_RB || rk == RK_PAIR_RB ||
                         (rk == RK_EMBED && g->smoother == BMG_RELAX_RBGS);
    int dst = inplace ? 0 : (flips & 1);
    if (l == minlevel) {
      hipLaunchKernelGGL((k_fill<T>), dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream, (T*)lv.da[dst],
                         lv.ncell, T(0));
    } else {
      ProfScope ps(g, PROF_PROLONG);
      Geo<DIM> gf{lv.n, lv.log2n};
      BCH bcc = level_bc<DIM>(g, l - 1);
      bool done = false;
      if constexpr (DIM == 2 && sizeof(T) == 8) {
        if (g->emb) {
          // bilinear_embed: injection next to empty coarse cells
          Geo<2> g2{lv.n, lv.log2n};
          hipLaunchKernelGGL(k_prolong_embed, dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream,
                             (const double*)g->L[l - 1].da[0], (double*)lv.da[dst], g2, bcc, g->emb->L[l - 1].cs);
          done = true;
        }
      }
      // one thread per coarse cell on the large 2D/3D levels
      if (done) {
      } else if (DIM >= 2 && lv.ncell >= (1u << 15)) {
        hipLaunchKernelGGL((k_prolong_children<DIM, T>), dim3(nblocks(g->L[l - 1].ncell)), dim3(kBlock), 0,
                           g->stream, (const T*)g->L[l - 1].da[0], (T*)lv.da[dst], gf, bcc);
      } else {
        hipLaunchKernelGGL((k_prolong<DIM, T>), dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream,
                           (const T*)g->L[l - 1].da[0], (T*)lv.da[dst], gf, bcc);
      }
    }
    TRY(check_launch());
    int cur = dst, rc = BMG_OK;
    for (int it = 0; it < nrelax;) {
      if (two && nrelax - it >= 2) {
        cur = sweep2<DIM, T>(g, l, cur, &rc);
        it += 2;
      } else {
        cur = sweep<DIM, T>(g, l, cur, &rc);
        it++;
      }
      TRY(rc);
    }
    if (cur != 0) return fail(BMG_ERR_ARG, "internal: sweep parity");
  }
  Level& lf = g->L[g->depth];
  if (defer && maxlevel == g->depth) {
    g->pending_correct = true;
    return BMG_OK;
  }
  {
    ProfScope ps(g, PROF_CORRECT);
    hipLaunchKernelGGL((k_correct<T>), dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, g->a,
                       (const T*)lf.da[0], lf.ncell);
  }
  return check_launch();
}

// launch k_corr_residual3 for the coefficient kinds
template <typename TO, typename TD>
void corr_residual3(bmg_grid* g, bool af, bool lf, unsigned nb, const TD* da, TO* res, double* r64, TO* coarse,
                    const Coef<double>& k, const Geo<3>& g3, const BCA& bc, double D, int nch) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nb), dim3(kSBlock), 0, g->stream, g->a, da, g->a2, g->b, res, r64, coarse, k, g3,
                       bc, D, nch, g->part);
  };
  if (af && lf) go(k_corr_residual3<TO, TD, true, true>);
  else if (af) go(k_corr_residual3<TO, TD, true, false>);
  else if (lf) go(k_corr_residual3<TO, TD, false, true>);
  else go(k_corr_residual3<TO, TD, false, false>);
}

// residual on the finest level; writes res in the cycle precision (or FP64
// when force64) and returns the max norm on the host
template <int DIM>
int residual_t(bmg_grid* g, double* maxres, bool force64, unsigned long long* nans, bool keep64 = false) {
  Level& lf = g->L[g->depth];
  Geo<DIM> geo{lf.n, lf.log2n};
  BCA bc;
  for (int d = 0; d < 3; d++) bc.periodic[d] = d < g->dim ? g->periodic[d] : 0;
  bool anyp = false;
  for (int d = 0; d < 3; d++) anyp |= bc.periodic[d] != 0;
  for (int b = 0; b < 6; b++) {
    bc.sigma[b] = g->bc_sigma[b];
    bc.c[b] = g->bc_c[b];
    // a one-cell grid with a periodic dimension: the periodic copy gives
    // every ghost the root value, box conditions included
    // (grid/multigrid.h:655-664)
    if (g->depth == 0 && anyp) {
      bc.sigma[b] = 1;
      bc.c[b] = nullptr;
    }
  }
  Coef<double> k = fine_coef64(g);
  double D = delta(g, g->depth);
  // x / D == x * (1/D) exactly when D is a power of two
  int ex;
  bc.pow2 = std::frexp(D, &ex) == 0.5;
  bc.invD = 1. / D;
  bool mixed = g->precision == BMG_MIXED && !force64;
  if constexpr (DIM == 2) {
    if (g->emb) {
      // the EMBED branch, FP64 (the solve refuses the mixed precision); it
      // does not restrict, so the cycle restricts res from the finest level
      ProfScope ps(g, PROF_RESIDUAL);
      g->coarse_res_fresh = false;
      hipLaunchKernelGGL(k_stat_reset, dim3(1), dim3(64), 0, g->stream, g->stat);
      hipLaunchKernelGGL(k_residual_embed, dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, g->a, g->b,
                         g->res64, k, geo, bc, D, emb_level(g, g->depth), g->stat);
      TRY(check_launch());
      Stat st;
      HIPCHECK(hipMemcpyAsync(&st, g->stat, sizeof st, hipMemcpyDeviceToHost, g->stream));
      HIPCHECK(hipStreamSynchronize(g->stream));
      double m;
      memcpy(&m, &st.maxbits, sizeof m);
      *maxres = m;
      if (nans) *nans += st.nans;
      return BMG_OK;
    }
  }
  if constexpr (DIM == 3) {
    if (fused_residual(g)) {
      // the pending correction, the residual and the first restriction in
      // one streaming pass; a + da goes to a2, which then becomes a
      const bool corr = g->pending_correct;
      if (corr && !g->a2) TRY(dalloc((void**)&g->a2, lf.ncell * sizeof(double)));
      // one block per CU (the kernel needs more than 128 registers): chunks
      // of 64 planes at 256^3, of 8 planes at 128^3
      const int nch = lf.n >= 256 ? lf.n / 64 : lf.n / 8;
      const unsigned nb = (unsigned)((lf.n / kSTK) * (lf.n / kSTJ) * nch);
      Geo<3> g3{lf.n, lf.log2n};
      {
        ProfScope ps(g, PROF_RESIDUAL);
        void* coarse = force64 ? nullptr : g->L[g->depth - 1].res;
        const bool af = k.lo[0] != nullptr, lamf = k.lam != nullptr;
        if (mixed) {
          const float* da = corr ? (const float*)lf.da[0] : nullptr;
          double* r64 = keep64 ? g->res64 : nullptr;
          corr_residual3<float, float>(g, af, lamf, nb, da, (float*)lf.res, r64, (float*)coarse, k, g3, bc, D, nch);
        } else {
          // FP64 cycle, or the FP64 residual of a mixed solve (no pending
          // correction then: the cycles' residuals consumed it)
          if (corr && g->precision == BMG_MIXED) return fail(BMG_ERR_ARG, "internal: pending FP32 correction");
          const double* da = corr ? (const double*)lf.da[0] : nullptr;
          corr_residual3<double, double>(g, af, lamf, nb, da, g->res64, nullptr, (double*)coarse, k, g3, bc, D, nch);
        }
        hipLaunchKernelGGL(k_reduce_stat, dim3(1), dim3(1024), 0, g->stream, g->part, (int)nb, g->stat);
      }
      TRY(check_launch());
      if (corr) {
        std::swap(g->a, g->a2);
        g->pending_correct = false;
      }
      if (!force64) g->coarse_res_fresh = true;
      Stat st;
      HIPCHECK(hipMemcpyAsync(&st, g->stat, sizeof st, hipMemcpyDeviceToHost, g->stream));
      HIPCHECK(hipStreamSynchronize(g->stream));
      double m;
      memcpy(&m, &st.maxbits, sizeof m);
      *maxres = m;
      if (nans) *nans += st.nans;
      return BMG_OK;
    }
  }
  // bricks of whole 2^DIM child groups: the fused residual + restriction
  bool fused = (DIM == 3 && lf.n >= 64) || (DIM == 2 && lf.n >= 128);
  if (fused) {
    ProfScope ps(g, PROF_RESIDUAL);
    // at most 4096 blocks (a multiple of 8), each folding several bricks
    unsigned nb = (unsigned)std::min<size_t>(lf.ncell / kBlock, 4096);
    void* coarse = force64 ? nullptr : g->L[g->depth - 1].res;
    if (mixed)
      hipLaunchKernelGGL((k_residual_restrict<DIM, float>), dim3(nb), dim3(kBlock), 0, g->stream, g->a, g->b,
                         (float*)lf.res, (float*)coarse, keep64 ? g->res64 : (double*)nullptr, k, geo, bc, D,
                         g->part);
    else
      hipLaunchKernelGGL((k_residual_restrict<DIM, double>), dim3(nb), dim3(kBlock), 0, g->stream, g->a, g->b,
                         g->res64, (double*)coarse, (double*)nullptr, k, geo, bc, D, g->part);
    hipLaunchKernelGGL(k_reduce_stat, dim3(1), dim3(1024), 0, g->stream, g->part, (int)nb, g->stat);
    if (!force64) g->coarse_res_fresh = true;
  } else {
    hipLaunchKernelGGL(k_stat_reset, dim3(1), dim3(64), 0, g->stream, g->stat);
    ProfScope ps(g, PROF_RESIDUAL);
    if (mixed)
      hipLaunchKernelGGL((k_residual<DIM, float>), dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, g->a,
                         g->b, (float*)lf.res, keep64 ? g->res64 : (double*)nullptr, k, geo, bc, D, g->stat);
    else
      hipLaunchKernelGGL((k_residual<DIM, double>), dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, g->a,
                         g->b, g->res64, (double*)nullptr, k, geo, bc, D, g->stat);
  }
  TRY(check_launch());
  Stat st;
  HIPCHECK(hipMemcpyAsync(&st, g->stat, sizeof st, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  double m;
  memcpy(&m, &st.maxbits, sizeof m);
  *maxres = m;
  if (nans) *nans += st.nans;
  return BMG_OK;
}

int sum_b(bmg_grid* g, double* out) {
  Level& lf = g->L[g->depth];
  int np = (int)std::min<size_t>((size_t)g->npartial, nblocks(lf.ncell));
  hipLaunchKernelGGL(k_sum_partial, dim3(np), dim3(kBlock), 0, g->stream, g->b, lf.ncell, g->partial);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, g->stream, g->partial, np, g->scalar);
  TRY(check_launch());
  HIPCHECK(hipMemcpyAsync(out, g->scalar, sizeof(double), hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

// restriction ({alpha, lambda}), poisson.h:418-420 in FP64, then the FP32
// copies of every level for the mixed cycle
template <int DIM>
int restrict_coefficients(bmg_grid* g) {
  for (int l = g->depth - 1; l >= 0; l--) {
    Level& lc = g->L[l];
    Level& lf = g->L[l + 1];
    Geo<DIM> gc{lc.n, lc.log2n};
    // lambda, the low faces and the high walls of the level in one launch
    CoefLevelArgs A{};
    if (g->lambda_field) {
      A.flam = lf.lam64;
      A.clam = lc.lam64;
    }
    if (g->alpha_field)
      for (int d = 0; d < DIM; d++) {
        A.flo[d] = lf.alo64[d];
        A.clo[d] = lc.alo64[d];
        A.fwall[d] = lf.awall64[d];
        A.cwall[d] = lc.awall64[d];
      }
    if (g->lambda_field || g->alpha_field)
      hipLaunchKernelGGL((k_restrict_coef_level<DIM>), dim3(nblocks(lc.ncell)), dim3(kBlock), 0, g->stream, A, gc);
    if (g->alpha_field) {
      // the periodic copy of face vectors reaches the components whose
      // companion carries the periodic BC (grid/multigrid.h:636-651,
      // grid/cartesian-common.h:903-917): x always, y only in 3D
      // (only the normal component's high wall is ever read).  On level 0
      // the copy sets every ghost of every component to the root's values
      // (grid/multigrid.h:655-664), so all the walls follow.
      bool any_copy = false;
      for (int d = 0; d < DIM; d++) {
        bool copies = DIM == 1 || d == 0 || (DIM == 3 && d == 1);
        any_copy |= g->periodic[d] && copies;
        if (g->periodic[d] && copies && l > 0)
          hipLaunchKernelGGL((k_wall_from_first<DIM>), dim3(nblocks(lc.nwall)), dim3(kBlock), 0, g->stream,
                             lc.alo64[d], lc.awall64[d], gc, d);
      }
      if (l == 0 && any_copy)
        for (int c = 0; c < DIM; c++)
          hipLaunchKernelGGL((k_wall_from_first<DIM>), dim3(1), dim3(kBlock), 0, g->stream, lc.alo64[c],
                             lc.awall64[c], gc, c);
    }
  }
  TRY(check_launch());
  g->coarse_coef_ready = false;
  return BMG_OK;
}

// FP32 copies of the coefficients for the mixed cycle
int refresh_coef_copies(bmg_grid* g) {
  if (g->coarse_coef_ready) return BMG_OK;
  if (g->precision == BMG_MIXED) {
    // every level's arrays, kConvSegs per launch
    ConvBatch B{};
    size_t blocks = 0;
    auto flush = [&]() {
      if (B.nseg) hipLaunchKernelGGL(k_convert_batch, dim3((unsigned)blocks), dim3(kBlock), 0, g->stream, B);
      B = ConvBatch{};
      blocks = 0;
    };
    auto add = [&](const double* src, void* dst, size_t n) {
      if (!n) return;
      const size_t nb = (n + (size_t)kBlock * kConvItems - 1) / ((size_t)kBlock * kConvItems);
      if (B.nseg == kConvSegs || blocks + nb > 0xffffffffu) flush();
      B.src[B.nseg] = src;
      B.dst[B.nseg] = (float*)dst;
      B.n[B.nseg] = (unsigned)n;
      B.first[B.nseg] = (unsigned)blocks;
      blocks += nb;
      B.first[++B.nseg] = (unsigned)blocks;
    };
    for (auto& lv : g->L) {
      if (lv.ncell > 0xffffffffu) return fail(BMG_ERR_ARG, "level of %zu cells is too large", (size_t)lv.ncell);
      for (int d = 0; d < g->dim && g->alpha_field; d++) {
        add(lv.alo64[d], lv.alo[d], lv.ncell);
        add(lv.awall64[d], lv.awall[d], lv.nwall);
      }
      if (g->lambda_field) add(lv.lam64, lv.lam, lv.ncell);
    }
    flush();
    TRY(check_launch());
  }
  g->coarse_coef_ready = true;
  return BMG_OK;
}

// --- dispatch on dimension / precision
template <int DIM, typename T>
int cycle_graph(bmg_grid* g, int nrelax, int minlevel, int maxlevel, bool defer) {
  // profiling records events around the kernels: plain launches then
  if (!g->use_graphs || g->prof) return mg_cycle_t<DIM, T>(g, nrelax, minlevel, maxlevel, defer);
  hipGraphExec_t& slot = g->cycle_graphs[std::make_tuple(nrelax, minlevel, maxlevel, (int)defer,
                                                         (int)g->coarse_res_fresh)];
  if (!slot) {
    // capturing runs the host side of the cycle (its flags) but no kernel
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    int rc = mg_cycle_t<DIM, T>(g, nrelax, minlevel, maxlevel, defer);
    hipError_t e = hipStreamEndCapture(g->stream, &graph);
    if (rc != BMG_OK) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return fail(BMG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(BMG_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    slot = exec;
  } else {
    // the host side of mg_cycle_t
    g->coarse_res_fresh = false;
    if (defer && maxlevel == g->depth) g->pending_correct = true;
  }
  HIPCHECK(hipGraphLaunch(slot, g->stream));
  return BMG_OK;
}

int dispatch_cycle(bmg_grid* g, int nrelax, int minlevel, int maxlevel, bool defer = false) {
  TRY(refresh_coef_copies(g));
  bool mixed = g->precision == BMG_MIXED;
  defer = defer && fused_residual(g);
  switch (g->dim) {
    case 1: return mixed ? mg_cycle_t<1, float>(g, nrelax, minlevel, maxlevel, false) : mg_cycle_t<1, double>(g, nrelax, minlevel, maxlevel, false);
    case 2: return mixed ? cycle_graph<2, float>(g, nrelax, minlevel, maxlevel, false) : cycle_graph<2, double>(g, nrelax, minlevel, maxlevel, false);
    default: return mixed ? cycle_graph<3, float>(g, nrelax, minlevel, maxlevel, defer) : cycle_graph<3, double>(g, nrelax, minlevel, maxlevel, defer);
  }
}

int dispatch_relax(bmg_grid* g, int l) {
  TRY(refresh_coef_copies(g));
  bool mixed = g->precision == BMG_MIXED;
  switch (g->dim) {
    case 1: return mixed ? relax_level<1, float>(g, l) : relax_level<1, double>(g, l);
    case 2: return mixed ? relax_level<2, float>(g, l) : relax_level<2, double>(g, l);
    default: return mixed ? relax_level<3, float>(g, l) : relax_level<3, double>(g, l);
  }
}

int dispatch_residual(bmg_grid* g, double* m, bool force64, unsigned long long* nans, bool keep64 = false) {
  switch (g->dim) {
    case 1: return residual_t<1>(g, m, force64, nans, keep64);
    case 2: return residual_t<2>(g, m, force64, nans, keep64);
    default: return residual_t<3>(g, m, force64, nans, keep64);
  }
}

int dispatch_restrict_coef(bmg_grid* g) {
  switch (g->dim) {
    case 1: return restrict_coefficients<1>(g);
    case 2: return restrict_coefficients<2>(g);
    default: return restrict_coefficients<3>(g);
  }
}

int check_grid(const bmg_grid* g) {
  if (!g) return fail(BMG_ERR_ARG, "null grid");
  return BMG_OK;
}

// copy between a user pointer and a device array
int copy_in(bmg_grid* g, void* dst, const double* src, size_t n, int on_device) {
  HIPCHECK(hipMemcpyAsync(dst, src, n * sizeof(double), on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice,
                          g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

// split a face component array (n+1 along d) into low faces and the wall
int set_face(bmg_grid* g, int d, int l, const double* src, int on_device) {
  Level& lv = g->L[l];
  int n = lv.n;
  size_t outer = 1, inner = 1;  // dims before d, dims after d
  for (int e = 0; e < g->dim; e++) {
    if (e < d) outer *= n;
    if (e > d) inner *= n;
  }
  // src layout: [outer][n+1][inner]; low = [outer][n][inner]; wall = [outer][inner]
  hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHECK(hipMemcpy2DAsync(lv.alo64[d], (size_t)n * inner * sizeof(double), src, (size_t)(n + 1) * inner * sizeof(double),
                            (size_t)n * inner * sizeof(double), outer, kind, g->stream));
  HIPCHECK(hipMemcpy2DAsync(lv.awall64[d], inner * sizeof(double), src + (size_t)n * inner,
                            (size_t)(n + 1) * inner * sizeof(double), inner * sizeof(double), outer, kind, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

int get_face(bmg_grid* g, int d, int l, double* dst, int on_device) {
  Level& lv = g->L[l];
  int n = lv.n;
  size_t outer = 1, inner = 1;
  for (int e = 0; e < g->dim; e++) {
    if (e < d) outer *= n;
    if (e > d) inner *= n;
  }
  hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  HIPCHECK(hipMemcpy2DAsync(dst, (size_t)(n + 1) * inner * sizeof(double), lv.alo64[d], (size_t)n * inner * sizeof(double),
                            (size_t)n * inner * sizeof(double), outer, kind, g->stream));
  HIPCHECK(hipMemcpy2DAsync(dst + (size_t)n * inner, (size_t)(n + 1) * inner * sizeof(double), lv.awall64[d],
                            inner * sizeof(double), inner * sizeof(double), outer, kind, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

int mg_solve_impl(bmg_grid* g, double tolerance, int nrelax, int minlevel, bmg_mgstats* out) {
  TRY(emb_check(g, minlevel));
  bmg_mgstats s;
  memset(&s, 0, sizeof s);
  unsigned long long nans = 0;
  TRY(sum_b(g, &s.sum));
  s.nrelax = nrelax > 0 ? nrelax : 4;
  double resb;
  TRY(dispatch_residual(g, &s.resb, false, &nans));
  resb = s.resa = s.resb;
  if (tolerance == 0.) tolerance = g->TOLERANCE;
  for (s.i = 0; s.i < g->NITERMAX && (s.i < g->NITERMIN || s.resa > tolerance); s.i++) {
    // the cycle's correction of a is applied by the residual kernel
    TRY(dispatch_cycle(g, s.nrelax, minlevel, g->depth, true));
    TRY(dispatch_residual(g, &s.resa, false, &nans));
    // nrelax adaptation, poisson.h:198-204
    if (s.resa > tolerance) {
      if (resb / s.resa < 1.2 && s.nrelax < 100) s.nrelax++;
      else if (resb / s.resa > 10 && s.nrelax > 2) s.nrelax--;
    }
    resb = s.resa;
  }
  s.minlevel = minlevel;
  if (s.resa > tolerance)
    fprintf(stderr,
            "WARNING: convergence for a not reached after %d iterations\n"
            "  res: %g sum: %g nrelax: %d\n",
            s.i, s.resa, s.sum, s.nrelax);
  // in mixed precision the finest-level residual is kept in FP64 too
  if (g->precision == BMG_MIXED) {
    double m;
    TRY(dispatch_residual(g, &m, true, nullptr));
  }
  prof_collect(g);
  if (out) *out = s;
  if (nans) return fail(BMG_ERR_NAN, "the residual held %llu NaN values", nans);
  return BMG_OK;
}

template <int DIM>
int coef_gradient_t(bmg_grid* g, const double* psi, double* const* galpha, double* glambda, int on_device) {
  Level& lf = g->L[g->depth];
  Geo<DIM> geo{lf.n, lf.log2n};
  BCA bc;
  memset(&bc, 0, sizeof bc);
  for (int b = 0; b < 6; b++) {
    bc.sigma[b] = g->bc_sigma[b];
    bc.c[b] = g->bc_c[b];
  }
  double D = delta(g, g->depth);
  size_t nface = (size_t)(lf.n + 1) * powd(lf.n, DIM - 1);
  const double* dpsi = psi;
  double *tmp_psi = nullptr, *tmp_out = nullptr;
  if (!on_device) {
    TRY(dalloc((void**)&tmp_psi, lf.ncell * sizeof(double)));
    HIPCHECK(hipMemcpy(tmp_psi, psi, lf.ncell * sizeof(double), hipMemcpyHostToDevice));
    dpsi = tmp_psi;
    TRY(dalloc((void**)&tmp_out, nface * sizeof(double)));
  }
  int rc = BMG_OK;
  for (int c = 0; c < DIM && rc == BMG_OK; c++) {
    if (!galpha[c]) continue;
    double* out = on_device ? galpha[c] : tmp_out;
    hipLaunchKernelGGL((k_coef_grad<DIM>), dim3(nblocks(nface)), dim3(kBlock), 0, g->stream, g->a, dpsi, out, geo,
                       bc, c, D);
    if (!on_device && hipMemcpyAsync(galpha[c], tmp_out, nface * sizeof(double), hipMemcpyDeviceToHost,
                                     g->stream) != hipSuccess)
      rc = fail(BMG_ERR_HIP, "copy of the alpha gradient failed");
    if (hipStreamSynchronize(g->stream) != hipSuccess) rc = fail(BMG_ERR_HIP, "coefficient gradient failed");
  }
  if (rc == BMG_OK && glambda) {
    double* out = on_device ? glambda : tmp_out;
    if (!on_device) {
      dfreet(tmp_out);
      rc = dalloc((void**)&tmp_out, lf.ncell * sizeof(double));
      out = tmp_out;
    }
    if (rc == BMG_OK) {
      hipLaunchKernelGGL(k_lambda_grad, dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, g->a, dpsi, out,
                         lf.ncell);
      if (!on_device && hipMemcpyAsync(glambda, tmp_out, lf.ncell * sizeof(double), hipMemcpyDeviceToHost,
                                       g->stream) != hipSuccess)
        rc = fail(BMG_ERR_HIP, "copy of the lambda gradient failed");
      if (hipStreamSynchronize(g->stream) != hipSuccess) rc = fail(BMG_ERR_HIP, "lambda gradient failed");
    }
  }
  dfreet(tmp_psi);
  dfreet(tmp_out);
  TRY(rc);
  return check_launch();
}

// project(), poisson.h:469-509 on the finest level: divergence of uf into
// b, poisson() for a with TOLERANCE/sq(dt), face update of uf
template <int DIM>
int project_t(bmg_grid* g, double* const* uf, double dt, int nrelax, bmg_mgstats* stats);

}  // namespace

// ------------------------------------------------------------------- C ABI

extern "C" {

const char* bmg_version(void) { return "bmg 0.1 (gfx950)"; }

const char* bmg_last_error(void) { return g_err.c_str(); }

int bmg_device_count(int* count) {
  HIPCHECK(hipGetDeviceCount(count));
  return BMG_OK;
}

}  // extern "C"

namespace {

// a grid of the given precision on its own stream (borrowed = null) or on
// the caller's stream `borrowed` (which it never destroys)
int grid_create(bmg_grid** out, int dimension, int depth, double L0, double X0, double Y0, double Z0,
                int precision, bool use_borrowed, hipStream_t borrowed) {
  if (!out) return fail(BMG_ERR_ARG, "null output pointer");
  if (dimension < 1 || dimension > 3) return fail(BMG_ERR_ARG, "dimension %d not in 1..3", dimension);
  int maxdepth = dimension == 3 ? 10 : dimension == 2 ? 15 : 28;
  if (depth < 0 || depth > maxdepth) return fail(BMG_ERR_ARG, "depth %d not in 0..%d", depth, maxdepth);
  if (!(L0 > 0.)) return fail(BMG_ERR_ARG, "L0 must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return fail(BMG_ERR_HIP, "no HIP device available: libbmg runs on the GPU only");
  if (precision != BMG_FP64 && precision != BMG_MIXED) return fail(BMG_ERR_ARG, "unknown precision %d", precision);
  bmg_grid* g = new bmg_grid();
  g->dim = dimension;
  g->depth = depth;
  g->precision = precision;
  g->L0 = L0;
  g->X0[0] = X0; g->X0[1] = Y0; g->X0[2] = Z0;
  g->L.resize(depth + 1);
  // tuning knob: the red-black kernel of the larger levels
  if (const char* e = getenv("BMG_RB_KERNEL")) {
    if (!strcmp(e, "inplace")) g->rb_override = RK_INPLACE_RB;
    else if (!strcmp(e, "fused")) g->rb_override = RK_FUSED_RB;
    else if (!strcmp(e, "stream") && dimension == 3) g->rb_override = RK_STREAM_RB;
    else if (!strcmp(e, "stream2") && dimension == 3) g->rb_override = RK_STREAM2_RB;
    else if (!strcmp(e, "pair") && dimension >= 2) g->rb_override = RK_PAIR_RB;
    else if (!strcmp(e, "stream3") && dimension == 3) g->rb_override = RK_STREAM3_RB;
  }
  if (const char* e = getenv("BMG_RB_REVERSE")) g->rb_reverse = e[0] != '0';
  if (const char* e = getenv("BMG_SWEEP2_MIN")) {
    const int v = atoi(e);
    g->sweep2_min = v >= 128 ? v : 0;
  }
  if (const char* e = getenv("BMG_GRAPHS")) g->use_graphs = e[0] != '0';
  int r = BMG_OK;
  // a blocking stream: it orders itself after work on the legacy default
  // stream, where callers (PyTorch included) usually prepare device inputs
  if (use_borrowed) {
    g->stream = borrowed;
  } else {
    if (hipStreamCreateWithFlags(&g->own_stream, hipStreamDefault) != hipSuccess) {
      delete g;
      return fail(BMG_ERR_HIP, "hipStreamCreate failed");
    }
    g->stream = g->own_stream;
  }
  for (int l = 0; l <= depth && r == BMG_OK; l++) {
    Level& lv = g->L[l];
    lv.n = 1 << l;
    lv.log2n = l;
    lv.ncell = powd(lv.n, dimension);
    lv.nwall = powd(lv.n, dimension - 1);
    for (int d = 0; d < dimension && r == BMG_OK; d++) {
      r = dalloc((void**)&lv.alo64[d], lv.ncell * sizeof(double));
      if (r == BMG_OK) r = dalloc((void**)&lv.awall64[d], lv.nwall * sizeof(double));
    }
    if (r == BMG_OK) r = dalloc((void**)&lv.lam64, lv.ncell * sizeof(double));
  }
  size_t nf = g->L[depth].ncell;
  if (r == BMG_OK) r = dalloc((void**)&g->a, nf * sizeof(double));
  if (r == BMG_OK) r = dalloc((void**)&g->b, nf * sizeof(double));
  if (r == BMG_OK) r = dalloc((void**)&g->res64, nf * sizeof(double));
  if (r == BMG_OK) r = dalloc((void**)&g->stat, sizeof(Stat));
  if (r == BMG_OK) r = dalloc((void**)&g->partial, g->npartial * sizeof(double));
  if (r == BMG_OK) r = dalloc((void**)&g->part, (nf / kBlock + 1) * sizeof(Part));
  if (r == BMG_OK) r = dalloc((void**)&g->scalar, sizeof(double));
  if (r == BMG_OK) r = alloc_levels(g);
  if (r == BMG_OK) {
    // every field starts at zero (reset (all, 0.), grid/multigrid.h:758)
    for (auto& lv : g->L) {
      size_t es = esize(g);
      (void)hipMemsetAsync(lv.da[0], 0, lv.ncell * es, g->stream);
      (void)hipMemsetAsync(lv.res, 0, lv.ncell * es, g->stream);
    }
    for (auto& lv : g->L) {
      for (int d = 0; d < dimension; d++) {
        (void)hipMemsetAsync(lv.alo64[d], 0, lv.ncell * sizeof(double), g->stream);
        (void)hipMemsetAsync(lv.awall64[d], 0, lv.nwall * sizeof(double), g->stream);
      }
      (void)hipMemsetAsync(lv.lam64, 0, lv.ncell * sizeof(double), g->stream);
    }
    (void)hipMemsetAsync(g->a, 0, nf * sizeof(double), g->stream);
    (void)hipMemsetAsync(g->b, 0, nf * sizeof(double), g->stream);
    if (hipStreamSynchronize(g->stream) != hipSuccess) r = fail(BMG_ERR_HIP, "initialisation failed");
  }
  if (r != BMG_OK) {
    std::string keep = g_err;
    bmg_grid_free(g);
    g_err = keep;
    return r;
  }
  *out = g;
  return BMG_OK;
}

}  // namespace

extern "C" {

int bmg_grid_new(bmg_grid** out, int dimension, int depth, double L0, double X0, double Y0, double Z0) {
  return grid_create(out, dimension, depth, L0, X0, Y0, Z0, BMG_FP64, false, nullptr);
}

int bmg_grid_free(bmg_grid* g) {
  if (!g) return BMG_OK;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  emb_free(g);
  for (auto& lv : g->L) {
    for (int s = 0; s < 2; s++) dfree(lv.da[s]);
    if (lv.res != (void*)g->res64) dfree(lv.res);
    for (int d = 0; d < 3; d++) {
      if (lv.alo[d] != (void*)lv.alo64[d]) dfree(lv.alo[d]);
      if (lv.awall[d] != (void*)lv.awall64[d]) dfree(lv.awall[d]);
      dfreet(lv.alo64[d]);
      dfreet(lv.awall64[d]);
    }
    if (lv.lam != (void*)lv.lam64) dfree(lv.lam);
    dfreet(lv.lam64);
  }
  drop_graphs(g);
  dfreet(g->a);
  dfreet(g->a2);
  dfreet(g->b);
  dfreet(g->res64);
  for (int b = 0; b < 6; b++) dfreet(g->bc_c[b]);
  dfreet(g->stat);
  dfreet(g->partial);
  dfreet(g->part);
  dfreet(g->scalar);
  for (double*& p : g->visc) dfreet(p);
  for (int c = 0; c < 3; c++) dfreet(g->visc_r[c]);
  prof_collect(g);
  for (auto e : g->ev_pool) (void)hipEventDestroy(e);
  if (g->own_stream) (void)hipStreamDestroy(g->own_stream);
  (void)hipGetLastError();  // teardown is best effort; leave no stale status
  delete g;
  return BMG_OK;
}

int bmg_grid_info(const bmg_grid* g, int* dimension, int* depth) {
  TRY(check_grid(g));
  if (dimension) *dimension = g->dim;
  if (depth) *depth = g->depth;
  return BMG_OK;
}

int bmg_set_precision(bmg_grid* g, int precision) {
  TRY(check_grid(g));
  if (precision != BMG_FP64 && precision != BMG_MIXED) return fail(BMG_ERR_ARG, "unknown precision %d", precision);
  if (precision == g->precision) return BMG_OK;
  // the cycle fields are scratch between cycles; a and b are kept
  g->precision = precision;
  TRY(alloc_levels(g));  // drops the graphs
  return BMG_OK;
}

int bmg_set_grid_kind(bmg_grid* g, int kind) {
  TRY(check_grid(g));
  if (kind != BMG_GRID_MULTIGRID && kind != BMG_GRID_TREE) return fail(BMG_ERR_ARG, "unknown grid kind %d", kind);
  if (kind == BMG_GRID_TREE && (g->periodic[0] || g->periodic[1] || g->periodic[2]))
    return fail(BMG_ERR_UNSUPPORTED, "periodic trees are not supported");
  // a full tree's Gauss-Seidel follows foreach_cell's Morton order
  // (grid/foreach_cell.h:66-123), not the multigrid's i-major one
  if (kind == BMG_GRID_TREE && g->smoother == BMG_RELAX_GS)
    return fail(BMG_ERR_UNSUPPORTED, "Gauss-Seidel in the Morton order of a full tree is not built; use JACOBI or RBGS");
  g->grid_kind = kind;
  drop_graphs(g);
  return BMG_OK;
}

int bmg_set_smoother(bmg_grid* g, int smoother) {
  TRY(check_grid(g));
  if (smoother != BMG_RELAX_GS && smoother != BMG_RELAX_JACOBI && smoother != BMG_RELAX_RBGS)
    return fail(BMG_ERR_ARG, "unknown smoother %d", smoother);
  if (smoother == BMG_RELAX_GS && g->grid_kind == BMG_GRID_TREE)
    return fail(BMG_ERR_UNSUPPORTED, "Gauss-Seidel in the Morton order of a full tree is not built; use JACOBI or RBGS");
  g->smoother = smoother;
  drop_graphs(g);
  return ensure_jacobi_buffers(g);
}

int bmg_set_tolerance(bmg_grid* g, double tolerance) {
  TRY(check_grid(g));
  g->TOLERANCE = tolerance;
  return BMG_OK;
}

int bmg_set_niter(bmg_grid* g, int nitermax, int nitermin) {
  TRY(check_grid(g));
  if (nitermax < 0 || nitermin < 0) return fail(BMG_ERR_ARG, "negative iteration count");
  g->NITERMAX = nitermax;
  g->NITERMIN = nitermin;
  return BMG_OK;
}

int bmg_set_stream(bmg_grid* g, void* stream) {
  TRY(check_grid(g));
  HIPCHECK(hipStreamSynchronize(g->stream));
  g->stream = stream ? (hipStream_t)stream : g->own_stream;
  return BMG_OK;
}

static int set_bc_common(bmg_grid* g, int b, int kind, int sigma, const double* c, double scale) {
  if (b < 0 || b >= 2 * g->dim) return fail(BMG_ERR_ARG, "direction %d out of range", b);
  drop_graphs(g);
  int d = b / 2;
  dfreet(g->bc_c[b]);
  if (kind == BMG_BC_PERIODIC) {
    g->periodic[d] = 1;
    g->bc_kind[2 * d] = g->bc_kind[2 * d + 1] = BMG_BC_PERIODIC;
    dfreet(g->bc_c[b ^ 1]);
    return BMG_OK;
  }
  if (g->periodic[d]) {
    g->periodic[d] = 0;
    g->bc_kind[b ^ 1] = BMG_BC_SYMMETRY;
    g->bc_sigma[b ^ 1] = 1;
  }
  g->bc_kind[b] = kind;
  g->bc_sigma[b] = sigma;
  if (c) {
    size_t nw = g->L[g->depth].nwall;
    std::vector<double> h(nw);
    for (size_t i = 0; i < nw; i++) h[i] = scale == 1. ? c[i] : scale * c[i];
    TRY(dalloc((void**)&g->bc_c[b], nw * sizeof(double)));
    HIPCHECK(hipMemcpy(g->bc_c[b], h.data(), nw * sizeof(double), hipMemcpyHostToDevice));
  }
  return BMG_OK;
}

int bmg_set_bc(bmg_grid* g, int b, int kind, const double* values) {
  TRY(check_grid(g));
  switch (kind) {
    case BMG_BC_SYMMETRY: return set_bc_common(g, b, kind, 1, nullptr, 1.);
    // dirichlet: 2.*(expr) - s  ->  c = 2 v, sigma = -1 (exact: 2v is exact and
    // c + (-1)*s rounds like c - s)
    case BMG_BC_DIRICHLET: return set_bc_common(g, b, kind, -1, values, 2.);
    // neumann: Delta*(expr) + s  ->  c = Delta g, sigma = +1
    case BMG_BC_NEUMANN: return set_bc_common(g, b, kind, 1, values, delta(g, g->depth));
    case BMG_BC_PERIODIC: return set_bc_common(g, b, kind, 1, nullptr, 1.);
    default: return fail(BMG_ERR_ARG, "unknown BC kind %d", kind);
  }
}

int bmg_set_bc_affine(bmg_grid* g, int b, int sigma, const double* c) {
  TRY(check_grid(g));
  if (sigma != 1 && sigma != -1 && sigma != 0) return fail(BMG_ERR_ARG, "sigma must be +1, -1 or 0");
  return set_bc_common(g, b, BMG_BC_AFFINE, sigma, c, 1.);
}

int bmg_set_embed(bmg_grid* g, const double* cs, const double* fs_x, const double* fs_y, const double* fs_z,
                  int kind, const double* bc, int third) {
  TRY(check_grid(g));
  (void)fs_z;
  if (kind == 0) {
    drop_graphs(g);
    emb_free(g);
    return BMG_OK;
  }
  if (g->dim != 2) return fail(BMG_ERR_UNSUPPORTED, "embedded boundaries are built for 2D grids");
  if (kind != BMG_EMBED_DIRICHLET && kind != BMG_EMBED_NEUMANN) return fail(BMG_ERR_ARG, "unknown embed kind %d", kind);
  if (!cs || !fs_x || !fs_y) return fail(BMG_ERR_ARG, "cs and fs are required");
  if (g->periodic[0] || g->periodic[1]) return fail(BMG_ERR_UNSUPPORTED, "embedded boundaries on periodic grids");
  HIPCHECK(hipStreamSynchronize(g->stream));
  drop_graphs(g);
  emb_free(g);
  EmbState* S = new EmbState();
  g->emb = S;
  S->kind = kind;
  S->third = third ? 1 : 0;
  const int D = g->depth;
  // the fractions of every level: restriction_average for cs, face_average
  // for fs (multigrid-common.h:26-63), children in foreach_child order
  std::vector<embgeo::Frac> F(D + 1);
  {
    embgeo::Frac& f = F[D];
    f.n = 1 << D;
    const size_t n = f.n;
    f.cs.assign(cs, cs + n * n);
    f.fx.assign(fs_x, fs_x + (n + 1) * n);
    f.fy.assign(fs_y, fs_y + n * (n + 1));
  }
  for (int l = D - 1; l >= 0; l--) {
    embgeo::Frac& f = F[l];
    const embgeo::Frac& h = F[l + 1];
    f.n = 1 << l;
    const int n = f.n, m = h.n;
    f.cs.resize((size_t)n * n);
    f.fx.resize((size_t)(n + 1) * n);
    f.fy.resize((size_t)n * (n + 1));
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        double sum = 0.;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++) sum += h.cs[(size_t)(2 * i + a) * m + 2 * j + b];
        f.cs[(size_t)i * n + j] = sum / 4;
      }
    for (int i = 0; i <= n; i++)
      for (int j = 0; j < n; j++) {
        double sum = 0.;
        for (int t = 0; t < 2; t++) sum += h.fx[(size_t)(2 * i) * m + 2 * j + t];
        f.fx[(size_t)i * n + j] = sum / 2.;
      }
    for (int i = 0; i < n; i++)
      for (int j = 0; j <= n; j++) {
        double sum = 0.;
        for (int t = 0; t < 2; t++) sum += h.fy[(size_t)(2 * i + t) * (m + 1) + 2 * j];
        f.fy[(size_t)i * (n + 1) + j] = sum / 2.;
      }
  }
  S->L.resize(D + 1);
  const bool dirichlet = kind == BMG_EMBED_DIRICHLET;
  int rc = BMG_OK;
  for (int l = 0; l <= D && rc == BMG_OK; l++) {
    const embgeo::Frac& f = F[l];
    const int n = f.n;
    std::vector<int> cell((size_t)n * n, -1);
    std::vector<EmbRec> rec;
    f.bad = false;
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        const double c = f.cs[(size_t)i * n + j];
        if (c > 0. && c < 1.) {
          cell[(size_t)i * n + j] = (int)rec.size();
          rec.push_back(embgeo::cut_record(f, i, j, dirichlet, nullptr, nullptr));
        }
      }
    EmbState::Lev& lv = S->L[l];
    lv.bad = f.bad;
    lv.nrec = (int)rec.size();
    rc = dalloc((void**)&lv.cell, cell.size() * sizeof(int));
    if (rc == BMG_OK) rc = dalloc((void**)&lv.rec, std::max<size_t>(rec.size(), 1) * sizeof(EmbRec));
    if (rc == BMG_OK) rc = dalloc((void**)&lv.cs, f.cs.size() * sizeof(double));
    if (rc == BMG_OK && (hipMemcpy(lv.cell, cell.data(), cell.size() * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
                         (rec.size() && hipMemcpy(lv.rec, rec.data(), rec.size() * sizeof(EmbRec),
                                                  hipMemcpyHostToDevice) != hipSuccess) ||
                         hipMemcpy(lv.cs, f.cs.data(), f.cs.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess))
      rc = fail(BMG_ERR_HIP, "upload of the embedded boundary failed");
  }
  // the finest level: the condition's values and the partial faces' codes
  const embgeo::Frac& f = F[D];
  const size_t n = f.n;
  std::vector<signed char> code[2] = {std::vector<signed char>((n + 1) * n, 0),
                                      std::vector<signed char>(n * (n + 1), 0)};
  f.bad = false;
  if (S->third)
    for (int d = 0; d < 2; d++)
      for (size_t fi = 0; fi <= n; fi++)
        for (size_t ft = 0; ft < n; ft++) {
          const size_t q = d == 0 ? fi * n + ft : ft * (n + 1) + fi;
          code[d][q] = embgeo::face_code(f, d, (int)fi, (int)ft);
        }
  S->fine_bad = f.bad;
  std::vector<double> bcv(n * n, 0.);
  if (bc) bcv.assign(bc, bc + n * n);
  if (rc == BMG_OK) rc = dalloc((void**)&S->bc, n * n * sizeof(double));
  for (int d = 0; d < 2 && rc == BMG_OK; d++) {
    rc = dalloc((void**)&S->fcode[d], code[d].size());
    if (rc == BMG_OK) rc = dalloc((void**)&S->fs[d], code[d].size() * sizeof(double));
    const std::vector<double>& fsv = d == 0 ? f.fx : f.fy;
    if (rc == BMG_OK && (hipMemcpy(S->fcode[d], code[d].data(), code[d].size(), hipMemcpyHostToDevice) != hipSuccess ||
                         hipMemcpy(S->fs[d], fsv.data(), fsv.size() * sizeof(double), hipMemcpyHostToDevice) !=
                             hipSuccess))
      rc = fail(BMG_ERR_HIP, "upload of the embedded boundary failed");
  }
  if (rc == BMG_OK &&
      hipMemcpy(S->bc, bcv.data(), bcv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    rc = fail(BMG_ERR_HIP, "upload of the embedded boundary failed");
  if (rc != BMG_OK) emb_free(g);
  return rc;
}

int bmg_embed_geometry(int n, const double* cs, const double* fs_x, const double* fs_y, double* px, double* py,
                       double* area, double* nx, double* ny) {
  if (n <= 0 || !cs || !fs_x || !fs_y) return fail(BMG_ERR_ARG, "bad arguments");
  embgeo::Frac f;
  f.n = n;
  const size_t m = n;
  f.cs.assign(cs, cs + m * m);
  f.fx.assign(fs_x, fs_x + (m + 1) * m);
  f.fy.assign(fs_y, fs_y + m * (m + 1));
  for (size_t q = 0; q < m * m; q++) {
    double v[5] = {0., 0., 0., 0., 0.};
    const double c = f.cs[q];
    if (c > 0. && c < 1.) {
      embgeo::V2 p, un;
      const EmbRec R = embgeo::cut_record(f, (int)(q / m), (int)(q % m), false, &p, &un);
      v[0] = p.x;
      v[1] = p.y;
      v[2] = R.area;
      v[3] = un.x;
      v[4] = un.y;
    }
    if (px) px[q] = v[0];
    if (py) py[q] = v[1];
    if (area) area[q] = v[2];
    if (nx) nx[q] = v[3];
    if (ny) ny[q] = v[4];
  }
  return BMG_OK;
}

int bmg_set_alpha_const(bmg_grid* g, double alpha) {
  TRY(check_grid(g));
  if (g->alpha_field || g->alpha_c != alpha) drop_graphs(g);
  g->alpha_field = false;
  g->alpha_c = alpha;
  g->coarse_coef_ready = false;
  return BMG_OK;
}

int bmg_set_lambda_const(bmg_grid* g, double lambda) {
  TRY(check_grid(g));
  if (g->lambda_field || g->lambda_c != lambda) drop_graphs(g);
  g->lambda_field = false;
  g->lambda_c = lambda;
  g->coarse_coef_ready = false;
  return BMG_OK;
}

int bmg_set_field(bmg_grid* g, int field, int level, const double* ptr, int on_device) {
  TRY(check_grid(g));
  if (!ptr) return fail(BMG_ERR_ARG, "null field pointer");
  if (level < 0 || level > g->depth) return fail(BMG_ERR_ARG, "level %d out of range", level);
  Level& lv = g->L[level];
  g->coarse_coef_ready = false;
  g->coarse_res_fresh = false;
  if (field >= BMG_F_ALPHA && field < BMG_F_ALPHA + g->dim) {
    TRY(set_face(g, field - BMG_F_ALPHA, level, ptr, on_device));
    if (!g->alpha_field) drop_graphs(g);
    g->alpha_field = true;
    return BMG_OK;
  }
  switch (field) {
    case BMG_F_A:
    case BMG_F_B:
      if (level != g->depth) return fail(BMG_ERR_ARG, "a and b live on the finest level only");
      return copy_in(g, field == BMG_F_A ? g->a : g->b, ptr, lv.ncell, on_device);
    case BMG_F_LAMBDA:
      TRY(copy_in(g, lv.lam64, ptr, lv.ncell, on_device));
      if (!g->lambda_field) drop_graphs(g);
      g->lambda_field = true;
      return BMG_OK;
    case BMG_F_RES:
    case BMG_F_DA: {
      void* dst = field == BMG_F_RES ? lv.res : lv.da[0];
      if (g->precision == BMG_FP64) return copy_in(g, dst, ptr, lv.ncell, on_device);
      // mixed: stage through the FP64 scratch of the finest residual
      TRY(copy_in(g, g->res64, ptr, lv.ncell, on_device));
      hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream, g->res64,
                         (float*)dst, lv.ncell);
      TRY(check_launch());
      HIPCHECK(hipStreamSynchronize(g->stream));
      return BMG_OK;
    }
    default: return fail(BMG_ERR_ARG, "unknown field %d", field);
  }
}

int bmg_get_field(bmg_grid* g, int field, int level, double* ptr, int on_device) {
  TRY(check_grid(g));
  if (!ptr) return fail(BMG_ERR_ARG, "null field pointer");
  if (level < 0 || level > g->depth) return fail(BMG_ERR_ARG, "level %d out of range", level);
  Level& lv = g->L[level];
  hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost;
  if (field >= BMG_F_ALPHA && field < BMG_F_ALPHA + g->dim) return get_face(g, field - BMG_F_ALPHA, level, ptr, on_device);
  const void* src = nullptr;
  switch (field) {
    case BMG_F_A:
    case BMG_F_B:
      if (level != g->depth) return fail(BMG_ERR_ARG, "a and b live on the finest level only");
      src = field == BMG_F_A ? g->a : g->b;
      break;
    case BMG_F_LAMBDA: src = lv.lam64; break;
    case BMG_F_RES:
      if (level == g->depth) { src = g->res64; break; }
      /* fall through */
    case BMG_F_DA: {
      const void* s = field == BMG_F_RES ? lv.res : lv.da[0];
      if (g->precision == BMG_FP64) { src = s; break; }
      double* tmp = nullptr;
      TRY(dalloc((void**)&tmp, lv.ncell * sizeof(double)));
      hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks(lv.ncell)), dim3(kBlock), 0, g->stream,
                         (const float*)s, tmp, lv.ncell);
      hipError_t e = hipMemcpyAsync(ptr, tmp, lv.ncell * sizeof(double), kind, g->stream);
      (void)hipStreamSynchronize(g->stream);
      (void)hipFree(tmp);
      if (e != hipSuccess) return fail(BMG_ERR_HIP, "copy failed: %s", hipGetErrorString(e));
      return BMG_OK;
    }
    default: return fail(BMG_ERR_ARG, "unknown field %d", field);
  }
  HIPCHECK(hipMemcpyAsync(ptr, src, lv.ncell * sizeof(double), kind, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

int bmg_restriction_alpha_lambda(bmg_grid* g) {
  TRY(check_grid(g));
  TRY(dispatch_restrict_coef(g));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

int bmg_residual(bmg_grid* g, double* maxres) {
  TRY(check_grid(g));
  TRY(emb_check(g, g->depth));
  double m = 0.;
  unsigned long long nans = 0;
  TRY(dispatch_residual(g, &m, false, &nans, true));
  if (maxres) *maxres = m;
  if (nans) return fail(BMG_ERR_NAN, "the residual held %llu NaN values", nans);
  return BMG_OK;
}

int bmg_relax(bmg_grid* g, int level) {
  TRY(check_grid(g));
  if (level < 0 || level > g->depth) return fail(BMG_ERR_ARG, "level %d out of range", level);
  TRY(emb_check(g, level));
  TRY(dispatch_relax(g, level));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

int bmg_mg_cycle(bmg_grid* g, int nrelax, int minlevel, int maxlevel) {
  TRY(check_grid(g));
  if (maxlevel < 0 || maxlevel > g->depth) return fail(BMG_ERR_ARG, "maxlevel %d out of range", maxlevel);
  TRY(emb_check(g, minlevel));
  if (nrelax < 0) return fail(BMG_ERR_ARG, "nrelax must be >= 0");
  TRY(dispatch_cycle(g, nrelax, minlevel, maxlevel));
  HIPCHECK(hipStreamSynchronize(g->stream));
  prof_collect(g);
  return BMG_OK;
}

int bmg_mg_solve(bmg_grid* g, double tolerance, int nrelax, int minlevel, bmg_mgstats* stats) {
  TRY(check_grid(g));
  return mg_solve_impl(g, tolerance, nrelax, minlevel, stats);
}

// poisson(), poisson.h:402-441
int bmg_poisson(bmg_grid* g, double tolerance, int nrelax, int minlevel, bmg_mgstats* stats) {
  TRY(check_grid(g));
  TRY(dispatch_restrict_coef(g));
  double defaultol = g->TOLERANCE;
  if (tolerance) g->TOLERANCE = tolerance;
  int r = mg_solve_impl(g, 0., nrelax, std::max(1, minlevel), stats);
  if (tolerance) g->TOLERANCE = defaultol;
  return r;
}

int bmg_project(bmg_grid* g, double* uf_x, double* uf_y, double* uf_z, double dt, int nrelax, int on_device,
                bmg_mgstats* stats) {
  TRY(check_grid(g));
  double* in[3] = {uf_x, uf_y, uf_z};
  for (int d = 0; d < g->dim; d++)
    if (!in[d]) return fail(BMG_ERR_ARG, "null face velocity component %d", d);
  // the reference's defaults: dt = 0 means 1, nrelax = 0 means 4
  if (dt == 0.) dt = 1.;
  if (nrelax == 0) nrelax = 4;
  if (!(dt > 0.) || nrelax < 0) return fail(BMG_ERR_ARG, "dt must be positive and nrelax non-negative");
  const Level& lf = g->L[g->depth];
  const size_t nface = (size_t)(lf.n + 1) * powd(lf.n, g->dim - 1);
  double* dev[3] = {nullptr, nullptr, nullptr};
  int rc = BMG_OK;
  if (!on_device) {
    for (int d = 0; d < g->dim && rc == BMG_OK; d++) {
      rc = dalloc((void**)&dev[d], nface * sizeof(double));
      if (rc == BMG_OK && hipMemcpy(dev[d], in[d], nface * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(BMG_ERR_HIP, "copy of the face velocity failed");
    }
  } else {
    for (int d = 0; d < g->dim; d++) dev[d] = in[d];
  }
  if (rc == BMG_OK) {
    switch (g->dim) {
      case 1: rc = project_t<1>(g, dev, dt, nrelax, stats); break;
      case 2: rc = project_t<2>(g, dev, dt, nrelax, stats); break;
      default: rc = project_t<3>(g, dev, dt, nrelax, stats); break;
    }
  }
  if (!on_device) {
    // the face update ran even when the solve reported a problem
    for (int d = 0; d < g->dim; d++) {
      if (dev[d] && hipMemcpy(in[d], dev[d], nface * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess &&
          rc == BMG_OK)
        rc = fail(BMG_ERR_HIP, "copy of the face velocity failed");
      dfreet(dev[d]);
    }
  }
  return rc;
}

}  // extern "C"

namespace {

// device copies of up to three host arrays of n doubles (or the device
// pointers themselves when on_device); copied back on request
struct DevTriple {
  double* p[3] = {nullptr, nullptr, nullptr};
  double* host[3] = {nullptr, nullptr, nullptr};
  bool owned = false;
  size_t n = 0;
  int take(double* const* in, int count, size_t n_, int on_device, bool copy_in) {
    n = n_;
    owned = !on_device;
    for (int d = 0; d < count; d++) {
      if (!in[d]) continue;
      if (on_device) {
        p[d] = in[d];
        continue;
      }
      host[d] = in[d];
      TRY(dalloc((void**)&p[d], n * sizeof(double)));
      if (copy_in) HIPCHECK(hipMemcpy(p[d], in[d], n * sizeof(double), hipMemcpyHostToDevice));
    }
    return BMG_OK;
  }
  int give_back() {
    if (!owned) return BMG_OK;
    for (int d = 0; d < 3; d++)
      if (p[d] && host[d]) HIPCHECK(hipMemcpy(host[d], p[d], n * sizeof(double), hipMemcpyDeviceToHost));
    return BMG_OK;
  }
  ~DevTriple() {
    if (owned)
      for (double*& q : p) dfreet(q);
  }
};

BCA fine_bca(const bmg_grid* g) {
  BCA bc;
  memset(&bc, 0, sizeof bc);
  for (int d = 0; d < 3; d++) bc.periodic[d] = d < g->dim ? g->periodic[d] : 0;
  for (int b = 0; b < 6; b++) {
    bc.sigma[b] = g->bc_sigma[b];
    bc.c[b] = g->bc_c[b];
  }
  return bc;
}

template <int DIM>
int centered_gradient_t(bmg_grid* g, double* const* acc, double* const* gout) {
  Level& lf = g->L[g->depth];
  Geo<DIM> geo{lf.n, lf.log2n};
  CenteredArgs A;
  memset(&A, 0, sizeof A);
  A.p = g->a;
  for (int d = 0; d < 3; d++) {
    A.alo[d] = g->alpha_field && d < DIM ? lf.alo64[d] : nullptr;
    A.awall[d] = g->alpha_field && d < DIM ? lf.awall64[d] : nullptr;
    A.acc[d] = d < DIM ? acc[d] : nullptr;
    A.gout[d] = d < DIM ? gout[d] : nullptr;
  }
  if (!g->alpha_field) A.alo[0] = nullptr;
  A.alpha_c = g->alpha_c;
  hipLaunchKernelGGL((k_centered_gradient<DIM>), dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, A, geo,
                     fine_bca(g), delta(g, g->depth));
  TRY(check_launch());
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

}  // namespace

extern "C" {

int bmg_centered_gradient(bmg_grid* g, const double* a_x, const double* a_y, const double* a_z, double* g_x,
                          double* g_y, double* g_z, int on_device) {
  TRY(check_grid(g));
  double* out[3] = {g_x, g_y, g_z};
  double* acc[3] = {(double*)a_x, (double*)a_y, (double*)a_z};
  int nacc = 0;
  for (int d = 0; d < g->dim; d++) {
    if (!out[d]) return fail(BMG_ERR_ARG, "null gradient component %d", d);
    nacc += acc[d] != nullptr;
  }
  if (nacc != 0 && nacc != g->dim) return fail(BMG_ERR_ARG, "acceleration: every component or none");
  const Level& lf = g->L[g->depth];
  const size_t nface = (size_t)(lf.n + 1) * powd(lf.n, g->dim - 1);
  DevTriple da, dg;
  TRY(da.take(acc, g->dim, nface, on_device, true));
  TRY(dg.take(out, g->dim, lf.ncell, on_device, false));
  switch (g->dim) {
    case 1: TRY(centered_gradient_t<1>(g, da.p, dg.p)); break;
    case 2: TRY(centered_gradient_t<2>(g, da.p, dg.p)); break;
    default: TRY(centered_gradient_t<3>(g, da.p, dg.p)); break;
  }
  return dg.give_back();
}

int bmg_correction(bmg_grid* g, double* u_x, double* u_y, double* u_z, const double* g_x, const double* g_y,
                   const double* g_z, double dt, int on_device) {
  TRY(check_grid(g));
  double* u[3] = {u_x, u_y, u_z};
  double* gv[3] = {(double*)g_x, (double*)g_y, (double*)g_z};
  for (int d = 0; d < g->dim; d++)
    if (!u[d] || !gv[d]) return fail(BMG_ERR_ARG, "null component %d", d);
  const size_t n = g->L[g->depth].ncell;
  DevTriple du, dgv;
  TRY(du.take(u, g->dim, n, on_device, true));
  TRY(dgv.take(gv, g->dim, n, on_device, true));
  for (int d = 0; d < g->dim; d++)
    hipLaunchKernelGGL(k_correction, dim3(nblocks(n)), dim3(kBlock), 0, g->stream, du.p[d], (const double*)dgv.p[d],
                       dt, n);
  TRY(check_launch());
  HIPCHECK(hipStreamSynchronize(g->stream));
  return du.give_back();
}

int bmg_coef_gradient(bmg_grid* g, const double* psi, double* galpha_x, double* galpha_y, double* galpha_z,
                      double* glambda, int on_device) {
  TRY(check_grid(g));
  if (!psi) return fail(BMG_ERR_ARG, "null adjoint state");
  for (int d = 0; d < g->dim; d++)
    if (g->periodic[d]) return fail(BMG_ERR_UNSUPPORTED, "coefficient gradients on periodic boundaries");
  double* ga[3] = {galpha_x, g->dim > 1 ? galpha_y : nullptr, g->dim > 2 ? galpha_z : nullptr};
  if (g->dim == 3) return coef_gradient_t<3>(g, psi, ga, glambda, on_device);
  if (g->dim == 2) return coef_gradient_t<2>(g, psi, ga, glambda, on_device);
  return coef_gradient_t<1>(g, psi, ga, glambda, on_device);
}

int bmg_synchronize(bmg_grid* g) {
  TRY(check_grid(g));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return BMG_OK;
}

int bmg_profile_enable(bmg_grid* g, int enable) {
  TRY(check_grid(g));
  prof_collect(g);
  g->prof = enable != 0;
  for (int k = 0; k < PROF_N; k++) {
    g->prof_ms[k] = 0.;
    g->prof_n[k] = 0;
  }
  return BMG_OK;
}

int bmg_profile_get(bmg_grid* g, int kind, double* total_ms, long* count) {
  TRY(check_grid(g));
  if (kind < 0 || kind >= PROF_N) return fail(BMG_ERR_ARG, "unknown profile kind %d", kind);
  prof_collect(g);
  if (total_ms) *total_ms = g->prof_ms[kind];
  if (count) *count = g->prof_n[kind];
  return BMG_OK;
}

// --------------------------------------------------------------------------
// poisson() on the reference's own storage.  The finest level's record
// array is uploaded once per call and a, b, alpha and lambda are gathered
// from it on the device; the whole solve runs on the GPU; a (and res when
// asked for) is scattered into the device copy, which goes back to the host
// in one copy (ghost cells and the other fields come back as they were: the
// caller marks `a` dirty so that the reference refills its ghosts on the
// next stencil).

}  // extern "C"

namespace {

struct ViewCache {
  bmg_grid* g = nullptr;
  int dim = -1, depth = -1;
  double L0 = 0., X0 = 0., Y0 = 0., Z0 = 0.;
  int precision = BMG_FP64, smoother = BMG_RELAX_RBGS;  // bmg_view_configure
  // device copy of the finest level's record array, and a dense box of
  // one field (both grown on demand, kept across calls)
  void* rec = nullptr;
  size_t rec_bytes = 0;
  double* box = nullptr;
  size_t box_n = 0;
};
ViewCache g_view_cache;

// a dense box of one field to or from the device record copy
int view_move(bmg_grid* g, ViewCache& vc, const bmg_multigrid_view* v, int h, const int ext[3], bool gather) {
  ViewBox B;
  B.nrec = v->datasize / sizeof(double);
  B.h = h;
  B.m = (1 << v->depth) + 4;
  size_t total = 1;
  for (int d = 0; d < 3; d++) {
    B.e[d] = d < v->dimension ? ext[d] : 1;
    total *= (size_t)B.e[d];
  }
  const unsigned nb = (unsigned)((total + kBlock - 1) / kBlock);
  double* rec = (double*)vc.rec;
  if (gather) {
    if (v->dimension == 3) hipLaunchKernelGGL((k_view_gather<3>), dim3(nb), dim3(kBlock), 0, g->stream, rec, B, vc.box);
    else if (v->dimension == 2) hipLaunchKernelGGL((k_view_gather<2>), dim3(nb), dim3(kBlock), 0, g->stream, rec, B, vc.box);
    else hipLaunchKernelGGL((k_view_gather<1>), dim3(nb), dim3(kBlock), 0, g->stream, rec, B, vc.box);
  } else {
    if (v->dimension == 3) hipLaunchKernelGGL((k_view_scatter<3>), dim3(nb), dim3(kBlock), 0, g->stream, rec, B, vc.box);
    else if (v->dimension == 2) hipLaunchKernelGGL((k_view_scatter<2>), dim3(nb), dim3(kBlock), 0, g->stream, rec, B, vc.box);
    else hipLaunchKernelGGL((k_view_scatter<1>), dim3(nb), dim3(kBlock), 0, g->stream, rec, B, vc.box);
  }
  return check_launch();
}



}  // namespace

extern "C" {

int bmg_poisson_view(const bmg_multigrid_view* v, const bmg_poisson_args* p, const bmg_view_bc* bc,
                     double TOLERANCE, int NITERMAX, int NITERMIN, bmg_mgstats* stats) {
  if (!v || !p || !bc || !v->d) return fail(BMG_ERR_ARG, "null view argument");
  int dim = v->dimension, depth = v->depth;
  if (dim < 1 || dim > 3) return fail(BMG_ERR_ARG, "dimension %d not in 1..3", dim);
  if (v->datasize % sizeof(double) || p->a < 0 || p->b < 0 || p->a >= v->nvarmax || p->b >= v->nvarmax)
    return fail(BMG_ERR_ARG, "bad field handles or datasize");
  ViewCache& vc = g_view_cache;
  if (!vc.g || vc.dim != dim || vc.depth != depth || vc.L0 != v->L0 || vc.X0 != v->X0 || vc.Y0 != v->Y0 ||
      vc.Z0 != v->Z0) {
    bmg_grid_free(vc.g);
    vc.g = nullptr;
    TRY(bmg_grid_new(&vc.g, dim, depth, v->L0, v->X0, v->Y0, v->Z0));
    vc.dim = dim; vc.depth = depth; vc.L0 = v->L0; vc.X0 = v->X0; vc.Y0 = v->Y0; vc.Z0 = v->Z0;
  }
  bmg_grid* g = vc.g;
  if (g->precision != vc.precision) TRY(bmg_set_precision(g, vc.precision));
  if (g->smoother != vc.smoother) TRY(bmg_set_smoother(g, vc.smoother));
  const int n = 1 << depth;
  int ext[3] = {n, n, n};
  // the finest level's records, uploaded once; every field the solve reads
  // is gathered from them on the device
  const size_t rec_bytes = powd((size_t)n + 4, dim) * v->datasize;
  const size_t box_n = powd(n, dim - 1) * (size_t)(n + 1);
  if (vc.rec_bytes < rec_bytes) {
    dfree(vc.rec);
    vc.rec_bytes = 0;
    TRY(dalloc(&vc.rec, rec_bytes));
    vc.rec_bytes = rec_bytes;
  }
  if (vc.box_n < box_n) {
    dfreet(vc.box);
    vc.box_n = 0;
    TRY(dalloc((void**)&vc.box, box_n * sizeof(double)));
    vc.box_n = box_n;
  }
  HIPCHECK(hipMemcpyAsync(vc.rec, v->d[depth], rec_bytes, hipMemcpyHostToDevice, g->stream));
  TRY(view_move(g, vc, v, p->a, ext, true));
  TRY(bmg_set_field(g, BMG_F_A, depth, vc.box, 1));
  TRY(view_move(g, vc, v, p->b, ext, true));
  TRY(bmg_set_field(g, BMG_F_B, depth, vc.box, 1));
  // alpha: unset (0) or constant handles -> constant; otherwise face fields
  bool aconst = true;
  for (int c = 0; c < dim; c++) aconst &= p->alpha[c] == 0 || p->alpha[c] >= v->nvarmax;
  if (aconst) {
    double val = 1.;  // unityf
    if (p->alpha[0] >= v->nvarmax) val = v->constant[p->alpha[0] - v->nvarmax];
    TRY(bmg_set_alpha_const(g, val));
  } else {
    for (int c = 0; c < dim; c++) {
      if (p->alpha[c] <= 0 || p->alpha[c] >= v->nvarmax)
        return fail(BMG_ERR_UNSUPPORTED, "alpha mixes constant and variable components");
      int fe[3] = {n, n, n};
      fe[c] = n + 1;
      TRY(view_move(g, vc, v, p->alpha[c], fe, true));
      TRY(bmg_set_field(g, BMG_F_ALPHA + c, depth, vc.box, 1));
    }
  }
  if (p->lambda <= 0 || p->lambda >= v->nvarmax) {
    double val = p->lambda >= v->nvarmax ? v->constant[p->lambda - v->nvarmax] : 0.;  // zeroc
    TRY(bmg_set_lambda_const(g, val));
  } else {
    TRY(view_move(g, vc, v, p->lambda, ext, true));
    TRY(bmg_set_field(g, BMG_F_LAMBDA, depth, vc.box, 1));
  }
  for (int b = 0; b < 2 * dim; b++) {
    if (bc->kind[b] == BMG_BC_PERIODIC) TRY(bmg_set_bc(g, b, BMG_BC_PERIODIC, nullptr));
    else if (bc->kind[b] == BMG_BC_AFFINE) TRY(bmg_set_bc_affine(g, b, bc->sigma[b], bc->c[b]));
    else return fail(BMG_ERR_ARG, "view BC kind must be AFFINE or PERIODIC");
  }
  TRY(bmg_set_tolerance(g, TOLERANCE));
  TRY(bmg_set_niter(g, NITERMAX, NITERMIN));
  bmg_mgstats s;
  int r = bmg_poisson(g, p->tolerance, p->nrelax, p->minlevel, &s);
  if (r != BMG_OK && r != BMG_ERR_NAN) return r;
  if (stats) *stats = s;
  // a (and res) into the device records, then the records back to the
  // host in one copy: every other byte goes back as it came
  TRY(bmg_get_field(g, BMG_F_A, depth, vc.box, 1));
  TRY(view_move(g, vc, v, p->a, ext, false));
  if (p->res >= 0 && p->res < v->nvarmax) {
    TRY(bmg_get_field(g, BMG_F_RES, depth, vc.box, 1));
    TRY(view_move(g, vc, v, p->res, ext, false));
  }
  HIPCHECK(hipMemcpyAsync(v->d[depth], vc.rec, rec_bytes, hipMemcpyDeviceToHost, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return r;
}

int bmg_view_configure(int precision, int smoother) {
  if (precision != BMG_FP64 && precision != BMG_MIXED) return fail(BMG_ERR_ARG, "unknown precision %d", precision);
  if (smoother != BMG_RELAX_GS && smoother != BMG_RELAX_JACOBI && smoother != BMG_RELAX_RBGS)
    return fail(BMG_ERR_ARG, "unknown smoother %d", smoother);
  g_view_cache.precision = precision;
  g_view_cache.smoother = smoother;
  return BMG_OK;
}

int bmg_poisson_noalloc(const bmg_noalloc_problem* q, double tolerance, int nrelax, int minlevel,
                        bmg_mgstats* stats) {
  if (!q || !q->a || !q->b) return fail(BMG_ERR_ARG, "null problem, a or b");
  char* base = nullptr;
  TRY(scratch_address(&base));
  // every device array of the call comes out of the scratch; the previous
  // call's contents are overwritten field by field before they are read
  Arena arena{base, kScratchBytes, 0};
  struct Scope {
    Arena* prev;
    explicit Scope(Arena* a) : prev(t_arena) { t_arena = a; }
    ~Scope() { t_arena = prev; }
  } scope(&arena);
  bmg_grid* g = nullptr;
  TRY(grid_create(&g, q->dimension, q->depth, q->L0, q->X0, q->Y0, q->Z0, q->precision, true,
                  (hipStream_t)q->stream));
  struct Free {
    bmg_grid* g;
    ~Free() { bmg_grid_free(g); }
  } release{g};
  g->use_graphs = false;  // instantiating a graph may allocate
  const int dim = q->dimension, n = 1 << q->depth;
  TRY(bmg_set_smoother(g, q->smoother));
  for (int b = 0; b < 2 * dim; b++) {
    if (q->bc_kind[b] == BMG_BC_PERIODIC) TRY(bmg_set_bc(g, b, BMG_BC_PERIODIC, nullptr));
    else if (q->bc_kind[b] == BMG_BC_AFFINE) TRY(bmg_set_bc_affine(g, b, q->bc_sigma[b], q->bc_c[b]));
    else return fail(BMG_ERR_ARG, "bc_kind[%d] must be BMG_BC_AFFINE or BMG_BC_PERIODIC", b);
  }
  TRY(bmg_set_field(g, BMG_F_A, q->depth, q->a, 1));
  TRY(bmg_set_field(g, BMG_F_B, q->depth, q->b, 1));
  int nalpha = 0;
  for (int c = 0; c < dim; c++) nalpha += q->alpha[c] != nullptr;
  if (nalpha != 0 && nalpha != dim) return fail(BMG_ERR_ARG, "alpha: every component or none");
  if (nalpha)
    for (int c = 0; c < dim; c++) TRY(bmg_set_field(g, BMG_F_ALPHA + c, q->depth, q->alpha[c], 1));
  else
    TRY(bmg_set_alpha_const(g, q->alpha_c));
  if (q->lambda) TRY(bmg_set_field(g, BMG_F_LAMBDA, q->depth, q->lambda, 1));
  else TRY(bmg_set_lambda_const(g, q->lambda_c));
  // the fused residual's second buffer of a, which bmg_poisson would
  // otherwise allocate during the first cycle
  if (dim == 3 && fused_residual(g) && !g->a2)
    TRY(dalloc((void**)&g->a2, g->L[g->depth].ncell * sizeof(double)));
  TRY(bmg_set_tolerance(g, q->TOLERANCE));
  TRY(bmg_set_niter(g, q->NITERMAX, q->NITERMIN));
  bmg_mgstats st;
  const int r = bmg_poisson(g, tolerance, nrelax, minlevel, &st);
  if (r != BMG_OK && r != BMG_ERR_NAN) return r;
  if (stats) *stats = st;
  HIPCHECK(hipMemcpyAsync(q->a, g->a, powd(n, dim) * sizeof(double), hipMemcpyDeviceToDevice, g->stream));
  HIPCHECK(hipStreamSynchronize(g->stream));
  return r;
}

int bmg_view_release(void) {
  ViewCache& vc = g_view_cache;
  if (vc.g) bmg_grid_free(vc.g);
  vc.g = nullptr;
  vc.dim = vc.depth = -1;
  dfree(vc.rec);
  vc.rec_bytes = 0;
  dfreet(vc.box);
  vc.box_n = 0;
  return BMG_OK;
}

}  // extern "C"

namespace {

template <int DIM>
int project_t(bmg_grid* g, double* const* uf, double dt, int nrelax, bmg_mgstats* stats) {
  Level& lf = g->L[g->depth];
  Geo<DIM> geo{lf.n, lf.log2n};
  const double D = delta(g, g->depth);
  hipLaunchKernelGGL((k_divergence<DIM>), dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, uf[0],
                     DIM >= 2 ? uf[1] : nullptr, DIM >= 3 ? uf[2] : nullptr, g->b, geo, dt * D);
  TRY(check_launch());
  g->coarse_res_fresh = false;
  const int rc = bmg_poisson(g, g->TOLERANCE / (dt * dt), nrelax, 0, stats);
  if (rc != BMG_OK && rc != BMG_ERR_NAN) return rc;
  BCA bc;
  memset(&bc, 0, sizeof bc);
  for (int d = 0; d < 3; d++) bc.periodic[d] = d < g->dim ? g->periodic[d] : 0;
  for (int b = 0; b < 6; b++) {
    bc.sigma[b] = g->bc_sigma[b];
    bc.c[b] = g->bc_c[b];
  }
  const size_t nface = (size_t)(lf.n + 1) * powd(lf.n, DIM - 1);
  for (int d = 0; d < DIM; d++) {
    const double* alo = g->alpha_field ? lf.alo64[d] : nullptr;
    const double* aw = g->alpha_field ? lf.awall64[d] : nullptr;
    hipLaunchKernelGGL((k_face_update<DIM>), dim3(nblocks(nface)), dim3(kBlock), 0, g->stream, uf[d], g->a, alo,
                       aw, g->alpha_c, geo, bc, d, dt, D);
  }
  TRY(check_launch());
  HIPCHECK(hipStreamSynchronize(g->stream));
  return rc;
}

}  // namespace

// ------------------------------------------------------------ viscosity()

namespace {

constexpr int kViscSlots = 16;
enum { VS_DA = 0, VS_DN = 3, VS_RES = 6, VS_MLO = 9, VS_MWALL = 12, VS_RHO = 15 };

double*& vbuf(bmg_grid* g, int l, int slot) { return g->visc[(size_t)l * kViscSlots + slot]; }

int visc_alloc(bmg_grid* g) {
  if (!g->visc.empty()) return BMG_OK;
  g->visc.assign((size_t)(g->depth + 1) * kViscSlots, nullptr);
  for (int l = 0; l <= g->depth; l++) {
    const Level& lv = g->L[l];
    for (int c = 0; c < g->dim; c++) {
      TRY(dalloc((void**)&vbuf(g, l, VS_DA + c), lv.ncell * sizeof(double)));
      TRY(dalloc((void**)&vbuf(g, l, VS_DN + c), lv.ncell * sizeof(double)));
      TRY(dalloc((void**)&vbuf(g, l, VS_RES + c), lv.ncell * sizeof(double)));
      TRY(dalloc((void**)&vbuf(g, l, VS_MLO + c), lv.ncell * sizeof(double)));
      TRY(dalloc((void**)&vbuf(g, l, VS_MWALL + c), std::max<size_t>(lv.nwall, 1) * sizeof(double)));
    }
    TRY(dalloc((void**)&vbuf(g, l, VS_RHO), lv.ncell * sizeof(double)));
  }
  for (int c = 0; c < g->dim; c++) TRY(dalloc((void**)&g->visc_r[c], g->L[g->depth].ncell * sizeof(double)));
  return BMG_OK;
}

// the viscous problem of one call: u (device, finest level), constant or
// field mu and rho, the components' wall sigmas
template <int DIM>
struct Visc {
  bmg_grid* g;
  double* u[3];
  bool mu_field, rho_field;
  double mu_c, rho_c, dt;
  int sig[3][6];

  ViscArgs args(int l, double* const* in, double* const* out, double* const* r) const {
    ViscArgs A;
    memset(&A, 0, sizeof A);
    for (int c = 0; c < DIM; c++) {
      A.u[c] = in[c];
      A.out[c] = out[c];
      A.r[c] = r[c];
      if (mu_field) {
        A.mlo[c] = vbuf(g, l, VS_MLO + c);
        A.mwall[c] = vbuf(g, l, VS_MWALL + c);
      }
      for (int b = 0; b < 6; b++) A.sig[c][b] = sig[c][b];
    }
    // periodic dimensions wrap; on level 0 the periodic copy gives every
    // ghost the root value (grid/multigrid.h:655-664)
    bool anyp = false;
    for (int d = 0; d < 3; d++) {
      A.periodic[d] = d < DIM ? g->periodic[d] : 0;
      anyp |= A.periodic[d] != 0;
    }
    if (l == 0 && anyp)
      for (int c = 0; c < 3; c++)
        for (int b = 0; b < 6; b++) A.sig[c][b] = 1;
    A.mu_c = mu_c;
    A.rho = rho_field ? vbuf(g, l, VS_RHO) : nullptr;
    A.rho_c = rho_c;
    A.dt = dt;
    A.D = delta(g, l);
    A.axi = DIM == 2 && g->axi;
    A.y0 = g->X0[1];
    return A;
  }
  double* const* lv(int l, int slot) const { return &g->visc[(size_t)l * kViscSlots + slot]; }

  // restriction ({mu, rho}) (viscosity.h:173-175)
  void restrict_coefficients() {
    for (int l = g->depth - 1; l >= 0; l--) {
      const Level& lc = g->L[l];
      Geo<DIM> gc{lc.n, lc.log2n};
      if (rho_field)
        hipLaunchKernelGGL((k_restrict<DIM, double>), dim3(nblocks(lc.ncell)), dim3(kBlock), 0, g->stream,
                           vbuf(g, l + 1, VS_RHO), vbuf(g, l, VS_RHO), gc);
      if (mu_field)
        for (int d = 0; d < DIM; d++) {
          hipLaunchKernelGGL((k_restrict_face_lo<DIM>), dim3(nblocks(lc.ncell)), dim3(kBlock), 0, g->stream,
                             vbuf(g, l + 1, VS_MLO + d), vbuf(g, l, VS_MLO + d), gc, d);
          hipLaunchKernelGGL((k_restrict_face_wall<DIM>), dim3(nblocks(std::max<size_t>(lc.nwall, 1))),
                             dim3(kBlock), 0, g->stream, vbuf(g, l + 1, VS_MWALL + d), vbuf(g, l, VS_MWALL + d),
                             lc.n, d);
        }
    }
  }

  int residual(double* maxres, unsigned long long* nans) {
    const Level& lf = g->L[g->depth];
    Geo<DIM> geo{lf.n, lf.log2n};
    hipLaunchKernelGGL(k_stat_reset, dim3(1), dim3(64), 0, g->stream, g->stat);
    hipLaunchKernelGGL((k_visc_residual<DIM>), dim3(std::min(nblocks(lf.ncell), 8192u)), dim3(kBlock), 0, g->stream,
                       args(g->depth, u, lv(g->depth, VS_RES), g->visc_r), geo, g->stat);
    TRY(check_launch());
    Stat st;
    HIPCHECK(hipMemcpyAsync(&st, g->stat, sizeof st, hipMemcpyDeviceToHost, g->stream));
    HIPCHECK(hipStreamSynchronize(g->stream));
    memcpy(maxres, &st.maxbits, sizeof(double));
    *nans += st.nans;
    return BMG_OK;
  }

  // viscosity_explicit (viscosity.h:180-190): resb = residual_viscosity
  // with b = u, then u += res
  int explicit_step(bmg_mgstats* out) {
    const Level& lf = g->L[g->depth];
    for (int c = 0; c < DIM; c++)
      HIPCHECK(hipMemcpyAsync(g->visc_r[c], u[c], lf.ncell * sizeof(double), hipMemcpyDeviceToDevice, g->stream));
    bmg_mgstats s;
    memset(&s, 0, sizeof s);
    unsigned long long nans = 0;
    TRY(residual(&s.resb, &nans));
    for (int c = 0; c < DIM; c++)
      hipLaunchKernelGGL((k_correct<double>), dim3(nblocks(lf.ncell)), dim3(kBlock), 0, g->stream, u[c],
                         (const double*)vbuf(g, g->depth, VS_RES + c), lf.ncell);
    TRY(check_launch());
    if (out) *out = s;
    return nans ? fail(BMG_ERR_NAN, "viscous residual is NaN") : BMG_OK;
  }

  // mg_cycle (poisson.h:32-90) with minlevel 0 and relax_viscosity in the
  // grid's smoother: red-black (default) or the reference's JACOBI variant
  int cycle(int nrelax) {
    const int D = g->depth;
    for (int l = D - 1; l >= 0; l--) {
      Geo<DIM> gc{g->L[l].n, g->L[l].log2n};
      for (int c = 0; c < DIM; c++)
        hipLaunchKernelGGL((k_restrict<DIM, double>), dim3(nblocks(g->L[l].ncell)), dim3(kBlock), 0, g->stream,
                           vbuf(g, l + 1, VS_RES + c), vbuf(g, l, VS_RES + c), gc);
    }
    for (int l = 0; l <= D; l++) {
      const Level& lv_ = g->L[l];
      Geo<DIM> gf{lv_.n, lv_.log2n};
      for (int c = 0; c < DIM; c++) {
        if (l == 0) {
          hipLaunchKernelGGL((k_fill<double>), dim3(nblocks(lv_.ncell)), dim3(kBlock), 0, g->stream,
                             vbuf(g, l, VS_DA + c), lv_.ncell, 0.);
        } else {
          BCH bcc;
          memset(&bcc, 0, sizeof bcc);
          bool anyp = false;
          for (int d = 0; d < 3; d++) {
            bcc.periodic[d] = d < DIM ? g->periodic[d] : 0;
            anyp |= bcc.periodic[d] != 0;
          }
          for (int b = 0; b < 6; b++) bcc.sigma[b] = (l - 1 == 0 && anyp) ? 1 : sig[c][b];
          hipLaunchKernelGGL((k_prolong<DIM, double>), dim3(nblocks(lv_.ncell)), dim3(kBlock), 0, g->stream,
                             vbuf(g, l - 1, VS_DA + c), vbuf(g, l, VS_DA + c), gf, bcc);
        }
      }
      // passes per sweep: Jacobi one (colour -1), Gauss-Seidel one (the
      // single cell of level 0 as colour 0), red-black two
      const bool jacobi = g->smoother == BMG_RELAX_JACOBI, gs = g->smoother == BMG_RELAX_GS;
      const int c0 = jacobi ? -1 : 0, c1 = jacobi || gs ? c0 + 1 : 2;
      for (int i = 0; i < nrelax; i++)
        for (int colour = c0; colour < c1; colour++) {
          const ViscArgs A = args(l, lv(l, VS_DA), lv(l, VS_DN), lv(l, VS_RES));
          if (lv_.n < 2) {
            hipLaunchKernelGGL((k_visc_one<DIM>), dim3(1), dim3(64), 0, g->stream, A, gf, colour);
          } else if (gs) {
            // hyperplanes in one workgroup up to 16384 cells per
            // hyperplane's index range, else one launch each
            const size_t count = DIM == 3 ? (size_t)lv_.n * lv_.n : (size_t)lv_.n;
            if (count <= 16384) {
              hipLaunchKernelGGL((k_visc_gs_block<DIM>), dim3(1), dim3(1024), 0, g->stream, A, gf);
            } else {
              const int smax = (DIM == 3 ? 7 : 3) * (lv_.n - 1);
              for (int s = 0; s <= smax; s++)
                hipLaunchKernelGGL((k_visc_gs_plane<DIM>), dim3(nblocks(count)), dim3(kBlock), 0, g->stream, A,
                                   gf, s);
            }
          } else if (jacobi)
            hipLaunchKernelGGL((k_visc_jacobi<DIM>), dim3(nblocks(lv_.ncell)), dim3(kBlock), 0, g->stream, A, gf);
          else
            hipLaunchKernelGGL((k_visc_rb<DIM>), dim3(nblocks(lv_.ncell / 2)), dim3(kBlock), 0, g->stream, A, gf,
                               colour);
          for (int c = 0; c < DIM; c++) std::swap(vbuf(g, l, VS_DA + c), vbuf(g, l, VS_DN + c));
        }
    }
    const size_t n = g->L[D].ncell;
    for (int c = 0; c < DIM; c++)
      hipLaunchKernelGGL((k_correct<double>), dim3(nblocks(n)), dim3(kBlock), 0, g->stream, u[c],
                         (const double*)vbuf(g, D, VS_DA + c), n);
    return check_launch();
  }

  // mg_solve (poisson.h:139-235) around residual_viscosity / relax_viscosity
  int solve(int nrelax, bmg_mgstats* out) {
    const Level& lf = g->L[g->depth];
    // r = u (viscosity.h:166-169) and the sum of b
    bmg_mgstats s;
    memset(&s, 0, sizeof s);
    for (int c = 0; c < DIM; c++) {
      HIPCHECK(hipMemcpyAsync(g->visc_r[c], u[c], lf.ncell * sizeof(double), hipMemcpyDeviceToDevice, g->stream));
      int np = (int)std::min<size_t>((size_t)g->npartial, nblocks(lf.ncell));
      hipLaunchKernelGGL(k_sum_partial, dim3(np), dim3(kBlock), 0, g->stream, g->visc_r[c], lf.ncell, g->partial);
      hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, g->stream, g->partial, np, g->scalar);
      TRY(check_launch());
      double sc = 0.;
      HIPCHECK(hipMemcpyAsync(&sc, g->scalar, sizeof(double), hipMemcpyDeviceToHost, g->stream));
      HIPCHECK(hipStreamSynchronize(g->stream));
      s.sum += sc;
    }
    restrict_coefficients();
    TRY(check_launch());
    s.nrelax = nrelax > 0 ? nrelax : 4;
    unsigned long long nans = 0;
    double resb;
    TRY(residual(&s.resb, &nans));
    resb = s.resa = s.resb;
    const double tolerance = g->TOLERANCE;
    for (s.i = 0; s.i < g->NITERMAX && (s.i < g->NITERMIN || s.resa > tolerance); s.i++) {
      TRY(cycle(s.nrelax));
      TRY(residual(&s.resa, &nans));
      if (s.resa > tolerance) {
        if (resb / s.resa < 1.2 && s.nrelax < 100)
          s.nrelax++;
        else if (resb / s.resa > 10 && s.nrelax > 2)
          s.nrelax--;
      }
      resb = s.resa;
    }
    s.minlevel = 0;
    if (out) *out = s;
    return nans ? fail(BMG_ERR_NAN, "viscous residual is NaN") : BMG_OK;
  }
};

// copy of a face field in the user layout (n+1 along d) into low faces and
// wall faces (as set_face does for alpha); along a periodic dimension face n
// is face 0, so the wall faces take face 0 and face n is not read
int visc_set_face(bmg_grid* g, int d, const double* src, int on_device) {
  const int l = g->depth;
  const Level& lv = g->L[l];
  const int n = lv.n;
  size_t outer = 1, inner = 1;
  for (int e = 0; e < g->dim; e++) {
    if (e < d) outer *= n;
    if (e > d) inner *= n;
  }
  hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  HIPCHECK(hipMemcpy2DAsync(vbuf(g, l, VS_MLO + d), (size_t)n * inner * sizeof(double), src,
                            (size_t)(n + 1) * inner * sizeof(double), (size_t)n * inner * sizeof(double), outer, kind,
                            g->stream));
  const double* wall = g->periodic[d] ? src : src + (size_t)n * inner;
  HIPCHECK(hipMemcpy2DAsync(vbuf(g, l, VS_MWALL + d), inner * sizeof(double), wall,
                            (size_t)(n + 1) * inner * sizeof(double), inner * sizeof(double), outer, kind, g->stream));
  return BMG_OK;
}

template <int DIM>
int viscosity_t(bmg_grid* g, double* const* u, const double* const* mu, double mu_c, const double* rho, double rho_c,
                double dt, const int (*sig)[6], int nrelax, int on_device, bmg_mgstats* stats, bool expl) {
  TRY(visc_alloc(g));
  Visc<DIM> V;
  V.g = g;
  for (int c = 0; c < 3; c++) V.u[c] = c < DIM ? u[c] : nullptr;
  V.mu_field = mu[0] != nullptr;
  V.rho_field = rho != nullptr;
  V.mu_c = mu_c;
  V.rho_c = rho_c;
  V.dt = dt;
  for (int c = 0; c < 3; c++)
    for (int b = 0; b < 6; b++) V.sig[c][b] = sig[c][b];
  const Level& lf = g->L[g->depth];
  if (V.mu_field)
    for (int d = 0; d < DIM; d++) TRY(visc_set_face(g, d, mu[d], on_device));
  if (V.rho_field)
    HIPCHECK(hipMemcpyAsync(vbuf(g, g->depth, VS_RHO), rho, lf.ncell * sizeof(double),
                            on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, g->stream));
  const int rc = expl ? V.explicit_step(stats) : V.solve(nrelax, stats);
  HIPCHECK(hipStreamSynchronize(g->stream));
  return rc;
}

}  // namespace

extern "C" {

int bmg_set_axi(bmg_grid* g, int axi) {
  TRY(check_grid(g));
  if (axi && g->dim != 2) return fail(BMG_ERR_UNSUPPORTED, "AXI is axisymmetric 2D");
  g->axi = axi != 0;
  return BMG_OK;
}

}  // extern "C"

namespace {

// bmg_viscosity and bmg_viscosity_explicit: checks, host copies, the step
int viscosity_entry(bmg_grid* g, double* u_x, double* u_y, double* u_z, const double* mu_x, const double* mu_y,
                    const double* mu_z, double mu_c, const double* rho, double rho_c, double dt, const int* sigma,
                    int nrelax, int on_device, bmg_mgstats* stats, bool expl) {
  TRY(check_grid(g));
  if (g->dim < 2) return fail(BMG_ERR_UNSUPPORTED, "viscosity() is built for 2D and 3D grids");
  double* u[3] = {u_x, u_y, u_z};
  const double* mu[3] = {mu_x, mu_y, mu_z};
  int nmu = 0;
  for (int d = 0; d < g->dim; d++) {
    if (!u[d]) return fail(BMG_ERR_ARG, "null velocity component %d", d);
    nmu += mu[d] != nullptr;
  }
  if (nmu != 0 && nmu != g->dim) return fail(BMG_ERR_ARG, "mu: every component or none");
  if (!(dt >= 0.) || nrelax < 0) return fail(BMG_ERR_ARG, "dt must be non-negative and nrelax non-negative");
  if (!rho && !(rho_c > 0.)) return fail(BMG_ERR_ARG, "rho must be positive");
  if (!nmu && !(mu_c >= 0.)) return fail(BMG_ERR_ARG, "mu must be non-negative");
  int sig[3][6];
  for (int c = 0; c < 3; c++)
    for (int b = 0; b < 6; b++) {
      // walls at rest by default: the normal component antisymmetric
      int v = sigma ? sigma[c * 6 + b] : (b / 2 == c ? -1 : 1);
      if (c < g->dim && b < 2 * g->dim && v != 1 && v != -1) return fail(BMG_ERR_ARG, "sigma must be +1 or -1");
      sig[c][b] = v == -1 ? -1 : 1;
    }
  const Level& lf = g->L[g->depth];
  double* dev[3] = {nullptr, nullptr, nullptr};
  int rc = BMG_OK;
  if (!on_device) {
    for (int d = 0; d < g->dim && rc == BMG_OK; d++) {
      rc = dalloc((void**)&dev[d], lf.ncell * sizeof(double));
      if (rc == BMG_OK && hipMemcpy(dev[d], u[d], lf.ncell * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(BMG_ERR_HIP, "copy of the velocity failed");
    }
  } else {
    for (int d = 0; d < g->dim; d++) dev[d] = u[d];
  }
  if (rc == BMG_OK)
    rc = g->dim == 2 ? viscosity_t<2>(g, dev, mu, mu_c, rho, rho_c, dt, sig, nrelax, on_device, stats, expl)
                     : viscosity_t<3>(g, dev, mu, mu_c, rho, rho_c, dt, sig, nrelax, on_device, stats, expl);
  if (!on_device) {
    for (int d = 0; d < g->dim; d++) {
      if (dev[d] && (rc == BMG_OK || rc == BMG_ERR_NAN) &&
          hipMemcpy(u[d], dev[d], lf.ncell * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = fail(BMG_ERR_HIP, "copy of the velocity failed");
      dfreet(dev[d]);
    }
  }
  return rc;
}

}  // namespace

extern "C" {

int bmg_viscosity(bmg_grid* g, double* u_x, double* u_y, double* u_z, const double* mu_x, const double* mu_y,
                  const double* mu_z, double mu_c, const double* rho, double rho_c, double dt, const int* sigma,
                  int nrelax, int on_device, bmg_mgstats* stats) {
  return viscosity_entry(g, u_x, u_y, u_z, mu_x, mu_y, mu_z, mu_c, rho, rho_c, dt, sigma, nrelax, on_device, stats,
                         false);
}

int bmg_viscosity_explicit(bmg_grid* g, double* u_x, double* u_y, double* u_z, const double* mu_x,
                           const double* mu_y, const double* mu_z, double mu_c, const double* rho, double rho_c,
                           double dt, const int* sigma, int on_device, bmg_mgstats* stats) {
  return viscosity_entry(g, u_x, u_y, u_z, mu_x, mu_y, mu_z, mu_c, rho, rho_c, dt, sigma, 0, on_device, stats, true);
}

}  // extern "C"

// ------------------------------------------------------------ dump / restore

namespace {

// struct DumpHeader, output.h:1049-1053 (coord = 3 doubles)
struct GridDumpHeader {
  double t;
  long len;
  int i, depth, npe, version;
  double n[3];
};
static_assert(sizeof(GridDumpHeader) == 56, "the reference's DumpHeader layout");
constexpr int kGridDumpVersion = 170901;

// a finest-level field restricted to every level on the GPU
// (restriction_average, children in foreach_child order), all levels copied
// to the host
template <int DIM>
int dump_levels(bmg_grid* g, const double* fine, std::vector<std::vector<double>>& host) {
  const int D = g->depth;
  std::vector<double*> dl(D + 1, nullptr);
  dl[D] = const_cast<double*>(fine);
  int rc = BMG_OK;
  for (int l = D - 1; l >= 0 && rc == BMG_OK; l--) {
    rc = dalloc((void**)&dl[l], g->L[l].ncell * sizeof(double));
    if (rc == BMG_OK) {
      Geo<DIM> gc{g->L[l].n, g->L[l].log2n};
      hipLaunchKernelGGL((k_restrict<DIM, double>), dim3(nblocks(g->L[l].ncell)), dim3(kBlock), 0, g->stream,
                         dl[l + 1], dl[l], gc);
      rc = check_launch();
    }
  }
  host.assign(D + 1, {});
  for (int l = 0; l <= D && rc == BMG_OK; l++) {
    host[l].resize(g->L[l].ncell);
    if (hipMemcpyAsync(host[l].data(), dl[l], g->L[l].ncell * sizeof(double), hipMemcpyDeviceToHost, g->stream) !=
        hipSuccess)
      rc = fail(BMG_ERR_HIP, "dump: copy failed");
  }
  if (rc == BMG_OK && hipStreamSynchronize(g->stream) != hipSuccess) rc = fail(BMG_ERR_HIP, "dump: sync failed");
  for (int l = 0; l < D; l++) dfreet(dl[l]);
  return rc;
}

size_t grid_index(int dim, int l, const int c[3]) {
  size_t n = (size_t)1 << l, idx = 0;
  for (int d = 0; d < dim; d++) idx = idx * n + (size_t)c[d];
  return idx;
}

// foreach_cell() on the multigrid: every cell of every level in pre-order
template <typename F>
int grid_preorder(int dim, int depth, int l, const int c[3], F&& f) {
  TRY(f(l, c));
  if (l == depth) return BMG_OK;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < (dim >= 2 ? 2 : 1); b++)
      for (int e = 0; e < (dim == 3 ? 2 : 1); e++) {
        const int ch[3] = {2 * c[0] + a, dim >= 2 ? 2 * c[1] + b : 0, dim == 3 ? 2 * c[2] + e : 0};
        TRY(grid_preorder(dim, depth, l + 1, ch, f));
      }
  return BMG_OK;
}

}  // namespace

extern "C" {

int bmg_grid_dump(bmg_grid* g, const char* path, const int* fields, const char* const* names, int nfields,
                  double time, int iter) {
  TRY(check_grid(g));
  if (!path || nfields < 0 || (nfields && (!fields || !names))) return fail(BMG_ERR_ARG, "null argument");
  for (int m = 0; m < nfields; m++)
    if (!names[m] || (fields[m] != BMG_F_A && fields[m] != BMG_F_B))
      return fail(BMG_ERR_ARG, "dump: field %d is not a or b", m);
  const int D = g->depth, dim = g->dim;
  std::vector<std::vector<std::vector<double>>> lv(nfields);
  for (int m = 0; m < nfields; m++) {
    const double* f = fields[m] == BMG_F_A ? g->a : g->b;
    TRY(dim == 1 ? dump_levels<1>(g, f, lv[m]) : dim == 2 ? dump_levels<2>(g, f, lv[m]) : dump_levels<3>(g, f, lv[m]));
  }
  std::vector<double> size(D + 1, 1.);
  for (int l = D - 1; l >= 0; l--) size[l] = 1. + (double)(1 << dim) * size[l + 1];
  std::string tmp = std::string(path) + "~";
  FILE* fp = fopen(tmp.c_str(), "wb");
  if (!fp) return fail(BMG_ERR_ARG, "dump: cannot write %s", tmp.c_str());
  GridDumpHeader h;
  memset(&h, 0, sizeof h);
  h.t = time;
  h.len = nfields + 1;
  h.i = iter;
  h.depth = D;
  h.npe = 1;
  h.version = kGridDumpVersion;
  bool ok = fwrite(&h, sizeof h, 1, fp) == 1;
  auto name = [&](const char* nm) {
    unsigned len = (unsigned)strlen(nm);
    ok = ok && fwrite(&len, sizeof len, 1, fp) == 1 && fwrite(nm, 1, len, fp) == len;
  };
  name("size");
  for (int m = 0; m < nfields; m++) name(names[m]);
  const double o[4] = {g->X0[0], g->X0[1], g->X0[2], g->L0};
  ok = ok && fwrite(o, sizeof(double), 4, fp) == 4;
  std::vector<char> buf;
  const size_t rl = 4 + 8 * (size_t)(nfields + 1);
  buf.reserve(rl * 65536 + rl);
  const int root[3] = {0, 0, 0};
  (void)grid_preorder(dim, D, 0, root, [&](int l, const int c[3]) {
    const unsigned flags = l == D ? 2u : 0u;
    const char* fb = reinterpret_cast<const char*>(&flags);
    buf.insert(buf.end(), fb, fb + 4);
    const char* sb = reinterpret_cast<const char*>(&size[l]);
    buf.insert(buf.end(), sb, sb + 8);
    const size_t idx = grid_index(dim, l, c);
    for (int m = 0; m < nfields; m++) {
      const char* vb = reinterpret_cast<const char*>(&lv[m][l][idx]);
      buf.insert(buf.end(), vb, vb + 8);
    }
    if (buf.size() >= rl * 65536) {
      ok = ok && fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
      buf.clear();
    }
    return BMG_OK;
  });
  ok = ok && (buf.empty() || fwrite(buf.data(), 1, buf.size(), fp) == buf.size());
  ok = (fclose(fp) == 0) && ok;
  if (!ok) {
    remove(tmp.c_str());
    return fail(BMG_ERR_ARG, "dump: write error on %s", tmp.c_str());
  }
  if (rename(tmp.c_str(), path) != 0) return fail(BMG_ERR_ARG, "dump: cannot rename %s", tmp.c_str());
  return BMG_OK;
}

int bmg_grid_restore(bmg_grid* g, const char* path, const int* fields, const char* const* names, int nfields,
                     double* time, int* iter) {
  TRY(check_grid(g));
  if (!path || nfields < 0 || (nfields && (!fields || !names))) return fail(BMG_ERR_ARG, "null argument");
  for (int m = 0; m < nfields; m++)
    if (!names[m] || (fields[m] != BMG_F_A && fields[m] != BMG_F_B))
      return fail(BMG_ERR_ARG, "restore: field %d is not a or b", m);
  FILE* fp = fopen(path, "rb");
  if (!fp) return fail(BMG_ERR_ARG, "restore: cannot open %s", path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(fp, fclose);
  GridDumpHeader h;
  if (fread(&h, sizeof h, 1, fp) != 1) return fail(BMG_ERR_ARG, "restore: expecting header");
  if (h.version != kGridDumpVersion)
    return fail(BMG_ERR_UNSUPPORTED, "restore: file version %d (reads %d)", h.version, kGridDumpVersion);
  if (h.depth != g->depth) return fail(BMG_ERR_ARG, "restore: file depth %d, grid depth %d", h.depth, g->depth);
  if (h.len < 1 || h.len > 100000) return fail(BMG_ERR_ARG, "restore: bad field count %ld", h.len);
  std::vector<std::string> fnames;
  for (long m = 0; m < h.len; m++) {
    unsigned len;
    if (fread(&len, sizeof len, 1, fp) != 1 || len > 4096) return fail(BMG_ERR_ARG, "restore: expecting len");
    std::string nm(len, '\0');
    if (len && fread(&nm[0], 1, len, fp) != len) return fail(BMG_ERR_ARG, "restore: expecting a name");
    fnames.push_back(nm);
  }
  double o[4];
  if (fread(o, sizeof(double), 4, fp) != 4) return fail(BMG_ERR_ARG, "restore: expecting coordinates");
  if (!(o[3] > 0.)) return fail(BMG_ERR_ARG, "restore: bad box size");
  std::vector<int> cols(nfields, -1);
  for (int m = 0; m < nfields; m++)
    for (long q = 1; q < h.len; q++)
      if (fnames[q] == names[m]) cols[m] = (int)q;
  const int D = g->depth, dim = g->dim;
  const size_t nf = g->L[D].ncell;
  std::vector<std::vector<double>> fine(nfields, std::vector<double>(nf, 0.));
  std::vector<double> rec(h.len);
  const int root[3] = {0, 0, 0};
  TRY(grid_preorder(dim, D, 0, root, [&](int l, const int c[3]) -> int {
    unsigned flags;
    if (fread(&flags, sizeof flags, 1, fp) != 1 || fread(rec.data(), sizeof(double), h.len, fp) != (size_t)h.len)
      return fail(BMG_ERR_ARG, "restore: the file ends before the grid does");
    if (((flags & 2u) != 0) != (l == D)) return fail(BMG_ERR_ARG, "restore: leaf flags do not match a multigrid");
    if (l == D) {
      const size_t idx = grid_index(dim, l, c);
      for (int m = 0; m < nfields; m++)
        if (cols[m] >= 0) fine[m][idx] = rec[cols[m]];
    }
    return BMG_OK;
  }));
  char extra;
  if (fread(&extra, 1, 1, fp) == 1) return fail(BMG_ERR_ARG, "restore: the file holds more records than the grid");
  if (o[3] != g->L0 || o[0] != g->X0[0] || o[1] != g->X0[1] || o[2] != g->X0[2]) {
    g->L0 = o[3];
    for (int d = 0; d < 3; d++) g->X0[d] = o[d];
    drop_graphs(g);
  }
  for (int m = 0; m < nfields; m++) TRY(bmg_set_field(g, fields[m], D, fine[m].data(), 0));
  if (time) *time = h.t;
  if (iter) *iter = h.i;
  return BMG_OK;
}

}  // extern "C"
