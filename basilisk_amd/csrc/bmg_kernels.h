// Device kernels of libbmg (gfx950).  Every level of the hierarchy is a set
// of SoA arrays of n^DIM cells in C order (i slowest, k fastest), with no
// ghost cells in memory: a kernel that reads past the edge of a level maps
// the index itself, which reproduces the ghost values that the reference's
// box_boundary_level / periodic_boundary_level_x would have written
// (grid/multigrid.h:543-701):
//   - periodic dimension: wrap around (both ghost layers are copies);
//   - otherwise ghost = c + sigma * s(nearest interior cell), and an edge or
//     corner ghost = product of the sigmas of its sides times the diagonal
//     interior value, which is what the reference's direction-ordered fill
//     leaves there for the homogeneous conditions of the correction field.
// Face coefficients of component d are stored as the low face of each cell
// (n^DIM values) plus the n^(DIM-1) faces on the high wall.
//
// All arithmetic follows the reference's evaluation order; the library is
// compiled with -ffp-contract=off so that the FP64 path reproduces the CPU
// oracle bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bmg {

constexpr int kBlock = 256;

// homogeneous boundary description of one level
struct BCH {
  int periodic[3];
  int sigma[6];  // right, left, top, bottom, front, back
  int tree;      // edge/corner ghosts by the tree rule (see bilinear_at)
};

// inhomogeneous boundary of `a` on the finest level: ghost = c + sigma * s,
// with the cell size (x / Delta is computed as x * invD when Delta is a
// power of two, which rounds identically)
struct BCA {
  int periodic[3];
  int sigma[6];
  const double* c[6];  // n^(DIM-1) values per side, or null for 0
  double invD;
  int pow2;
};

// cell indices fit in 32 bits (at most 2^30 cells per level)
using idx_t = uint32_t;

// two adjacent values as one vector load/store
template <typename T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};


template <int DIM>
struct Geo {
  int n, log2n;
  __device__ __forceinline__ void decode(idx_t idx, int c[3]) const {
    const idx_t m = (idx_t)(n - 1);
    if (DIM == 3) {
      c[2] = (int)(idx & m);
      c[1] = (int)((idx >> log2n) & m);
      c[0] = (int)(idx >> (2 * log2n));
    } else if (DIM == 2) {
      c[2] = 0;
      c[1] = (int)(idx & m);
      c[0] = (int)(idx >> log2n);
    } else {
      c[0] = (int)idx;
      c[1] = c[2] = 0;
    }
  }
  __device__ __forceinline__ idx_t stride(int d) const {
    if (DIM == 3) return d == 0 ? (idx_t)1 << (2 * log2n) : d == 1 ? (idx_t)1 << log2n : 1u;
    if (DIM == 2) return d == 0 ? (idx_t)1 << log2n : 1u;
    return 1u;
  }
  __device__ __forceinline__ idx_t index(const int c[3]) const {
    if (DIM == 3) return ((((idx_t)c[0] << log2n) + (idx_t)c[1]) << log2n) + (idx_t)c[2];
    if (DIM == 2) return ((idx_t)c[0] << log2n) + (idx_t)c[1];
    return (idx_t)c[0];
  }
  // index on the wall normal to d (tangential coordinates in order)
  __device__ __forceinline__ idx_t wall(int d, const int c[3]) const {
    if (DIM == 3)
      return d == 0 ? ((idx_t)c[1] << log2n) + c[2] : d == 1 ? ((idx_t)c[0] << log2n) + c[2]
                                                           : ((idx_t)c[0] << log2n) + c[1];
    if (DIM == 2) return d == 0 ? (idx_t)c[1] : (idx_t)c[0];
    return 0u;
  }
  __device__ __forceinline__ idx_t ncell() const { return (idx_t)1 << (DIM * log2n); }
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of logical blocks so that neighbouring rows and planes of
// a stencil share one L2.
__device__ __forceinline__ idx_t logical_block() {
  idx_t b = blockIdx.x, nb = gridDim.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  return b;
}

// index of the neighbour of cell c along d (dir +1 or -1): the cell itself
// when the neighbour is a non-periodic ghost.  `ghost` tells which.
template <int DIM>
__device__ __forceinline__ idx_t nbr_index(const Geo<DIM>& g, const int* periodic, idx_t idx, const int c[3],
                                           int d, int dir, bool& ghost) {
  const idx_t s = g.stride(d);
  const bool inside = dir > 0 ? c[d] < g.n - 1 : c[d] > 0;
  const idx_t wrapoff = (idx_t)(g.n - 1) * s;
  const idx_t step = dir > 0 ? idx + s : idx - s;
  const idx_t wrapped = dir > 0 ? idx - wrapoff : idx + wrapoff;
  ghost = !inside && !periodic[d];
  return inside ? step : (periodic[d] ? wrapped : idx);
}

// value of field f at the neighbour of cell c along d (dir +1 or -1), with
// the homogeneous ghost mapping (branch-free: a ghost reads the cell itself
// and scales it by sigma)
template <int DIM, typename T>
__device__ __forceinline__ T nbr_h(const T* __restrict__ f, const Geo<DIM>& g, const BCH& bc, idx_t idx,
                                   const int c[3], int d, int dir, T self) {
  bool ghost;
  idx_t iq = nbr_index<DIM>(g, bc.periodic, idx, c, d, dir, ghost);
  T v = ghost ? self : f[iq];
  return ghost ? T(bc.sigma[2 * d + (dir < 0)]) * v : v;
}

// alpha on the high face of cell c along d (one load from a selected address)
template <int DIM, typename T>
__device__ __forceinline__ T alpha_hi(const T* __restrict__ lo, const T* __restrict__ wall, const Geo<DIM>& g,
                                      idx_t idx, const int c[3], int d) {
  const T* p = c[d] < g.n - 1 ? lo + idx + g.stride(d) : wall + g.wall(d, c);
  return *p;
}

template <typename T>
struct Coef {
  const T* lo[3];
  const T* wall[3];
  T alpha;        // used when lo[0] is null (constant alpha, unityf)
  const T* lam;   // null: constant lam_c
  T lam_c;
};

// n/d of the relax step, poisson.h:303-319
template <int DIM, typename T>
__device__ __forceinline__ T relax_value(const T* __restrict__ da, const T* __restrict__ res, const Coef<T>& k,
                                         const Geo<DIM>& g, const BCH& bc, idx_t idx, const int c[3], T self,
                                         T D2) {
  T lam = k.lam ? k.lam[idx] : k.lam_c;
  T nn = -D2 * res[idx], dd = -lam * D2;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    T ah, al;
    if (k.lo[0]) {
      ah = alpha_hi<DIM, T>(k.lo[d], k.wall[d], g, idx, c, d);
      al = k.lo[d][idx];
    } else {
      ah = al = k.alpha;
    }
    T ap = nbr_h<DIM, T>(da, g, bc, idx, c, d, +1, self);
    T am = nbr_h<DIM, T>(da, g, bc, idx, c, d, -1, self);
    nn += ah * ap + al * am;
    dd += ah + al;
  }
  return nn / dd;
}

// weighted Jacobi sweep, poisson.h:293-294, 303-328: out = (a + 2 n/d)/3
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_relax_jacobi(const T* __restrict__ da, T* __restrict__ out,
                                                         const T* __restrict__ res, Coef<T> k,
                                                         Geo<DIM> g, BCH bc, T D2) {
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int c[3];
  g.decode(idx, c);
  T self = da[idx];
  T cval = relax_value<DIM, T>(da, res, k, g, bc, idx, c, self, D2);
  out[idx] = (self + T(2) * cval) / T(3);
}

// --- the reference's default smoother: in-place Gauss-Seidel ------------
//
// relax() without -DJACOBI (poisson.h:296, 303-320) overwrites each cell as
// foreach_level visits it, i-major with j and then k inside
// (grid/multigrid.h:191-215), after one ghost fill (the ghosts keep the
// values of the sweep's start).  A cell therefore reads the new values of
// its neighbours at a lower index along any dimension and the old values of
// the others and of every ghost.  Those dependencies only link cells of
// neighbouring hyperplanes i+j+k = s-1 and s+1, so the sweep runs hyperplane
// by hyperplane, dold -> dnew, with the result of the serial loop bit for
// bit: lower neighbours come from dnew, upper ones and ghosts (mirrored or
// periodic) from dold.

// n/d of cell c in that order
template <int DIM, typename T>
__device__ __forceinline__ T gs_value(const T* __restrict__ dold, const T* dnew, const T* __restrict__ res,
                                      const Coef<T>& k, const Geo<DIM>& g, const BCH& bc, idx_t idx,
                                      const int c[3], T D2) {
  const T self = dold[idx];
  T lam = k.lam ? k.lam[idx] : k.lam_c;
  T nn = -D2 * res[idx], dd = -lam * D2;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    T ah, al;
    if (k.lo[0]) {
      ah = alpha_hi<DIM, T>(k.lo[d], k.wall[d], g, idx, c, d);
      al = k.lo[d][idx];
    } else {
      ah = al = k.alpha;
    }
    T ap = nbr_h<DIM, T>(dold, g, bc, idx, c, d, +1, self);
    T am = c[d] > 0 ? dnew[idx - g.stride(d)] : nbr_h<DIM, T>(dold, g, bc, idx, c, d, -1, self);
    nn += ah * ap + al * am;
    dd += ah + al;
  }
  return nn / dd;
}

// cells of hyperplane s: first index i in [i0, i1], (i - i0, j) flattened
// over rows of n in 3D (k = s - i - j, masked when outside the level)
template <int DIM>
__device__ __forceinline__ void gs_plane_range(const Geo<DIM>& g, int s, int& i0, idx_t& count) {
  const int hi = (DIM - 1) * (g.n - 1);
  i0 = s - hi > 0 ? s - hi : 0;
  const int i1 = s < g.n - 1 ? s : g.n - 1;
  count = i1 < i0 ? 0u : (idx_t)(i1 - i0 + 1) * (DIM == 3 ? (idx_t)g.n : 1u);
}

template <int DIM>
__device__ __forceinline__ bool gs_plane_cell(const Geo<DIM>& g, int s, int i0, idx_t t, int c[3]) {
  if (DIM == 3) {
    c[0] = i0 + (int)(t >> g.log2n);
    c[1] = (int)(t & (idx_t)(g.n - 1));
    c[2] = s - c[0] - c[1];
    return c[2] >= 0 && c[2] < g.n;
  }
  if (DIM == 2) {
    c[0] = i0 + (int)t;
    c[1] = s - c[0];
    c[2] = 0;
    return true;
  }
  c[0] = s;
  c[1] = c[2] = 0;
  return true;
}

// one hyperplane per launch (levels whose hyperplanes fill many workgroups)
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_gs_plane(const T* __restrict__ dold, T* dnew, const T* __restrict__ res,
                                                     Coef<T> k, Geo<DIM> g, BCH bc, T D2, int s) {
  int i0;
  idx_t count;
  gs_plane_range<DIM>(g, s, i0, count);
  const idx_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  int c[3];
  if (!gs_plane_cell<DIM>(g, s, i0, t, c)) return;
  const idx_t idx = g.index(c);
  dnew[idx] = gs_value<DIM, T>(dold, dnew, res, k, g, bc, idx, c, D2);
}

// the whole sweep in one workgroup, a barrier between hyperplanes (all
// waves of a workgroup share the CU's vector L1, so the barrier makes the
// new values of hyperplane s visible to hyperplane s+1)
template <int DIM, typename T>
__device__ __forceinline__ void gs_sweep_block(const T* __restrict__ dold, T* dnew, const T* __restrict__ res,
                                               const Coef<T>& k, const Geo<DIM>& g, const BCH& bc, T D2) {
  const int smax = DIM * (g.n - 1);
  for (int s = 0; s <= smax; s++) {
    int i0;
    idx_t count;
    gs_plane_range<DIM>(g, s, i0, count);
    for (idx_t t = threadIdx.x; t < count; t += (idx_t)blockDim.x) {
      int c[3];
      if (!gs_plane_cell<DIM>(g, s, i0, t, c)) continue;
      const idx_t idx = g.index(c);
      dnew[idx] = gs_value<DIM, T>(dold, dnew, res, k, g, bc, idx, c, D2);
    }
    __syncthreads();
  }
}

template <int DIM, typename T>
__global__ __launch_bounds__(1024) void k_gs_block(const T* __restrict__ dold, T* dnew, const T* __restrict__ res,
                                                   Coef<T> k, Geo<DIM> g, BCH bc, T D2) {
  gs_sweep_block<DIM, T>(dold, dnew, res, k, g, bc, D2);
}

// a full red-black sweep in one launch, dold -> dnew: a red cell is relaxed
// from the old field; a black cell relaxes its red neighbours on the fly
// (from the old field, exactly as the red pass does) and then itself from
// those values, so the result equals the two-colour sweep bit for bit.
// A ghost neighbour of a black cell is sigma times its own old value.
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_relax_rb_fused(const T* __restrict__ dold, T* __restrict__ dnew,
                                                           const T* __restrict__ res, Coef<T> k, Geo<DIM> g,
                                                           BCH bc, T D2) {
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int c[3];
  g.decode(idx, c);
  T self = dold[idx];
  if (((c[0] + c[1] + c[2]) & 1) == 0) {
    dnew[idx] = relax_value<DIM, T>(dold, res, k, g, bc, idx, c, self, D2);
    return;
  }
  T lam = k.lam ? k.lam[idx] : k.lam_c;
  T nn = -D2 * res[idx], dd = -lam * D2;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    T v[2];
#pragma unroll
    for (int side = 0; side < 2; side++) {
      int dir = side == 0 ? 1 : -1;
      int q[3] = {c[0], c[1], c[2]};
      q[d] += dir;
      bool ghost = false;
      if (q[d] < 0 || q[d] >= g.n) {
        if (bc.periodic[d]) q[d] = q[d] < 0 ? q[d] + g.n : q[d] - g.n;
        else ghost = true;
      }
      if (ghost) {
        v[side] = T(bc.sigma[2 * d + side]) * self;
      } else {
        idx_t iq = g.index(q);
        v[side] = relax_value<DIM, T>(dold, res, k, g, bc, iq, q, dold[iq], D2);
      }
    }
    T ah, al;
    if (k.lo[0]) {
      ah = alpha_hi<DIM, T>(k.lo[d], k.wall[d], g, idx, c, d);
      al = k.lo[d][idx];
    } else {
      ah = al = k.alpha;
    }
    nn += ah * v[0] + al * v[1];
    dd += ah + al;
  }
  dnew[idx] = nn / dd;
}

// one colour of a red-black Gauss-Seidel sweep, in place; one thread per
// cell of that colour (colour = parity of i+j+k)
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_relax_rb(T* __restrict__ da, const T* __restrict__ res,
                                                     Coef<T> k, Geo<DIM> g, BCH bc, T D2, int colour,
                                                     int reverse) {
  // reverse: walk the blocks from the far end, so that a pass starts on the
  // data the previous pass touched last (still in the memory-side cache)
  idx_t lb = logical_block();
  if (reverse) lb = gridDim.x - 1 - lb;
  idx_t t = lb * kBlock + threadIdx.x;
  idx_t half = g.n > 1 ? g.ncell() / 2 : (colour == 0 ? 1u : 0u);
  if (t >= half) return;
  idx_t idx;
  if (g.n > 1) {
    // rows along the last dimension hold n/2 cells of each colour
    idx_t row = t >> (g.log2n - 1);
    idx_t kk = t & (idx_t)((g.n >> 1) - 1);
    int c0[3];
    g.decode(row << g.log2n, c0);
    int par = DIM == 3 ? (c0[0] + c0[1]) & 1 : DIM == 2 ? c0[0] & 1 : 0;
    idx = (row << g.log2n) + 2 * kk + (idx_t)((colour + par) & 1);
  } else {
    idx = 0;
  }
  int c[3];
  g.decode(idx, c);
  T self = da[idx];
  da[idx] = relax_value<DIM, T>(da, res, k, g, bc, idx, c, self, D2);
}

// In-place colour pass over pairs of adjacent cells along the last
// dimension: every thread owns the aligned pair (k, k+1) of a row, loads the
// pair and the neighbouring rows' pairs as vectors, and updates the one cell
// of the pair that has the pass colour.  The arithmetic is relax_value's, in
// the same order; only the memory access differs from k_relax_rb (contiguous
// vector loads instead of every other element).
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_relax_rb_pair(T* __restrict__ da, const T* __restrict__ res,
                                                          Coef<T> k, Geo<DIM> g, BCH bc, T D2, int colour,
                                                          int reverse) {
  using V2 = typename Vec2<T>::type;
  idx_t lb = logical_block();
  if (reverse) lb = gridDim.x - 1 - lb;
  const idx_t p = lb * kBlock + threadIdx.x;
  if (p >= g.ncell() / 2) return;
  const idx_t i0 = 2 * p;
  int c[3];
  g.decode(i0, c);
  const int e = (colour + c[0] + (DIM == 3 ? c[1] : 0)) & 1;  // the last coordinate of i0 is even
  const idx_t t = i0 + e;
  c[DIM - 1] += e;
  const int n = g.n;
  auto pick = [&](const T* f, idx_t at) {
    V2 v = *reinterpret_cast<const V2*>(f + at);
    return e ? v.y : v.x;
  };
  const V2 pr = *reinterpret_cast<const V2*>(da + i0);
  const T self = e ? pr.y : pr.x;
  const T lam = k.lam ? k.lam[t] : k.lam_c;
  T nn = -D2 * pick(res, i0), dd = -lam * D2;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const idx_t s = g.stride(d);
    const bool last = d == DIM - 1;
    T ah, al;
    if (k.lo[0]) {
      if (!last) {
        al = pick(k.lo[d], i0);
        ah = c[d] < n - 1 ? pick(k.lo[d], i0 + s) : k.wall[d][g.wall(d, c)];
      } else {
        const V2 lp = *reinterpret_cast<const V2*>(k.lo[d] + i0);
        al = e ? lp.y : lp.x;
        ah = e == 0 ? lp.y : (c[d] < n - 1 ? k.lo[d][t + 1] : k.wall[d][g.wall(d, c)]);
      }
    } else {
      ah = al = k.alpha;
    }
    T ap, am;
    const T sp = T(bc.sigma[2 * d]) * self, sm = T(bc.sigma[2 * d + 1]) * self;
    const idx_t wrap = (idx_t)(n - 1) * s;
    if (!last) {
      ap = c[d] < n - 1 ? pick(da, i0 + s) : (bc.periodic[d] ? pick(da, i0 - wrap) : sp);
      am = c[d] > 0 ? pick(da, i0 - s) : (bc.periodic[d] ? pick(da, i0 + wrap) : sm);
    } else {
      ap = e == 0 ? pr.y : (c[d] < n - 1 ? da[t + 1] : (bc.periodic[d] ? da[t - wrap] : sp));
      am = e == 1 ? pr.x : (c[d] > 0 ? da[t - 1] : (bc.periodic[d] ? da[t + wrap] : sm));
    }
    nn += ah * ap + al * am;
    dd += ah + al;
  }
  da[t] = nn / dd;
}

// coarse = mean of the 2^DIM children, restriction_average
// (grid/multigrid-common.h:26-32) in foreach_child order (k fastest)
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(const T* __restrict__ fine, T* __restrict__ coarse,
                                                     Geo<DIM> gc) {
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gc.ncell()) return;
  int c[3];
  gc.decode(idx, c);
  Geo<DIM> gf{gc.n * 2, gc.log2n + 1};
  int f[3] = {2 * c[0], 2 * c[1], 2 * c[2]};
  idx_t b = gf.index(f);
  T sum = T(0);
  if (DIM == 1) {
    sum += fine[b];
    sum += fine[b + 1];
  } else if (DIM == 2) {
    idx_t sx = gf.stride(0);
    sum += fine[b];
    sum += fine[b + 1];
    sum += fine[b + sx];
    sum += fine[b + sx + 1];
  } else {
    idx_t sx = gf.stride(0), sy = gf.stride(1);
    sum += fine[b];
    sum += fine[b + 1];
    sum += fine[b + sy];
    sum += fine[b + sy + 1];
    sum += fine[b + sx];
    sum += fine[b + sx + 1];
    sum += fine[b + sx + sy];
    sum += fine[b + sx + sy + 1];
  }
  coarse[idx] = sum / T(1 << DIM);
}

// value of a coarse-level field at coarse coordinates q (possibly one cell
// outside), with the homogeneous ghost mapping
template <int DIM, typename T>
__device__ __forceinline__ T coarse_at(const T* __restrict__ f, const Geo<DIM>& g, const BCH& bc, int q[3]) {
  T sign = T(1);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    if (q[d] < 0) {
      if (bc.periodic[d]) q[d] += g.n;
      else { q[d] = 0; sign *= T(bc.sigma[2 * d + 1]); }
    } else if (q[d] >= g.n) {
      if (bc.periodic[d]) q[d] -= g.n;
      else { q[d] = g.n - 1; sign *= T(bc.sigma[2 * d]); }
    }
  }
  return sign * f[g.index(q)];
}

// index contributions of the two coarse cells along one dimension of the
// bilinear stencil: the parent p and its neighbour p+o (o = -1 or +1),
// mapped like a ghost cell (wrapped, or the parent itself with sign sigma)
template <int DIM, typename T>
__device__ __forceinline__ void axis_pair(const Geo<DIM>& gc, const BCH& bc, int d, int p, int o, idx_t& off0,
                                          idx_t& off1, T& sg, bool& gh) {
  const idx_t s = gc.stride(d);
  const int q = p + o;
  const bool out = q < 0 || q >= gc.n;
  const int qw = q < 0 ? q + gc.n : q - gc.n;
  const int qm = out ? (bc.periodic[d] ? qw : p) : q;
  gh = out && !bc.periodic[d];
  sg = gh ? T(bc.sigma[2 * d + (o < 0)]) : T(1);
  off0 = (idx_t)p * s;
  off1 = (idx_t)qm * s;
}

// a ghost cell outside along two or three dimensions, from the value v of
// the interior cell diagonal to it.  Multigrid (grid/multigrid.h:543-594,
// one reflection per direction): the product of the sigmas times v.  Tree
// (grid/tree.h:1168-1258): b1 + b2 - v, or b1 + b2 + b3 - 2 v, with
// b = sigma v per crossed side (sums of +-v, so exact in any order).
template <typename T>
__device__ __forceinline__ T corner2(T v, T sa, T sb, bool tree) {
  return tree ? (sb * v + sa * v) - v : (sa * sb) * v;
}
template <typename T>
__device__ __forceinline__ T corner3(T v, T sx, T sy, T sz, bool gx, bool gy, bool gz, bool tree) {
  int k = (int)gx + (int)gy + (int)gz;
  if (!tree || k < 2) return ((sx * sy) * sz) * v;
  if (k == 2) return ((gz ? sz * v : T(0)) + (gy ? sy * v : T(0)) + (gx ? sx * v : T(0))) - v;
  return ((sz * v + sy * v) + sx * v) - T(2) * v;
}

// bilinear value at fine cell c from the coarse level,
// grid/multigrid-common.h:124-140, child offsets grid/multigrid.h:175-188
template <int DIM, typename T>
__device__ __forceinline__ T bilinear_at(const T* __restrict__ coarse, const Geo<DIM>& gc, const BCH& bcc,
                                         const int c[3]) {
  idx_t x0, x1, y0 = 0, y1 = 0, z0 = 0, z1 = 0;
  T sx, sy = T(1), sz = T(1);
  bool gx, gy = false, gz = false;
  const bool tree = bcc.tree != 0;
  axis_pair<DIM, T>(gc, bcc, 0, c[0] >> 1, 2 * (c[0] & 1) - 1, x0, x1, sx, gx);
  if (DIM >= 2) axis_pair<DIM, T>(gc, bcc, 1, c[1] >> 1, 2 * (c[1] & 1) - 1, y0, y1, sy, gy);
  if (DIM >= 3) axis_pair<DIM, T>(gc, bcc, 2, c[2] >> 1, 2 * (c[2] & 1) - 1, z0, z1, sz, gz);
  if (DIM == 1) {
    return (T(3) * coarse[x0] + sx * coarse[x1]) / T(4);
  } else if (DIM == 2) {
    T c00 = coarse[x0 + y0];
    T c10 = sx * coarse[x1 + y0];
    T c01 = sy * coarse[x0 + y1];
    T c11 = (gx && gy) ? corner2(coarse[x1 + y1], sx, sy, tree) : (sx * sy) * coarse[x1 + y1];
    return (T(9) * c00 + T(3) * (c10 + c01) + c11) / T(16);
  } else {
    T c000 = coarse[x0 + y0 + z0];
    T c100 = sx * coarse[x1 + y0 + z0];
    T c010 = sy * coarse[x0 + y1 + z0];
    T c001 = sz * coarse[x0 + y0 + z1];
    T c110 = (gx && gy) ? corner2(coarse[x1 + y1 + z0], sx, sy, tree) : (sx * sy) * coarse[x1 + y1 + z0];
    T c101 = (gx && gz) ? corner2(coarse[x1 + y0 + z1], sx, sz, tree) : (sx * sz) * coarse[x1 + y0 + z1];
    T c011 = (gy && gz) ? corner2(coarse[x0 + y1 + z1], sy, sz, tree) : (sy * sz) * coarse[x0 + y1 + z1];
    T c111 = corner3(coarse[x1 + y1 + z1], sx, sy, sz, gx, gy, gz, tree);
    return (T(27) * c000 + T(9) * (c100 + c010 + c001) + T(3) * (c110 + c101 + c011) + c111) / T(64);
  }
}

template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_prolong(const T* __restrict__ coarse, T* __restrict__ fine,
                                                    Geo<DIM> gf, BCH bcc) {
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gf.ncell()) return;
  int c[3];
  gf.decode(idx, c);
  Geo<DIM> gc{gf.n >> 1, gf.log2n - 1};
  fine[idx] = bilinear_at<DIM, T>(coarse, gc, bcc, c);
}

// the same prolongation, one thread per coarse cell: its 2^DIM children
// (bilinear_at each, sharing the coarse loads), the two children along the
// last dimension written as one vector
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_prolong_children(const T* __restrict__ coarse, T* __restrict__ fine,
                                                             Geo<DIM> gf, BCH bcc) {
  using V2 = typename Vec2<T>::type;
  Geo<DIM> gc{gf.n >> 1, gf.log2n - 1};
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gc.ncell()) return;
  int cc[3];
  gc.decode(idx, cc);
  const int nb = DIM == 3 ? 4 : 2;  // children groups along the other dimensions
#pragma unroll
  for (int m = 0; m < nb; m++) {
    int c[3] = {2 * cc[0], 2 * cc[1], 2 * cc[2]};
    if (DIM == 3) {
      c[0] += m >> 1;
      c[1] += m & 1;
    } else {
      c[0] += m;
    }
    V2 v;
    v.x = bilinear_at<DIM, T>(coarse, gc, bcc, c);
    c[DIM - 1] += 1;
    v.y = bilinear_at<DIM, T>(coarse, gc, bcc, c);
    c[DIM - 1] -= 1;
    *reinterpret_cast<V2*>(fine + gf.index(c)) = v;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T* __restrict__ f, size_t n, T v) {
  size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) f[idx] = v;
}

// --- the reference's own storage (grid/multigrid.h:16-73) ------------------
//
// Level l of the reference holds (n+4)^DIM cell records (GHOSTS = 2) of
// `nrec` doubles each, in C order; field handle h is double h of a record.
// These kernels move one field between a device copy of such a record
// array and a dense box of extents e[0..DIM) (a cell field: n per
// dimension; a face component c: n+1 along c, face i of a row being the
// low face of cell i, so face n reads the record of the ghost cell n).
struct ViewBox {
  size_t nrec;   // doubles per record (datasize / 8)
  int h;         // field handle
  int m;         // n + 4
  int e[3];      // box extents (1 beyond DIM)
};

__device__ __forceinline__ size_t view_record(const ViewBox& B, int DIMS, size_t q, int c[3]) {
  const size_t e1 = (size_t)B.e[1], e2 = (size_t)B.e[2];
  c[2] = (int)(q % e2);
  c[1] = (int)((q / e2) % e1);
  c[0] = (int)(q / (e1 * e2));
  const size_t m = (size_t)B.m;
  if (DIMS == 3) return ((size_t)(c[0] + 2) * m + (size_t)(c[1] + 2)) * m + (size_t)(c[2] + 2);
  if (DIMS == 2) return (size_t)(c[0] + 2) * m + (size_t)(c[1] + 2);
  return (size_t)(c[0] + 2);
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_view_gather(const double* __restrict__ rec, ViewBox B,
                                                        double* __restrict__ out) {
  const size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= (size_t)B.e[0] * B.e[1] * B.e[2]) return;
  int c[3];
  out[q] = rec[view_record(B, DIM, q, c) * B.nrec + (size_t)B.h];
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_view_scatter(double* __restrict__ rec, ViewBox B,
                                                         const double* __restrict__ in) {
  const size_t q = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (q >= (size_t)B.e[0] * B.e[1] * B.e[2]) return;
  int c[3];
  rec[view_record(B, DIM, q, c) * B.nrec + (size_t)B.h] = in[q];
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_convert(const TI* __restrict__ in, TO* __restrict__ out, size_t n) {
  size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) out[idx] = (TO)in[idx];
}

// a += da on the finest level, poisson.h:84-89
template <typename T>
__global__ __launch_bounds__(kBlock) void k_correct(double* __restrict__ a, const T* __restrict__ da, size_t n) {
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx < n) a[idx] += (double)da[idx];
}

// --- coefficient gradients (adjoint) ----------------------------------
//
// With L(a) = lambda a + div(alpha grad a) (ghosts from the BCs of `a`) and
// the adjoint state psi solving the same operator with homogeneous BCs and
// right-hand side dJ/da, dJ/dp = -psi^T dL/dp at fixed a:
//   interior face between cells l and r:  (psi_r - psi_l)/D * (a_r - a_l)/D
//   wall face of cell i:                  -psi_i * ((ghost(a_i) - a_i)/D)/D
//   lambda of cell i:                     -psi_i a_i
// One thread per face of component c, in the user layout (n+1 along c).

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_coef_grad(const double* __restrict__ a, const double* __restrict__ psi,
                                                      double* __restrict__ out, Geo<DIM> g, BCA bc, int c,
                                                      double D) {
  size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t nf = (size_t)(g.n + 1);
  for (int d = 1; d < DIM; d++) nf *= (size_t)g.n;
  if (t >= nf) return;
  int q[3] = {0, 0, 0};
  size_t r = t;
  for (int d = DIM - 1; d >= 0; d--) {
    size_t m = (size_t)(d == c ? g.n + 1 : g.n);
    q[d] = (int)(r % m);
    r /= m;
  }
  const int ic = q[c];
  int cl[3] = {q[0], q[1], q[2]}, cr[3] = {q[0], q[1], q[2]};
  cl[c] = ic - 1;
  double v;
  if (ic > 0 && ic < g.n) {
    idx_t il = g.index(cl), ir = g.index(cr);
    v = ((psi[ir] - psi[il]) / D) * ((a[ir] - a[il]) / D);
  } else {
    const bool low = ic == 0;
    const int* cc = low ? cr : cl;
    idx_t i = g.index(cc);
    int side = 2 * c + (low ? 1 : 0);
    const double* cv = bc.c[side];
    double ghost = (cv ? cv[g.wall(c, cc)] : 0.) + (double)bc.sigma[side] * a[i];
    v = -psi[i] * (((ghost - a[i]) / D) / D);
  }
  out[t] = v;
}

__global__ __launch_bounds__(kBlock) void k_lambda_grad(const double* __restrict__ a, const double* __restrict__ psi,
                                                        double* __restrict__ out, size_t n) {
  size_t t = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (t < n) out[t] = -psi[t] * a[t];
}

// --- reductions -------------------------------------------------------

// statistics slots written by the residual kernel
struct Stat {
  unsigned long long maxbits;  // max |res| as the bits of a non-negative double
  unsigned long long nans;     // number of NaN residuals
};

__global__ void k_stat_reset(Stat* s) {
  if (threadIdx.x == 0) {
    s->maxbits = 0ull;
    s->nans = 0ull;
  }
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// max over the block, then one atomic per block; NaN never wins a max
// (like `if (fabs(res[]) > maxres)` in poisson.h:386) but is counted
__device__ __forceinline__ void block_max_to(Stat* st, double v) {
  __shared__ double part[kBlock / 64];
  v = wave_max(v);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) part[w] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = part[0];
    for (int i = 1; i < kBlock / 64; i++) m = fmax(m, part[i]);
    if (m > 0.) atomicMax(&st->maxbits, (unsigned long long)__double_as_longlong(m));
  }
}

// residual of one cell of `a` on the finest level, poisson.h:373-389
// (!TREE): b - lambda a + sum_d (alpha[] (a[] - a[-1])/D - alpha[1] (a[1] - a[])/D)/D,
// with the inhomogeneous ghost of `a` (c + sigma a) on the walls
template <int DIM>
__device__ __forceinline__ double residual_cell(const double* __restrict__ a, const double* __restrict__ b,
                                                const Coef<double>& k, const Geo<DIM>& g, const BCA& bc,
                                                idx_t idx, const int c[3], double D) {
  const double a0 = a[idx];
  const double lam = k.lam ? k.lam[idx] : k.lam_c;
  double r = b[idx] - lam * a0;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    bool gp, gm;
    const idx_t ip = nbr_index<DIM>(g, bc.periodic, idx, c, d, +1, gp);
    const idx_t im = nbr_index<DIM>(g, bc.periodic, idx, c, d, -1, gm);
    const double vp = a[ip], vm = a[im];
    double cp = 0., cm = 0.;
    if (gp && bc.c[2 * d]) cp = bc.c[2 * d][g.wall(d, c)];
    if (gm && bc.c[2 * d + 1]) cm = bc.c[2 * d + 1][g.wall(d, c)];
    const double ap = gp ? cp + double(bc.sigma[2 * d]) * vp : vp;
    const double am = gm ? cm + double(bc.sigma[2 * d + 1]) * vm : vm;
    double ah, al;
    if (k.lo[0]) {
      ah = alpha_hi<DIM, double>(k.lo[d], k.wall[d], g, idx, c, d);
      al = k.lo[d][idx];
    } else {
      ah = al = k.alpha;
    }
    if (bc.pow2) r += (al * ((a0 - am) * bc.invD) - ah * ((ap - a0) * bc.invD)) * bc.invD;
    else r += (al * ((a0 - am) / D) - ah * ((ap - a0) / D)) / D;
  }
  return r;
}

// residual of `a` on the finest level, writes res in the cycle precision TO
template <int DIM, typename TO>
__global__ __launch_bounds__(kBlock) void k_residual(const double* __restrict__ a, const double* __restrict__ b,
                                                     TO* __restrict__ res, double* __restrict__ res64,
                                                     Coef<double> k, Geo<DIM> g, BCA bc, double D, Stat* st) {
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  double m = 0.;
  if (idx < g.ncell()) {
    int c[3];
    g.decode(idx, c);
    double r = residual_cell<DIM>(a, b, k, g, bc, idx, c, D);
    res[idx] = (TO)r;
    if (res64) res64[idx] = r;
    if (r != r) atomicAdd(&st->nans, 1ull);
    else m = fabs(r);
  }
  block_max_to(st, m);
}

// deterministic two-stage sum (sum of b, poisson.h:163-167)
__global__ __launch_bounds__(kBlock) void k_sum_partial(const double* __restrict__ x, size_t n,
                                                        double* __restrict__ part) {
  __shared__ double sh[kBlock / 64];
  double v = 0.;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) v += x[i];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.;
    for (int i = 0; i < kBlock / 64; i++) s += sh[i];
    part[blockIdx.x] = s;
  }
}

__global__ void k_sum_final(const double* __restrict__ part, int np, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.;
    for (int i = 0; i < np; i++) s += part[i];
    *out = s;
  }
}

// --- project(), poisson.h:469-509 ----------------------------------------
//
// Face velocities in the user layout: component d has n+1 faces along d,
// C order.

template <int DIM>
__device__ __forceinline__ size_t uface(int n, int d, const int c[3]) {
  const size_t e1 = DIM >= 2 ? (size_t)n + (d == 1) : 1, e2 = DIM >= 3 ? (size_t)n + (d == 2) : 1;
  return ((size_t)c[0] * e1 + c[1]) * e2 + c[2];
}

// div = 0; div += uf.x[1] - uf.x[] per dimension (x, y, z); div /= dt*Delta
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_divergence(const double* __restrict__ u0, const double* __restrict__ u1,
                                                       const double* __restrict__ u2, double* __restrict__ b,
                                                       Geo<DIM> g, double dtD) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int c[3];
  g.decode(idx, c);
  const double* u[3] = {u0, u1, u2};
  double div = 0.;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    int h[3] = {c[0], c[1], c[2]};
    h[d] += 1;
    div += u[d][uface<DIM>(g.n, d, h)] - u[d][uface<DIM>(g.n, d, c)];
  }
  b[idx] = div / dtD;
}

// uf.x[] -= dt*alpha.x[]*face_gradient_x(p, 0) on every face of component
// d, walls included; face_gradient_x(a,i) = (a[i] - a[i-1])/Delta
// (common.h:1288) with the ghosts of p = a from its boundary conditions
// (c + sigma s, or the periodic copy)
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_face_update(double* __restrict__ u, const double* __restrict__ a,
                                                        const double* __restrict__ alo,
                                                        const double* __restrict__ awall, double alpha_c,
                                                        Geo<DIM> g, BCA bc, int d, double dt, double D) {
  const int n = g.n;
  const size_t e1 = DIM >= 2 ? (size_t)n + (d == 1) : 1, e2 = DIM >= 3 ? (size_t)n + (d == 2) : 1;
  const size_t e0 = (size_t)n + (d == 0);
  const size_t f = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (f >= e0 * e1 * e2) return;
  int c[3] = {(int)(f / (e1 * e2)), (int)((f / e2) % e1), (int)(f % e2)};
  const int i = c[d];
  int q[3] = {c[0], c[1], c[2]};
  q[d] = i < n ? i : n - 1;
  const double ahi = a[g.index(q)];
  q[d] = i > 0 ? i - 1 : 0;
  const double alow = a[g.index(q)];
  const idx_t w = g.wall(d, c);
  double ph, pl;
  if (i < n) {
    ph = ahi;
  } else if (bc.periodic[d]) {
    q[d] = 0;
    ph = a[g.index(q)];
  } else {
    const double cv = bc.c[2 * d] ? bc.c[2 * d][w] : 0.;
    ph = cv + double(bc.sigma[2 * d]) * alow;
  }
  if (i > 0) {
    pl = alow;
  } else if (bc.periodic[d]) {
    q[d] = n - 1;
    pl = a[g.index(q)];
  } else {
    const double cv = bc.c[2 * d + 1] ? bc.c[2 * d + 1][w] : 0.;
    pl = cv + double(bc.sigma[2 * d + 1]) * ahi;
  }
  double alpha = alpha_c;
  if (alo) {
    q[d] = i;
    alpha = i < n ? alo[g.index(q)] : awall[w];
  }
  u[f] = u[f] - dt * alpha * ((ph - pl) / D);
}

// --- the projection event of navier-stokes/centered.h ------------------
//
// centered_gradient (centered.h:387-406): on every face
//   gf.x[] = fm.x[]*a.x[] - alpha.x[]*(p[] - p[-1])/Delta
// (fm = unityf, a the face acceleration, p the grid's `a` with its
// boundary conditions), then in every cell
//   g.x[] = (gf.x[] + gf.x[1])/(fm.x[] + fm.x[1] + SEPS)
// with SEPS = 0 (common.h:1287).  A cell evaluates its two faces with the
// expression the face loop uses, so shared faces get identical values.

// p on both sides of the face with index i (0..n) along d of the row of
// cell c: the ghost values at the walls (c + sigma s) or the periodic copy
template <int DIM>
__device__ __forceinline__ void face_sides(const double* __restrict__ p, const Geo<DIM>& g, const BCA& bc, int d,
                                           const int c[3], int i, double& ph, double& pl) {
  const int n = g.n;
  int q[3] = {c[0], c[1], c[2]};
  q[d] = i < n ? i : n - 1;
  const double ahi = p[g.index(q)];
  q[d] = i > 0 ? i - 1 : 0;
  const double alow = p[g.index(q)];
  const idx_t w = g.wall(d, c);
  if (i < n) {
    ph = ahi;
  } else if (bc.periodic[d]) {
    q[d] = 0;
    ph = p[g.index(q)];
  } else {
    const double cv = bc.c[2 * d] ? bc.c[2 * d][w] : 0.;
    ph = cv + double(bc.sigma[2 * d]) * alow;
  }
  if (i > 0) {
    pl = alow;
  } else if (bc.periodic[d]) {
    q[d] = n - 1;
    pl = p[g.index(q)];
  } else {
    const double cv = bc.c[2 * d + 1] ? bc.c[2 * d + 1][w] : 0.;
    pl = cv + double(bc.sigma[2 * d + 1]) * ahi;
  }
}

// index of face i along d in the user face layout (n+1 along d)
template <int DIM>
__device__ __forceinline__ size_t uface_at(int n, int d, const int c[3], int i) {
  int q[3] = {c[0], c[1], c[2]};
  q[d] = i;
  size_t e1 = DIM >= 2 ? (size_t)n + (d == 1) : 1, e2 = DIM >= 3 ? (size_t)n + (d == 2) : 1;
  return ((size_t)q[0] * e1 + (size_t)(DIM >= 2 ? q[1] : 0)) * e2 + (size_t)(DIM >= 3 ? q[2] : 0);
}

struct CenteredArgs {
  const double* p;
  const double* alo[3];
  const double* awall[3];
  double alpha_c;
  const double* acc[3];  // face acceleration (user face layout) or null for zerof
  double* gout[3];
};

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_centered_gradient(CenteredArgs A, Geo<DIM> g, BCA bc, double D) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int c[3];
  g.decode(idx, c);
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    double gf[2];
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int i = c[d] + h;
      double ph, pl;
      face_sides<DIM>(A.p, g, bc, d, c, i, ph, pl);
      double alpha = A.alpha_c;
      if (A.alo[0]) {
        int q[3] = {c[0], c[1], c[2]};
        q[d] = i;
        alpha = i < g.n ? A.alo[d][g.index(q)] : A.awall[d][g.wall(d, c)];
      }
      const double acc = A.acc[d] ? A.acc[d][uface_at<DIM>(g.n, d, c, i)] : 0.;
      gf[h] = 1. * acc - alpha * (ph - pl) / D;
    }
    A.gout[d][idx] = (gf[0] + gf[1]) / (1. + 1. + 0.);
  }
}

// correction (centered.h:322-327): u.x[] += dt*g.x[]
__global__ __launch_bounds__(kBlock) void k_correction(double* __restrict__ u, const double* __restrict__ gx,
                                                       double dt, size_t n) {
  const size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) u[idx] += dt * gx[idx];
}

// --- face coefficients (FP64, once per poisson() call) ------------------

// coarse low faces of component d: face_average, grid/multigrid-common.h:42-58,
// with the tangential order that the reference's dimension rotation gives
// each component
template <int DIM>
__device__ __forceinline__ void restrict_face_lo_at(const double* __restrict__ fine, double* __restrict__ coarse,
                                                    const Geo<DIM>& gc, int d, idx_t idx) {
  int c[3];
  gc.decode(idx, c);
  Geo<DIM> gf{gc.n * 2, gc.log2n + 1};
  int f[3] = {2 * c[0], 2 * c[1], 2 * c[2]};
  size_t b = gf.index(f);
  if (DIM == 1) {
    coarse[idx] = fine[b];
  } else if (DIM == 2) {
    size_t t = gf.stride(1 - d);  // the tangential dimension
    coarse[idx] = (fine[b] + fine[b + t]) / 2.;
  } else {
    // order of the two tangential offsets (u then v) per component:
    // x: (j,k) = 00,10,01,11; y: (i,k) = 00,01,10,11 -> k first; z: (i,j) = 00,10,01,11
    size_t su, sv;
    if (d == 0) { su = gf.stride(1); sv = gf.stride(2); }
    else if (d == 1) { su = gf.stride(2); sv = gf.stride(0); }
    else { su = gf.stride(0); sv = gf.stride(1); }
    coarse[idx] = (fine[b] + fine[b + su] + fine[b + sv] + fine[b + su + sv]) / 4.;
  }
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_restrict_face_lo(const double* __restrict__ fine, double* __restrict__ coarse,
                                                             Geo<DIM> gc, int d) {
  idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gc.ncell()) return;
  restrict_face_lo_at<DIM>(fine, coarse, gc, d, idx);
}

// coarse high-wall faces of component d from the fine high-wall faces
template <int DIM>
__device__ __forceinline__ void restrict_face_wall_at(const double* __restrict__ fine, double* __restrict__ coarse,
                                                      int nc, int d, size_t idx) {
  int nf = 2 * nc;
  if (DIM == 1) {
    coarse[0] = fine[0];
  } else if (DIM == 2) {
    coarse[idx] = (fine[2 * idx] + fine[2 * idx + 1]) / 2.;
  } else {
    // wall coordinates (t0, t1) in (i,j,k) order with t1 fastest
    int t0 = (int)(idx / nc), t1 = (int)(idx % nc);
    size_t b = (size_t)(2 * t0) * nf + 2 * t1;
    size_t s0 = nf, s1 = 1;
    // x: tangential (j,k): u=j (s0), v=k (s1); y: (i,k): u=k (s1), v=i (s0);
    // z: (i,j): u=i (s0), v=j (s1)
    size_t su = d == 1 ? s1 : s0, sv = d == 1 ? s0 : s1;
    coarse[idx] = (fine[b] + fine[b + su] + fine[b + sv] + fine[b + su + sv]) / 4.;
  }
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_restrict_face_wall(const double* __restrict__ fine, double* __restrict__ coarse,
                                                               int nc, int d) {
  size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t nw = DIM == 3 ? (size_t)nc * nc : DIM == 2 ? (size_t)nc : 1;
  if (idx >= nw) return;
  restrict_face_wall_at<DIM>(fine, coarse, nc, d, idx);
}

// every coefficient restriction from level l+1 to level l in one launch
// (one thread per coarse cell; nwall <= ncell): lambda (restriction_average),
// the low faces and the high walls of each component -- the arithmetic of
// k_restrict, k_restrict_face_lo and k_restrict_face_wall, element for element
struct CoefLevelArgs {
  const double* flo[3];  // null: constant alpha
  double* clo[3];
  const double* fwall[3];
  double* cwall[3];
  const double* flam;  // null: constant lambda
  double* clam;
};

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_restrict_coef_level(CoefLevelArgs A, Geo<DIM> gc) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gc.ncell()) return;
  if (A.flam) {
    int c[3];
    gc.decode(idx, c);
    Geo<DIM> gf{gc.n * 2, gc.log2n + 1};
    int f[3] = {2 * c[0], 2 * c[1], 2 * c[2]};
    idx_t b = gf.index(f);
    double sum = 0.;
    if (DIM == 1) {
      sum += A.flam[b];
      sum += A.flam[b + 1];
    } else if (DIM == 2) {
      idx_t sx = gf.stride(0);
      sum += A.flam[b];
      sum += A.flam[b + 1];
      sum += A.flam[b + sx];
      sum += A.flam[b + sx + 1];
    } else {
      idx_t sx = gf.stride(0), sy = gf.stride(1);
      sum += A.flam[b];
      sum += A.flam[b + 1];
      sum += A.flam[b + sy];
      sum += A.flam[b + sy + 1];
      sum += A.flam[b + sx];
      sum += A.flam[b + sx + 1];
      sum += A.flam[b + sx + sy];
      sum += A.flam[b + sx + sy + 1];
    }
    A.clam[idx] = sum / double(1 << DIM);
  }
  if (A.flo[0]) {
    const size_t nw = DIM == 3 ? (size_t)gc.n * gc.n : DIM == 2 ? (size_t)gc.n : 1;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      restrict_face_lo_at<DIM>(A.flo[d], A.clo[d], gc, d, idx);
      if (idx < nw) restrict_face_wall_at<DIM>(A.fwall[d], A.cwall[d], gc.n, d, idx);
    }
  }
}

// FP64 -> FP32 copies of many arrays in one launch: segment s covers blocks
// first[s] .. first[s+1]-1, kConvItems elements per thread (strided by the
// block size, so each load instruction stays coalesced)
constexpr int kConvSegs = 32;
constexpr int kConvItems = 8;
struct ConvBatch {
  const double* src[kConvSegs];
  float* dst[kConvSegs];
  unsigned n[kConvSegs];
  unsigned first[kConvSegs + 1];
  int nseg;
};

__global__ __launch_bounds__(kBlock) void k_convert_batch(ConvBatch B) {
  const unsigned blk = blockIdx.x;
  // the segment of this block: the number of later segments starting at or
  // before it (independent loads of the increasing starts, not a chain)
  int s = 0;
#pragma unroll
  for (int q = 1; q < kConvSegs; q++) s += (q < B.nseg && blk >= B.first[q]) ? 1 : 0;
  const size_t i0 = (size_t)(blk - B.first[s]) * (kBlock * kConvItems) + threadIdx.x;
  const double* __restrict__ src = B.src[s];
  float* __restrict__ dst = B.dst[s];
  const size_t n = B.n[s];
  double v[kConvItems];
#pragma unroll
  for (int k = 0; k < kConvItems; k++) {
    const size_t i = i0 + (size_t)k * kBlock;
    v[k] = i < n ? src[i] : 0.;
  }
#pragma unroll
  for (int k = 0; k < kConvItems; k++) {
    const size_t i = i0 + (size_t)k * kBlock;
    if (i < n) dst[i] = (float)v[k];
  }
}

// wall faces of component d copied from the first plane of low faces (the
// periodic copy of face vectors on coarse levels)
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_wall_from_first(const double* __restrict__ lo, double* __restrict__ wall,
                                                            Geo<DIM> g, int d) {
  size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t nw = DIM == 3 ? (size_t)g.n * g.n : DIM == 2 ? (size_t)g.n : 1;
  if (idx >= nw) return;
  int c[3] = {0, 0, 0};
  if (DIM == 3) {
    int t0 = (int)(idx / g.n), t1 = (int)(idx % g.n);
    if (d == 0) { c[1] = t0; c[2] = t1; }
    else if (d == 1) { c[0] = t0; c[2] = t1; }
    else { c[0] = t0; c[1] = t1; }
  } else if (DIM == 2) {
    c[1 - d] = (int)idx;
  }
  c[d] = 0;
  wall[idx] = lo[g.index(c)];
}

// --- residual fused with the first restriction ---------------------------
//
// One block covers a brick of 2 x 2 x 64 cells (3D) or 2 x 128 (2D), i.e.
// whole groups of 2^DIM children: it computes their residuals, writes them,
// and writes the 32 (3D) or 64 (2D) coarse residuals that restriction_average
// makes of them, summed in foreach_child order.  The per-block max |res| and
// NaN count go to `part`; k_reduce_stat folds them (no contended atomics).

struct Part {
  double max;
  unsigned long long nans;
};

template <int DIM, typename TO>
__global__ __launch_bounds__(kBlock) void k_residual_restrict(const double* __restrict__ a, const double* __restrict__ b,
                                                              TO* __restrict__ res, TO* __restrict__ coarse,
                                                              double* __restrict__ res64, Coef<double> k, Geo<DIM> g,
                                                              BCA bc, double D, Part* __restrict__ part) {
  __shared__ TO sh[kBlock];
  __shared__ double smax[kBlock / 64];
  __shared__ unsigned long long snan[kBlock / 64];
  const int t = threadIdx.x;
  const idx_t nbricks = g.ncell() / kBlock;
  double macc = 0.;
  unsigned long long nacc = 0;
  // grid-stride over the bricks (the grid is capped, one partial per block)
  for (idx_t bid = logical_block(); bid < nbricks; bid += gridDim.x) {
    int c[3];
    if (DIM == 3) {
      int nkb = g.n >> 6, nj2 = g.n >> 1;
      int bk = (int)(bid % nkb), bj = (int)((bid / nkb) % nj2), bi = (int)(bid / ((idx_t)nkb * nj2));
      c[0] = 2 * bi + (t >> 7);
      c[1] = 2 * bj + ((t >> 6) & 1);
      c[2] = 64 * bk + (t & 63);
    } else {
      int nkb = g.n >> 7;
      int bk = (int)(bid % nkb), bi = (int)(bid / nkb);
      c[0] = 2 * bi + (t >> 7);
      c[1] = 128 * bk + (t & 127);
      c[2] = 0;
    }
    idx_t idx = g.index(c);
    double r = residual_cell<DIM>(a, b, k, g, bc, idx, c, D);
    TO ro = (TO)r;
    res[idx] = ro;
    if (res64) res64[idx] = r;
    if (r != r) nacc++;
    else macc = fmax(macc, fabs(r));
    if (coarse) {
      sh[t] = ro;
      __syncthreads();
      Geo<DIM> gc{g.n >> 1, g.log2n - 1};
      if (DIM == 3 && t < 32) {
        // children (a,b,c) of coarse cell t: sh[a*128 + b*64 + 2t + c]
        TO sum = TO(0);
        sum += sh[2 * t];
        sum += sh[2 * t + 1];
        sum += sh[64 + 2 * t];
        sum += sh[64 + 2 * t + 1];
        sum += sh[128 + 2 * t];
        sum += sh[128 + 2 * t + 1];
        sum += sh[192 + 2 * t];
        sum += sh[192 + 2 * t + 1];
        int cc[3] = {c[0] >> 1, c[1] >> 1, (c[2] & ~63) / 2 + t};
        coarse[gc.index(cc)] = sum / TO(8);
      } else if (DIM == 2 && t < 64) {
        TO sum = TO(0);
        sum += sh[2 * t];
        sum += sh[2 * t + 1];
        sum += sh[128 + 2 * t];
        sum += sh[128 + 2 * t + 1];
        int cc[3] = {c[0] >> 1, (c[1] & ~127) / 2 + t, 0};
        coarse[gc.index(cc)] = sum / TO(4);
      }
      __syncthreads();
    }
  }
  macc = wave_max(macc);
  for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o, 64);
  int lane = t & 63, w = t >> 6;
  if (lane == 0) {
    smax[w] = macc;
    snan[w] = nacc;
  }
  __syncthreads();
  if (t == 0) {
    double mm = smax[0];
    unsigned long long ns = snan[0];
    for (int i = 1; i < kBlock / 64; i++) {
      mm = fmax(mm, smax[i]);
      ns += snan[i];
    }
    part[blockIdx.x] = Part{mm, ns};
  }
}

__global__ __launch_bounds__(1024) void k_reduce_stat(const Part* __restrict__ part, int np, Stat* st) {
  __shared__ double smax[16];
  __shared__ unsigned long long snan[16];
  double m = 0.;
  unsigned long long ns = 0;
  for (int i = threadIdx.x; i < np; i += 1024) {
    m = fmax(m, part[i].max);
    ns += part[i].nans;
  }
  m = wave_max(m);
  for (int o = 32; o > 0; o >>= 1) ns += __shfl_xor(ns, o, 64);
  if ((threadIdx.x & 63) == 0) {
    smax[threadIdx.x >> 6] = m;
    snan[threadIdx.x >> 6] = ns;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < 16; i++) {
      m = fmax(smax[0], smax[i]);
      smax[0] = m;
      snan[0] += snan[i];
    }
    st->maxbits = (unsigned long long)__double_as_longlong(smax[0]);
    st->nans = snan[0];
  }
}

// --- streaming red-black sweep, 3D ---------------------------------------
//
// One full red-black Gauss-Seidel sweep in a single pass, reading `dold` and
// writing `dnew` (the two buffers never alias).  A block owns a tile of
// kTJ x kTK cells in (j,k) and a run of `chunk` planes along i, which it
// walks plane by plane keeping four planes of the field in LDS (a 2-cell
// halo in j and k):
//   - red cells of plane s are relaxed on the tile plus a 1-cell ring, from
//     old black values only;
//   - black cells of plane s-1 are then relaxed on the tile from the new red
//     values of planes s-2, s-1 and s;
//   - plane s-1 of the tile is written out.
// Red cells on the ring and on the planes just outside the chunk are
// recomputed from `dold` instead of being read from the block that owns
// them, so blocks never exchange data and the values are exactly those of
// the two-colour sweep (red first, ghost fill, black).
//
// The field plane s+2 is loaded one step ahead into registers; the
// coefficients of a step are loaded before its first barrier.

constexpr int kTJ = 8, kTK = 64, kRJ = kTJ + 4, kRK = kTK + 4;
constexpr int kRJ1 = kTJ + 2, kRK1 = kTK + 2;
constexpr int kNLoad = (kRJ * kRK + kBlock - 1) / kBlock;    // field values per thread
constexpr int kNRed = (kRJ1 * kRK1 + kBlock - 1) / kBlock;   // red positions per thread
constexpr int kNBlack = (kTJ * kTK) / kBlock;                // black positions per thread

// coefficients of one cell: res, lambda and the six face alphas
template <typename T>
struct CellCoef {
  T res, lam, al[3], ah[3];
};

template <typename T, bool ALPHAF, bool LAMF>
__device__ __forceinline__ void fetch_coef(CellCoef<T>& cf, const T* __restrict__ res, const Coef<T>& k,
                                           const Geo<3>& g, const int c[3]) {
  size_t idx = ((size_t)c[0] * g.n + c[1]) * g.n + c[2];
  cf.res = res[idx];
  cf.lam = LAMF ? k.lam[idx] : k.lam_c;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (ALPHAF) {
      cf.al[d] = k.lo[d][idx];
      cf.ah[d] = alpha_hi<3, T>(k.lo[d], k.wall[d], g, idx, c, d);
    } else {
      cf.al[d] = cf.ah[d] = k.alpha;
    }
  }
}

template <typename T, bool ALPHAF, bool LAMF>
__global__ __launch_bounds__(kBlock) void k_rb_stream3(const T* __restrict__ dold, T* __restrict__ dnew,
                                                       const T* __restrict__ res, Coef<T> k, Geo<3> g, BCH bc,
                                                       T D2, int chunk) {
  __shared__ T buf[4][kRJ][kRK];
  const int n = g.n, tid = threadIdx.x;
  size_t b = logical_block();
  const int nkt = n / kTK, njt = n / kTJ;
  const int kt = (int)(b % nkt), jt = (int)((b / nkt) % njt), ct = (int)(b / ((size_t)nkt * njt));
  const int i0 = ct * chunk, j0 = jt * kTJ, k0 = kt * kTK;
  const int i1 = i0 + chunk;

  auto inrange = [&](int d, int v) { return bc.periodic[d] || (v >= 0 && v < n); };
  auto wrap = [&](int v) { return v < 0 ? v + n : (v >= n ? v - n : v); };
  auto slot = [&](int s) { return (s - i0 + 2) & 3; };

  // field plane s over the halo region: registers <- global
  auto fetch_plane = [&](int s, T* v) {
#pragma unroll
    for (int m = 0; m < kNLoad; m++) {
      int q = tid + m * kBlock;
      v[m] = T(0);
      if (q < kRJ * kRK && inrange(0, s)) {
        int pj = q / kRK, pk = q - pj * kRK;
        int gj = j0 + pj - 2, gk = k0 + pk - 2;
        if (inrange(1, gj) && inrange(2, gk)) v[m] = dold[((size_t)wrap(s) * n + wrap(gj)) * n + wrap(gk)];
      }
    }
  };
  auto commit_plane = [&](int s, const T* v) {
    if (!inrange(0, s)) return;
    T(*P)[kRK] = buf[slot(s)];
#pragma unroll
    for (int m = 0; m < kNLoad; m++) {
      int q = tid + m * kBlock;
      if (q < kRJ * kRK) {
        int pj = q / kRK, pk = q - pj * kRK;
        P[pj][pk] = v[m];
      }
    }
  };
  // position m of this thread in the red region (ring included) / the tile
  auto red_pos = [&](int m, int& gj, int& gk) {
    int q = tid + m * kBlock;
    if (q >= kRJ1 * kRK1) return false;
    int pj = q / kRK1, pk = q - pj * kRK1;
    gj = j0 + pj - 1;
    gk = k0 + pk - 1;
    return inrange(1, gj) && inrange(2, gk);
  };
  auto black_pos = [&](int m, int& gj, int& gk) {
    int q = tid + m * kBlock;
    int pj = q / kTK, pk = q - pj * kTK;
    gj = j0 + pj;
    gk = k0 + pk;
  };
  auto fetch_red = [&](int s, CellCoef<T>* cf) {
#pragma unroll
    for (int m = 0; m < kNRed; m++) {
      int gj, gk;
      if (inrange(0, s) && red_pos(m, gj, gk) && !((s + gj + gk) & 1)) {
        int c[3] = {wrap(s), wrap(gj), wrap(gk)};
        fetch_coef<T, ALPHAF, LAMF>(cf[m], res, k, g, c);
      }
    }
  };
  auto fetch_black = [&](int s, CellCoef<T>* cf) {
#pragma unroll
    for (int m = 0; m < kNBlack; m++) {
      int gj, gk;
      black_pos(m, gj, gk);
      if (s >= i0 && s < i1 && ((s + gj + gk) & 1)) {
        int c[3] = {s, gj, gk};
        fetch_coef<T, ALPHAF, LAMF>(cf[m], res, k, g, c);
      }
    }
  };
  // n/d at unwrapped position (s, gj, gk), neighbours from the LDS planes
  auto relax_at = [&](int s, int gj, int gk, const CellCoef<T>& cf) -> T {
    const int pj = gj - j0 + 2, pk = gk - k0 + 2;
    const int sl = slot(s);
    const T self = buf[sl][pj][pk];
    T nn = -D2 * cf.res, dd = -cf.lam * D2;
    T ap[3], am[3];
    ap[0] = inrange(0, s + 1) ? buf[slot(s + 1)][pj][pk] : T(bc.sigma[0]) * self;
    am[0] = inrange(0, s - 1) ? buf[slot(s - 1)][pj][pk] : T(bc.sigma[1]) * self;
    ap[1] = inrange(1, gj + 1) ? buf[sl][pj + 1][pk] : T(bc.sigma[2]) * self;
    am[1] = inrange(1, gj - 1) ? buf[sl][pj - 1][pk] : T(bc.sigma[3]) * self;
    ap[2] = inrange(2, gk + 1) ? buf[sl][pj][pk + 1] : T(bc.sigma[4]) * self;
    am[2] = inrange(2, gk - 1) ? buf[sl][pj][pk - 1] : T(bc.sigma[5]) * self;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      nn += cf.ah[d] * ap[d] + cf.al[d] * am[d];
      dd += cf.ah[d] + cf.al[d];
    }
    return nn / dd;
  };

  T pre_da[kNLoad];
  {
    T tmp[kNLoad];
    fetch_plane(i0 - 2, tmp);
    commit_plane(i0 - 2, tmp);
    fetch_plane(i0 - 1, tmp);
    commit_plane(i0 - 1, tmp);
  }
  fetch_plane(i0, pre_da);
  for (int s = i0 - 1; s <= i1; s++) {
    commit_plane(s + 1, pre_da);
    // loads: the field plane for the next step, the coefficients for this one
    if (s + 2 <= i1 + 1) fetch_plane(s + 2, pre_da);
    CellCoef<T> cur_red[kNRed], cur_blk[kNBlack];
    fetch_red(s, cur_red);
    fetch_black(s - 1, cur_blk);
    __syncthreads();
    // red on plane s over the tile and its 1-cell ring
    if (inrange(0, s)) {
#pragma unroll
      for (int m = 0; m < kNRed; m++) {
        int gj, gk;
        if (red_pos(m, gj, gk) && !((s + gj + gk) & 1))
          buf[slot(s)][gj - j0 + 2][gk - k0 + 2] = relax_at(s, gj, gk, cur_red[m]);
      }
    }
    __syncthreads();
    // black on plane s-1 over the tile, then write the plane out
    const int q1 = s - 1;
    if (q1 >= i0 && q1 < i1) {
#pragma unroll
      for (int m = 0; m < kNBlack; m++) {
        int gj, gk;
        black_pos(m, gj, gk);
        T v;
        if ((q1 + gj + gk) & 1) {
          v = relax_at(q1, gj, gk, cur_blk[m]);
          buf[slot(q1)][gj - j0 + 2][gk - k0 + 2] = v;
        } else {
          v = buf[slot(q1)][gj - j0 + 2][gk - k0 + 2];
        }
        dnew[((size_t)q1 * n + gj) * n + gk] = v;
      }
    }
    __syncthreads();
  }
}

// --- streaming red-black sweep, 3D, colour-mapped threads ---------------
//
// The schedule and the arithmetic of k_rb_stream3 (red of plane s over the
// tile and its 1-cell ring from old black values, black of plane s-1 over
// the tile from the new red values), with two changes:
//   - every thread owns cells of one colour: in the tile of 8 x 64 cells the
//     pair (j, 2m), (j, 2m+1) -- one black, one red -- written out with one
//     vector store, and up to two of the 330 red cells of the ring region;
//     no thread slot is predicated off by the colour;
//   - the coefficients of step s+1 are loaded into registers while step s is
//     computed, so their latency overlaps the LDS work.

constexpr int kRedRow = kRK1 / 2;                 // red cells per ring row (33)
constexpr int kRedCells = kRJ1 * kRedRow;         // 330
constexpr int kNRed2 = (kRedCells + kBlock - 1) / kBlock;  // 2

template <typename T, bool ALPHAF, bool LAMF>
__global__ __launch_bounds__(kBlock) void k_rb_stream3b(const T* __restrict__ dold, T* __restrict__ dnew,
                                                        const T* __restrict__ res, Coef<T> k, Geo<3> g, BCH bc,
                                                        T D2, int chunk) {
  __shared__ T buf[4][kRJ][kRK];
  const int n = g.n, tid = threadIdx.x;
  size_t b = logical_block();
  const int nkt = n / kTK, njt = n / kTJ;
  const int kt = (int)(b % nkt), jt = (int)((b / nkt) % njt), ct = (int)(b / ((size_t)nkt * njt));
  const int i0 = ct * chunk, j0 = jt * kTJ, k0 = kt * kTK;
  const int i1 = i0 + chunk;

  auto inrange = [&](int d, int v) { return bc.periodic[d] || (v >= 0 && v < n); };
  auto wrap = [&](int v) { return v < 0 ? v + n : (v >= n ? v - n : v); };
  auto slot = [&](int s) { return (s - i0 + 2) & 3; };

  auto fetch_plane = [&](int s, T* v) {
#pragma unroll
    for (int m = 0; m < kNLoad; m++) {
      int q = tid + m * kBlock;
      v[m] = T(0);
      if (q < kRJ * kRK && inrange(0, s)) {
        int pj = q / kRK, pk = q - pj * kRK;
        int gj = j0 + pj - 2, gk = k0 + pk - 2;
        if (inrange(1, gj) && inrange(2, gk)) v[m] = dold[((size_t)wrap(s) * n + wrap(gj)) * n + wrap(gk)];
      }
    }
  };
  auto commit_plane = [&](int s, const T* v) {
    if (!inrange(0, s)) return;
    T(*P)[kRK] = buf[slot(s)];
#pragma unroll
    for (int m = 0; m < kNLoad; m++) {
      int q = tid + m * kBlock;
      if (q < kRJ * kRK) {
        int pj = q / kRK, pk = q - pj * kRK;
        P[pj][pk] = v[m];
      }
    }
  };
  // red cell m of this thread on plane s: row of the ring region and the k
  // of the red parity (s + j + k even); false when outside the domain
  auto red_cell = [&](int s, int m, int& gj, int& gk) {
    int q = tid + m * kBlock;
    if (q >= kRedCells) return false;
    int r = q / kRedRow, mm = q - r * kRedRow;
    gj = j0 - 1 + r;
    const int kb = k0 - 1;
    gk = kb + 2 * mm + ((s + gj + kb) & 1);
    return inrange(1, gj) && inrange(2, gk);
  };
  // the pair of tile cells of this thread: row j, cells k and k+1
  const int pr = tid / (kTK / 2), pm = tid - pr * (kTK / 2);
  const int bj = j0 + pr, bk = k0 + 2 * pm;
  // which of the pair is black on plane q: (q + bj + bk + e) odd
  auto black_e = [&](int q) { return (q + bj + bk + 1) & 1; };

  auto fetch_red = [&](int s, CellCoef<T>* cf) {
#pragma unroll
    for (int m = 0; m < kNRed2; m++) {
      int gj, gk;
      if (inrange(0, s) && red_cell(s, m, gj, gk)) {
        int c[3] = {wrap(s), wrap(gj), wrap(gk)};
        fetch_coef<T, ALPHAF, LAMF>(cf[m], res, k, g, c);
      }
    }
  };
  auto fetch_black = [&](int s, CellCoef<T>& cf) {
    if (s >= i0 && s < i1) {
      int c[3] = {s, bj, bk + black_e(s)};
      fetch_coef<T, ALPHAF, LAMF>(cf, res, k, g, c);
    }
  };
  auto relax_at = [&](int s, int gj, int gk, const CellCoef<T>& cf) -> T {
    const int pj = gj - j0 + 2, pk = gk - k0 + 2;
    const int sl = slot(s);
    const T self = buf[sl][pj][pk];
    T nn = -D2 * cf.res, dd = -cf.lam * D2;
    T ap[3], am[3];
    ap[0] = inrange(0, s + 1) ? buf[slot(s + 1)][pj][pk] : T(bc.sigma[0]) * self;
    am[0] = inrange(0, s - 1) ? buf[slot(s - 1)][pj][pk] : T(bc.sigma[1]) * self;
    ap[1] = inrange(1, gj + 1) ? buf[sl][pj + 1][pk] : T(bc.sigma[2]) * self;
    am[1] = inrange(1, gj - 1) ? buf[sl][pj - 1][pk] : T(bc.sigma[3]) * self;
    ap[2] = inrange(2, gk + 1) ? buf[sl][pj][pk + 1] : T(bc.sigma[4]) * self;
    am[2] = inrange(2, gk - 1) ? buf[sl][pj][pk - 1] : T(bc.sigma[5]) * self;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      nn += cf.ah[d] * ap[d] + cf.al[d] * am[d];
      dd += cf.ah[d] + cf.al[d];
    }
    return nn / dd;
  };

  T pre_da[kNLoad];
  {
    T tmp[kNLoad];
    fetch_plane(i0 - 2, tmp);
    commit_plane(i0 - 2, tmp);
    fetch_plane(i0 - 1, tmp);
    commit_plane(i0 - 1, tmp);
  }
  fetch_plane(i0, pre_da);
  CellCoef<T> cur_red[kNRed2], cur_blk, nxt_red[kNRed2], nxt_blk;
  fetch_red(i0 - 1, cur_red);
  fetch_black(i0 - 2, cur_blk);
  for (int s = i0 - 1; s <= i1; s++) {
    commit_plane(s + 1, pre_da);
    if (s + 2 <= i1 + 1) fetch_plane(s + 2, pre_da);
    fetch_red(s + 1, nxt_red);
    fetch_black(s, nxt_blk);
    __syncthreads();
    if (inrange(0, s)) {
#pragma unroll
      for (int m = 0; m < kNRed2; m++) {
        int gj, gk;
        if (red_cell(s, m, gj, gk)) buf[slot(s)][gj - j0 + 2][gk - k0 + 2] = relax_at(s, gj, gk, cur_red[m]);
      }
    }
    __syncthreads();
    const int q1 = s - 1;
    if (q1 >= i0 && q1 < i1) {
      const int e = black_e(q1);
      T v[2];
      v[e] = relax_at(q1, bj, bk + e, cur_blk);
      v[1 - e] = buf[slot(q1)][bj - j0 + 2][bk + (1 - e) - k0 + 2];
      using V2 = typename Vec2<T>::type;
      V2 out;
      out.x = v[0];
      out.y = v[1];
      *reinterpret_cast<V2*>(dnew + ((size_t)q1 * n + bj) * n + bk) = out;
    }
    __syncthreads();
#pragma unroll
    for (int m = 0; m < kNRed2; m++) cur_red[m] = nxt_red[m];
    cur_blk = nxt_blk;
  }
}

// --- raw buffer access ---------------------------------------------------
//
// A buffer resource (base address, size in bytes) per array or plane; the
// hardware range check makes a load at an offset outside [0, size) return
// zero without touching memory, so loads need neither clamping nor branches,
// and the per-lane address is one 32-bit offset shared by every array.

using Rsrc = __amdgpu_buffer_rsrc_t;
// the builtins' own vector types (GCC vector_size, not ext_vector_type):
// values are moved in and out with bit casts only
typedef unsigned int bv2 __attribute__((vector_size(8)));
typedef unsigned int bv4 __attribute__((vector_size(16)));
// dword 3 of a gfx9 buffer resource: 32-bit data format
constexpr int kBufWord3 = 0x00020000;
// an offset past the end of every resource
constexpr uint32_t kOOB = 0x80000000u;

__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, kBufWord3);
}
// loads and stores at byte offset off (per lane) + soff (wave-uniform, e.g.
// a plane); a resource covers at most 2^31 bytes, so kOOB + soff never wraps
template <typename T>
__device__ __forceinline__ T bld1(Rsrc r, uint32_t off, uint32_t soff = 0);
template <>
__device__ __forceinline__ float bld1<float>(Rsrc r, uint32_t off, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, (int)soff, 0));
}
template <>
__device__ __forceinline__ double bld1<double>(Rsrc r, uint32_t off, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0));
}
template <typename T>
__device__ __forceinline__ typename Vec2<T>::type bld2(Rsrc r, uint32_t off, uint32_t soff = 0);
template <>
__device__ __forceinline__ float2 bld2<float>(Rsrc r, uint32_t off, uint32_t soff) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, (int)soff, 0));
}
template <>
__device__ __forceinline__ double2 bld2<double>(Rsrc r, uint32_t off, uint32_t soff) {
  return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, (int)soff, 0));
}
__device__ __forceinline__ void bst2(Rsrc r, uint32_t off, uint32_t soff, float2 v) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(bv2, v), r, (int)off, (int)soff, 0);
}
__device__ __forceinline__ void bst2(Rsrc r, uint32_t off, uint32_t soff, double2 v) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(bv4, v), r, (int)off, (int)soff, 0);
}
// bitwise or: combines two loads of which at most one was in range
__device__ __forceinline__ float bor(float a, float b) {
  return __builtin_bit_cast(float, __builtin_bit_cast(unsigned, a) | __builtin_bit_cast(unsigned, b));
}
__device__ __forceinline__ double bor(double a, double b) {
  return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, a) | __builtin_bit_cast(unsigned long long, b));
}
// --- end of raw buffer access

// --- streaming red-black sweep, 3D, coefficients streamed per column ---
//
// One full red-black sweep in one pass, dold -> dnew, for levels with walls
// in every direction (no periodic wrap).  The schedule is k_rb_stream3's:
// step s relaxes the red cells of plane s, then the black cells of plane
// s-1, from the field planes held in LDS (eight slots, so that loading a
// plane needs no barrier of its own).  The halo cells outside the level hold
// their ghost values, sigma times the old value of the cell they mirror --
// the value relax_value gives a ghost neighbour -- so that every neighbour is
// a plain LDS read.  The tile is TJ x 64 cells (TJ = 32 in FP32, 16 in
// FP64, where registers allow only 512 threads) with one thread per pair of
// cells (j, k), (j, k+1), k even.  On plane s the red cell of a
// pair and on plane s-1 its black cell sit at the same k, so the coefficients
// of a pair on plane p -- res, lambda and the faces, loaded as vectors --
// serve the red cell on step p and the black cell on step p+1.  The red cells
// of the 1-cell ring around the tile, which the black cells on the tile's
// edge need, are relaxed by the first TJ + 66 threads (two waves), with the
// coefficients relax_value reads; like k_rb_stream3, blocks recompute them
// instead of exchanging data.  Same arithmetic as relax_value.
//
// One barrier per step (between the loads and the red cells).  Everything
// a step reads from global memory is loaded during the step before, through
// buffer resources of the plane (a plane outside the level
// gets an empty resource), and the three coefficient records rotate their
// roles through a loop unrolled by three, so that no register copy and no
// branch makes the compiler wait for a load early.

// tiles of the streaming kernels: 16 x 64 cells, 512 threads (the residual
// kernel below; the sweep's FP64 tile), and the sweep's FP32 tile of 32 rows
constexpr int kSTJ = 16, kSTK = 64, kSBlock = 512;
constexpr int kSTJ32 = 32;
constexpr int kSRingThreads = 128;  // the waves that relax the sweep's ring

// coefficients of a pair of cells (k, k+1) on one plane p
template <typename T>
struct PairCoef {
  typename Vec2<T>::type res, lam, ay, ayh, az, axl, axh;  // axh: x faces of plane p+1 (or the wall)
  T azh;  // high z face of cell k+1: low face of cell k+2, or the wall
};


template <typename T>
__device__ __forceinline__ T relax_nd(T D2, T res, T lam, const T al[3], const T ah[3], const T ap[3],
                                      const T am[3]) {
  T nn = -D2 * res, dd = -lam * D2;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    nn += ah[d] * ap[d] + al[d] * am[d];
    dd += ah[d] + al[d];
  }
  return nn / dd;
}

// at most 128 registers: 16 waves per CU (FP32, one block of 1024 threads);
// FP64: one block of 512 threads
template <typename T, bool ALPHAF, bool LAMF, int TJ>
__global__ __launch_bounds__(TJ * kSTK / 2) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 4 : 2)))
void k_rb_stream3c(const T* __restrict__ dold, T* __restrict__ dnew, const T* __restrict__ res, Coef<T> k,
                   Geo<3> g, BCH bc, T D2, int nchunks) {
  using V2 = typename Vec2<T>::type;
  constexpr int kSTJ = TJ, kSBlock = TJ * kSTK / 2;
  constexpr int kSRJ = kSTJ + 4, kSRK = kSTK + 4;              // field region, 2-cell halo
  constexpr int kSNRing = kSTJ + kSTK + 2;                      // red ring cells per plane
  static_assert(kSNRing <= kSRingThreads, "ring cells fit two waves");
  constexpr int kSRegion2 = kSRJ * kSRK / 2;                    // vector slots of the field region
  constexpr int kSNLoad = (kSRegion2 + kSBlock - 1) / kSBlock;  // vector field loads per thread
  __shared__ __align__(16) T buf[8][kSRJ][kSRK];
  const int n = g.n, tid = threadIdx.x;
  const size_t b = logical_block();
  const int nkt = n / kSTK, njt = n / kSTJ;
  const int kt = (int)(b % nkt), jt = (int)((b / nkt) % njt), ct = (int)(b / ((size_t)nkt * njt));
  const int i0 = (int)((size_t)ct * n / nchunks), i1 = (int)((size_t)(ct + 1) * n / nchunks);
  const int j0 = jt * kSTJ, k0 = kt * kSTK;
  // the pair of this thread: row bj, cells bk and bk+1
  const int pr = tid >> 5, pm = tid & 31;
  const int bj = j0 + pr, bk = k0 + 2 * pm;
  const int pj = pr + 2, pk = 2 * pm + 2;  // LDS position of cell bk
  auto slot = [&](int p) { return (p - i0 + 16) & 7; };
  auto in = [&](int v) { return v >= 0 && v < n; };
  auto clampn = [&](int v) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); };
  constexpr uint32_t es = sizeof(T);
  const size_t plane = (size_t)n * n;
  const uint32_t pbytes = (uint32_t)(plane * es), rbytes = (uint32_t)n * es;
  const uint32_t voff = ((uint32_t)bj * n + bk) * es;  // the pair within a plane
  // one resource per array; a plane is selected by the uniform offset
  const uint32_t vbytes = (uint32_t)((size_t)n * plane * es);
  const Rsrc Rold = make_rsrc(dold, vbytes), Rnew = make_rsrc(dnew, vbytes), Rres = make_rsrc(res, vbytes);
  const Rsrc Rlam = make_rsrc(LAMF ? k.lam : nullptr, LAMF ? vbytes : 0u);
  const Rsrc Rl0 = make_rsrc(ALPHAF ? k.lo[0] : nullptr, ALPHAF ? vbytes : 0u);
  const Rsrc Rl1 = make_rsrc(ALPHAF ? k.lo[1] : nullptr, ALPHAF ? vbytes : 0u);
  const Rsrc Rl2 = make_rsrc(ALPHAF ? k.lo[2] : nullptr, ALPHAF ? vbytes : 0u);
  const Rsrc Rw0 = make_rsrc(ALPHAF ? k.wall[0] : nullptr, ALPHAF ? pbytes : 0u);
  const Rsrc Rw1 = make_rsrc(ALPHAF ? k.wall[1] : nullptr, ALPHAF ? pbytes : 0u);
  const Rsrc Rw2 = make_rsrc(ALPHAF ? k.wall[2] : nullptr, ALPHAF ? pbytes : 0u);
  // offsets of plane p (clamped into the level) in the n^3 arrays and of its
  // row in the y and z walls
  auto pso = [&](int p) { return (uint32_t)clampn(p) * pbytes; };
  auto wso = [&](int p) { return (uint32_t)clampn(p) * rbytes; };

  // field plane p over the tile and its 2-cell halo; outside the level, the
  // cells the ghosts mirror (plane, row and pair clamped into the level)
  auto fetch_plane = [&](int p, V2* v) {
    const uint32_t so = pso(p);
#pragma unroll
    for (int m = 0; m < kSNLoad; m++) {
      const int q = tid + m * kSBlock;
      const int row = q / (kSRK / 2), c2 = q - row * (kSRK / 2);
      const int gj = clampn(j0 - 2 + row), gk = k0 - 2 + 2 * c2;
      const int mk = gk < 0 ? 0 : (gk > n - 2 ? n - 2 : gk);
      v[m] = bld2<T>(Rold, q < kSRegion2 ? ((uint32_t)gj * n + mk) * es : kOOB, so);
    }
  };
  // ... stored with the ghost rule: sigma times the mirrored cell
  auto commit_plane = [&](int p, const V2* v) {
    T(*P)[kSRK] = buf[slot(p)];
#pragma unroll
    for (int m = 0; m < kSNLoad; m++) {
      const int q = tid + m * kSBlock;
      if (q < kSRegion2) {
        const int row = q / (kSRK / 2), c2 = q - row * (kSRK / 2);
        const int gj = j0 - 2 + row, gk = k0 - 2 + 2 * c2;
        V2 w = v[m];
        if (p < 0 || p >= n || gj < 0 || gj >= n || gk < 0 || gk >= n) {
          const int sx = p < 0 ? 1 : 0, sy = gj < 0 ? 3 : 2;
          if (p < 0 || p >= n) {
            w.x = T(bc.sigma[sx]) * w.x;
            w.y = T(bc.sigma[sx]) * w.y;
          } else if (gj < 0 || gj >= n) {
            w.x = T(bc.sigma[sy]) * w.x;
            w.y = T(bc.sigma[sy]) * w.y;
          } else if (gk < 0) {
            w.y = T(bc.sigma[5]) * w.x;  // cell -1 mirrors cell 0
          } else {
            w.x = T(bc.sigma[4]) * w.y;  // cell n mirrors cell n-1
          }
        }
        *reinterpret_cast<V2*>(&P[row][2 * c2]) = w;
      }
    }
  };
  // coefficients of the pair on plane p (clamped into the level); the high
  // x faces of the last plane are the wall (n x n, (j,k))
  auto fetch_pair = [&](int p, PairCoef<T>& c) {
    const uint32_t so = pso(p), sw = wso(p);
    c.res = bld2<T>(Rres, voff, so);
    if (LAMF) c.lam = bld2<T>(Rlam, voff, so);
    if (ALPHAF) {
      const bool lastj = bj == n - 1, lastk = bk + 2 == n, lastp = clampn(p) == n - 1;
      c.ay = bld2<T>(Rl1, voff, so);
      const V2 yl = bld2<T>(Rl1, lastj ? kOOB : voff + rbytes, so), yw = bld2<T>(Rw1, lastj ? bk * es : kOOB, sw);
      c.ayh.x = bor(yl.x, yw.x);
      c.ayh.y = bor(yl.y, yw.y);
      c.az = bld2<T>(Rl2, voff, so);
      c.azh = bor(bld1<T>(Rl2, lastk ? kOOB : voff + 2 * es, so), bld1<T>(Rw2, lastk ? bj * es : kOOB, sw));
      c.axl = bld2<T>(Rl0, voff, so);
      c.axh = bld2<T>(lastp ? Rw0 : Rl0, voff, lastp ? 0u : so + pbytes);
    }
  };
  // ring cell of this thread on plane s (threads 0..81)
  const int rq = tid < kSNRing ? tid : kSNRing;
  auto ring_cell = [&](int s, int& gj, int& gk) {
    if (rq < 2 * (kSTK / 2 + 1)) {
      const int half = kSTK / 2 + 1;  // 33 red cells on a ring row
      gj = rq < half ? j0 - 1 : j0 + kSTJ;
      const int m = rq < half ? rq : rq - half;
      gk = k0 - 1 + 2 * m + ((s + gj + k0 - 1) & 1);
    } else {
      const int qq = rq - 2 * (kSTK / 2 + 1);
      gk = qq < kSTJ / 2 ? k0 - 1 : k0 + kSTK;
      const int m = qq < kSTJ / 2 ? qq : qq - kSTJ / 2;
      gj = j0 + 2 * m + ((s + j0 + gk) & 1);
    }
    return rq < kSNRing && in(gj) && in(gk);
  };
  // coefficients of this thread's ring cell on plane s (relax_value's reads)
  auto fetch_ring = [&](int s, CellCoef<T>& cf) {
    int gj, gk;
    const bool ok = ring_cell(s, gj, gk);
    const uint32_t off = ok ? ((uint32_t)gj * n + gk) * es : kOOB;
    const uint32_t so = pso(s), sw = wso(s);
    const bool lastp = clampn(s) == n - 1;
    cf.res = bld1<T>(Rres, off, so);
    cf.lam = LAMF ? bld1<T>(Rlam, off, so) : k.lam_c;
    if (ALPHAF) {
      cf.al[0] = bld1<T>(Rl0, off, so);
      cf.ah[0] = bld1<T>(lastp ? Rw0 : Rl0, off, lastp ? 0u : so + pbytes);
      cf.al[1] = bld1<T>(Rl1, off, so);
      cf.ah[1] = bor(bld1<T>(Rl1, ok && gj < n - 1 ? off + rbytes : kOOB, so),
                     bld1<T>(Rw1, ok && gj == n - 1 ? gk * es : kOOB, sw));
      cf.al[2] = bld1<T>(Rl2, off, so);
      cf.ah[2] = bor(bld1<T>(Rl2, ok && gk < n - 1 ? off + es : kOOB, so),
                     bld1<T>(Rw2, ok && gk == n - 1 ? gj * es : kOOB, sw));
    } else {
#pragma unroll
      for (int d = 0; d < 3; d++) cf.al[d] = cf.ah[d] = k.alpha;
    }
  };
  // neighbours of the cell at LDS (qj, qk) of plane s (ghosts included)
  auto neighbours = [&](int s, int qj, int qk, T* ap, T* am) {
    const int sl = slot(s);
    ap[0] = buf[slot(s + 1)][qj][qk];
    am[0] = buf[slot(s - 1)][qj][qk];
    ap[1] = buf[sl][qj + 1][qk];
    am[1] = buf[sl][qj - 1][qk];
    ap[2] = buf[sl][qj][qk + 1];
    am[2] = buf[sl][qj][qk - 1];
  };
  auto pick = [](V2 v, int e) { return e ? v.y : v.x; };
  // relax cell e of the pair on plane s with the coefficients cp of plane s
  auto relax_pair_cell = [&](int s, int e, const PairCoef<T>& cp) {
    const int qk = pk + e;
    T ap[3], am[3], al[3], ah[3];
    neighbours(s, pj, qk, ap, am);
    if (ALPHAF) {
      al[0] = pick(cp.axl, e);
      ah[0] = pick(cp.axh, e);
      al[1] = pick(cp.ay, e);
      ah[1] = pick(cp.ayh, e);
      al[2] = pick(cp.az, e);
      ah[2] = e ? cp.azh : cp.az.y;
    } else {
      al[0] = al[1] = al[2] = ah[0] = ah[1] = ah[2] = k.alpha;
    }
    return relax_nd<T>(D2, pick(cp.res, e), LAMF ? pick(cp.lam, e) : k.lam_c, al, ah, ap, am);
  };

  V2 dreg[kSNLoad];
  // step s: red of plane s (coefficients cur, ring rcur), black of plane
  // s-1 (coefficients prev) when `black`; loads plane s+2 of the field and
  // the coefficients of plane s+1 (next, rnext)
  auto step = [&](int s, bool black, const PairCoef<T>& prev, const PairCoef<T>& cur, PairCoef<T>& next,
                  const CellCoef<T>& rcur, CellCoef<T>& rnext) {
    commit_plane(s + 1, dreg);
    fetch_plane(s + 2, dreg);
    fetch_pair(s + 1, next);
    if (tid < kSRingThreads) fetch_ring(s + 1, rnext);
    __syncthreads();
    const int e = (s + bj + bk) & 1;
    if (in(s)) {
      buf[slot(s)][pj][pk + e] = relax_pair_cell(s, e, cur);
      int gj, gk;
      if (tid < kSRingThreads && s >= i0 && s < i1 && ring_cell(s, gj, gk)) {
        const int qj = gj - j0 + 2, qk = gk - k0 + 2;
        T ap[3], am[3];
        neighbours(s, qj, qk, ap, am);
        buf[slot(s)][qj][qk] = relax_nd<T>(D2, rcur.res, rcur.lam, rcur.al, rcur.ah, ap, am);
      }
    }
    // no barrier here: the black cell of plane s-1 reads the red cells of
    // planes s and s-2 at its own (j, k), which this thread wrote itself, and
    // the red cells of plane s-1 around it, written during the step before
    // (ordered by this step's barrier)
    if (black) {
      // black of plane s-1 at the k of this step's red cell; the pair is
      // written out with the red value of plane s-1
      const int q1 = s - 1;
      V2 out;
      const T v = relax_pair_cell(q1, e, prev);
      const T r = buf[slot(q1)][pj][pk + 1 - e];
      out.x = e ? r : v;
      out.y = e ? v : r;
      bst2(Rnew, voff, (uint32_t)q1 * pbytes, out);
    }
  };

  // prologue: planes i0-2 and i0-1 in LDS, plane i0 in registers
  {
    V2 tmp[kSNLoad];
    fetch_plane(i0 - 2, tmp);
    commit_plane(i0 - 2, tmp);
    fetch_plane(i0 - 1, tmp);
    commit_plane(i0 - 1, tmp);
  }
  fetch_plane(i0, dreg);
  PairCoef<T> A{}, B{}, C{};
  CellCoef<T> RA{}, RB{}, RC{};
  fetch_pair(i0 - 1, A);
  step(i0 - 1, false, C, A, B, RC, RA);  // RA <- ring of plane i0
  step(i0, false, A, B, C, RA, RB);
  for (int s = i0 + 1; s <= i1; s += 3) {
    step(s, true, B, C, A, RB, RC);
    if (s + 1 > i1) break;
    step(s + 1, true, C, A, B, RC, RA);
    if (s + 2 > i1) break;
    step(s + 2, true, A, B, C, RA, RB);
  }
}

// --- two red-black sweeps in one pass (temporal blocking), 3D, FP32 ------
//
// Two consecutive red-black sweeps, A then B, of a level with walls in
// every direction, dold -> dnew, in one pass over the coefficients: each
// plane's res, lambda and faces are loaded once and serve four half-sweeps.
// Step s of the stream along i relaxes, for the cells whose column parity is
// red on plane s, A's red cells of plane s, A's black cells of plane s-1,
// B's red cells of plane s-2 and B's black cells of plane s-3.  Each reads
// its plane neighbours in its own column, written by the same thread earlier
// in the step or two steps before, and its in-plane neighbours, which have
// the other parity and were written during the step before (ordered by the
// step's one barrier).  The coefficients of a pair of cells on plane p serve
// the four stages of steps p .. p+3, so a thread keeps the records of planes
// s+1 .. s-3 of its pair and rotates their roles through a loop unrolled by
// five.
//
// B's black cells on the tile's edge need B's red cells one cell outside the
// tile, which need A's black cells two cells out, which need A's red cells
// three cells out: 5 extra waves own the pairs of a ring of width 3 around
// the 16 x 64 tile and relax, per step, the stages their cell's distance to
// the tile allows; and a chunk of planes [i0, i1) runs its stages from three
// planes below to three above it.  Blocks recompute these rings instead of
// exchanging data.  The field region in LDS has a halo of 4 cells.
//
// Ghosts: a sweep reads, across a wall, sigma times the value its cell had
// before the sweep.  The planes hold sigma times dold there when loaded, so
// an A stage that writes a cell next to a wall also writes sigma times the
// new value into the ghost position, which no other cell reads.
//
// The arithmetic is relax_nd's: the result equals two sweeps of
// k_rb_stream3c bit for bit.
constexpr int kTTJ = 16, kTTK = 64, kTH = 4;                      // tile rows, columns, LDS halo
constexpr int kTInner = kTTJ * kTTK / 2;                          // pairs of the tile
constexpr int kTRingPairs = 6 * ((kTTK + 2 * kTH) / 2) + 4 * kTTJ;  // pairs of the ring
constexpr int kTBlock = (kTInner + kTRingPairs + 63) / 64 * 64;

template <typename T, bool ALPHAF, bool LAMF>
__global__ __launch_bounds__(kTBlock) void k_rb_stream3t(const T* __restrict__ dold, T* __restrict__ dnew,
                                                         const T* __restrict__ res, Coef<T> k, Geo<3> g, BCH bc,
                                                         T D2, int nchunks) {
  using V2 = typename Vec2<T>::type;
  constexpr int RJ = kTTJ + 2 * kTH, RK = kTTK + 2 * kTH;  // field region in LDS
  constexpr int HP = RK / 2;                               // pairs per region row
  constexpr int Region2 = RJ * HP;
  constexpr int NLoad = (Region2 + kTBlock - 1) / kTBlock;
  __shared__ __align__(16) T buf[8][RJ][RK];
  const int n = g.n, tid = threadIdx.x;
  const size_t b = logical_block();
  const int nkt = n / kTTK, njt = n / kTTJ;
  const int kt = (int)(b % nkt), jt = (int)((b / nkt) % njt), ct = (int)(b / ((size_t)nkt * njt));
  const int i0 = (int)((size_t)ct * n / nchunks), i1 = (int)((size_t)(ct + 1) * n / nchunks);
  const int j0 = jt * kTTJ, k0 = kt * kTTK;

  // the pair of this thread, (pj, pk) and (pj, pk+1): the tile's pairs, then
  // the ring's (three rows below and above the tile over the region's width,
  // and the two pairs left and right of each tile row)
  int pj = j0, pk = k0;
  const bool valid = tid < kTInner + kTRingPairs;
  if (tid < kTInner) {
    pj = j0 + (tid >> 5);
    pk = k0 + 2 * (tid & 31);
  } else if (valid) {
    const int u = tid - kTInner;
    if (u < 6 * HP) {
      const int r = u / HP, c = u - r * HP;
      pj = r < 3 ? j0 - 3 + r : j0 + kTTJ + r - 3;
      pk = k0 - kTH + 2 * c;
    } else {
      const int v = u - 6 * HP, side = v & 3;
      pj = j0 + (v >> 2);
      pk = side < 2 ? k0 - kTH + 2 * side : k0 + kTTK + 2 * (side - 2);
    }
  }
  const bool pair_in = valid && pj >= 0 && pj < n && pk >= 0 && pk < n;
  // how far each cell of the pair lies outside the tile (more than 3: idle)
  auto dist = [](int x, int lo, int len) { return x < lo ? lo - x : (x >= lo + len ? x - (lo + len - 1) : 0); };
  const int wj = dist(pj, j0, kTTJ);
  const int w0 = pair_in ? max(wj, dist(pk, k0, kTTK)) : 99;
  const int w1 = pair_in ? max(wj, dist(pk + 1, k0, kTTK)) : 99;
  const int qj = pj - j0 + kTH, qk0 = pk - k0 + kTH;  // LDS position of cell pk

  const int s0 = i0 - kTH;  // lowest plane loaded
  auto slot = [&](int p) { return (p - s0 + 16) & 7; };
  auto clampn = [&](int v) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); };
  constexpr uint32_t es = sizeof(T);
  const size_t plane = (size_t)n * n;
  const uint32_t pbytes = (uint32_t)(plane * es), rbytes = (uint32_t)n * es;
  const uint32_t voff = pair_in ? ((uint32_t)pj * n + pk) * es : kOOB;
  const uint32_t vbytes = (uint32_t)((size_t)n * plane * es);
  const Rsrc Rold = make_rsrc(dold, vbytes), Rnew = make_rsrc(dnew, vbytes), Rres = make_rsrc(res, vbytes);
  const Rsrc Rlam = make_rsrc(LAMF ? k.lam : nullptr, LAMF ? vbytes : 0u);
  const Rsrc Rl0 = make_rsrc(ALPHAF ? k.lo[0] : nullptr, ALPHAF ? vbytes : 0u);
  const Rsrc Rl1 = make_rsrc(ALPHAF ? k.lo[1] : nullptr, ALPHAF ? vbytes : 0u);
  const Rsrc Rl2 = make_rsrc(ALPHAF ? k.lo[2] : nullptr, ALPHAF ? vbytes : 0u);
  const Rsrc Rw0 = make_rsrc(ALPHAF ? k.wall[0] : nullptr, ALPHAF ? pbytes : 0u);
  const Rsrc Rw1 = make_rsrc(ALPHAF ? k.wall[1] : nullptr, ALPHAF ? pbytes : 0u);
  const Rsrc Rw2 = make_rsrc(ALPHAF ? k.wall[2] : nullptr, ALPHAF ? pbytes : 0u);
  auto pso = [&](int p) { return (uint32_t)clampn(p) * pbytes; };
  auto wso = [&](int p) { return (uint32_t)clampn(p) * rbytes; };

  // field plane p over the region; outside the level, the cells the ghosts
  // mirror (plane, row and pair clamped into the level)
  auto fetch_plane = [&](int p, V2* v) {
    const uint32_t so = pso(p);
#pragma unroll
    for (int m = 0; m < NLoad; m++) {
      const int q = tid + m * kTBlock;
      const int row = q / HP, c2 = q - row * HP;
      const int gj = clampn(j0 - kTH + row), gk = k0 - kTH + 2 * c2;
      const int mk = gk < 0 ? 0 : (gk > n - 2 ? n - 2 : gk);
      v[m] = bld2<T>(Rold, q < Region2 ? ((uint32_t)gj * n + mk) * es : kOOB, so);
    }
  };
  // ... stored with the ghost rule: sigma times the mirrored cell
  auto commit_plane = [&](int p, const V2* v) {
    T(*P)[RK] = buf[slot(p)];
#pragma unroll
    for (int m = 0; m < NLoad; m++) {
      const int q = tid + m * kTBlock;
      if (q < Region2) {
        const int row = q / HP, c2 = q - row * HP;
        const int gj = j0 - kTH + row, gk = k0 - kTH + 2 * c2;
        V2 w = v[m];
        if (p < 0 || p >= n) {
          const T sx = T(bc.sigma[p < 0 ? 1 : 0]);
          w.x = sx * w.x;
          w.y = sx * w.y;
        } else if (gj < 0 || gj >= n) {
          const T sy = T(bc.sigma[gj < 0 ? 3 : 2]);
          w.x = sy * w.x;
          w.y = sy * w.y;
        } else if (gk < 0) {
          w.y = T(bc.sigma[5]) * w.x;  // cell -1 mirrors cell 0
        } else if (gk >= n) {
          w.x = T(bc.sigma[4]) * w.y;  // cell n mirrors cell n-1
        }
        *reinterpret_cast<V2*>(&P[row][2 * c2]) = w;
      }
    }
  };
  // coefficients of the pair on plane p (clamped into the level)
  auto fetch_pair = [&](int p, PairCoef<T>& c) {
    const uint32_t so = pso(p), sw = wso(p);
    c.res = bld2<T>(Rres, voff, so);
    if (LAMF) c.lam = bld2<T>(Rlam, voff, so);
    if (ALPHAF) {
      const bool lastj = pair_in && pj == n - 1, lastk = pair_in && pk + 2 == n, lastp = clampn(p) == n - 1;
      c.ay = bld2<T>(Rl1, voff, so);
      const V2 yl = bld2<T>(Rl1, lastj ? kOOB : voff + rbytes, so), yw = bld2<T>(Rw1, lastj ? pk * es : kOOB, sw);
      c.ayh.x = bor(yl.x, yw.x);
      c.ayh.y = bor(yl.y, yw.y);
      c.az = bld2<T>(Rl2, voff, so);
      c.azh = bor(bld1<T>(Rl2, lastk ? kOOB : voff + 2 * es, so), bld1<T>(Rw2, lastk ? pj * es : kOOB, sw));
      c.axl = bld2<T>(Rl0, voff, so);
      c.axh = bld2<T>(lastp ? Rw0 : Rl0, voff, lastp ? 0u : so + pbytes);
    }
  };
  auto pick = [](V2 v, int e) { return e ? v.y : v.x; };
  // relax cell e of the pair on plane p with the pair's coefficients cp
  auto relax_at = [&](int p, int e, const PairCoef<T>& cp) {
    const int qk = qk0 + e, sl = slot(p);
    T ap[3], am[3], al[3], ah[3];
    ap[0] = buf[slot(p + 1)][qj][qk];
    am[0] = buf[slot(p - 1)][qj][qk];
    ap[1] = buf[sl][qj + 1][qk];
    am[1] = buf[sl][qj - 1][qk];
    ap[2] = buf[sl][qj][qk + 1];
    am[2] = buf[sl][qj][qk - 1];
    if (ALPHAF) {
      al[0] = pick(cp.axl, e);
      ah[0] = pick(cp.axh, e);
      al[1] = pick(cp.ay, e);
      ah[1] = pick(cp.ayh, e);
      al[2] = pick(cp.az, e);
      ah[2] = e ? cp.azh : cp.az.y;
    } else {
      al[0] = al[1] = al[2] = ah[0] = ah[1] = ah[2] = k.alpha;
    }
    return relax_nd<T>(D2, pick(cp.res, e), LAMF ? pick(cp.lam, e) : k.lam_c, al, ah, ap, am);
  };
  // an A value: the cell, and sigma times it in the ghosts that mirror it
  auto put_a = [&](int p, int e, T v) {
    const int qk = qk0 + e, kc = pk + e;
    T(*P)[RK] = buf[slot(p)];
    P[qj][qk] = v;
    if (pj == 0) P[qj - 1][qk] = T(bc.sigma[3]) * v;
    if (pj == n - 1) P[qj + 1][qk] = T(bc.sigma[2]) * v;
    if (kc == 0) P[qj][qk - 1] = T(bc.sigma[5]) * v;
    if (kc == n - 1) P[qj][qk + 1] = T(bc.sigma[4]) * v;
    if (p == 0) buf[slot(-1)][qj][qk] = T(bc.sigma[1]) * v;
    if (p == n - 1) buf[slot(n)][qj][qk] = T(bc.sigma[0]) * v;
  };
  // the planes each stage covers, clipped to the level
  const int lo3 = max(0, i0 - 3), hi3 = min(n - 1, i1 + 2);
  const int lo2 = max(0, i0 - 2), hi2 = min(n - 1, i1 + 1);
  const int lo1 = max(0, i0 - 1), hi1 = min(n - 1, i1);

  V2 dreg[NLoad];
  // step s with the records of planes s (r0), s-1 (r1), s-2 (r2), s-3 (r3);
  // loads plane s+2 of the field and the record of plane s+1 (nx)
  auto step = [&](int s, const PairCoef<T>& r3, const PairCoef<T>& r2, const PairCoef<T>& r1,
                  const PairCoef<T>& r0, PairCoef<T>& nx) {
    commit_plane(s + 1, dreg);
    fetch_plane(s + 2, dreg);
    fetch_pair(s + 1, nx);
    __syncthreads();
    const int e = (s + pj + pk) & 1;  // the cell of the pair that is red on plane s
    const int w = e ? w1 : w0;
    if (w <= 3 && s >= lo3 && s <= hi3) put_a(s, e, relax_at(s, e, r0));
    if (w <= 2 && s - 1 >= lo2 && s - 1 <= hi2) put_a(s - 1, e, relax_at(s - 1, e, r1));
    if (w <= 1 && s - 2 >= lo1 && s - 2 <= hi1) buf[slot(s - 2)][qj][qk0 + e] = relax_at(s - 2, e, r2);
    if (w == 0 && s - 3 >= i0 && s - 3 < i1) {
      // B's black cell of plane s-3, written out with B's red one
      const int q = s - 3;
      const T v = relax_at(q, e, r3);
      const T r = buf[slot(q)][qj][qk0 + 1 - e];
      V2 out;
      out.x = e ? r : v;
      out.y = e ? v : r;
      bst2(Rnew, voff, (uint32_t)q * pbytes, out);
    }
  };

  {
    V2 tmp[NLoad];
    fetch_plane(i0 - 4, tmp);
    commit_plane(i0 - 4, tmp);
    fetch_plane(i0 - 3, tmp);
    commit_plane(i0 - 3, tmp);
  }
  fetch_plane(i0 - 2, dreg);
  PairCoef<T> R0{}, R1{}, R2{}, R3{}, R4{};
  fetch_pair(i0 - 3, R0);
  // steps i0-3 .. i1+2; the record of plane s is R[(s - i0 + 3) % 5]
  for (int s = i0 - 3; s <= i1 + 2; s += 5) {
    step(s, R2, R3, R4, R0, R1);
    if (s + 1 > i1 + 2) break;
    step(s + 1, R3, R4, R0, R1, R2);
    if (s + 2 > i1 + 2) break;
    step(s + 2, R4, R0, R1, R2, R3);
    if (s + 3 > i1 + 2) break;
    step(s + 3, R0, R1, R2, R3, R4);
    if (s + 4 > i1 + 2) break;
    step(s + 4, R1, R2, R3, R4, R0);
  }
}

// --- correction, residual and first restriction in one pass, 3D --------
//
// a_new = a + da on the finest level (the correction that ends mg_cycle,
// poisson.h:84-89; none when da is null), the residual of a_new (as
// residual_cell, poisson.h:373-389) and its restriction to the next level
// (restriction_average, children in foreach_child order), streaming along i
// like k_rb_stream3c: a block owns a 16 x 64 tile of (j,k) and a chunk of an
// even number of planes.  a_new of planes p-1..p+1 sits in LDS with a 1-cell
// halo, recomputed from a and da by every block that needs it, so blocks never
// exchange data; a_new goes to a second array, never in place.  Periodic
// directions wrap the halo.  Partial max |res| and NaN counts per block go to
// `part` as in k_residual_restrict.

constexpr int kRRJ = kSTJ + 2, kRRK = kSTK + 4;           // a_new region: 1-row halo, pairs from k0-2
constexpr int kRRegion2 = kRRJ * kRRK / 2;                 // vector slots (612)
constexpr int kRNLoad = (kRRegion2 + kSBlock - 1) / kSBlock;  // 2

// what the residual of a pair of cells (k, k+1) on one plane reads besides a
// (the x-wall values are read on the two wall planes only)
struct ResCoef {
  double2 b, lam, axl, axh, ay, ayh, az;
  double azh;
  double2 cy, cz;  // values of the y and z walls the cells touch (0 elsewhere)
};

template <typename TO>
struct Vec2Of;
template <>
struct Vec2Of<float> {
  using type = float2;
};
template <>
struct Vec2Of<double> {
  using type = double2;
};

template <typename TO, typename TD, bool ALPHAF, bool LAMF>
__global__ __launch_bounds__(kSBlock) __attribute__((amdgpu_waves_per_eu(2)))
void k_corr_residual3(const double* __restrict__ a, const TD* __restrict__ da, double* __restrict__ anew,
                      const double* __restrict__ b, TO* __restrict__ res, double* __restrict__ res64,
                      TO* __restrict__ coarse, Coef<double> k, Geo<3> g, BCA bc, double D, int nchunks,
                      Part* __restrict__ part) {
  using TD2 = typename Vec2<TD>::type;
  using TO2 = typename Vec2Of<TO>::type;
  __shared__ __align__(16) double A[4][kRRJ][kRRK];
  __shared__ double smax[kSBlock / 64];
  __shared__ unsigned long long snan[kSBlock / 64];
  const int n = g.n, tid = threadIdx.x;
  const size_t blk = logical_block();
  const int nkt = n / kSTK, njt = n / kSTJ;
  const int kt = (int)(blk % nkt), jt = (int)((blk / nkt) % njt), ct = (int)(blk / ((size_t)nkt * njt));
  // chunks of whole coarse planes
  const int i0 = 2 * (int)((size_t)ct * (n / 2) / nchunks), i1 = 2 * (int)((size_t)(ct + 1) * (n / 2) / nchunks);
  const int j0 = jt * kSTJ, k0 = kt * kSTK;
  const int pr = tid >> 5, pm = tid & 31;
  const int bj = j0 + pr, bk = k0 + 2 * pm;
  const int qj = pr + 1, qk = 2 * pm + 2;  // LDS position of cell bk
  const bool corr = da != nullptr;
  const bool px = bc.periodic[0] != 0, py = bc.periodic[1] != 0, pz = bc.periodic[2] != 0;
  auto slot = [&](int p) { return (p - i0 + 8) & 3; };
  auto wrapv = [&](int v) { return v < 0 ? v + n : (v >= n ? v - n : v); };
  const size_t plane = (size_t)n * n;
  const uint32_t pb8 = (uint32_t)(plane * 8), rb8 = (uint32_t)n * 8, vb8 = (uint32_t)(plane * n * 8);
  const uint32_t cell0 = (uint32_t)bj * n + bk;  // the pair within a plane
  auto clampn = [&](int v) { return v < 0 ? 0 : (v > n - 1 ? n - 1 : v); };
  // one resource per array (empty for a null array); a plane is selected by
  // the uniform offset of plane p clamped into the level
  auto arr3 = [&](const void* arr, uint32_t es) { return make_rsrc(arr, arr ? (uint32_t)(plane * n * es) : 0u); };
  auto arr2 = [&](const double* arr) { return make_rsrc(arr, arr ? pb8 : 0u); };
  const Rsrc Ra = arr3(a, 8), Rd = arr3(corr ? (const void*)da : nullptr, sizeof(TD)), Rb = arr3(b, 8);
  const Rsrc Rlam = arr3(LAMF ? k.lam : nullptr, 8);
  const Rsrc Rl0 = arr3(ALPHAF ? k.lo[0] : nullptr, 8), Rl1 = arr3(ALPHAF ? k.lo[1] : nullptr, 8);
  const Rsrc Rl2 = arr3(ALPHAF ? k.lo[2] : nullptr, 8);
  const Rsrc Rw0 = arr2(ALPHAF ? k.wall[0] : nullptr), Rw1 = arr2(ALPHAF ? k.wall[1] : nullptr);
  const Rsrc Rw2 = arr2(ALPHAF ? k.wall[2] : nullptr);
  const Rsrc Rc2 = arr2(py ? nullptr : bc.c[2]), Rc3 = arr2(py ? nullptr : bc.c[3]);
  const Rsrc Rc4 = arr2(pz ? nullptr : bc.c[4]), Rc5 = arr2(pz ? nullptr : bc.c[5]);
  (void)vb8;

  // a (and da) of plane p over the tile and its 1-cell halo
  auto fetch_region = [&](int p, double2* va, TD2* vd) {
    const uint32_t pp = (uint32_t)clampn(px ? wrapv(p) : p);
#pragma unroll
    for (int m = 0; m < kRNLoad; m++) {
      const int q = tid + m * kSBlock;
      const int row = q / (kRRK / 2), c2 = q - row * (kRRK / 2);
      const int gj = j0 - 1 + row, gk = k0 - 2 + 2 * c2;
      const bool ok = q < kRRegion2 && (py || (gj >= 0 && gj < n)) && (pz || (gk >= 0 && gk < n));
      const uint32_t c = (uint32_t)wrapv(gj) * n + wrapv(gk);
      va[m] = bld2<double>(Ra, ok ? c * 8 : kOOB, pp * pb8);
      vd[m] = bld2<TD>(Rd, ok ? c * (uint32_t)sizeof(TD) : kOOB, pp * (uint32_t)(plane * sizeof(TD)));
    }
  };
  auto commit_region = [&](int p, const double2* va, const TD2* vd) {
    if (!px && (p < 0 || p >= n)) return;
    double(*P)[kRRK] = A[slot(p)];
#pragma unroll
    for (int m = 0; m < kRNLoad; m++) {
      const int q = tid + m * kSBlock;
      if (q < kRRegion2) {
        const int row = q / (kRRK / 2), c2 = q - row * (kRRK / 2);
        double2 v = va[m];
        if (corr) {
          v.x += (double)vd[m].x;
          v.y += (double)vd[m].y;
        }
        *reinterpret_cast<double2*>(&P[row][2 * c2]) = v;
      }
    }
  };
  // the coefficients of the pair on plane p (clamped into the level)
  auto fetch_coef3 = [&](int p, ResCoef& c) {
    const uint32_t off = cell0 * 8, so = (uint32_t)clampn(p) * pb8, sw = (uint32_t)clampn(p) * rb8;
    c.b = bld2<double>(Rb, off, so);
    if (LAMF) c.lam = bld2<double>(Rlam, off, so);
    if (ALPHAF) {
      const bool lastj = bj == n - 1, lastk = bk + 2 == n, lastp = clampn(p) == n - 1;
      c.axl = bld2<double>(Rl0, off, so);
      c.axh = bld2<double>(lastp ? Rw0 : Rl0, off, lastp ? 0u : so + pb8);
      c.ay = bld2<double>(Rl1, off, so);
      const double2 yl = bld2<double>(Rl1, lastj ? kOOB : off + rb8, so);
      const double2 yw = bld2<double>(Rw1, lastj ? bk * 8 : kOOB, sw);
      c.ayh.x = bor(yl.x, yw.x);
      c.ayh.y = bor(yl.y, yw.y);
      c.az = bld2<double>(Rl2, off, so);
      c.azh = bor(bld1<double>(Rl2, lastk ? kOOB : off + 16, so), bld1<double>(Rw2, lastk ? bj * 8 : kOOB, sw));
    }
    // boundary values: y walls are rows of lanes, z walls single cells
    const bool yp = !py && bj == n - 1, ym = !py && bj == 0;
    const double2 y1 = bld2<double>(Rc2, yp ? bk * 8 : kOOB, sw);
    const double2 y2 = bld2<double>(Rc3, ym ? bk * 8 : kOOB, sw);
    c.cy.x = bor(y1.x, y2.x);
    c.cy.y = bor(y1.y, y2.y);
    c.cz.x = bld1<double>(Rc5, !pz && bk == 0 ? bj * 8 : kOOB, sw);
    c.cz.y = bld1<double>(Rc4, !pz && bk + 1 == n - 1 ? bj * 8 : kOOB, sw);
  };
  auto pick = [](double2 v, int e) { return e ? v.y : v.x; };
  auto term = [&](double r, double al, double ah, double a0, double am, double ap) {
    if (bc.pow2) return r + (al * ((a0 - am) * bc.invD) - ah * ((ap - a0) * bc.invD)) * bc.invD;
    return r + (al * ((a0 - am) / D) - ah * ((ap - a0) / D)) / D;
  };

  double macc = 0.;
  unsigned long long nacc = 0;
  TO acc = TO(0);
  double2 ra[kRNLoad];
  TD2 rd[kRNLoad];
  // step p: residual of plane p (coefficients cur); loads the region of
  // plane p+2 and the coefficients of plane p+1 (next)
  auto step = [&](int p, const ResCoef& cur, ResCoef& next) {
    commit_region(p + 1, ra, rd);
    fetch_region(p + 2, ra, rd);
    fetch_coef3(p + 1, next);
    // the x-wall values, on the two wall planes only
    double2 cx;
    cx.x = cx.y = 0.;
    if (!px && (p == 0 || p == n - 1)) cx = bld2<double>(arr2(p == n - 1 ? bc.c[0] : bc.c[1]), cell0 * 8);
    __syncthreads();
    const int s = slot(p), sp = slot(p + 1), sm = slot(p - 1);
    TO ro[2];
    double a0s[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      const int kk = bk + e;
      const double a0 = A[s][qj][qk + e];
      a0s[e] = a0;
      const double lam = LAMF ? pick(cur.lam, e) : k.lam_c;
      double r = pick(cur.b, e) - lam * a0;
      {
        const bool gp = !px && p == n - 1, gm = !px && p == 0;
        const double vp = gp ? a0 : A[sp][qj][qk + e], vm = gm ? a0 : A[sm][qj][qk + e];
        const double ap = gp ? pick(cx, e) + double(bc.sigma[0]) * vp : vp;
        const double am = gm ? pick(cx, e) + double(bc.sigma[1]) * vm : vm;
        const double ah = ALPHAF ? pick(cur.axh, e) : k.alpha, al = ALPHAF ? pick(cur.axl, e) : k.alpha;
        r = term(r, al, ah, a0, am, ap);
      }
      {
        const bool gp = !py && bj == n - 1, gm = !py && bj == 0;
        const double vp = gp ? a0 : A[s][qj + 1][qk + e], vm = gm ? a0 : A[s][qj - 1][qk + e];
        const double ap = gp ? pick(cur.cy, e) + double(bc.sigma[2]) * vp : vp;
        const double am = gm ? pick(cur.cy, e) + double(bc.sigma[3]) * vm : vm;
        const double ah = ALPHAF ? pick(cur.ayh, e) : k.alpha, al = ALPHAF ? pick(cur.ay, e) : k.alpha;
        r = term(r, al, ah, a0, am, ap);
      }
      {
        const bool gp = !pz && kk == n - 1, gm = !pz && kk == 0;
        const double vp = gp ? a0 : A[s][qj][qk + e + 1], vm = gm ? a0 : A[s][qj][qk + e - 1];
        const double ap = gp ? pick(cur.cz, e) + double(bc.sigma[4]) * vp : vp;
        const double am = gm ? pick(cur.cz, e) + double(bc.sigma[5]) * vm : vm;
        const double ah = ALPHAF ? (e ? cur.azh : cur.az.y) : k.alpha, al = ALPHAF ? pick(cur.az, e) : k.alpha;
        r = term(r, al, ah, a0, am, ap);
      }
      ro[e] = (TO)r;
      if (r != r) nacc++;
      else macc = fmax(macc, fabs(r));
      if (res64) res64[(size_t)p * plane + cell0 + e] = r;
    }
    TO2 o;
    o.x = ro[0];
    o.y = ro[1];
    *reinterpret_cast<TO2*>(res + (size_t)p * plane + cell0) = o;
    if (corr) {
      double2 w;
      w.x = a0s[0];
      w.y = a0s[1];
      *reinterpret_cast<double2*>(anew + (size_t)p * plane + cell0) = w;
    }
    if (coarse) {
      // children (p&1, row parity, k parity) summed in foreach_child order;
      // the odd row of a pair of rows is 32 lanes up
      const TO u0 = __shfl_down(ro[0], 32, 64), u1 = __shfl_down(ro[1], 32, 64);
      if ((pr & 1) == 0) {
        if ((p & 1) == 0) acc = TO(0);
        acc += ro[0];
        acc += ro[1];
        acc += u0;
        acc += u1;
        if (p & 1) {
          const int nc = n / 2;
          coarse[((size_t)(p / 2) * nc + bj / 2) * nc + bk / 2] = acc / TO(8);
        }
      }
    }
  };

  // prologue: planes i0-1 and i0 in LDS, plane i0+1 in registers
  {
    double2 ta[kRNLoad];
    TD2 td[kRNLoad];
    fetch_region(i0 - 1, ta, td);
    commit_region(i0 - 1, ta, td);
    fetch_region(i0, ta, td);
    commit_region(i0, ta, td);
  }
  fetch_region(i0 + 1, ra, rd);
  ResCoef C0{}, C1{};
  fetch_coef3(i0, C0);
  for (int p = i0; p < i1; p += 2) {
    step(p, C0, C1);
    step(p + 1, C1, C0);
  }
  // block partials of max |res| and the NaN count
  macc = wave_max(macc);
  for (int o = 32; o > 0; o >>= 1) nacc += __shfl_xor(nacc, o, 64);
  const int lane = tid & 63, w = tid >> 6;
  if (lane == 0) {
    smax[w] = macc;
    snan[w] = nacc;
  }
  __syncthreads();
  if (tid == 0) {
    double mm = smax[0];
    unsigned long long ns = snan[0];
    for (int i = 1; i < kSBlock / 64; i++) {
      mm = fmax(mm, smax[i]);
      ns += snan[i];
    }
    part[blockIdx.x] = Part{mm, ns};
  }
}

// --- all coarse levels in one workgroup ----------------------------------
//
// Levels minlevel..top of the cycle (each at most 4096 cells) run in one
// block of 1024 threads: restriction of res from `top` down to minlevel,
// then from minlevel up to top the zero / bilinear initial guess and nrelax
// sweeps, with a block barrier between steps.  Same arithmetic as the
// per-level kernels.

// levels 0..12 at most (4096 cells in 1D)
constexpr int kCoarseLevels = 13;

template <typename T, int DIM>
struct CoarseLevel {
  T* da;
  T* da2;  // Jacobi output
  T* res;
  Coef<T> k;
  Geo<DIM> g;
  BCH bc;
  T D2;
};

template <typename T, int DIM>
struct CoarseArgs {
  CoarseLevel<T, DIM> L[kCoarseLevels];
  int minlevel, top, nrelax, jacobi, gs;  // gs: the in-place Gauss-Seidel order, through da2
};

template <int DIM, typename T>
__device__ void coarse_relax_cell(const CoarseLevel<T, DIM>& L, idx_t idx, int colour_or_jacobi) {
  int c[3];
  L.g.decode(idx, c);
  T self = L.da[idx];
  T v = relax_value<DIM, T>(L.da, L.res, L.k, L.g, L.bc, idx, c, self, L.D2);
  if (colour_or_jacobi < 0) L.da2[idx] = (self + T(2) * v) / T(3);
  else L.da[idx] = v;
}

template <int DIM, typename T>
__global__ __launch_bounds__(1024) void k_coarse_cycle(CoarseArgs<T, DIM> A) {
  const int t = threadIdx.x;
  for (int l = A.top - 1; l >= A.minlevel; l--) {
    const CoarseLevel<T, DIM>& C = A.L[l];
    const CoarseLevel<T, DIM>& F = A.L[l + 1];
    idx_t nc = C.g.ncell();
    for (idx_t idx = t; idx < nc; idx += (idx_t)blockDim.x) {
      int c[3];
      C.g.decode(idx, c);
      int f[3] = {2 * c[0], 2 * c[1], 2 * c[2]};
      idx_t b = F.g.index(f);
      T sum = T(0);
      if (DIM == 1) {
        sum += F.res[b];
        sum += F.res[b + 1];
      } else if (DIM == 2) {
        idx_t sx = F.g.stride(0);
        sum += F.res[b];
        sum += F.res[b + 1];
        sum += F.res[b + sx];
        sum += F.res[b + sx + 1];
      } else {
        idx_t sx = F.g.stride(0), sy = F.g.stride(1);
        sum += F.res[b];
        sum += F.res[b + 1];
        sum += F.res[b + sy];
        sum += F.res[b + sy + 1];
        sum += F.res[b + sx];
        sum += F.res[b + sx + 1];
        sum += F.res[b + sx + sy];
        sum += F.res[b + sx + sy + 1];
      }
      C.res[idx] = sum / T(1 << DIM);
    }
    __syncthreads();
  }
  for (int l = A.minlevel; l <= A.top; l++) {
    const CoarseLevel<T, DIM>& L = A.L[l];
    idx_t nl = L.g.ncell();
    if (l == A.minlevel) {
      for (idx_t idx = t; idx < nl; idx += (idx_t)blockDim.x) L.da[idx] = T(0);
    } else {
      const CoarseLevel<T, DIM>& C = A.L[l - 1];
      for (idx_t idx = t; idx < nl; idx += (idx_t)blockDim.x) {
        int c[3];
        L.g.decode(idx, c);
        T v = bilinear_at<DIM, T>(C.da, C.g, C.bc, c);
        L.da[idx] = v;
      }
    }
    __syncthreads();
    for (int it = 0; it < A.nrelax; it++) {
      if (A.gs) {
        gs_sweep_block<DIM, T>(L.da, L.da2, L.res, L.k, L.g, L.bc, L.D2);
        for (idx_t idx = t; idx < nl; idx += (idx_t)blockDim.x) L.da[idx] = L.da2[idx];
        __syncthreads();
      } else if (A.jacobi) {
        for (idx_t idx = t; idx < nl; idx += (idx_t)blockDim.x) coarse_relax_cell<DIM, T>(L, idx, -1);
        __syncthreads();
        for (idx_t idx = t; idx < nl; idx += (idx_t)blockDim.x) L.da[idx] = L.da2[idx];
        __syncthreads();
      } else {
        for (int colour = 0; colour < 2; colour++) {
          for (idx_t idx = t; idx < nl; idx += (idx_t)blockDim.x) {
            int c[3];
            L.g.decode(idx, c);
            if (((c[0] + c[1] + c[2]) & 1) == colour) coarse_relax_cell<DIM, T>(L, idx, colour);
          }
          __syncthreads();
        }
      }
    }
  }
}

}  // namespace bmg
