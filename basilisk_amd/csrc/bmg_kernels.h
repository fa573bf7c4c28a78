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
};

// inhomogeneous boundary of `a` on the finest level: ghost = c + sigma * s
struct BCA {
  int periodic[3];
  int sigma[6];
  const double* c[6];  // n^(DIM-1) values per side, or null for 0
};

template <int DIM>
struct Geo {
  int n, log2n;
  __device__ __forceinline__ void decode(size_t idx, int c[3]) const {
    if (DIM == 3) {
      c[2] = (int)(idx & (size_t)(n - 1));
      c[1] = (int)((idx >> log2n) & (size_t)(n - 1));
      c[0] = (int)(idx >> (2 * log2n));
    } else if (DIM == 2) {
      c[2] = 0;
      c[1] = (int)(idx & (size_t)(n - 1));
      c[0] = (int)(idx >> log2n);
    } else {
      c[0] = (int)idx;
      c[1] = c[2] = 0;
    }
  }
  __device__ __forceinline__ size_t stride(int d) const {
    if (DIM == 3) return d == 0 ? (size_t)n * n : d == 1 ? (size_t)n : 1;
    if (DIM == 2) return d == 0 ? (size_t)n : 1;
    return 1;
  }
  __device__ __forceinline__ size_t index(const int c[3]) const {
    if (DIM == 3) return ((size_t)c[0] * n + c[1]) * n + c[2];
    if (DIM == 2) return (size_t)c[0] * n + c[1];
    return (size_t)c[0];
  }
  // index on the wall normal to d (tangential coordinates in order)
  __device__ __forceinline__ size_t wall(int d, const int c[3]) const {
    if (DIM == 3) return d == 0 ? (size_t)c[1] * n + c[2] : d == 1 ? (size_t)c[0] * n + c[2] : (size_t)c[0] * n + c[1];
    if (DIM == 2) return d == 0 ? (size_t)c[1] : (size_t)c[0];
    return 0;
  }
  __device__ __forceinline__ size_t ncell() const {
    return DIM == 3 ? (size_t)n * n * n : DIM == 2 ? (size_t)n * n : (size_t)n;
  }
};

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of logical blocks so that neighbouring rows and planes of
// a stencil share one L2.
__device__ __forceinline__ size_t logical_block() {
  size_t b = blockIdx.x, nb = gridDim.x;
  if ((nb & 7) == 0) b = (b & 7) * (nb >> 3) + (b >> 3);
  return b;
}

// value of field f at the neighbour of cell c along d (dir +1 or -1), with
// the homogeneous ghost mapping
template <int DIM, typename T>
__device__ __forceinline__ T nbr_h(const T* __restrict__ f, const Geo<DIM>& g, const BCH& bc,
                                   size_t idx, const int c[3], int d, int dir, T self) {
  size_t s = g.stride(d);
  if (dir > 0) {
    if (c[d] < g.n - 1) return f[idx + s];
    if (bc.periodic[d]) return f[idx - (size_t)(g.n - 1) * s];
    return T(bc.sigma[2 * d]) * self;
  } else {
    if (c[d] > 0) return f[idx - s];
    if (bc.periodic[d]) return f[idx + (size_t)(g.n - 1) * s];
    return T(bc.sigma[2 * d + 1]) * self;
  }
}

// alpha on the high face of cell c along d
template <int DIM, typename T>
__device__ __forceinline__ T alpha_hi(const T* __restrict__ lo, const T* __restrict__ wall,
                                      const Geo<DIM>& g, size_t idx, const int c[3], int d) {
  return c[d] < g.n - 1 ? lo[idx + g.stride(d)] : wall[g.wall(d, c)];
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
__device__ __forceinline__ T relax_value(const T* __restrict__ da, const T* __restrict__ res,
                                         const Coef<T>& k, const Geo<DIM>& g, const BCH& bc,
                                         size_t idx, const int c[3], T self, T D2) {
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
  size_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int c[3];
  g.decode(idx, c);
  T self = da[idx];
  T cval = relax_value<DIM, T>(da, res, k, g, bc, idx, c, self, D2);
  out[idx] = (self + T(2) * cval) / T(3);
}

// one colour of a red-black Gauss-Seidel sweep, in place; one thread per
// cell of that colour (colour = parity of i+j+k)
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_relax_rb(T* __restrict__ da, const T* __restrict__ res,
                                                     Coef<T> k, Geo<DIM> g, BCH bc, T D2, int colour) {
  size_t t = logical_block() * kBlock + threadIdx.x;
  size_t half = g.n > 1 ? g.ncell() / 2 : (colour == 0 ? 1 : 0);
  if (t >= half) return;
  size_t idx;
  if (g.n > 1) {
    // rows along the last dimension hold n/2 cells of each colour
    size_t row = t >> (g.log2n - 1);
    int kk = (int)(t & (size_t)((g.n >> 1) - 1));
    int c0[3];
    g.decode(row << g.log2n, c0);
    int par = DIM == 3 ? (c0[0] + c0[1]) & 1 : DIM == 2 ? c0[0] & 1 : 0;
    idx = (row << g.log2n) + 2 * kk + ((colour + par) & 1);
  } else {
    idx = 0;
  }
  int c[3];
  g.decode(idx, c);
  T self = da[idx];
  da[idx] = relax_value<DIM, T>(da, res, k, g, bc, idx, c, self, D2);
}

// coarse = mean of the 2^DIM children, restriction_average
// (grid/multigrid-common.h:26-32) in foreach_child order (k fastest)
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_restrict(const T* __restrict__ fine, T* __restrict__ coarse,
                                                     Geo<DIM> gc) {
  size_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gc.ncell()) return;
  int c[3];
  gc.decode(idx, c);
  Geo<DIM> gf{gc.n * 2, gc.log2n + 1};
  int f[3] = {2 * c[0], 2 * c[1], 2 * c[2]};
  size_t b = gf.index(f);
  T sum = T(0);
  if (DIM == 1) {
    sum += fine[b];
    sum += fine[b + 1];
  } else if (DIM == 2) {
    size_t sx = gf.stride(0);
    sum += fine[b];
    sum += fine[b + 1];
    sum += fine[b + sx];
    sum += fine[b + sx + 1];
  } else {
    size_t sx = gf.stride(0), sy = gf.stride(1);
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

// fine = bilinear interpolation of the coarse level, grid/multigrid-common.h:124-140,
// child offsets grid/multigrid.h:175-188
template <int DIM, typename T>
__global__ __launch_bounds__(kBlock) void k_prolong(const T* __restrict__ coarse, T* __restrict__ fine,
                                                    Geo<DIM> gf, BCH bcc) {
  size_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gf.ncell()) return;
  int c[3];
  gf.decode(idx, c);
  Geo<DIM> gc{gf.n >> 1, gf.log2n - 1};
  int p[3] = {c[0] >> 1, c[1] >> 1, c[2] >> 1};
  int o[3] = {2 * (c[0] & 1) - 1, 2 * (c[1] & 1) - 1, 2 * (c[2] & 1) - 1};
  T v;
  if (DIM == 1) {
    int q0[3] = {p[0], 0, 0}, q1[3] = {p[0] + o[0], 0, 0};
    v = (T(3) * coarse_at<DIM, T>(coarse, gc, bcc, q0) + coarse_at<DIM, T>(coarse, gc, bcc, q1)) / T(4);
  } else if (DIM == 2) {
    int q00[3] = {p[0], p[1], 0}, q10[3] = {p[0] + o[0], p[1], 0};
    int q01[3] = {p[0], p[1] + o[1], 0}, q11[3] = {p[0] + o[0], p[1] + o[1], 0};
    T c00 = coarse_at<DIM, T>(coarse, gc, bcc, q00);
    T c10 = coarse_at<DIM, T>(coarse, gc, bcc, q10);
    T c01 = coarse_at<DIM, T>(coarse, gc, bcc, q01);
    T c11 = coarse_at<DIM, T>(coarse, gc, bcc, q11);
    v = (T(9) * c00 + T(3) * (c10 + c01) + c11) / T(16);
  } else {
    int q000[3] = {p[0], p[1], p[2]};
    int q100[3] = {p[0] + o[0], p[1], p[2]};
    int q010[3] = {p[0], p[1] + o[1], p[2]};
    int q001[3] = {p[0], p[1], p[2] + o[2]};
    int q110[3] = {p[0] + o[0], p[1] + o[1], p[2]};
    int q101[3] = {p[0] + o[0], p[1], p[2] + o[2]};
    int q011[3] = {p[0], p[1] + o[1], p[2] + o[2]};
    int q111[3] = {p[0] + o[0], p[1] + o[1], p[2] + o[2]};
    T c000 = coarse_at<DIM, T>(coarse, gc, bcc, q000);
    T c100 = coarse_at<DIM, T>(coarse, gc, bcc, q100);
    T c010 = coarse_at<DIM, T>(coarse, gc, bcc, q010);
    T c001 = coarse_at<DIM, T>(coarse, gc, bcc, q001);
    T c110 = coarse_at<DIM, T>(coarse, gc, bcc, q110);
    T c101 = coarse_at<DIM, T>(coarse, gc, bcc, q101);
    T c011 = coarse_at<DIM, T>(coarse, gc, bcc, q011);
    T c111 = coarse_at<DIM, T>(coarse, gc, bcc, q111);
    v = (T(27) * c000 + T(9) * (c100 + c010 + c001) + T(3) * (c110 + c101 + c011) + c111) / T(64);
  }
  fine[idx] = v;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(T* __restrict__ f, size_t n, T v) {
  size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) f[idx] = v;
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_convert(const TI* __restrict__ in, TO* __restrict__ out, size_t n) {
  size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (idx < n) out[idx] = (TO)in[idx];
}

// a += da on the finest level, poisson.h:84-89
template <typename T>
__global__ __launch_bounds__(kBlock) void k_correct(double* __restrict__ a, const T* __restrict__ da, size_t n) {
  size_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx < n) a[idx] += (double)da[idx];
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

// residual of `a` on the finest level, poisson.h:373-389 (!TREE), with the
// inhomogeneous ghost of `a`; writes res in the cycle precision TO
template <int DIM, typename TO>
__global__ __launch_bounds__(kBlock) void k_residual(const double* __restrict__ a, const double* __restrict__ b,
                                                     TO* __restrict__ res, double* __restrict__ res64,
                                                     Coef<double> k, Geo<DIM> g, BCA bc, double D, Stat* st) {
  size_t idx = logical_block() * kBlock + threadIdx.x;
  double m = 0.;
  if (idx < g.ncell()) {
    int c[3];
    g.decode(idx, c);
    double a0 = a[idx];
    double lam = k.lam ? k.lam[idx] : k.lam_c;
    double r = b[idx] - lam * a0;
#pragma unroll
    for (int d = 0; d < DIM; d++) {
      size_t s = g.stride(d);
      double ap, am;
      if (c[d] < g.n - 1) ap = a[idx + s];
      else if (bc.periodic[d]) ap = a[idx - (size_t)(g.n - 1) * s];
      else {
        const double* cv = bc.c[2 * d];
        ap = (cv ? cv[g.wall(d, c)] : 0.) + double(bc.sigma[2 * d]) * a0;
      }
      if (c[d] > 0) am = a[idx - s];
      else if (bc.periodic[d]) am = a[idx + (size_t)(g.n - 1) * s];
      else {
        const double* cv = bc.c[2 * d + 1];
        am = (cv ? cv[g.wall(d, c)] : 0.) + double(bc.sigma[2 * d + 1]) * a0;
      }
      double ah, al;
      if (k.lo[0]) {
        ah = alpha_hi<DIM, double>(k.lo[d], k.wall[d], g, idx, c, d);
        al = k.lo[d][idx];
      } else {
        ah = al = k.alpha;
      }
      r += (al * ((a0 - am) / D) - ah * ((ap - a0) / D)) / D;
    }
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

// --- face coefficients (FP64, once per poisson() call) ------------------

// coarse low faces of component d: face_average, grid/multigrid-common.h:42-58,
// with the tangential order that the reference's dimension rotation gives
// each component
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_restrict_face_lo(const double* __restrict__ fine, double* __restrict__ coarse,
                                                             Geo<DIM> gc, int d) {
  size_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gc.ncell()) return;
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

// coarse high-wall faces of component d from the fine high-wall faces
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_restrict_face_wall(const double* __restrict__ fine, double* __restrict__ coarse,
                                                               int nc, int d) {
  size_t idx = (size_t)blockIdx.x * kBlock + threadIdx.x;
  size_t nw = DIM == 3 ? (size_t)nc * nc : DIM == 2 ? (size_t)nc : 1;
  if (idx >= nw) return;
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

}  // namespace bmg
