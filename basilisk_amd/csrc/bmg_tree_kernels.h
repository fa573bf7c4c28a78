// Device kernels of the adaptive-tree path of libbmg.
//
// A tree level is stored as a compact array of "slots": the active cells of
// the level in the reference's cache order (pre-order / Morton), then the
// halo cells (children of the prolongation leaves of the level above), then
// every allocated boundary (ghost) cell.  All levels are concatenated, so a
// field is one array indexed by the global slot.  The host builds index
// tables (neighbours, children, 3^dim rings, boundary rules, face lists) once
// per mesh; every kernel is a gather over those tables.
//
// Arithmetic follows the reference expression by expression (relax
// poisson.h:303-320, residual 355-372, bilinear multigrid-common.h:124-140,
// restriction_average 26-32, face_average 42-58, halo_face
// tree-common.h:373-409, refine_face_x 464-511, box rules tree.h:1139-1258)
// and libbmg is compiled with -ffp-contract=off, so the FP64 path equals the
// CPU oracle bit for bit.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace bmgt {

constexpr int kB = 256;
constexpr int kSeg = 24;

// A launch over several groups of slots ("segments"): each segment is a
// contiguous slot range [start, start + count) when list is null, otherwise
// the slots list[0 .. count).  Blocks are assigned to segments by prefix.
struct Segs {
  int n;
  int first_block[kSeg + 1];
  int start[kSeg];
  int count[kSeg];
  int abase[kSeg];  // active-table index of slot lvoff (the level's first slot)
  int lvoff[kSeg];
  double D[kSeg];   // Delta of the segment's level
  const int* list[kSeg];
  // colour lists: the 2 dim neighbour slots of list entry t at nbl + 2 dim t
  // (the rows of the neighbour table in list order, read contiguously)
  const int* nbl[kSeg];
};

__device__ __forceinline__ int seg_of(const Segs& S) {
  int s = 0;
  while (s + 1 < S.n && (int)blockIdx.x >= S.first_block[s + 1]) s++;
  return s;
}

// slot handled by this thread, or -1
__device__ __forceinline__ int seg_slot(const Segs& S, int s) {
  int t = ((int)blockIdx.x - S.first_block[s]) * kB + (int)threadIdx.x;
  if (t >= S.count[s]) return -1;
  return S.list[s] ? S.list[s][t] : S.start[s] + t;
}

// one boundary (ghost) cell: normal rule (nterm 1), 2D diagonal (2), 3D
// diagonal (3) or no neighbour (0, the reference's `undefined`)
struct BoxRule {
  int ghost, src;
  int nterm;
  int side[3];
  int ti[3];  // index of the face value in the side's table
};

// boundary kinds of the field being filled
enum { BK_A = 0, BK_HOM = 1, BK_SYM = 2 };
enum { K_SYMMETRY = 0, K_DIRICHLET = 1, K_NEUMANN = 2 };

struct BCT {
  int kind[6];
  const double* v[6];
};

template <typename T>
__device__ __forceinline__ T bc_term(const BCT& bc, int bk, int side, int ti, double D, T s) {
  int kind = bk == BK_SYM ? K_SYMMETRY : bc.kind[side];
  if (kind == K_SYMMETRY) return s;
  if (bk == BK_HOM) return kind == K_DIRICHLET ? -s : s;
  double v = bc.v[side] ? bc.v[side][ti] : 0.;
  if (kind == K_DIRICHLET) return (T)(2. * v - (double)s);
  return (T)(D * v + (double)s);
}

template <typename T>
__device__ __forceinline__ void t_box_one(const BoxRule& q, T* __restrict__ f, const BCT& bc, int bk, double D) {
  T s = q.nterm ? f[q.src] : T(0);
  T v;
  if (q.nterm == 1)
    v = bc_term<T>(bc, bk, q.side[0], q.ti[0], D, s);
  else if (q.nterm == 2)
    v = bc_term<T>(bc, bk, q.side[0], q.ti[0], D, s) + bc_term<T>(bc, bk, q.side[1], q.ti[1], D, s) - s;
  else if (q.nterm == 3)
    v = bc_term<T>(bc, bk, q.side[0], q.ti[0], D, s) + bc_term<T>(bc, bk, q.side[1], q.ti[1], D, s) +
        bc_term<T>(bc, bk, q.side[2], q.ti[2], D, s) - T(2) * s;
  else
    v = (T)__builtin_nan("");
  f[q.ghost] = v;
}

template <typename T>
__global__ __launch_bounds__(kB) void k_t_box(const BoxRule* __restrict__ R, int n, T* __restrict__ f, BCT bc, int bk,
                                              double D) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  const BoxRule q = R[r];
  t_box_one<T>(q, f, bc, bk, D);
}

// restriction_average over the 2^dim children (child block in foreach_child
// order) of each record
template <typename T, int DIM>
__device__ __forceinline__ void t_restrict_one(T* __restrict__ f, int self, int child0) {
  const T* c = f + child0;
  T sum = T(0);
#pragma unroll
  for (int m = 0; m < (1 << DIM); m++) sum += c[m];
  f[self] = sum / T(1 << DIM);
}

template <typename T, int DIM>
__global__ __launch_bounds__(kB) void k_t_restrict(const int* __restrict__ self, const int* __restrict__ child0, int n,
                                                   T* __restrict__ f) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  t_restrict_one<T, DIM>(f, self[r], child0[r]);
}

// bilinear (grid/multigrid-common.h:124-140) of the 2^dim children of a
// record, in foreach_child order, from its 3^dim ring in src
template <typename T, int DIM>
__device__ __forceinline__ void t_bilinear_children(const int* __restrict__ g, const T* src, T* v) {
  if (DIM == 2) {
    T r[9];
#pragma unroll
    for (int m = 0; m < 9; m++) r[m] = src[g[m]];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        int cx = 2 * a - 1, cy = 2 * b - 1;
        T c0 = r[4], c1 = r[(1 + cx) * 3 + 1], c2 = r[3 + 1 + cy], c3 = r[(1 + cx) * 3 + 1 + cy];
        v[a * 2 + b] = (T(9) * c0 + T(3) * (c1 + c2) + c3) / T(16);
      }
  } else {
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++)
#pragma unroll
        for (int e = 0; e < 2; e++) {
          int cx = 2 * a - 1, cy = 2 * b - 1, cz = 2 * e - 1;
#define RG(x, y, z) src[g[((1 + (x)) * 3 + 1 + (y)) * 3 + 1 + (z)]]
          T c000 = RG(0, 0, 0), c100 = RG(cx, 0, 0), c010 = RG(0, cy, 0), c001 = RG(0, 0, cz);
          T c110 = RG(cx, cy, 0), c101 = RG(cx, 0, cz), c011 = RG(0, cy, cz), c111 = RG(cx, cy, cz);
#undef RG
          v[(a * 2 + b) * 2 + e] =
              (T(27) * c000 + T(9) * (c100 + c010 + c001) + T(3) * (c110 + c101 + c011) + c111) / T(64);
        }
  }
}

// bilinear into the 2^dim children of each record from its 3^dim ring
template <typename T, int DIM>
__device__ __forceinline__ void t_prolong_one(const int* __restrict__ g, int child0, T* f) {
  T v[1 << DIM];
  t_bilinear_children<T, DIM>(g, f, v);
#pragma unroll
  for (int m = 0; m < (1 << DIM); m++) f[child0 + m] = v[m];
}

template <typename T, int DIM>
__global__ __launch_bounds__(kB) void k_t_prolong(const int* __restrict__ ring, const int* __restrict__ child0, int n,
                                                  T* __restrict__ f) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  constexpr int NR = DIM == 2 ? 9 : 27;
  t_prolong_one<T, DIM>(ring + (size_t)r * NR, child0[r], f);
}

// per-dimension coefficient pointers of relax (null = constant)
template <typename T>
struct RelaxCoef {
  const T* alpha[3];
  T alpha_c;
  const T* lam;
  T lam_c;
};

// the value n/d of poisson.h:303-320 at slot gs (active index ai)
template <typename T, int DIM>
__device__ __forceinline__ T t_relax_value(const T* __restrict__ da, const T* __restrict__ res,
                                           const int* __restrict__ nb, const RelaxCoef<T>& k, int gs, T D2) {
  T n = -D2 * res[gs], d = -(k.lam ? k.lam[gs] : k.lam_c) * D2;
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    int lo = nb[2 * c], hi = nb[2 * c + 1];
    T ah = k.alpha[c] ? k.alpha[c][hi] : k.alpha_c, al = k.alpha[c] ? k.alpha[c][gs] : k.alpha_c;
    n += ah * da[hi] + al * da[lo];
    d += ah + al;
  }
  return n / d;
}

// relaxation of the cells of foreach_level_or_leaf(l).  colour < 0: all
// cells, value written to out (Jacobi's c); colour 0/1: in place, cells of
// that colour only; 2: in place, every cell of the segments (lists of one
// colour)
template <typename T, int DIM>
__global__ __launch_bounds__(kB) void k_t_relax(Segs S, T* __restrict__ da, T* __restrict__ out,
                                                const T* __restrict__ res, const int* __restrict__ nbr,
                                                const uint8_t* __restrict__ colour, RelaxCoef<T> k, int sel) {
  int s = seg_of(S);
  int gs = seg_slot(S, s);
  if (gs < 0) return;
  int ai = S.abase[s] + (gs - S.lvoff[s]);
  if ((sel == 0 || sel == 1) && colour[ai] != sel) return;
  T D = (T)S.D[s];
  const int t = ((int)blockIdx.x - S.first_block[s]) * kB + (int)threadIdx.x;
  const int* nb = S.nbl[s] ? S.nbl[s] + (size_t)t * 2 * DIM : nbr + (size_t)ai * 2 * DIM;
  T v = t_relax_value<T, DIM>(da, res, nb, k, gs, D * D);
  if (sel < 0)
    out[gs] = v;
  else
    da[gs] = v;
}

template <typename T>
__global__ __launch_bounds__(kB) void k_t_jacobi(Segs S, T* __restrict__ da, const T* __restrict__ c) {
  int s = seg_of(S);
  int gs = seg_slot(S, s);
  if (gs < 0) return;
  da[gs] = (da[gs] + T(2) * c[gs]) / T(3);
}

template <typename T>
__global__ __launch_bounds__(kB) void k_t_fill(Segs S, T* __restrict__ f, T v) {
  int s = seg_of(S);
  int gs = seg_slot(S, s);
  if (gs < 0) return;
  f[gs] = v;
}

// face fluxes of the residual, g.x[] = alpha.x[] * ((a[] - a[-1]) / Delta)
struct FaceOp {
  int gs, gm, d;
};

__global__ __launch_bounds__(kB) void k_t_faces(const FaceOp* __restrict__ F, int n, const double* __restrict__ a,
                                                double* g0, double* g1, double* g2, const double* al0,
                                                const double* al1, const double* al2, double alpha_c, double D) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  FaceOp q = F[r];
  double* g = q.d == 0 ? g0 : q.d == 1 ? g1 : g2;
  const double* al = q.d == 0 ? al0 : q.d == 1 ? al1 : al2;
  double alpha = al ? al[q.gs] : alpha_c;
  g[q.gs] = alpha * ((a[q.gs] - a[q.gm]) / D);
}

// mean of 2 (2D) or 4 (3D) fine faces into one coarse face (halo_face,
// face_average); the fine list is in the reference's summation order
struct FaceAvg {
  int target, d;
  int fine[4];
};

template <int DIM>
__global__ __launch_bounds__(kB) void k_t_face_avg(const FaceAvg* __restrict__ F, int n, double* f0, double* f1,
                                                   double* f2) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  FaceAvg q = F[r];
  double* f = q.d == 0 ? f0 : q.d == 1 ? f1 : f2;
  double s = 0.;
#pragma unroll
  for (int m = 0; m < (1 << (DIM - 1)); m++) s += f[q.fine[m]];
  f[q.target] = s / (double)(1 << (DIM - 1));
}

// refine_face_x for one prolongation leaf and one component d.  v[...] are
// slots at the leaf's level: self, self -+ t1, self -+ t2, next (= self + d),
// next -+ t1, next -+ t2; out[] the children faces (o, j, k) with o the
// position along d (0, 1, 2)
struct FaceProlong {
  int d;
  int flags;  // 1: low side written, 2: high side written
  int self, t1m, t1p, t2m, t2p;
  int next, nt1m, nt1p, nt2m, nt2p;
  int out[3][4];
};

template <int DIM>
__global__ __launch_bounds__(kB) void k_t_face_prolong(const FaceProlong* __restrict__ F, int n, double* f0,
                                                       double* f1, double* f2) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  const FaceProlong& q = F[r];
  double* v = q.d == 0 ? f0 : q.d == 1 ? f1 : f2;
  double vs = v[q.self], vn = v[q.next];
  if (q.flags & 1) {
    double g1 = (v[q.t1p] - v[q.t1m]) / 8.;
    if (DIM == 3) {
      double g2 = (v[q.t2p] - v[q.t2m]) / 8.;
      for (int j = 0; j <= 1; j++)
        for (int k = 0; k <= 1; k++) v[q.out[0][j * 2 + k]] = vs + (2 * j - 1) * g1 + (2 * k - 1) * g2;
    } else
      for (int j = 0; j <= 1; j++) v[q.out[0][j]] = vs + (2 * j - 1) * g1;
  }
  if (q.flags & 2) {
    double g1 = (v[q.nt1p] - v[q.nt1m]) / 8.;
    if (DIM == 3) {
      double g2 = (v[q.nt2p] - v[q.nt2m]) / 8.;
      for (int j = 0; j <= 1; j++)
        for (int k = 0; k <= 1; k++) v[q.out[2][j * 2 + k]] = vn + (2 * j - 1) * g1 + (2 * k - 1) * g2;
    } else
      for (int j = 0; j <= 1; j++) v[q.out[2][j]] = vn + (2 * j - 1) * g1;
  }
  {
    double g1 = (v[q.t1p] - v[q.t1m] + v[q.nt1p] - v[q.nt1m]) / 16.;
    if (DIM == 3) {
      double g2 = (v[q.t2p] - v[q.t2m] + v[q.nt2p] - v[q.nt2m]) / 16.;
      for (int j = 0; j <= 1; j++)
        for (int k = 0; k <= 1; k++)
          v[q.out[1][j * 2 + k]] = (vs + vn) / 2. + (2 * j - 1) * g1 + (2 * k - 1) * g2;
    } else
      for (int j = 0; j <= 1; j++) v[q.out[1][j]] = (vs + vn) / 2. + (2 * j - 1) * g1;
  }
}

// tangential faces of ghost cells (the face-vector part of box_boundary_level
// with symmetry BCs): a gather of pre-pass values, then a scatter
struct FaceCopy {
  int target, src, d;  // src < 0: zero
};

__global__ __launch_bounds__(kB) void k_t_face_gather(const FaceCopy* __restrict__ F, int n, const double* f0,
                                                      const double* f1, const double* f2, double* __restrict__ tmp) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  FaceCopy q = F[r];
  const double* f = q.d == 0 ? f0 : q.d == 1 ? f1 : f2;
  tmp[r] = q.src >= 0 ? f[q.src] : 0.;
}

__global__ __launch_bounds__(kB) void k_t_face_scatter(const FaceCopy* __restrict__ F, int n, double* f0, double* f1,
                                                       double* f2, const double* __restrict__ tmp) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  FaceCopy q = F[r];
  double* f = q.d == 0 ? f0 : q.d == 1 ? f1 : f2;
  f[q.target] = tmp[r];
}

// values given in face-cache order (a foreach_face assignment)
// --- project() on the tree (poisson.h:469-509) ---------------------------

// face-cache values out of the per-dimension slot arrays (k_t_face_set's
// inverse)
__global__ __launch_bounds__(kB) void k_t_face_get(const int* __restrict__ gs, const int* __restrict__ dd, long n,
                                                   double* __restrict__ v, const double* f0, const double* f1,
                                                   const double* f2) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  const double* f = dd[r] == 0 ? f0 : dd[r] == 1 ? f1 : f2;
  v[r] = f[gs[r]];
}

// divergence at the leaves: div = 0; div += uf.x[1] - uf.x[] per
// dimension; div /= dt*Delta -> b
template <int DIM>
__global__ __launch_bounds__(kB) void k_t_leaf_div(Segs S, const double* u0, const double* u1, const double* u2,
                                                   const int* __restrict__ nbr, double dt, double* __restrict__ b) {
  int s = seg_of(S);
  int gs = seg_slot(S, s);
  if (gs < 0) return;
  int ai = S.abase[s] + (gs - S.lvoff[s]);
  const int* nb = nbr + (size_t)ai * 2 * DIM;
  const double* u[3] = {u0, u1, u2};
  double div = 0.;
#pragma unroll
  for (int d = 0; d < DIM; d++) div += u[d][nb[2 * d + 1]] - u[d][gs];
  div /= dt * S.D[s];
  b[gs] = div;
}

// the face update on the faces of one level:
// uf.x[] -= dt*alpha.x[]*face_gradient_x (p, 0)
__global__ __launch_bounds__(kB) void k_t_face_update(const FaceOp* __restrict__ F, int n,
                                                      const double* __restrict__ p, double* u0, double* u1,
                                                      double* u2, const double* al0, const double* al1,
                                                      const double* al2, double alpha_c, double dt, double D) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  FaceOp q = F[r];
  double* u = q.d == 0 ? u0 : q.d == 1 ? u1 : u2;
  const double* al = q.d == 0 ? al0 : q.d == 1 ? al1 : al2;
  double alpha = al ? al[q.gs] : alpha_c;
  u[q.gs] -= dt * alpha * ((p[q.gs] - p[q.gm]) / D);
}

__global__ __launch_bounds__(kB) void k_t_face_set(const int* __restrict__ gs, const int* __restrict__ dd, long n,
                                                   const double* __restrict__ v, double* f0, double* f1, double* f2) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  double* f = dd[r] == 0 ? f0 : dd[r] == 1 ? f1 : f2;
  f[gs[r]] = v[r];
}

// statistics of the residual
struct TStat {
  unsigned long long maxbits;
  unsigned long long nans;
};

__global__ void k_t_stat_reset(TStat* s) {
  if (threadIdx.x == 0) {
    s->maxbits = 0ull;
    s->nans = 0ull;
  }
}

__device__ __forceinline__ double t_wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// leaf residual: res = b - lambda a - sum_d (g.x[1] - g.x[]) / Delta
template <int DIM, typename TO>
__global__ __launch_bounds__(kB) void k_t_leaf_res(Segs S, const double* __restrict__ a, const double* __restrict__ b,
                                                   const double* g0, const double* g1, const double* g2,
                                                   const int* __restrict__ nbr, const double* __restrict__ lam,
                                                   double lam_c, double* __restrict__ res64, TO* __restrict__ res,
                                                   TStat* st) {
  __shared__ double smax[kB / 64];
  __shared__ unsigned long long snan;
  if (threadIdx.x == 0) snan = 0ull;
  __syncthreads();
  int s = seg_of(S);
  int gs = seg_slot(S, s);
  double m = 0.;
  if (gs >= 0) {
    int ai = S.abase[s] + (gs - S.lvoff[s]);
    const int* nb = nbr + (size_t)ai * 2 * DIM;
    double D = S.D[s];
    double r = b[gs] - (lam ? lam[gs] : lam_c) * a[gs];
    const double* g[3] = {g0, g1, g2};
#pragma unroll
    for (int d = 0; d < DIM; d++) r -= (g[d][nb[2 * d + 1]] - g[d][gs]) / D;
    res64[gs] = r;
    res[gs] = (TO)r;
    if (r != r)
      atomicAdd(&snan, 1ull);
    else
      m = fabs(r);
  }
  m = t_wave_max(m);
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = smax[0];
    for (int i = 1; i < kB / 64; i++) mm = fmax(mm, smax[i]);
    if (mm > 0.) atomicMax(&st->maxbits, (unsigned long long)__double_as_longlong(mm));
    if (snan) atomicAdd(&st->nans, snan);
  }
}

template <typename T>
__global__ __launch_bounds__(kB) void k_t_correct(const int* __restrict__ leaves, long n, double* __restrict__ a,
                                                  const T* __restrict__ da) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  int gs = leaves[r];
  a[gs] += (double)da[gs];
}

// leaf values in cache order <-> slots
__global__ __launch_bounds__(kB) void k_t_leaf_scatter(const int* __restrict__ leaves, long n,
                                                       const double* __restrict__ v, double* __restrict__ f) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r < n) f[leaves[r]] = v[r];
}

template <typename T>
__global__ __launch_bounds__(kB) void k_t_leaf_gather(const int* __restrict__ leaves, long n, const T* __restrict__ f,
                                                      double* __restrict__ v) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r < n) v[r] = (double)f[leaves[r]];
}

template <typename TI, typename TO>
__global__ __launch_bounds__(kB) void k_t_convert(const TI* __restrict__ in, TO* __restrict__ out, long n) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r < n) out[r] = (TO)in[r];
}

// sum of b over the leaves: per-block partial sums, then one block
__global__ __launch_bounds__(kB) void k_t_sum_partial(const int* __restrict__ leaves, long n,
                                                      const double* __restrict__ b, double* __restrict__ part) {
  __shared__ double sp[kB / 64];
  double s = 0.;
  for (long r = (long)blockIdx.x * kB + threadIdx.x; r < n; r += (long)gridDim.x * kB) s += b[leaves[r]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sp[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = sp[0];
    for (int i = 1; i < kB / 64; i++) t += sp[i];
    part[blockIdx.x] = t;
  }
}

__global__ void k_t_sum_final(const double* __restrict__ part, int np, double* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.;
    for (int i = 0; i < np; i++) s += part[i];
    *out = s;
  }
}

// --- small operations batched into one workgroup --------------------------
//
// A tree cycle is a long chain of dependent operations, most of them on a
// few hundred cells (the ghost cells of a level, the coarse levels).  Each
// one as its own launch costs a couple of microseconds whatever its size, so
// consecutive small operations are queued and run by one 1024-thread
// workgroup, in order, with a barrier between them.  The operations and the
// arithmetic are those of the standalone kernels above.

enum { TOP_RESTRICT = 0, TOP_PROLONG = 1, TOP_BOX = 2, TOP_RELAX = 3, TOP_JACOBI = 4, TOP_FILL = 5 };
constexpr int kOpsBlock = 1024;
constexpr int kMaxOps = 250;

struct TOp {
  unsigned char code, field;
  signed char arg;  // prolong: halo; box: bc kind; relax: colour (-1: Jacobi value to tmp)
  unsigned char level;
};

struct TOps {
  int n;
  TOp op[kMaxOps];
};

// per-level extents of the index tables
struct TLevelInfo {
  int off, nA, abase, rec_off, n_nonleaf, n_prolong, box_off, box_n, leaf_off, leaf_n;
};

template <typename T>
struct TreeArgs {
  const TLevelInfo* lv;
  const int *rec_self, *rec_child0, *rec_ring, *nbr, *lvl_leaves;
  const uint8_t* colour;
  const BoxRule* box;
  T* fld[3];  // fields an operation can name (da, res, tmp; or `a` alone)
  const T* res;
  RelaxCoef<T> k;
  BCT bc;
  double L0;
};

template <typename T, int DIM>
__device__ __forceinline__ void t_ops_cells(const TreeArgs<T>& A, const TOp& op) {
  const int l = op.level;
  T* da = A.fld[0];
  for (int m = l; m >= 0; m--) {
    const TLevelInfo L = A.lv[m];
    const int n = m == l ? L.nA : L.leaf_n;
    const double Dm = A.L0 * (1. / (1 << m));
    for (int i = threadIdx.x; i < n; i += (int)blockDim.x) {
      const int gs = m == l ? L.off + i : A.lvl_leaves[L.leaf_off + i];
      const int ai = L.abase + (gs - L.off);
      if (op.code == TOP_FILL) {
        A.fld[op.field][gs] = T(0);
      } else if (op.code == TOP_JACOBI) {
        da[gs] = (da[gs] + T(2) * A.fld[2][gs]) / T(3);
      } else {
        if (op.arg >= 0 && A.colour[ai] != op.arg) continue;
        const T D = (T)Dm;
        T v = t_relax_value<T, DIM>(da, A.res, A.nbr + (size_t)ai * 2 * DIM, A.k, gs, D * D);
        if (op.arg < 0)
          A.fld[2][gs] = v;
        else
          da[gs] = v;
      }
    }
  }
}

template <typename T, int DIM>
__global__ __launch_bounds__(kOpsBlock) void k_t_ops(TOps P, TreeArgs<T> A) {
  constexpr int NR = DIM == 2 ? 9 : 27;
  for (int o = 0; o < P.n; o++) {
    const TOp op = P.op[o];
    const TLevelInfo L = A.lv[op.level];
    T* f = A.fld[op.field];
    switch (op.code) {
      case TOP_RESTRICT:
        for (int r = threadIdx.x; r < L.n_nonleaf; r += (int)blockDim.x)
          t_restrict_one<T, DIM>(f, A.rec_self[L.rec_off + r], A.rec_child0[L.rec_off + r]);
        break;
      case TOP_PROLONG: {
        const int r0 = L.rec_off + (op.arg ? L.n_nonleaf : 0);
        const int n = op.arg ? L.n_prolong : L.n_nonleaf;
        for (int r = threadIdx.x; r < n; r += (int)blockDim.x)
          t_prolong_one<T, DIM>(A.rec_ring + (size_t)(r0 + r) * NR, A.rec_child0[r0 + r], f);
        break;
      }
      case TOP_BOX: {
        const double D = A.L0 * (1. / (1 << op.level));
        for (int r = threadIdx.x; r < L.box_n; r += (int)blockDim.x) {
          const BoxRule q = A.box[L.box_off + r];
          t_box_one<T>(q, f, A.bc, op.arg, D);
        }
        break;
      }
      default:
        t_ops_cells<T, DIM>(A, op);
    }
    __syncthreads();
  }
}

// --- adapt_wavelet -------------------------------------------------------

// the wavelet estimate of adapt_wavelet (grid/tree-common.h:215-223): for
// every child of a refined cell, |value - bilinear prolongation of the
// parent's ring|
template <int DIM>
__global__ __launch_bounds__(kB) void k_t_wavelet(const int* __restrict__ ring, const int* __restrict__ child0, int n,
                                                  const double* __restrict__ f, double* __restrict__ err) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  constexpr int NR = DIM == 2 ? 9 : 27;
  double v[1 << DIM];
  t_bilinear_children<double, DIM>(ring + (size_t)r * NR, f, v);
  const int c0 = child0[r];
#pragma unroll
  for (int m = 0; m < (1 << DIM); m++) err[c0 + m] = fabs(f[c0 + m] - v[m]);
}

// a field carried across a mesh change: each slot of the new layout takes
// the value of the same cell in the old one (prev < 0: a new cell, zero)
__global__ __launch_bounds__(kB) void k_t_remap(const int* __restrict__ prev, long n, const double* __restrict__ old,
                                                double* __restrict__ out) {
  long s = (long)blockIdx.x * kB + threadIdx.x;
  if (s >= n) return;
  const int p = prev[s];
  out[s] = p >= 0 ? old[p] : 0.0;
}

// the children of the cells refined by adapt_wavelet: refine_bilinear of the
// parent's ring in the old layout (ring: old slots) into the new slots
template <int DIM>
__global__ __launch_bounds__(kB) void k_t_refine_remap(const int* __restrict__ ring, const int* __restrict__ child0,
                                                       int n, const double* __restrict__ old,
                                                       double* __restrict__ out) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  constexpr int NR = DIM == 2 ? 9 : 27;
  double v[1 << DIM];
  t_bilinear_children<double, DIM>(ring + (size_t)r * NR, old, v);
  const int c0 = child0[r];
#pragma unroll
  for (int m = 0; m < (1 << DIM); m++) out[c0 + m] = v[m];
}

}  // namespace bmgt

// --- the adjoint of poisson() on the tree ----------------------------------
//
// The tree operator r(x) = -lambda x - sum_d (g_d[hi] - g_d[])/Delta (g =
// halo_face of alpha (x[] - x[-1])/Delta, x after its tree boundary) is not
// symmetric, so its adjoint runs the transposes of the forward stages in
// reverse order: each stage that wrote y = f(x) becomes xbar += (df/dx)^T
// ybar, ybar = 0.  Where several writes reach one value, the kernels gather
// them through inverse tables in the forward order (face ops in face-cache
// order, box rules in boundary-cache order, bilinear terms in record,
// child, term order, relaxed cells in active-cache order), so every sum is
// accumulated in one fixed order; where at most two terms reach a value
// (the leaf divergence's two faces) an atomic add is exact in either order.
namespace bmgt {

constexpr int kAdjLevels = 24;

struct AdjLv {
  int nl;                   // levels 0 .. nl-1
  int off[kAdjLevels];      // first slot of the level
  int nslot[kAdjLevels];    // slots of the level
  int nA[kAdjLevels];       // active cells of the level (first in its slots)
  int abase[kAdjLevels];    // active index of the level's first active cell
  double D[kAdjLevels];     // Delta of the level
};

__device__ __forceinline__ int adj_level_of_slot(const AdjLv& L, int s) {
  int l = 0;
  while (l + 1 < L.nl && s >= L.off[l + 1]) l++;
  return l;
}
__device__ __forceinline__ int adj_level_of_ai(const AdjLv& L, int ai) {
  int l = 0;
  while (l + 1 < L.nl && ai >= L.abase[l + 1]) l++;
  return l;
}

// one bilinear term of a record's child: ring position and weight numerator
struct ProEnt {
  int rec;
  unsigned char child, term, pad0, pad1;
};

// homogeneous (sym = 0) or symmetry (sym = 1) factor of a box rule:
// ghost = k * src
__device__ __forceinline__ double box_factor(const BoxRule& q, const BCT& bc, int sym) {
  double h[3];
#pragma unroll
  for (int m = 0; m < 3; m++) h[m] = sym ? 1. : (bc.kind[q.side[m]] == K_DIRICHLET ? -1. : 1.);
  if (q.nterm == 1) return h[0];
  if (q.nterm == 2) return h[0] + h[1] - 1.;
  return h[0] + h[1] + h[2] - 2.;
}

// y (leaf order) -> xb[leaf] = -lambda y, gb_d[leaf] += y/Delta,
// gb_d[hi] -= y/Delta
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_div_T(AdjLv L, const int* __restrict__ leaf_gs, long nleaf,
                                                 const double* __restrict__ y, const int* __restrict__ nbr,
                                                 const double* __restrict__ lam, double lam_c, double* xb,
                                                 double* g0, double* g1, double* g2) {
  long n = (long)blockIdx.x * kB + threadIdx.x;
  if (n >= nleaf) return;
  const int gs = leaf_gs[n];
  const int l = adj_level_of_slot(L, gs);
  const int ai = L.abase[l] + (gs - L.off[l]);
  const int* nb = nbr + (size_t)ai * 2 * DIM;
  const double yv = y[n], D = L.D[l];
  xb[gs] = -(lam ? lam[gs] : lam_c) * yv;
  double* g[3] = {g0, g1, g2};
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const double tt = yv / D;
    atomicAdd(&g[d][gs], tt);
    atomicAdd(&g[d][nb[2 * d + 1]], -tt);
  }
}

// halo_face transposed: each target's adjoint to its fine faces, then 0
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_halo_face_T(const FaceAvg* __restrict__ F, int n, double* f0, double* f1,
                                                       double* f2) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  FaceAvg q = F[r];
  double* f = q.d == 0 ? f0 : q.d == 1 ? f1 : f2;
  const double v = f[q.target];
#pragma unroll
  for (int m = 0; m < (1 << (DIM - 1)); m++) f[q.fine[m]] += v / (double)(1 << (DIM - 1));
  f[q.target] = 0.;
}

// the face fluxes transposed: per slot, c = (alpha gb)/Delta of each face op
// whose gs (+c) or gm (-c) it is, in face op order
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_faces_T(AdjLv L, int nslot, const FaceOp* __restrict__ fops,
                                                   const int* __restrict__ fop_gs, const int* __restrict__ fop_gm,
                                                   const double* al0, const double* al1, const double* al2,
                                                   double alpha_c, const double* g0, const double* g1,
                                                   const double* g2, double* xb) {
  int s = blockIdx.x * kB + threadIdx.x;
  if (s >= nslot) return;
  int op[2 * DIM];
  int k = 0;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    const int a = fop_gs[s * DIM + d], b = fop_gm[s * DIM + d];
    if (a >= 0) op[k++] = a * 2;        // +c
    if (b >= 0) op[k++] = b * 2 + 1;    // -c
  }
  if (!k) return;
  // face op order (insertion sort of at most 2 dim entries)
  for (int i = 1; i < k; i++) {
    const int v = op[i];
    int j = i - 1;
    while (j >= 0 && op[j] > v) {
      op[j + 1] = op[j];
      j--;
    }
    op[j + 1] = v;
  }
  const double D = L.D[adj_level_of_slot(L, s)];
  const double* al[3] = {al0, al1, al2};
  const double* g[3] = {g0, g1, g2};
  double v = xb[s];
  for (int i = 0; i < k; i++) {
    const FaceOp q = fops[op[i] >> 1];
    const double alpha = al[q.d] ? al[q.d][q.gs] : alpha_c;
    const double c = (alpha * g[q.d][q.gs]) / D;
    if (op[i] & 1)
      v -= c;
    else
      v += c;
  }
  xb[s] = v;
}

// box rules of one level transposed, gather part: each source slot of the
// level adds k * x[ghost] of the rules naming it, in rule order
__global__ __launch_bounds__(kB) void k_ta_box_T(int s0, int n, const int* __restrict__ inv_off,
                                                 const int* __restrict__ inv, const BoxRule* __restrict__ R, BCT bc,
                                                 int sym, double* x) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  const int s = s0 + r;
  const int b = inv_off[s], e = inv_off[s + 1];
  if (b == e) return;
  double v = x[s];
  for (int i = b; i < e; i++) {
    const BoxRule q = R[inv[i]];
    v += box_factor(q, bc, sym) * x[q.ghost];
  }
  x[s] = v;
}

__global__ __launch_bounds__(kB) void k_ta_box_zero(const BoxRule* __restrict__ R, int n, double* x) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  x[R[r].ghost] = 0.;
}

// bilinear transposed, gather part: each ring slot of the level adds
// (w * x[child])/den over the terms naming it, of the records in [r0, r1)
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_bilinear_T(int s0, int n, const int* __restrict__ inv_off,
                                                      const ProEnt* __restrict__ inv,
                                                      const int* __restrict__ child0, int r0, int r1, double* x) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  const int s = s0 + r;
  const int b = inv_off[s], e = inv_off[s + 1];
  if (b == e) return;
  double v = x[s];
  bool any = false;
  for (int i = b; i < e; i++) {
    const ProEnt p = inv[i];
    if (p.rec < r0 || p.rec >= r1) continue;
    const double cb = x[child0[p.rec] + p.child];
    double w;
    if (DIM == 2)
      w = p.term == 0 ? 9. : (p.term == 3 ? 1. : 3.);
    else
      w = p.term == 0 ? 27. : (p.term <= 3 ? 9. : (p.term <= 6 ? 3. : 1.));
    v += (w * cb) / (DIM == 2 ? 16. : 64.);
    any = true;
  }
  if (any) x[s] = v;
}

// the children of the records [r0, r1) cleared (after their bilinear_T)
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_children_zero(const int* __restrict__ child0, int r0, int n, double* x) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  const int c = child0[r0 + r];
#pragma unroll
  for (int m = 0; m < (1 << DIM); m++) x[c + m] = 0.;
}

// restriction transposed: each parent's adjoint to its children, then 0
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_restrict_T(const int* __restrict__ self, const int* __restrict__ child0,
                                                      int n, double* x) {
  int r = blockIdx.x * kB + threadIdx.x;
  if (r >= n) return;
  const double pb = x[self[r]];
  const int c = child0[r];
#pragma unroll
  for (int m = 0; m < (1 << DIM); m++) x[c + m] += pb / (double)(1 << DIM);
  x[self[r]] = 0.;
}

// the relax denominator of the cell at active index ai (slot gs)
template <int DIM>
__device__ __forceinline__ double adj_relax_den(const int* __restrict__ nbr, int ai, int gs, const double* const* al,
                                                double alpha_c, const double* lam, double lam_c, double D2) {
  const int* nb = nbr + (size_t)ai * 2 * DIM;
  double d = -(lam ? lam[gs] : lam_c) * D2;
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    const double ah = al[c] ? al[c][nb[2 * c + 1]] : alpha_c, alo = al[c] ? al[c][gs] : alpha_c;
    d += ah + alo;
  }
  return d;
}

// one red-black pass of relax (level lrel, colour) transposed, gather part:
// each slot of levels 0..lrel adds (alpha v)/d of the relaxing cells that
// read it, in active order
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_relax_T(AdjLv L, int nrecv, const int* __restrict__ inv_off,
                                                   const int* __restrict__ inv, const int* __restrict__ nbr,
                                                   const uint8_t* __restrict__ colour,
                                                   const uint8_t* __restrict__ slot_leaf, int lrel, int col,
                                                   const double* al0, const double* al1, const double* al2,
                                                   double alpha_c, const double* lam, double lam_c, double* db) {
  int s = blockIdx.x * kB + threadIdx.x;
  if (s >= nrecv) return;
  const int b = inv_off[s], e = inv_off[s + 1];
  if (b == e) return;
  const double* al[3] = {al0, al1, al2};
  double v = db[s];
  bool any = false;
  for (int i = b; i < e; i++) {
    const int ai = inv[i] >> 3, dir = inv[i] & 7;
    if (colour[ai] != col) continue;
    const int m = adj_level_of_ai(L, ai);
    const int gs = L.off[m] + (ai - L.abase[m]);
    if (m > lrel || (m < lrel && !slot_leaf[gs])) continue;
    const double D2 = L.D[m] * L.D[m];
    const double den = adj_relax_den<DIM>(nbr, ai, gs, al, alpha_c, lam, lam_c, D2);
    const int c = dir >> 1;
    const double a = (dir & 1) ? (al[c] ? al[c][nbr[(size_t)ai * 2 * DIM + dir]] : alpha_c)
                               : (al[c] ? al[c][gs] : alpha_c);
    v += (a * db[gs]) / den;
    any = true;
  }
  if (any) db[s] = v;
}

// ... and the relaxing cells themselves: rb += (-Delta^2 v)/d, db = 0
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_relax_cells_T(Segs S, const int* __restrict__ nbr, const double* al0,
                                                         const double* al1, const double* al2, double alpha_c,
                                                         const double* lam, double lam_c, double* db, double* rb) {
  int sg = seg_of(S);
  int gs = seg_slot(S, sg);
  if (gs < 0) return;
  const int ai = S.abase[sg] + (gs - S.lvoff[sg]);
  const double* al[3] = {al0, al1, al2};
  const double D2 = S.D[sg] * S.D[sg];
  const double den = adj_relax_den<DIM>(nbr, ai, gs, al, alpha_c, lam, lam_c, D2);
  rb[gs] += (-D2 * db[gs]) / den;
  db[gs] = 0.;
}

// values at the leaves, leaf order <-> slots
__global__ __launch_bounds__(kB) void k_ta_leaf_get(const int* __restrict__ leaf_gs, long n, const double* x,
                                                    double* out) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r < n) out[r] = x[leaf_gs[r]];
}
__global__ __launch_bounds__(kB) void k_ta_leaf_set(const int* __restrict__ leaf_gs, long n, const double* in,
                                                    double* x) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r < n) x[leaf_gs[r]] = in[r];
}
__global__ __launch_bounds__(kB) void k_ta_leaf_mark(const int* __restrict__ leaf_gs, long n, uint8_t* mark) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r < n) mark[leaf_gs[r]] = 1;
}

// psi += m; rho = g + t; max |rho| into st
__global__ __launch_bounds__(kB) void k_ta_update(long n, double* psi, const double* m, const double* g,
                                                  const double* tt, double* rho, TStat* st, int step) {
  __shared__ double smax[kB / 64];
  long r = (long)blockIdx.x * kB + threadIdx.x;
  double v = 0.;
  if (r < n) {
    if (step == 0) {
      psi[r] += m[r];
    } else {
      const double x = g[r] + tt[r];
      rho[r] = x;
      v = fabs(x);
    }
  }
  if (step == 0) return;
  v = t_wave_max(v);
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = smax[0];
    for (int i = 1; i < kB / 64; i++) mm = fmax(mm, smax[i]);
    if (mm > 0.) atomicMax(&st->maxbits, (unsigned long long)__double_as_longlong(mm));
    if (mm != mm) atomicAdd(&st->nans, 1ull);
  }
}

// dJ/dalpha per face value: gb_d[gs] (a[gs] - a[gm])/Delta
template <int DIM>
__global__ __launch_bounds__(kB) void k_ta_galpha(AdjLv L, long nv, const int* __restrict__ fv_gs,
                                                  const int* __restrict__ fv_d, const int* __restrict__ fop_gs,
                                                  const FaceOp* __restrict__ fops, const double* a, const double* g0,
                                                  const double* g1, const double* g2, double* out) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r >= nv) return;
  const int gs = fv_gs[r], d = fv_d[r];
  const FaceOp q = fops[fop_gs[gs * DIM + d]];
  const double D = L.D[adj_level_of_slot(L, gs)];
  const double* g = d == 0 ? g0 : d == 1 ? g1 : g2;
  out[r] = g[gs] * ((a[gs] - a[q.gm]) / D);
}

// dJ/dlambda = -psi a at the leaves
__global__ __launch_bounds__(kB) void k_ta_glambda(const int* __restrict__ leaf_gs, long n, const double* psi,
                                                   const double* a, double* out) {
  long r = (long)blockIdx.x * kB + threadIdx.x;
  if (r < n) out[r] = -psi[r] * a[leaf_gs[r]];
}

}  // namespace bmgt
