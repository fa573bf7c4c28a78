// libbmg: kernels of viscosity() on the uniform multigrid grid, the second
// client of mg_solve (viscosity.h:27-178), FP64.
//
//   relax_viscosity     viscosity.h:27-96 (lambda = 1; with AXI, 2D, the y
//                       component's lambda = 1 + dt/rho (mu.x[] + mu.x[1] +
//                       mu.y[] + mu.y[0,1])/2/y^2, viscosity.h:12-14): the
//                       default in-place Gauss-Seidel in the reference's
//                       i-major order (hyperplanes, see k_visc_gs_block), the JACOBI
//                       variant, w from the old iterate for every component,
//                       then u = (u + 2 w)/3; or red-black Gauss-Seidel (an
//                       extension standing in for the sequential in-place
//                       default): the cells with i+j+k of one parity take w,
//                       read from the state before the pass, the others
//                       keep their value (one pass per colour, out of place:
//                       the cross terms read same-colour diagonal cells)
//   residual_viscosity  viscosity.h:98-160, the "naive" (!TREE) form
//
// Component c couples to the other components e through the cross terms of
// div(2 mu D(u)); e runs over c+1, c+2 (mod dim), the order the reference's
// dimension rotation gives each component.  Ghosts: every component is
// sigma times the mirrored cell on each wall (walls at rest: antisymmetry
// -1 for the normal component, symmetry +1 for the tangential ones, or -1
// for no slip); an edge or corner ghost is the product of the crossed
// sides' sigmas times the diagonal interior cell, the value the
// multigrid's direction-by-direction ghost fill leaves there
// (grid/multigrid.h:543-594).  Every expression keeps the reference's
// evaluation order; with -ffp-contract=off the results equal the oracle's
// bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include "bmg_kernels.h"

namespace bmg {

struct ViscArgs {
  const double* u[3];     // iterate (da on the cycle's levels, u for the residual)
  double* out[3];         // relax: new iterate; residual: res
  const double* r[3];     // right-hand side (res of the level, or r = u_old)
  const double* mlo[3];   // mu: low face of each cell per component (null: mu_c)
  const double* mwall[3]; // mu: faces on the high wall
  double mu_c;
  const double* rho;      // null: rho_c
  double rho_c;
  int sig[3][6];          // per component: right, left, top, bottom, front, back
  int periodic[3];        // periodic dimensions (their neighbours wrap around)
  double dt, D;
  int axi;                // 2D axisymmetric: x the axis, y the radius
  double y0;              // the box's lower y (cell centres at (j + 1/2) D + y0)
};

// the values a cell's stencils read: component m at offset o (|o_d| <= 1),
// slot (o_x+1)*9 + (o_y+1)*3 + (o_z+1) in 3D, (o_x+1)*3 + (o_y+1) in 2D.
// Component m is read at the centre and its face neighbours (as the
// component being relaxed) and, for every other component c, at o_c = +-1
// with o_m in {-1, 0, 1} (the cross terms): 15 values in 3D, 7 in 2D.
template <int DIM>
struct VNb {
  static constexpr int NS = DIM == 3 ? 27 : 9;
  double v[DIM][NS];
  static __device__ __forceinline__ constexpr int slot(int ox, int oy, int oz) {
    return DIM == 3 ? (ox + 1) * 9 + (oy + 1) * 3 + (oz + 1) : (ox + 1) * 3 + (oy + 1);
  }
  static __device__ __forceinline__ constexpr bool needed(int m, int ox, int oy, int oz) {
    const int o[3] = {ox, oy, oz};
    int nz = 0;
    for (int d = 0; d < 3; d++) nz += o[d] != 0;
    if (nz <= 1) return true;  // centre and face neighbours
    if (nz > 2 || o[m] == 0) return false;
    return true;  // o_m = +-1 and one other axis +-1
  }
};

// per-axis addressing of the 3^dim neighbourhood of cell x: the clamped
// offset along each axis, and which of its -1 / +1 neighbours are ghosts
// (their factor, a component's sigma of that side, is applied at the load)
template <int DIM>
struct VAxes {
  idx_t off[3][3];
  int lo, hi;  // bit d: the -1 / +1 neighbour along d is a ghost
  __device__ __forceinline__ void init(const Geo<DIM>& g, const int x[3], const int* periodic) {
    lo = hi = 0;
#pragma unroll
    for (int d = 0; d < 3; d++)
#pragma unroll
      for (int o = -1; o <= 1; o++) {
        int q = d < DIM ? x[d] + o : 0;
        if (d < DIM && periodic[d]) q = q < 0 ? q + g.n : (q >= g.n ? q - g.n : q);
        const bool l = d < DIM && q < 0, h = d < DIM && q >= g.n;
        if (l) lo |= 1 << d;
        if (h) hi |= 1 << d;
        q = l ? 0 : h ? g.n - 1 : q;
        off[d][o + 1] = d < DIM ? (idx_t)q * g.stride(d) : 0u;
      }
  }
  // the ghost factor of component m at offset o along d
  __device__ __forceinline__ int factor(const ViscArgs& A, int m, int d, int o) const {
    return o < 0 ? ((lo >> d) & 1 ? A.sig[m][2 * d + 1] : 1) : o > 0 ? ((hi >> d) & 1 ? A.sig[m][2 * d] : 1) : 1;
  }
};

template <int DIM>
__device__ __forceinline__ void visc_load(const ViscArgs& A, const VAxes<DIM>& X, VNb<DIM>& N) {
#pragma unroll
  for (int m = 0; m < DIM; m++)
#pragma unroll
    for (int ox = -1; ox <= 1; ox++)
#pragma unroll
      for (int oy = -1; oy <= 1; oy++)
#pragma unroll
        for (int oz = (DIM == 3 ? -1 : 0); oz <= (DIM == 3 ? 1 : 0); oz++)
          if (VNb<DIM>::needed(m, ox, oy, oz)) {
            const idx_t a = X.off[0][ox + 1] + X.off[1][oy + 1] + X.off[2][oz + 1];
            const int s = X.factor(A, m, 0, ox) * X.factor(A, m, 1, oy) * X.factor(A, m, 2, oz);
            N.v[m][VNb<DIM>::slot(ox, oy, oz)] = (double)s * A.u[m][a];
          }
}

// the same values for the Gauss-Seidel sweep (relax_viscosity without
// -DJACOBI: w = u, each cell overwritten as the i-major loop visits it): a
// neighbour before this cell in that order (the first non-zero offset is
// -1) that lies inside the level has its new value (A.out), every other one
// and every ghost (wrapped or mirrored) the sweep's start value (A.u)
template <int DIM>
__device__ __forceinline__ void visc_load_gs(const ViscArgs& A, const VAxes<DIM>& X, const int x[3], int n,
                                             VNb<DIM>& N) {
#pragma unroll
  for (int m = 0; m < DIM; m++)
#pragma unroll
    for (int ox = -1; ox <= 1; ox++)
#pragma unroll
      for (int oy = -1; oy <= 1; oy++)
#pragma unroll
        for (int oz = (DIM == 3 ? -1 : 0); oz <= (DIM == 3 ? 1 : 0); oz++)
          if (VNb<DIM>::needed(m, ox, oy, oz)) {
            const int o[3] = {ox, oy, oz};
            const int first = ox ? ox : oy ? oy : oz;
            bool inside = true;
#pragma unroll
            for (int d = 0; d < DIM; d++)
              if (o[d]) inside = inside && x[d] + o[d] >= 0 && x[d] + o[d] < n;
            const idx_t a = X.off[0][ox + 1] + X.off[1][oy + 1] + X.off[2][oz + 1];
            const int s = X.factor(A, m, 0, ox) * X.factor(A, m, 1, oy) * X.factor(A, m, 2, oz);
            const double* src = first < 0 && inside ? A.out[m] : A.u[m];
            N.v[m][VNb<DIM>::slot(ox, oy, oz)] = (double)s * src[a];
          }
}

// mu on the low (0) and high (1) face of cell idx along each axis
template <int DIM>
__device__ __forceinline__ void visc_mu(const ViscArgs& A, const Geo<DIM>& g, idx_t idx, const int x[3],
                                        double mu[3][2]) {
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    if (A.mlo[d]) {
      mu[d][0] = A.mlo[d][idx];
      mu[d][1] = alpha_hi<DIM, double>(A.mlo[d], A.mwall[d], g, idx, x, d);
    } else {
      mu[d][0] = mu[d][1] = A.mu_c;
    }
  }
}

// the bracket of relax_viscosity (RES = false: the neighbour terms of the
// numerator) or of residual_viscosity (RES = true) for component c
template <int DIM, bool RES>
__device__ __forceinline__ double visc_terms(const VNb<DIM>& N, const double mu[3][2], int c) {
  auto U = [&](int comp, int d1, int s1, int d2, int s2) {
    int o[3] = {0, 0, 0};
    o[d1] += s1;
    o[d2] += s2;
    return N.v[comp][VNb<DIM>::slot(o[0], o[1], o[2])];
  };
  const double u0 = N.v[c][VNb<DIM>::slot(0, 0, 0)];
  const double mch = mu[c][1], mcl = mu[c][0];
  double S;
  if (RES)
    S = 2. * mch * (U(c, c, 1, c, 0) - u0) - 2. * mcl * (u0 - U(c, c, -1, c, 0));
  else
    S = 2. * mch * U(c, c, 1, c, 0) + 2. * mcl * U(c, c, -1, c, 0);
#pragma unroll
  for (int k = 1; k < DIM; k++) {
    const int e = (c + k) % DIM;
    const double meh = mu[e][1], mel = mu[e][0];
    const double up = (U(e, c, 1, e, 0) + U(e, c, 1, e, 1)) / 4.;
    const double um = (U(e, c, -1, e, 0) + U(e, c, -1, e, 1)) / 4.;
    const double lp = (U(e, c, 1, e, -1) + U(e, c, 1, e, 0)) / 4.;
    const double lm = (U(e, c, -1, e, -1) + U(e, c, -1, e, 0)) / 4.;
    if (RES) {
      S = S + meh * (U(c, e, 1, e, 0) - u0 + up - um);
      S = S - mel * (u0 - U(c, e, -1, e, 0) + lp - lm);
    } else {
      S = S + meh * (U(c, e, 1, e, 0) + up - um);
      S = S - mel * (-U(c, e, -1, e, 0) + lp - lm);
    }
  }
  return S;
}

// lambda of component c (viscosity.h:12-23): 1, or with AXI for the y
// component 1 + dt/rho (mu.x[] + mu.x[1] + mu.y[] + mu.y[0,1])/2/y^2 at the
// cell's centre, (j + 1/2) Delta + Y0 as the reference's `y`
template <int DIM>
__device__ __forceinline__ double visc_lambda(const ViscArgs& A, const double mu[3][2], double dtr, const int x[3],
                                              int c) {
  if (DIM != 2 || !A.axi || c != 1) return 1.;
  const double y = ((double)x[1] + 0.5) * A.D + A.y0;
  return 1. + dtr * (mu[0][0] + mu[0][1] + mu[1][0] + mu[1][1]) / 2. / (y * y);
}

template <int DIM>
__device__ __forceinline__ double visc_diag(const double mu[3][2], int c) {
  double s = 2. * mu[c][1] + 2. * mu[c][0];
#pragma unroll
  for (int k = 1; k < DIM; k++) {
    const int e = (c + k) % DIM;
    s = s + mu[e][1] + mu[e][0];
  }
  return s;
}

// w of relax_viscosity for every component of cell idx (GS: the values of
// the Gauss-Seidel sweep)
template <int DIM, bool GS = false>
__device__ __forceinline__ void visc_target(const ViscArgs& A, const Geo<DIM>& g, idx_t idx, const int x[3],
                                            double w[3]) {
  VAxes<DIM> X;
  X.init(g, x, A.periodic);
  VNb<DIM> N;
  if (GS)
    visc_load_gs<DIM>(A, X, x, g.n, N);
  else
    visc_load<DIM>(A, X, N);
  double mu[3][2];
  visc_mu<DIM>(A, g, idx, x, mu);
  const double D2 = A.D * A.D;
  const double dtr = A.dt / (A.rho ? A.rho[idx] : A.rho_c);
#pragma unroll
  for (int c = 0; c < DIM; c++)
    w[c] = (dtr * visc_terms<DIM, false>(N, mu, c) + A.r[c][idx] * D2) /
           (D2 * visc_lambda<DIM>(A, mu, dtr, x, c) + dtr * visc_diag<DIM>(mu, c));
}

// one Jacobi sweep of relax_viscosity over a level: out = (u + 2 w)/3
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_visc_jacobi(ViscArgs A, Geo<DIM> g) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int x[3];
  g.decode(idx, x);
  double w[3];
  visc_target<DIM>(A, g, idx, x, w);
#pragma unroll
  for (int c = 0; c < DIM; c++) A.out[c][idx] = (A.u[c][idx] + 2. * w[c]) / 3.;
}

// one red-black pass: the cells whose i+j+k has the parity `colour` take w,
// read from the pass input, the others are copied.  One thread per pair of
// cells along the last dimension (one of each parity, so no wave diverges);
// levels with n >= 2.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_visc_rb(ViscArgs A, Geo<DIM> g, int colour) {
  const idx_t t = logical_block() * kBlock + threadIdx.x;
  if (t >= g.ncell() / 2) return;
  const idx_t pair = 2 * t;  // the pair's first cell (k even)
  int x[3];
  g.decode(pair, x);
  const int par = (x[0] + x[1] + x[2]) & 1;
  const idx_t idx = par == colour ? pair : pair + 1;
  const idx_t other = par == colour ? pair + 1 : pair;
  x[DIM - 1] += par == colour ? 0 : 1;
  double w[3];
  visc_target<DIM>(A, g, idx, x, w);
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    A.out[c][idx] = w[c];
    A.out[c][other] = A.u[c][other];
  }
}

// The Gauss-Seidel sweep: a cell reads the new values of the cells before
// it in the i-major loop and the old values of the others; its stencils
// reach at most one cell along two axes, so with w = 4i + 2j + k (3D) or
// 2i + j (2D) every cell before (i, j, k) in that order has a smaller w and
// every cell after it a larger one: the cells of one hyperplane w = s are
// independent and the sweep runs hyperplane by hyperplane, A.u -> A.out,
// equal to the serial loop bit for bit.  Thread t of hyperplane s: (i, j) =
// (t / n, t % n) and k = s - 4i - 2j in 3D, i = t and j = s - 2i in 2D.
template <int DIM>
__device__ __forceinline__ bool visc_gs_cell(const Geo<DIM>& g, int s, idx_t t, int x[3]) {
  if (DIM == 3) {
    x[0] = (int)(t >> g.log2n);
    x[1] = (int)(t & (idx_t)(g.n - 1));
    x[2] = s - 4 * x[0] - 2 * x[1];
    return x[2] >= 0 && x[2] < g.n;
  }
  x[0] = (int)t;
  x[1] = s - 2 * x[0];
  x[2] = 0;
  return x[1] >= 0 && x[1] < g.n;
}

template <int DIM>
__device__ __forceinline__ void visc_gs_at(const ViscArgs& A, const Geo<DIM>& g, int s, idx_t t) {
  int x[3];
  if (!visc_gs_cell<DIM>(g, s, t, x)) return;
  const idx_t idx = g.index(x);
  double w[3];
  visc_target<DIM, true>(A, g, idx, x, w);
#pragma unroll
  for (int c = 0; c < DIM; c++) A.out[c][idx] = w[c];
}

// the hyperplanes (7 (n-1) + 1 in 3D, 3 (n-1) + 1 in 2D) in one workgroup
// with a barrier between them
template <int DIM>
__global__ __launch_bounds__(1024) void k_visc_gs_block(ViscArgs A, Geo<DIM> g) {
  const int smax = (DIM == 3 ? 7 : 3) * (g.n - 1);
  const idx_t count = DIM == 3 ? (idx_t)g.n * g.n : (idx_t)g.n;
  for (int s = 0; s <= smax; s++) {
    for (idx_t t = threadIdx.x; t < count; t += blockDim.x) visc_gs_at<DIM>(A, g, s, t);
    __syncthreads();
  }
}

// one hyperplane per launch (large levels)
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_visc_gs_plane(ViscArgs A, Geo<DIM> g, int s) {
  const idx_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= (DIM == 3 ? (idx_t)g.n * g.n : (idx_t)g.n)) return;
  visc_gs_at<DIM>(A, g, s, t);
}

// the single cell of level 0: either parity, the same arithmetic
template <int DIM>
__global__ void k_visc_one(ViscArgs A, Geo<DIM> g, int colour) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int x[3] = {0, 0, 0};
  double w[3];
  visc_target<DIM>(A, g, 0, x, w);
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    const double u0 = A.u[c][0];
    A.out[c][0] = colour < 0 ? (u0 + 2. * w[c]) / 3. : colour == 0 ? w[c] : u0;
  }
}

// residual_viscosity on the finest level: res = r - u + dt/rho S / D^2,
// max |res| over the components to st (NaN counted)
// (a grid-stride loop over a capped grid: one atomic max per block)
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_visc_residual(ViscArgs A, Geo<DIM> g, Stat* st) {
  double m = 0.;
  const idx_t n = g.ncell(), stride = (idx_t)gridDim.x * kBlock;
  for (idx_t idx = logical_block() * kBlock + threadIdx.x; idx < n; idx += stride) {
    int x[3];
    g.decode(idx, x);
    VAxes<DIM> X;
    X.init(g, x, A.periodic);
    VNb<DIM> N;
    visc_load<DIM>(A, X, N);
    double mu[3][2];
    visc_mu<DIM>(A, g, idx, x, mu);
    const double dtr = A.dt / (A.rho ? A.rho[idx] : A.rho_c);
#pragma unroll
    for (int c = 0; c < DIM; c++) {
      const double u0 = N.v[c][VNb<DIM>::slot(0, 0, 0)];
      const double v =
          (A.r[c][idx] - visc_lambda<DIM>(A, mu, dtr, x, c) * u0) + dtr * visc_terms<DIM, true>(N, mu, c) / (A.D * A.D);
      A.out[c][idx] = v;
      if (v != v) atomicAdd(&st->nans, 1ull);
      else m = fmax(m, fabs(v));
    }
  }
  block_max_to(st, m);
}

}  // namespace bmg
