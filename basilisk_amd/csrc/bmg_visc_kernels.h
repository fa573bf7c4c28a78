// libbmg: kernels of viscosity() on the uniform multigrid grid, the second
// client of mg_solve (viscosity.h:27-178), FP64.
//
//   relax_viscosity     viscosity.h:27-96 (non-AXI, lambda = 1): the JACOBI
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
  double dt, D;
};

// component field f at cell x + o (|o_d| <= 1), with the wall ghosts
template <int DIM>
__device__ __forceinline__ double v_at(const double* __restrict__ f, const Geo<DIM>& g, const int* sig,
                                       const int x[3], const int o[3]) {
  int q[3] = {x[0] + o[0], x[1] + o[1], x[2] + o[2]};
  double s = 1.0;
#pragma unroll
  for (int d = 0; d < DIM; d++) {
    if (q[d] < 0) {
      q[d] = 0;
      s *= (double)sig[2 * d + 1];
    } else if (q[d] >= g.n) {
      q[d] = g.n - 1;
      s *= (double)sig[2 * d];
    }
  }
  return s * f[g.index(q)];
}

template <int DIM>
__device__ __forceinline__ double mu_face(const ViscArgs& A, const Geo<DIM>& g, idx_t idx, const int x[3], int d,
                                          bool hi) {
  if (!A.mlo[d]) return A.mu_c;
  return hi ? alpha_hi<DIM, double>(A.mlo[d], A.mwall[d], g, idx, x, d) : A.mlo[d][idx];
}

// the bracket of relax_viscosity (RES = false: the neighbour terms of the
// numerator) or of residual_viscosity (RES = true) for component c
template <int DIM, bool RES>
__device__ __forceinline__ double visc_terms(const ViscArgs& A, const Geo<DIM>& g, idx_t idx, const int x[3],
                                             int c, double u0) {
  auto U = [&](int comp, int d1, int s1, int d2, int s2) {
    int o[3] = {0, 0, 0};
    o[d1] += s1;
    o[d2] += s2;
    return v_at<DIM>(A.u[comp], g, A.sig[comp], x, o);
  };
  const double mch = mu_face<DIM>(A, g, idx, x, c, true), mcl = mu_face<DIM>(A, g, idx, x, c, false);
  double S;
  if (RES)
    S = 2. * mch * (U(c, c, 1, c, 0) - u0) - 2. * mcl * (u0 - U(c, c, -1, c, 0));
  else
    S = 2. * mch * U(c, c, 1, c, 0) + 2. * mcl * U(c, c, -1, c, 0);
#pragma unroll
  for (int k = 1; k < DIM; k++) {
    const int e = (c + k) % DIM;
    const double meh = mu_face<DIM>(A, g, idx, x, e, true), mel = mu_face<DIM>(A, g, idx, x, e, false);
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

template <int DIM>
__device__ __forceinline__ double visc_diag(const ViscArgs& A, const Geo<DIM>& g, idx_t idx, const int x[3], int c) {
  double s = 2. * mu_face<DIM>(A, g, idx, x, c, true) + 2. * mu_face<DIM>(A, g, idx, x, c, false);
#pragma unroll
  for (int k = 1; k < DIM; k++) {
    const int e = (c + k) % DIM;
    s = s + mu_face<DIM>(A, g, idx, x, e, true) + mu_face<DIM>(A, g, idx, x, e, false);
  }
  return s;
}

// one sweep of relax_viscosity over a level: colour < 0, Jacobi,
// out = (u + 2 w)/3; colour 0 or 1, the red-black pass of that parity,
// out = w there and u elsewhere
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_visc_relax(ViscArgs A, Geo<DIM> g, int colour) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int x[3];
  g.decode(idx, x);
  const bool upd = colour < 0 || ((x[0] + x[1] + x[2]) & 1) == colour;
  const double D2 = A.D * A.D;
  const double dtr = A.dt / (A.rho ? A.rho[idx] : A.rho_c);
#pragma unroll
  for (int c = 0; c < DIM; c++) {
    const double u0 = A.u[c][idx];
    double v = u0;
    if (upd) {
      const double w = (dtr * visc_terms<DIM, false>(A, g, idx, x, c, u0) + A.r[c][idx] * D2) /
                       (D2 * 1. + dtr * visc_diag<DIM>(A, g, idx, x, c));
      v = colour < 0 ? (u0 + 2. * w) / 3. : w;
    }
    A.out[c][idx] = v;
  }
}

// residual_viscosity on the finest level: res = r - u + dt/rho S / D^2,
// max |res| over the components to st (NaN counted)
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_visc_residual(ViscArgs A, Geo<DIM> g, Stat* st) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  double m = 0.;
  if (idx < g.ncell()) {
    int x[3];
    g.decode(idx, x);
    const double dtr = A.dt / (A.rho ? A.rho[idx] : A.rho_c);
#pragma unroll
    for (int c = 0; c < DIM; c++) {
      const double u0 = A.u[c][idx];
      const double v = (A.r[c][idx] - 1. * u0) + dtr * visc_terms<DIM, true>(A, g, idx, x, c, u0) / (A.D * A.D);
      A.out[c][idx] = v;
      if (v != v) atomicAdd(&st->nans, 1ull);
      else m = fmax(m, fabs(v));
    }
  }
  block_max_to(st, m);
}

}  // namespace bmg
