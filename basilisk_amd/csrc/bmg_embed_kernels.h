// Embedded boundaries (embed.h) for the 2D multigrid grid, FP64: the EMBED
// branch of relax and residual (poisson.h:277-279, 309-318, 351-352,
// 380-385), embed_flux (embed.h:654-702) with dirichlet_gradient
// (embed.h:371-457), the embedded face gradients of the residual when the
// field's `third` attribute is set (embed.h:61-76, 179-195), and the
// multigrid prolongation bilinear_embed (embed.h:738-757).
//
// Everything that depends only on the volume and face fractions -- the
// fragment's length and centroid, the normal, which interpolation stencil
// dirichlet_gradient takes and where -- is worked out on the host once per
// set-up (EmbRec per cut cell, a code per partial face); the kernels
// evaluate the field-dependent part with the reference's expressions in its
// order.
#pragma once

#include "bmg_kernels.h"

namespace bmg {

constexpr double kSEPS = 1e-30;  // embed.h:38-39

// one cut cell (0 < cs < 1) of a level
struct EmbRec {
  int mode;          // dirichlet_gradient: 0 degenerate, 1 second order, 2 third order
  int off[2][3];     // cell offsets of s[i,j-1], s[i,j], s[i,j+1] for the two points
  double y1[2];      // interpolation coordinate of each point
  double d[2];       // distance of each point; mode 0: max(1e-3, |p.x/n.x|)
  double area, fa;   // fragment length; sum of the cell's four face fractions
};

struct EmbLevel {
  const int* cell;               // per cell: record index, or -1
  const EmbRec* rec;
  const double* cs;              // per cell (bilinear_embed)
  const double* bc;              // finest level: the condition's value per cell
  const signed char* fcode[2];   // finest level, per face: 0 plain, +-1 the j of the embedded gradient
  const double* fs[2];           // finest level face fractions
  int dirichlet;                 // the a[embed] condition is Dirichlet (else Neumann)
};

__device__ __forceinline__ double emb_quadratic(double x, double a1, double a2, double a3) {
  return ((a1) * ((x)-1.) + (a3) * ((x) + 1.)) * (x) / 2. - (a2) * ((x)-1.) * ((x) + 1.);
}

// embed_flux of field s at a cell (record R; s(o) is the value at flat
// offset o from the cell): returns the coefficient of the cell's own value
// and sets *val; bc, dirichlet: the condition there
template <class S>
__device__ __forceinline__ double embed_flux_dev(const EmbRec& R, const S& s, double bc, int dirichlet,
                                                 double mua, double D, double* val) {
  *val = 0.;
  double grad = bc;
  if (!grad && !dirichlet) return 0.;
  double coef = 0.;
  if (dirichlet) {
    if (R.mode == 0) {
      coef = -1. / (R.d[0] * D);
      grad = bc / (R.d[0] * D);
    } else {
      const double v0 = emb_quadratic(R.y1[0], s(R.off[0][0]), s(R.off[0][1]), s(R.off[0][2]));
      if (R.mode == 2) {
        const double v1 = emb_quadratic(R.y1[1], s(R.off[1][0]), s(R.off[1][1]), s(R.off[1][2]));
        grad = (R.d[1] * (bc - v0) / R.d[0] - R.d[0] * (bc - v1) / R.d[1]) / ((R.d[1] - R.d[0]) * D);
      } else {
        grad = (bc - v0) / (R.d[0] * D);
      }
    }
  }
  *val = -mua / (R.fa + kSEPS) * grad * R.area / D;
  return -mua / (R.fa + kSEPS) * coef * R.area / D;
}

// mua = sum over dimensions of alpha.x[] + alpha.x[1] (embed_flux's order)
__device__ __forceinline__ double emb_mua(const Coef<double>& k, const Geo<2>& g, idx_t idx, const int c[3]) {
  double mua = 0.;
#pragma unroll
  for (int d = 0; d < 2; d++) {
    double ah, al;
    if (k.lo[0]) {
      ah = alpha_hi<2, double>(k.lo[d], k.wall[d], g, idx, c, d);
      al = k.lo[d][idx];
    } else {
      ah = al = k.alpha;
    }
    mua += al + ah;
  }
  return mua;
}

// one relax pass with the EMBED branch, src -> dst: colour 0 / 1 relaxes
// the cells of that parity from src and copies the others (a cut cell's
// stencil reaches cells of its own parity, so a pass reads only the state
// before it); colour -1: weighted Jacobi, dst = (src + 2 n/d)/3.  A cell
// whose d is zero gets 0, and its right-hand side res is zeroed
// (poisson.h:315-318).  da carries the homogeneous condition: value 0.
__global__ __launch_bounds__(kBlock) void k_relax_embed(const double* __restrict__ src, double* __restrict__ dst,
                                                        double* __restrict__ res, Coef<double> k, Geo<2> g, BCH bc,
                                                        double D2, double D, int colour, EmbLevel E) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= g.ncell()) return;
  int c[3];
  g.decode(idx, c);
  const double self = src[idx];
  if (colour >= 0 && ((c[0] + c[1]) & 1) != colour) {
    dst[idx] = self;
    return;
  }
  const double lam = k.lam ? k.lam[idx] : k.lam_c;
  double nn = -D2 * res[idx], dd = -lam * D2;
#pragma unroll
  for (int d = 0; d < 2; d++) {
    double ah, al;
    if (k.lo[0]) {
      ah = alpha_hi<2, double>(k.lo[d], k.wall[d], g, idx, c, d);
      al = k.lo[d][idx];
    } else {
      ah = al = k.alpha;
    }
    const double ap = nbr_h<2, double>(src, g, bc, idx, c, d, +1, self);
    const double am = nbr_h<2, double>(src, g, bc, idx, c, d, -1, self);
    nn += ah * ap + al * am;
    dd += ah + al;
  }
  const int r = E.cell[idx];
  if (r >= 0) {
    double cv;
    const double e = embed_flux_dev(E.rec[r], [&](int o) { return src[idx + o]; }, 0., E.dirichlet,
                                    emb_mua(k, g, idx, c), D, &cv);
    nn -= cv * D2;
    dd += e * D2;
  }
  double v;
  if (!dd) {
    v = 0.;
    res[idx] = 0.;
  } else {
    v = nn / dd;
  }
  dst[idx] = colour >= 0 ? v : (self + 2. * v) / 3.;
}

// relax() without -DJACOBI (poisson.h:296, 303-320) with the EMBED branch:
// each cell is overwritten as the i-major loop visits it, so a cell reads
// the new values of the cells before it in that order and the old values of
// the others, of itself and of the ghosts.  The stencil of a cut cell
// (embed_flux's interpolation) reaches two cells along each axis, so the
// cells of the hyperplane w = 3i + j depend only on hyperplanes below w
// (a cell before (i, j) has w' <= w - 1) and the sweep runs hyperplane by
// hyperplane in one workgroup, dold -> dnew, equal to the serial loop bit
// for bit: a flat offset o < 0 is a cell before this one (dnew), o >= 0 one
// after it or itself (dold).  A cell whose d is zero gets 0 and its res 0.
__global__ __launch_bounds__(1024) void k_gs_embed_block(const double* __restrict__ dold, double* dnew,
                                                         double* __restrict__ res, Coef<double> k, Geo<2> g,
                                                         BCH bc, double D2, double D, EmbLevel E) {
  const int n = g.n, smax = 4 * (n - 1);
  for (int s = 0; s <= smax; s++) {
    // the cells (i, s - 3i) with 0 <= s - 3i < n
    const int ilo = s - (n - 1) <= 0 ? 0 : (s - (n - 1) + 2) / 3, ihi = min(n - 1, s / 3);
    for (int i = ilo + (int)threadIdx.x; i <= ihi; i += (int)blockDim.x) {
      const int c[3] = {i, s - 3 * i, 0};
      const idx_t idx = g.index(c);
      auto at = [&](int o) { return o < 0 ? dnew[(int)idx + o] : dold[(int)idx + o]; };
      const double self = dold[idx];
      const double lam = k.lam ? k.lam[idx] : k.lam_c;
      double nn = -D2 * res[idx], dd = -lam * D2;
#pragma unroll
      for (int d = 0; d < 2; d++) {
        double ah, al;
        if (k.lo[0]) {
          ah = alpha_hi<2, double>(k.lo[d], k.wall[d], g, idx, c, d);
          al = k.lo[d][idx];
        } else {
          ah = al = k.alpha;
        }
        const double ap = nbr_h<2, double>(dold, g, bc, idx, c, d, +1, self);
        const double am = c[d] > 0 ? dnew[idx - g.stride(d)] : nbr_h<2, double>(dold, g, bc, idx, c, d, -1, self);
        nn += ah * ap + al * am;
        dd += ah + al;
      }
      const int r = E.cell[idx];
      if (r >= 0) {
        double cv;
        const double e = embed_flux_dev(E.rec[r], at, 0., E.dirichlet, emb_mua(k, g, idx, c), D, &cv);
        nn -= cv * D2;
        dd += e * D2;
      }
      if (!dd) {
        dnew[idx] = 0.;
        res[idx] = 0.;
      } else {
        dnew[idx] = nn / dd;
      }
    }
    __syncthreads();
  }
}

// face_gradient_x (a, h) at cell c along d, h = 0 low face, 1 high face:
// the embedded form on the faces the host coded (third, partial face, face
// condition met), else (a[h] - a[h-1])/Delta with the box ghosts of a
__device__ __forceinline__ double emb_face_gradient(const double* __restrict__ a, const EmbLevel& E,
                                                    const Geo<2>& g, const BCA& bc, idx_t idx, const int c[3],
                                                    int d, int h, double D) {
  const int n = g.n;
  // the face: index along d is c[d] + h; faces of direction d are (n+1) x n
  const int fi = c[d] + h, ft = c[1 - d];
  const size_t fidx = d == 0 ? (size_t)fi * n + ft : (size_t)ft * (n + 1) + fi;
  const int code = E.fcode[d] ? E.fcode[d][fidx] : 0;
  if (code) {
    const double fsf = E.fs[d][fidx];
    const idx_t sd = g.stride(d), st = g.stride(1 - d);
    const idx_t i0 = h ? idx + sd : idx;  // cell on the high side of the face
    const idx_t im = i0 - sd;             // and on its low side
    const idx_t j0 = code > 0 ? i0 + st : i0 - st, jm = code > 0 ? im + st : im - st;
    return ((1. + fsf) * (a[i0] - a[im]) + (1. - fsf) * (a[j0] - a[jm])) / (2. * D);
  }
  // plain: the neighbour across the face, with the inhomogeneous ghost
  bool gh;
  const int dir = h ? +1 : -1;
  const idx_t q = nbr_index<2>(g, bc.periodic, idx, c, d, dir, gh);
  double v = a[q];
  if (gh) {
    const int side = h ? 2 * d : 2 * d + 1;
    const double cc = bc.c[side] ? bc.c[side][g.wall(d, c)] : 0.;
    v = cc + double(bc.sigma[side]) * v;
  }
  const double a0 = a[idx];
  return h ? (v - a0) / D : (a0 - v) / D;
}

// the residual with the EMBED branch on the finest level (poisson.h:373-389
// with embed.h's face gradients and embedded flux), FP64, with the max norm
__global__ __launch_bounds__(kBlock) void k_residual_embed(const double* __restrict__ a,
                                                           const double* __restrict__ b, double* __restrict__ res,
                                                           Coef<double> k, Geo<2> g, BCA bc, double D,
                                                           EmbLevel E, Stat* st) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  double m = 0.;
  if (idx < g.ncell()) {
    int c[3];
    g.decode(idx, c);
    const double a0 = a[idx];
    const double lam = k.lam ? k.lam[idx] : k.lam_c;
    double r = b[idx] - lam * a0;
#pragma unroll
    for (int d = 0; d < 2; d++) {
      double ah, al;
      if (k.lo[0]) {
        ah = alpha_hi<2, double>(k.lo[d], k.wall[d], g, idx, c, d);
        al = k.lo[d][idx];
      } else {
        ah = al = k.alpha;
      }
      r += (al * emb_face_gradient(a, E, g, bc, idx, c, d, 0, D) - ah * emb_face_gradient(a, E, g, bc, idx, c, d, 1, D)) /
           D;
    }
    const int q = E.cell[idx];
    if (q >= 0) {
      double cv;
      const double e = embed_flux_dev(E.rec[q], [&](int o) { return a[idx + o]; }, E.bc[idx], E.dirichlet,
                                    emb_mua(k, g, idx, c), D, &cv);
      r += cv - e * a0;
    }
    res[idx] = r;
    if (r != r) atomicAdd(&st->nans, 1ull);
    else m = fabs(r);
  }
  block_max_to(st, m);
}

// bilinear_embed: injection when one of the coarse cells bilinear would use
// is empty (cs = 0; coarse ghosts of cs are the symmetric mirror)
__global__ __launch_bounds__(kBlock) void k_prolong_embed(const double* __restrict__ coarse,
                                                          double* __restrict__ fine, Geo<2> gf, BCH bcc,
                                                          const double* __restrict__ cs_coarse) {
  const idx_t idx = logical_block() * kBlock + threadIdx.x;
  if (idx >= gf.ncell()) return;
  int c[3];
  gf.decode(idx, c);
  const int nc = gf.n >> 1;
  const int pi = c[0] >> 1, pj = c[1] >> 1, cx = 2 * (c[0] & 1) - 1, cy = 2 * (c[1] & 1) - 1;
  auto csat = [&](int i, int j) {
    i = i < 0 ? 0 : (i >= nc ? nc - 1 : i);
    j = j < 0 ? 0 : (j >= nc ? nc - 1 : j);
    return cs_coarse[(size_t)i * nc + j];
  };
  if (!csat(pi, pj) || !csat(pi + cx, pj) || !csat(pi, pj + cy) || !csat(pi + cx, pj + cy)) {
    fine[idx] = coarse[(size_t)pi * nc + pj];
    return;
  }
  Geo<2> gc{nc, gf.log2n - 1};
  fine[idx] = bilinear_at<2, double>(coarse, gc, bcc, c);
}

}  // namespace bmg
