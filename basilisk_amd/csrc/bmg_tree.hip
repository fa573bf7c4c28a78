// libbmg adaptive-tree path: the host-side quadtree/octree (2:1-balanced
// refinement and the reference's caches), its compilation into per-level
// slot tables in GPU memory, and the solver loop over the kernels of
// bmg_tree_kernels.h, behind the bmg_tree_* entry points of include/bmg.h.
//
// Reference semantics (paths relative to the reference's src/):
//   Cell, flags, neighbours      grid/tree.h:24-52, 398-399
//   init_grid                    grid/tree.h:1488-1584
//   refine_cell, refine()        grid/tree-common.h:21-67, 308-326
//   increment_neighbors          grid/tree.h:1032-1041, alloc_children 929-972
//   update_cache_f (FBOUNDARY)   grid/tree.h:556-684, foreach_cell.h:66-135
//   foreach_level_or_leaf        grid/tree.h:735-740
//   tree_boundary_level          grid/tree-common.h:587-716
//   box_boundary_level           grid/tree.h:1139-1363
//   halo_face, refine_face_x     grid/tree-common.h:373-409, 464-511
//   relax, residual (TREE)       poisson.h:271-338, 355-372
//   mg_cycle, mg_solve, poisson  poisson.h:32-90, 139-235, 402-441
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <atomic>
#include <chrono>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "../../include/bmg.h"
#include "bmg_internal.h"
#include "bmg_tree_kernels.h"

using namespace bmgt;

namespace {

constexpr int GH = 2;
constexpr int kMaxDepth = 16;
// cell flags; the last four are adapt_wavelet's (grid/tree-common.h:197-198,
// 201, 218: refined, too_fine, too_coarse, just_fine)
enum : uint8_t {
  C_ACTIVE = 1, C_LEAF = 2, C_MARK = 4, C_ALLOC = 8,
  C_REFINED = 16, C_TOO_FINE = 32, C_TOO_COARSE = 64, C_JUST_FINE = 128
};

// host tables whose every entry the loops that build them write:
// default-initialised, so no serial zero fill precedes the host threads
template <typename T>
struct NoFillAlloc : std::allocator<T> {
  template <typename U>
  struct rebind {
    using other = NoFillAlloc<U>;
  };
  NoFillAlloc() = default;
  template <typename U>
  NoFillAlloc(const NoFillAlloc<U>&) noexcept {}
  template <typename U, typename... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
  template <typename U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
};
template <typename T>
using table = std::vector<T, NoFillAlloc<T>>;

// f(begin, end) over [0, n) on up to 16 host threads, at least `grain`
// items each; f reads the topology through find() only, which writes nothing
template <typename F>
void par_for(size_t n, F&& f, size_t grain = 32768) {
  unsigned nt = std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
  nt = (unsigned)std::min<size_t>(nt, n / grain);
  if (nt <= 1) {
    f((size_t)0, n);
    return;
  }
  std::vector<std::thread> th;
  for (unsigned k = 0; k < nt; k++) th.emplace_back([&, k] { f(n * k / nt, n * (k + 1) / nt); });
  for (auto& x : th) x.join();
}

struct Cell {
  uint8_t flags = 0;
  int8_t pid = 0;
  uint16_t nb = 0;
  int32_t slot = -1;
};

struct IJK {
  int i, j, k;
  int& operator[](int d) { return d == 0 ? i : d == 1 ? j : k; }
};

// one level of the tree: a dense directory of 8^dim bricks of cells
struct TLevel {
  int dim = 2, n = 0, len = 0, nbd = 0;
  std::vector<int32_t> dir;
  std::vector<std::unique_ptr<Cell[]>> bricks;
  void init(int l, int d) {
    dim = d;
    n = 1 << l;
    len = n + 2 * GH;
    nbd = (len + 7) >> 3;
    size_t nd = (size_t)nbd * nbd * (dim == 3 ? nbd : 1);
    dir.assign(nd, -1);
  }
  size_t bsize() const { return dim == 3 ? 512 : 64; }
  size_t bindex(int i, int j, int k) const {
    return dim == 3 ? ((size_t)(i >> 3) * nbd + (j >> 3)) * nbd + (k >> 3) : (size_t)(i >> 3) * nbd + (j >> 3);
  }
  size_t cindex(int i, int j, int k) const {
    return dim == 3 ? (size_t)(((i & 7) << 6) | ((j & 7) << 3) | (k & 7)) : (size_t)(((i & 7) << 3) | (j & 7));
  }
  bool inside(int i, int j, int k) const {
    return i >= 0 && i < len && j >= 0 && j < len && (dim == 2 || (k >= 0 && k < len));
  }
  Cell* find(int i, int j, int k) {
    if (!inside(i, j, k)) return nullptr;
    int32_t b = dir[bindex(i, j, k)];
    if (b < 0) return nullptr;
    Cell* c = &bricks[b][cindex(i, j, k)];
    return (c->flags & C_ALLOC) ? c : nullptr;
  }
  Cell& make(int i, int j, int k) {
    size_t bi = bindex(i, j, k);
    if (dir[bi] < 0) {
      dir[bi] = (int32_t)bricks.size();
      bricks.emplace_back(new Cell[bsize()]);
    }
    return bricks[dir[bi]][cindex(i, j, k)];
  }
};

struct Face {
  int level;
  IJK c;
  int flags;
};

// cells refined by adapt_wavelet, each with the slots (in the layout before
// the mesh change) of its 3^dim ring, which refine_bilinear reads
struct RefineLog {
  std::vector<int> level;
  std::vector<IJK> cell;
  std::vector<int> ring;
  bool ok = true;
};

// host-side topology, following the reference's tree
struct Topo {
  int dim = 2, depth = 0;
  std::vector<TLevel> L;
  std::vector<long> nc;  // allocated groups of children per level (Layer.nc)
  bool dirty = true;
  // caches
  std::vector<std::vector<IJK>> active, prolong, bnd;
  std::vector<int> leaf_level;
  std::vector<IJK> leaf_c;
  std::vector<Face> faces;

  Cell& at(int l, int i, int j, int k) {
    Cell* c = L[l].find(i, j, k);
    if (!c) {
      static Cell none;
      none = Cell();
      return none;
    }
    return *c;
  }
  bool alloc(int l, int i, int j, int k) { return L[l].find(i, j, k) != nullptr; }

  void init(int d) {
    dim = d;
    depth = 0;
    L.clear();
    L.resize(1);
    L[0].init(0, dim);
    nc.assign(1, 0);
    int kmax = dim == 3 ? 2 * GH : 0;
    for (int i = 0; i <= 2 * GH; i++)
      for (int j = 0; j <= 2 * GH; j++)
        for (int k = 0; k <= kmax; k++) {
          Cell& c = L[0].make(i, j, k);
          c.flags = C_ALLOC;
          int a = i - GH, b = j - GH, e = dim == 3 ? k - GH : 0;
          // right=0, left=1, top=2, bottom=3, front=4, back=5
          c.pid = (int8_t)(a > 0 ? -1 : a < 0 ? -2 : b > 0 ? -3 : b < 0 ? -4 : e > 0 ? -5 : e < 0 ? -6 : 0);
        }
    Cell& root = L[0].make(GH, GH, dim == 3 ? GH : 0);
    root.flags = C_ALLOC | C_LEAF | C_ACTIVE;
    root.pid = 0;
    dirty = true;
  }

  void add_level() {
    L.emplace_back();
    nc.push_back(0);
    depth++;
    L[depth].init(depth, dim);
  }

  int kext() const { return dim == 3 ? 2 : 1; }

  void alloc_children(int l, IJK c) {
    if (l == depth)
      add_level();
    else if (alloc(l + 1, 2 * c.i - GH, 2 * c.j - GH, dim == 3 ? 2 * c.k - GH : 0))
      return;
    int8_t pid = at(l, c.i, c.j, c.k).pid;
    nc[l + 1]++;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < kext(); e++) {
          Cell& ch = L[l + 1].make(2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0);
          ch.flags = C_ALLOC;
          ch.nb = 0;
          ch.pid = pid;
          ch.slot = -1;
        }
  }

  void increment_neighbors(int l, IJK c) {
    dirty = true;
    if (at(l, c.i, c.j, c.k).nb++ == 0) alloc_children(l, c);
    int e0 = dim == 3 ? -1 : 0, e1 = dim == 3 ? 1 : 0;
    for (int a = -1; a <= 1; a++)
      for (int b = -1; b <= 1; b++)
        for (int e = e0; e <= e1; e++) {
          IJK n = {c.i + a, c.j + b, c.k + e};
          if (at(l, n.i, n.j, n.k).nb++ == 0) alloc_children(l, n);
        }
    at(l, c.i, c.j, c.k).nb--;
  }

  // refine_cell with the 2:1 rule; the coarser neighbours it refines get
  // `flag`; with a log, every refined cell is recorded with its ring's slots
  void refine_cell(int l, IJK c, uint8_t flag = 0, RefineLog* log = nullptr) {
    if (l > 0) {
      int ch[3] = {2 * ((c.i + GH) % 2) - 1, 2 * ((c.j + GH) % 2) - 1, dim == 3 ? 2 * ((c.k + GH) % 2) - 1 : 1};
      int pc[3] = {(c.i + GH) / 2, (c.j + GH) / 2, dim == 3 ? (c.k + GH) / 2 : 0};
      for (int a = 0; a != 2 * ch[0]; a += ch[0])
        for (int b = 0; b != 2 * ch[1]; b += ch[1])
          for (int e = 0; e != (dim == 3 ? 2 * ch[2] : 1); e += (dim == 3 ? ch[2] : 1)) {
            IJK p = {pc[0] + a, pc[1] + b, dim == 3 ? pc[2] + e : 0};
            Cell& pcell = at(l - 1, p.i, p.j, p.k);
            if (pcell.pid >= 0 && (pcell.flags & C_LEAF)) {
              refine_cell(l - 1, p, flag, log);
              at(l - 1, p.i, p.j, p.k).flags |= flag;
            }
          }
    }
    increment_neighbors(l, c);
    if (log) {
      log->level.push_back(l);
      log->cell.push_back(c);
      int e0 = dim == 3 ? -1 : 0, e1 = dim == 3 ? 1 : 0;
      for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
          for (int e = e0; e <= e1; e++) {
            Cell* n = L[l].find(c.i + a, c.j + b, c.k + e);
            int s = n ? n->slot : -1;
            if (s < 0) log->ok = false;
            log->ring.push_back(s);
          }
    }
    Cell& cell = at(l, c.i, c.j, c.k);
    uint8_t cflag = (cell.flags & C_ACTIVE) ? (C_ACTIVE | C_LEAF) : C_LEAF;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < kext(); e++)
          at(l + 1, 2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0).flags |= cflag;
    at(l, c.i, c.j, c.k).flags &= (uint8_t)~C_LEAF;
  }

  // free_children, grid/tree.h:990-1004: the deepest layer goes with its last
  // group of children (update_depth (-1))
  bool free_children(int l, IJK c) {
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < kext(); e++) {
          Cell* ch = L[l + 1].find(2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0);
          if (!ch) return false;
          ch->flags = 0;
          ch->nb = 0;
        }
    if (--nc[l + 1] == 0) {
      if (l + 1 != depth) return false;
      L.pop_back();
      nc.pop_back();
      depth--;
    }
    return true;
  }

  // decrement_neighbors, grid/tree.h:1043-1059
  bool decrement_neighbors(int l, IJK c) {
    dirty = true;
    bool ok = true;
    int e0 = dim == 3 ? -1 : 0, e1 = dim == 3 ? 1 : 0;
    for (int a = -1; a <= 1; a++)
      for (int b = -1; b <= 1; b++)
        for (int e = e0; e <= e1; e++) {
          IJK n = {c.i + a, c.j + b, c.k + e};
          Cell* p = L[l].find(n.i, n.j, n.k);
          if (p && --p->nb == 0) ok = free_children(l, n) && ok;
        }
    Cell& cc = at(l, c.i, c.j, c.k);
    if (cc.nb)
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int e = 0; e < kext(); e++) {
            Cell& ch = at(l + 1, 2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0);
            ch.flags = C_ALLOC;
            ch.pid = cc.pid;
          }
    return ok;
  }

  // coarsen_cell, grid/tree-common.h:98-131 (the fields are handled on the
  // GPU): refused when a child has a refined neighbour
  int coarsen_cell(int l, IJK c) {
    Cell& cc = at(l, c.i, c.j, c.k);
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < kext(); e++) {
          Cell* ch = L[l + 1].find(2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0);
          if (!ch) return -1;
          if (ch->nb || (ch->pid < 0 && ch->pid != cc.pid)) return 0;
        }
    cc.flags |= C_LEAF;
    return decrement_neighbors(l, c) ? 1 : -1;
  }

  // adapt_wavelet's passes over the topology (grid/tree-common.h:196-294);
  // err[i][slot] is the wavelet estimate of estimated scalar i at each
  // child of a refined cell, max[i] its tolerance
  struct Adapt {
    std::vector<const std::vector<double>*> err;
    std::vector<double> max;
    int maxlevel = 0, minlevel = 1;
    long nf = 0, nc = 0;
    RefineLog log;
    bool ok = true;
  };

  // the refinement pass: a pre-order foreach_cell() in which `continue`
  // skips a cell's children
  void adapt_refine(int l, IJK c, Adapt& A) {
    Cell* cp = L[l].find(c.i, c.j, c.k);
    if (!cp || !(cp->flags & C_ACTIVE)) return;
    if (cp->flags & C_LEAF) {
      if (cp->flags & C_TOO_COARSE) {
        cp->flags &= (uint8_t)~C_TOO_COARSE;
        refine_cell(l, c, C_REFINED, &A.log);
        A.nf++;
      }
      return;
    }
    if (cp->flags & C_REFINED) {  // refined earlier in this pass: skip its children
      cp->flags &= (uint8_t)~C_TOO_COARSE;
      return;
    }
    const int lev = l + 1;
    Cell* ch[8];
    int m = 0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < kext(); e++, m++) {
          ch[m] = L[l + 1].find(2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0);
          if (!ch[m] || ch[m]->slot < 0 || ch[m]->slot >= (long)A.err[0]->size()) {
            A.ok = false;
            return;
          }
        }
    // each estimated scalar in turn updates the children's flags
    for (size_t i = 0; i < A.err.size(); i++) {
      const std::vector<double>& err = *A.err[i];
      const double max = A.max[i];
      for (m = 0; m < (1 << dim); m++) {
        uint8_t& fl = ch[m]->flags;
        const double e = err[ch[m]->slot];
        if (e > max && lev < A.maxlevel) {
          fl &= (uint8_t)~C_TOO_FINE;
          fl |= C_TOO_COARSE;
        } else if ((e <= max / 1.5 || lev > A.maxlevel) && !(fl & (C_TOO_COARSE | C_JUST_FINE))) {
          if (lev >= A.minlevel) fl |= C_TOO_FINE;
        } else if (!(fl & C_TOO_COARSE)) {
          fl &= (uint8_t)~C_TOO_FINE;
          fl |= C_JUST_FINE;
        }
      }
    }
    for (m = 0; m < (1 << dim); m++) {
      uint8_t& fl = ch[m]->flags;
      fl &= (uint8_t)~C_JUST_FINE;
      if (!(fl & C_LEAF)) {
        fl &= (uint8_t)~C_TOO_COARSE;
        if (lev >= A.maxlevel) fl |= C_TOO_FINE;
      } else if (!(fl & C_ACTIVE))
        fl &= (uint8_t)~C_TOO_COARSE;
    }
    if (l < depth)
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int e = 0; e < kext(); e++)
            adapt_refine(l + 1, {2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0}, A);
  }

  // one level l of the coarsening passes
  void adapt_coarsen(int m, IJK c, int l, Adapt& A) {
    if (m > depth) return;
    Cell* cp = L[m].find(c.i, c.j, c.k);
    if (!cp) return;
    if (!is_boundary(*cp)) {
      if (m == l) {
        if (!(cp->flags & C_LEAF)) {
          if (cp->flags & C_REFINED)
            cp->flags &= (uint8_t)~(C_REFINED | C_TOO_FINE);
          else if (cp->flags & C_TOO_FINE) {
            int r = coarsen_cell(m, c);
            if (r < 0) A.ok = false;
            if (r > 0) A.nc++;
            cp->flags &= (uint8_t)~C_TOO_FINE;  // do not coarsen the parent
          }
        }
        if (cp->flags & C_TOO_FINE)
          cp->flags &= (uint8_t)~C_TOO_FINE;
        else if (m > 0) {
          Cell* pp = L[m - 1].find((c.i + GH) / 2, (c.j + GH) / 2, dim == 3 ? (c.k + GH) / 2 : 0);
          if (pp && (pp->flags & C_TOO_FINE)) pp->flags &= (uint8_t)~C_TOO_FINE;
        }
        return;
      } else if (cp->flags & C_LEAF)
        return;
    }
    if (m < depth)
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int e = 0; e < kext(); e++)
            adapt_coarsen(m + 1, {2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0}, l, A);
  }

  // the non-leaf cells foreach_cell() visits (pre-order, not descending
  // into leaves)
  void nonleaf(int l, IJK c, std::vector<int>& lev, std::vector<IJK>& cells) {
    Cell* cp = L[l].find(c.i, c.j, c.k);
    if (!cp || (cp->flags & C_LEAF)) return;
    lev.push_back(l);
    cells.push_back(c);
    if (l < depth)
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int e = 0; e < kext(); e++)
            nonleaf(l + 1, {2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0}, lev, cells);
  }

  // one level l of unrefine()'s coarsening passes (grid/tree-common.h:356-369)
  void unrefine_pass(int m, IJK c, int l, long& nc, bool& ok) {
    if (m > depth) return;
    Cell* cp = L[m].find(c.i, c.j, c.k);
    if (!cp || (cp->flags & C_LEAF)) return;
    if (m == l) {
      if (cp->flags & C_TOO_FINE) {
        int r = coarsen_cell(m, c);
        if (r < 0) ok = false;
        if (r > 0) nc++;
        cp->flags &= (uint8_t)~C_TOO_FINE;
      }
      return;
    }
    if (m < depth)
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int e = 0; e < kext(); e++)
            unrefine_pass(m + 1, {2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0}, l, nc, ok);
  }

  bool is_boundary(const Cell& c) const { return c.pid < 0; }
  bool is_leaf(const Cell& c) const { return c.flags & C_LEAF; }
  bool is_refined(const Cell& c) const { return !(c.flags & C_LEAF) && c.nb && c.pid >= 0; }
  bool is_prolongation(const Cell& c) const { return !(c.flags & C_LEAF) && !c.nb && c.pid >= 0; }

  // the cache entries of one stretch of the pre-order traversal; `bnd`
  // holds candidates (a boundary cell may repeat: the merge keeps its first
  // occurrence, as the marks of a single traversal do)
  struct CachePart {
    std::vector<std::vector<IJK>> active, prolong, bnd;
    std::vector<int> leaf_level;
    std::vector<IJK> leaf_c;
    std::vector<Face> faces;
    int root_level = -1;  // a subtree still to traverse from this root
    IJK root{0, 0, 0};
  };
  void size_part(CachePart& P) const {
    P.active.assign(depth + 1, {});
    P.prolong.assign(depth + 1, {});
    P.bnd.assign(depth + 1, {});
  }
  // a cell, or a default (unallocated) cell; reads only
  const Cell& cell_or_none(int l, int i, int j, int k) {
    static const Cell none{};
    const Cell* c = L[l].find(i, j, k);
    return c ? *c : none;
  }

  // the entries of cell c itself (grid/tree.h:556-684's caches); true when
  // its children follow in the traversal.  Reads the topology only.
  bool visit_self(int l, IJK c, CachePart& P) {
    Cell* cp = L[l].find(c.i, c.j, c.k);
    if (!cp) return false;
    const Cell& cell = *cp;
    if (cell.pid == 0 && (cell.flags & C_ACTIVE)) P.active[l].push_back(c);
    // only a cell next to the box has ghost cells in its ring
    const int hi = (1 << l) + GH - 1;
    const bool edge = c.i <= GH || c.i >= hi || c.j <= GH || c.j >= hi || (dim == 3 && (c.k <= GH || c.k >= hi));
    if (edge && !is_boundary(cell)) {
      int e0 = dim == 3 ? -1 : 0, e1 = dim == 3 ? 1 : 0;
      for (int a = -1; a <= 1; a++)
        for (int b = -1; b <= 1; b++)
          for (int e = e0; e <= e1; e++) {
            const Cell* n = L[l].find(c.i + a, c.j + b, c.k + e);
            if (n && is_boundary(*n)) P.bnd[l].push_back({c.i + a, c.j + b, c.k + e});
          }
    }
    if (is_leaf(cell)) {
      if (cell.pid == 0) {
        P.leaf_level.push_back(l);
        P.leaf_c.push_back(c);
        int flags = 0;
        for (int d = 0; d < dim; d++) {
          IJK m = c;
          m[d]--;
          const Cell& mc = cell_or_none(l, m.i, m.j, m.k);
          if (is_boundary(mc) || is_prolongation(mc) || is_leaf(mc)) flags |= 1 << d;
        }
        if (flags) P.faces.push_back({l, c, flags});
        for (int d = 0; d < dim; d++) {
          IJK p = c;
          p[d]++;
          const Cell& pc = cell_or_none(l, p.i, p.j, p.k);
          if (is_boundary(pc) || is_prolongation(pc) || (pc.pid != 0 && is_leaf(pc)))
            P.faces.push_back({l, p, 1 << d});
        }
        if (cell.nb > 0) P.prolong[l].push_back(c);
      }
      return false;
    }
    return l < depth;
  }
  IJK child(IJK c, int a, int b, int e) const {
    return {2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0};
  }
  // pre-order traversal of the subtree of c into P
  void visit_part(int l, IJK c, CachePart& P) {
    if (!visit_self(l, c, P)) return;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < kext(); e++) visit_part(l + 1, child(c, a, b, e), P);
  }
  // pre-order traversal above level `split`: the cells there go to the last
  // part; each cell of level `split` gets a part of its own, traversed later,
  // and a new part follows it
  void visit_top(int l, IJK c, std::vector<CachePart>& parts, int split) {
    if (l == split) {
      if (!L[l].find(c.i, c.j, c.k)) return;
      parts.emplace_back();
      parts.back().root_level = l;
      parts.back().root = c;
      parts.emplace_back();
      size_part(parts.back());
      return;
    }
    if (!visit_self(l, c, parts.back())) return;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < kext(); e++) visit_top(l + 1, child(c, a, b, e), parts, split);
  }

  // the caches in pre-order; the subtrees below level 3 (3D) or 4 (2D) are
  // traversed on up to 16 host threads and concatenated in traversal order,
  // which is the order of a single traversal
  void update() {
    if (!dirty) return;
    const int split = dim == 3 ? 3 : 4;
    const IJK root = {GH, GH, dim == 3 ? GH : 0};
    std::vector<CachePart> parts(1);
    size_part(parts[0]);
    if (dim >= 2 && depth > split)
      visit_top(0, root, parts, split);
    else
      visit_part(0, root, parts[0]);
    std::vector<size_t> subs;
    for (size_t q = 0; q < parts.size(); q++)
      if (parts[q].root_level >= 0) subs.push_back(q);
    if (!subs.empty()) {
      std::atomic<size_t> next{0};
      auto work = [&]() {
        for (size_t q; (q = next.fetch_add(1)) < subs.size();) {
          CachePart& P = parts[subs[q]];
          size_part(P);
          visit_part(P.root_level, P.root, P);
        }
      };
      unsigned nt = std::min<unsigned>(16u, std::max(1u, std::thread::hardware_concurrency()));
      nt = (unsigned)std::min<size_t>(nt, subs.size());
      std::vector<std::thread> th;
      for (unsigned k = 1; k < nt; k++) th.emplace_back(work);
      work();
      for (auto& x : th) x.join();
    }
    // concatenation in part order: offsets first, then the copies on host
    // threads; the boundary candidates keep their first occurrences
    const size_t np = parts.size();
    std::vector<std::vector<size_t>> aoff(depth + 1, std::vector<size_t>(np + 1, 0)), poff = aoff;
    std::vector<size_t> loff(np + 1, 0), foff(np + 1, 0);
    for (size_t q = 0; q < np; q++) {
      for (int l = 0; l <= depth; l++) {
        aoff[l][q + 1] = aoff[l][q] + parts[q].active[l].size();
        poff[l][q + 1] = poff[l][q] + parts[q].prolong[l].size();
      }
      loff[q + 1] = loff[q] + parts[q].leaf_c.size();
      foff[q + 1] = foff[q] + parts[q].faces.size();
    }
    active.assign(depth + 1, {});
    prolong.assign(depth + 1, {});
    bnd.assign(depth + 1, {});
    for (int l = 0; l <= depth; l++) {
      active[l].resize(aoff[l][np]);
      prolong[l].resize(poff[l][np]);
    }
    leaf_level.assign(loff[np], 0);
    leaf_c.resize(loff[np]);
    faces.resize(foff[np]);
    par_for(np, [&](size_t q0, size_t q1) {
      for (size_t q = q0; q < q1; q++) {
        const CachePart& P = parts[q];
        for (int l = 0; l <= depth; l++) {
          std::copy(P.active[l].begin(), P.active[l].end(), active[l].begin() + aoff[l][q]);
          std::copy(P.prolong[l].begin(), P.prolong[l].end(), prolong[l].begin() + poff[l][q]);
        }
        std::copy(P.leaf_level.begin(), P.leaf_level.end(), leaf_level.begin() + loff[q]);
        std::copy(P.leaf_c.begin(), P.leaf_c.end(), leaf_c.begin() + loff[q]);
        std::copy(P.faces.begin(), P.faces.end(), faces.begin() + foff[q]);
      }
    }, 1);
    for (auto& P : parts) {
      for (int l = 0; l <= depth; l++)
        for (auto& g : P.bnd[l]) {
          Cell& n = at(l, g.i, g.j, g.k);
          if (!(n.flags & C_MARK)) {
            bnd[l].push_back(g);
            n.flags |= C_MARK;
          }
        }
      P = CachePart();
    }
    for (int l = 0; l <= depth; l++)
      for (auto& c : bnd[l]) at(l, c.i, c.j, c.k).flags &= (uint8_t)~C_MARK;
    dirty = false;
  }

  bool full() {
    update();
    return (long)leaf_c.size() == (1L << (depth * dim));
  }
};

// the compiled tree in GPU memory
struct DevIndex {
  int depth = 0, dim = 2;
  long nslot = 0, nleaf = 0;
  std::vector<int> off, nA, abase;                 // per level
  std::vector<int> rec_off, n_nonleaf, n_prolong;  // parent records per level
  std::vector<int> box_off, box_n;                 // box rules per level
  std::vector<int> face_off, face_n;               // residual face ops per level
  std::vector<int> fr_off, fr_n;                   // alpha face restriction per level
  std::vector<int> fp_off, fp_n;                   // refine_face_x ops per level
  std::vector<int> fc_off, fc_n;                   // ghost tangential face copies per level
  std::vector<int> leaf_off, leaf_n;               // leaves of each level (slot lists)
  int nhf = 0;                                     // halo_face ops
  long nfaceval = 0;
  int *nbr = nullptr, *rec_self = nullptr, *rec_child0 = nullptr, *rec_ring = nullptr;
  uint8_t* colour = nullptr;
  int *leaf_gs = nullptr, *lvl_leaves = nullptr;
  // per level and colour: the active slots, then the leaf slots (red-black
  // passes launch over one colour only)
  int* col_list = nullptr;
  int* col_nbr = nullptr;  // the neighbour-table rows of col_list's entries, in its order
  std::vector<int> col_off, col_n, lcol_off, lcol_n;  // [2 * level + colour]
  BoxRule* box = nullptr;
  FaceOp* fops = nullptr;
  FaceAvg *hf = nullptr, *fr = nullptr;
  FaceProlong* fp = nullptr;
  FaceCopy* fc = nullptr;
  int *fv_gs = nullptr, *fv_d = nullptr;
  double* fc_tmp = nullptr;
  int maxfc = 0;
  TLevelInfo* lvinfo = nullptr;  // per-level extents, for the batched kernel
  std::vector<int> bc_off;  // per level offset into each side's face table
};

void dfree_(void* p) {
  if (p) (void)hipFree(p);
}

template <typename T, typename A>
int upload(T** dst, const std::vector<T, A>& v) {
  size_t bytes = std::max<size_t>(v.size() * sizeof(T), 8);
  hipError_t e = hipMalloc((void**)dst, bytes);
  if (e != hipSuccess) return bmg::fail(BMG_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  if (!v.empty()) HIPCHECK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return BMG_OK;
}

unsigned nb_(long n) { return (unsigned)((n + kB - 1) / kB); }

}  // namespace

// user scalars of a tree (BMG_F_S .. BMG_F_S + kUserScalars - 1)
constexpr int kUserScalars = 4;
// largest operation the batched small-operation kernel takes by default
constexpr long kSmallOp = 1024;

struct bmg_tree {
  int dim = 2;
  double L0 = 1., X0[3] = {0., 0., 0.};
  Topo topo;
  // solver settings
  int precision = BMG_FP64, smoother = BMG_RELAX_RBGS;
  double TOLERANCE = 1e-3;
  int NITERMAX = 100, NITERMIN = 1;
  int bc_kind[6] = {0, 0, 0, 0, 0, 0};
  std::vector<double> bc_host[6];  // face values of every level, concatenated
  bool alpha_field = false, lambda_field = false;
  double alpha_c = 1., lambda_c = 0.;
  // device
  bool compiled = false;
  DevIndex ix;
  double *a = nullptr, *b = nullptr, *res64 = nullptr, *g[3] = {nullptr, nullptr, nullptr};
  // user scalars BMG_F_S + k, which adapt_wavelet may estimate and which
  // follow the mesh; us[0] always exists, the others once set
  double* us[kUserScalars] = {};
  double* uf[3] = {nullptr, nullptr, nullptr};  // project()'s face velocity per slot (allocated on use)
  // the adjoint's inverse tables and scratch (built on first use per mesh)
  struct Adj {
    bool built = false;
    int *fop_gs = nullptr, *fop_gm = nullptr;      // per slot and dimension: face op with that gs / gm
    int *box_off = nullptr, *box_inv = nullptr;    // per source slot: its box rules, rule order
    int* pro_off = nullptr;                        // per ring slot: bilinear terms naming it
    bmgt::ProEnt* pro_inv = nullptr;
    int *nbr_off = nullptr, *nbr_inv = nullptr;    // per slot: (active index, direction) reading it
    uint8_t* slot_leaf = nullptr;
    double *xb = nullptr, *gb[3] = {nullptr, nullptr, nullptr}, *db = nullptr, *rb = nullptr;
    double *lv0 = nullptr, *lv1 = nullptr, *lv2 = nullptr, *lv3 = nullptr, *lv4 = nullptr;  // leaf-order vectors
  } adj;
  bool us_set[kUserScalars] = {};
  double *al64[3] = {nullptr, nullptr, nullptr}, *lam64 = nullptr;
  void *da = nullptr, *res = nullptr, *tmp = nullptr, *alT[3] = {nullptr, nullptr, nullptr}, *lamT = nullptr;
  double* bcv[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  bool coef_ready = false;
  TStat* stat = nullptr;
  double *part = nullptr, *scalar = nullptr;
  hipStream_t own_stream = nullptr, stream = nullptr;
  // field values kept across a recompilation, in leaf order
  std::vector<double> keep_a, keep_b, keep_lam, keep_us[kUserScalars];
  // when set, compile() records each new slot's slot in the previous layout
  std::vector<int>* prev_slots = nullptr;
  // HIP graphs of the cycle (per nrelax, minlevel, maxlevel) and of the
  // residual's kernels: the topology and every kernel argument are fixed
  // until a setting or the mesh changes, when they are dropped
  bool use_graphs = true;
  int ops_block = kOpsBlock;  // threads of the batched small-operation kernel (BMG_OPS_BLOCK)
  bool batching = true;  // small operations run batched in one workgroup (BMG_TREE_BATCH=0: off)
  long small_op = kSmallOp;  // largest batched operation (BMG_TREE_SMALL_OP overrides)
  std::map<std::tuple<int, int, int>, hipGraphExec_t> cycle_graphs;
  hipGraphExec_t residual_graph = nullptr;
  // relax-kernel timing (bmg_tree_profile_*): HIP events around every
  // k_t_relax launch on the tree's stream, with the cells each relaxes;
  // the cycle runs without graphs meanwhile
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<long, std::pair<hipEvent_t, hipEvent_t>>> ev_used;
  double prof_ms = 0.;
  long prof_n = 0, prof_cells = 0;
};

namespace {

long seg_cells(const Segs& S) {
  long n = 0;
  for (int i = 0; i < S.n; i++) n += S.count[i];
  return n;
}

// events around one relax launch of `cells` cells (no-op unless profiling)
struct TProf {
  bmg_tree* t;
  hipEvent_t e1 = nullptr;
  TProf(bmg_tree* t_, long cells) : t(t_) {
    if (!t->prof) return;
    hipEvent_t e0;
    if (t->ev_pool.size() >= 2) {
      e0 = t->ev_pool.back(); t->ev_pool.pop_back();
      e1 = t->ev_pool.back(); t->ev_pool.pop_back();
    } else {
      (void)hipEventCreate(&e0);
      (void)hipEventCreate(&e1);
    }
    (void)hipEventRecord(e0, t->stream);
    t->ev_used.push_back({cells, {e0, e1}});
  }
  ~TProf() {
    if (t->prof) (void)hipEventRecord(e1, t->stream);
  }
};

void tprof_collect(bmg_tree* t) {
  if (t->ev_used.empty()) return;
  (void)hipStreamSynchronize(t->stream);
  for (auto& u : t->ev_used) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, u.second.first, u.second.second);
    t->prof_ms += ms;
    t->prof_n += 1;
    t->prof_cells += u.first;
    t->ev_pool.push_back(u.second.first);
    t->ev_pool.push_back(u.second.second);
  }
  t->ev_used.clear();
}

}  // namespace

namespace {

size_t tsize(const bmg_tree* t) { return t->precision == BMG_MIXED ? sizeof(float) : sizeof(double); }

void drop_graphs(bmg_tree* t) {
  for (auto& kv : t->cycle_graphs)
    if (kv.second) (void)hipGraphExecDestroy(kv.second);
  t->cycle_graphs.clear();
  if (t->residual_graph) (void)hipGraphExecDestroy(t->residual_graph);
  t->residual_graph = nullptr;
}

void free_device(bmg_tree* t) {
  drop_graphs(t);
  DevIndex& x = t->ix;
  for (void* p : {(void*)x.nbr, (void*)x.rec_self, (void*)x.rec_child0, (void*)x.rec_ring, (void*)x.colour,
                  (void*)x.leaf_gs, (void*)x.lvl_leaves, (void*)x.col_list, (void*)x.col_nbr, (void*)x.box, (void*)x.fops, (void*)x.hf, (void*)x.fr,
                  (void*)x.fp, (void*)x.fc, (void*)x.fv_gs, (void*)x.fv_d, (void*)x.fc_tmp, (void*)x.lvinfo})
    dfree_(p);
  x = DevIndex();
  bool mixed = t->precision == BMG_MIXED;
  dfree_(t->a); dfree_(t->b); dfree_(t->res64);
  for (auto& p : t->us) {
    dfree_(p);
    p = nullptr;
  }
  for (int d = 0; d < 3; d++) {
    dfree_(t->g[d]);
    dfree_(t->al64[d]);
    if (mixed) dfree_(t->alT[d]);
    t->g[d] = t->al64[d] = nullptr;
    t->alT[d] = nullptr;
  }
  dfree_(t->lam64);
  for (int d = 0; d < 3; d++) {
    dfree_(t->uf[d]);
    t->uf[d] = nullptr;
  }
  {
    auto& A = t->adj;
    for (void* p : {(void*)A.fop_gs, (void*)A.fop_gm, (void*)A.box_off, (void*)A.box_inv, (void*)A.pro_off,
                    (void*)A.pro_inv, (void*)A.nbr_off, (void*)A.nbr_inv, (void*)A.slot_leaf, (void*)A.xb,
                    (void*)A.gb[0], (void*)A.gb[1], (void*)A.gb[2], (void*)A.db, (void*)A.rb, (void*)A.lv0,
                    (void*)A.lv1, (void*)A.lv2, (void*)A.lv3, (void*)A.lv4})
      dfree_(p);
    A = bmg_tree::Adj();
  }
  if (mixed) { dfree_(t->lamT); dfree_(t->res); }
  dfree_(t->da); dfree_(t->tmp);
  for (int s = 0; s < 6; s++) { dfree_(t->bcv[s]); t->bcv[s] = nullptr; }
  dfree_(t->stat); dfree_(t->part); dfree_(t->scalar);
  t->a = t->b = t->res64 = t->lam64 = nullptr;
  t->da = t->res = t->tmp = t->lamT = nullptr;
  t->stat = nullptr; t->part = t->scalar = nullptr;
  t->compiled = false;
  t->coef_ready = false;
  // the frees above are best effort; do not leave their status behind for
  // the next launch check
  (void)hipGetLastError();
}

int dalloc_(void** p, size_t bytes) {
  if (!bytes) bytes = 8;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return bmg::fail(BMG_ERR_NOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  return hipMemset(*p, 0, bytes) == hipSuccess ? BMG_OK : bmg::fail(BMG_ERR_HIP, "hipMemset failed");
}

// tangential index of the face of interior cell c on side s at level l
int bc_ti(int dim, int l, int side, const int c[3]) {
  int d = side / 2, n = 1 << l, ti = 0;
  for (int e = 0; e < dim; e++)
    if (e != d) ti = ti * n + (c[e] - GH);
  return ti;
}

int bid(const Cell& c) { return -c.pid - 1; }

// the tangential child offsets of face averages per component
const int kFtan[3][4][3] = {
    {{0, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}},
    {{0, 0, 0}, {0, 0, 1}, {1, 0, 0}, {1, 0, 1}},
    {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0}},
};
const int kFtan2[2][2][3] = {{{0, 0, 0}, {0, 1, 0}}, {{0, 0, 0}, {1, 0, 0}}};

int ftan(int dim, int d, int m, int e) { return dim == 3 ? kFtan[d][m][e] : kFtan2[d][m][e]; }

// host wall time of the phases of a rebuild, printed to stderr when
// BMG_TREE_TIMING is set (diagnostics only; GPU phases are synchronised
// first when it is)
struct PhaseClock {
  bool on = getenv("BMG_TREE_TIMING") != nullptr;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  void mark(const char* what, hipStream_t s = nullptr) {
    if (!on) return;
    if (s) (void)hipStreamSynchronize(s);
    const auto t1 = std::chrono::steady_clock::now();
    fprintf(stderr, "[bmg tree] %s: %.1f ms\n", what, std::chrono::duration<double, std::milli>(t1 - t0).count());
    t0 = t1;
  }
};

// compile the topology into slot tables and allocate the fields
int compile(bmg_tree* t) {
  PhaseClock clk;
  Topo& T = t->topo;
  T.update();
  clk.mark("compile: caches");
  const int dim = t->dim, D = T.depth;
  if (D + 1 > kSeg) return bmg::fail(BMG_ERR_UNSUPPORTED, "tree depth %d exceeds %d levels", D, kSeg - 1);
  // keep leaf values across recompilation
  free_device(t);
  DevIndex& x = t->ix;
  x.depth = D;
  x.dim = dim;
  x.off.assign(D + 2, 0);
  x.nA.assign(D + 1, 0);
  x.abase.assign(D + 1, 0);
  // slot assignment: active, halo children, ghost blocks
  std::vector<std::vector<IJK>> ghosts(D + 1);
  long next = 0;
  if (t->prev_slots) t->prev_slots->clear();
  auto take = [&](Cell* c) {
    if (t->prev_slots) t->prev_slots->push_back(c->slot);
    c->slot = (int)next++;
  };
  // n cells taking the next n slots in order (distinct cells, so the host
  // threads write distinct slots); false when cell_of(q) finds no cell
  auto take_range = [&](size_t n, auto&& cell_of) {
    const long base = next;
    if (t->prev_slots) t->prev_slots->resize((size_t)base + n);
    std::atomic<bool> found{true};
    par_for(n, [&](size_t q0, size_t q1) {
      for (size_t q = q0; q < q1; q++) {
        Cell* c = cell_of(q);
        if (!c) {
          found.store(false, std::memory_order_relaxed);
          continue;
        }
        if (t->prev_slots) (*t->prev_slots)[(size_t)base + q] = c->slot;
        c->slot = (int)(base + (long)q);
      }
    });
    next += (long)n;
    return found.load();
  };
  const int nch = 1 << dim;  // children per cell, in foreach_child order
  for (int l = 0; l <= D; l++) {
    TLevel& Lv = T.L[l];
    x.off[l] = (int)next;
    x.abase[l] = l ? x.abase[l - 1] + x.nA[l - 1] : 0;
    if (!take_range(T.active[l].size(), [&](size_t q) {
          const IJK& c = T.active[l][q];
          return Lv.find(c.i, c.j, c.k);
        }))
      return bmg::fail(BMG_ERR_ARG, "tree: missing active cell");
    x.nA[l] = (int)T.active[l].size();
    if (l > 0) {
      const std::vector<IJK>& pl = T.prolong[l - 1];
      if (!take_range(pl.size() * nch, [&](size_t q) {
            const IJK& p = pl[q / nch];
            const int m = (int)(q % nch);
            const int a = dim == 3 ? m >> 2 : m >> 1, b = dim == 3 ? (m >> 1) & 1 : m & 1, e = dim == 3 ? m & 1 : 0;
            return Lv.find(2 * p.i - GH + a, 2 * p.j - GH + b, dim == 3 ? 2 * p.k - GH + e : 0);
          }))
        return bmg::fail(BMG_ERR_ARG, "tree: missing halo child");
      for (auto& gp : ghosts[l - 1]) {
        Cell* gc = T.L[l - 1].find(gp.i, gp.j, gp.k);
        if (gc->nb == 0) continue;
        for (int a = 0; a < 2; a++)
          for (int b = 0; b < 2; b++)
            for (int e = 0; e < T.kext(); e++) {
              IJK c = {2 * gp.i - GH + a, 2 * gp.j - GH + b, dim == 3 ? 2 * gp.k - GH + e : 0};
              Cell* ch = Lv.find(c.i, c.j, c.k);
              if (!ch) return bmg::fail(BMG_ERR_ARG, "tree: missing ghost child");
              take(ch);
              ghosts[l].push_back(c);
            }
      }
    } else {
      int kmax = dim == 3 ? 2 * GH : 0;
      for (int i = 0; i <= 2 * GH; i++)
        for (int j = 0; j <= 2 * GH; j++)
          for (int k = 0; k <= kmax; k++) {
            Cell* c = Lv.find(i, j, k);
            if (c->pid < 0) {
              take(c);
              ghosts[0].push_back({i, j, k});
            }
          }
    }
    // every allocated cell has a slot
    std::atomic<long> count{0};
    std::atomic<bool> slotted{true};
    const long lo = x.off[l], hi = next;
    par_for(Lv.bricks.size(), [&](size_t b0, size_t b1) {
      long cnt = 0;
      for (size_t b = b0; b < b1; b++)
        for (size_t q = 0; q < Lv.bsize(); q++)
          if (Lv.bricks[b][q].flags & C_ALLOC) {
            cnt++;
            if (Lv.bricks[b][q].slot < lo || Lv.bricks[b][q].slot >= hi) slotted.store(false, std::memory_order_relaxed);
          }
      count += cnt;
    }, 64);
    if (!slotted) return bmg::fail(BMG_ERR_ARG, "tree: unslotted cell at level %d", l);
    if (count != next - x.off[l]) return bmg::fail(BMG_ERR_ARG, "tree: slot count mismatch at level %d", l);
  }
  x.off[D + 1] = (int)next;
  x.nslot = next;
  const long NS = next;
  auto S = [&](int l, int i, int j, int k) -> int {
    Cell* c = T.L[l].find(i, j, k);
    return c ? c->slot : -1;
  };
  // any slot out of range in a table is an error (tables are filled by
  // several host threads)
  std::atomic<bool> ok{true};
  auto chk = [&](int v) {
    if (v < 0 || v >= NS) ok.store(false, std::memory_order_relaxed);
    return v;
  };
  clk.mark("compile: slots");
  // active neighbours and colours
  long NA = x.abase[D] + x.nA[D];
  table<int> nbr((size_t)NA * 2 * dim);
  table<uint8_t> colour(NA);
  for (int l = 0; l <= D; l++)
    par_for((size_t)x.nA[l], [&](size_t q0, size_t q1) {
      for (size_t q = q0; q < q1; q++) {
        IJK c = T.active[l][q];
        size_t ai = (size_t)x.abase[l] + q;
        for (int d = 0; d < dim; d++) {
          int o[3] = {0, 0, 0};
          o[d] = 1;
          nbr[ai * 2 * dim + 2 * d] = chk(S(l, c.i - o[0], c.j - o[1], c.k - o[2]));
          nbr[ai * 2 * dim + 2 * d + 1] = chk(S(l, c.i + o[0], c.j + o[1], c.k + o[2]));
        }
        colour[ai] = (uint8_t)((c.i + c.j + c.k) & 1);
      }
    });
  if (!ok) return bmg::fail(BMG_ERR_ARG, "tree: active cell without neighbours");
  // the non-leaf active cells of each level (indices into active[l])
  clk.mark("compile: tables: neighbours");
  std::vector<std::vector<int>> nonleaf(D + 1);
  for (int l = 0; l <= D; l++) {
    std::vector<uint8_t> isnl((size_t)x.nA[l]);
    par_for((size_t)x.nA[l], [&](size_t q0, size_t q1) {
      for (size_t q = q0; q < q1; q++) {
        const IJK& c = T.active[l][q];
        isnl[q] = !T.is_leaf(*T.L[l].find(c.i, c.j, c.k));
      }
    });
    for (int q = 0; q < x.nA[l]; q++)
      if (isnl[q]) nonleaf[l].push_back(q);
  }
  // parent records: non-leaf active cells, then prolongation leaves
  const int NR = dim == 3 ? 27 : 9;
  x.rec_off.assign(D + 1, 0);
  x.n_nonleaf.assign(D + 1, 0);
  x.n_prolong.assign(D + 1, 0);
  size_t nrec = 0;
  for (int l = 0; l <= D; l++) {
    x.rec_off[l] = (int)nrec;
    x.n_nonleaf[l] = (int)nonleaf[l].size();
    x.n_prolong[l] = (int)T.prolong[l].size();
    nrec += (size_t)x.n_nonleaf[l] + x.n_prolong[l];
  }
  table<int> rself(nrec), rchild(nrec), rring(nrec * NR);
  auto add_rec = [&](size_t r, int l, IJK c) {
    rself[r] = chk(S(l, c.i, c.j, c.k));
    int c0 = chk(S(l + 1, 2 * c.i - GH, 2 * c.j - GH, dim == 3 ? 2 * c.k - GH : 0));
    rchild[r] = c0;
    // the kernels address the children as child0 + (foreach_child index)
    int m = 0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < T.kext(); e++, m++)
          if (S(l + 1, 2 * c.i - GH + a, 2 * c.j - GH + b, dim == 3 ? 2 * c.k - GH + e : 0) != c0 + m)
            ok.store(false, std::memory_order_relaxed);
    int e0 = dim == 3 ? -1 : 0, e1 = dim == 3 ? 1 : 0;
    size_t w = r * NR;
    for (int a = -1; a <= 1; a++)
      for (int b = -1; b <= 1; b++)
        for (int e = e0; e <= e1; e++) rring[w++] = chk(S(l, c.i + a, c.j + b, c.k + e));
  };
  for (int l = 0; l <= D; l++) {
    const size_t r0 = x.rec_off[l], np = x.n_nonleaf[l];
    par_for(np, [&](size_t k0, size_t k1) {
      for (size_t k = k0; k < k1; k++) add_rec(r0 + k, l, T.active[l][nonleaf[l][k]]);
    });
    par_for(T.prolong[l].size(), [&](size_t k0, size_t k1) {
      for (size_t k = k0; k < k1; k++) add_rec(r0 + np + k, l, T.prolong[l][k]);
    });
  }
  if (!ok) return bmg::fail(BMG_ERR_ARG, "tree: parent record with an unallocated ring");
  // leaves: global cache order and per level
  x.nleaf = (long)T.leaf_c.size();
  clk.mark("compile: tables: parent records");
  table<int> leaf_gs(x.nleaf);
  std::vector<std::vector<int>> lv(D + 1);
  par_for((size_t)x.nleaf, [&](size_t n0, size_t n1) {
    for (size_t n = n0; n < n1; n++) {
      IJK c = T.leaf_c[n];
      leaf_gs[n] = chk(S(T.leaf_level[n], c.i, c.j, c.k));
    }
  });
  for (long n = 0; n < x.nleaf; n++) lv[T.leaf_level[n]].push_back(leaf_gs[n]);
  std::vector<int> lvl_leaves;
  x.leaf_off.assign(D + 1, 0);
  x.leaf_n.assign(D + 1, 0);
  for (int l = 0; l <= D; l++) {
    x.leaf_off[l] = (int)lvl_leaves.size();
    x.leaf_n[l] = (int)lv[l].size();
    lvl_leaves.insert(lvl_leaves.end(), lv[l].begin(), lv[l].end());
  }
  clk.mark("compile: tables: leaves");
  // box rules (boundary cache order)
  x.bc_off.assign(D + 2, 0);
  for (int l = 0; l <= D; l++) {
    long m = 1;
    for (int e = 1; e < dim; e++) m *= (1L << l);
    x.bc_off[l + 1] = x.bc_off[l] + (int)m;
  }
  std::vector<BoxRule> box;
  x.box_off.assign(D + 1, 0);
  x.box_n.assign(D + 1, 0);
  for (int l = 0; l <= D; l++) {
    x.box_off[l] = (int)box.size();
    for (auto& g : T.bnd[l]) {
      BoxRule r;
      memset(&r, 0, sizeof r);
      r.ghost = chk(S(l, g.i, g.j, g.k));
      r.src = -1;
      const int gc[3] = {g.i, g.j, g.k};
      Cell& gcell = T.at(l, g.i, g.j, g.k);
      bool done = false;
      for (int d = 0; d < dim && !done; d++)
        for (int s = -1; s <= 1 && !done; s += 2) {
          int c[3] = {gc[0], gc[1], gc[2]};
          c[d] += s;
          Cell* p = T.L[l].find(c[0], c[1], c[2]);
          if (p && p->pid >= 0) {
            r.nterm = 1;
            r.src = p->slot;
            r.side[0] = bid(gcell);
            r.ti[0] = x.bc_off[l] + bc_ti(dim, l, r.side[0], c);
            done = true;
          }
        }
      for (int pl = 0; pl < (dim == 3 ? 3 : 1) && !done; pl++) {
        int da = pl, db = (pl + 1) % dim;
        for (int i = -1; i <= 1 && !done; i += 2)
          for (int j = -1; j <= 1 && !done; j += 2) {
            int nn[3] = {gc[0], gc[1], gc[2]}, n1[3] = {gc[0], gc[1], gc[2]}, n2[3] = {gc[0], gc[1], gc[2]};
            nn[da] += i; nn[db] += j; n1[da] += i; n2[db] += j;
            Cell* pn = T.L[l].find(nn[0], nn[1], nn[2]);
            Cell* p1 = T.L[l].find(n1[0], n1[1], n1[2]);
            Cell* p2 = T.L[l].find(n2[0], n2[1], n2[2]);
            if (pn && pn->pid >= 0 && p1 && p1->pid < 0 && p2 && p2->pid < 0) {
              r.nterm = 2;
              r.src = pn->slot;
              r.side[0] = bid(*p1);
              r.side[1] = bid(*p2);
              r.ti[0] = x.bc_off[l] + bc_ti(dim, l, r.side[0], nn);
              r.ti[1] = x.bc_off[l] + bc_ti(dim, l, r.side[1], nn);
              done = true;
            }
          }
      }
      if (dim == 3)
        for (int i = -1; i <= 1 && !done; i += 2)
          for (int j = -1; j <= 1 && !done; j += 2)
            for (int k = -1; k <= 1 && !done; k += 2) {
              Cell* p0 = T.L[l].find(gc[0] + i, gc[1] + j, gc[2] + k);
              Cell& p1 = T.at(l, gc[0] + i, gc[1] + j, gc[2]);
              Cell& p2 = T.at(l, gc[0] + i, gc[1], gc[2] + k);
              Cell& p3 = T.at(l, gc[0], gc[1] + j, gc[2] + k);
              if (p0 && p0->pid >= 0 && p1.pid < 0 && p2.pid < 0 && p3.pid < 0) {
                int n0[3] = {gc[0] + i, gc[1] + j, gc[2] + k};
                r.nterm = 3;
                r.src = p0->slot;
                r.side[0] = bid(p1);
                r.side[1] = bid(p2);
                r.side[2] = bid(p3);
                for (int m = 0; m < 3; m++) r.ti[m] = x.bc_off[l] + bc_ti(dim, l, r.side[m], n0);
                done = true;
              }
            }
      if (!done) r.nterm = 0;
      if (r.nterm) chk(r.src);
      box.push_back(r);
    }
    x.box_n[l] = (int)box.size() - x.box_off[l];
  }
  clk.mark("compile: tables: box rules");
  // residual face ops per level, and the face-cache value map (entry order
  // then dimension); the positions come from prefix counts of the flag bits
  const size_t NF = T.faces.size();
  table<size_t> gpos(NF), lpos(NF);
  x.face_off.assign(D + 1, 0);
  x.face_n.assign(D + 1, 0);
  {
    size_t g = 0;
    for (size_t e = 0; e < NF; e++) {
      gpos[e] = g;
      const int nb = __builtin_popcount((unsigned)T.faces[e].flags & ((1u << dim) - 1));
      g += nb;
      x.face_n[T.faces[e].level] += nb;
    }
    x.nfaceval = (long)g;
    for (int l = 1; l <= D; l++) x.face_off[l] = x.face_off[l - 1] + x.face_n[l - 1];
    std::vector<size_t> cur(x.face_off.begin(), x.face_off.end());
    for (size_t e = 0; e < NF; e++) {
      lpos[e] = cur[T.faces[e].level];
      cur[T.faces[e].level] += __builtin_popcount((unsigned)T.faces[e].flags & ((1u << dim) - 1));
    }
  }
  table<int> fv_gs(x.nfaceval), fv_d(x.nfaceval);
  table<FaceOp> fops(x.nfaceval);
  par_for(NF, [&](size_t e0, size_t e1) {
    for (size_t e = e0; e < e1; e++) {
      const Face& f = T.faces[e];
      int l = f.level;
      const int fc[3] = {f.c.i, f.c.j, f.c.k};
      size_t g = gpos[e], lp = lpos[e];
      for (int d = 0; d < dim; d++)
        if (f.flags & (1 << d)) {
          int m[3] = {fc[0], fc[1], fc[2]};
          m[d]--;
          FaceOp op = {chk(S(l, fc[0], fc[1], fc[2])), chk(S(l, m[0], m[1], m[2])), d};
          fops[lp++] = op;
          fv_gs[g] = op.gs;
          fv_d[g++] = d;
        }
    }
  });
  clk.mark("compile: tables: face ops");
  // fine-face averages: halo_face (all levels in one launch) and the face
  // restriction of alpha (per level)
  auto fine_avg = [&](int l, const int c[3], int d, int hi, int target) {
    FaceAvg fa;
    memset(&fa, 0, sizeof fa);
    fa.target = target;
    fa.d = d;
    int base[3] = {2 * c[0] - GH, 2 * c[1] - GH, dim == 3 ? 2 * c[2] - GH : 0};
    base[d] += 2 * hi;
    for (int m = 0; m < (1 << (dim - 1)); m++)
      fa.fine[m] = chk(S(l + 1, base[0] + ftan(dim, d, m, 0), base[1] + ftan(dim, d, m, 1),
                         dim == 3 ? base[2] + ftan(dim, d, m, 2) : 0));
    return fa;
  };
  // halo faces of the prolongation leaves, levels from the depth up: per
  // leaf a bit (2d + side) for each refined face neighbour, then the records
  // in (leaf, d, side) order at prefix-sum positions
  std::vector<FaceAvg> hf;
  for (int l = D - 1; l >= 0; l--) {
    const size_t np = T.prolong[l].size();
    std::vector<uint8_t> mask(np);
    par_for(np, [&](size_t k0, size_t k1) {
      for (size_t k = k0; k < k1; k++) {
        const IJK& p = T.prolong[l][k];
        const int c[3] = {p.i, p.j, p.k};
        uint8_t b = 0;
        for (int d = 0; d < dim; d++) {
          int m[3] = {c[0], c[1], c[2]}, q[3] = {c[0], c[1], c[2]};
          m[d]--;
          q[d]++;
          const Cell* cm = T.L[l].find(m[0], m[1], m[2]);
          const Cell* cq = T.L[l].find(q[0], q[1], q[2]);
          if (cm && T.is_refined(*cm)) b |= (uint8_t)(1u << (2 * d));
          if (cq && T.is_refined(*cq)) b |= (uint8_t)(1u << (2 * d + 1));
        }
        mask[k] = b;
      }
    });
    std::vector<size_t> pos(np + 1);
    pos[0] = hf.size();
    for (size_t k = 0; k < np; k++) pos[k + 1] = pos[k] + (size_t)__builtin_popcount(mask[k]);
    hf.resize(pos[np]);
    par_for(np, [&](size_t k0, size_t k1) {
      for (size_t k = k0; k < k1; k++) {
        const IJK& p = T.prolong[l][k];
        const int c[3] = {p.i, p.j, p.k};
        size_t w = pos[k];
        for (int d = 0; d < dim; d++) {
          int q[3] = {c[0], c[1], c[2]};
          q[d]++;
          if (mask[k] & (1u << (2 * d))) hf[w++] = fine_avg(l, c, d, 0, chk(S(l, c[0], c[1], c[2])));
          if (mask[k] & (1u << (2 * d + 1))) hf[w++] = fine_avg(l, c, d, 1, chk(S(l, q[0], q[1], q[2])));
        }
      }
    });
  }
  x.nhf = (int)hf.size();
  x.fr_off.assign(D + 1, 0);
  x.fr_n.assign(D + 1, 0);
  size_t nfr = 0;
  for (int l = 0; l <= D; l++) {
    x.fr_off[l] = (int)nfr;
    x.fr_n[l] = (int)nonleaf[l].size() * 2 * dim;
    nfr += x.fr_n[l];
  }
  table<FaceAvg> fr(nfr);
  for (int l = 0; l <= D; l++)
    par_for(nonleaf[l].size(), [&](size_t k0, size_t k1) {
      for (size_t k = k0; k < k1; k++) {
        const IJK& p = T.active[l][nonleaf[l][k]];
        const int c[3] = {p.i, p.j, p.k};
        size_t w = x.fr_off[l] + k * 2 * dim;
        for (int d = 0; d < dim; d++) {
          int q[3] = {c[0], c[1], c[2]};
          q[d]++;
          fr[w++] = fine_avg(l, c, d, 0, chk(S(l, c[0], c[1], c[2])));
          fr[w++] = fine_avg(l, c, d, 1, chk(S(l, q[0], q[1], q[2])));
        }
      }
    });
  // refine_face_x of the prolongation leaves
  x.fp_off.assign(D + 1, 0);
  x.fp_n.assign(D + 1, 0);
  size_t nfp = 0;
  for (int l = 0; l < D; l++) {
    x.fp_off[l] = (int)nfp;
    x.fp_n[l] = (int)T.prolong[l].size() * dim;
    nfp += x.fp_n[l];
  }
  table<FaceProlong> fp(nfp);
  for (int l = 0; l < D; l++)
    par_for(T.prolong[l].size(), [&](size_t k0, size_t k1) {
    for (size_t kp = k0; kp < k1; kp++) {
      const IJK& p = T.prolong[l][kp];
      const int c[3] = {p.i, p.j, p.k};
      for (int d = 0; d < dim; d++) {
        int t1 = (d + 1) % dim, t2 = (d + 2) % dim;
        auto sl = [&](int od, int o1, int o2) {
          int q[3] = {c[0], c[1], c[2]};
          q[d] += od;
          q[t1] += o1;
          if (dim == 3) q[t2] += o2;
          return chk(S(l, q[0], q[1], q[2]));
        };
        FaceProlong r;
        memset(&r, 0, sizeof r);
        r.d = d;
        int m[3] = {c[0], c[1], c[2]}, q[3] = {c[0], c[1], c[2]};
        m[d]--;
        q[d]++;
        const Cell* cm = T.L[l].find(m[0], m[1], m[2]);
        const Cell* cq = T.L[l].find(q[0], q[1], q[2]);
        const bool rm = cm && T.is_refined(*cm), rq = cq && T.is_refined(*cq);
        r.flags = (!rm ? 1 : 0) | ((!rq && cq && cq->nb) ? 2 : 0);
        r.self = sl(0, 0, 0);
        r.t1m = sl(0, -1, 0);
        r.t1p = sl(0, 1, 0);
        r.t2m = dim == 3 ? sl(0, 0, -1) : r.self;
        r.t2p = dim == 3 ? sl(0, 0, 1) : r.self;
        r.next = sl(1, 0, 0);
        r.nt1m = sl(1, -1, 0);
        r.nt1p = sl(1, 1, 0);
        r.nt2m = dim == 3 ? sl(1, 0, -1) : r.next;
        r.nt2p = dim == 3 ? sl(1, 0, 1) : r.next;
        int base[3] = {2 * c[0] - GH, 2 * c[1] - GH, dim == 3 ? 2 * c[2] - GH : 0};
        for (int o = 0; o < 3; o++) {
          if ((o == 0 && !(r.flags & 1)) || (o == 2 && !(r.flags & 2))) continue;
          for (int j = 0; j < 2; j++)
            for (int k = 0; k < (dim == 3 ? 2 : 1); k++) {
              int f[3] = {base[0], base[1], base[2]};
              f[d] += o;
              f[t1] += j;
              if (dim == 3) f[t2] += k;
              r.out[o][dim == 3 ? j * 2 + k : j] = chk(S(l + 1, f[0], f[1], f[2]));
            }
        }
        fp[(size_t)x.fp_off[l] + kp * dim + d] = r;
      }
    }
    });
  clk.mark("compile: tables: face averages and refine_face");
  // ghost tangential faces: sequential semantics resolved to pre-pass sources
  std::vector<FaceCopy> fc;
  x.fc_off.assign(D + 1, 0);
  x.fc_n.assign(D + 1, 0);
  for (int l = 0; l <= D; l++) {
    x.fc_off[l] = (int)fc.size();
    // (target, source) pairs of this pass in order of first write, and the
    // position of each target among them
    std::vector<std::vector<std::pair<int, int>>> w(dim);
    std::vector<std::unordered_map<int, size_t>> wpos(dim);
    for (auto& g : T.bnd[l]) {
      const int gc[3] = {g.i, g.j, g.k};
      int gs = chk(S(l, gc[0], gc[1], gc[2]));
      for (int d = 0; d < dim; d++) {
        int m[3] = {gc[0], gc[1], gc[2]};
        m[d]--;
        Cell* pm = T.L[l].find(m[0], m[1], m[2]);
        if (pm && pm->pid >= 0) continue;
        int t1 = (d + 1) % dim, t2 = (d + 2) % dim;
        int found = -1;
        for (int j = -1; j <= 1 && found < 0; j += 2) {
          for (int pass = 0; pass < (dim == 3 ? 2 : 1) && found < 0; pass++) {
            int tt = pass == 0 ? t1 : t2;
            int a[3] = {gc[0], gc[1], gc[2]};
            a[tt] += j;
            int bb[3] = {a[0], a[1], a[2]};
            bb[d]--;
            Cell* pa = T.L[l].find(a[0], a[1], a[2]);
            Cell* pb = T.L[l].find(bb[0], bb[1], bb[2]);
            if ((pa && pa->pid >= 0) || (pb && pb->pid >= 0)) found = chk(S(l, a[0], a[1], a[2]));
          }
        }
        int src = found;
        if (found >= 0) {
          auto it = wpos[d].find(found);
          if (it != wpos[d].end()) src = w[d][it->second].second;  // written earlier in this pass
        }
        // later writes of the same target replace earlier ones
        auto it = wpos[d].find(gs);
        if (it != wpos[d].end()) {
          w[d][it->second].second = src;
        } else {
          wpos[d].emplace(gs, w[d].size());
          w[d].push_back({gs, src});
        }
      }
    }
    for (int d = 0; d < dim; d++)
      for (auto& pr : w[d]) fc.push_back({pr.first, pr.second, d});
    x.fc_n[l] = (int)fc.size() - x.fc_off[l];
    x.maxfc = std::max(x.maxfc, x.fc_n[l]);
  }
  if (!ok) return bmg::fail(BMG_ERR_ARG, "tree: index table refers to an unallocated cell");
  clk.mark("compile: tables");
  // upload
  TRY(upload(&x.nbr, nbr));
  TRY(upload(&x.colour, colour));
  TRY(upload(&x.rec_self, rself));
  TRY(upload(&x.rec_child0, rchild));
  TRY(upload(&x.rec_ring, rring));
  TRY(upload(&x.leaf_gs, leaf_gs));
  TRY(upload(&x.lvl_leaves, lvl_leaves));
  {
    std::vector<int> cl;
    x.col_off.assign(2 * (D + 1), 0);
    x.col_n.assign(2 * (D + 1), 0);
    x.lcol_off.assign(2 * (D + 1), 0);
    x.lcol_n.assign(2 * (D + 1), 0);
    for (int l = 0; l <= D; l++)
      for (int c = 0; c < 2; c++) {
        x.col_off[2 * l + c] = (int)cl.size();
        for (int q = 0; q < x.nA[l]; q++)
          if (colour[(size_t)x.abase[l] + q] == c) cl.push_back(x.off[l] + q);
        x.col_n[2 * l + c] = (int)cl.size() - x.col_off[2 * l + c];
        x.lcol_off[2 * l + c] = (int)cl.size();
        for (int q = 0; q < x.leaf_n[l]; q++) {
          const int gs = lvl_leaves[(size_t)x.leaf_off[l] + q];
          if (colour[(size_t)x.abase[l] + (gs - x.off[l])] == c) cl.push_back(gs);
        }
        x.lcol_n[2 * l + c] = (int)cl.size() - x.lcol_off[2 * l + c];
      }
    TRY(upload(&x.col_list, cl));
    // the neighbour rows of the list entries (level m of slot gs: the one
    // whose slot range holds it), so that a colour pass reads them in order
    std::vector<int> cnb(cl.size() * 2 * dim);
    std::vector<int8_t> slot_level(x.nslot, -1);
    for (int l = 0; l <= D; l++)
      for (int q = 0; q < x.nA[l]; q++) slot_level[(size_t)x.off[l] + q] = (int8_t)l;
    std::atomic<bool> bad{false};
    par_for(cl.size(), [&](size_t e0, size_t e1) {
      for (size_t e = e0; e < e1; e++) {
        const int gs = cl[e], m = gs >= 0 && gs < (int)x.nslot ? slot_level[gs] : -1;
        if (m < 0) {
          bad = true;
          return;
        }
        const size_t ai = (size_t)x.abase[m] + (gs - x.off[m]);
        for (int k = 0; k < 2 * dim; k++) cnb[e * 2 * dim + k] = nbr[ai * 2 * dim + k];
      }
    });
    if (bad) return bmg::fail(BMG_ERR_ARG, "tree: colour list entry outside the active cells");
    TRY(upload(&x.col_nbr, cnb));
  }
  TRY(upload(&x.box, box));
  TRY(upload(&x.fops, fops));
  TRY(upload(&x.hf, hf));
  TRY(upload(&x.fr, fr));
  TRY(upload(&x.fp, fp));
  TRY(upload(&x.fc, fc));
  TRY(upload(&x.fv_gs, fv_gs));
  TRY(upload(&x.fv_d, fv_d));
  {
    std::vector<TLevelInfo> li(D + 1);
    for (int l = 0; l <= D; l++)
      li[l] = {x.off[l], x.nA[l], x.abase[l], x.rec_off[l], x.n_nonleaf[l], x.n_prolong[l],
               x.box_off[l], x.box_n[l], x.leaf_off[l], x.leaf_n[l]};
    TRY(upload(&x.lvinfo, li));
  }
  TRY(dalloc_((void**)&x.fc_tmp, sizeof(double) * std::max(1, x.maxfc)));
  // fields
  bool mixed = t->precision == BMG_MIXED;
  size_t ts = tsize(t);
  TRY(dalloc_((void**)&t->a, NS * 8));
  TRY(dalloc_((void**)&t->b, NS * 8));
  for (int k = 0; k < kUserScalars; k++)
    if (k == 0 || t->us_set[k]) TRY(dalloc_((void**)&t->us[k], NS * 8));
  TRY(dalloc_((void**)&t->res64, NS * 8));
  TRY(dalloc_((void**)&t->lam64, NS * 8));
  for (int d = 0; d < dim; d++) {
    TRY(dalloc_((void**)&t->g[d], NS * 8));
    TRY(dalloc_((void**)&t->al64[d], NS * 8));
    if (mixed)
      TRY(dalloc_(&t->alT[d], NS * ts));
    else
      t->alT[d] = t->al64[d];
  }
  if (mixed) {
    TRY(dalloc_(&t->lamT, NS * ts));
    TRY(dalloc_(&t->res, NS * ts));
  } else {
    t->lamT = t->lam64;
    t->res = t->res64;
  }
  TRY(dalloc_(&t->da, NS * ts));
  TRY(dalloc_(&t->tmp, NS * ts));
  for (int s = 0; s < 2 * dim; s++) {
    TRY(dalloc_((void**)&t->bcv[s], sizeof(double) * x.bc_off[D + 1]));
    // the levels given; deeper ones (a tree grown since) read zero
    size_t nv = std::min<size_t>(t->bc_host[s].size(), (size_t)x.bc_off[D + 1]);
    if (nv) HIPCHECK(hipMemcpy(t->bcv[s], t->bc_host[s].data(), sizeof(double) * nv, hipMemcpyHostToDevice));
  }
  TRY(dalloc_((void**)&t->stat, sizeof(TStat)));
  TRY(dalloc_((void**)&t->part, sizeof(double) * 1024));
  TRY(dalloc_((void**)&t->scalar, sizeof(double)));
  t->compiled = true;
  // restore leaf values set before the mesh changed
  auto restore = [&](std::vector<double>& v, double* f) -> int {
    if ((long)v.size() != x.nleaf) return BMG_OK;
    double* tmp = nullptr;
    TRY(dalloc_((void**)&tmp, v.size() * 8));
    HIPCHECK(hipMemcpy(tmp, v.data(), v.size() * 8, hipMemcpyHostToDevice));
    k_t_leaf_scatter<<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, tmp, f);
    HIPCHECK(hipStreamSynchronize(t->stream));
    dfree_(tmp);
    return BMG_OK;
  };
  TRY(restore(t->keep_a, t->a));
  TRY(restore(t->keep_b, t->b));
  for (int k = 0; k < kUserScalars; k++)
    if (t->us[k]) TRY(restore(t->keep_us[k], t->us[k]));
  TRY(restore(t->keep_lam, t->lam64));
  clk.mark("compile: upload and fields", t->stream);
  return BMG_OK;
}

// the tree's topology lives on the host; the GPU is needed from the first
// solver call on (no CPU execution path)
int ensure(bmg_tree* t) {
  if (!t->stream) {
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
      return bmg::fail(BMG_ERR_HIP, "no HIP device available");
    HIPCHECK(hipStreamCreateWithFlags(&t->own_stream, hipStreamDefault));
    t->stream = t->own_stream;
  }
  if (t->topo.dirty || !t->compiled) return compile(t);
  return BMG_OK;
}

// ------------------------------------------------------------ launches

template <typename T>
T* F(void* p) {
  return (T*)p;
}

// segments: the active cells of level l then the leaves of levels l-1..0
Segs level_or_leaf(const bmg_tree* t, int l) {
  const DevIndex& x = t->ix;
  Segs S;
  memset(&S, 0, sizeof S);
  int blocks = 0;
  for (int m = l; m >= 0; m--) {
    int s = S.n++;
    S.first_block[s] = blocks;
    S.abase[s] = x.abase[m];
    S.lvoff[s] = x.off[m];
    S.D[s] = t->L0 * (1. / (1 << m));
    if (m == l) {
      S.start[s] = x.off[m];
      S.count[s] = x.nA[m];
      S.list[s] = nullptr;
    } else {
      S.count[s] = x.leaf_n[m];
      S.list[s] = x.lvl_leaves + x.leaf_off[m];
    }
    blocks += (int)nb_(S.count[s]);
  }
  S.first_block[S.n] = blocks;
  return S;
}

// the same cells of one colour only (lists built at compilation)
Segs level_or_leaf_colour(const bmg_tree* t, int l, int c) {
  const DevIndex& x = t->ix;
  Segs S;
  memset(&S, 0, sizeof S);
  int blocks = 0;
  for (int m = l; m >= 0; m--) {
    int s = S.n++;
    S.first_block[s] = blocks;
    S.abase[s] = x.abase[m];
    S.lvoff[s] = x.off[m];
    S.D[s] = t->L0 * (1. / (1 << m));
    const int off = m == l ? x.col_off[2 * m + c] : x.lcol_off[2 * m + c];
    S.count[s] = m == l ? x.col_n[2 * m + c] : x.lcol_n[2 * m + c];
    S.list[s] = x.col_list + off;
    S.nbl[s] = x.col_nbr + (size_t)off * 2 * t->dim;
    blocks += (int)nb_(S.count[s]);
  }
  S.first_block[S.n] = blocks;
  return S;
}

// segments: the leaves of every level
Segs all_leaves(const bmg_tree* t) {
  const DevIndex& x = t->ix;
  Segs S;
  memset(&S, 0, sizeof S);
  int blocks = 0;
  for (int m = 0; m <= x.depth; m++) {
    int s = S.n++;
    S.first_block[s] = blocks;
    S.abase[s] = x.abase[m];
    S.lvoff[s] = x.off[m];
    S.D[s] = t->L0 * (1. / (1 << m));
    S.count[s] = x.leaf_n[m];
    S.list[s] = x.lvl_leaves + x.leaf_off[m];
    blocks += (int)nb_(S.count[s]);
  }
  S.first_block[S.n] = blocks;
  return S;
}

double delta(const bmg_tree* t, int l) { return t->L0 * (1. / (1 << l)); }

BCT bct(const bmg_tree* t) {
  BCT b;
  for (int s = 0; s < 6; s++) {
    b.kind[s] = t->bc_kind[s];
    b.v[s] = t->bcv[s];
  }
  return b;
}

template <typename T>
void box(bmg_tree* t, T* f, int bk, int l) {
  const DevIndex& x = t->ix;
  int n = x.box_n[l];
  if (n) k_t_box<T><<<nb_(n), kB, 0, t->stream>>>(x.box + x.box_off[l], n, f, bct(t), bk, delta(t, l));
}

template <typename T, int DIM>
void restrict_level(bmg_tree* t, T* f, int l) {
  const DevIndex& x = t->ix;
  int n = x.n_nonleaf[l];
  if (n)
    k_t_restrict<T, DIM>
        <<<nb_(n), kB, 0, t->stream>>>(x.rec_self + x.rec_off[l], x.rec_child0 + x.rec_off[l], n, f);
}

template <typename T, int DIM>
void prolong_records(bmg_tree* t, T* f, int l, bool halo) {
  const DevIndex& x = t->ix;
  int r0 = x.rec_off[l] + (halo ? x.n_nonleaf[l] : 0);
  int n = halo ? x.n_prolong[l] : x.n_nonleaf[l];
  const int NR = DIM == 3 ? 27 : 9;
  if (n)
    k_t_prolong<T, DIM><<<nb_(n), kB, 0, t->stream>>>(x.rec_ring + (size_t)r0 * NR, x.rec_child0 + r0, n, f);
}

template <typename T>
RelaxCoef<T> relax_coef(bmg_tree* t) {
  RelaxCoef<T> k;
  for (int d = 0; d < 3; d++) k.alpha[d] = t->alpha_field && d < t->dim ? F<T>(t->alT[d]) : nullptr;
  k.alpha_c = (T)t->alpha_c;
  k.lam = t->lambda_field ? F<T>(t->lamT) : nullptr;
  k.lam_c = (T)t->lambda_c;
  return k;
}

template <typename T, int DIM>
int relax(bmg_tree* t, int l) {
  const DevIndex& x = t->ix;
  Segs S = level_or_leaf(t, l);
  unsigned nbk = (unsigned)S.first_block[S.n];
  if (!nbk) return BMG_OK;
  RelaxCoef<T> k = relax_coef<T>(t);
  T* da = F<T>(t->da);
  T* res = F<T>(t->res);
  if (t->smoother == BMG_RELAX_JACOBI) {
    {
      TProf tp(t, seg_cells(S));
      k_t_relax<T, DIM><<<nbk, kB, 0, t->stream>>>(S, da, F<T>(t->tmp), res, x.nbr, x.colour, k, -1);
    }
    k_t_jacobi<T><<<nbk, kB, 0, t->stream>>>(S, da, F<T>(t->tmp));
  } else if (t->smoother == BMG_RELAX_RBGS) {
    for (int c = 0; c < 2; c++) {
      if (c) for (int m = 0; m <= l; m++) box<T>(t, da, BK_HOM, m);
      Segs SC = level_or_leaf_colour(t, l, c);
      unsigned nbc = (unsigned)SC.first_block[SC.n];
      if (nbc) {
        TProf tp(t, seg_cells(SC));
        k_t_relax<T, DIM><<<nbc, kB, 0, t->stream>>>(SC, da, nullptr, res, x.nbr, x.colour, k, 2);
      }
    }
  } else
    return bmg::fail(BMG_ERR_UNSUPPORTED, "the in-place Gauss-Seidel smoother is sequential; use JACOBI or RBGS");
  return BMG_OK;
}

// Operations of at most small_op elements (default kSmallOp) are queued and
// run by one workgroup (k_t_ops); a larger one first flushes the queue, then
// runs as its own kernel, so the order of the chain is kept.

template <typename T, int DIM>
struct Batcher {
  bmg_tree* t;
  const DevIndex& x;
  TOps P;
  TreeArgs<T> A;
  Batcher(bmg_tree* t_, T* f0, T* f1, T* f2) : t(t_), x(t_->ix) {
    P.n = 0;
    memset(&A, 0, sizeof A);
    A.lv = x.lvinfo;
    A.rec_self = x.rec_self;
    A.rec_child0 = x.rec_child0;
    A.rec_ring = x.rec_ring;
    A.nbr = x.nbr;
    A.lvl_leaves = x.lvl_leaves;
    A.colour = x.colour;
    A.box = x.box;
    A.fld[0] = f0;
    A.fld[1] = f1;
    A.fld[2] = f2;
    A.res = f1;
    A.k = relax_coef<T>(t);
    A.bc = bct(t);
    A.L0 = t->L0;
  }
  bool small(long n) const { return t->batching && n <= t->small_op; }
  void flush() {
    if (P.n) k_t_ops<T, DIM><<<1, t->ops_block, 0, t->stream>>>(P, A);
    P.n = 0;
  }
  void push(int code, int level, int arg, int field) {
    if (P.n == kMaxOps) flush();
    P.op[P.n++] = TOp{(unsigned char)code, (unsigned char)field, (signed char)arg, (unsigned char)level};
  }
  void box(int f, int bk, int l) {
    int n = x.box_n[l];
    if (!n) return;
    if (small(n)) return push(TOP_BOX, l, bk, f);
    flush();
    k_t_box<T><<<nb_(n), kB, 0, t->stream>>>(x.box + x.box_off[l], n, A.fld[f], A.bc, bk, delta(t, l));
  }
  void restrict_level(int f, int l) {
    int n = x.n_nonleaf[l];
    if (!n) return;
    if (small(n)) return push(TOP_RESTRICT, l, 0, f);
    flush();
    k_t_restrict<T, DIM><<<nb_(n), kB, 0, t->stream>>>(x.rec_self + x.rec_off[l], x.rec_child0 + x.rec_off[l], n,
                                                       A.fld[f]);
  }
  void prolong(int f, int l, bool halo) {
    int r0 = x.rec_off[l] + (halo ? x.n_nonleaf[l] : 0);
    int n = halo ? x.n_prolong[l] : x.n_nonleaf[l];
    if (!n) return;
    if (small(n)) return push(TOP_PROLONG, l, halo ? 1 : 0, f);
    flush();
    const int NR = DIM == 3 ? 27 : 9;
    k_t_prolong<T, DIM><<<nb_(n), kB, 0, t->stream>>>(x.rec_ring + (size_t)r0 * NR, x.rec_child0 + r0, n, A.fld[f]);
  }
  // tree_boundary_level ({f}, l) (l < 0: depth)
  void tree_boundary(int f, int bk, int l) {
    int depth = l < 0 ? x.depth : l;
    if (t->topo.full()) {
      box(f, bk, depth);
      return;
    }
    for (int m = depth - 1; m >= 0; m--) restrict_level(f, m);
    box(f, bk, 0);
    for (int i = 0; i < depth; i++) {
      prolong(f, i, true);
      box(f, bk, i + 1);
    }
  }
  long cells(int l) const {
    long n = x.nA[l];
    for (int m = 0; m < l; m++) n += x.leaf_n[m];
    return n;
  }
  void fill(int l) {
    if (small(cells(l))) return push(TOP_FILL, l, 0, 0);
    flush();
    Segs S = level_or_leaf(t, l);
    if (S.first_block[S.n]) k_t_fill<T><<<(unsigned)S.first_block[S.n], kB, 0, t->stream>>>(S, A.fld[0], T(0));
  }
  int relax(int l) {
    if (t->smoother != BMG_RELAX_JACOBI && t->smoother != BMG_RELAX_RBGS)
      return bmg::fail(BMG_ERR_UNSUPPORTED,
                       "the in-place Gauss-Seidel smoother is sequential; use JACOBI or RBGS");
    const bool jac = t->smoother == BMG_RELAX_JACOBI;
    if (small(cells(l))) {
      if (jac) {
        push(TOP_RELAX, l, -1, 0);
        push(TOP_JACOBI, l, 0, 0);
      } else {
        push(TOP_RELAX, l, 0, 0);
        for (int m = 0; m <= l; m++) box(0, BK_HOM, m);
        push(TOP_RELAX, l, 1, 0);
      }
      return BMG_OK;
    }
    flush();
    Segs S = level_or_leaf(t, l);
    unsigned nbk = (unsigned)S.first_block[S.n];
    if (!nbk) return BMG_OK;
    T* da = A.fld[0];
    if (jac) {
      {
        TProf tp(t, seg_cells(S));
        k_t_relax<T, DIM><<<nbk, kB, 0, t->stream>>>(S, da, A.fld[2], A.res, x.nbr, x.colour, A.k, -1);
      }
      k_t_jacobi<T><<<nbk, kB, 0, t->stream>>>(S, da, A.fld[2]);
    } else {
      for (int c = 0; c < 2; c++) {
        if (c) {
          for (int m = 0; m <= l; m++) box(0, BK_HOM, m);
          flush();
        }
        Segs SC = level_or_leaf_colour(t, l, c);
        unsigned nbc = (unsigned)SC.first_block[SC.n];
        if (nbc) {
          TProf tp(t, seg_cells(SC));
          k_t_relax<T, DIM><<<nbc, kB, 0, t->stream>>>(SC, da, nullptr, A.res, x.nbr, x.colour, A.k, 2);
        }
      }
    }
    return BMG_OK;
  }
  // restriction (res) of mg_cycle (tree_restriction), on field f
  void restriction_res(int f) {
    int D = x.depth;
    if (t->topo.full()) {
      box(f, BK_SYM, D);
      for (int l = D - 1; l >= 0; l--) {
        restrict_level(f, l);
        box(f, BK_SYM, l);
      }
    } else
      tree_boundary(f, BK_SYM, -1);
  }
};

// residual (TREE form) in FP64; res in the cycle precision as well
template <typename T, int DIM>
void residual_kernels(bmg_tree* t) {
  const DevIndex& x = t->ix;
  {
    Batcher<double, DIM> B(t, t->a, nullptr, nullptr);
    B.tree_boundary(0, BK_A, -1);
    B.flush();
  }
  const double* al[3] = {nullptr, nullptr, nullptr};
  if (t->alpha_field)
    for (int d = 0; d < DIM; d++) al[d] = t->al64[d];
  for (int l = 0; l <= x.depth; l++) {
    int n = x.face_n[l];
    if (n)
      k_t_faces<<<nb_(n), kB, 0, t->stream>>>(x.fops + x.face_off[l], n, t->a, t->g[0], t->g[1], t->g[2], al[0],
                                              al[1], al[2], t->alpha_c, delta(t, l));
  }
  if (x.nhf) k_t_face_avg<DIM><<<nb_(x.nhf), kB, 0, t->stream>>>(x.hf, x.nhf, t->g[0], t->g[1], t->g[2]);
  k_t_stat_reset<<<1, 64, 0, t->stream>>>(t->stat);
  Segs S = all_leaves(t);
  const double* lam = t->lambda_field ? t->lam64 : nullptr;
  k_t_leaf_res<DIM, T><<<(unsigned)S.first_block[S.n], kB, 0, t->stream>>>(
      S, t->a, t->b, t->g[0], t->g[1], t->g[2], x.nbr, lam, t->lambda_c, t->res64, F<T>(t->res), t->stat);
}

// runs `launches` (kernel launches on t->stream only) as a captured graph,
// instantiated on first use and replayed afterwards
template <typename F>
int run_graph(bmg_tree* t, hipGraphExec_t* slot, F launches) {
  if (!t->use_graphs) return launches();
  if (!*slot) {
    hipGraph_t graph = nullptr;
    HIPCHECK(hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal));
    int rc = launches();
    hipError_t e = hipStreamEndCapture(t->stream, &graph);
    if (rc != BMG_OK) {
      if (e == hipSuccess && graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return bmg::fail(BMG_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return bmg::fail(BMG_ERR_HIP, "graph instantiation failed: %s", hipGetErrorString(e));
    *slot = exec;
  }
  HIPCHECK(hipGraphLaunch(*slot, t->stream));
  return BMG_OK;
}

template <typename T, int DIM>
int residual(bmg_tree* t, double* maxres) {
  TRY(run_graph(t, &t->residual_graph, [&]() {
    residual_kernels<T, DIM>(t);
    return BMG_OK;
  }));
  TStat st;
  HIPCHECK(hipMemcpyAsync(&st, t->stat, sizeof st, hipMemcpyDeviceToHost, t->stream));
  HIPCHECK(hipStreamSynchronize(t->stream));
  HIPCHECK(hipGetLastError());
  double m;
  memcpy(&m, &st.maxbits, sizeof m);
  *maxres = m;
  return st.nans ? bmg::fail(BMG_ERR_NAN, "residual: %llu NaN values", st.nans) : BMG_OK;
}

// restriction ({alpha, lambda}) of poisson() on the tree, in FP64, then the
// copies in the cycle precision
template <typename T, int DIM>
int coefficients(bmg_tree* t) {
  const DevIndex& x = t->ix;
  if (t->coef_ready) return BMG_OK;
  bool full = t->topo.full();
  int D = x.depth;
  auto faces_restrict = [&](int l) {
    int n = x.fr_n[l];
    if (n) k_t_face_avg<DIM><<<nb_(n), kB, 0, t->stream>>>(x.fr + x.fr_off[l], n, t->al64[0], t->al64[1], t->al64[2]);
  };
  auto faces_box = [&](int l) {
    int n = x.fc_n[l];
    if (!n) return;
    k_t_face_gather<<<nb_(n), kB, 0, t->stream>>>(x.fc + x.fc_off[l], n, t->al64[0], t->al64[1], t->al64[2],
                                                  x.fc_tmp);
    k_t_face_scatter<<<nb_(n), kB, 0, t->stream>>>(x.fc + x.fc_off[l], n, t->al64[0], t->al64[1], t->al64[2],
                                                   x.fc_tmp);
  };
  if (t->alpha_field && x.nhf)
    k_t_face_avg<DIM><<<nb_(x.nhf), kB, 0, t->stream>>>(x.hf, x.nhf, t->al64[0], t->al64[1], t->al64[2]);
  if (full) {
    if (t->lambda_field) box<double>(t, t->lam64, BK_SYM, D);
    if (t->alpha_field) faces_box(D);
    for (int l = D - 1; l >= 0; l--) {
      if (t->lambda_field) restrict_level<double, DIM>(t, t->lam64, l);
      if (t->alpha_field) faces_restrict(l);
      if (t->lambda_field) box<double>(t, t->lam64, BK_SYM, l);
      if (t->alpha_field) faces_box(l);
    }
  } else {
    for (int l = D - 1; l >= 0; l--) {
      if (t->lambda_field) restrict_level<double, DIM>(t, t->lam64, l);
      if (t->alpha_field) faces_restrict(l);
    }
    if (t->lambda_field) box<double>(t, t->lam64, BK_SYM, 0);
    if (t->alpha_field) faces_box(0);
    for (int i = 0; i < D; i++) {
      if (t->lambda_field) prolong_records<double, DIM>(t, t->lam64, i, true);
      if (t->alpha_field && x.fp_n[i])
        k_t_face_prolong<DIM><<<nb_(x.fp_n[i]), kB, 0, t->stream>>>(x.fp + x.fp_off[i], x.fp_n[i], t->al64[0],
                                                                    t->al64[1], t->al64[2]);
      if (t->lambda_field) box<double>(t, t->lam64, BK_SYM, i + 1);
      if (t->alpha_field) faces_box(i + 1);
    }
  }
  if (t->precision == BMG_MIXED) {
    long NS = x.nslot;
    if (t->alpha_field)
      for (int d = 0; d < DIM; d++)
        k_t_convert<double, float><<<nb_(NS), kB, 0, t->stream>>>(t->al64[d], F<float>(t->alT[d]), NS);
    if (t->lambda_field) k_t_convert<double, float><<<nb_(NS), kB, 0, t->stream>>>(t->lam64, F<float>(t->lamT), NS);
  }
  HIPCHECK(hipGetLastError());
  t->coef_ready = true;
  return BMG_OK;
}

template <typename T, int DIM>
int cycle_kernels(bmg_tree* t, int nrelax, int minlevel, int maxlevel) {
  const DevIndex& x = t->ix;
  T* da = F<T>(t->da);
  Batcher<T, DIM> B(t, da, F<T>(t->res), F<T>(t->tmp));
  B.restriction_res(1);
  minlevel = std::min(minlevel, maxlevel);
  for (int l = minlevel; l <= maxlevel; l++) {
    if (l == minlevel)
      B.fill(l);
    else
      B.prolong(0, l - 1, false);
    B.tree_boundary(0, BK_HOM, l);
    for (int i = 0; i < nrelax; i++) {
      TRY(B.relax(l));
      B.tree_boundary(0, BK_HOM, l);
    }
  }
  B.flush();
  k_t_correct<T><<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, t->a, da);
  HIPCHECK(hipGetLastError());
  return BMG_OK;
}

template <typename T, int DIM>
int mg_cycle(bmg_tree* t, int nrelax, int minlevel, int maxlevel) {
  TRY(coefficients<T, DIM>(t));
  if (!t->use_graphs || t->prof) {
    TRY((cycle_kernels<T, DIM>(t, nrelax, minlevel, maxlevel)));
    tprof_collect(t);
    return BMG_OK;
  }
  hipGraphExec_t& slot = t->cycle_graphs[std::make_tuple(nrelax, minlevel, maxlevel)];
  return run_graph(t, &slot, [&]() { return cycle_kernels<T, DIM>(t, nrelax, minlevel, maxlevel); });
}

template <int DIM>
int dispatch_cycle(bmg_tree* t, int nrelax, int minlevel, int maxlevel) {
  return t->precision == BMG_MIXED ? mg_cycle<float, DIM>(t, nrelax, minlevel, maxlevel)
                                   : mg_cycle<double, DIM>(t, nrelax, minlevel, maxlevel);
}

int cycle(bmg_tree* t, int nrelax, int minlevel, int maxlevel) {
  return t->dim == 3 ? dispatch_cycle<3>(t, nrelax, minlevel, maxlevel)
                     : dispatch_cycle<2>(t, nrelax, minlevel, maxlevel);
}

int resid(bmg_tree* t, double* m) {
  if (t->dim == 3)
    return t->precision == BMG_MIXED ? residual<float, 3>(t, m) : residual<double, 3>(t, m);
  return t->precision == BMG_MIXED ? residual<float, 2>(t, m) : residual<double, 2>(t, m);
}

int coefs(bmg_tree* t) {
  if (t->dim == 3)
    return t->precision == BMG_MIXED ? coefficients<float, 3>(t) : coefficients<double, 3>(t);
  return t->precision == BMG_MIXED ? coefficients<float, 2>(t) : coefficients<double, 2>(t);
}

int sum_b(bmg_tree* t, double* out) {
  const DevIndex& x = t->ix;
  int np = (int)std::min<long>(1024, std::max<long>(1, (long)nb_(x.nleaf)));
  k_t_sum_partial<<<np, kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, t->b, t->part);
  k_t_sum_final<<<1, 64, 0, t->stream>>>(t->part, np, t->scalar);
  HIPCHECK(hipMemcpyAsync(out, t->scalar, sizeof(double), hipMemcpyDeviceToHost, t->stream));
  HIPCHECK(hipStreamSynchronize(t->stream));
  return BMG_OK;
}

int mg_solve_impl(bmg_tree* t, double tolerance, int nrelax, int minlevel, bmg_mgstats* out) {
  bmg_mgstats s;
  memset(&s, 0, sizeof s);
  TRY(sum_b(t, &s.sum));
  s.nrelax = nrelax > 0 ? nrelax : 4;
  double resb;
  int rc = resid(t, &resb);
  if (rc != BMG_OK && rc != BMG_ERR_NAN) return rc;
  int nan = rc == BMG_ERR_NAN;
  s.resb = s.resa = resb;
  if (tolerance == 0.) tolerance = t->TOLERANCE;
  for (s.i = 0; s.i < t->NITERMAX && (s.i < t->NITERMIN || s.resa > tolerance); s.i++) {
    TRY(cycle(t, s.nrelax, minlevel, t->ix.depth));
    rc = resid(t, &s.resa);
    if (rc != BMG_OK && rc != BMG_ERR_NAN) return rc;
    nan |= rc == BMG_ERR_NAN;
    if (s.resa > tolerance) {
      if (resb / s.resa < 1.2 && s.nrelax < 100)
        s.nrelax++;
      else if (resb / s.resa > 10 && s.nrelax > 2)
        s.nrelax--;
    }
    resb = s.resa;
    if (nan) break;
  }
  s.minlevel = minlevel;
  if (out) *out = s;
  return nan ? bmg::fail(BMG_ERR_NAN, "residual is NaN") : BMG_OK;
}

int set_leaf_values(bmg_tree* t, double* f, const double* v, int on_device) {
  const DevIndex& x = t->ix;
  const double* src = v;
  double* tmp = nullptr;
  if (!on_device) {
    TRY(dalloc_((void**)&tmp, x.nleaf * 8));
    HIPCHECK(hipMemcpy(tmp, v, x.nleaf * 8, hipMemcpyHostToDevice));
    src = tmp;
  }
  k_t_leaf_scatter<<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, src, f);
  HIPCHECK(hipStreamSynchronize(t->stream));
  dfree_(tmp);
  return BMG_OK;
}

template <typename T>
int get_leaf_values(bmg_tree* t, const T* f, double* v, int on_device) {
  const DevIndex& x = t->ix;
  double* dst = v;
  double* tmp = nullptr;
  if (!on_device) {
    TRY(dalloc_((void**)&tmp, x.nleaf * 8));
    dst = tmp;
  }
  k_t_leaf_gather<T><<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, f, dst);
  HIPCHECK(hipStreamSynchronize(t->stream));
  if (!on_device) HIPCHECK(hipMemcpy(v, tmp, x.nleaf * 8, hipMemcpyDeviceToHost));
  dfree_(tmp);
  return BMG_OK;
}

// ------------------------------------------------------------ adapt_wavelet

// the fields that follow the mesh (adapt_wavelet's default list, all
// fields): the user scalars that exist, a and b, in groups of three for the
// batcher, with their boundary kinds
struct FieldGroup {
  double* f[3];
  int bk[3];
};
std::vector<FieldGroup> mesh_fields(double* const us[kUserScalars], double* a, double* b) {
  std::vector<FieldGroup> g(1);
  g[0] = FieldGroup{{us[0], a, b}, {BK_SYM, BK_A, BK_SYM}};
  bool extra = false;
  for (int k = 1; k < kUserScalars; k++) extra |= us[k] != nullptr;
  if (extra) g.push_back(FieldGroup{{us[1], us[2], us[3]}, {BK_SYM, BK_SYM, BK_SYM}});
  static_assert(kUserScalars == 4, "one group of three extra user scalars");
  return g;
}
double* const* user_scalars(bmg_tree* t) { return t->us; }

// the arrays of the fields that follow the mesh (user scalars, then a and
// b), taken out of the tree before a recompilation, which allocates new
// ones (user scalars that are set); the leaf values kept for a plain
// recompilation are dropped
void take_mesh_fields(bmg_tree* t, double* olds[kUserScalars + 2]) {
  for (int k = 0; k < kUserScalars; k++) {
    olds[k] = t->us_set[k] ? t->us[k] : nullptr;
    if (!olds[k]) dfree_(t->us[k]);
    t->us[k] = nullptr;
    t->keep_us[k].clear();
  }
  olds[kUserScalars] = t->a;
  olds[kUserScalars + 1] = t->b;
  t->a = t->b = nullptr;
  t->keep_a.clear();
  t->keep_b.clear();
  t->keep_lam.clear();
}

// adapt_wavelet's boundary (list) and restriction (slist) for the fields
// that follow the mesh: every level restricted and its ghosts and halos
// filled (a, b and the scalars not estimated are also restricted on a full
// tree, where the reference leaves their coarse levels alone: coarsen_cell
// recomputes the same averages)
template <int DIM>
void adapt_boundary(bmg_tree* t) {
  for (const FieldGroup& G : mesh_fields(user_scalars(t), t->a, t->b)) {
    Batcher<double, DIM> B(t, G.f[0], G.f[1], G.f[2]);
    if (t->topo.full()) {
      const int D = t->ix.depth;
      for (int f = 0; f < 3; f++) {
        if (!G.f[f]) continue;
        B.box(f, G.bk[f], D);
        for (int l = D - 1; l >= 0; l--) {
          B.restrict_level(f, l);
          B.box(f, G.bk[f], l);
        }
      }
    } else {
      for (int f = 0; f < 3; f++)
        if (G.f[f]) B.tree_boundary(f, G.bk[f], -1);
    }
    B.flush();
  }
}

// the wavelet estimate of scalar s at every child of a refined cell, by slot
template <int DIM>
int adapt_estimate(bmg_tree* t, const double* s, std::vector<double>& err_h) {
  const DevIndex& x = t->ix;
  double* err = nullptr;
  TRY(dalloc_((void**)&err, x.nslot * 8));
  const int NR = DIM == 3 ? 27 : 9;
  for (int l = 0; l <= x.depth; l++) {
    int n = x.n_nonleaf[l];
    if (n)
      k_t_wavelet<DIM><<<nb_(n), kB, 0, t->stream>>>(x.rec_ring + (size_t)x.rec_off[l] * NR,
                                                     x.rec_child0 + x.rec_off[l], n, s, err);
  }
  err_h.resize(x.nslot);
  hipError_t e = hipMemcpyAsync(err_h.data(), err, x.nslot * 8, hipMemcpyDeviceToHost, t->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(t->stream);
  dfree_(err);
  if (e != hipSuccess) return bmg::fail(BMG_ERR_HIP, "wavelet estimate: %s", hipGetErrorString(e));
  return BMG_OK;
}

// the fields that follow the mesh in the new layout: the old value of every
// cell that stays, the bilinear of the old ring in the children of the
// refined cells, then the tree boundary (mpi_boundary_update (list)).
// olds: the old arrays of the user scalars, then a and b (null: none)
template <int DIM>
int adapt_remap(bmg_tree* t, double* const olds[kUserScalars + 2], long old_nslot, const std::vector<int>& prev,
                const RefineLog& log) {
  const DevIndex& x = t->ix;
  Topo& T = t->topo;
  // every slot the kernels read lies in the old arrays
  if ((long)prev.size() != x.nslot) return bmg::fail(BMG_ERR_ARG, "tree adaptation: slot map size");
  for (int p : prev)
    if (p >= old_nslot) return bmg::fail(BMG_ERR_ARG, "tree adaptation: slot map out of range");
  for (int p : log.ring)
    if (p < 0 || p >= old_nslot) return bmg::fail(BMG_ERR_ARG, "tree adaptation: ring outside the old layout");
  const size_t nr = log.level.size();
  std::vector<int> child0(nr);
  for (size_t r = 0; r < nr; r++) {
    const int l = log.level[r];
    const IJK c = log.cell[r];
    int m = 0;
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        for (int e = 0; e < T.kext(); e++, m++) {
          Cell* ch = T.L[l + 1].find(2 * c.i - GH + a, 2 * c.j - GH + b, DIM == 3 ? 2 * c.k - GH + e : 0);
          if (!ch || (m > 0 && ch->slot != child0[r] + m))
            return bmg::fail(BMG_ERR_ARG, "tree adaptation: children of a refined cell not in one block");
          if (m == 0) child0[r] = ch->slot;
        }
  }
  int *dprev = nullptr, *dring = nullptr, *dchild = nullptr;
  int rc = upload(&dprev, prev);
  if (rc == BMG_OK) rc = upload(&dring, log.ring);
  if (rc == BMG_OK) rc = upload(&dchild, child0);
  if (rc == BMG_OK) {
    double* news[kUserScalars + 2];
    for (int k = 0; k < kUserScalars; k++) news[k] = t->us[k];
    news[kUserScalars] = t->a;
    news[kUserScalars + 1] = t->b;
    for (int f = 0; f < kUserScalars + 2; f++) {
      if (!olds[f] || !news[f]) continue;
      k_t_remap<<<nb_(x.nslot), kB, 0, t->stream>>>(dprev, x.nslot, olds[f], news[f]);
      if (nr) k_t_refine_remap<DIM><<<nb_((long)nr), kB, 0, t->stream>>>(dring, dchild, (int)nr, olds[f], news[f]);
    }
    for (const FieldGroup& G : mesh_fields(user_scalars(t), t->a, t->b)) {
      Batcher<double, DIM> B(t, G.f[0], G.f[1], G.f[2]);
      for (int f = 0; f < 3; f++)
        if (G.f[f]) B.tree_boundary(f, G.bk[f], -1);
      B.flush();
    }
    hipError_t e = hipStreamSynchronize(t->stream);
    if (e != hipSuccess) rc = bmg::fail(BMG_ERR_HIP, "tree adaptation: %s", hipGetErrorString(e));
  }
  dfree_(dprev);
  dfree_(dring);
  dfree_(dchild);
  return rc;
}

// ------------------------------------------------------------ dump / restore

// struct DumpHeader, output.h:1049-1053 (coord = 3 doubles)
struct DumpHeaderC {
  double t;
  long len;
  int i, depth, npe, version;
  double n[3];
};
static_assert(sizeof(DumpHeaderC) == 56, "the reference's DumpHeader layout");
constexpr int kDumpVersion = 170901;
constexpr unsigned kDumpLeaf = 2;  // the `leaf` flag, grid/tree.h:31-34

// the active cells in foreach_cell() order (pre-order, stopping at leaves)
void dump_preorder(Topo& T, int l, IJK c, std::vector<Cell*>& out) {
  Cell* cp = T.L[l].find(c.i, c.j, c.k);
  if (!cp) return;
  out.push_back(cp);
  if (T.is_leaf(*cp) || l >= T.depth) return;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < T.kext(); e++)
        dump_preorder(T, l + 1, {2 * c.i - GH + a, 2 * c.j - GH + b, T.dim == 3 ? 2 * c.k - GH + e : 0}, out);
}

// subtree sizes in pre-order (subtree_size (size, false),
// grid/multigrid-common.h:422-445: a leaf 1, a parent 1 + its children)
size_t dump_sizes(const std::vector<Cell*>& cells, int dim, size_t q, std::vector<double>& size) {
  if (cells[q]->flags & C_LEAF) {
    size[q] = 1.;
    return q + 1;
  }
  double sum = 1.;
  size_t r = q + 1;
  for (int m = 0; m < (1 << dim); m++) {
    const size_t nxt = dump_sizes(cells, dim, r, size);
    sum += size[r];
    r = nxt;
  }
  size[q] = sum;
  return r;
}

double* field_ptr(bmg_tree* t, int field) {
  if (field >= BMG_F_S && field < BMG_F_S + kUserScalars) return t->us[field - BMG_F_S];
  return field == BMG_F_A ? t->a : field == BMG_F_B ? t->b : nullptr;
}

// restore(): the pre-order traversal of the growing tree, one record per
// visited cell, refining a leaf whose record is not a leaf; each cell's
// record index is kept in its slot until the tree is compiled
int restore_visit(Topo& T, int l, IJK c, FILE* fp, size_t reclen, std::vector<double>& rec, long& nrec,
                  const std::vector<int>& cols, std::vector<std::vector<double>>& vals) {
  Cell* cp = T.L[l].find(c.i, c.j, c.k);
  if (!cp) return BMG_OK;
  unsigned flags;
  if (fread(&flags, sizeof flags, 1, fp) != 1 || fread(rec.data(), sizeof(double), reclen, fp) != reclen)
    return bmg::fail(BMG_ERR_ARG, "restore: the file ends before the tree does");
  cp->slot = (int32_t)nrec++;
  for (size_t m = 0; m < cols.size(); m++) vals[m].push_back(cols[m] >= 0 ? rec[cols[m]] : 0.);
  if (!(flags & kDumpLeaf) && T.is_leaf(*cp)) {
    if (l + 1 >= kSeg || l + 1 > kMaxDepth) return bmg::fail(BMG_ERR_UNSUPPORTED, "restore: tree deeper than supported");
    T.refine_cell(l, c);
  }
  if (T.is_leaf(*cp) || l >= T.depth) return BMG_OK;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < T.kext(); e++)
        TRY(restore_visit(T, l + 1, {2 * c.i - GH + a, 2 * c.j - GH + b, T.dim == 3 ? 2 * c.k - GH + e : 0}, fp,
                          reclen, rec, nrec, cols, vals));
  return BMG_OK;
}

}  // namespace

// ------------------------------------------------------------ C ABI

extern "C" {

int bmg_tree_new(bmg_tree** out, int dimension, double L0, double X0, double Y0, double Z0) {
  if (!out) return bmg::fail(BMG_ERR_ARG, "null handle pointer");
  *out = nullptr;
  if (dimension != 2 && dimension != 3) return bmg::fail(BMG_ERR_ARG, "trees are 2D or 3D (got %d)", dimension);
  if (!(L0 > 0.)) return bmg::fail(BMG_ERR_ARG, "L0 must be positive");
  bmg_tree* t = new bmg_tree();
  t->dim = dimension;
  t->L0 = L0;
  t->X0[0] = X0;
  t->X0[1] = Y0;
  t->X0[2] = Z0;
  t->topo.init(dimension);
  const char* gr = getenv("BMG_GRAPHS");
  t->use_graphs = !(gr && gr[0] == '0');
  if (const char* ob = getenv("BMG_OPS_BLOCK")) {
    int v = atoi(ob);
    if (v >= 64 && v <= kOpsBlock && v % 64 == 0) t->ops_block = v;
  }
  const char* tb = getenv("BMG_TREE_BATCH");
  t->batching = !(tb && tb[0] == '0');
  if (const char* so = getenv("BMG_TREE_SMALL_OP")) {
    const long v = atol(so);
    if (v >= 64 && v <= (1L << 20)) t->small_op = v;
  }
  *out = t;
  return BMG_OK;
}

int bmg_tree_free(bmg_tree* t) {
  if (!t) return BMG_OK;
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  free_device(t);
  for (auto& u : t->ev_used) {
    (void)hipEventDestroy(u.second.first);
    (void)hipEventDestroy(u.second.second);
  }
  for (hipEvent_t e : t->ev_pool) (void)hipEventDestroy(e);
  if (t->own_stream) (void)hipStreamDestroy(t->own_stream);
  (void)hipGetLastError();
  delete t;
  return BMG_OK;
}

int bmg_tree_info(bmg_tree* t, int* depth, long* nleaves, long* ncells) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  t->topo.update();
  if (depth) *depth = t->topo.depth;
  if (nleaves) *nleaves = (long)t->topo.leaf_c.size();
  if (ncells) {
    long n = 0;
    for (auto& L : t->topo.L)
      for (auto& br : L.bricks)
        for (size_t q = 0; q < L.bsize(); q++) n += (br[q].flags & C_ALLOC) ? 1 : 0;
    *ncells = n;
  }
  return BMG_OK;
}

int bmg_tree_leaves(bmg_tree* t, int* level, double* xyz) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  Topo& T = t->topo;
  T.update();
  for (size_t n = 0; n < T.leaf_c.size(); n++) {
    int l = T.leaf_level[n];
    double D = t->L0 * (1. / (1 << l));
    if (level) level[n] = l;
    if (xyz) {
      xyz[3 * n] = (T.leaf_c[n].i - GH + 0.5) * D + t->X0[0];
      xyz[3 * n + 1] = (T.leaf_c[n].j - GH + 0.5) * D + t->X0[1];
      xyz[3 * n + 2] = t->dim == 3 ? (T.leaf_c[n].k - GH + 0.5) * D + t->X0[2] : 0.;
    }
  }
  return BMG_OK;
}

int bmg_tree_refine(bmg_tree* t, const unsigned char* mark, long* nrefined) {
  if (!t || !mark) return bmg::fail(BMG_ERR_ARG, "null argument");
  Topo& T = t->topo;
  T.update();
  // on a compiled tree the fields follow the mesh: boundary (all), then
  // refine_cell (point, all) logs each refined cell's ring for the remap
  const bool carry = t->compiled && !T.dirty;
  if (carry) {
    if (t->dim == 2)
      adapt_boundary<2>(t);
    else
      adapt_boundary<3>(t);
    HIPCHECK(hipStreamSynchronize(t->stream));
  }
  std::vector<int> ll = T.leaf_level;
  std::vector<IJK> lc = T.leaf_c;
  RefineLog log;
  long nr = 0;
  for (size_t q = 0; q < lc.size(); q++) {
    if (!mark[q]) continue;
    Cell& c = T.at(ll[q], lc[q].i, lc[q].j, lc[q].k);
    if (!(c.flags & C_LEAF)) continue;
    if (ll[q] + 1 >= kSeg || ll[q] + 1 > kMaxDepth)
      return bmg::fail(BMG_ERR_UNSUPPORTED, "refinement beyond level %d", std::min(kSeg - 1, kMaxDepth));
    T.refine_cell(ll[q], lc[q], 0, carry ? &log : nullptr);
    nr++;
  }
  if (nr) {
    T.dirty = true;
    t->coef_ready = false;
    if (carry) {
      if (!log.ok) {
        t->compiled = false;
        return bmg::fail(BMG_ERR_ARG, "tree refinement met an inconsistent tree");
      }
      double* olds[kUserScalars + 2];
      take_mesh_fields(t, olds);
      const long old_nslot = t->ix.nslot;
      std::vector<int> prev;
      t->prev_slots = &prev;
      int rc = compile(t);
      t->prev_slots = nullptr;
      if (rc == BMG_OK)
        rc = t->dim == 2 ? adapt_remap<2>(t, olds, old_nslot, prev, log) : adapt_remap<3>(t, olds, old_nslot, prev, log);
      for (double* p : olds) dfree_(p);
      if (rc != BMG_OK) return rc;
    } else {
      t->compiled = false;
      t->keep_a.clear();
      t->keep_b.clear();
      t->keep_lam.clear();
      for (int k = 0; k < kUserScalars; k++) {
        t->keep_us[k].clear();
        t->us_set[k] = false;
      }
    }
  }
  if (nrefined) *nrefined = nr;
  return BMG_OK;
}

int bmg_tree_adapt_wavelet_list(bmg_tree* t, int n, const int* fields, const double* max, int maxlevel,
                                int minlevel, int* nf, int* nc) {
  if (!t || !fields || !max) return bmg::fail(BMG_ERR_ARG, "null argument");
  if (n < 1 || n > kUserScalars) return bmg::fail(BMG_ERR_ARG, "1 to %d estimated scalars (got %d)", kUserScalars, n);
  for (int i = 0; i < n; i++) {
    if (fields[i] < BMG_F_S || fields[i] >= BMG_F_S + kUserScalars)
      return bmg::fail(BMG_ERR_ARG, "field %d is not a user scalar", fields[i]);
    if (!(max[i] >= 0.)) return bmg::fail(BMG_ERR_ARG, "max must be a non-negative number");
  }
  const int lim = std::min(kSeg - 1, kMaxDepth);
  if (maxlevel < 0 || maxlevel > lim) return bmg::fail(BMG_ERR_ARG, "maxlevel %d outside 0..%d", maxlevel, lim);
  TRY(ensure(t));
  for (int i = 0; i < n; i++)
    if (!t->us_set[fields[i] - BMG_F_S])
      return bmg::fail(BMG_ERR_ARG, "set the estimated scalar (field %d) first", fields[i]);
  Topo& T = t->topo;
  PhaseClock clk;
  std::vector<std::vector<double>> err(n);
  if (t->dim == 2)
    adapt_boundary<2>(t);
  else
    adapt_boundary<3>(t);
  for (int i = 0; i < n; i++) {
    const double* s = t->us[fields[i] - BMG_F_S];
    TRY(t->dim == 2 ? adapt_estimate<2>(t, s, err[i]) : adapt_estimate<3>(t, s, err[i]));
  }
  clk.mark("adapt: estimate", t->stream);
  Topo::Adapt A;
  for (int i = 0; i < n; i++) {
    A.err.push_back(&err[i]);
    A.max.push_back(max[i]);
  }
  A.maxlevel = maxlevel;
  A.minlevel = minlevel < 1 ? 1 : minlevel;
  const IJK root = {GH, GH, t->dim == 3 ? GH : 0};
  T.adapt_refine(0, root, A);
  for (int l = T.depth; l >= 0; l--) T.adapt_coarsen(0, root, l, A);
  clk.mark("adapt: refine and coarsen passes");
  if (!A.ok || !A.log.ok) {
    t->compiled = false;
    return bmg::fail(BMG_ERR_ARG, "tree adaptation met an inconsistent tree");
  }
  if (A.nf || A.nc) {
    T.dirty = true;
    // the fields that follow the mesh survive the recompilation in their
    // old layout
    double* olds[kUserScalars + 2];
    take_mesh_fields(t, olds);
    const long old_nslot = t->ix.nslot;
    std::vector<int> prev;
    t->prev_slots = &prev;
    clk.mark("adapt: before compile");
    int rc = compile(t);
    t->prev_slots = nullptr;
    clk.mark("adapt: compile");
    if (rc == BMG_OK)
      rc = t->dim == 2 ? adapt_remap<2>(t, olds, old_nslot, prev, A.log)
                       : adapt_remap<3>(t, olds, old_nslot, prev, A.log);
    for (double* p : olds) dfree_(p);
    clk.mark("adapt: remap", t->stream);
    if (rc != BMG_OK) return rc;
    t->coef_ready = false;
  }
  if (nf) *nf = (int)A.nf;
  if (nc) *nc = (int)A.nc;
  return BMG_OK;
}

int bmg_tree_adapt_wavelet(bmg_tree* t, double max, int maxlevel, int minlevel, int* nf, int* nc) {
  const int s = BMG_F_S;
  return bmg_tree_adapt_wavelet_list(t, 1, &s, &max, maxlevel, minlevel, nf, nc);
}

int bmg_tree_nonleaf(bmg_tree* t, long* n, int* level, double* xyz) {
  if (!t || !n) return bmg::fail(BMG_ERR_ARG, "null argument");
  Topo& T = t->topo;
  std::vector<int> lev;
  std::vector<IJK> cells;
  T.nonleaf(0, {GH, GH, t->dim == 3 ? GH : 0}, lev, cells);
  *n = (long)cells.size();
  for (size_t q = 0; q < cells.size(); q++) {
    const double D = t->L0 * (1. / (1 << lev[q]));
    if (level) level[q] = lev[q];
    if (xyz) {
      xyz[3 * q] = (cells[q].i - GH + 0.5) * D + t->X0[0];
      xyz[3 * q + 1] = (cells[q].j - GH + 0.5) * D + t->X0[1];
      xyz[3 * q + 2] = t->dim == 3 ? (cells[q].k - GH + 0.5) * D + t->X0[2] : 0.;
    }
  }
  return BMG_OK;
}

int bmg_tree_unrefine(bmg_tree* t, const unsigned char* mark, long* ncoarsened) {
  if (!t || !mark) return bmg::fail(BMG_ERR_ARG, "null argument");
  TRY(ensure(t));
  Topo& T = t->topo;
  std::vector<int> lev;
  std::vector<IJK> cells;
  const IJK root = {GH, GH, t->dim == 3 ? GH : 0};
  T.nonleaf(0, root, lev, cells);
  // the fields hold their restriction where coarsen_cell reads it
  if (t->dim == 2)
    adapt_boundary<2>(t);
  else
    adapt_boundary<3>(t);
  HIPCHECK(hipStreamSynchronize(t->stream));
  for (size_t q = 0; q < cells.size(); q++)
    if (mark[q]) T.at(lev[q], cells[q].i, cells[q].j, cells[q].k).flags |= C_TOO_FINE;
  long nc = 0;
  bool ok = true;
  for (int l = T.depth; l >= 0; l--) T.unrefine_pass(0, root, l, nc, ok);
  if (!ok) {
    t->compiled = false;
    return bmg::fail(BMG_ERR_ARG, "unrefine met an inconsistent tree");
  }
  if (nc) {
    T.dirty = true;
    double* olds[kUserScalars + 2];
    take_mesh_fields(t, olds);
    const long old_nslot = t->ix.nslot;
    std::vector<int> prev;
    t->prev_slots = &prev;
    int rc = compile(t);
    t->prev_slots = nullptr;
    RefineLog none;
    if (rc == BMG_OK)
      rc = t->dim == 2 ? adapt_remap<2>(t, olds, old_nslot, prev, none) : adapt_remap<3>(t, olds, old_nslot, prev, none);
    for (double* p : olds) dfree_(p);
    if (rc != BMG_OK) return rc;
    t->coef_ready = false;
  }
  if (ncoarsened) *ncoarsened = nc;
  return BMG_OK;
}

int bmg_tree_dump(bmg_tree* t, const char* path, const int* fields, const char* const* names, int nfields,
                  double time, int iter) {
  if (!t || !path || (nfields && (!fields || !names)) || nfields < 0)
    return bmg::fail(BMG_ERR_ARG, "null argument");
  TRY(ensure(t));
  for (int m = 0; m < nfields; m++)
    if (!names[m] || !field_ptr(t, fields[m]))
      return bmg::fail(BMG_ERR_ARG, "dump: field %d is not a, b or a user scalar that is set", m);
  // non-leaf cells hold the restriction of their leaves
  if (t->dim == 2)
    adapt_boundary<2>(t);
  else
    adapt_boundary<3>(t);
  const DevIndex& x = t->ix;
  std::vector<std::vector<double>> host(nfields, std::vector<double>(x.nslot));
  for (int m = 0; m < nfields; m++)
    HIPCHECK(hipMemcpyAsync(host[m].data(), field_ptr(t, fields[m]), x.nslot * 8, hipMemcpyDeviceToHost, t->stream));
  HIPCHECK(hipStreamSynchronize(t->stream));
  Topo& T = t->topo;
  std::vector<Cell*> cells;
  dump_preorder(T, 0, {GH, GH, t->dim == 3 ? GH : 0}, cells);
  std::vector<double> size(cells.size());
  if (dump_sizes(cells, t->dim, 0, size) != cells.size()) return bmg::fail(BMG_ERR_ARG, "dump: inconsistent tree");
  std::string tmp = std::string(path) + "~";
  FILE* fp = fopen(tmp.c_str(), "wb");
  if (!fp) return bmg::fail(BMG_ERR_ARG, "dump: cannot write %s", tmp.c_str());
  DumpHeaderC h;
  memset(&h, 0, sizeof h);
  h.t = time;
  h.len = nfields + 1;
  h.i = iter;
  h.depth = T.depth;
  h.npe = 1;
  h.version = kDumpVersion;
  bool ok = fwrite(&h, sizeof h, 1, fp) == 1;
  auto name = [&](const char* nm) {
    unsigned len = (unsigned)strlen(nm);
    ok = ok && fwrite(&len, sizeof len, 1, fp) == 1 && fwrite(nm, 1, len, fp) == len;
  };
  name("size");
  for (int m = 0; m < nfields; m++) name(names[m]);
  const double o[4] = {t->X0[0], t->X0[1], t->X0[2], t->L0};
  ok = ok && fwrite(o, sizeof(double), 4, fp) == 4;
  const size_t rl = 4 + 8 * (size_t)(nfields + 1);
  std::vector<char> buf;
  buf.reserve(rl * 65536);
  for (size_t q = 0; q < cells.size() && ok; q++) {
    const unsigned flags = (cells[q]->flags & C_LEAF) ? kDumpLeaf : 0u;
    const char* fb = reinterpret_cast<const char*>(&flags);
    buf.insert(buf.end(), fb, fb + 4);
    const char* sb = reinterpret_cast<const char*>(&size[q]);
    buf.insert(buf.end(), sb, sb + 8);
    for (int m = 0; m < nfields; m++) {
      const char* vb = reinterpret_cast<const char*>(&host[m][cells[q]->slot]);
      buf.insert(buf.end(), vb, vb + 8);
    }
    if (buf.size() >= rl * 65536 || q + 1 == cells.size()) {
      ok = fwrite(buf.data(), 1, buf.size(), fp) == buf.size();
      buf.clear();
    }
  }
  ok = (fclose(fp) == 0) && ok;
  if (!ok) {
    remove(tmp.c_str());
    return bmg::fail(BMG_ERR_ARG, "dump: write error on %s", tmp.c_str());
  }
  if (rename(tmp.c_str(), path) != 0) return bmg::fail(BMG_ERR_ARG, "dump: cannot rename %s", tmp.c_str());
  return BMG_OK;
}

int bmg_tree_restore(bmg_tree* t, const char* path, const int* fields, const char* const* names, int nfields,
                     double* time, int* iter) {
  if (!t || !path || (nfields && (!fields || !names)) || nfields < 0)
    return bmg::fail(BMG_ERR_ARG, "null argument");
  for (int m = 0; m < nfields; m++)
    if (!names[m] || (fields[m] != BMG_F_A && fields[m] != BMG_F_B &&
                      !(fields[m] >= BMG_F_S && fields[m] < BMG_F_S + kUserScalars)))
      return bmg::fail(BMG_ERR_ARG, "restore: field %d is not a, b or a user scalar", m);
  FILE* fp = fopen(path, "rb");
  if (!fp) return bmg::fail(BMG_ERR_ARG, "restore: cannot open %s", path);
  std::unique_ptr<FILE, int (*)(FILE*)> guard(fp, fclose);
  DumpHeaderC h;
  if (fread(&h, sizeof h, 1, fp) != 1) return bmg::fail(BMG_ERR_ARG, "restore: expecting header");
  if (h.version != kDumpVersion)
    return bmg::fail(BMG_ERR_UNSUPPORTED, "restore: file version %d (reads %d)", h.version, kDumpVersion);
  if (h.len < 1 || h.len > 100000) return bmg::fail(BMG_ERR_ARG, "restore: bad field count %ld", h.len);
  std::vector<std::string> fnames;
  for (long m = 0; m < h.len; m++) {
    unsigned len;
    if (fread(&len, sizeof len, 1, fp) != 1 || len > 4096) return bmg::fail(BMG_ERR_ARG, "restore: expecting len");
    std::string nm(len, '\0');
    if (len && fread(&nm[0], 1, len, fp) != len) return bmg::fail(BMG_ERR_ARG, "restore: expecting a name");
    fnames.push_back(nm);
  }
  double o[4];
  if (fread(o, sizeof(double), 4, fp) != 4) return bmg::fail(BMG_ERR_ARG, "restore: expecting coordinates");
  if (!(o[3] > 0.)) return bmg::fail(BMG_ERR_ARG, "restore: bad box size");
  // column of each requested field (the subtree size, column 0, is skipped;
  // a name the file lacks restores as zero)
  std::vector<int> cols(nfields, -1);
  for (int m = 0; m < nfields; m++)
    for (long q = 1; q < h.len; q++)
      if (fnames[q] == names[m]) cols[m] = (int)q;
  // a new tree from init_grid (1)
  free_device(t);
  t->topo.init(t->dim);
  t->L0 = o[3];
  for (int d = 0; d < 3; d++) t->X0[d] = o[d];
  t->keep_a.clear();
  t->keep_b.clear();
  t->keep_lam.clear();
  for (int k = 0; k < kUserScalars; k++) {
    t->keep_us[k].clear();
    t->us_set[k] = false;
  }
  t->coef_ready = false;
  Topo& T = t->topo;
  std::vector<double> rec(h.len);
  std::vector<std::vector<double>> vals(nfields);
  long nrec = 0;
  int rc = restore_visit(T, 0, {GH, GH, t->dim == 3 ? GH : 0}, fp, (size_t)h.len, rec, nrec, cols, vals);
  if (rc != BMG_OK) {
    t->topo.init(t->dim);
    return rc;
  }
  char extra;
  if (fread(&extra, 1, 1, fp) == 1) {
    t->topo.init(t->dim);
    return bmg::fail(BMG_ERR_ARG, "restore: the file holds more records than the tree");
  }
  T.dirty = true;
  T.update();
  std::vector<int> recs(T.leaf_c.size());
  for (size_t n = 0; n < T.leaf_c.size(); n++) {
    const IJK c = T.leaf_c[n];
    recs[n] = T.at(T.leaf_level[n], c.i, c.j, c.k).slot;
  }
  TRY(ensure(t));
  std::vector<double> leafv(recs.size());
  for (int m = 0; m < nfields; m++) {
    for (size_t n = 0; n < recs.size(); n++) leafv[n] = vals[m][recs[n]];
    TRY(bmg_tree_set_leaf_field(t, fields[m], leafv.data(), 0));
  }
  if (time) *time = h.t;
  if (iter) *iter = h.i;
  return BMG_OK;
}

int bmg_tree_cache(bmg_tree* t, int which, int level, int* n, int* ijk) {
  if (!t || !n) return bmg::fail(BMG_ERR_ARG, "null argument");
  Topo& T = t->topo;
  T.update();
  if (level < 0 || level > T.depth || which < 0 || which > 2) return bmg::fail(BMG_ERR_ARG, "bad cache request");
  auto& c = which == 0 ? T.active[level] : which == 1 ? T.prolong[level] : T.bnd[level];
  if (ijk)
    for (size_t q = 0; q < c.size(); q++) {
      ijk[3 * q] = c[q].i;
      ijk[3 * q + 1] = c[q].j;
      ijk[3 * q + 2] = c[q].k;
    }
  *n = (int)c.size();
  return BMG_OK;
}

int bmg_tree_faces(bmg_tree* t, long* n, int* out) {
  if (!t || !n) return bmg::fail(BMG_ERR_ARG, "null argument");
  Topo& T = t->topo;
  T.update();
  if (out)
    for (size_t q = 0; q < T.faces.size(); q++) {
      out[5 * q] = T.faces[q].level;
      out[5 * q + 1] = T.faces[q].c.i;
      out[5 * q + 2] = T.faces[q].c.j;
      out[5 * q + 3] = T.faces[q].c.k;
      out[5 * q + 4] = T.faces[q].flags;
    }
  *n = (long)T.faces.size();
  return BMG_OK;
}

int bmg_tree_set_precision(bmg_tree* t, int precision) {
  if (!t || (precision != BMG_FP64 && precision != BMG_MIXED)) return bmg::fail(BMG_ERR_ARG, "bad precision");
  if (precision != t->precision) {
    // keep the leaf values of a, b and lambda across the reallocation
    if (t->compiled) {
      t->keep_a.resize(t->ix.nleaf);
      t->keep_b.resize(t->ix.nleaf);
      TRY(get_leaf_values<double>(t, t->a, t->keep_a.data(), 0));
      TRY(get_leaf_values<double>(t, t->b, t->keep_b.data(), 0));
      for (int k = 0; k < kUserScalars; k++)
        if (t->us_set[k]) {
          t->keep_us[k].resize(t->ix.nleaf);
          TRY(get_leaf_values<double>(t, t->us[k], t->keep_us[k].data(), 0));
        }
      if (t->lambda_field) {
        t->keep_lam.resize(t->ix.nleaf);
        TRY(get_leaf_values<double>(t, t->lam64, t->keep_lam.data(), 0));
      }
      if (t->alpha_field) return bmg::fail(BMG_ERR_ARG, "set the precision before alpha");
      free_device(t);
    }
    t->precision = precision;
  }
  return BMG_OK;
}

int bmg_tree_set_smoother(bmg_tree* t, int smoother) {
  if (!t || smoother < BMG_RELAX_GS || smoother > BMG_RELAX_RBGS) return bmg::fail(BMG_ERR_ARG, "bad smoother");
  if (smoother == BMG_RELAX_GS)
    return bmg::fail(BMG_ERR_UNSUPPORTED, "the in-place Gauss-Seidel smoother is sequential; use JACOBI or RBGS");
  t->smoother = smoother;
  drop_graphs(t);
  return BMG_OK;
}

int bmg_tree_set_tolerance(bmg_tree* t, double tol) {
  if (!t || !(tol >= 0.)) return bmg::fail(BMG_ERR_ARG, "bad tolerance");
  t->TOLERANCE = tol;
  return BMG_OK;
}

int bmg_tree_set_niter(bmg_tree* t, int nitermax, int nitermin) {
  if (!t || nitermax < 0 || nitermin < 0) return bmg::fail(BMG_ERR_ARG, "bad iteration limits");
  t->NITERMAX = nitermax;
  t->NITERMIN = nitermin;
  return BMG_OK;
}

int bmg_tree_set_bc(bmg_tree* t, int side, int kind, const double* values) {
  if (!t || side < 0 || side >= 2 * t->dim) return bmg::fail(BMG_ERR_ARG, "bad side");
  if (kind != BMG_BC_SYMMETRY && kind != BMG_BC_DIRICHLET && kind != BMG_BC_NEUMANN)
    return bmg::fail(BMG_ERR_UNSUPPORTED, "trees take symmetry, Dirichlet or Neumann sides");
  Topo& T = t->topo;
  T.update();
  long m = 0;
  for (int l = 0; l <= T.depth; l++) {
    long p = 1;
    for (int e = 1; e < t->dim; e++) p *= (1L << l);
    m += p;
  }
  t->bc_kind[side] = kind;
  drop_graphs(t);
  t->bc_host[side].assign(m, 0.);
  if (values) std::copy(values, values + m, t->bc_host[side].begin());
  if (t->compiled) {
    if ((long)t->ix.bc_off[t->ix.depth + 1] != m) return bmg::fail(BMG_ERR_ARG, "tree changed");
    HIPCHECK(hipMemcpy(t->bcv[side], t->bc_host[side].data(), sizeof(double) * m, hipMemcpyHostToDevice));
  }
  return BMG_OK;
}

int bmg_tree_set_leaf_field(bmg_tree* t, int field, const double* values, int on_device) {
  if (!t || !values) return bmg::fail(BMG_ERR_ARG, "null argument");
  TRY(ensure(t));
  if (field == BMG_F_A) return set_leaf_values(t, t->a, values, on_device);
  if (field == BMG_F_B) return set_leaf_values(t, t->b, values, on_device);
  if (field >= BMG_F_S && field < BMG_F_S + kUserScalars) {
    const int k = field - BMG_F_S;
    if (!t->us[k]) TRY(dalloc_((void**)&t->us[k], std::max<long>(1, t->ix.nslot) * 8));
    t->us_set[k] = true;
    return set_leaf_values(t, t->us[k], values, on_device);
  }
  if (field == BMG_F_LAMBDA) {
    t->lambda_field = true;
    t->coef_ready = false;
    drop_graphs(t);
    return set_leaf_values(t, t->lam64, values, on_device);
  }
  return bmg::fail(BMG_ERR_ARG, "field %d cannot be set on a tree", field);
}

int bmg_tree_get_leaf_field(bmg_tree* t, int field, double* values, int on_device) {
  if (!t || !values) return bmg::fail(BMG_ERR_ARG, "null argument");
  TRY(ensure(t));
  bool mixed = t->precision == BMG_MIXED;
  if (field >= BMG_F_S && field < BMG_F_S + kUserScalars) {
    if (!t->us[field - BMG_F_S]) return bmg::fail(BMG_ERR_ARG, "user scalar %d is not set", field - BMG_F_S);
    return get_leaf_values<double>(t, t->us[field - BMG_F_S], values, on_device);
  }
  switch (field) {
    case BMG_F_A: return get_leaf_values<double>(t, t->a, values, on_device);
    case BMG_F_B: return get_leaf_values<double>(t, t->b, values, on_device);
    case BMG_F_RES: return get_leaf_values<double>(t, t->res64, values, on_device);
    case BMG_F_DA:
      return mixed ? get_leaf_values<float>(t, F<float>(t->da), values, on_device)
                   : get_leaf_values<double>(t, F<double>(t->da), values, on_device);
    case BMG_F_LAMBDA: return get_leaf_values<double>(t, t->lam64, values, on_device);
  }
  return bmg::fail(BMG_ERR_ARG, "field %d cannot be read on a tree", field);
}

}  // extern "C"

namespace {

// project() (poisson.h:469-509) on the tree: vals are the face velocity's
// values on the face cache (device memory), updated in place
template <int DIM>
int tree_project_t(bmg_tree* t, double* vals, double dt, int nrelax, bmg_mgstats* stats) {
  const DevIndex& x = t->ix;
  for (int d = 0; d < DIM; d++)
    if (!t->uf[d]) TRY(dalloc_((void**)&t->uf[d], std::max<long>(1, x.nslot) * sizeof(double)));
  double* u1 = DIM > 1 ? t->uf[1] : nullptr;
  double* u2 = DIM > 2 ? t->uf[2] : nullptr;
  if (x.nfaceval)
    k_t_face_set<<<nb_(x.nfaceval), kB, 0, t->stream>>>(x.fv_gs, x.fv_d, x.nfaceval, vals, t->uf[0], u1, u2);
  // the divergence's stencil reads uf.x[1]: halo_face first (the fine
  // fluxes averaged onto the coarse faces of resolution jumps)
  if (x.nhf) k_t_face_avg<DIM><<<nb_(x.nhf), kB, 0, t->stream>>>(x.hf, x.nhf, t->uf[0], u1, u2);
  Segs S = all_leaves(t);
  if (S.first_block[S.n])
    k_t_leaf_div<DIM><<<(unsigned)S.first_block[S.n], kB, 0, t->stream>>>(S, t->uf[0], u1, u2, x.nbr, dt, t->b);
  HIPCHECK(hipGetLastError());
  const int rc = bmg_tree_poisson(t, t->TOLERANCE / (dt * dt), nrelax, 0, stats);
  if (rc != BMG_OK && rc != BMG_ERR_NAN) return rc;
  // the face update's stencil reads p[-1]: the full tree boundary of p
  {
    Batcher<double, DIM> B(t, t->a, nullptr, nullptr);
    B.tree_boundary(0, BK_A, -1);
    B.flush();
  }
  const double* al[3] = {nullptr, nullptr, nullptr};
  if (t->alpha_field)
    for (int d = 0; d < DIM; d++) al[d] = t->al64[d];
  for (int l = 0; l <= x.depth; l++) {
    const int n = x.face_n[l];
    if (n)
      k_t_face_update<<<nb_(n), kB, 0, t->stream>>>(x.fops + x.face_off[l], n, t->a, t->uf[0], u1, u2, al[0], al[1],
                                                    al[2], t->alpha_c, dt, delta(t, l));
  }
  if (x.nfaceval)
    k_t_face_get<<<nb_(x.nfaceval), kB, 0, t->stream>>>(x.fv_gs, x.fv_d, x.nfaceval, vals, t->uf[0], u1, u2);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(t->stream));
  return rc;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// --- the adjoint of poisson() on the tree --------------------------------

template <typename T>
int download(std::vector<T>& v, const T* dev, size_t n) {
  v.resize(n);
  if (n) HIPCHECK(hipMemcpy(v.data(), dev, n * sizeof(T), hipMemcpyDeviceToHost));
  return BMG_OK;
}

// counting sort of (key, value) pairs by key, stable: CSR offsets over
// nkeys and the values in their original order per key
template <typename V>
void csr(int nkeys, const std::vector<int>& key, const std::vector<V>& val, std::vector<int>& off,
         std::vector<V>& out) {
  off.assign(nkeys + 1, 0);
  for (int k : key) off[k + 1]++;
  for (int i = 0; i < nkeys; i++) off[i + 1] += off[i];
  out.resize(val.size());
  std::vector<int> cur(off.begin(), off.end() - 1);
  for (size_t i = 0; i < key.size(); i++) out[cur[key[i]]++] = val[i];
}

int nslot_of(const DevIndex& x, int l) { return (int)((l < x.depth ? x.off[l + 1] : x.nslot) - x.off[l]); }

AdjLv adj_levels(const bmg_tree* t) {
  const DevIndex& x = t->ix;
  AdjLv L;
  memset(&L, 0, sizeof L);
  L.nl = x.depth + 1;
  for (int l = 0; l <= x.depth; l++) {
    L.off[l] = x.off[l];
    L.nslot[l] = nslot_of(x, l);
    L.nA[l] = x.nA[l];
    L.abase[l] = x.abase[l];
    L.D[l] = t->L0 * (1. / (1 << l));
  }
  return L;
}

// the inverse tables of the transposed stages, from the compiled tree
int adj_build(bmg_tree* t) {
  auto& A = t->adj;
  if (A.built) return BMG_OK;
  const DevIndex& x = t->ix;
  const int dim = t->dim, D = x.depth, NR = dim == 3 ? 27 : 9, NCH = 1 << dim;
  if (D + 1 > kAdjLevels) return bmg::fail(BMG_ERR_UNSUPPORTED, "the adjoint handles trees of depth < %d", kAdjLevels);
  if (x.nslot >= (1L << 30)) return bmg::fail(BMG_ERR_UNSUPPORTED, "too many slots for the adjoint tables");
  const int NS = (int)x.nslot;
  // face ops: which op has a slot as gs / gm, per dimension
  std::vector<FaceOp> F;
  TRY(download(F, x.fops, (size_t)x.nfaceval));
  std::vector<int> fgs((size_t)NS * dim, -1), fgm((size_t)NS * dim, -1);
  for (size_t o = 0; o < F.size(); o++) {
    int& a = fgs[(size_t)F[o].gs * dim + F[o].d];
    int& b = fgm[(size_t)F[o].gm * dim + F[o].d];
    if (a >= 0 || b >= 0) return bmg::fail(BMG_ERR_ARG, "internal: a slot is the face of two ops");
    a = b = (int)o;
  }
  // box rules by source (rule order)
  const int nrules = x.box_off[D] + x.box_n[D];
  std::vector<BoxRule> R;
  TRY(download(R, x.box, (size_t)nrules));
  std::vector<int> rk, rv, roff, rinv;
  for (int r = 0; r < nrules; r++)
    if (R[r].nterm) {
      rk.push_back(R[r].src);
      rv.push_back(r);
    }
  csr(NS, rk, rv, roff, rinv);
  // bilinear terms by ring slot (record, child, term order)
  const int nrec = x.rec_off[D] + x.n_nonleaf[D] + x.n_prolong[D];
  std::vector<int> ring;
  TRY(download(ring, x.rec_ring, (size_t)nrec * NR));
  std::vector<int> pk;
  std::vector<bmgt::ProEnt> pv;
  pk.reserve((size_t)nrec * NCH * NCH);
  pv.reserve((size_t)nrec * NCH * NCH);
  for (int r = 0; r < nrec; r++)
    for (int m = 0; m < NCH; m++) {
      int a = dim == 3 ? m >> 2 : m >> 1, b = dim == 3 ? (m >> 1) & 1 : m & 1, e = dim == 3 ? m & 1 : 0;
      int cx = 2 * a - 1, cy = 2 * b - 1, cz = 2 * e - 1;
      static const int T3[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
      static const int T2[4][2] = {{0, 0}, {1, 0}, {0, 1}, {1, 1}};
      for (int k = 0; k < NCH; k++) {
        int pos;
        if (dim == 3)
          pos = ((1 + T3[k][0] * cx) * 3 + 1 + T3[k][1] * cy) * 3 + 1 + T3[k][2] * cz;
        else
          pos = (1 + T2[k][0] * cx) * 3 + 1 + T2[k][1] * cy;
        const int slot = ring[(size_t)r * NR + pos];
        if (slot < 0 || slot >= NS) return bmg::fail(BMG_ERR_ARG, "internal: ring slot out of range");
        pk.push_back(slot);
        pv.push_back(bmgt::ProEnt{r, (unsigned char)m, (unsigned char)k, 0, 0});
      }
    }
  std::vector<int> poff;
  std::vector<bmgt::ProEnt> pinv;
  csr(NS, pk, pv, poff, pinv);
  // neighbour reads of relax by receiving slot (active order)
  const int nact = x.abase[D] + x.nA[D];
  if ((long)nact * 8 >= (1L << 31)) return bmg::fail(BMG_ERR_UNSUPPORTED, "too many cells for the adjoint tables");
  std::vector<int> nb;
  TRY(download(nb, x.nbr, (size_t)nact * 2 * dim));
  std::vector<int> nk, nv, noff, ninv;
  nk.reserve(nb.size());
  nv.reserve(nb.size());
  for (int ai = 0; ai < nact; ai++)
    for (int dir = 0; dir < 2 * dim; dir++) {
      nk.push_back(nb[(size_t)ai * 2 * dim + dir]);
      nv.push_back(ai * 8 + dir);
    }
  csr(NS, nk, nv, noff, ninv);
  TRY(upload(&A.fop_gs, fgs));
  TRY(upload(&A.fop_gm, fgm));
  TRY(upload(&A.box_off, roff));
  TRY(upload(&A.box_inv, rinv));
  TRY(upload(&A.pro_off, poff));
  TRY(upload(&A.pro_inv, pinv));
  TRY(upload(&A.nbr_off, noff));
  TRY(upload(&A.nbr_inv, ninv));
  TRY(dalloc_((void**)&A.slot_leaf, std::max(1, NS)));
  HIPCHECK(hipMemset(A.slot_leaf, 0, std::max(1, NS)));
  if (x.nleaf) k_ta_leaf_mark<<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, A.slot_leaf);
  for (double** p : {&A.xb, &A.gb[0], &A.gb[1], &A.gb[2], &A.db, &A.rb})
    TRY(dalloc_((void**)p, std::max(1, NS) * sizeof(double)));
  for (double** p : {&A.lv0, &A.lv1, &A.lv2, &A.lv3, &A.lv4})
    TRY(dalloc_((void**)p, std::max<long>(1, x.nleaf) * sizeof(double)));
  HIPCHECK(hipStreamSynchronize(t->stream));
  A.built = true;
  return BMG_OK;
}

template <int DIM>
struct AdjOps {
  bmg_tree* t;
  const DevIndex& x;
  AdjLv L;
  const double* al[3] = {nullptr, nullptr, nullptr};
  const double* lam;
  BCT bc;
  explicit AdjOps(bmg_tree* t_) : t(t_), x(t_->ix), L(adj_levels(t_)), lam(t_->lambda_field ? t_->lam64 : nullptr),
                                  bc(bct(t_)) {
    if (t->alpha_field)
      for (int d = 0; d < DIM; d++) al[d] = t->al64[d];
  }
  void box_T(double* f, int l, int sym) {
    const int n = nslot_of(x, l);
    if (n) k_ta_box_T<<<nb_(n), kB, 0, t->stream>>>(x.off[l], n, t->adj.box_off, t->adj.box_inv, x.box, bc, sym, f);
    if (x.box_n[l]) k_ta_box_zero<<<nb_(x.box_n[l]), kB, 0, t->stream>>>(x.box + x.box_off[l], x.box_n[l], f);
  }
  // bilinear transposed for the records [r0, r0 + nr) of level l
  void bilinear_T(double* f, int l, int r0, int nr) {
    if (!nr) return;
    const int n = nslot_of(x, l);
    k_ta_bilinear_T<DIM><<<nb_(n), kB, 0, t->stream>>>(x.off[l], n, t->adj.pro_off, t->adj.pro_inv, x.rec_child0,
                                                       r0, r0 + nr, f);
    k_ta_children_zero<DIM><<<nb_(nr), kB, 0, t->stream>>>(x.rec_child0, r0, nr, f);
  }
  void restrict_T(double* f, int m) {
    const int n = x.n_nonleaf[m];
    if (n)
      k_ta_restrict_T<DIM><<<nb_(n), kB, 0, t->stream>>>(x.rec_self + x.rec_off[m], x.rec_child0 + x.rec_off[m], n,
                                                         f);
  }
  // tree_boundary_level (f, lev) transposed (lev < 0: depth)
  void boundary_T(double* f, int lev, int sym) {
    const int depth = lev < 0 ? x.depth : lev;
    if (t->topo.full()) {
      box_T(f, depth, sym);
      return;
    }
    for (int i = depth - 1; i >= 0; i--) {
      box_T(f, i + 1, sym);
      bilinear_T(f, i, x.rec_off[i] + x.n_nonleaf[i], x.n_prolong[i]);
    }
    box_T(f, 0, sym);
    for (int m = 0; m < depth; m++) restrict_T(f, m);
  }
  // y (leaf order) -> gb (residual_T and halo_face_T), xb[leaves] = -lambda y
  void faces_of(const double* y) {
    auto& A = t->adj;
    const size_t bytes = (size_t)std::max<long>(1, x.nslot) * sizeof(double);
    (void)hipMemsetAsync(A.xb, 0, bytes, t->stream);
    for (int d = 0; d < 3; d++) (void)hipMemsetAsync(A.gb[d], 0, bytes, t->stream);
    if (x.nleaf)
      k_ta_div_T<DIM><<<nb_(x.nleaf), kB, 0, t->stream>>>(L, x.leaf_gs, x.nleaf, y, x.nbr, lam, t->lambda_c, A.xb,
                                                          A.gb[0], A.gb[1], A.gb[2]);
    if (x.nhf) k_ta_halo_face_T<DIM><<<nb_(x.nhf), kB, 0, t->stream>>>(x.hf, x.nhf, A.gb[0], A.gb[1], A.gb[2]);
  }
  // out = r^T y (leaf order, device)
  void residual_T(const double* y, double* out) {
    auto& A = t->adj;
    faces_of(y);
    if (x.nslot)
      k_ta_faces_T<DIM><<<nb_(x.nslot), kB, 0, t->stream>>>(L, (int)x.nslot, x.fops, A.fop_gs, A.fop_gm, al[0], al[1],
                                                            al[2], t->alpha_c, A.gb[0], A.gb[1], A.gb[2], A.xb);
    boundary_T(A.xb, -1, 0);
    if (x.nleaf) k_ta_leaf_get<<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, A.xb, out);
  }
  // one red-black pass of relax (level l, colour) transposed
  void relax_T(int l, int col) {
    auto& A = t->adj;
    const int nrecv = x.off[l] + nslot_of(x, l);
    if (nrecv)
      k_ta_relax_T<DIM><<<nb_(nrecv), kB, 0, t->stream>>>(L, nrecv, A.nbr_off, A.nbr_inv, x.nbr, x.colour,
                                                          A.slot_leaf, l, col, al[0], al[1], al[2], t->alpha_c, lam,
                                                          t->lambda_c, A.db);
    Segs S = level_or_leaf_colour(t, l, col);
    if (S.first_block[S.n])
      k_ta_relax_cells_T<DIM><<<(unsigned)S.first_block[S.n], kB, 0, t->stream>>>(S, x.nbr, al[0], al[1], al[2],
                                                                                  t->alpha_c, lam, t->lambda_c,
                                                                                  A.db, A.rb);
  }
  // out = M^T y: the transposed V-cycle (red-black relax) from da = 0
  void cycle_T(const double* y, double* out, int nrelax, int minlevel, int maxlevel) {
    auto& A = t->adj;
    const size_t bytes = (size_t)std::max<long>(1, x.nslot) * sizeof(double);
    (void)hipMemsetAsync(A.db, 0, bytes, t->stream);
    (void)hipMemsetAsync(A.rb, 0, bytes, t->stream);
    if (x.nleaf) k_ta_leaf_set<<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, y, A.db);
    minlevel = std::min(minlevel, maxlevel);
    for (int l = maxlevel; l >= minlevel; l--) {
      for (int it = 0; it < nrelax; it++) {
        boundary_T(A.db, l, 0);
        relax_T(l, 1);
        for (int m = l; m >= 0; m--) box_T(A.db, m, 0);
        relax_T(l, 0);
      }
      boundary_T(A.db, l, 0);
      if (l == minlevel) {
        Segs S = level_or_leaf(t, l);
        if (S.first_block[S.n]) k_t_fill<double><<<(unsigned)S.first_block[S.n], kB, 0, t->stream>>>(S, A.db, 0.);
      } else {
        bilinear_T(A.db, l - 1, x.rec_off[l - 1], x.n_nonleaf[l - 1]);
      }
    }
    if (t->topo.full()) {
      for (int l = 0; l < x.depth; l++) {
        box_T(A.rb, l, 1);
        restrict_T(A.rb, l);
      }
      box_T(A.rb, x.depth, 1);
    } else {
      boundary_T(A.rb, -1, 1);
    }
    if (x.nleaf) k_ta_leaf_get<<<nb_(x.nleaf), kB, 0, t->stream>>>(x.leaf_gs, x.nleaf, A.rb, out);
  }
};

template <int DIM>
int tree_adjoint_t(bmg_tree* t, const double* g, double tolerance, int nitermax, int nrelax, double* psi,
                   double* galpha, double* glambda, bmg_mgstats* st) {
  const DevIndex& x = t->ix;
  auto& A = t->adj;
  AdjOps<DIM> O(t);
  const long N = x.nleaf;
  const unsigned nbl = nb_(std::max<long>(1, N));
  double *rho = A.lv0, *tmp = A.lv1;
  HIPCHECK(hipMemcpyAsync(rho, g, N * sizeof(double), hipMemcpyDeviceToDevice, t->stream));
  HIPCHECK(hipMemsetAsync(psi, 0, N * sizeof(double), t->stream));
  // resb = max|g|
  k_t_stat_reset<<<1, 64, 0, t->stream>>>(t->stat);
  HIPCHECK(hipMemsetAsync(tmp, 0, N * sizeof(double), t->stream));
  k_ta_update<<<nbl, kB, 0, t->stream>>>(N, psi, tmp, g, tmp, rho, t->stat, 1);
  TStat ts;
  HIPCHECK(hipMemcpyAsync(&ts, t->stat, sizeof ts, hipMemcpyDeviceToHost, t->stream));
  HIPCHECK(hipStreamSynchronize(t->stream));
  double resb, resa;
  memcpy(&resb, &ts.maxbits, sizeof resb);
  resa = resb;
  int it;
  for (it = 0; it < nitermax && (it < 1 || resa > tolerance); it++) {
    O.cycle_T(rho, tmp, nrelax, 1, x.depth);
    k_ta_update<<<nbl, kB, 0, t->stream>>>(N, psi, tmp, g, tmp, rho, t->stat, 0);
    O.residual_T(psi, tmp);
    k_t_stat_reset<<<1, 64, 0, t->stream>>>(t->stat);
    k_ta_update<<<nbl, kB, 0, t->stream>>>(N, psi, tmp, g, tmp, rho, t->stat, 1);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(&ts, t->stat, sizeof ts, hipMemcpyDeviceToHost, t->stream));
    HIPCHECK(hipStreamSynchronize(t->stream));
    if (ts.nans) return bmg::fail(BMG_ERR_NAN, "adjoint: NaN in the residual");
    memcpy(&resa, &ts.maxbits, sizeof resa);
  }
  if (galpha && x.nfaceval) {
    {
      Batcher<double, DIM> B(t, t->a, nullptr, nullptr);
      B.tree_boundary(0, BK_A, -1);
      B.flush();
    }
    O.faces_of(psi);
    k_ta_galpha<DIM><<<nb_(x.nfaceval), kB, 0, t->stream>>>(O.L, x.nfaceval, x.fv_gs, x.fv_d, A.fop_gs, x.fops, t->a,
                                                            A.gb[0], A.gb[1], A.gb[2], galpha);
  }
  if (glambda && N) k_ta_glambda<<<nb_(N), kB, 0, t->stream>>>(x.leaf_gs, N, psi, t->a, glambda);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(t->stream));
  if (st) {
    memset(st, 0, sizeof *st);
    st->i = it;
    st->resb = resb;
    st->resa = resa;
    st->nrelax = nrelax;
    st->minlevel = 1;
  }
  return BMG_OK;
}

// a leaf-order (or face-value) array in device memory: the pointer itself
// when on_device, else a device copy (copied back by give)
struct LeafBuf {
  double* dev = nullptr;
  double* host = nullptr;
  bool owned = false;
  long n = 0;
  int take(double* p, long n_, int on_device, bool copy_in) {
    n = n_;
    if (!p) return BMG_OK;
    if (on_device) {
      dev = p;
      return BMG_OK;
    }
    host = p;
    owned = true;
    TRY(dalloc_((void**)&dev, std::max<long>(1, n) * sizeof(double)));
    if (copy_in && n) HIPCHECK(hipMemcpy(dev, p, n * sizeof(double), hipMemcpyHostToDevice));
    return BMG_OK;
  }
  int give() {
    if (owned && host && n) HIPCHECK(hipMemcpy(host, dev, n * sizeof(double), hipMemcpyDeviceToHost));
    return BMG_OK;
  }
  ~LeafBuf() {
    if (owned) dfree_(dev);
  }
};

int adj_ready(bmg_tree* t) {
  TRY(ensure(t));
  TRY(coefs(t));
  return adj_build(t);
}

}  // namespace

extern "C" {

int bmg_tree_residual_transpose(bmg_tree* t, const double* y, double* out, int on_device) {
  if (!t || !y || !out) return bmg::fail(BMG_ERR_ARG, "null argument");
  TRY(adj_ready(t));
  LeafBuf Y, O;
  TRY(Y.take((double*)y, t->ix.nleaf, on_device, true));
  TRY(O.take(out, t->ix.nleaf, on_device, false));
  if (t->dim == 3) AdjOps<3>(t).residual_T(Y.dev, O.dev);
  else AdjOps<2>(t).residual_T(Y.dev, O.dev);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(t->stream));
  return O.give();
}

int bmg_tree_cycle_transpose(bmg_tree* t, const double* y, double* out, int nrelax, int minlevel, int on_device) {
  if (!t || !y || !out || nrelax < 0 || minlevel < 0) return bmg::fail(BMG_ERR_ARG, "bad argument");
  TRY(adj_ready(t));
  LeafBuf Y, O;
  TRY(Y.take((double*)y, t->ix.nleaf, on_device, true));
  TRY(O.take(out, t->ix.nleaf, on_device, false));
  if (t->dim == 3) AdjOps<3>(t).cycle_T(Y.dev, O.dev, nrelax, minlevel, t->ix.depth);
  else AdjOps<2>(t).cycle_T(Y.dev, O.dev, nrelax, minlevel, t->ix.depth);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipStreamSynchronize(t->stream));
  return O.give();
}

int bmg_tree_adjoint(bmg_tree* t, const double* g, double tolerance, int nitermax, int nrelax, double* psi,
                     double* galpha, double* glambda, int on_device, bmg_mgstats* stats) {
  if (!t || !g || !psi) return bmg::fail(BMG_ERR_ARG, "null argument");
  if (nitermax < 0) return bmg::fail(BMG_ERR_ARG, "nitermax must be non-negative");
  if (nrelax <= 0) nrelax = 4;
  TRY(adj_ready(t));
  const DevIndex& x = t->ix;
  LeafBuf G, P, GA, GL;
  TRY(G.take((double*)g, x.nleaf, on_device, true));
  TRY(P.take(psi, x.nleaf, on_device, false));
  TRY(GA.take(galpha, x.nfaceval, on_device, false));
  TRY(GL.take(glambda, x.nleaf, on_device, false));
  const int rc = t->dim == 3 ? tree_adjoint_t<3>(t, G.dev, tolerance, nitermax, nrelax, P.dev, GA.dev, GL.dev, stats)
                             : tree_adjoint_t<2>(t, G.dev, tolerance, nitermax, nrelax, P.dev, GA.dev, GL.dev, stats);
  if (rc != BMG_OK) return rc;
  TRY(P.give());
  TRY(GA.give());
  return GL.give();
}

int bmg_tree_project(bmg_tree* t, double* uf, double dt, int nrelax, int on_device, bmg_mgstats* stats) {
  if (!t || !uf) return bmg::fail(BMG_ERR_ARG, "null argument");
  if (dt == 0.) dt = 1.;
  if (nrelax == 0) nrelax = 4;
  if (!(dt > 0.) || nrelax < 0) return bmg::fail(BMG_ERR_ARG, "dt must be positive and nrelax non-negative");
  TRY(ensure(t));
  const DevIndex& x = t->ix;
  double* vals = uf;
  double* tmp = nullptr;
  if (!on_device) {
    TRY(dalloc_((void**)&tmp, std::max<long>(1, x.nfaceval) * 8));
    HIPCHECK(hipMemcpy(tmp, uf, x.nfaceval * 8, hipMemcpyHostToDevice));
    vals = tmp;
  }
  const int rc = t->dim == 3 ? tree_project_t<3>(t, vals, dt, nrelax, stats)
                             : tree_project_t<2>(t, vals, dt, nrelax, stats);
  if (!on_device) {
    if ((rc == BMG_OK || rc == BMG_ERR_NAN) && hipMemcpy(uf, tmp, x.nfaceval * 8, hipMemcpyDeviceToHost) != hipSuccess) {
      dfree_(tmp);
      return bmg::fail(BMG_ERR_HIP, "copy of the face velocity failed");
    }
    dfree_(tmp);
  }
  return rc;
}

int bmg_tree_set_alpha_faces(bmg_tree* t, const double* values, int on_device) {
  if (!t || !values) return bmg::fail(BMG_ERR_ARG, "null argument");
  TRY(ensure(t));
  const DevIndex& x = t->ix;
  const double* src = values;
  double* tmp = nullptr;
  if (!on_device) {
    TRY(dalloc_((void**)&tmp, std::max<long>(1, x.nfaceval) * 8));
    HIPCHECK(hipMemcpy(tmp, values, x.nfaceval * 8, hipMemcpyHostToDevice));
    src = tmp;
  }
  if (x.nfaceval)
    k_t_face_set<<<nb_(x.nfaceval), kB, 0, t->stream>>>(x.fv_gs, x.fv_d, x.nfaceval, src, t->al64[0], t->al64[1],
                                                        t->al64[2]);
  HIPCHECK(hipStreamSynchronize(t->stream));
  dfree_(tmp);
  t->alpha_field = true;
  t->coef_ready = false;
  drop_graphs(t);
  return BMG_OK;
}

int bmg_tree_set_alpha_const(bmg_tree* t, double alpha) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  t->alpha_field = false;
  t->alpha_c = alpha;
  t->coef_ready = false;
  drop_graphs(t);
  return BMG_OK;
}

int bmg_tree_set_lambda_const(bmg_tree* t, double lambda) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  t->lambda_field = false;
  t->lambda_c = lambda;
  t->coef_ready = false;
  drop_graphs(t);
  return BMG_OK;
}

int bmg_tree_restriction_alpha_lambda(bmg_tree* t) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  TRY(ensure(t));
  t->coef_ready = false;
  TRY(coefs(t));
  HIPCHECK(hipStreamSynchronize(t->stream));
  return BMG_OK;
}

int bmg_tree_residual(bmg_tree* t, double* maxres) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  TRY(ensure(t));
  double m = 0.;
  int rc = resid(t, &m);
  if (maxres) *maxres = m;
  return rc;
}

int bmg_tree_relax(bmg_tree* t, int level) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  TRY(ensure(t));
  if (level < 0 || level > t->ix.depth) return bmg::fail(BMG_ERR_ARG, "bad level %d", level);
  TRY(coefs(t));
  int rc = t->dim == 3 ? (t->precision == BMG_MIXED ? relax<float, 3>(t, level) : relax<double, 3>(t, level))
                       : (t->precision == BMG_MIXED ? relax<float, 2>(t, level) : relax<double, 2>(t, level));
  TRY(rc);
  HIPCHECK(hipStreamSynchronize(t->stream));
  HIPCHECK(hipGetLastError());
  return BMG_OK;
}

int bmg_tree_mg_cycle(bmg_tree* t, int nrelax, int minlevel, int maxlevel) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  TRY(ensure(t));
  if (minlevel < 0 || maxlevel < 0 || maxlevel > t->ix.depth) return bmg::fail(BMG_ERR_ARG, "bad levels");
  TRY(cycle(t, nrelax, minlevel, maxlevel));
  HIPCHECK(hipStreamSynchronize(t->stream));
  return BMG_OK;
}

int bmg_tree_mg_solve(bmg_tree* t, double tolerance, int nrelax, int minlevel, bmg_mgstats* stats) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  TRY(ensure(t));
  if (minlevel < 0) return bmg::fail(BMG_ERR_ARG, "bad minlevel");
  return mg_solve_impl(t, tolerance, nrelax, minlevel, stats);
}

int bmg_tree_poisson(bmg_tree* t, double tolerance, int nrelax, int minlevel, bmg_mgstats* stats) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  TRY(ensure(t));
  t->coef_ready = false;
  TRY(coefs(t));
  double defaultol = t->TOLERANCE;
  if (tolerance) t->TOLERANCE = tolerance;
  int rc = mg_solve_impl(t, 0., nrelax, std::max(1, minlevel), stats);
  if (tolerance) t->TOLERANCE = defaultol;
  return rc;
}

int bmg_tree_synchronize(bmg_tree* t) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  HIPCHECK(hipStreamSynchronize(t->stream));
  return BMG_OK;
}

int bmg_tree_profile_enable(bmg_tree* t, int enable) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  tprof_collect(t);
  t->prof = enable != 0;
  t->prof_ms = 0.;
  t->prof_n = t->prof_cells = 0;
  return BMG_OK;
}

int bmg_tree_profile_get(bmg_tree* t, double* total_ms, long* launches, long* cells) {
  if (!t) return bmg::fail(BMG_ERR_ARG, "null tree");
  tprof_collect(t);
  if (total_ms) *total_ms = t->prof_ms;
  if (launches) *launches = t->prof_n;
  if (cells) *cells = t->prof_cells;
  return BMG_OK;
}

}  // extern "C"
