// Grid tree, leaf-cell and neighbour/ghost topology on the device.
//
// Reference behaviour restated here (all integer/bit-exact):
//   grid/generation.hpp:39-60      MinLevel: breadth-first refinement, children
//                                  appended at node_end() (implementation.hpp:320-364)
//   implementation.hpp:457-543     find_samelvl_neighbor (up/down tree walk)
//   grid/grid.hpp:344-374          cell_length_at_level / cell_coordinates
//   grid/grid.hpp:601-626          is_cut_by (vertex signed distances)
//   solver.hpp:876-930             create_local_cells (leaf order, initialDomain)
//   solver.hpp:935-977, 843-861    create_ghost_cells / create_ghost_cell
//   solver.hpp:811-834             sort_gc (canonical = stable order)
//   solver.hpp:918-928             distances
//   coupling.hpp:24-39             local_bndry_ids (CHT partner map)
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

#include "internal.hpp"

namespace hom3g {

namespace {

constexpr int64_t kInv = INT64_MAX;

__host__ __device__ inline int64_t no_nodes_upto(int nd, int level) {
  // total nodes of a full tree with levels [0, level]
  int64_t tot = 0, lvl = 1;
  for (int l = 0; l <= level; ++l) { tot += lvl; lvl <<= nd; }
  return tot;
}

// containers/hierarchical/implementation.hpp:46-57 in closed form: moving
// from child cpos towards neighbour position npos stays inside the parent
// iff the child's bit along the direction points the other way.
__device__ inline int rel_sibling(int cpos, int npos) {
  const int d = npos >> 1;
  const int want = (npos & 1) ? 0 : 1;
  return (((cpos >> d) & 1) == want) ? (cpos ^ (1 << d)) : -1;
}

__global__ void k_uniform_tree(int nd, int L, int64_t nn, int64_t* parent, int64_t* first_child,
                               int32_t* level) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int l = 0;
  int64_t off = 0, cnt = 1;
  while (n >= off + cnt) { off += cnt; cnt <<= nd; ++l; }
  const int64_t k = n - off;
  const int64_t off_prev = (l > 0) ? off - (cnt >> nd) : 0;
  parent[n] = (l > 0) ? off_prev + (k >> nd) : kInv;
  first_child[n] = (l < L) ? off + cnt + (k << nd) : kInv;
  level[n] = l;
}

__global__ void k_levels(int64_t nn, const int64_t* parent, int32_t* level) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int l = 0;
  int64_t p = parent[n];
  while (p != kInv) { p = parent[p]; ++l; }
  level[n] = l;
}

__device__ int64_t samelvl_neighbor(const int64_t* parent, const int64_t* first_child, int64_t n, int pos) {
  if (parent[n] == kInv) return kInv;
  int trav[32];
  int ntr = 0;
  int64_t cur = n;
  int pip;
  while (true) {
    const int64_t p = parent[cur];
    pip = (int)(cur - first_child[p]);
    if (rel_sibling(pip, pos) >= 0) break;
    if (parent[p] == kInv) return kInv;
    trav[ntr++] = pip;
    cur = p;
  }
  cur = first_child[parent[cur]] + rel_sibling(pip, pos);
  const int opp = pos ^ 1;
  for (int i = ntr - 1; i >= 0; --i) {
    const int64_t fc = first_child[cur];
    if (fc == kInv) return kInv;
    cur = fc + rel_sibling(trav[i], opp);
  }
  return cur;
}

__global__ void k_node_neighbors(int npos, int64_t nn, const int64_t* parent, const int64_t* first_child,
                                 int64_t* out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nn * npos) return;
  const int64_t n = t / npos;
  const int pos = (int)(t % npos);
  out[t] = samelvl_neighbor(parent, first_child, n, pos);
}

// cell_coordinates (grid.hpp:358-374): x(n) = x(parent) + crp * (0.25 * len(parent))
__device__ void node_coords(int nd, double root_len, const double* root_x, const int64_t* parent,
                            const int64_t* first_child, int64_t n, double* x, double* len) {
  int pips[32];
  int l = 0;
  int64_t cur = n;
  while (parent[cur] != kInv) {
    const int64_t p = parent[cur];
    pips[l++] = (int)(cur - first_child[p]);
    cur = p;
  }
  for (int d = 0; d < nd; ++d) x[d] = root_x[d];
  for (int k = l - 1, plev = 0; k >= 0; --k, ++plev) {
    const double plen = root_len / (double)(1ll << plev);
    const double rel = 0.25 * plen;
    for (int d = 0; d < nd; ++d) x[d] = x[d] + (double)(((pips[k] >> d) & 1) ? 1 : -1) * rel;
  }
  *len = root_len / (double)(1ll << l);
}

struct RootInfo {
  int nd;
  double root_len;
  double root_x[3];
};

__global__ void k_node_geometry(RootInfo R, int64_t nn, const int64_t* parent, const int64_t* first_child,
                                double* x_rowmajor, double* len) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  double x[3], l;
  node_coords(R.nd, R.root_len, R.root_x, parent, first_child, n, x, &l);
  for (int d = 0; d < R.nd; ++d) x_rowmajor[n * R.nd + d] = x[d];
  len[n] = l;
}

// geometry/implicit.hpp:39-103 on the device
__device__ double sdf(const hom3g_geometry& g, const double* x, int nd) {
  double r;
  if (g.kind == HOM3G_GEOM_EDGE) {
    double t[3];
    for (int d = 0; d < nd; ++d) t[d] = g.b[d] * (x[d] - g.a[d]);
    r = (nd == 2) ? t[0] + t[1] : t[0] + (t[1] + t[2]);
  } else if (g.kind == HOM3G_GEOM_SQUARE) {
    double dist[3], l2[3];
    for (int d = 0; d < nd; ++d) l2[d] = 0.5 * g.b[d];
    for (int d = 0; d < nd; ++d) dist[d] = g.a[d] - l2[d] - x[d];
    for (int d = 0; d < nd; ++d)
      if (x[d] > g.a[d]) dist[d] = x[d] - (g.a[d] + l2[d]);
    int np = 0;
    for (int d = 0; d < nd; ++d) np += dist[d] > 0 ? 1 : 0;
    if (np == nd) {
      double s[3];
      for (int d = 0; d < nd; ++d) s[d] = dist[d] * dist[d];
      r = sqrt(nd == 2 ? s[0] + s[1] : s[0] + (s[1] + s[2]));
    } else if (np < 2) {
      r = dist[0];
      for (int d = 1; d < nd; ++d) r = r < dist[d] ? dist[d] : r;
    } else if (nd == 3) {
      if (dist[0] < 0) r = sqrt(dist[1] * dist[1] + dist[2] * dist[2]);
      else if (dist[1] < 0) r = sqrt(dist[0] * dist[0] + dist[2] * dist[2]);
      else if (dist[2] < 0) r = sqrt(dist[0] * dist[0] + dist[1] * dist[1]);
      else r = __builtin_nan("");
    } else {
      r = __builtin_nan("");
    }
  } else if (g.kind == HOM3G_GEOM_SPHERE) {
    double s[3];
    for (int d = 0; d < nd; ++d) { const double t = x[d] - g.a[d]; s[d] = t * t; }
    r = sqrt(nd == 2 ? s[0] + s[1] : s[0] + (s[1] + s[2])) - g.b[0];
  } else {
    r = __builtin_nan("");
  }
  return g.invert ? -1 * r : r;
}

__device__ bool is_cut_by(const hom3g_geometry& g, int nd, const double* xc, double len) {
  const double half = 0.5 * len;
  bool all_pos = true, all_neg = true;
  for (int v = 0; v < (1 << nd); ++v) {
    double xv[3];
    for (int d = 0; d < nd; ++d) xv[d] = xc[d] + half * (double)(((v >> d) & 1) ? 1 : -1);
    const double s = sdf(g, xv, nd);
    all_pos = all_pos && (s > 0);
    all_neg = all_neg && (s < 0);
  }
  return !(all_pos || all_neg);
}

struct DomainSpec {
  hom3g_geometry g;
  int sign;
  int use_mask;
};

// leaf flag + domain test per node (solver.hpp:884-886)
__global__ void k_leaf_flags(RootInfo R, DomainSpec D, int64_t nn, const int64_t* parent,
                             const int64_t* first_child, const int64_t* leaf_rank, const uint8_t* mask,
                             int64_t* in_flag) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  if (first_child[n] != kInv) { in_flag[n] = 0; return; }
  if (D.use_mask) { in_flag[n] = mask[leaf_rank[n]] ? 1 : 0; return; }
  if (D.sign == 0) { in_flag[n] = 1; return; }
  double x[3], l;
  node_coords(R.nd, R.root_len, R.root_x, parent, first_child, n, x, &l);
  const double s = sdf(D.g, x, R.nd);
  in_flag[n] = D.sign > 0 ? (s > 0. ? 1 : 0) : (s < 0. ? 1 : 0);
}

__global__ void k_is_leaf(int64_t nn, const int64_t* first_child, int64_t* flag) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  flag[n] = first_child[n] == kInv ? 1 : 0;
}

__global__ void k_make_cells(RootInfo R, int64_t nn, const int64_t* parent, const int64_t* first_child,
                             const int64_t* in_flag, const int64_t* cell_of, int64_t* node2cell,
                             int64_t* node_idx, double* xc, double* len, int64_t ld) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  if (!in_flag[n]) { node2cell[n] = kInv; return; }
  const int64_t c = cell_of[n];
  node2cell[n] = c;
  node_idx[c] = n;
  double x[3], l;
  node_coords(R.nd, R.root_len, R.root_x, parent, first_child, n, x, &l);
  for (int d = 0; d < R.nd; ++d) xc[d * ld + c] = x[d];
  len[c] = l;
}

// Solver neighbours (solver.hpp:899-908).  Same-level leaf -> its cell; on
// refined grids (2:1 extension) a same-level neighbour with children is a
// face towards its 2^(nd-1) children on the shared face (flagged in
// is_fine[c * npos + pos], resolved after a scan), and a missing
// same-level neighbour whose parent-level neighbour is a leaf is a face
// towards that coarser leaf.
__global__ void k_cell_neighbors(int nd, int64_t nc, const int64_t* parent, const int64_t* first_child,
                                 const int64_t* node_idx, const int64_t* node2cell, int32_t* nghbr,
                                 int64_t ld, int64_t* is_fine, int* err) {
  const int npos = 2 * nd;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nc * npos) return;
  const int64_t c = t / npos;
  const int pos = (int)(t % npos);
  const int64_t n = node_idx[c];
  const int64_t nn = samelvl_neighbor(parent, first_child, n, pos);
  int32_t val = -1;
  int64_t fine = 0;
  if (nn != kInv && first_child[nn] == kInv) {
    const int64_t cid = node2cell[nn];
    val = (cid == kInv) ? -1 : (int32_t)cid;
  } else if (nn != kInv) {
    const int d = pos >> 1, want = (pos & 1) ? 0 : 1;
    int nvalid = 0, k = 0;
    for (int cp = 0; cp < (1 << nd); ++cp) {
      if (((cp >> d) & 1) != want) continue;
      const int64_t ch = first_child[nn] + cp;
      if (first_child[ch] != kInv) atomicOr(err, 1);  // not 2:1 balanced
      nvalid += (node2cell[ch] != kInv) ? 1 : 0;
      ++k;
    }
    if (nvalid == k) fine = 1;
    else if (nvalid != 0) atomicOr(err, 2);  // domain boundary cuts a refined face
  } else if (parent[n] != kInv) {
    const int64_t p = parent[n];
    const int64_t pn = samelvl_neighbor(parent, first_child, p, pos);
    const int d = pos >> 1, side = pos & 1;
    const int pip = (int)(n - first_child[p]);
    if (((pip >> d) & 1) == side && pn != kInv && first_child[pn] == kInv) {
      const int64_t cid = node2cell[pn];
      val = (cid == kInv) ? -1 : (int32_t)cid;
    }
  }
  nghbr[pos * ld + c] = val;
  is_fine[t] = fine;
}

__global__ void k_fine_faces(int nd, int64_t nc, const int64_t* parent, const int64_t* first_child,
                             const int64_t* node_idx, const int64_t* node2cell, const int64_t* is_fine,
                             const int64_t* fidx, int32_t* nghbr, int64_t ld, int32_t* fine) {
  const int npos = 2 * nd;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nc * npos || !is_fine[t]) return;
  const int64_t c = t / npos;
  const int pos = (int)(t % npos);
  const int64_t nn = samelvl_neighbor(parent, first_child, node_idx[c], pos);
  const int64_t f = fidx[t];
  const int d = pos >> 1, want = (pos & 1) ? 0 : 1;
  const int K = 1 << (nd - 1);
  int k = 0;
  for (int cp = 0; cp < (1 << nd); ++cp) {
    if (((cp >> d) & 1) != want) continue;
    fine[f * K + k++] = (int32_t)node2cell[first_child[nn] + cp];
  }
  nghbr[pos * ld + c] = (int32_t)(-(2 + f));
}

// sets bit 4 of *flags if the cells' leaves are not all on one level
__global__ void k_level_range(int64_t nc, const int64_t* node_idx, const int32_t* level, int* flags) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nc) return;
  if (level[node_idx[c]] != level[node_idx[0]]) atomicOr(flags, 4);
}

// ---- grid refinement passes (2:1 extension) ---------------------------------
// A pass flags leaves on the tree as it stands at the start of the pass and
// refines them in increasing node order: the children of the k-th flagged
// node are nodes end + 2^nd * k ... (children appended at the end, as in
// implementation.hpp:320-364).
__global__ void k_refine_flags(RootInfo R, int max_level, hom3g_geometry g, double band, int64_t nn,
                               const int64_t* parent, const int64_t* first_child, const int32_t* level,
                               int64_t* flag) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  int64_t f = 0;
  if (first_child[n] == kInv && level[n] < max_level) {
    double x[3], h;
    node_coords(R.nd, R.root_len, R.root_x, parent, first_child, n, x, &h);
    const double s = sdf(g, x, R.nd);
    const double reach = band + 0.5 * h * sqrt((double)R.nd);
    f = (fabs(s) <= reach) ? 1 : 0;
  }
  flag[n] = f;
}

// 2:1 balance: a refined node whose same-level face neighbour is missing
// flags the coarser leaf covering that position.
__global__ void k_balance_flags(int nd, int64_t nn, const int64_t* parent, const int64_t* first_child,
                                int64_t* flag) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= nn) return;
  if (first_child[n] == kInv || parent[n] == kInv) return;
  for (int pos = 0; pos < 2 * nd; ++pos) {
    if (samelvl_neighbor(parent, first_child, n, pos) != kInv) continue;
    for (int64_t cur = parent[n]; cur != kInv; cur = parent[cur]) {
      const int64_t nbp = samelvl_neighbor(parent, first_child, cur, pos);
      if (nbp != kInv) {
        if (first_child[nbp] == kInv) flag[nbp] = 1;
        break;
      }
    }
  }
}

__global__ void k_zero_i64(int64_t n, int64_t* a) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) a[i] = 0;
}

__global__ void k_refine_apply(int nd, int64_t end, const int64_t* flag, const int64_t* rank, int64_t* parent,
                               int64_t* first_child, int32_t* level) {
  const int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (n >= end || !flag[n]) return;
  const int nch = 1 << nd;
  const int64_t first = end + (int64_t)nch * rank[n];
  first_child[n] = first;
  for (int c = 0; c < nch; ++c) {
    parent[first + c] = n;
    first_child[first + c] = kInv;
    level[first + c] = level[n] + 1;
  }
}

// ghost count for one boundary (solver.hpp:940-967)
__global__ void k_ghost_count(hom3g_geometry g, int nd, int64_t nc, const int32_t* nghbr, int64_t ld,
                              const double* xc, const double* len, int64_t* cnt, uint8_t* posmask) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nc) return;
  double x[3];
  for (int d = 0; d < nd; ++d) x[d] = xc[d * ld + c];
  const double l = len[c];
  int k = 0;
  uint8_t m = 0;
  if (is_cut_by(g, nd, x, l)) {
    for (int pos = 0; pos < 2 * nd; ++pos) {
      if (nghbr[pos * ld + c] != -1) continue;
      double xn[3];
      for (int d = 0; d < nd; ++d) xn[d] = x[d] + (double)((pos >> 1) == d ? ((pos & 1) ? 1 : -1) : 0) * l;
      if (sdf(g, xn, nd) < 0.) { ++k; m |= (uint8_t)(1u << pos); }
    }
  }
  cnt[c] = k;
  posmask[c] = m;
}

struct GhostTmp {
  int32_t* bndry;
  int32_t* pos;
  int32_t* bc;
};

__global__ void k_ghost_create(int nd, int64_t nc, int64_t ghost_base, int bc, const int64_t* off,
                               const uint8_t* posmask, int32_t* nghbr, int64_t ld, GhostTmp T) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= nc) return;
  const uint8_t m = posmask[c];
  if (!m) return;
  int64_t g = off[c];  // index within this boundary's ghosts
  for (int pos = 0; pos < 2 * nd; ++pos) {
    if (!((m >> pos) & 1)) continue;
    nghbr[pos * ld + c] = (int32_t)(nc + ghost_base + g);
    T.bndry[g] = (int32_t)c;
    T.pos[g] = pos;
    T.bc[g] = bc;
    ++g;
  }
}

struct GhostFinal {
  const int32_t* bndry;
  const int32_t* pos;
  const int32_t* bc;
};

__global__ void k_ghost_finalize(int nd, int64_t n_int, int64_t ng, GhostFinal T, int32_t* nghbr, int64_t ld,
                                 double* xc, double* len, int64_t* node_idx, int32_t* bc_idx,
                                 int32_t* g_bndry, int32_t* g_pos, int32_t* g_hookbc) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int64_t c = n_int + g;
  const int b = T.bndry[g];
  const int pos = T.pos[g];
  for (int p = 0; p < 2 * nd; ++p) nghbr[p * ld + c] = -1;
  nghbr[(pos ^ 1) * ld + c] = b;
  const double l = len[b];
  // ghost_cell_coordinates (solver.hpp:787-807): x_b + len_b * rel(pos)
  for (int d = 0; d < nd; ++d)
    xc[d * ld + c] = xc[d * ld + b] + l * (double)((pos >> 1) == d ? ((pos & 1) ? 1 : -1) : 0);
  len[c] = l;
  node_idx[c] = kInv;
  bc_idx[c] = T.bc[g];
  g_bndry[g] = b;
  g_pos[g] = pos;
  g_hookbc[g] = T.bc[g];
}

__global__ void k_internal_bcidx(int64_t n_int, int32_t* bc_idx) {
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c < n_int) bc_idx[c] = INT32_MAX;
}

// distances (solver.hpp:918-928): Euclidean norm in Eigen's summation order
__global__ void k_distances(int nd, int64_t nc, const int32_t* nghbr, int64_t ld, const double* xc,
                            double* dist) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nc * 2 * nd) return;
  const int64_t c = t % nc;
  const int pos = (int)(t / nc);
  const int n = nghbr[pos * ld + c];
  if (n < 0) { dist[pos * ld + c] = __builtin_nan(""); return; }
  double s[3];
  for (int d = 0; d < nd; ++d) { const double v = xc[d * ld + c] - xc[d * ld + n]; s[d] = v * v; }
  dist[pos * ld + c] = sqrt(nd == 2 ? s[0] + s[1] : s[0] + (s[1] + s[2]));
}

// periodic image: walk from the boundary cell away from the ghost to the
// last internal cell of the row
__global__ void k_periodic_images(int64_t n_int, int64_t ng, const int32_t* g_bndry, const int32_t* g_pos,
                                  const int32_t* nghbr, int64_t ld, const double* xc, const double* dist_unused,
                                  int32_t* g_image, double* g_dist, int nd) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int b = g_bndry[g];
  const int walk = g_pos[g] ^ 1;  // direction from ghost towards the boundary cell
  int cur = b;
  while (true) {
    const int n = nghbr[walk * ld + cur];
    if (n < 0 || n >= n_int) break;
    cur = n;
  }
  g_image[g] = cur;
  // cell_dx(bndry, ghost)
  const int64_t c = n_int + g;
  double s[3];
  for (int d = 0; d < nd; ++d) { const double v = xc[d * ld + b] - xc[d * ld + c]; s[d] = v * v; }
  g_dist[g] = sqrt(nd == 2 ? s[0] + s[1] : s[0] + (s[1] + s[2]));
}

// CHT partner (coupling.hpp:24-39): same-level grid neighbour of the
// boundary node on the ghost side, mapped to the other solver's cells.
__global__ void k_partners(int64_t ng, const int32_t* g_bndry, const int32_t* g_pos, const int32_t* g_bc,
                           const int* is_cht, const int64_t* node_idx, const int64_t* parent,
                           const int64_t* first_child, const int64_t* const* other_n2c, int32_t* g_partner,
                           int* err) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= ng) return;
  const int bc = g_bc[g];
  if (!is_cht[bc]) { g_partner[g] = -1; return; }
  const int64_t node0 = node_idx[g_bndry[g]];
  const int64_t node1 = samelvl_neighbor(parent, first_child, node0, g_pos[g]);
  int64_t cell = (node1 == kInv) ? kInv : other_n2c[bc][node1];
  if (cell == kInv) { g_partner[g] = -1; atomicAdd(err, 1); return; }
  g_partner[g] = (int32_t)cell;
}

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

template <class T>
void exclusive_scan(Ctx& C, const T* in, T* out, int64_t n) {
  size_t bytes = 0;
  H3_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, (int)n, C.stream));
  DevBuf<char> tmp;
  tmp.alloc(bytes ? bytes : 1);
  H3_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, (int)n, C.stream));
  H3_HIP(hipStreamSynchronize(C.stream));
}

RootInfo root_info(const GridDev& G) {
  RootInfo R;
  R.nd = G.nd;
  R.root_len = G.root_len;
  for (int d = 0; d < 3; ++d) R.root_x[d] = G.root_x[d];
  return R;
}

}  // namespace

void grid_generate_uniform(Ctx& C, GridDev& G, int level) {
  H3_REQUIRE(level >= 0 && level <= 20, "min_level out of range");
  const int64_t nn = no_nodes_upto(G.nd, level);
  G.nnodes = nn;
  G.parent.alloc(nn);
  G.first_child.alloc(nn);
  G.level.alloc(nn);
  hipLaunchKernelGGL(k_uniform_tree, dim3(nblk(nn)), dim3(256), 0, C.stream, G.nd, level, nn, G.parent.p,
                     G.first_child.p, G.level.p);
  H3_LAUNCH_CHECK();
  G.uniform_level = level;
}

template <class T>
static void grow(DevBuf<T>& b, int64_t keep, int64_t count, hipStream_t st) {
  DevBuf<T> nb;
  nb.alloc(count);
  if (keep > 0) H3_HIP(hipMemcpyAsync(nb.p, b.p, keep * sizeof(T), hipMemcpyDeviceToDevice, st));
  std::swap(b.p, nb.p);
  std::swap(b.n, nb.n);
}

// one pass: refine the flagged nodes of [0, end); returns how many
static int64_t refine_flagged(Ctx& C, GridDev& G, DevBuf<int64_t>& flag) {
  const int64_t end = G.nnodes;
  DevBuf<int64_t> rank;
  rank.alloc(end);
  exclusive_scan(C, flag.p, rank.p, end);
  int64_t lr = 0, lf = 0;
  rank.download(&lr, 1, C.stream, end - 1);
  flag.download(&lf, 1, C.stream, end - 1);
  H3_HIP(hipStreamSynchronize(C.stream));
  const int64_t k = lr + lf;
  if (k == 0) return 0;
  const int64_t nn = end + ((int64_t)1 << G.nd) * k;
  grow(G.parent, end, nn, C.stream);
  grow(G.first_child, end, nn, C.stream);
  grow(G.level, end, nn, C.stream);
  hipLaunchKernelGGL(k_refine_apply, dim3(nblk(end)), dim3(256), 0, C.stream, G.nd, end, flag.p, rank.p,
                     G.parent.p, G.first_child.p, G.level.p);
  H3_LAUNCH_CHECK();
  H3_HIP(hipStreamSynchronize(C.stream));
  G.nnodes = nn;
  G.uniform_level = -1;
  return k;
}

int64_t grid_refine_region(Ctx& C, GridDev& G, int max_level, const hom3g_geometry& g, double band) {
  H3_REQUIRE(max_level >= 0 && max_level <= 20, "max_level out of range");
  int64_t total = 0;
  for (;;) {
    DevBuf<int64_t> flag;
    flag.alloc(G.nnodes);
    hipLaunchKernelGGL(k_refine_flags, dim3(nblk(G.nnodes)), dim3(256), 0, C.stream, root_info(G), max_level, g,
                       band, G.nnodes, G.parent.p, G.first_child.p, G.level.p, flag.p);
    H3_LAUNCH_CHECK();
    const int64_t k = refine_flagged(C, G, flag);
    if (k == 0) break;
    total += k;
  }
  return total;
}

int64_t grid_balance(Ctx& C, GridDev& G) {
  int64_t total = 0;
  for (;;) {
    DevBuf<int64_t> flag;
    flag.alloc(G.nnodes);
    hipLaunchKernelGGL(k_zero_i64, dim3(nblk(G.nnodes)), dim3(256), 0, C.stream, G.nnodes, flag.p);
    H3_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_balance_flags, dim3(nblk(G.nnodes)), dim3(256), 0, C.stream, G.nd, G.nnodes, G.parent.p,
                       G.first_child.p, flag.p);
    H3_LAUNCH_CHECK();
    const int64_t k = refine_flagged(C, G, flag);
    if (k == 0) break;
    total += k;
  }
  return total;
}

void grid_compute_levels(Ctx& C, GridDev& G) {
  G.level.alloc(G.nnodes);
  hipLaunchKernelGGL(k_levels, dim3(nblk(G.nnodes)), dim3(256), 0, C.stream, G.nnodes, G.parent.p, G.level.p);
  H3_LAUNCH_CHECK();
}

void grid_neighbors(Ctx& C, GridDev& G, int64_t* host_out) {
  const int npos = 2 * G.nd;
  DevBuf<int64_t> out;
  out.alloc(G.nnodes * npos);
  hipLaunchKernelGGL(k_node_neighbors, dim3(nblk(G.nnodes * npos)), dim3(256), 0, C.stream, npos, G.nnodes,
                     G.parent.p, G.first_child.p, out.p);
  H3_LAUNCH_CHECK();
  out.download(host_out, G.nnodes * npos, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
}

void grid_geometry(Ctx& C, GridDev& G, double* x_rowmajor, double* len) {
  DevBuf<double> x, l;
  x.alloc(G.nnodes * G.nd);
  l.alloc(G.nnodes);
  hipLaunchKernelGGL(k_node_geometry, dim3(nblk(G.nnodes)), dim3(256), 0, C.stream, root_info(G), G.nnodes,
                     G.parent.p, G.first_child.p, x.p, l.p);
  H3_LAUNCH_CHECK();
  if (x_rowmajor) x.download(x_rowmajor, G.nnodes * G.nd, C.stream);
  if (len) l.download(len, G.nnodes, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
}

// Replicates the reference's apply_bcs range computation (solver.hpp:561-586)
// including its behaviour when a boundary has no ghost cells: each ghost is
// assigned the condition whose range covers it.
static std::vector<int32_t> apply_ranges(const std::vector<int32_t>& gbc, int nbc) {
  const int64_t N = (int64_t)gbc.size();
  std::vector<int32_t> out(N, -1);
  if (N == 0) return out;
  int bcIdx = gbc[0];
  auto find = [&](int64_t from) {
    for (int64_t c = from; c < N; ++c)
      if (gbc[c] == bcIdx) return c;
    return N;
  };
  int64_t first = find(0), last = first;
  for (int k = 0; k < nbc; ++k) {
    ++bcIdx;
    first = last;
    last = find(first);
    for (int64_t g = first; g < last; ++g) out[g] = k;
  }
  return out;
}

void solver_build(Ctx& C, GridDev& G, SolverDev& S) {
  const int nd = S.nd, npos = 2 * nd;
  const int64_t nn = G.nnodes;
  const RootInfo R = root_info(G);
  // 1) leaves in node order passing the domain test
  DevBuf<int64_t> flag, rank, inflag, cellof;
  flag.alloc(nn);
  rank.alloc(nn);
  inflag.alloc(nn);
  cellof.alloc(nn);
  DevBuf<uint8_t> mask;
  DomainSpec D;
  D.g = S.domain_geom;
  D.sign = S.domain_sign;
  D.use_mask = S.domain_mask.empty() ? 0 : 1;
  if (D.use_mask) {
    hipLaunchKernelGGL(k_is_leaf, dim3(nblk(nn)), dim3(256), 0, C.stream, nn, G.first_child.p, flag.p);
    H3_LAUNCH_CHECK();
    exclusive_scan(C, flag.p, rank.p, nn);
    mask.alloc(S.domain_mask.size());
    mask.upload(S.domain_mask.data(), S.domain_mask.size(), C.stream);
  }
  hipLaunchKernelGGL(k_leaf_flags, dim3(nblk(nn)), dim3(256), 0, C.stream, R, D, nn, G.parent.p,
                     G.first_child.p, rank.p, mask.p, inflag.p);
  H3_LAUNCH_CHECK();
  exclusive_scan(C, inflag.p, cellof.p, nn);
  int64_t last_off = 0, last_flag = 0;
  cellof.download(&last_off, 1, C.stream, nn - 1);
  inflag.download(&last_flag, 1, C.stream, nn - 1);
  H3_HIP(hipStreamSynchronize(C.stream));
  const int64_t n_int = last_off + last_flag;
  H3_REQUIRE(n_int > 0, "solver domain contains no cells");
  H3_REQUIRE(n_int < INT32_MAX / 2, "too many cells for 32-bit device ids");

  // temporary internal-cell arrays (ld = n_int)
  DevBuf<int32_t> ngh;
  DevBuf<double> xc, len;
  DevBuf<int64_t> nidx;
  ngh.alloc((size_t)npos * n_int);
  xc.alloc((size_t)nd * n_int);
  len.alloc(n_int);
  nidx.alloc(n_int);
  S.node2cell.alloc(nn);
  hipLaunchKernelGGL(k_make_cells, dim3(nblk(nn)), dim3(256), 0, C.stream, R, nn, G.parent.p, G.first_child.p,
                     inflag.p, cellof.p, S.node2cell.p, nidx.p, xc.p, len.p, n_int);
  H3_LAUNCH_CHECK();
  DevBuf<int64_t> isfine, fidx;
  DevBuf<int> d_err;
  isfine.alloc((size_t)npos * n_int);
  fidx.alloc((size_t)npos * n_int);
  d_err.alloc(1);
  {
    int zero = 0;
    d_err.upload(&zero, 1, C.stream);
  }
  hipLaunchKernelGGL(k_cell_neighbors, dim3(nblk(n_int * npos)), dim3(256), 0, C.stream, nd, n_int,
                     G.parent.p, G.first_child.p, nidx.p, S.node2cell.p, ngh.p, n_int, isfine.p, d_err.p);
  H3_LAUNCH_CHECK();
  exclusive_scan(C, isfine.p, fidx.p, (int64_t)npos * n_int);
  {
    int err = 0;
    int64_t lo = 0, lf = 0;
    d_err.download(&err, 1, C.stream);
    fidx.download(&lo, 1, C.stream, (size_t)npos * n_int - 1);
    isfine.download(&lf, 1, C.stream, (size_t)npos * n_int - 1);
    H3_HIP(hipStreamSynchronize(C.stream));
    H3_REQUIRE(!(err & 1), "grid is not 2:1 balanced (call hom3g_grid_balance)");
    H3_REQUIRE(!(err & 2), "domain boundary cuts a refined face");
    S.nfine = lo + lf;
  }
  const int K = 1 << (nd - 1);
  S.fine.alloc(S.nfine > 0 ? (size_t)S.nfine * K : 1);
  if (S.nfine > 0) {
    hipLaunchKernelGGL(k_fine_faces, dim3(nblk(n_int * npos)), dim3(256), 0, C.stream, nd, n_int, G.parent.p,
                       G.first_child.p, nidx.p, S.node2cell.p, isfine.p, fidx.p, ngh.p, n_int, S.fine.p);
    H3_LAUNCH_CHECK();
  }
  {
    // leaves of more than one level?
    hipLaunchKernelGGL(k_level_range, dim3(nblk(n_int)), dim3(256), 0, C.stream, n_int, nidx.p, G.level.p,
                       d_err.p);
    H3_LAUNCH_CHECK();
    int flags = 0;
    d_err.download(&flags, 1, C.stream);
    H3_HIP(hipStreamSynchronize(C.stream));
    S.nonuniform = (flags & 4) != 0;
    H3_REQUIRE(!S.nonuniform || (S.P.physics == HOM3G_EULER && S.P.flux == HOM3G_FLUX_AUSM &&
                                 S.P.recon == HOM3G_RECON_NONE),
               "refined grids support first-order Euler AUSM only");
  }
  isfine.release(); fidx.release();
  flag.release(); rank.release(); inflag.release(); cellof.release();

  // 2) ghosts, boundary by boundary (creation order = canonical sort order)
  std::vector<int32_t> h_bndry, h_pos, h_bc;
  DevBuf<int64_t> cnt, off;
  DevBuf<uint8_t> pm;
  cnt.alloc(n_int);
  off.alloc(n_int);
  pm.alloc(n_int);
  int64_t ng = 0;
  DevBuf<int32_t> tb, tp, tc;
  for (size_t b = 0; b < S.bcs.size(); ++b) {
    hipLaunchKernelGGL(k_ghost_count, dim3(nblk(n_int)), dim3(256), 0, C.stream, S.bcs[b].geom, nd, n_int,
                       ngh.p, (int64_t)n_int, xc.p, len.p, cnt.p, pm.p);
    H3_LAUNCH_CHECK();
    exclusive_scan(C, cnt.p, off.p, n_int);
    int64_t lo = 0, lc = 0;
    off.download(&lo, 1, C.stream, n_int - 1);
    cnt.download(&lc, 1, C.stream, n_int - 1);
    H3_HIP(hipStreamSynchronize(C.stream));
    const int64_t nb = lo + lc;
    if (nb == 0) continue;
    tb.alloc(nb); tp.alloc(nb); tc.alloc(nb);
    GhostTmp T{tb.p, tp.p, tc.p};
    hipLaunchKernelGGL(k_ghost_create, dim3(nblk(n_int)), dim3(256), 0, C.stream, nd, n_int, ng,
                       (int)b, off.p, pm.p, ngh.p, (int64_t)n_int, T);
    H3_LAUNCH_CHECK();
    std::vector<int32_t> hb(nb), hp(nb), hc(nb);
    tb.download(hb.data(), nb, C.stream);
    tp.download(hp.data(), nb, C.stream);
    tc.download(hc.data(), nb, C.stream);
    H3_HIP(hipStreamSynchronize(C.stream));
    h_bndry.insert(h_bndry.end(), hb.begin(), hb.end());
    h_pos.insert(h_pos.end(), hp.begin(), hp.end());
    h_bc.insert(h_bc.end(), hc.begin(), hc.end());
    ng += nb;
  }
  H3_REQUIRE((size_t)S.bcs.size() <= (size_t)kMaxBC, "too many boundary conditions");

  // 3) final arrays with ld = cap
  const int64_t ncells = n_int + ng;
  int64_t cap = S.P.max_cells > 0 ? S.P.max_cells : ncells;
  if (S.P.max_cells > 0 && ncells > S.P.max_cells)
    throw Error(HOM3G_ERR_CAPACITY, "maxNoCells exceeded (" + std::to_string(ncells) + " cells)");
  cap = ((cap + 63) / 64) * 64;
  S.cap = cap;
  S.n_int = n_int;
  S.n_ghost = ng;
  S.ncells = ncells;
  S.nghbr.alloc((size_t)npos * cap);
  S.x_center.alloc((size_t)nd * cap);
  S.length.alloc(cap);
  S.node_idx.alloc(cap);
  S.bc_idx.alloc(cap);
  S.dist.alloc((size_t)npos * cap);
  for (int p = 0; p < npos; ++p)
    H3_HIP(hipMemcpy2DAsync(S.nghbr.p + (size_t)p * cap, sizeof(int32_t), ngh.p + (size_t)p * n_int,
                            sizeof(int32_t), sizeof(int32_t), n_int, hipMemcpyDeviceToDevice, C.stream));
  for (int d = 0; d < nd; ++d)
    H3_HIP(hipMemcpyAsync(S.x_center.p + (size_t)d * cap, xc.p + (size_t)d * n_int, n_int * sizeof(double),
                          hipMemcpyDeviceToDevice, C.stream));
  H3_HIP(hipMemcpyAsync(S.length.p, len.p, n_int * sizeof(double), hipMemcpyDeviceToDevice, C.stream));
  H3_HIP(hipMemcpyAsync(S.node_idx.p, nidx.p, n_int * sizeof(int64_t), hipMemcpyDeviceToDevice, C.stream));
  hipLaunchKernelGGL(k_internal_bcidx, dim3(nblk(n_int)), dim3(256), 0, C.stream, n_int, S.bc_idx.p);
  H3_LAUNCH_CHECK();
  const int64_t ngA = ng > 0 ? ng : 1;
  S.g_bndry.alloc(ngA); S.g_pos.alloc(ngA); S.g_hookbc.alloc(ngA); S.g_applybc.alloc(ngA);
  S.g_image.alloc(ngA); S.g_partner.alloc(ngA); S.g_dist.alloc(ngA);
  H3_HIP(hipMemsetAsync(S.g_partner.p, 0xff, ngA * sizeof(int32_t), C.stream));  // -1: no partner yet
  if (ng > 0) {
    DevBuf<int32_t> fb, fp, fc;
    fb.alloc(ng); fp.alloc(ng); fc.alloc(ng);
    fb.upload(h_bndry.data(), ng, C.stream);
    fp.upload(h_pos.data(), ng, C.stream);
    fc.upload(h_bc.data(), ng, C.stream);
    GhostFinal T{fb.p, fp.p, fc.p};
    hipLaunchKernelGGL(k_ghost_finalize, dim3(nblk(ng)), dim3(256), 0, C.stream, nd, n_int, ng, T, S.nghbr.p,
                       cap, S.x_center.p, S.length.p, S.node_idx.p, S.bc_idx.p, S.g_bndry.p, S.g_pos.p,
                       S.g_hookbc.p);
    H3_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_periodic_images, dim3(nblk(ng)), dim3(256), 0, C.stream, n_int, ng, S.g_bndry.p,
                       S.g_pos.p, S.nghbr.p, cap, S.x_center.p, (const double*)nullptr, S.g_image.p,
                       S.g_dist.p, nd);
    H3_LAUNCH_CHECK();
    std::vector<int32_t> ap = apply_ranges(h_bc, (int)S.bcs.size());
    S.g_applybc.upload(ap.data(), ng, C.stream);
    H3_HIP(hipStreamSynchronize(C.stream));
  }
  hipLaunchKernelGGL(k_distances, dim3(nblk(ncells * npos)), dim3(256), 0, C.stream, nd, ncells, S.nghbr.p,
                     cap, S.x_center.p, S.dist.p);
  H3_LAUNCH_CHECK();
  H3_HIP(hipStreamSynchronize(C.stream));

  // brick-kernel eligibility: full uniform grid, cell id == Morton index,
  // one edge boundary per face that is periodic or a zero-gradient Neumann.
  S.level = G.uniform_level;
  S.length.download(&S.h_uniform, 1, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  S.brick_ok = false;
  S.uses_vortex_bc = false;
  for (auto& bc : S.bcs) if (bc.kind == HOM3G_BC_ISENTROPIC_VORTEX) S.uses_vortex_bc = true;
  if (G.uniform_level >= 3 && n_int == (int64_t(1) << (nd * G.uniform_level)) &&
      (int)S.bcs.size() == 2 * nd) {
    bool ok = true;
    bool seen[3][2] = {{false, false}, {false, false}, {false, false}};
    for (auto& bc : S.bcs) {
      if (bc.geom.kind != HOM3G_GEOM_EDGE || bc.geom.invert) { ok = false; break; }
      int dir = -1, side = -1;
      for (int d = 0; d < nd; ++d) {
        if (bc.geom.b[d] == 1.0) { dir = d; side = 0; }
        else if (bc.geom.b[d] == -1.0) { dir = d; side = 1; }
        else if (bc.geom.b[d] != 0.0) ok = false;
      }
      if (dir < 0 || seen[dir][side]) { ok = false; break; }
      seen[dir][side] = true;
      if (bc.kind == HOM3G_BC_PERIODIC) S.face_mode[dir][side] = 1;
      else if (bc.kind == HOM3G_BC_NEUMANN) {
        for (int k = 0; k < bc.nparams; ++k) if (bc.params[k] != 0.0) ok = false;
        S.face_mode[dir][side] = 0;
      } else ok = false;
    }
    for (int d = 0; d < nd && ok; ++d)
      if (S.face_mode[d][0] != S.face_mode[d][1]) ok = false;
    // expected ghost count 2nd * n^(nd-1)
    if (ok && ng != (int64_t)2 * nd * (int64_t(1) << ((nd - 1) * G.uniform_level))) ok = false;
    S.brick_ok = ok && brick_supported(S);
  }
  if (!S.brick_ok && G.uniform_level >= 0 && brick_supported(S) && brick_table_build(C, S)) {
    S.brick_ok = true;
    S.brick_table = true;
  }
  S.built = true;
}

void solver_build_partners(Ctx& C, SolverDev& S) {
  if (S.n_ghost == 0) return;
  GridDev& G = *C.grids[S.grid_id];
  std::vector<int> is_cht(kMaxBC, 0);
  std::vector<const int64_t*> n2c(kMaxBC, nullptr);
  bool any = false;
  for (size_t b = 0; b < S.bcs.size(); ++b) {
    const int k = S.bcs[b].kind;
    if (k == HOM3G_BC_CHT_FLUID || k == HOM3G_BC_CHT_SOLID || k == HOM3G_BC_EULER_COUPLING) {
      H3_REQUIRE(S.bcs[b].other_sid >= 0 && S.bcs[b].other_sid < (int)C.solvers.size() &&
                     C.solvers[S.bcs[b].other_sid],
                 "CHT boundary without a coupled solver");
      SolverDev& O = *C.solvers[S.bcs[b].other_sid];
      H3_REQUIRE(O.built, "coupled solver not built");
      H3_REQUIRE(O.grid_id == S.grid_id, "coupled solvers must share the grid");
      is_cht[b] = 1;
      n2c[b] = O.node2cell.p;
      any = true;
    }
  }
  if (!any) return;
  DevBuf<int> d_is;
  DevBuf<const int64_t*> d_n2c;
  DevBuf<int> d_err;
  d_is.alloc(kMaxBC);
  d_n2c.alloc(kMaxBC);
  d_err.alloc(1);
  d_is.upload(is_cht.data(), kMaxBC, C.stream);
  d_n2c.upload(n2c.data(), kMaxBC, C.stream);
  int zero = 0;
  d_err.upload(&zero, 1, C.stream);
  hipLaunchKernelGGL(k_partners, dim3(nblk(S.n_ghost)), dim3(256), 0, C.stream, S.n_ghost, S.g_bndry.p,
                     S.g_pos.p, S.g_hookbc.p, d_is.p, S.node_idx.p, G.parent.p, G.first_child.p, d_n2c.p,
                     S.g_partner.p, d_err.p);
  H3_LAUNCH_CHECK();
  int err = 0;
  d_err.download(&err, 1, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  H3_REQUIRE(err == 0, "CHT interface ghost without a partner cell in the coupled solver");
}

}  // namespace hom3g
