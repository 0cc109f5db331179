// CPU ORACLE — test infrastructure only (see oracle.hpp).
// Restates the reference's hierarchical tree, Cartesian grid and implicit
// geometries.
#include "oracle.hpp"
#include <algorithm>
#include <stdexcept>

namespace orc {

// grid/grid.hpp:41-59 child_rel_pos stencils (x fastest)
int child_rel_pos(int nd, int cpos, int d) {
  (void)nd;
  return ((cpos >> d) & 1) ? 1 : -1;
}

// grid/grid.hpp:64-80 nghbr_rel_pos stencils: pos 2d -> -1 along d, 2d+1 -> +1
int nghbr_rel_pos(int nd, int npos, int d) {
  (void)nd;
  if (npos / 2 != d) return 0;
  return (npos % 2) ? 1 : -1;
}

// containers/hierarchical/implementation.hpp:46-57 (rel_sibling_positions_arr)
int rel_sibling_position(int cpos, int npos) {
  static const int I = -1;
  static const int tbl[8][6] = {
      {I, 1, I, 2, I, 4}, {0, I, I, 3, I, 5}, {I, 3, 0, I, I, 6}, {2, I, 1, I, I, 7},
      {I, 5, I, 6, 0, I}, {4, I, I, 7, 1, I}, {I, 7, 4, I, 2, I}, {6, I, 5, I, 3, I}};
  return tbl[cpos][npos];
}

// geometry/implicit.hpp:39-103 (+ adaptors::Inverted :119-132)
Num signed_distance(const GeomDesc& g, const Num* x, int nd) {
  Num r;
  switch (g.kind) {
    case GEOM_EDGE: {  // implicit.hpp:101-102: normal.dot(x - point)
      Num t[3];
      for (int d = 0; d < nd; ++d) t[d] = g.b[d] * (x[d] - g.a[d]);
      r = esum(t, nd);
      break;
    }
    case GEOM_SQUARE: {  // implicit.hpp:57-84
      Num dist[3], l2[3];
      for (int d = 0; d < nd; ++d) l2[d] = 0.5 * g.b[d];
      for (int d = 0; d < nd; ++d) dist[d] = g.a[d] - l2[d] - x[d];
      for (int d = 0; d < nd; ++d)
        if (x[d] > g.a[d]) dist[d] = x[d] - (g.a[d] + l2[d]);
      int npos_ = 0;
      for (int d = 0; d < nd; ++d) npos_ += dist[d] > 0 ? 1 : 0;
      if (npos_ == nd) {
        Num s[3];
        for (int d = 0; d < nd; ++d) s[d] = dist[d] * dist[d];
        r = std::sqrt(esum(s, nd));
      } else if (npos_ < 2) {
        r = dist[0];
        for (int d = 1; d < nd; ++d) r = std::max(r, dist[d]);
      } else if (nd == 3) {
        if (dist[0] < 0) r = std::sqrt(std::pow(dist[1], 2) + std::pow(dist[2], 2));
        else if (dist[1] < 0) r = std::sqrt(std::pow(dist[0], 2) + std::pow(dist[2], 2));
        else if (dist[2] < 0) r = std::sqrt(std::pow(dist[0], 2) + std::pow(dist[1], 2));
        else r = invalid_num();  // reference TERMINATEs here (implicit.hpp:83)
      } else {
        r = invalid_num();
      }
      break;
    }
    case GEOM_SPHERE: {  // implicit.hpp:44-45: (x - xc).norm() - r
      Num s[3];
      for (int d = 0; d < nd; ++d) { Num t = x[d] - g.a[d]; s[d] = t * t; }
      r = std::sqrt(esum(s, nd)) - g.b[0];
      break;
    }
    default:
      r = invalid_num();
  }
  return g.invert ? -1 * r : r;
}

// grid/grid.hpp:108-132 RootCell(min, max)
Grid::Grid(int nd_, const Num* xmin, const Num* xmax, int max_solvers_)
    : nd(nd_), max_solvers(max_solvers_) {
  root_length = std::numeric_limits<Num>::epsilon();
  for (int d = 0; d < nd; ++d) {
    const Num lengthD = xmax[d] - xmin[d];
    root_length = std::max(root_length, lengthD);
    root_coords[d] = xmin[d] + 0.5 * lengthD;
  }
  // implementation.hpp:848-856: root node
  parent.push_back(kInvalidInd);
  first_child.push_back(kInvalidInd);
  node2cells.assign(max_solvers, std::vector<Ind>(1, kInvalidInd));
}

// implementation.hpp:255-259.  For a leaf the reference adds pos to
// invalid<Ind>() (signed overflow); only pos 0 of a leaf is well defined.
// The oracle returns invalid for any child of a leaf.
Ind Grid::child(Ind n, int pos) const {
  const Ind fc = first_child[n];
  return valid(fc) ? fc + pos : kInvalidInd;
}

// implementation.hpp:320-364 refine_node: children appended at node_end()
Ind Grid::refine(Ind n) {
  if (!is_leaf(n)) throw std::runtime_error("refine: node is not a leaf");
  const Ind first = size();
  for (int c = 0; c < nchild(); ++c) {
    parent.push_back(n);
    first_child.push_back(kInvalidInd);
    for (auto& m : node2cells) m.push_back(kInvalidInd);
  }
  first_child[n] = first;
  return first;
}

// implementation.hpp:427-441
int Grid::level(Ind n) const {
  int l = 0;
  n = parent[n];
  while (valid(n)) { n = parent[n]; ++l; }
  return l;
}

// implementation.hpp:271-280 (linear search over the parent's children)
int Grid::position_in_parent(Ind n) const {
  const Ind p = parent[n];
  for (int c = 0; c < nchild(); ++c)
    if (child(p, c) == n) return c;
  return nchild();
}

// implementation.hpp:457-543 find_samelvl_neighbor
Ind Grid::find_samelvl_neighbor(Ind n, int pos) const {
  if (is_root(n)) return kInvalidInd;
  int traversed[32];
  int ntrav = 0;
  Ind cur = n;
  while (true) {
    const Ind p = parent[cur];
    const int pip = position_in_parent(cur);
    if (rel_sibling_position(pip, pos) >= 0) break;
    if (is_root(p)) return kInvalidInd;
    traversed[ntrav++] = pip;
    cur = p;
  }
  cur = child(parent[cur], rel_sibling_position(position_in_parent(cur), pos));
  for (int i = ntrav - 1; i >= 0; --i) {
    const int next_pos = rel_sibling_position(traversed[i], opposite_position(pos));
    const Ind next = child(cur, next_pos);
    if (!valid(next)) return kInvalidInd;
    cur = next;
  }
  return cur;
}

// grid.hpp:344-346: rootCell_.length / std::pow(2, level)
Num Grid::cell_length_at_level(Ind lvl) const {
  return root_length / std::pow(2, lvl);
}

// grid.hpp:358-374 (recursive; relLength = 0.25 * parent length)
void Grid::cell_coordinates(Ind n, Num* x) const {
  const int lvl = level(n);
  if (lvl != 0) {
    const Ind p = parent[n];
    const int pip = position_in_parent(n);
    const Num relLength = 0.25 * cell_length(p);
    Num xp[3];
    cell_coordinates(p, xp);
    for (int d = 0; d < nd; ++d)
      x[d] = xp[d] + static_cast<Num>(child_rel_pos(nd, pip, d)) * relLength;
  } else {
    for (int d = 0; d < nd; ++d) x[d] = root_coords[d];
  }
}

// grid.hpp:379-390
void Grid::neighbor_coordinates(Ind n, int pos, Num* x) const {
  Num xn[3];
  cell_coordinates(n, xn);
  const Num length = cell_length(n);
  for (int d = 0; d < nd; ++d) x[d] = xn[d] + nghbr_rel_pos(nd, pos, d) * length;
}

// grid.hpp:601-626 (vertices from cell_vertices_coords :414-434)
bool Grid::is_cut_by(Ind n, const GeomDesc& g) const {
  Num xc[3];
  cell_coordinates(n, xc);
  const Num half = 0.5 * cell_length(n);
  bool all_pos = true, all_neg = true;
  for (int v = 0; v < nchild(); ++v) {
    Num xv[3];
    for (int d = 0; d < nd; ++d) xv[d] = xc[d] + half * static_cast<Num>(child_rel_pos(nd, v, d));
    const Num s = signed_distance(g, xv, nd);
    all_pos = all_pos && (s > 0);
    all_neg = all_neg && (s < 0);
  }
  return !(all_pos || all_neg);
}

// grid/generation.hpp:45-59 MinLevel: passes over the nodes that exist at the
// start of each pass, refining leaves whose level differs from the target.
void Grid::min_level(int target) {
  bool done = false;
  while (!done) {
    done = true;
    const Ind end = size();
    for (Ind n = 0; n < end; ++n) {
      if (is_leaf(n) && level(n) != target) {
        refine(n);
        done = false;
      }
    }
  }
}

// Extension (no reference counterpart): refine leaves whose cell is cut by
// the band |sdf| <= band around g, up to max_level, in the same pass
// structure as MinLevel (children appended at the end of the node list).
void Grid::refine_region(int max_level, const GeomDesc& g, Num band) {
  bool done = false;
  while (!done) {
    done = true;
    const Ind end = size();
    for (Ind n = 0; n < end; ++n) {
      if (!is_leaf(n) || level(n) >= max_level) continue;
      Num xc[3];
      cell_coordinates(n, xc);
      const Num h = cell_length(n);
      const Num s = signed_distance(g, xc, nd);
      // cell-centre distance within band + half diagonal
      const Num reach = band + 0.5 * h * std::sqrt(static_cast<Num>(nd));
      if (std::fabs(s) <= reach) { refine(n); done = false; }
    }
  }
}

// Extension: 2:1 balance.  A pass flags, on the tree as it stands at the
// start of the pass, every leaf that is the coarser face neighbour of a
// refined node two or more levels finer, then refines the flagged leaves in
// increasing node order (children appended at the end, as in MinLevel).
// Passes repeat until nothing is flagged.
void Grid::balance_2to1() {
  bool done = false;
  while (!done) {
    done = true;
    const Ind end = size();
    std::vector<uint8_t> flag(end, 0);
    for (Ind n = 0; n < end; ++n) {
      // n has children; every face neighbour of n at its own level must exist
      if (is_leaf(n) || is_root(n)) continue;
      for (int pos = 0; pos < npos(); ++pos) {
        if (valid(find_samelvl_neighbor(n, pos))) continue;
        // walk up to the coarser leaf covering that position
        for (Ind cur = parent[n]; valid(cur); cur = parent[cur]) {
          const Ind nbp = find_samelvl_neighbor(cur, pos);
          if (valid(nbp)) {
            if (is_leaf(nbp)) flag[nbp] = 1;
            break;
          }
        }
      }
    }
    for (Ind n = 0; n < end; ++n)
      if (flag[n]) { refine(n); done = false; }
  }
}

uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// U[0,1) with 53 random bits, keyed by (seed, stream, counter)
Num prng_uniform(uint64_t seed, uint64_t stream, uint64_t counter) {
  const uint64_t k = splitmix64(seed ^ splitmix64(stream ^ splitmix64(counter)));
  return static_cast<Num>(k >> 11) * (1.0 / 9007199254740992.0);
}

uint64_t node_morton(const Grid& g, Ind n) {
  uint64_t m = 0;
  int shift = 0;
  while (!g.is_root(n)) {
    m |= static_cast<uint64_t>(g.position_in_parent(n)) << shift;
    shift += g.nd;
    n = g.parent[n];
  }
  return m;
}

uint64_t morton_encode(int nd, const uint32_t* ijk) {
  uint64_t m = 0;
  for (int b = 0; b < 21; ++b)
    for (int d = 0; d < nd; ++d)
      m |= static_cast<uint64_t>((ijk[d] >> b) & 1u) << (b * nd + d);
  return m;
}

}  // namespace orc
