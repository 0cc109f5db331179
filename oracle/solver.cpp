// CPU ORACLE — test infrastructure only (see oracle.hpp).
// Restates src/solver/fv/solver.hpp and the euler/cns/heat physics,
// boundary conditions and coupling, in the reference's operation order.
#include "oracle.hpp"
#include <algorithm>
#include <stdexcept>

namespace orc {

// ---------------------------------------------------------------------------
// Quantities: euler/quantities.hpp:131-145, cns/quantities.hpp:23-30, 113-115
// ---------------------------------------------------------------------------
Quantities::Quantities(const PhysDesc& p)
    : gamma(p.gamma), Minfty(p.Minfty), ReInfty(p.ReInfty), Pr0(p.Pr0) {
  sOverTref = (p.T0 != 0) ? p.Sutherland0 / p.T0 : 0.0;
  sOverTrefP1 = sOverTref + 1.0;
  Re0 = (p.physics == PHYS_CNS) ? ReInfty * mu_infinity() / (rho_infinity() * u_infinity()) : 0.0;
}
// boost::units pow<-1>(x) = 1/x and pow<2>(x) = x*x (static integer powers)
Num Quantities::T_infinity() const { return 1. / (1 + 0.5 * (gammaM1()) * (Minfty * Minfty)); }
Num Quantities::rho_infinity() const { return std::pow(T_infinity(), 1. / gammaM1()); }
Num Quantities::u_infinity() const { return Minfty * std::sqrt(T_infinity()); }
Num Quantities::p_infinity() const { return std::pow(T_infinity(), gamma / gammaM1()) / gamma; }

// ---------------------------------------------------------------------------
// Solver construction / storage (container.hpp:125-172)
// ---------------------------------------------------------------------------
Solver::Solver(Grid& g, int sidx, const PhysDesc& p)
    : grid(g), solver_idx(sidx), P(p), qty(p) {
  nd = p.nd;
  nvars = (p.physics == PHYS_HEAT) ? 1 : nd + 2;
  npos = 2 * nd;
  cap = p.max_cells > 0 ? p.max_cells : 1024;
}

static void grow_cols(std::vector<Num>& a, int ncols, Ind oldcap, Ind newcap, Num fill) {
  std::vector<Num> b((size_t)ncols * newcap, fill);
  for (int k = 0; k < ncols; ++k)
    std::copy(a.begin() + (size_t)k * oldcap, a.begin() + (size_t)k * oldcap + oldcap,
              b.begin() + (size_t)k * newcap);
  a.swap(b);
}
static void grow_cols(std::vector<Ind>& a, int ncols, Ind oldcap, Ind newcap, Ind fill) {
  std::vector<Ind> b((size_t)ncols * newcap, fill);
  for (int k = 0; k < ncols; ++k)
    std::copy(a.begin() + (size_t)k * oldcap, a.begin() + (size_t)k * oldcap + oldcap,
              b.begin() + (size_t)k * newcap);
  a.swap(b);
}

// sequential/implementation.hpp:246-251 push_cell + container.hpp:157-172 reset_cell
Ind Solver::push_cell() {
  if (lhs.empty()) {
    lhs.assign((size_t)nvars * cap, invalid_num());
    rhs.assign((size_t)nvars * cap, invalid_num());
    stg.assign((size_t)nvars * cap, invalid_num());
    x_center.assign((size_t)nd * cap, invalid_num());
    distances.assign((size_t)npos * cap, invalid_num());
    length.assign(cap, invalid_num());
    neighbors.assign((size_t)npos * cap, kInvalidInd);
    node_idx.assign(cap, kInvalidInd);
    bc_idx.assign(cap, kInvalidSInd);
  }
  if (ncells == cap) {
    if (P.max_cells > 0) throw std::runtime_error("maxNoCells exceeded");
    const Ind nc = cap * 2;
    grow_cols(lhs, nvars, cap, nc, invalid_num());
    grow_cols(rhs, nvars, cap, nc, invalid_num());
    grow_cols(stg, nvars, cap, nc, invalid_num());
    grow_cols(x_center, nd, cap, nc, invalid_num());
    grow_cols(distances, npos, cap, nc, invalid_num());
    grow_cols(length, 1, cap, nc, invalid_num());
    grow_cols(neighbors, npos, cap, nc, kInvalidInd);
    grow_cols(node_idx, 1, cap, nc, kInvalidInd);
    bc_idx.resize(nc, kInvalidSInd);
    cap = nc;
  }
  return ncells++;
}

// solver.hpp:133-138
void Solver::initialize() {
  // solver.hpp:867-874 init_solver_variables
  dt_ = 0; time_ = 0; tEnd_ = P.time_end; step_ = 0; forced_dt_ = 0; forced_dt_step_ = kInvalidInd;
  create_local_cells();
  if (nonuniform && (P.physics != PHYS_EULER || P.flux != FLUX_AUSM || P.recon != RECON_NONE))
    throw std::runtime_error("refined grids support first-order Euler AUSM only");
  impose_initial_condition();
  apply_perturbation();
  set_dt();
}

// solver.hpp:876-930
void Solver::create_local_cells() {
  Num xc[3];
  Ind leaf_rank = 0;
  for (Ind n = 0; n < grid.size(); ++n) {
    if (!grid.is_leaf(n)) continue;
    grid.cell_coordinates(n, xc);
    bool in = true;
    if (!domain_mask.empty()) in = domain_mask[leaf_rank] != 0;
    else if (domain_sign > 0) in = signed_distance(domain_geom, xc, nd) > 0.;
    else if (domain_sign < 0) in = signed_distance(domain_geom, xc, nd) < 0.;
    ++leaf_rank;
    if (!in) continue;
    const Ind c = push_cell();
    node_idx[c] = n;
    grid.node2cells[solver_idx][n] = c;
    for (int d = 0; d < nd; ++d) x_center[(Ind)d * cap + c] = xc[d];
    length[c] = grid.cell_length(n);
  }
  // neighbours: node neighbours -> local cell ids (solver.hpp:899-908).
  // On a uniform grid every neighbour is a same-level leaf or missing, as in
  // the reference.  2:1 refinement extension: a same-level neighbour that
  // has children gives a face towards its 2^(nd-1) children on the shared
  // face; a missing same-level neighbour whose parent-level neighbour is a
  // leaf gives a face towards that coarser leaf.
  const auto& n2c = grid.node2cells[solver_idx];
  fine.clear();
  nonuniform = false;
  {
    int lvl0 = -1;
    for (Ind c = 0; c < ncells; ++c) {
      const int l = grid.level(node_idx[c]);
      if (lvl0 < 0) lvl0 = l;
      else if (l != lvl0) nonuniform = true;
    }
  }
  for (Ind c = 0; c < ncells; ++c) {
    const Ind n = node_idx[c];
    for (int pos = 0; pos < npos; ++pos) {
      const Ind nn = grid.find_samelvl_neighbor(n, pos);
      Ind val = kInvalidInd;
      if (valid(nn) && grid.is_leaf(nn)) {
        val = n2c[nn];
      } else if (valid(nn)) {
        // finer neighbours: children of nn whose bit along d faces c
        const int d = pos / 2, want = (pos & 1) ? 0 : 1;
        Ind ids[4];
        int k = 0, nvalid = 0;
        for (int cp = 0; cp < grid.nchild(); ++cp) {
          if (((cp >> d) & 1) != want) continue;
          const Ind ch = grid.child(nn, cp);
          if (!grid.is_leaf(ch)) throw std::runtime_error("grid is not 2:1 balanced");
          ids[k++] = n2c[ch];
          nvalid += valid(n2c[ch]) ? 1 : 0;
        }
        if (nvalid == k) {
          val = -(2 + (Ind)(fine.size() / nfine_per()));
          for (int j = 0; j < k; ++j) fine.push_back(ids[j]);
        } else if (nvalid != 0) {
          throw std::runtime_error("domain boundary cuts a refined face");
        }
      } else if (!grid.is_root(n)) {
        // coarser neighbour: the parent's same-level neighbour, if a leaf
        const Ind pn = grid.find_samelvl_neighbor(grid.parent[n], pos);
        const int d = pos / 2, side = pos & 1;
        const bool on_face = ((grid.position_in_parent(n) >> d) & 1) == side;
        if (on_face && valid(pn) && grid.is_leaf(pn)) val = n2c[pn];
      }
      nghbr(c, pos) = val;
    }
  }
  firstGC = ncells;
  create_ghost_cells();
  sort_gc();
  // distances (solver.hpp:918-928)
  for (Ind c = 0; c < ncells; ++c)
    for (int pos = 0; pos < npos; ++pos) {
      const Ind n = nghbr(c, pos);
      if (!valid(n) || is_fine_ref(n)) continue;
      distances[(Ind)pos * cap + c] = cell_dx(c, n);
    }
}

// solver.hpp:935-977
void Solver::create_ghost_cells() {
  const Ind nleaf = ncells;
  SInd bcid = 0;
  for (auto& bc : bcs) {
    for (Ind c = 0; c < nleaf; ++c) {
      const Ind n = node_idx[c];
      if (!valid(n)) continue;
      if (!grid.is_cut_by(n, bc.geom)) continue;
      int missing[6], nm = 0;
      for (int pos = 0; pos < npos; ++pos)
        if (nghbr(c, pos) == kInvalidInd) missing[nm++] = pos;
      for (int k = 0; k < nm; ++k) {
        Num xn[3];
        grid.neighbor_coordinates(n, missing[k], xn);
        if (signed_distance(bc.geom, xn, nd) < 0.) {
          const Ind g = create_ghost_cell(c, missing[k]);
          bc_idx[g] = bcid;
        }
      }
    }
    ++bcid;
  }
}

// solver.hpp:843-861
Ind Solver::create_ghost_cell(Ind b, int pos) {
  const Ind g = push_cell();
  node_idx[g] = kInvalidInd;
  nghbr(b, pos) = g;
  nghbr(g, opposite_position(pos)) = b;
  // ghost_cell_coordinates (solver.hpp:787-807): x_b + len_b * rel(pos)
  const Num len = length[b];
  for (int d = 0; d < nd; ++d)
    x_center[(Ind)d * cap + g] = x_center[(Ind)d * cap + b] + len * static_cast<Num>(nghbr_rel_pos(nd, pos, d));
  length[g] = length[b];
  return g;
}

// solver.hpp:811-834.  ghost_order 0: canonical (creation order, which is
// already sorted by bc_idx; a stable sort is the identity).  ghost_order 1:
// this toolchain's std::sort with the reference's comparator.
void Solver::sort_gc() {
  Ind first = kInvalidInd;
  for (Ind c = 0; c < ncells; ++c)
    if (bc_idx[c] != kInvalidSInd) { first = c; break; }
  if (!valid(first)) return;
  if (ghost_order == 1) {
    const Ind ng = ncells - first;
    std::vector<Ind> perm(ng);
    for (Ind i = 0; i < ng; ++i) perm[i] = first + i;
    std::sort(perm.begin(), perm.end(), [&](Ind a, Ind b) { return bc_idx[a] < bc_idx[b]; });
    auto permute = [&](auto& arr, int ncols) {
      using T = typename std::decay<decltype(arr[0])>::type;
      std::vector<T> tmp((size_t)ng);
      for (int k = 0; k < ncols; ++k) {
        for (Ind i = 0; i < ng; ++i) tmp[i] = arr[(size_t)k * cap + perm[i]];
        for (Ind i = 0; i < ng; ++i) arr[(size_t)k * cap + first + i] = tmp[i];
      }
    };
    permute(lhs, nvars); permute(rhs, nvars); permute(x_center, nd);
    permute(distances, npos); permute(length, 1); permute(neighbors, npos);
    permute(node_idx, 1);
    std::vector<SInd> t(ng);
    for (Ind i = 0; i < ng; ++i) t[i] = bc_idx[perm[i]];
    for (Ind i = 0; i < ng; ++i) bc_idx[first + i] = t[i];
  }
  for (Ind g = first; g < ncells; ++g)
    for (int pos = 0; pos < npos; ++pos) {
      const Ind n = nghbr(g, pos);
      if (valid(n)) nghbr(n, opposite_position(pos)) = g;
    }
}

// ---------------------------------------------------------------------------
// Initial conditions
// ---------------------------------------------------------------------------
// euler/quantities.hpp:153-166 cv
void Solver::cv_of(const Num* pvs, Num* cvs) const {
  const Num r = pvs[nd];
  cvs[nd] = r;
  Num velMag2 = 0;
  for (int d = 0; d < nd; ++d) {
    cvs[d] = r * pvs[d];
    velMag2 += std::pow(pvs[d], 2);
  }
  cvs[nd + 1] = pvs[nd + 1] / (qty.gamma - 1.0) + 0.5 * r * velMag2;
}
// euler/quantities.hpp:173-186 pv
void Solver::pv_of(const Num* cvs, Num* pvs) const {
  pvs[nd] = cvs[nd];
  Num velMag2 = 0;
  for (int d = 0; d < nd; ++d) {
    pvs[d] = cvs[d] / cvs[nd];
    velMag2 += std::pow(pvs[d], 2);
  }
  pvs[nd + 1] = (qty.gamma - 1.0) * (cvs[nd + 1] - 0.5 * cvs[nd] * velMag2);
}

static void cv_gamma(int nd, const Num* pvs, Num gamma, Num* cvs) {
  const Num r = pvs[nd];
  cvs[nd] = r;
  Num velMag2 = 0;
  for (int d = 0; d < nd; ++d) {
    cvs[d] = r * pvs[d];
    velMag2 += std::pow(pvs[d], 2);
  }
  cvs[nd + 1] = pvs[nd + 1] / (gamma - 1.0) + 0.5 * r * velMag2;
}

// euler/initial_conditions.hpp:19-45 isentropic_vortex (2D only)
static void isentropic_vortex(const Num* x, Num t, Num cx, Num cy, Num* out) {
  const Num beta = 5;
  const Num gamma = 1.4;
  const Num x0[2] = {cx, cy};
  const Num u = 1, v = 1;
  const Num x_t = x[0] - u * t;
  const Num y_t = x[1] - v * t;
  const Num r = std::sqrt(std::pow(x_t - x0[0], 2) + std::pow(y_t - x0[1], 2));
  const Num eT1mr2 = std::exp(1 - std::pow(r, 2));
  const Num eT21mr2 = std::exp(2 * (1 - std::pow(r, 2)));
  Num pv[4];
  pv[0] = u - beta * eT1mr2 * (y_t - x0[1]) / (2 * kPi);
  pv[1] = v + beta * eT1mr2 * (x_t - x0[0]) / (2 * kPi);
  pv[2] = std::pow((1 - ((gamma - 1) * std::pow(beta, 2) * eT21mr2) / (16 * gamma * std::pow(kPi, 2))),
                   1 / (gamma - 1));
  pv[3] = std::pow(pv[2], gamma);
  cv_gamma(2, pv, 1.4, out);
}

void Solver::ic_point(const Num* x, Num* out) const {
  if (ic_kind < 0 || ic_kind > IC_CALLBACK) throw std::runtime_error("unknown IC");
  const int min_params[] = {nvars, 9, 1, nd + 1 + 2 * nvars, 3, 2 + 4 * nvars, 1, 4, nvars + nvars * nd, 0};
  if ((int)ic_params.size() < min_params[ic_kind] || (ic_kind == IC_CALLBACK && !ic_cb))
    throw std::runtime_error("initial condition parameters missing");
  const Num* P_ = ic_params.data();
  switch (ic_kind) {
    case IC_CONSTANT_PV: cv_of(P_, out); return;
    case IC_SHOCK_TUBE: {  // euler/initial_conditions.hpp:57-94
      const int dir = (int)P_[0];
      const Num angle = P_[1], x0 = P_[2];
      const Num rhoL = P_[3], umagL = P_[4], pL = P_[5], rhoR = P_[6], umagR = P_[7], pR = P_[8];
      const Num alpha = angle * (2 * kPi) / 360;
      const Num gamma = 1.4;
      Num x_rot[3] = {0, 0, 0};
      x_rot[0] = std::cos(alpha) * x[0] - std::sin(alpha) * x[1];
      x_rot[1] = std::sin(alpha) * x[0] + std::cos(alpha) * x[1];
      Num uL[3] = {0, 0, 0}, uR[3] = {0, 0, 0};
      uL[0] = umagL * std::cos(alpha); uL[1] = umagL * std::sin(alpha);
      uR[0] = umagR * std::cos(alpha); uR[1] = umagR * std::sin(alpha);
      const Num pos = (alpha == 0) ? x0 : x0 * std::sqrt(2.);
      Num pv[5];
      if (x_rot[dir] < pos) {
        pv[nd] = rhoL; for (int d = 0; d < nd; ++d) pv[d] = uL[d]; pv[nd + 1] = pL;
      } else {
        pv[nd] = rhoR; for (int d = 0; d < nd; ++d) pv[d] = uR[d]; pv[nd + 1] = pR;
      }
      cv_gamma(nd, pv, gamma, out);
      return;
    }
    case IC_ISENTROPIC_VORTEX: {
      const Num t = P_[0];
      const Num cx = ic_params.size() > 2 ? P_[1] : -5.0;
      const Num cy = ic_params.size() > 2 ? P_[2] : -5.0;
      isentropic_vortex(x, t, cx, cy, out);
      return;
    }
    case IC_SPHERE: {  // euler_fv_solver_test.cpp:222-245 (explosion) / blast
      // params: centre[nd], radius, pv_inside[nvars], pv_outside[nvars]
      Num s[3];
      for (int d = 0; d < nd; ++d) { const Num t = x[d] - P_[d]; s[d] = t * t; }
      const Num r = std::sqrt(esum(s, nd));
      const Num* pv = (r < P_[nd]) ? P_ + nd + 1 : P_ + nd + 1 + nvars;
      cv_of(pv, out);
      return;
    }
    case IC_TGV: {  // params: V0, rho_inf, p_inf (3D, [0,2pi]^3)
      const Num V0 = P_[0], rinf = P_[1], pinf = P_[2];
      Num pv[5];
      pv[0] = V0 * std::sin(x[0]) * std::cos(x[1]) * std::cos(x[2]);
      pv[1] = -V0 * std::cos(x[0]) * std::sin(x[1]) * std::cos(x[2]);
      pv[2] = 0.0;
      pv[3] = rinf;
      pv[4] = pinf + rinf * V0 * V0 / 16.0 * (std::cos(2 * x[0]) + std::cos(2 * x[1])) * (std::cos(2 * x[2]) + 2.0);
      cv_of(pv, out);
      return;
    }
    case IC_FOUR_QUADRANT: {  // params: xs, ys, then 4 x pv[nvars] (q0: x<xs,y<ys; q1: x>=xs,y<ys; q2; q3)
      const int q = (x[0] >= P_[0] ? 1 : 0) + (x[1] >= P_[1] ? 2 : 0);
      cv_of(P_ + 2 + q * nvars, out);
      return;
    }
    case IC_HEAT_CONSTANT: out[0] = P_[0]; return;
    case IC_HEAT_SINE: {  // T = a + b * prod_d sin(k (x_d - x0))
      Num v = P_[1];
      for (int d = 0; d < nd; ++d) v *= std::sin(P_[2] * (x[d] - P_[3]));
      out[0] = P_[0] + v;
      return;
    }
    case IC_LINEAR_PV: {  // pv_v = pv0_v + sum_d G[v][d] x_d
      Num pv[5];
      for (int v = 0; v < nvars; ++v) {
        Num a = P_[v];
        for (int d = 0; d < nd; ++d) a += P_[nvars + v * nd + d] * x[d];
        pv[v] = a;
      }
      if (P.physics == PHYS_HEAT) out[0] = pv[0]; else cv_of(pv, out);
      return;
    }
    case IC_CALLBACK: {
      // the callback takes doubles whatever the oracle computes in
      double xd[3] = {0, 0, 0}, od[5] = {0, 0, 0, 0, 0};
      for (int d = 0; d < nd; ++d) xd[d] = (double)x[d];
      ic_cb(xd, od, ic_ctx);
      for (int v = 0; v < nvars; ++v) out[v] = od[v];
      return;
    }
    default: throw std::runtime_error("unknown IC");
  }
}

// solver.hpp:980-992 + quadrature/quadrature.hpp:14-74
void Solver::impose_initial_condition() {
  const int ngp = 1 << nd;
  const Num gpv = 1 / std::sqrt(3);
  for (Ind c = 0; c < n_internal(); ++c) {
    const Num len = grid.cell_length(node_idx[c]);
    const Num dx2 = 0.5 * len;
    Num result[5] = {0, 0, 0, 0, 0};
    Num vals[8][5];
    for (int gp = 0; gp < ngp; ++gp) {
      Num xg[3];
      for (int d = 0; d < nd; ++d) {
        const Num g = ((gp >> d) & 1) ? gpv : -1 / std::sqrt(3);
        xg[d] = x_center[(Ind)d * cap + c] + dx2 * g;
      }
      ic_point(xg, vals[gp]);
    }
    for (int gp = 0; gp < ngp; ++gp)
      for (int v = 0; v < nvars; ++v) result[v] = result[v] + vals[gp][v];
    // geometry/cell_cartesian.hpp: volume<nd>(dx) = dx * volume<nd-1>(dx)
    const Num V = (nd == 2) ? len * len : len * (len * len);
    for (int v = 0; v < nvars; ++v) {
      const Num integ = result[v] * V / static_cast<Num>(ngp);
      lhs[(Ind)v * cap + c] = integ / V;
    }
  }
}

// Seeded density perturbation (extension): rho *= 1 + amp * (2U - 1) at
// fixed velocity and pressure, U = prng_uniform(seed, level, Morton index of
// the cell's node at its level), so the field does not depend on cell order.
void Solver::apply_perturbation() {
  if (perturb_amp == 0 || P.physics == PHYS_HEAT) return;
  for (Ind c = 0; c < n_internal(); ++c) {
    const Ind n = node_idx[c];
    const Num U = prng_uniform(perturb_seed, (uint64_t)grid.level(n), node_morton(grid, n));
    const Num eps = perturb_amp * (2 * U - 1);
    Num cvs[5], pvs[5];
    for (int v = 0; v < nvars; ++v) cvs[v] = Q(lhs.data(), c, v);
    pv_of(cvs, pvs);
    pvs[nd] = pvs[nd] * (1 + eps);
    cv_of(pvs, cvs);
    for (int v = 0; v < nvars; ++v) Q(lhs.data(), c, v) = cvs[v];
  }
}

// ---------------------------------------------------------------------------
// Physics helpers (euler/physics.hpp:83-149)
// ---------------------------------------------------------------------------
Num Solver::u_mag2(const Num* Qb, Ind c) const {
  Num t[3];
  for (int d = 0; d < nd; ++d) { const Num ud = u(Qb, c, d); t[d] = ud * ud; }
  return esum(t, nd);  // Eigen squaredNorm
}
Num Solver::p(const Num* Qb, Ind c) const {
  return qty.gammaM1() * (rho_E(Qb, c) - 0.5 * rho(Qb, c) * u_mag2(Qb, c));
}
Num Solver::a(const Num* Qb, Ind c) const { return std::sqrt(qty.gamma * p(Qb, c) / rho(Qb, c)); }
Num Solver::T(const Num* Qb, Ind c) const {
  if (P.physics == PHYS_HEAT) return Q(Qb, c, 0);
  return qty.gamma * p(Qb, c) / rho(Qb, c);
}

// solver.hpp:590-592 + geometry/algorithms.hpp:10-11
Num Solver::cell_dx(Ind a_, Ind b_) const {
  Num t[3];
  for (int d = 0; d < nd; ++d) {
    const Num s = x_center[(Ind)d * cap + a_] - x_center[(Ind)d * cap + b_];
    t[d] = s * s;
  }
  return std::sqrt(esum(t, nd));
}

// solver.hpp:627-636
int Solver::which_neighbor(Ind c, Ind n) const {
  for (int pos = 0; pos < npos; ++pos)
    if (nghbr(c, pos) == n) return pos;
  return kInvalidSInd;
}

// solver.hpp:298-313
Solver::BndryInfo Solver::boundary_info(Ind g) const {
  int bpos = kInvalidSInd;
  Ind bidx = kInvalidInd;
  for (int pos = 0; pos < npos; ++pos)
    if (valid(nghbr(g, pos))) { bpos = pos; bidx = nghbr(g, pos); break; }
  return {bidx, bpos, g, opposite_position(bpos)};
}

// Periodic image of a ghost: the internal cell at the far end of the
// boundary cell's row, opposite to the ghost (extension).
Ind Solver::periodic_image(Ind g) const {
  const BndryInfo bi = boundary_info(g);
  Ind cur = bi.bndryIdx;
  const int dir = bi.bndryPos;  // position of bndry w.r.t. ghost = walking direction
  while (true) {
    const Ind n = nghbr(cur, dir);
    if (!valid(n) || is_fine_ref(n) || is_ghost(n)) return cur;
    cur = n;
  }
}

// solver.hpp:596-623
Num Solver::slope(const Num* Qb, Ind c, int v, int dir) const {
  const Ind nn = nghbr(c, 2 * dir);
  const Ind np = nghbr(c, 2 * dir + 1);
  if (valid(nn) && valid(np)) return (Q(Qb, np, v) - Q(Qb, nn, v)) / (2. * length[c]);
  if (valid(nn)) return slope(Qb, nn, v, dir);
  if (valid(np)) return slope(Qb, np, v, dir);
  const BndryInfo bi = boundary_info(c);
  return bc_slope(Qb, bcs[bc_idx[c]], c, bi.bndryIdx, v, dir);
}

// boundary_condition.hpp:63-67 default hook; wall overrides
// cns/boundary_conditions.hpp:143-147, 230-234; periodic: image slope.
Num Solver::bc_slope(const Num* Qb, const BC& bc, Ind ghost, Ind bndry, int v, int dir) const {
  switch (bc.kind) {
    case BC_ADIABATIC_NOSLIP:
    case BC_ISOTHERMAL_NOSLIP:
    case BC_CHT_FLUID: {
      const Num tmp = slope(Qb, bndry, v, dir);
      return v < nd ? -tmp : tmp;
    }
    case BC_PERIODIC: return slope(Qb, periodic_image(ghost), v, dir);
    default: return slope(Qb, bndry, v, dir);
  }
}

// solver.hpp:656-678
Num Solver::surface_slope(const Num* Qb, Ind l, Ind r, int v, int sdir) const {
  const int pos = which_neighbor(l, r);
  const int ndir = pos / 2;
  if (ndir == sdir) {
    const Num dx = length[r];
    return (Q(Qb, r, v) - Q(Qb, l, v)) / dx;
  }
  return 0.5 * (slope(Qb, r, v, sdir) + slope(Qb, l, v, sdir));
}

// ---------------------------------------------------------------------------
// Numerical fluxes
// ---------------------------------------------------------------------------
// euler/physics.hpp:297-310
static inline Num m_int(int s_, Num M) {
  const Num s = static_cast<Num>(s_);
  return std::abs(M) > 1 ? 0.5 * (M + s * std::abs(M)) : s * 0.25 * std::pow(M + s, 2.);
}
static inline Num p_int(int s_, Num M) {
  const Num s = static_cast<Num>(s_);
  return std::abs(M) > 1 ? 0.5 * (M + s * std::abs(M)) / M : 0.25 * std::pow(M + s, 2.) * (2. - s * M);
}
Num ausm_m_split(int s, Num M) { return m_int(s, M); }
Num ausm_p_split(int s, Num M) { return p_int(s, M); }

// AUSM of two conserved states (euler/physics.hpp:276-325), used for the
// cell-based flux and for MUSCL-reconstructed face states.
void Solver::ausm_states(const Num* QL, const Num* QR, int d, Num* f) const {
  auto prim = [&](const Num* q, Num& ud, Num& p_, Num& a_) {
    const Num r = q[nd];
    Num t[3];
    for (int k = 0; k < nd; ++k) { const Num uk = q[k] / r; t[k] = uk * uk; }
    p_ = qty.gammaM1() * (q[nd + 1] - 0.5 * r * esum(t, nd));
    a_ = std::sqrt(qty.gamma * p_ / r);
    ud = q[d] / r;
  };
  Num uL, pL, aL, uR, pR, aR;
  prim(QL, uL, pL, aL);
  prim(QR, uR, pR, aR);
  const Num ML = uL / aL;
  const Num MR = uR / aR;
  const Num m_i = m_int(+1, ML) + m_int(-1, MR);
  const Num p_i = p_int(+1, ML) * pL + p_int(-1, MR) * pR;
  const Num* q = m_i >= 0 ? QL : QR;
  const Num pc = m_i >= 0 ? pL : pR;
  const Num ac = m_i >= 0 ? aL : aR;
  // theta_: (rho u, rho, rho E + p) * a
  for (int k = 0; k < nd; ++k) f[k] = q[k] * ac;
  f[nd] = q[nd] * ac;
  f[nd + 1] = (q[nd + 1] + pc) * ac;
  for (int v = 0; v < nvars; ++v) f[v] *= m_i;
  f[d] += p_i;
}

void Solver::ausm(const Num* Qb, Ind l, Ind r, int d, Num* f) const {
  Num QL[5], QR[5];
  for (int v = 0; v < nvars; ++v) { QL[v] = Q(Qb, l, v); QR[v] = Q(Qb, r, v); }
  if (P.recon == RECON_MUSCL) {
    // extension: limited linear reconstruction of conserved variables
    // second neighbours; across a periodic ghost they continue at the image
    Ind ll = nghbr(l, 2 * d);
    Ind rr = nghbr(r, 2 * d + 1);
    if (!valid(ll) && is_ghost(l) && bcs[bc_idx[l]].kind == BC_PERIODIC) ll = nghbr(periodic_image(l), 2 * d);
    if (!valid(rr) && is_ghost(r) && bcs[bc_idx[r]].kind == BC_PERIODIC) rr = nghbr(periodic_image(r), 2 * d + 1);
    auto lim = [&](Num a_, Num b_) -> Num {
      if (P.limiter == LIM_ZERO) return 0.0;
      if (a_ * b_ <= 0) return 0.0;
      if (P.limiter == LIM_MINMOD) return std::fabs(a_) < std::fabs(b_) ? a_ : b_;
      return 2.0 * a_ * b_ / (a_ + b_);
    };
    for (int v = 0; v < nvars; ++v) {
      const Num ql = Q(Qb, l, v), qr = Q(Qb, r, v);
      const Num sl = valid(ll) ? lim(ql - Q(Qb, ll, v), qr - ql) : 0.0;
      const Num sr = valid(rr) ? lim(qr - ql, Q(Qb, rr, v) - qr) : 0.0;
      QL[v] = ql + 0.5 * sl;
      QR[v] = qr - 0.5 * sr;
    }
  }
  ausm_states(QL, QR, d, f);
}

// euler/physics.hpp:98-113, 262-267
void Solver::llf(const Num* Qb, Ind l, Ind r, int d, Num dx, Num dt, Num* f) const {
  auto flux = [&](Ind c, Num* fo) {
    const Num ud = u(Qb, c, d);
    const Num pc = p(Qb, c);
    fo[nd] = rho_u(Qb, c, d);
    for (int i = 0; i < nd; ++i) fo[i] = rho_u(Qb, c, i) * ud;
    fo[d] += pc;
    fo[nd + 1] = ud * (rho_E(Qb, c) + pc);
  };
  Num fl[5], fr[5];
  flux(l, fl);
  flux(r, fr);
  for (int v = 0; v < nvars; ++v)
    f[v] = 0.5 * (fl[v] + fr[v] + dx / dt * (Q(Qb, l, v) - Q(Qb, r, v)));
}

// cns/physics.hpp:221-338 (three_point viscous flux), with the helper
// functions restated one-to-one (values recomputed as in the reference).
void Solver::viscous(const Num* Qb, Ind l, Ind r, int dir, Num* f) const {
  const int RHO = nd, RHOE = nd + 1;
  auto Q_srfc = [&](int v) { return 0.5 * (Q(Qb, l, v) + Q(Qb, r, v)); };
  auto rho_srfc = [&]() { return Q_srfc(RHO); };
  auto rho_u_srfc = [&](int d) { return Q_srfc(d); };
  auto rho_E_srfc = [&]() { return Q_srfc(RHOE); };
  auto u_srfc = [&](int d) { return rho_u_srfc(d) / rho_srfc(); };
  auto umag2_srfc = [&]() {
    Num umag2 = 0;
    for (int d = 0; d < nd; ++d) umag2 += std::pow(u_srfc(d), 2.);
    return umag2;
  };
  auto p_srfc = [&]() { return qty.gammaM1() * (rho_E_srfc() - 0.5 * rho_srfc() * umag2_srfc()); };
  auto grad_rho_srfc = [&](int d) { return surface_slope(Qb, l, r, RHO, d); };
  auto grad_rho_u_srfc = [&](int c, int d) { return surface_slope(Qb, l, r, c, d); };
  auto grad_rho_E_srfc = [&](int d) { return surface_slope(Qb, l, r, RHOE, d); };
  auto grad_u_srfc = [&](int c, int d) {
    const Num rho_s = rho_srfc();
    return (grad_rho_u_srfc(c, d) * rho_s - grad_rho_srfc(d) * rho_u_srfc(c)) / std::pow(rho_s, 2.);
  };
  auto grad_p_srfc = [&](int dd) {
    const Num rho_s = rho_srfc();
    Num grad_rho_umag2 = grad_rho_srfc(dd) * umag2_srfc();
    for (int d = 0; d < nd; ++d) grad_rho_umag2 += rho_s * 2. * u_srfc(d) * grad_u_srfc(d, dd);
    return qty.gammaM1() * (grad_rho_E_srfc(dd) - 0.5 * grad_rho_umag2);
  };
  // mu at the surface: temperature from surface-averaged p and rho
  const Num p_avg = 0.5 * (p(Qb, l) + p(Qb, r));
  const Num rho_avg = 0.5 * (rho(Qb, l) + rho(Qb, r));
  const Num mu_srfc = qty.sutherland(qty.gamma * p_avg / rho_avg);
  const Num gradP = grad_p_srfc(dir);
  const Num gradRho = grad_rho_srfc(dir);
  // heat_flux (cns/physics.hpp:101-105)
  const Num rho_s = rho_srfc();
  const Num ps = p_srfc();
  const Num q = -mu_srfc * qty.gamma / (qty.gammaM1() * qty.Pr0 * rho_s) * (gradP - ps / rho_s * gradRho);
  // shear_stress (cns/physics.hpp:110-123): nabla_u(i,j) = slope_u(j, i)
  Num nab[3][3];
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j) nab[i][j] = grad_u_srfc(j, i);
  Num diag[3];
  for (int i = 0; i < nd; ++i) diag[i] = nab[i][i];
  const Num div_u = esum(diag, nd) * 2. / 3.;
  Num tau_col[3];
  for (int i = 0; i < nd; ++i)
    tau_col[i] = -mu_srfc * (nab[i][dir] + nab[dir][i] - div_u * (i == dir ? 1.0 : 0.0));
  for (int v = 0; v < nvars; ++v) f[v] = 0;
  for (int d = 0; d < nd; ++d) {
    f[d] = 1. / qty.Re0 * tau_col[d];
    f[RHOE] += u_srfc(d) * tau_col[d];
  }
  f[RHOE] = 1. / qty.Re0 * (f[RHOE] + q);
}

// euler/physics.hpp:175-178, cns/physics.hpp:154-161, heat/physics.hpp:67-71, 113-119
void Solver::compute_num_flux(const Num* Qb, Ind l, Ind r, int d, Num dx, Num dt, Num* out) const {
  if (P.physics == PHYS_HEAT) {
    out[0] = 1 / dx * (Q(Qb, l, 0) - Q(Qb, r, 0));
    return;
  }
  if (P.flux == FLUX_LLF) llf(Qb, l, r, d, dx, dt, out);
  else ausm(Qb, l, r, d, out);
  if (P.physics == PHYS_CNS) {
    Num fv[5];
    viscous(Qb, l, r, d, fv);
    for (int v = 0; v < nvars; ++v) out[v] = out[v] + fv[v];
  }
}

// solver.hpp:439-460.  2:1 refinement extension: a face towards finer
// cells is the mean of the sub-face fluxes, summed pairwise
// ((F0 + F1) + (F2 + F3) in 3D, F0 + F1 in 2D) and scaled by 2^(1-nd); the
// fine cells see their full face.  Equal sub-face fluxes sum exactly, so a
// uniform state stays uniform.
void Solver::num_flux(const Num* Qb, Ind c, Num* out) const {
  Num result[5] = {0, 0, 0, 0, 0};
  const Num dx = length[c];
  auto side_flux = [&](Ind other, bool minus, int d, Num* f) {
    if (!is_fine_ref(other)) {
      if (minus) compute_num_flux(Qb, other, c, d, dx, dt_, f);
      else compute_num_flux(Qb, c, other, d, dx, dt_, f);
      return;
    }
    const Ind* fc = &fine[(size_t)fine_index(other) * nfine_per()];
    Num sub[4][5];
    for (int k = 0; k < nfine_per(); ++k) {
      if (minus) compute_num_flux(Qb, fc[k], c, d, dx, dt_, sub[k]);
      else compute_num_flux(Qb, c, fc[k], d, dx, dt_, sub[k]);
    }
    for (int v = 0; v < nvars; ++v) {
      const Num s = (nd == 2) ? sub[0][v] + sub[1][v] : (sub[0][v] + sub[1][v]) + (sub[2][v] + sub[3][v]);
      f[v] = s * ((nd == 2) ? 0.5 : 0.25);
    }
  };
  for (int d = 0; d < nd; ++d) {
    const Ind m = nghbr(c, 2 * d);
    const Ind p_ = nghbr(c, 2 * d + 1);
    Num fm[5], fp[5];
    side_flux(m, true, d, fm);
    side_flux(p_, false, d, fp);
    for (int v = 0; v < nvars; ++v) result[v] += (fm[v] - fp[v]);
  }
  for (int v = 0; v < nvars; ++v) out[v] = dt_ / dx * result[v];
}

// physics compute_source_term: zero in the reference; the heat extension
// returns dt * S_c (per-cell volumetric source).
void Solver::source_term(Ind c, Num* out) const {
  for (int v = 0; v < nvars; ++v) out[v] = 0;
  if (P.physics == PHYS_HEAT && !source.empty()) out[0] = dt_ * source[c];
}

// euler/physics.hpp:182-196, cns/physics.hpp:167-187, heat/physics.hpp:82-85
Num Solver::cell_dt(const Num* Qb, Ind c) const {
  const Num h = length[c];
  if (P.physics == PHYS_HEAT) return P.cfl * 0.5 * std::pow(h, 2.);
  const Num a_ = a(Qb, c);
  Num S_max = std::numeric_limits<Num>::lowest();
  for (int d = 0; d < nd; ++d) S_max = std::max(S_max, std::abs(u(Qb, c, d)) + a_);
  const Num dt_inv = P.cfl * h / S_max;
  if (P.physics == PHYS_EULER) return dt_inv;
  const Num mu_c = qty.sutherland(T(Qb, c));
  const Num S_coeff = mu_c / (rho(Qb, c) * h);
  const Num S_viscous = std::max(S_coeff / qty.Re0, S_coeff / (qty.Re0 * qty.Pr0 * qty.gammaM1()));
  const Num dt_viscous = P.cfl_viscous * h / (2 * S_viscous);
  return std::min(dt_inv, dt_viscous);
}

// solver.hpp:511-517.  The cell loops below carry OpenMP pragmas for the
// threaded timing build only (liboracle_fast.so); every build used as the
// checker is compiled without OpenMP and runs them in order.
Num Solver::compute_dt() const {
  Num dt = std::numeric_limits<Num>::max();
  const Ind ni = n_internal();
#pragma omp parallel for reduction(min : dt) schedule(static)
  for (Ind c = 0; c < ni; ++c) dt = std::min(dt, cell_dt(lhs.data(), c));
  return dt;
}

// solver.hpp:541-551
void Solver::set_dt() {
  if (time_ + dt_ < tEnd_) {
    if (step_ == forced_dt_step_) dt_ = forced_dt_;
    else dt_ = compute_dt();
  } else {
    dt_ = tEnd_ - time_;
  }
}

// solver.hpp:140-145
void Solver::solve() {
  apply_bcs(lhs.data());
  set_dt();
  evolve();
  time_ += dt_;
  ++step_;
}

// solver.hpp:469-494 (+ SSP-RK3 extension)
void Solver::evolve() {
  const Ind ni = n_internal();
  Num nf[5], src[5];
  Num* L = lhs.data();
  Num* R = rhs.data();
  Num* S = stg.data();
  if (P.ti == TI_EULER_FWD) {
#pragma omp parallel for private(nf, src) schedule(static)
    for (Ind c = 0; c < ni; ++c) {
      num_flux(L, c, nf);
      source_term(c, src);
      for (int v = 0; v < nvars; ++v) Q(R, c, v) = Q(L, c, v) + nf[v] + src[v];
    }
#pragma omp parallel for schedule(static)
    for (Ind c = 0; c < ni; ++c)
      for (int v = 0; v < nvars; ++v) Q(L, c, v) = Q(R, c, v);
  } else if (P.ti == TI_HEUN_RK2) {
#pragma omp parallel for private(nf, src) schedule(static)
    for (Ind c = 0; c < ni; ++c) {
      num_flux(L, c, nf);
      for (int v = 0; v < nvars; ++v) Q(R, c, v) = Q(L, c, v) + nf[v];
    }
    apply_bcs(R);
#pragma omp parallel for private(nf, src) schedule(static)
    for (Ind c = 0; c < ni; ++c) {
      num_flux(R, c, nf);
      for (int v = 0; v < nvars; ++v) Q(L, c, v) = 0.5 * (Q(L, c, v) + Q(R, c, v) + nf[v]);
    }
  } else {  // SSP-RK3 (Shu-Osher); stage 1 is the forward-Euler stage
#pragma omp parallel for private(nf, src) schedule(static)
    for (Ind c = 0; c < ni; ++c) {
      num_flux(L, c, nf);
      source_term(c, src);
      for (int v = 0; v < nvars; ++v) Q(R, c, v) = Q(L, c, v) + nf[v] + src[v];
    }
    apply_bcs(R);
#pragma omp parallel for private(nf, src) schedule(static)
    for (Ind c = 0; c < ni; ++c) {
      num_flux(R, c, nf);
      source_term(c, src);
      for (int v = 0; v < nvars; ++v) Q(S, c, v) = 0.75 * Q(L, c, v) + 0.25 * (Q(R, c, v) + nf[v] + src[v]);
    }
    apply_bcs(S);
#pragma omp parallel for private(nf, src) schedule(static)
    for (Ind c = 0; c < ni; ++c) {
      num_flux(S, c, nf);
      source_term(c, src);
      for (int v = 0; v < nvars; ++v)
        Q(L, c, v) = (1. / 3.) * Q(L, c, v) + (2. / 3.) * (Q(S, c, v) + nf[v] + src[v]);
    }
  }
}

// ---------------------------------------------------------------------------
// Boundary conditions (solver.hpp:561-586 + per-physics Conditions)
// ---------------------------------------------------------------------------
Ind Solver::cht_partner(Ind ghost, const BC& bc) const {
  // coupling.hpp:24-39 local_bndry_ids
  const BndryInfo bi = boundary_info(ghost);
  const Ind node0 = node_idx[bi.bndryIdx];
  const Ind node1 = grid.find_samelvl_neighbor(node0, bi.ghostPos);
  if (!valid(node1)) throw std::runtime_error("CHT partner node not found");
  const Ind b1 = grid.node2cells[bc.other->solver_idx][node1];
  if (!valid(b1)) throw std::runtime_error("CHT partner cell not found");
  return b1;
}

Num Solver::cht_surface_T(Ind ghost, const BC& bc) const {
  const BndryInfo bi = boundary_info(ghost);
  const Ind b1 = cht_partner(ghost, bc);
  Num T_fluid, T_solid;
  if (P.physics == PHYS_HEAT) {
    T_solid = T(lhs.data(), bi.bndryIdx);
    T_fluid = bc.other->T(bc.other->lhs.data(), b1);
  } else {
    T_fluid = T(lhs.data(), bi.bndryIdx);
    T_solid = bc.other->T(bc.other->lhs.data(), b1);
  }
  if (bc.kratio == 1.0) return 0.5 * (T_fluid + T_solid);
  return (T_fluid + bc.kratio * T_solid) / (1.0 + bc.kratio);
}

void Solver::apply_bcs(Num* buf) {
  const Ind N = ncells;
  Ind firstB = N;
  for (Ind c = 0; c < N; ++c) if (bc_idx[c] != kInvalidSInd) { firstB = c; break; }
  if (firstB == N) throw std::runtime_error("no boundary cells found!");
  SInd bcIdx = bc_idx[firstB];
  auto find = [&](Ind from) { for (Ind c = from; c < N; ++c) if (bc_idx[c] == bcIdx) return c; return N; };
  Ind first = find(0);
  Ind last = first;
  for (size_t k = 0; k < bcs.size(); ++k) {
    const BC& bc = bcs[k];
    ++bcIdx;
    first = last;
    last = find(first);
    for (Ind g = first; g < last; ++g) {
      const Ind b = boundary_info(g).bndryIdx;
      const Num* pp = bc.params.data();
      switch (bc.kind) {
        case BC_EULER_NEUMANN: {
          const Num dx = cell_dx(b, g);
          for (int v = 0; v < nvars; ++v) {
            const Num gv = bc.params.empty() ? 0.0 : pp[v];
            Q(buf, g, v) = Q(buf, b, v) - dx * gv;
          }
          break;
        }
        case BC_EULER_DIRICHLET:
          for (int v = 0; v < nvars; ++v) Q(buf, g, v) = 2.0 * pp[v] - Q(buf, b, v);
          break;
        case BC_ISENTROPIC_VORTEX: {
          Num xg[3], xb[3], qg[4], qb[4];
          for (int d = 0; d < nd; ++d) { xg[d] = x_center[(Ind)d * cap + g]; xb[d] = x_center[(Ind)d * cap + b]; }
          const Num cx = bc.params.size() >= 2 ? pp[0] : -5.0;
          const Num cy = bc.params.size() >= 2 ? pp[1] : -5.0;
          isentropic_vortex(xg, time_, cx, cy, qg);
          isentropic_vortex(xb, time_, cx, cy, qb);
          for (int v = 0; v < nvars; ++v) { Q(buf, g, v) = qg[v]; Q(buf, b, v) = qb[v]; }
          break;
        }
        case BC_CNS_INFLOW: case BC_CNS_OUTFLOW: case BC_ADIABATIC_NOSLIP: case BC_ADIABATIC_SLIP:
        case BC_ISOTHERMAL_NOSLIP: case BC_CHT_FLUID: {
          Num cvb[5], pvb[5], pvg[5], cvg[5];
          for (int v = 0; v < nvars; ++v) cvb[v] = Q(buf, b, v);
          pv_of(cvb, pvb);
          const int RHO = nd, PP = nd + 1;
          if (bc.kind == BC_CNS_INFLOW) {
            for (int d = 0; d < nd; ++d) pvg[d] = 2.0 * pp[d] - pvb[d];
            pvg[RHO] = 2.0 * pp[RHO] - pvb[RHO];
            pvg[PP] = pvb[PP] - 0 * 0;
          } else if (bc.kind == BC_CNS_OUTFLOW) {
            for (int d = 0; d < nd; ++d) pvg[d] = pvb[d] - 0 * 0;
            pvg[RHO] = pvb[RHO] - 0 * 0;
            pvg[PP] = 2.0 * pp[PP] - pvb[PP];
          } else if (bc.kind == BC_ADIABATIC_NOSLIP) {
            for (int d = 0; d < nd; ++d) pvg[d] = 2.0 * 0 - pvb[d];
            pvg[RHO] = pvb[RHO] - 0 * 0;
            pvg[PP] = pvb[PP] - 0 * 0;
          } else if (bc.kind == BC_ADIABATIC_SLIP) {
            const int normal = boundary_info(g).bndryPos / 2;
            for (int d = 0; d < nd; ++d) pvg[d] = pvb[d] - 0 * 0;
            pvg[normal] = 2.0 * 0 - pvb[normal];
            pvg[RHO] = pvb[RHO] - 0 * 0;
            pvg[PP] = pvb[PP] - 0 * 0;
          } else {  // isothermal no-slip / CHT fluid side
            const Num Ts = (bc.kind == BC_CHT_FLUID) ? cht_surface_T(g, bc) : pp[0];
            for (int d = 0; d < nd; ++d) pvg[d] = 2.0 * 0 - pvb[d];
            const Num p_srfc = pvb[PP];
            const Num rho_srfc = p_srfc * qty.gamma / Ts;
            pvg[RHO] = 2.0 * rho_srfc - pvb[RHO];
            pvg[PP] = pvb[PP] - 0 * 0;
          }
          cv_of(pvg, cvg);
          for (int v = 0; v < nvars; ++v) Q(buf, g, v) = cvg[v];
          break;
        }
        case BC_HEAT_DIRICHLET:
          Q(buf, g, 0) = 2.0 * pp[0] - Q(buf, b, 0);
          break;
        case BC_CHT_SOLID:
          Q(buf, g, 0) = 2.0 * cht_surface_T(g, bc) - Q(buf, b, 0);
          break;
        case BC_HEAT_NEUMANN:
          Q(buf, g, 0) = Q(buf, b, 0) - cell_dx(b, g) * pp[0];
          break;
        case BC_PERIODIC: {
          const Ind im = periodic_image(g);
          for (int v = 0; v < nvars; ++v) Q(buf, g, v) = Q(buf, im, v);
          break;
        }
        case BC_EULER_COUPLING: {
          // euler/coupling_conditions.hpp:24-31: the ghost takes the other
          // solver's state at the partner cell, in the buffer of the same
          // role (lhs_tag -> other's lhs, rhs_tag -> other's rhs; the SSP-RK3
          // stage buffer -> the other's stage buffer)
          const Ind b1 = cht_partner(g, bc);
          const Solver& o = *bc.other;
          const Num* ob = (buf == lhs.data()) ? o.lhs.data() : (buf == rhs.data() ? o.rhs.data() : o.stg.data());
          for (int v = 0; v < nvars; ++v) Q(buf, g, v) = o.Q(ob, b1, v);
          break;
        }
        default: throw std::runtime_error("unknown BC kind");
      }
    }
  }
}

void Solver::conservation_sums(Num* out) const {
  for (int v = 0; v < nvars; ++v) {
    Num s = 0;
    for (Ind c = 0; c < n_internal(); ++c) {
      const Num h = length[c];
      const Num V = (nd == 2) ? h * h : h * (h * h);
      s += Q(lhs.data(), c, v) * V;
    }
    out[v] = s;
  }
}

// coupling.hpp:47-63
void couple_cns_heat(Num Tr0, Solver& cns, Solver& heat) {
  const Num scaling = Tr0 * cns.qty.Pr0 * cns.qty.Re0;
  const Num min_dt_cns = cns.min_dt();
  const Num min_dt_heat = heat.min_dt() / scaling;
  const Num dt = std::min(min_dt_cns, min_dt_heat);
  cns.force_dt(dt);
  heat.force_dt(dt * scaling);
}

// coupling::apply with a prescribed fluid dt (extension for the adjoint:
// the dt schedule of a reference run, frozen so that perturbed runs take
// the same steps)
void couple_cns_heat_fixed(Num Tr0, Num dt, Solver& cns, Solver& heat) {
  const Num scaling = Tr0 * cns.qty.Pr0 * cns.qty.Re0;
  cns.force_dt(dt);
  heat.force_dt(dt * scaling);
}

// Wall heat flux through the conjugate interface, seen from the solid:
// sum over the solid's CHT ghosts of A (T_b - T_g) / h with A = h^(nd-1) and
// T_g = 2 T_s - T_b (heat/coupling_conditions.hpp:24-37), i.e.
// h^(nd-2) * 2 (T_b - T_s); positive when heat leaves the solid.
Num cht_wall_heat_flux(const Solver& heat) {
  Num q = 0;
  for (Ind g = heat.n_internal(); g < heat.ncells; ++g) {
    const BC& bc = heat.bcs[heat.bc_idx[g]];
    if (bc.kind != BC_CHT_SOLID) continue;
    const Solver::BndryInfo bi = heat.boundary_info(g);
    const Num Tb = heat.T(heat.lhs.data(), bi.bndryIdx);
    const Num Ts = heat.cht_surface_T(g, bc);
    const Num h = heat.length[bi.bndryIdx];
    const Num w = (heat.nd == 2) ? 1.0 : h;
    q += w * (2 * (Tb - Ts));
  }
  return q;
}

// coupling.hpp:66-73
void couple_same(Solver& s1, Solver& s2) {
  const Num dt = std::min(s1.min_dt(), s2.min_dt());
  s1.force_dt(dt);
  s2.force_dt(dt);
}

}  // namespace orc
