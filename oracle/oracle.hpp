// CPU ORACLE — test infrastructure only.
//
// A faithful C++17 restatement of gnzlbg/hom3's finite-volume hot path
// (grid tree, leaf/neighbour/ghost topology, Euler/CNS/heat physics,
// boundary conditions, CHT coupling, explicit time integration), written
// without Eigen/Boost so it builds with plain g++ (-O2 -ffp-contract=off,
// no fast-math).  It is the checker for the HIP product path: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// The product library (hom3_amd/libhom3g.so) never links or calls it.
//
// Every function cites the reference file:line it follows.  Arithmetic is
// kept in the reference's operation order, including Eigen's fixed-size
// reduction order (redux_novec_unroller: a0 + (a1 + a2) for 3 terms),
// std::pow(x, 2.) where the reference writes it, and the explicit
// left-to-right accumulations of hand-written loops.
//
// Parity status per function is recorded in the oracle section of the
// design notes; functions whose only check is the ULP-level order of an
// Eigen/libm call are marked "parity unpinned (ULP)".
#pragma once
#include <cstdint>
#include <cmath>
#include <limits>
#include <memory>
#include <string>
#include <vector>
#include <functional>

namespace orc {

#ifdef ORACLE_LONG_DOUBLE
using Num = long double;  // checker build for finite differences at small steps
#else
using Num = double;
#endif
using Ind = int64_t;
using SInd = int32_t;

// misc/constants.hpp:42-63: invalid<T>() = numeric max for integers, NaN for Num
constexpr Ind kInvalidInd = std::numeric_limits<Ind>::max();
constexpr SInd kInvalidSInd = std::numeric_limits<SInd>::max();
inline Num invalid_num() { return std::numeric_limits<Num>::quiet_NaN(); }
inline bool valid(Ind i) { return i != kInvalidInd; }

// misc/constants.hpp:67: boost::math::constants::pi<double>()
constexpr Num kPi = 3.141592653589793238462643383279502884;

// Eigen fixed-size redux order (redux_novec_unroller): sizes 1..3.
inline Num esum(const Num* a, int n) {
  if (n == 1) return a[0];
  if (n == 2) return a[0] + a[1];
  return a[0] + (a[1] + a[2]);
}

// ---------------------------------------------------------------------------
// Geometry: implicit signed-distance functions (geometry/implicit.hpp:39-103)
// plus the Inverted adaptor (:439-452).  Lowered to POD descriptors so both
// the oracle and the product consume the same description.
// ---------------------------------------------------------------------------
enum GeomKind : int32_t { GEOM_NONE = 0, GEOM_EDGE = 1, GEOM_SQUARE = 2, GEOM_SPHERE = 3 };

struct GeomDesc {      // mirrors hom3g_geometry in include/hom3g.h
  int32_t kind;
  int32_t invert;      // 1: signed distance multiplied by -1 (adaptors::Inverted)
  double a[3];         // EDGE: point;  SQUARE: centre;  SPHERE: centre
  double b[3];         // EDGE: normal; SQUARE: lengths; SPHERE: b[0] = radius
                       // (double in every build: the struct crosses the C API)
};

Num signed_distance(const GeomDesc& g, const Num* x, int nd);

// ---------------------------------------------------------------------------
// Hierarchical tree + Cartesian grid (containers/hierarchical/implementation.hpp,
// grid/grid.hpp, grid/generation.hpp)
// ---------------------------------------------------------------------------
struct Grid {
  int nd;
  Num root_length;        // grid.hpp:106-132 RootCell
  Num root_coords[3];
  std::vector<Ind> parent, first_child;
  std::vector<std::vector<Ind>> node2cells;   // per solver (implementation.hpp:795)
  int max_solvers;

  Grid(int nd_, const Num* xmin, const Num* xmax, int max_solvers_);
  Ind size() const { return (Ind)parent.size(); }
  int nchild() const { return 1 << nd; }
  int npos() const { return 2 * nd; }
  bool is_leaf(Ind n) const { return !valid(first_child[n]); }
  bool is_root(Ind n) const { return !valid(parent[n]); }
  Ind child(Ind n, int pos) const;
  Ind refine(Ind n);
  int level(Ind n) const;
  int position_in_parent(Ind n) const;
  Ind find_samelvl_neighbor(Ind n, int pos) const;
  Num cell_length_at_level(Ind level) const;
  Num cell_length(Ind n) const { return cell_length_at_level(level(n)); }
  void cell_coordinates(Ind n, Num* x) const;
  void neighbor_coordinates(Ind n, int pos, Num* x) const;
  bool is_cut_by(Ind n, const GeomDesc& g) const;
  void min_level(int level);                       // generation.hpp:45-59
  void refine_region(int max_level, const GeomDesc& g, Num band);  // extension
  void balance_2to1();                                             // extension
};

int child_rel_pos(int nd, int cpos, int d);   // grid.hpp:41-59
int nghbr_rel_pos(int nd, int npos, int d);   // grid.hpp:64-80
int rel_sibling_position(int cpos, int npos); // implementation.hpp:46-57
inline int opposite_position(int p) { return p ^ 1; }  // implementation.hpp:60-61

// ---------------------------------------------------------------------------
// Solver configuration
// ---------------------------------------------------------------------------
enum Physics : int32_t { PHYS_EULER = 0, PHYS_CNS = 1, PHYS_HEAT = 2 };
enum FluxKind : int32_t { FLUX_AUSM = 0, FLUX_LLF = 1, FLUX_THREE_POINT = 2 };
enum TimeInt : int32_t { TI_EULER_FWD = 0, TI_HEUN_RK2 = 1, TI_SSP_RK3 = 2 };
enum Recon : int32_t { RECON_NONE = 0, RECON_MUSCL = 1 };
enum Limiter : int32_t { LIM_ZERO = 0, LIM_MINMOD = 1, LIM_VANLEER = 2 };

struct PhysDesc {       // mirrors hom3g_physics in include/hom3g.h (double in every build)
  int32_t physics, nd, flux, ti, recon, limiter;
  double gamma, cfl, cfl_viscous;
  double Rspecific, T0, rho0, Minfty;     // euler/quantities.hpp:23-29
  double ReInfty, Pr0, Sutherland0;       // cns/quantities.hpp:23-30
  double diffusivity, T_ref;              // heat/quantities.hpp:14-17
  double time_end;
  Ind max_cells;
};

// Nondimensional quantities (euler/quantities.hpp:131-145, cns/quantities.hpp:23-30, 113-115)
struct Quantities {
  Num gamma, Minfty, ReInfty, Pr0, sOverTref, sOverTrefP1, Re0;
  explicit Quantities(const PhysDesc& p);
  Num gammaM1() const { return gamma - 1.; }
  Num T_infinity() const;
  Num rho_infinity() const;
  Num u_infinity() const;
  Num p_infinity() const;
  Num mu_infinity() const { return sutherland(T_infinity()); }
  Num sutherland(Num T) const { return std::pow(T, 1.5) * sOverTrefP1 / (T + sOverTref); }
};

// Boundary condition kinds (one Condition class each in the reference)
enum BCKind : int32_t {
  BC_EULER_NEUMANN = 0,       // euler/boundary_conditions.hpp:17-38 (also cns::bc::Neumann)
  BC_EULER_DIRICHLET = 1,     // euler/boundary_conditions.hpp:42-60
  BC_ISENTROPIC_VORTEX = 2,   // euler/boundary_conditions.hpp:65-89
  BC_CNS_INFLOW = 3,          // cns/boundary_conditions.hpp:28-65
  BC_CNS_OUTFLOW = 4,         // cns/boundary_conditions.hpp:72-105
  BC_ADIABATIC_NOSLIP = 5,    // cns/boundary_conditions.hpp:116-150
  BC_ADIABATIC_SLIP = 6,      // cns/boundary_conditions.hpp:156-193
  BC_ISOTHERMAL_NOSLIP = 7,   // cns/boundary_conditions.hpp:199-238
  BC_HEAT_DIRICHLET = 8,      // heat/boundary_conditions.hpp:15-37
  BC_HEAT_NEUMANN = 9,        // heat/boundary_conditions.hpp:40-63
  BC_PERIODIC = 10,           // extension: ghost = periodic image, slope hook = image slope
  BC_CHT_FLUID = 11,          // cns/coupling_conditions.hpp:21-43
  BC_CHT_SOLID = 12,          // heat/coupling_conditions.hpp:18-38
  BC_EULER_COUPLING = 13,     // euler/coupling_conditions.hpp:17-35
};

struct Solver;

struct BC {
  std::string name;
  GeomDesc geom;
  int kind;
  std::vector<Num> params;     // constant distributions (see apply())
  Solver* other = nullptr;     // coupled solver (CHT)
  Num kratio = 1.0;            // CHT extension: k_solid / k_fluid
};

// Initial conditions (function-valued property "initialCondition")
enum ICKind : int32_t {
  IC_CONSTANT_PV = 0,      // constant primitive state
  IC_SHOCK_TUBE = 1,       // euler/initial_conditions.hpp:57-94
  IC_ISENTROPIC_VORTEX = 2,// euler/initial_conditions.hpp:19-45 (centre param; reference: (-5,-5))
  IC_SPHERE = 3,           // explosion test / blast: inside/outside states
  IC_TGV = 4,              // Taylor-Green vortex (extension)
  IC_FOUR_QUADRANT = 5,    // four-quadrant Riemann problem (extension)
  IC_HEAT_CONSTANT = 6,
  IC_HEAT_SINE = 7,        // T = a + b*prod sin(k x_d) (tests)
  IC_LINEAR_PV = 8,        // pv = pv0 + G x (tests)
  IC_CALLBACK = 9,
};
typedef void (*ic_callback_t)(const double* x, double* out_cv, void* ctx);

struct Solver {
  Grid& grid;
  int solver_idx;
  PhysDesc P;
  Quantities qty;
  int nd, nvars, npos;
  Ind cap;
  // fv::Container (container.hpp:58-66), SoA column-major with ld = cap
  std::vector<Num> lhs, rhs, stg, x_center, distances, length;
  std::vector<Ind> neighbors, node_idx;
  std::vector<SInd> bc_idx;
  Ind ncells = 0, firstGC = kInvalidInd;
  std::vector<BC> bcs;
  std::vector<Num> source;            // per internal cell volumetric source (heat extension)
  // 2:1 refinement extension.  A face towards 2^(nd-1) finer leaves is
  // stored in the neighbour table as -(2 + f); fine[f * nfine_per() + k]
  // holds the finer cells in ascending child position.
  std::vector<Ind> fine;
  bool nonuniform = false;
  int nfine_per() const { return 1 << (nd - 1); }
  static bool is_fine_ref(Ind n) { return n < 0; }
  static Ind fine_index(Ind n) { return -(n + 2); }
  // seeded density perturbation applied after the initial condition
  uint64_t perturb_seed = 0; Num perturb_amp = 0;
  // domain mask
  GeomDesc domain_geom{}; int domain_sign = 0;
  std::vector<uint8_t> domain_mask;   // optional explicit mask over leaf order
  // IC
  int ic_kind = IC_CONSTANT_PV; std::vector<Num> ic_params; ic_callback_t ic_cb = nullptr; void* ic_ctx = nullptr;
  int ghost_order = 0;                 // 0 canonical (stable), 1 std::sort
  // time (solver.hpp:406-416)
  Num dt_ = 0, time_ = 0, tEnd_ = 0, forced_dt_ = 0; Ind step_ = 0, forced_dt_step_ = kInvalidInd;

  Solver(Grid& g, int sidx, const PhysDesc& p);
  Num& Q(Num* buf, Ind c, int v) { return buf[(Ind)v * cap + c]; }
  Num Q(const Num* buf, Ind c, int v) const { return buf[(Ind)v * cap + c]; }
  Ind& nghbr(Ind c, int p) { return neighbors[(Ind)p * cap + c]; }
  Ind nghbr(Ind c, int p) const { return neighbors[(Ind)p * cap + c]; }
  Ind n_internal() const { return valid(firstGC) ? firstGC : ncells; }
  bool is_ghost(Ind c) const { return !valid(node_idx[c]); }

  // initialization (solver.hpp:133-138, 867-992)
  void initialize();
  void create_local_cells();
  void create_ghost_cells();
  Ind push_cell();
  Ind create_ghost_cell(Ind b, int pos);
  void sort_gc();
  void impose_initial_condition();
  void apply_perturbation();
  void ic_point(const Num* x, Num* out) const;

  // stepping (solver.hpp:140-145, 469-551)
  void solve();
  Num compute_dt() const;
  Num min_dt() const { return compute_dt(); }
  void force_dt(Num dt) { forced_dt_step_ = step_; forced_dt_ = dt; }
  void set_dt();
  void evolve();
  void apply_bcs(Num* buf);
  void num_flux(const Num* Qb, Ind c, Num* out) const;
  void compute_num_flux(const Num* Qb, Ind l, Ind r, int d, Num dx, Num dt, Num* out) const;
  void source_term(Ind c, Num* out) const;

  // physics helpers (euler/physics.hpp:53-165)
  Num rho(const Num* Qb, Ind c) const { return Q(Qb, c, nd); }
  Num rho_u(const Num* Qb, Ind c, int d) const { return Q(Qb, c, d); }
  Num rho_E(const Num* Qb, Ind c) const { return Q(Qb, c, nd + 1); }
  Num u(const Num* Qb, Ind c, int d) const { return rho_u(Qb, c, d) / rho(Qb, c); }
  Num u_mag2(const Num* Qb, Ind c) const;
  Num p(const Num* Qb, Ind c) const;
  Num a(const Num* Qb, Ind c) const;
  Num T(const Num* Qb, Ind c) const;
  Num cell_dx(Ind a, Ind b) const;
  int which_neighbor(Ind c, Ind n) const;
  struct BndryInfo { Ind bndryIdx; int bndryPos; Ind ghostIdx; int ghostPos; };
  BndryInfo boundary_info(Ind g) const;
  Num slope(const Num* Qb, Ind c, int v, int dir) const;
  Num surface_slope(const Num* Qb, Ind l, Ind r, int v, int sdir) const;
  Num bc_slope(const Num* Qb, const BC& bc, Ind ghost, Ind bndry, int v, int dir) const;
  Num cell_dt(const Num* Qb, Ind c) const;
  void ausm_states(const Num* QL, const Num* QR, int d, Num* f) const;
  void ausm(const Num* Qb, Ind l, Ind r, int d, Num* f) const;
  void llf(const Num* Qb, Ind l, Ind r, int d, Num dx, Num dt, Num* f) const;
  void viscous(const Num* Qb, Ind l, Ind r, int d, Num* f) const;
  void pv_of(const Num* cvs, Num* pvs) const;   // euler/quantities.hpp:173-186
  void cv_of(const Num* pvs, Num* cvs) const;   // euler/quantities.hpp:153-166
  Num cht_surface_T(Ind ghost, const BC& bc) const;
  Ind cht_partner(Ind ghost, const BC& bc) const;
  Ind periodic_image(Ind ghost) const;

  // conservation sums over internal cells (extension; FP64 left-to-right)
  void conservation_sums(Num* out) const;
};

// AUSM split Mach number m_int_<s> and pressure weight p_int_<s>
// (euler/physics.hpp:297-310), s = +1 or -1
Num ausm_m_split(int s, Num M);
Num ausm_p_split(int s, Num M);

// coupling.hpp:47-63
void couple_cns_heat(Num Tr0, Solver& cns, Solver& heat);
void couple_cns_heat_fixed(Num Tr0, Num dt, Solver& cns, Solver& heat);
Num cht_wall_heat_flux(const Solver& heat);
// coupling.hpp:66-73
void couple_same(Solver& s1, Solver& s2);

// counter-based PRNG (extension; SplitMix64)
uint64_t splitmix64(uint64_t x);
Num prng_uniform(uint64_t seed, uint64_t stream, uint64_t counter);  // U[0,1)
uint64_t morton_encode(int nd, const uint32_t* ijk);
// Morton index of node n among the nodes of its level (child positions from
// the root down, the root-level digit most significant)
uint64_t node_morton(const Grid& g, Ind n);

}  // namespace orc
