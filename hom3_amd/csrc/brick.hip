// Brick kernels: the residual + Runge-Kutta stage update for full uniform
// grids (every leaf at one level, cell id = Morton index), with periodic or
// zero-gradient (Neumann) faces.
//
// Layout.  The reference orders the leaves of a MinLevel grid breadth-first
// with children in Z-order (grid/generation.hpp:45-59,
// containers/hierarchical/implementation.hpp:320-364), so solver cell c is the
// Morton index of its integer coordinates (x bit 0).  A Morton-aligned block
// of B^nd consecutive cells is therefore a B^nd brick: the interior of a brick
// is one contiguous, fully coalesced SoA load per variable.
//
// One workgroup = one brick (3D: 8^3 cells / 512 threads, 2D: 16^2 / 256).
//  1. load the brick + halo (faces at depth H, edges at depth 1 for the
//     viscous tangential slopes) into LDS; halo cells are wrapped (periodic)
//     or clamped (zero-gradient), which reproduces the ghost values and the
//     slope hooks of the reference's ghost cells for these two conditions;
//  2. per-position primitives (p, 1/(rho a)) in LDS, formed as the values load;
//  3. per direction d: every face of the brick is evaluated ONCE into LDS
//     (the reference evaluates each face from both sides,
//     solver.hpp:439-460, with identical values): each thread the face
//     below its cell, single waves the faces on the brick's upper boundary
//     (see TailCfg); then every cell accumulates (F_m - F_p) in d-ascending
//     order like num_flux;
//  4. the stage combination (forward Euler / Heun / SSP-RK3) is written
//     back coalesced; optionally the next step's CFL minimum is fused in.
// Bricks are mapped to workgroups XCD-aware: consecutive workgroup ids are
// dispatched round-robin over the 8 XCDs, so workgroup w processes brick
// (w % 8) * (nb / 8) + w / 8 and every XCD sweeps one contiguous Morton range
// (an octant), keeping neighbouring bricks' halos in that XCD's L2.
#include "brick_geo.hpp"

namespace hom3g {

namespace {

struct BrickArgs {
  const void* in;     // stage input state (TIn)
  const double* base; // U^n for the RK combinations (FP64)
  void* out;          // stage output (TOut)
  long long ld;
  int level;          // grid side n = 2^level
  int periodic[3];    // per direction: 1 periodic, 0 clamp
  double h;
  double inv_h;       // 1 / h (host-computed: no division per thread)
  int kind;           // stage combination (see stepper.hip StageArgs)
  int fuse_dt;
  int advance;        // last stage of the step: also time += dt, ++step
  const unsigned* halo;  // tiled bricks: per brick, the halo cells (see brick_table_build)
};

template <class T> __device__ __forceinline__ T ld_state(const void* p, long long i);
template <> __device__ __forceinline__ double ld_state<double>(const void* p, long long i) {
  return static_cast<const double*>(p)[i];
}
template <> __device__ __forceinline__ float ld_state<float>(const void* p, long long i) {
  return static_cast<const float*>(p)[i];
}

// Face buffers.  Along each direction a brick has B^nd "minus" faces (one
// per cell, towards its lower neighbour) and B^(nd-1) faces on its upper
// boundary (the tail).  Every thread evaluates the minus face of its cell;
// the tails go to NSLOT slots and are evaluated by one wave each, NSLOT of
// them in the first direction's pass (the remaining ones in the next pass),
// so that no pass waits for one wave evaluating a second face layer.  Two
// slots for the 3D FP64 brick keep its LDS within half a CU (two resident
// workgroups).
template <int ND, bool MUSCL, class TC>
struct TailCfg {
  static constexpr int NSLOT = (ND == 3 && !MUSCL && sizeof(TC) == 8) ? 2 : ND;
  // pass in which the tail of direction D is evaluated, its wave and slot
  static constexpr int pass(int D) { return D < NSLOT ? 0 : D - NSLOT + 1; }
  static constexpr int wave(int D) { return D < NSLOT ? D : D - NSLOT; }
  static constexpr int slot(int D) { return D % NSLOT; }
};

template <int ND, int PHYS, bool MUSCL, class TC>
struct Smem {
  using G = Geo<ND, PHYS, MUSCL>;
  TC q[G::NV][G::NPOS];
  TC p[MUSCL ? 1 : G::NPOS];
  TC iar[MUSCL ? 1 : G::NPOS];  // 1 / (rho a)
  TC fm[G::NV][BrickCfg<ND>::NT];                    // minus faces, at fm_slot(Morton cell)
  TC ft[TailCfg<ND, MUSCL, TC>::NSLOT][G::NV][G::TANG];  // upper-boundary faces
  double red[BrickCfg<ND>::NT / 64];
  // tiled bricks: per face-halo entry j (closed-form order), bit j % 64 of
  // gmask[j / 64][0] is set when the entry is a ghost and of [1] when its
  // slope hook negates the velocity slopes (walls)
  unsigned long long gmask[(2 * ND * G::TANG + 63) / 64][2];
};

// Face-pass lane order: in the face passes thread t evaluates the face
// below cell face_cell(t) rather than below its own (Morton) cell.  The
// LDS position of cell (x, y, z) is x + E y + E^2 z plus a constant; for
// E = 10 (3D brick, halo depth 1) it is x + 10 y + 4 z modulo 32, so the
// 32 cells {x = 0..7, z = p, p + 2, p + 4, p + 6} of one y-plane and z
// parity occupy 32 distinct values modulo 32 and the 8-byte reads of a
// half-wave (at the cells and at constant offsets from them) hit distinct
// banks; for E = 18 (2D, depth 1) the rows y and y + 8 do.  Other layouts
// keep the Morton order.
template <int ND, int PHYS, bool MUSCL>
__device__ __forceinline__ void face_cell(int t, int* c) {
  using G = Geo<ND, PHYS, MUSCL>;
  if constexpr (ND == 3 && G::E == 10) {
    c[0] = t & 7;
    c[1] = t >> 6;
    c[2] = 2 * ((t >> 3) & 3) + ((t >> 5) & 1);
  } else if constexpr (ND == 2 && G::E == 18) {
    c[0] = t & 15;
    c[1] = (t >> 5) + 8 * ((t >> 4) & 1);
    c[2] = 0;
  } else {
    c[2] = 0;
    demorton<ND>((unsigned)t, c);
  }
}

// LDS slot of the minus-face flux of Morton cell m.  With the face-pass
// lane order above, the Morton indices a half-wave writes take only 4
// distinct values modulo 16 (8-byte slots: 4-way bank conflicts); an XOR of
// two higher index bits into bits 1-2 (3D: bits 5-6; 2D: bits 4 and 6 into
// bits 1 and 3) spreads them over all 16, and keeps the accumulate reads
// (at m and at the Morton neighbour along D) conflict-free.  A bijection on
// the brick's cells: the source bits are left unchanged.
template <int ND, int PHYS, bool MUSCL>
__device__ __forceinline__ int fm_slot(unsigned m) {
  using G = Geo<ND, PHYS, MUSCL>;
  if constexpr (ND == 3 && G::E == 10) return (int)(m ^ ((m >> 4) & 6u));
  else if constexpr (ND == 2 && G::E == 18) return (int)(m ^ ((m >> 3) & 10u));
  else return (int)m;
}

// face flux along D between LDS positions iL, iR (neighbours along D).
// hook (tiled bricks): 0 for a face between two cells; bit 0: iL is a
// ghost, bit 1: iR is a ghost, bit 2: that ghost's slope hook is a wall's.
// A ghost's tangential slopes are those of its boundary cell, the cell on
// the other side of the face, with the velocity components negated at
// walls (solver.hpp:596-623, cns/boundary_conditions.hpp:143-147), so the
// face's tangential gradient is twice the boundary cell's, or zero.
template <int ND, int PHYS, bool MUSCL, class TC, int D>
__device__ __forceinline__ void brick_face(const Smem<ND, PHYS, MUSCL, TC>& s, int iL, int iR,
                                           const PhysConst& pc, TC inv_h, TC q_coef,
                                           TC inv_Re0, TC* f, int hook = 0) {
  using G = Geo<ND, PHYS, MUSCL>;
  constexpr int NV = G::NV;
  const TC gamma = TC(pc.gamma), gm1 = TC(pc.gm1);
  TC qL[NV], qR[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) { qL[v] = s.q[v][iL]; qR[v] = s.q[v][iR]; }
  if constexpr (MUSCL) {
    constexpr int st = (D == 0) ? 1 : (D == 1 ? G::E : G::E * G::E);
    TC aL[NV], aR[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const TC sl = limit(pc.limiter, qL[v] - s.q[v][iL - st], qR[v] - qL[v]);
      const TC sr = limit(pc.limiter, qR[v] - qL[v], s.q[v][iR + st] - qR[v]);
      aL[v] = qL[v] + TC(0.5) * sl;
      aR[v] = qR[v] - TC(0.5) * sr;
    }
    const Prim<ND, TC> wL = primitives_rcp<ND>(aL, gamma, gm1);
    const Prim<ND, TC> wR = primitives_rcp<ND>(aR, gamma, gm1);
    ausm_flux_rcp<ND, D, TC>(aL, wL, aR, wR, f);
  } else {
    const TC pL = s.p[iL], pR = s.p[iR];
    const TC iarL = s.iar[iL], iarR = s.iar[iR];
    // a = (gamma p / rho) / a = gamma p / (rho a); M = u / a = (rho u) / (rho a)
    const TC aL = gamma * pL * iarL;
    const TC aR = gamma * pR * iarR;
    ausm_flux_fast<ND, D, TC>(qL, qL[D], pL, aL, iarL, qR, qR[D], pR, aR, iarR, f);
    if constexpr (G::VISC) {
      TC Gn[NV];
      TC Gt[NV][ND - 1];
#pragma unroll
      for (int v = 0; v < NV; ++v) Gn[v] = qR[v] - qL[v];
#pragma unroll
      for (int t = 0; t < ND - 1; ++t) {
        const int sd = (t < D) ? t : t + 1;
        const int st = (sd == 0) ? 1 : (sd == 1 ? G::E : G::E * G::E);
#pragma unroll
        for (int v = 0; v < NV; ++v)
          Gt[v][t] = (s.q[v][iR + st] - s.q[v][iR - st]) + (s.q[v][iL + st] - s.q[v][iL - st]);
      }
      if (hook) {
        const int ib = (hook & 1) ? iR : iL;
#pragma unroll
        for (int t = 0; t < ND - 1; ++t) {
          const int sd = (t < D) ? t : t + 1;
          const int st = (sd == 0) ? 1 : (sd == 1 ? G::E : G::E * G::E);
#pragma unroll
          for (int v = 0; v < NV; ++v) {
            const TC d = s.q[v][ib + st] - s.q[v][ib - st];
            Gt[v][t] = ((hook & 4) && v < ND) ? TC(0) : d + d;
          }
        }
      }
      TC fv[NV];
      viscous_flux_fast<ND, D, TC>(qL, pL, qR, pR, Gn, Gt, pc, q_coef, inv_Re0, inv_h, fv);
#pragma unroll
      for (int v = 0; v < NV; ++v) f[v] = f[v] + fv[v];
    }
  }
}

// tangential index of a cell / face position across direction D
template <int ND, int B, int D>
__device__ __forceinline__ int tang_index(const int* c) {
  if constexpr (ND == 3) {
    if constexpr (D == 0) return c[1] + B * c[2];
    else if constexpr (D == 1) return c[0] + B * c[2];
    else return c[0] + B * c[1];
  } else {
    return (D == 0) ? c[1] : c[0];
  }
}

// upper-boundary face tg of direction D (lane tg of the wave that owns it)
// ghost bits of face-halo entry j (tiled bricks): bit 0 ghost, bit 1 wall hook
template <int ND, int PHYS, bool MUSCL, class TC>
__device__ __forceinline__ int ghost_bits(const Smem<ND, PHYS, MUSCL, TC>& s, int j) {
  const unsigned long long b = 1ull << (j & 63);
  return ((s.gmask[j >> 6][0] & b) ? 1 : 0) | ((s.gmask[j >> 6][1] & b) ? 2 : 0);
}

template <int ND, int PHYS, bool MUSCL, class TC, bool TABLE, int D>
__device__ __forceinline__ void brick_tail(Smem<ND, PHYS, MUSCL, TC>& s, const PhysConst& pc, TC inv_h,
                                           TC q_coef, TC inv_Re0, int tg) {
  using G = Geo<ND, PHYS, MUSCL>;
  constexpr int NV = G::NV;
  constexpr int B = G::B;
  constexpr int st = (D == 0) ? 1 : (D == 1 ? G::E : G::E * G::E);
  constexpr int slot = TailCfg<ND, MUSCL, TC>::slot(D);
  if (tg >= G::TANG) return;
  int c[3];
  if constexpr (ND == 3) {
    const int ta = tg % B, tb = tg / B;
    if constexpr (D == 0) { c[0] = B; c[1] = ta; c[2] = tb; }
    else if constexpr (D == 1) { c[0] = ta; c[1] = B; c[2] = tb; }
    else { c[0] = ta; c[1] = tb; c[2] = B; }
  } else {
    if constexpr (D == 0) { c[0] = B; c[1] = tg; }
    else { c[0] = tg; c[1] = B; }
    c[2] = 0;
  }
  const int iR = G::lin(c);
  int hook = 0;
  if constexpr (TABLE && G::VISC) {
    // the tail's R cell is face-halo entry (2 D + 1) TANG + tg
    const int gb = ghost_bits(s, (2 * D + 1) * G::TANG + tg);
    hook = (gb & 1) ? (2 | ((gb & 2) ? 4 : 0)) : 0;
  }
  TC f[NV];
  brick_face<ND, PHYS, MUSCL, TC, D>(s, iR - st, iR, pc, inv_h, q_coef, inv_Re0, f, hook);
#pragma unroll
  for (int v = 0; v < NV; ++v) s.ft[slot][v][tg] = f[v];
}

// tails whose pass is P: direction D2 on wave wave(D2)
template <int ND, int PHYS, bool MUSCL, class TC, bool TABLE, int P, int D2>
__device__ __forceinline__ void brick_tails_from(Smem<ND, PHYS, MUSCL, TC>& s, const PhysConst& pc, TC inv_h,
                                                 TC q_coef, TC inv_Re0, int t) {
  if constexpr (D2 < ND) {
    using TCf = TailCfg<ND, MUSCL, TC>;
    if constexpr (TCf::pass(D2) == P) {
      if ((t >> 6) == TCf::wave(D2))
        brick_tail<ND, PHYS, MUSCL, TC, TABLE, D2>(s, pc, inv_h, q_coef, inv_Re0, t & 63);
    }
    brick_tails_from<ND, PHYS, MUSCL, TC, TABLE, P, D2 + 1>(s, pc, inv_h, q_coef, inv_Re0, t);
  }
}

template <int ND, int PHYS, bool MUSCL, class TC, bool TABLE, int D>
__device__ __forceinline__ void brick_direction(Smem<ND, PHYS, MUSCL, TC>& s, const PhysConst& pc, TC inv_h,
                                                TC q_coef, TC inv_Re0, const int* cell,
                                                TC* acc) {
  using G = Geo<ND, PHYS, MUSCL>;
  constexpr int NV = G::NV;
  constexpr int B = G::B;
  constexpr int st = (D == 0) ? 1 : (D == 1 ? G::E : G::E * G::E);
  const int t = threadIdx.x;
  // the minus face of cell face_cell(t) (Morton index fm_idx), then the
  // tails scheduled in this pass
  {
    int fcell[3] = {0, 0, 0};
    face_cell<ND, PHYS, MUSCL>(t, fcell);
    const int fm_idx = fm_slot<ND, PHYS, MUSCL>((unsigned)morton<ND>(fcell));
    const int iR = G::lin(fcell);
    int hook = 0;
    if constexpr (TABLE && G::VISC) {
      // a cell on the brick's lower side along D: its L cell is face-halo
      // entry 2 D TANG + (tangential index)
      if (fcell[D] == 0) {
        const int gb = ghost_bits(s, 2 * D * G::TANG + tang_index<ND, B, D>(fcell));
        hook = (gb & 1) ? (1 | ((gb & 2) ? 4 : 0)) : 0;
      }
    }
    TC f[NV];
    brick_face<ND, PHYS, MUSCL, TC, D>(s, iR - st, iR, pc, inv_h, q_coef, inv_Re0, f, hook);
#pragma unroll
    for (int v = 0; v < NV; ++v) s.fm[v][fm_idx] = f[v];
  }
  brick_tails_from<ND, PHYS, MUSCL, TC, TABLE, D, 0>(s, pc, inv_h, q_coef, inv_Re0, t);
  __syncthreads();
  // acc += F(minus face) - F(plus face); the plus face is the minus face of
  // the next cell along D (Morton increment along the axis) or a tail.  (The
  // always-true guard keeps this block apart for the scheduler, which
  // otherwise raises the register count past four waves per SIMD.)
  if (t < BrickCfg<ND>::NT) {
    const unsigned mask = (ND == 3 ? 0x49u : 0x55u) << D;
    const unsigned m = (unsigned)t;
    const unsigned nm = (((m | ~mask) + 1u) & mask) | (m & ~mask);
    const bool top = cell[D] == B - 1;
    const TC* fp = top ? &s.ft[TailCfg<ND, MUSCL, TC>::slot(D)][0][tang_index<ND, B, D>(cell)]
                       : &s.fm[0][top ? 0 : fm_slot<ND, PHYS, MUSCL>(nm)];
    const int pstride = top ? G::TANG : BrickCfg<ND>::NT;
    const int own = fm_slot<ND, PHYS, MUSCL>(m);
#pragma unroll
    for (int v = 0; v < NV; ++v) acc[v] += (s.fm[v][own] - fp[v * pstride]);
  }
  __syncthreads();
}

// p = (g-1)(rho E - rho |u|^2 / 2) and 1/(rho a), a = sqrt(g p / rho), of
// one position, from its conserved state
template <int ND, int PHYS, bool MUSCL, class TC>
__device__ __forceinline__ void put_prims(Smem<ND, PHYS, MUSCL, TC>& s, int i, const TC* q, const PhysConst& pc) {
  TC p, ia;
  brick_prims<ND, TC>(q, TC(pc.gamma), TC(pc.gm1), p, ia);
  s.p[i] = p;
  s.iar[i] = ia;
}

template <int ND, int PHYS, bool MUSCL, class TC, class TIn, class TOut, bool TABLE>
__global__ void __launch_bounds__(BrickCfg<ND>::NT) k_brick(BrickArgs A, PhysConst pc, DevScalars* scal) {
  static_assert(!(TABLE && MUSCL), "tiled bricks have halos of depth 1");
  using G = Geo<ND, PHYS, MUSCL>;
  constexpr int NV = G::NV;
  constexpr int B = G::B;
  constexpr int NT = BrickCfg<ND>::NT;
  constexpr int H = G::H;
  constexpr int E = G::E;
  __shared__ Smem<ND, PHYS, MUSCL, TC> s;

  const int t = threadIdx.x;
  const long long nb = (long long)gridDim.x;
  // XCD-aware brick order (8 XCDs, round-robin dispatch)
  long long brick = blockIdx.x;
  if ((nb & 7) == 0) brick = (long long)(blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
  const int n = 1 << A.level;
  int b0[3] = {0, 0, 0};
  demorton<ND>((unsigned)brick, b0);
#pragma unroll
  for (int d = 0; d < ND; ++d) b0[d] *= B;

  // 1) interior: thread t <-> Morton cell t of the brick (coalesced), and
  // halo positions (faces to depth H; edges to depth 1 for the viscous
  // slopes) from the compile-time list: the Morton index of a halo cell is
  // the per-axis Morton bits of its coordinate, which only take 5 values
  // per axis for the whole brick (inside, -1, -2, +B, +B+1; wrapped or
  // clamped at the grid boundary).  The primitives (p, 1/(rho a)) of every
  // position a face reads (interior and face halos) are formed from the
  // loaded registers.
  int cell[3] = {0, 0, 0};
  demorton<ND>((unsigned)t, cell);
  const long long gc = brick * (long long)NT + t;
  {
    using HL = HaloList<ND, PHYS, MUSCL>;
    const int ic = G::lin(cell);
    // per axis, the Morton bits of the 5 coordinates a halo can take
    unsigned P[3][5] = {{0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}, {0, 0, 0, 0, 0}};
    auto make_P = [&]() {
#pragma unroll
      for (int d = 0; d < ND; ++d) {
        const int off[5] = {0, -1, -2, B, B + 1};
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          int x = b0[d] + off[k];
          if (A.periodic[d]) x &= (n - 1);
          else x = x < 0 ? 0 : (x >= n ? n - 1 : x);
          const unsigned sp = (ND == 3) ? part1by2((unsigned)x) : part1by1((unsigned)x);
          P[d][k] = sp << d;
        }
      }
    };
    if constexpr (G::H == 1) {
      // the interior loads go out first (they need no halo arithmetic),
      // then this thread's halo entry in closed form (no table read ahead of
      // the loads): all values are in flight before the first wait.
      // Threads past the end of the list load its last entry again and
      // discard it (no branch around the loads).
      TC qi[NV], qh[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) qi[v] = TC(ld_state<TIn>(A.in, v * A.ld + gc));
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (!TABLE) make_P();
      const bool has_h = t < HL::N;
      const int jh = has_h ? t : HL::N - 1;
      int hc[3] = {0, 0, 0};
      halo_closed_form<ND, PHYS, MUSCL>(jh, hc);
      long long gh;
      if constexpr (TABLE) {
        // tiled bricks: the cell (internal, ghost or periodic image) from
        // the brick's halo table, and its ghost bits for the face passes
        // (the table holds the viscous halo list: face halos first)
        const unsigned e = A.halo[brick * HaloList<ND, HOM3G_CNS, false>::N + jh];
        gh = (long long)(e & 0x3fffffffu);
        const unsigned code = e >> 30;
        const bool face = t < 2 * ND * G::TANG;
        const unsigned long long g0 = __ballot(face && (code & 1u) != 0);
        const unsigned long long g1 = __ballot(face && code == 3u);
        if ((t & 63) == 0 && (t >> 6) < (2 * ND * G::TANG + 63) / 64) {
          s.gmask[t >> 6][0] = g0;
          s.gmask[t >> 6][1] = g1;
        }
      } else {
        unsigned m = 0;
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          const int x = hc[d];
          const unsigned sp = (ND == 3) ? (unsigned)((x & 1) | ((x & 2) << 2) | ((x & 4) << 4))
                                        : (unsigned)((x & 1) | ((x & 2) << 1) | ((x & 4) << 2) | ((x & 8) << 3));
          m |= x < 0 ? P[d][1] : (x >= B ? P[d][3] : (P[d][0] | (sp << d)));
        }
        gh = (long long)m;
      }
      const int ih = G::lin(hc);
#pragma unroll
      for (int v = 0; v < NV; ++v) qh[v] = TC(ld_state<TIn>(A.in, v * A.ld + gh));
      // keep the consumers below the loads
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int v = 0; v < NV; ++v) s.q[v][ic] = qi[v];
      if constexpr (!MUSCL) put_prims<ND, PHYS, MUSCL, TC>(s, ic, qi, pc);
      if (has_h) {
#pragma unroll
        for (int v = 0; v < NV; ++v) s.q[v][ih] = qh[v];
        // face halos come first
        if constexpr (!MUSCL) {
          if (t < HL::NF) put_prims<ND, PHYS, MUSCL, TC>(s, ih, qh, pc);
        }
      }
      static_assert(HL::N <= NT, "one halo entry per thread");
    } else {
      // deeper halos (MUSCL): the interior, then the list entries
      {
        TC q[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          q[v] = TC(ld_state<TIn>(A.in, v * A.ld + gc));
          s.q[v][ic] = q[v];
        }
        if constexpr (!MUSCL) put_prims<ND, PHYS, MUSCL, TC>(s, ic, q, pc);
      }
      make_P();
      for (int j = t; j < HL::N; j += NT) {
        const unsigned e = kHalo<ND, PHYS, MUSCL>.pos[j];
        unsigned m = kHalo<ND, PHYS, MUSCL>.local[j];
#pragma unroll
        for (int d = 0; d < ND; ++d) {
          const unsigned code = (e >> (12 + 3 * d)) & 7u;
          m |= code == 0 ? P[d][0] : (code == 1 ? P[d][1] : (code == 2 ? P[d][2] : (code == 3 ? P[d][3] : P[d][4])));
        }
        const int i = (int)(e & 0xfffu);
        const long long gi = (long long)m;
        TC q[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          q[v] = TC(ld_state<TIn>(A.in, v * A.ld + gi));
          s.q[v][i] = q[v];
        }
        if constexpr (!MUSCL) {
          if (j < HL::NF) put_prims<ND, PHYS, MUSCL, TC>(s, i, q, pc);
        }
      }
    }
  }
  __syncthreads();

  // 2) faces once per direction, accumulated in d-ascending order
  const TC inv_h = TC(A.inv_h);
  const TC q_coef = TC(pc.q_coef);
  const TC inv_Re0 = TC(pc.inv_Re0);
  TC acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = TC(0);
  brick_direction<ND, PHYS, MUSCL, TC, TABLE, 0>(s, pc, inv_h, q_coef, inv_Re0, cell, acc);
  // U^n of this cell for the stage combination, loaded one pass ahead of
  // its use (read once: non-temporal)
  double bq[NV];
  auto load_base = [&]() {
#pragma unroll
    for (int v = 0; v < NV; ++v)
      bq[v] = (A.kind >= 1 && A.kind <= 3) ? __builtin_nontemporal_load(&A.base[v * A.ld + gc]) : 0.0;
  };
  if constexpr (ND == 3) {
    brick_direction<ND, PHYS, MUSCL, TC, TABLE, 1>(s, pc, inv_h, q_coef, inv_Re0, cell, acc);
    load_base();
    brick_direction<ND, PHYS, MUSCL, TC, TABLE, 2>(s, pc, inv_h, q_coef, inv_Re0, cell, acc);
  } else {
    load_base();
    brick_direction<ND, PHYS, MUSCL, TC, TABLE, 1>(s, pc, inv_h, q_coef, inv_Re0, cell, acc);
  }

  // 3) stage combination (FP64 accumulation of the conserved variables)
  const double dtdx = scal->dtdx;  // dt / h, formed with dt
  double o[NV];
  const int ic = G::lin(cell);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double nf = dtdx * (double)acc[v];
    const double q = (double)s.q[v][ic];
    switch (A.kind) {
      case 0: o[v] = q + nf; break;
      case 1: o[v] = 0.5 * (bq[v] + q + nf); break;
      case 2: o[v] = 0.75 * bq[v] + 0.25 * (q + nf); break;
      case 3: o[v] = (1. / 3.) * bq[v] + (2. / 3.) * (q + nf); break;
      default: o[v] = q + nf; break;
    }
    // written once, read by the next stage kernel: non-temporal
    __builtin_nontemporal_store(TOut(o[v]), &static_cast<TOut*>(A.out)[v * A.ld + gc]);
  }
  if (A.fuse_dt) {
    double qd[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) qd[v] = (double)TOut(o[v]);
    double cd;
    if constexpr (sizeof(TC) == 4) cd = (double)cell_dt_fast_f32<ND, PHYS == HOM3G_CNS>(qd, (float)A.h, pc);
    else cd = cell_dt_fast<ND, PHYS == HOM3G_CNS>(qd, A.h, pc);
    for (int off = 32; off > 0; off >>= 1) cd = ref_min(cd, __shfl_xor(cd, off, 64));
    if ((t & 63) == 0) s.red[t >> 6] = cd;
    __syncthreads();
    if (t == 0) {
      double m = s.red[0];
      for (int k = 1; k < NT / 64; ++k) m = ref_min(m, s.red[k]);
      atomic_lower(&scal->dtmin, m);
    }
  }
  // advance (solver.hpp:502-505) fused into the step's last stage: the stage
  // kernels read no scalar but dt, so one thread may move time and step
  if (A.advance && blockIdx.x == 0 && t == 0) {
    scal->time += scal->dt;
    scal->step += 1;
  }
}

template <int ND, int PHYS, bool MUSCL, class TC, class TIn, class TOut>
void launch(Ctx& C, SolverDev& S, const BrickArgs& A) {
  const long long ncell = S.n_int;
  const long long nb = ncell / BrickCfg<ND>::NT;
  C.prof_mark();
  if constexpr (!MUSCL) {
    if (S.brick_table) {
      hipLaunchKernelGGL((k_brick<ND, PHYS, MUSCL, TC, TIn, TOut, true>), dim3((unsigned)nb), dim3(BrickCfg<ND>::NT),
                         0, C.stream, A, S.pc, S.scal.p);
      H3_LAUNCH_CHECK();
      C.prof_mark();
      return;
    }
  }
  hipLaunchKernelGGL((k_brick<ND, PHYS, MUSCL, TC, TIn, TOut, false>), dim3((unsigned)nb), dim3(BrickCfg<ND>::NT), 0,
                     C.stream, A, S.pc, S.scal.p);
  H3_LAUNCH_CHECK();
  C.prof_mark();
}

template <int ND, int PHYS, bool MUSCL>
void dispatch_precision(Ctx& C, SolverDev& S, const BrickArgs& A, bool in32, bool out32) {
  if (S.P.precision == HOM3G_FP32_STATE) {
    if (in32 && out32) launch<ND, PHYS, MUSCL, float, float, float>(C, S, A);
    else if (in32) launch<ND, PHYS, MUSCL, float, float, double>(C, S, A);
    else if (out32) launch<ND, PHYS, MUSCL, float, double, float>(C, S, A);
    else launch<ND, PHYS, MUSCL, float, double, double>(C, S, A);
  } else {
    launch<ND, PHYS, MUSCL, double, double, double>(C, S, A);
  }
}

// host mirrors of the brick's Morton helpers
int host_morton(int nd, const int* c) {
  int m = 0;
  for (int b = 0; b < 5; ++b)
    for (int d = 0; d < nd; ++d) m |= ((c[d] >> b) & 1) << (nd * b + d);
  return m;
}

template <int ND>
bool table_build_nd(Ctx& C, SolverDev& S) {
  using G = Geo<ND, HOM3G_CNS, false>;  // halo list of the viscous flux (a superset of the Euler one)
  using HL = HaloList<ND, HOM3G_CNS, false>;
  constexpr int B = G::B, NT = BrickCfg<ND>::NT, NH = HL::N;
  const int64_t n_int = S.n_int, cap = S.cap, ng = S.n_ghost;
  if (n_int == 0 || n_int % NT != 0 || S.cap >= (int64_t(1) << 30)) return false;
  const int npos = 2 * ND;
  std::vector<int32_t> nb((size_t)npos * cap), gb(ng > 0 ? ng : 1), gh(ng > 0 ? ng : 1), gi(ng > 0 ? ng : 1);
  S.nghbr.download(nb.data(), nb.size(), C.stream);
  if (ng > 0) {
    S.g_bndry.download(gb.data(), ng, C.stream);
    S.g_hookbc.download(gh.data(), ng, C.stream);
    S.g_image.download(gi.data(), ng, C.stream);
  }
  H3_HIP(hipStreamSynchronize(C.stream));
  auto N = [&](int64_t c, int dir) -> int64_t { return nb[(size_t)dir * cap + c]; };
  // a periodic ghost stands for its image cell (its value and its slopes)
  auto resolve = [&](int64_t id) -> int64_t {
    if (id >= n_int && id < n_int + ng && S.bcs[gh[id - n_int]].kind == HOM3G_BC_PERIODIC) return gi[id - n_int];
    return id;
  };
  auto wall = [&](int kind) {
    return kind == HOM3G_BC_ADIABATIC_NOSLIP || kind == HOM3G_BC_ISOTHERMAL_NOSLIP || kind == HOM3G_BC_CHT_FLUID;
  };
  const int64_t nbrick = n_int / NT;
  std::vector<uint32_t> tab((size_t)nbrick * NH);
  int lc[NT][3];
  for (int t = 0; t < NT; ++t) {
    lc[t][2] = 0;
    for (int d = 0; d < ND; ++d) {
      int x = 0;
      for (int b = 0; b < 5; ++b) x |= ((t >> (ND * b + d)) & 1) << b;
      lc[t][d] = x;
    }
  }
  for (int64_t k = 0; k < nbrick; ++k) {
    const int64_t c0 = k * NT;
    auto cell_at = [&](const int* c) { return c0 + host_morton(ND, c); };
    // the brick's cells are a B^nd block in Morton order: every internal
    // neighbour link matches
    for (int t = 0; t < NT; ++t)
      for (int d = 0; d < ND; ++d) {
        int c[3] = {lc[t][0], lc[t][1], lc[t][2]};
        if (c[d] + 1 < B) {
          int cn[3] = {c[0], c[1], c[2]};
          cn[d] += 1;
          if (N(c0 + t, 2 * d + 1) != cell_at(cn)) return false;
          if (N(cell_at(cn), 2 * d) != c0 + t) return false;
        }
      }
    for (int j = 0; j < NH; ++j) {
      int hc[3] = {0, 0, 0};
      halo_closed_form<ND, HOM3G_CNS, false>(j, hc);
      int in[3] = {hc[0], hc[1], hc[2]};
      int outd[2] = {-1, -1}, nout = 0;
      for (int d = 0; d < ND; ++d)
        if (hc[d] < 0 || hc[d] >= B) {
          outd[nout++] = d;
          in[d] = hc[d] < 0 ? 0 : B - 1;
        }
      const int64_t cin = cell_at(in);
      auto dir_of = [&](int d) { return 2 * d + (hc[d] < 0 ? 0 : 1); };
      uint32_t e;
      if (nout == 1) {
        const int64_t raw = N(cin, dir_of(outd[0]));
        if (raw < 0) return false;  // no neighbour or ghost, or finer cells
        const int64_t id = resolve(raw);
        uint32_t code = 0;
        if (id >= n_int) code = wall(S.bcs[gh[id - n_int]].kind) ? 3u : 1u;
        e = (uint32_t)id | (code << 30);
      } else {
        // an edge cell: reached through either face neighbour; a path
        // through a ghost is not taken (the faces next to that ghost use its
        // slope hook instead of this cell)
        int64_t via[2] = {-1, -1};
        for (int p = 0; p < 2; ++p) {
          const int64_t n1 = N(cin, dir_of(outd[p]));
          if (n1 < 0) return false;
          const int64_t r1 = resolve(n1);
          if (r1 >= n_int) continue;
          const int64_t n2 = N(r1, dir_of(outd[1 - p]));
          if (n2 < 0) return false;
          via[p] = resolve(n2);
        }
        if (via[0] >= 0 && via[1] >= 0 && via[0] != via[1]) return false;  // concave corner
        const int64_t id = via[0] >= 0 ? via[0] : via[1];
        e = id >= 0 ? (uint32_t)id : ((uint32_t)c0 | (2u << 30));  // 2: not read
      }
      tab[(size_t)k * NH + j] = e;
    }
  }
  S.brick_halo.alloc(tab.size());
  S.brick_halo.upload(tab.data(), tab.size(), C.stream);
  // reverse map for the adjoint's gather: per cell, the halo entries
  // (brick * NH + j) that read it, in increasing entry order
  const int64_t nc = n_int + ng;
  std::vector<int32_t> rp((size_t)nc + 1, 0), src;
  for (size_t e = 0; e < tab.size(); ++e)
    if ((tab[e] >> 30) != 2u) ++rp[(tab[e] & 0x3fffffffu) + 1];
  for (int64_t c = 0; c < nc; ++c) rp[c + 1] += rp[c];
  src.resize((size_t)rp[nc]);
  std::vector<int32_t> pos(rp.begin(), rp.end() - 1);
  for (size_t e = 0; e < tab.size(); ++e)
    if ((tab[e] >> 30) != 2u) src[(size_t)pos[tab[e] & 0x3fffffffu]++] = (int32_t)e;
  S.brick_rev_rp.alloc(rp.size());
  S.brick_rev_rp.upload(rp.data(), rp.size(), C.stream);
  std::vector<int32_t> tl;
  for (int64_t c = 0; c < nc; ++c)
    if (rp[c + 1] > rp[c]) tl.push_back((int32_t)c);
  S.brick_rev_nt = (int64_t)tl.size();
  S.brick_rev_tl.alloc(std::max<size_t>(1, tl.size()));
  if (!tl.empty()) S.brick_rev_tl.upload(tl.data(), tl.size(), C.stream);
  S.brick_rev_src.alloc(std::max<size_t>(1, src.size()));
  if (!src.empty()) S.brick_rev_src.upload(src.data(), src.size(), C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  return true;
}

}  // namespace

bool brick_supported(const SolverDev& S) {
  if (S.P.physics == HOM3G_HEAT) return false;
  if (S.P.physics == HOM3G_CNS && S.P.recon != HOM3G_RECON_NONE) return false;
  if (S.P.physics == HOM3G_EULER && S.P.flux == HOM3G_FLUX_LLF && S.P.recon == HOM3G_RECON_NONE) return false;
  if (S.P.flux == HOM3G_FLUX_LLF) return false;
  if (S.nd == 3 && S.level < 3) return false;
  if (S.nd == 2 && S.level < 4) return false;
  if (S.nd == 3 && S.level > 10) return false;
  if (S.nd == 2 && S.level > 15) return false;
  return true;
}

// Tiled bricks: a uniform solver whose internal cells fall into full B^nd
// bricks, each a Morton-ordered run of consecutive cell ids (the leaf order
// of a domain made of whole bricks), with any boundary conditions.  Each
// brick's halo cells come from a table built here from the neighbour map:
// internal cells, ghosts (filled before each stage by the ghost kernel) or,
// for periodic ghosts, their image cells.  Not for MUSCL, host callbacks,
// the isentropic-vortex condition (it rewrites boundary cells) or grids
// where an edge cell of a brick is two different ghosts (concave corners).
bool brick_table_build(Ctx& C, SolverDev& S) {
  if (S.nonuniform || S.P.recon != HOM3G_RECON_NONE || S.has_host_bc || S.uses_vortex_bc) return false;
  if (S.P.physics == HOM3G_HEAT || S.P.flux == HOM3G_FLUX_LLF) return false;
  if (S.level < 0 || (S.nd == 3 && S.level > 10) || (S.nd == 2 && S.level > 15)) return false;
  for (const BCDesc& b : S.bcs)
    if (b.kind == HOM3G_BC_EULER_COUPLING) return false;
  return S.nd == 3 ? table_build_nd<3>(C, S) : table_build_nd<2>(C, S);
}

// stage: see StageArgs kinds; in/out are the device buffers for this stage
void brick_stage(Ctx& C, SolverDev& S, int stage, const void* in, void* out, bool fuse_dt, bool advance) {
  BrickArgs A;
  A.in = in;
  A.base = S.lhs.p;
  A.out = out;
  A.ld = S.cap;
  A.level = S.level;
  for (int d = 0; d < 3; ++d) A.periodic[d] = (d < S.nd) ? S.face_mode[d][0] : 0;
  A.h = S.h_uniform;
  A.inv_h = 1.0 / S.h_uniform;
  A.kind = stage;
  A.fuse_dt = fuse_dt ? 1 : 0;
  A.advance = advance ? 1 : 0;
  A.halo = S.brick_table ? S.brick_halo.p : nullptr;
  const bool in32 = (in == (const void*)S.rhs32.p || in == (const void*)S.stg32.p) && in != nullptr;
  const bool out32 = (out == (void*)S.rhs32.p || out == (void*)S.stg32.p) && out != nullptr;
  const bool muscl = S.P.recon == HOM3G_RECON_MUSCL;
  if (S.nd == 3) {
    if (S.P.physics == HOM3G_CNS) dispatch_precision<3, HOM3G_CNS, false>(C, S, A, in32, out32);
    else if (muscl) dispatch_precision<3, HOM3G_EULER, true>(C, S, A, in32, out32);
    else dispatch_precision<3, HOM3G_EULER, false>(C, S, A, in32, out32);
  } else {
    if (S.P.physics == HOM3G_CNS) dispatch_precision<2, HOM3G_CNS, false>(C, S, A, in32, out32);
    else if (muscl) dispatch_precision<2, HOM3G_EULER, true>(C, S, A, in32, out32);
    else dispatch_precision<2, HOM3G_EULER, false>(C, S, A, in32, out32);
  }
}

}  // namespace hom3g
