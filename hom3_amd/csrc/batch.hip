// Batched solves: many independent first-order Euler instances of uneven
// size (each with its own grid, 2:1 refinement and time step) advanced by
// one persistent kernel per RK stage, with the whole step captured in a
// hipGraph and the launch shape picked by a timing autotuner.
//
// There is no batching in the reference; each instance follows the single
// solver's step exactly (solver.hpp:140-145: set_dt, evolve, advance) with
// its own dt, time and step, so a batch of one equals hom3g_solver_solve.
//
// Layout.  The instances' internal cells are concatenated, each instance
// keeping its own (Morton leaf) order: one SoA state array per stage buffer
// with ld = total cells.  Ghost cells are resolved into the neighbour table
// at build time: a zero-gradient Neumann ghost holds a copy of its boundary
// cell (boundary_condition.hpp:70-77 with g = 0) and a periodic ghost a copy
// of its image, so the face to that ghost reads the boundary cell or the
// image directly and no ghost-fill pass is needed.
//
// Work items are tiles of consecutive cells of one instance.  A fixed grid
// of workgroups sweeps the item list (persistent kernel); the final stage
// also reduces each item's CFL bound of the new state into its instance's
// running minimum (atomic min on an order-preserving integer encoding, exact
// and order independent).
#include <algorithm>
#include <cstring>

#include "internal.hpp"

namespace hom3g {

namespace {

inline unsigned nblk(int64_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

struct BatchArgs {
  const BatchItem* items;
  long long nitems;
  const void* in;
  const double* base;
  void* out;
  long long ld;
  const int32_t* nghbr;
  const int32_t* fine;
  const double* len;
  DevScalars* scal;
  int kind;     // 0: in+nf  1: 0.5*((base+in)+nf)  2: 0.75 base + 0.25 (in+nf)  3: 1/3 base + 2/3 (in+nf)
  int fuse_dt;  // reduce the CFL bound of the new state per instance
};

template <class T> __device__ __forceinline__ T ld_as(const void* p, long long i);
template <> __device__ __forceinline__ double ld_as<double>(const void* p, long long i) {
  return static_cast<const double*>(p)[i];
}
template <> __device__ __forceinline__ float ld_as<float>(const void* p, long long i) {
  return static_cast<const float*>(p)[i];
}

template <int ND, class TC, class TIn>
__device__ __forceinline__ void load_cell(const void* in, long long ld, long long c, TC* q) {
#pragma unroll
  for (int v = 0; v < ND + 2; ++v) q[v] = TC(ld_as<TIn>(in, v * ld + c));
}

// AUSM through one side of cell c along direction D (euler/physics.hpp:276-325):
// towards one cell, or the mean of the pairwise-summed sub-face fluxes of a
// face towards finer cells (nb <= -2)
template <int ND, int D, class TC, class TIn>
__device__ __forceinline__ void side(const BatchArgs& A, const PhysConst& pc, const TC* qc, const Prim<ND, TC>& wc,
                                     int nb, bool minus, TC* f) {
  constexpr int NV = ND + 2;
  const TC gamma = TC(pc.gamma), gm1 = TC(pc.gm1);
  if (nb >= 0) {
    TC qn[NV];
    load_cell<ND, TC, TIn>(A.in, A.ld, nb, qn);
    const Prim<ND, TC> wn = primitives_rcp<ND>(qn, gamma, gm1);
    if (minus) ausm_flux_rcp<ND, D, TC>(qn, wn, qc, wc, f);
    else ausm_flux_rcp<ND, D, TC>(qc, wc, qn, wn, f);
    return;
  }
  constexpr int K = 1 << (ND - 1);
  const int32_t* fc = A.fine + (long long)(-(nb + 2)) * K;
  TC sub[K][NV];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    TC qn[NV];
    load_cell<ND, TC, TIn>(A.in, A.ld, fc[k], qn);
    const Prim<ND, TC> wn = primitives_rcp<ND>(qn, gamma, gm1);
    if (minus) ausm_flux_rcp<ND, D, TC>(qn, wn, qc, wc, sub[k]);
    else ausm_flux_rcp<ND, D, TC>(qc, wc, qn, wn, sub[k]);
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const TC sm = (ND == 2) ? sub[0][v] + sub[1][v] : (sub[0][v] + sub[1][v]) + (sub[2 % K][v] + sub[3 % K][v]);
    f[v] = sm * TC((ND == 2) ? 0.5 : 0.25);
  }
}

// one side with the same-level neighbour's state already loaded (qn); a
// face towards finer cells gathers its sub-cells here
template <int ND, int D, class TC, class TIn>
__device__ __forceinline__ void side_pre(const BatchArgs& A, const PhysConst& pc, const TC* qc, const Prim<ND, TC>& wc,
                                         int nb, const TC* qn, bool minus, TC* f) {
  if (nb >= 0) {
    const Prim<ND, TC> wn = primitives_rcp<ND>(qn, TC(pc.gamma), TC(pc.gm1));
    if (minus) ausm_flux_rcp<ND, D, TC>(qn, wn, qc, wc, f);
    else ausm_flux_rcp<ND, D, TC>(qc, wc, qn, wn, f);
  } else {
    side<ND, D, TC, TIn>(A, pc, qc, wc, nb, minus, f);
  }
}

template <int ND, int D, class TC, class TIn>
__device__ __forceinline__ void direction_pre(const BatchArgs& A, const PhysConst& pc, const TC* qc,
                                              const Prim<ND, TC>& wc, const int* nb, const TC (*qn)[ND + 2],
                                              TC* acc) {
  constexpr int NV = ND + 2;
  TC fm[NV], fp[NV];
  side_pre<ND, D, TC, TIn>(A, pc, qc, wc, nb[2 * D], qn[2 * D], true, fm);
  side_pre<ND, D, TC, TIn>(A, pc, qc, wc, nb[2 * D + 1], qn[2 * D + 1], false, fp);
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] += (fm[v] - fp[v]);
}

template <int ND, int D, class TC, class TIn>
__device__ __forceinline__ void direction(const BatchArgs& A, const PhysConst& pc, long long c, const TC* qc,
                                          const Prim<ND, TC>& wc, TC* acc) {
  constexpr int NV = ND + 2;
  const int m = A.nghbr[(2 * D) * A.ld + c];
  const int p = A.nghbr[(2 * D + 1) * A.ld + c];
  TC fm[NV], fp[NV];
  side<ND, D, TC, TIn>(A, pc, qc, wc, m, true, fm);
  side<ND, D, TC, TIn>(A, pc, qc, wc, p, false, fp);
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] += (fm[v] - fp[v]);
}

template <int ND, class TC, class TIn, class TOut>
__global__ void __launch_bounds__(512) k_batch_stage(BatchArgs A, PhysConst pc) {
  constexpr int NV = ND + 2;
  __shared__ double red[16];
  for (long long it = blockIdx.x; it < A.nitems; it += gridDim.x) {
    const BatchItem I = A.items[it];
    const double dt = A.scal[I.inst].dt;
    double cmin = 1.7976931348623157e308;
    for (long long c = I.begin + threadIdx.x; c < I.end; c += blockDim.x) {
      // the cell, its neighbour ids and the same-level neighbours' states,
      // all in flight before the first flux
      TC qc[NV];
      load_cell<ND, TC, TIn>(A.in, A.ld, c, qc);
      int nb[2 * ND];
#pragma unroll
      for (int p = 0; p < 2 * ND; ++p) nb[p] = A.nghbr[p * A.ld + c];
      TC qn[2 * ND][NV];
#pragma unroll
      for (int p = 0; p < 2 * ND; ++p) load_cell<ND, TC, TIn>(A.in, A.ld, nb[p] >= 0 ? (long long)nb[p] : c, qn[p]);
      const Prim<ND, TC> wc = primitives_rcp<ND>(qc, TC(pc.gamma), TC(pc.gm1));
      TC acc[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) acc[v] = TC(0);
      direction_pre<ND, 0, TC, TIn>(A, pc, qc, wc, nb, qn, acc);
      direction_pre<ND, 1, TC, TIn>(A, pc, qc, wc, nb, qn, acc);
      if constexpr (ND == 3) direction_pre<ND, 2, TC, TIn>(A, pc, qc, wc, nb, qn, acc);
      const double h = A.len[c];
      const double dtdx = dt / h;
      double o[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const double nf = dtdx * (double)acc[v];
        const double q = (double)qc[v];
        switch (A.kind) {
          case 0: o[v] = q + nf; break;
          case 1: o[v] = 0.5 * (A.base[v * A.ld + c] + q + nf); break;
          case 2: o[v] = 0.75 * A.base[v * A.ld + c] + 0.25 * (q + nf); break;
          case 3: o[v] = (1. / 3.) * A.base[v * A.ld + c] + (2. / 3.) * (q + nf); break;
          default: o[v] = q + nf; break;
        }
        static_cast<TOut*>(A.out)[v * A.ld + c] = TOut(o[v]);
      }
      if (A.fuse_dt) {
        double qd[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) qd[v] = (double)TOut(o[v]);
        const Prim<ND, double> w = primitives<ND>(qd, pc.gamma, pc.gm1);
        cmin = ref_min(cmin, cell_dt_euler<ND, double>(w, h, pc.cfl));
      }
    }
    if (A.fuse_dt) {
      for (int o = 32; o > 0; o >>= 1) cmin = ref_min(cmin, __shfl_xor(cmin, o, 64));
      const int nw = (blockDim.x + 63) >> 6;
      if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cmin;
      __syncthreads();
      if (threadIdx.x == 0) {
        double m = red[0];
        for (int k = 1; k < nw; ++k) m = ref_min(m, red[k]);
        atomic_lower(&A.scal[I.inst].dtmin, m);
      }
      __syncthreads();
    }
  }
}

// compute_dt (solver.hpp:511-517) of the current FP64 state, per instance
template <int ND>
__global__ void __launch_bounds__(512) k_batch_dt(const BatchItem* items, long long nitems, const double* U,
                                                   long long ld, const double* len, DevScalars* scal,
                                                   PhysConst pc) {
  constexpr int NV = ND + 2;
  __shared__ double red[16];
  for (long long it = blockIdx.x; it < nitems; it += gridDim.x) {
    const BatchItem I = items[it];
    double cmin = 1.7976931348623157e308;
    for (long long c = I.begin + threadIdx.x; c < I.end; c += blockDim.x) {
      double q[NV];
#pragma unroll
      for (int v = 0; v < NV; ++v) q[v] = U[v * ld + c];
      const Prim<ND, double> w = primitives<ND>(q, pc.gamma, pc.gm1);
      cmin = ref_min(cmin, cell_dt_euler<ND, double>(w, len[c], pc.cfl));
    }
    for (int o = 32; o > 0; o >>= 1) cmin = ref_min(cmin, __shfl_xor(cmin, o, 64));
    const int nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cmin;
    __syncthreads();
    if (threadIdx.x == 0) {
      double m = red[0];
      for (int k = 1; k < nw; ++k) m = ref_min(m, red[k]);
      atomic_lower(&scal[I.inst].dtmin, m);
    }
    __syncthreads();
  }
}

// per instance: set_dt (solver.hpp:541-551) and reset of the running minimum
__global__ void k_batch_set_dt(int n, DevScalars* s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  DevScalars& S = s[i];
  const double dtm = ordered_to_dbl(S.dtmin);
  if (S.time + S.dt < S.tEnd) S.dt = (S.step == S.forced_step) ? S.forced_dt : dtm;
  else S.dt = S.tEnd - S.time;
  S.dtmin = dbl_to_ordered(1.7976931348623157e308);
}

// per instance: advance (solver.hpp:502-505)
__global__ void k_batch_advance(int n, DevScalars* s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  s[i].time += s[i].dt;
  s[i].step += 1;
}

__global__ void k_batch_reset(int n, DevScalars* s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) s[i].dtmin = dbl_to_ordered(1.7976931348623157e308);
}

// forward Euler's lhs = rhs copy (solver.hpp:473-476)
__global__ void k_batch_copy(const double* __restrict__ a, double* __restrict__ b, long long n) {
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    b[i] = a[i];
}

// neighbour table of one instance into the batch: cells + offset, ghosts
// resolved to the boundary cell (Neumann) or the periodic image, faces
// towards finer cells + fine-face offset
__global__ void k_batch_nghbr(int npos, long long n_int, long long cap, const int32_t* src, const int32_t* g_apply,
                              const int32_t* g_image, const int32_t* bc_kind, long long off, long long foff,
                              long long ld, int32_t* dst, int* err) {
  const long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (t >= n_int * npos) return;
  const long long c = t % n_int;
  const int pos = (int)(t / n_int);
  const int v = src[pos * cap + c];
  int r;
  if (v >= n_int) {
    const long long g = v - n_int;
    const int a = g_apply[g];
    const int k = (a >= 0 && a < kMaxBC) ? bc_kind[a] : -1;
    if (k == HOM3G_BC_NEUMANN) r = (int)(off + c);
    else if (k == HOM3G_BC_PERIODIC) r = (int)(off + g_image[g]);
    else { r = (int)(off + c); atomicOr(err, 1); }
  } else if (v >= 0) {
    r = (int)(off + v);
  } else if (v <= -2) {
    r = (int)(-(2 + (foff + (-(v + 2)))));
  } else {
    r = (int)(off + c);
    atomicOr(err, 2);
  }
  dst[pos * ld + off + c] = r;
}

__global__ void k_batch_fine(long long n, const int32_t* src, long long off, int32_t* dst) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (int32_t)(src[i] + off);
}

template <int ND, class TC, class TIn, class TOut>
void launch_stage_t(Ctx& C, BatchDev& B, const BatchArgs& A) {
  hipLaunchKernelGGL((k_batch_stage<ND, TC, TIn, TOut>), dim3(B.grid), dim3(B.block), 0, C.stream, A, B.pc);
  H3_LAUNCH_CHECK();
}

template <int ND>
void launch_stage_nd(Ctx& C, BatchDev& B, const BatchArgs& A, bool in32, bool out32) {
  if (!in32 && !out32) launch_stage_t<ND, double, double, double>(C, B, A);
  else if (!in32 && out32) launch_stage_t<ND, float, double, float>(C, B, A);
  else if (in32 && out32) launch_stage_t<ND, float, float, float>(C, B, A);
  else launch_stage_t<ND, float, float, double>(C, B, A);
}

void launch_stage(Ctx& C, BatchDev& B, int kind, const void* in, bool in32, const double* base, void* out,
                  bool out32, bool fuse) {
  BatchArgs A;
  A.items = B.items.p;
  A.nitems = (long long)B.h_items.size();
  A.in = in;
  A.base = base;
  A.out = out;
  A.ld = B.ncell;
  A.nghbr = B.nghbr.p;
  A.fine = B.fine.p;
  A.len = B.len.p;
  A.scal = B.scal.p;
  A.kind = kind;
  A.fuse_dt = fuse ? 1 : 0;
  if (B.nd == 2) launch_stage_nd<2>(C, B, A, in32, out32);
  else launch_stage_nd<3>(C, B, A, in32, out32);
}

void launch_dt(Ctx& C, BatchDev& B) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(B.grid), dim3(B.block), 0, C.stream, (const BatchItem*)B.items.p,
                       (long long)B.h_items.size(), (const double*)B.U.p, (long long)B.ncell,
                       (const double*)B.len.p, B.scal.p, B.pc);
    H3_LAUNCH_CHECK();
  };
  if (B.nd == 2) go(k_batch_dt<2>);
  else go(k_batch_dt<3>);
}

void make_items(BatchDev& B, int tile) {
  B.h_items.clear();
  for (int i = 0; i < B.ninst; ++i)
    for (int64_t b = B.off[i]; b < B.off[i] + B.nint[i]; b += tile) {
      BatchItem I;
      I.inst = i;
      I.pad = 0;
      I.begin = b;
      I.end = std::min<int64_t>(b + tile, B.off[i] + B.nint[i]);
      B.h_items.push_back(I);
    }
  B.items.alloc(B.h_items.size());
  B.items.upload(B.h_items.data(), B.h_items.size(), nullptr);
  H3_HIP(hipDeviceSynchronize());
  B.tile = tile;
}

void drop_graph(BatchDev& B) {
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
}

// One step of every instance (solver.hpp:140-145): [compute_dt] set_dt,
// evolve (solver.hpp:469-499 + SSP-RK3), advance.  After the first step the
// CFL bound of the new state comes fused out of the last stage.
void enqueue_step(Ctx& C, BatchDev& B) {
  const int nthr = 256;
  hipLaunchKernelGGL(k_batch_set_dt, dim3(nblk(B.ninst, nthr)), dim3(nthr), 0, C.stream, B.ninst, B.scal.p);
  H3_LAUNCH_CHECK();
  const bool f32 = B.precision == HOM3G_FP32_STATE;
  void* s1 = f32 ? (void*)B.S1f.p : (void*)B.S1d.p;
  void* s2 = f32 ? (void*)B.S2f.p : (void*)B.S2d.p;
  if (B.ti == HOM3G_TI_SSP_RK3) {
    launch_stage(C, B, 0, B.U.p, false, nullptr, s1, f32, false);
    launch_stage(C, B, 2, s1, f32, B.U.p, s2, f32, false);
    launch_stage(C, B, 3, s2, f32, B.U.p, B.U.p, false, true);
  } else if (B.ti == HOM3G_TI_HEUN_RK2) {
    launch_stage(C, B, 0, B.U.p, false, nullptr, s1, f32, false);
    launch_stage(C, B, 1, s1, f32, B.U.p, B.U.p, false, true);
  } else {
    launch_stage(C, B, 0, B.U.p, false, nullptr, B.S1d.p, false, false);
    const long long n = (long long)B.nvars * B.ncell;
    hipLaunchKernelGGL(k_batch_copy, dim3(2048), dim3(256), 0, C.stream, (const double*)B.S1d.p, B.U.p, n);
    H3_LAUNCH_CHECK();
    launch_dt(C, B);
  }
  hipLaunchKernelGGL(k_batch_advance, dim3(nblk(B.ninst, nthr)), dim3(nthr), 0, C.stream, B.ninst, B.scal.p);
  H3_LAUNCH_CHECK();
}

void build_graph(Ctx& C, BatchDev& B) {
  if (B.exec) return;
  hipGraph_t g = nullptr;
  H3_HIP(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
  try {
    enqueue_step(C, B);
  } catch (...) {
    (void)hipStreamEndCapture(C.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  H3_HIP(hipStreamEndCapture(C.stream, &g));
  B.graph = g;
  H3_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
}

void ensure_dtmin(Ctx& C, BatchDev& B) {
  if (B.dtmin_current) return;
  hipLaunchKernelGGL(k_batch_reset, dim3(nblk(B.ninst)), dim3(256), 0, C.stream, B.ninst, B.scal.p);
  H3_LAUNCH_CHECK();
  launch_dt(C, B);
  B.dtmin_current = true;
}

}  // namespace

void batch_create(Ctx& C, const std::vector<SolverDev*>& inst, int precision, BatchDev& B) {
  H3_REQUIRE(!inst.empty(), "empty batch");
  const SolverDev& S0 = *inst[0];
  B.nd = S0.nd;
  B.nvars = S0.nvars;
  B.ninst = (int)inst.size();
  B.precision = precision;
  B.ti = S0.P.ti;
  B.pc = S0.pc;
  for (SolverDev* S : inst) {
    H3_REQUIRE(S->built, "batch instance not built");
    H3_REQUIRE(S->P.physics == HOM3G_EULER && S->P.flux == HOM3G_FLUX_AUSM && S->P.recon == HOM3G_RECON_NONE,
               "batched instances must be first-order Euler AUSM");
    H3_REQUIRE(S->nd == B.nd && S->P.ti == B.ti && S->P.gamma == S0.P.gamma && S->P.cfl == S0.P.cfl,
               "batched instances must share nd, time integration, gamma and CFL");
    for (const BCDesc& b : S->bcs) {
      bool ok = b.kind == HOM3G_BC_PERIODIC;
      if (b.kind == HOM3G_BC_NEUMANN) {
        ok = true;
        for (int k = 0; k < b.nparams; ++k) ok = ok && b.params[k] == 0.0;
      }
      H3_REQUIRE(ok, "batched instances support zero-gradient Neumann and periodic boundaries only");
    }
  }
  B.off.resize(B.ninst);
  B.nint.resize(B.ninst);
  int64_t N = 0, NF = 0;
  std::vector<int64_t> foff(B.ninst);
  for (int i = 0; i < B.ninst; ++i) {
    B.off[i] = N;
    B.nint[i] = inst[i]->n_int;
    foff[i] = NF;
    N += inst[i]->n_int;
    NF += inst[i]->nfine;
  }
  H3_REQUIRE(N < INT32_MAX / 2, "too many cells in the batch for 32-bit ids");
  B.ncell = N;
  B.nfine = NF;
  const int npos = 2 * B.nd, K = 1 << (B.nd - 1), nv = B.nvars;
  B.U.alloc((size_t)nv * N);
  if (precision == HOM3G_FP32_STATE) {
    B.S1f.alloc((size_t)nv * N);
    B.S2f.alloc((size_t)nv * N);
    if (B.ti == HOM3G_TI_EULER_FWD) B.S1d.alloc((size_t)nv * N);
  } else {
    B.S1d.alloc((size_t)nv * N);
    B.S2d.alloc((size_t)nv * N);
  }
  B.nghbr.alloc((size_t)npos * N);
  B.fine.alloc(NF > 0 ? (size_t)NF * K : 1);
  B.len.alloc(N);
  B.scal.alloc(B.ninst);
  DevBuf<int> err;
  err.alloc(1);
  int zero = 0;
  err.upload(&zero, 1, C.stream);
  DevBuf<int32_t> kinds;
  kinds.alloc(kMaxBC);
  for (int i = 0; i < B.ninst; ++i) {
    SolverDev& S = *inst[i];
    const int64_t n = S.n_int, off = B.off[i];
    H3_HIP(hipMemcpy2DAsync(B.U.p + off, (size_t)N * sizeof(double), S.lhs.p, (size_t)S.cap * sizeof(double),
                            (size_t)n * sizeof(double), nv, hipMemcpyDeviceToDevice, C.stream));
    H3_HIP(hipMemcpyAsync(B.len.p + off, S.length.p, n * sizeof(double), hipMemcpyDeviceToDevice, C.stream));
    H3_HIP(hipMemcpyAsync(B.scal.p + i, S.scal.p, sizeof(DevScalars), hipMemcpyDeviceToDevice, C.stream));
    int32_t hk[kMaxBC];
    for (int b = 0; b < kMaxBC; ++b) hk[b] = b < (int)S.bcs.size() ? S.bcs[b].kind : -1;
    kinds.upload(hk, kMaxBC, C.stream);
    hipLaunchKernelGGL(k_batch_nghbr, dim3(nblk(n * npos)), dim3(256), 0, C.stream, npos, (long long)n,
                       (long long)S.cap, (const int32_t*)S.nghbr.p, (const int32_t*)S.g_applybc.p,
                       (const int32_t*)S.g_image.p, (const int32_t*)kinds.p, (long long)off, (long long)foff[i],
                       (long long)N, B.nghbr.p, err.p);
    H3_LAUNCH_CHECK();
    if (S.nfine > 0) {
      hipLaunchKernelGGL(k_batch_fine, dim3(nblk(S.nfine * K)), dim3(256), 0, C.stream, (long long)S.nfine * K,
                         (const int32_t*)S.fine.p, (long long)off, B.fine.p + foff[i] * K);
      H3_LAUNCH_CHECK();
    }
    H3_HIP(hipStreamSynchronize(C.stream));  // kinds is reused
  }
  int e = 0;
  err.download(&e, 1, C.stream);
  H3_HIP(hipStreamSynchronize(C.stream));
  H3_REQUIRE(e == 0, "batched instance has a face without neighbour or an unsupported boundary");
  // default launch shape until autotuned
  int dev = 0, ncu = 0;
  H3_HIP(hipGetDevice(&dev));
  H3_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  B.ncu = ncu > 0 ? ncu : 256;
  B.block = 256;
  B.grid = B.ncu * 4;
  make_items(B, 1024);
  B.dtmin_current = false;
}

void batch_set_config(Ctx& C, BatchDev& B, int block, int tile, int grid) {
  H3_REQUIRE(block >= 64 && block <= 512 && block % 64 == 0, "block must be a multiple of 64 in [64, 512]");
  H3_REQUIRE(tile >= block, "tile must hold at least one cell per thread");
  H3_REQUIRE(grid >= 1, "grid must be positive");
  (void)C;
  B.block = block;
  B.grid = grid;
  if (tile != B.tile) make_items(B, tile);
  drop_graph(B);
}

void batch_solve(Ctx& C, BatchDev& B, int nsteps, bool use_graph) {
  H3_REQUIRE(nsteps >= 0, "nsteps must be non-negative");
  if (nsteps == 0) return;
  ensure_dtmin(C, B);
  if (!use_graph) {
    for (int s = 0; s < nsteps; ++s) enqueue_step(C, B);
    return;
  }
  build_graph(C, B);
  for (int s = 0; s < nsteps; ++s) H3_HIP(hipGraphLaunch(B.exec, C.stream));
}

// Times candidate launch shapes on stage 1 of the current state with the
// current dt (written to a stage buffer: neither the state nor the
// instances' scalars change) and keeps the fastest.
void batch_autotune(Ctx& C, BatchDev& B, int* best_block, int* best_tile, int* best_grid, float* best_ms) {
  const bool f32 = B.precision == HOM3G_FP32_STATE;
  void* s1 = f32 ? (void*)B.S1f.p : (void*)B.S1d.p;
  const int blocks[] = {128, 256, 512};
  const int tiles_per_block[] = {1, 4, 16};
  // up to 64 workgroups per CU: on a large single grid about one item per
  // workgroup, the persistent sweep degenerating to a plain launch
  const int wgs_per_cu[] = {2, 4, 8, 16, 64};
  float bms = 1e30f;
  int bb = B.block, bt = B.tile, bg = B.grid;
  for (int blk : blocks)
    for (int tpb : tiles_per_block)
      for (int w : wgs_per_cu) {
        const int tile = blk * tpb;
        const int grid = B.ncu * w;
        batch_set_config(C, B, blk, tile, grid);
        launch_stage(C, B, 0, B.U.p, false, nullptr, s1, f32, false);  // warm
        float tmin = 1e30f;
        for (int r = 0; r < 3; ++r) {
          H3_HIP(hipEventRecord(C.ev0, C.stream));
          launch_stage(C, B, 0, B.U.p, false, nullptr, s1, f32, false);
          H3_HIP(hipEventRecord(C.ev1, C.stream));
          H3_HIP(hipEventSynchronize(C.ev1));
          float ms = 0;
          H3_HIP(hipEventElapsedTime(&ms, C.ev0, C.ev1));
          tmin = std::min(tmin, ms);
        }
        if (tmin < bms) { bms = tmin; bb = blk; bt = tile; bg = grid; }
      }
  batch_set_config(C, B, bb, bt, bg);
  if (best_block) *best_block = bb;
  if (best_tile) *best_tile = bt;
  if (best_grid) *best_grid = bg;
  if (best_ms) *best_ms = bms;
}

// events around nsteps steps (graph or stream launches), ms per step
float batch_time_solve(Ctx& C, BatchDev& B, int nsteps, bool use_graph) {
  ensure_dtmin(C, B);
  if (use_graph) build_graph(C, B);  // capture outside the timed region
  H3_HIP(hipEventRecord(C.ev0, C.stream));
  batch_solve(C, B, nsteps, use_graph);
  H3_HIP(hipEventRecord(C.ev1, C.stream));
  H3_HIP(hipEventSynchronize(C.ev1));
  float ms = 0;
  H3_HIP(hipEventElapsedTime(&ms, C.ev0, C.ev1));
  return nsteps > 0 ? ms / nsteps : 0.f;
}

void batch_download(Ctx& C, BatchDev& B, int inst, double* out) {
  H3_REQUIRE(inst >= 0 && inst < B.ninst, "instance out of range");
  const int64_t n = B.nint[inst];
  for (int v = 0; v < B.nvars; ++v)
    H3_HIP(hipMemcpyAsync(out + (size_t)v * n, B.U.p + (size_t)v * B.ncell + B.off[inst], n * sizeof(double),
                          hipMemcpyDeviceToHost, C.stream));
  H3_HIP(hipStreamSynchronize(C.stream));
}

void batch_scalars(Ctx& C, BatchDev& B, int inst, DevScalars* out) {
  H3_REQUIRE(inst >= 0 && inst < B.ninst, "instance out of range");
  H3_HIP(hipMemcpyAsync(out, B.scal.p + inst, sizeof(DevScalars), hipMemcpyDeviceToHost, C.stream));
  H3_HIP(hipStreamSynchronize(C.stream));
}

void batch_release(BatchDev& B) { drop_graph(B); }

}  // namespace hom3g
