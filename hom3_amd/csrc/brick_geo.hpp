// Geometry of the brick kernels (forward stages in brick.hip, reverse
// stages in adjoint.hip): brick sizes, Morton helpers, the LDS layout of a
// brick with its halo and the halo lists.
#pragma once
#include "internal.hpp"

namespace hom3g {
namespace {

template <int ND> struct BrickCfg;
template <> struct BrickCfg<3> { static constexpr int B = 8, NT = 512, LOG = 3; };
template <> struct BrickCfg<2> { static constexpr int B = 16, NT = 256, LOG = 4; };

__device__ __forceinline__ unsigned part1by1(unsigned x) {
  x &= 0x0000ffffu;
  x = (x | (x << 8)) & 0x00ff00ffu;
  x = (x | (x << 4)) & 0x0f0f0f0fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
__device__ __forceinline__ unsigned compact1by1(unsigned x) {
  x &= 0x55555555u;
  x = (x | (x >> 1)) & 0x33333333u;
  x = (x | (x >> 2)) & 0x0f0f0f0fu;
  x = (x | (x >> 4)) & 0x00ff00ffu;
  x = (x | (x >> 8)) & 0x0000ffffu;
  return x;
}
__device__ __forceinline__ unsigned part1by2(unsigned x) {
  x &= 0x000003ffu;
  x = (x | (x << 16)) & 0xff0000ffu;
  x = (x | (x << 8)) & 0x0300f00fu;
  x = (x | (x << 4)) & 0x030c30c3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__device__ __forceinline__ unsigned compact1by2(unsigned x) {
  x &= 0x09249249u;
  x = (x | (x >> 2)) & 0x030c30c3u;
  x = (x | (x >> 4)) & 0x0300f00fu;
  x = (x | (x >> 8)) & 0xff0000ffu;
  x = (x | (x >> 16)) & 0x000003ffu;
  return x;
}

template <int ND>
__device__ __forceinline__ long long morton(const int* g) {
  if constexpr (ND == 3) return (long long)(part1by2(g[0]) | (part1by2(g[1]) << 1) | (part1by2(g[2]) << 2));
  else return (long long)(part1by1(g[0]) | (part1by1(g[1]) << 1));
}
template <int ND>
__device__ __forceinline__ void demorton(unsigned m, int* g) {
  if constexpr (ND == 3) { g[0] = compact1by2(m); g[1] = compact1by2(m >> 1); g[2] = compact1by2(m >> 2); }
  else { g[0] = compact1by1(m); g[1] = compact1by1(m >> 1); }
}


template <int ND, int PHYS, bool MUSCL>
struct Geo {
  static constexpr int B = BrickCfg<ND>::B;
  static constexpr int H = MUSCL ? 2 : 1;
  static constexpr int E = B + 2 * H;
  static constexpr int NPOS = (ND == 3) ? E * E * E : E * E;
  static constexpr int NV = ND + 2;
  static constexpr int NFACE = (B + 1) * (ND == 3 ? B * B : B);
  static constexpr int TANG = (ND == 3) ? B * B : B;
  static constexpr bool VISC = (PHYS == HOM3G_CNS);
  __host__ __device__ static __forceinline__ int lin(int x, int y, int z) {
    if constexpr (ND == 3) return ((z + H) * E + (y + H)) * E + (x + H);
    else return (y + H) * E + (x + H);
  }
  __host__ __device__ static __forceinline__ int lin(const int* c) {
    if constexpr (ND == 3) return lin(c[0], c[1], c[2]);
    else return lin(c[0], c[1], 0);
  }
  template <int D>
  __device__ static __forceinline__ int stride() {
    if constexpr (D == 0) return 1;
    else if constexpr (D == 1) return E;
    else return E * E;
  }
};

// Halo positions a brick reads, built at compile time: the face halos
// (depth H per face) first, then, for the viscous flux, the edge cells at
// depth 1.  Entry bits 0-11: LDS position; bits 12-26: per-axis offset code
// (3 bits each; 0 inside the brick, 1: -1, 2: -2, 3: +B, 4: +B+1); bits
// 27-31 unused; the Morton bits of the inside coordinates are stored apart.
constexpr unsigned spread_bits(int v, int nd, int axis) {
  unsigned m = 0;
  for (int b = 0; b < 16; ++b)
    if ((v >> b) & 1) m |= 1u << (nd * b + axis);
  return m;
}

template <int ND, int PHYS, bool MUSCL>
struct HaloList {
  using G = Geo<ND, PHYS, MUSCL>;
  static constexpr int classify(int i, int* hc) {
    const int E = G::E, H = G::H, B = G::B;
    hc[0] = i % E - H;
    hc[1] = (ND == 3 ? (i / E) % E : i / E) - H;
    hc[2] = (ND == 3) ? i / (E * E) - H : 0;
    int nout = 0, depth = 0;
    for (int d = 0; d < ND; ++d) {
      const int o = hc[d] < 0 ? -hc[d] : (hc[d] >= B ? hc[d] - B + 1 : 0);
      nout += o > 0 ? 1 : 0;
      depth = depth > o ? depth : o;
    }
    if (nout == 1) return 1;                       // face halo
    if (G::VISC && nout == 2 && depth == 1) return 2;  // edge halo
    return 0;
  }
  static constexpr int count(int kind) {
    int n = 0;
    for (int i = 0; i < G::NPOS; ++i) {
      int hc[3] = {0, 0, 0};
      n += classify(i, hc) == kind ? 1 : 0;
    }
    return n;
  }
  static constexpr int NF = count(1), NE = count(2), N = NF + NE;
  unsigned pos[N];
  unsigned local[N];
  constexpr HaloList() : pos(), local() {
    int k = 0;
    for (int kind = 1; kind <= 2; ++kind)
      for (int i = 0; i < G::NPOS; ++i) {
        int hc[3] = {0, 0, 0};
        if (classify(i, hc) != kind) continue;
        unsigned e = (unsigned)i, l = 0;
        for (int d = 0; d < ND; ++d) {
          const int B = G::B;
          unsigned code = 0;
          if (hc[d] == -1) code = 1;
          else if (hc[d] == -2) code = 2;
          else if (hc[d] == B) code = 3;
          else if (hc[d] == B + 1) code = 4;
          else l |= spread_bits(hc[d], ND, d);
          e |= code << (12 + 3 * d);
        }
        pos[k] = e;
        local[k] = l;
        ++k;
      }
  }
};

template <int ND, int PHYS, bool MUSCL>
__constant__ HaloList<ND, PHYS, MUSCL> kHalo = HaloList<ND, PHYS, MUSCL>();

// Halo position j (halo depth 1) in closed form, the same set as HaloList in
// another order: the face halos face by face (axis d, low then high side,
// the two other axes in increasing order; j < 2 nd B^(nd-1)), then the edge
// cells of the viscous slopes (axis pairs (0,1), (0,2), (1,2), sides
// low/high, along the free axis; in 2D the four corner cells).  hc: the
// brick coordinates, -1 or B on the halo axes.
template <int ND, int PHYS, bool MUSCL>
__host__ __device__ __forceinline__ void halo_closed_form(int j, int* hc) {
  using G = Geo<ND, PHYS, MUSCL>;
  constexpr int B = G::B;
  constexpr int NFH = 2 * ND * G::TANG;
  if (j < NFH) {
    const int f = j / G::TANG, k = j % G::TANG;
    const int d = f >> 1;
    hc[d] = (f & 1) ? B : -1;
    const int a1 = (d == 0) ? 1 : 0;
    const int a2 = (d == 2) ? 1 : 2;
    hc[a1] = k % B;
    if constexpr (ND == 3) hc[a2] = k / B;
  } else {
    constexpr int L = (ND == 3) ? B : 1;
    const int e = (j - NFH) / L, k = (j - NFH) % L;
    const int pr = e >> 2, sides = e & 3;
    const int d1 = (pr == 2) ? 1 : 0, d2 = (pr == 0) ? 1 : 2;
    hc[d1] = (sides & 1) ? B : -1;
    hc[d2 < ND ? d2 : 1] = (sides & 2) ? B : -1;
    if constexpr (ND == 3) hc[3 - d1 - d2] = k;
  }
}

}  // namespace
}  // namespace hom3g
