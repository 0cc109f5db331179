// Host-side check of the brick kernel's halo enumeration: the closed-form
// halo positions (halo_closed_form, used by the depth-1 load phase) are the
// same set as the compile-time halo list, with the face halos first.
#include "../../hom3_amd/csrc/brick.hip"
#include <cstdio>
#include <set>
#include <tuple>
using namespace hom3g;
template <int ND, int PHYS, bool MUSCL>
int check(const char* name) {
  using G = Geo<ND, PHYS, MUSCL>;
  constexpr HaloList<ND, PHYS, MUSCL> hl{};
  std::set<int> a_face, a_edge, b_face, b_edge;
  for (int j = 0; j < hl.N; ++j) {
    const int pos = (int)(hl.pos[j] & 0xfffu);
    (j < hl.NF ? a_face : a_edge).insert(pos);
    int hc[3] = {0, 0, 0};
    halo_closed_form<ND, PHYS, MUSCL>(j, hc);
    for (int d = 0; d < ND; ++d)
      if (hc[d] < -1 || hc[d] > G::B) { printf("%s: j=%d out of range\n", name, j); return 1; }
    (j < hl.NF ? b_face : b_edge).insert(G::lin(hc));
  }
  const bool ok = a_face == b_face && a_edge == b_edge && (int)b_face.size() == hl.NF &&
                  (int)b_edge.size() == hl.N - hl.NF;
  printf("%s: N=%d NF=%d %s\n", name, hl.N, hl.NF, ok ? "OK" : "MISMATCH");
  return ok ? 0 : 1;
}
int main() {
  int bad = 0;
  bad += check<3, HOM3G_CNS, false>("3D CNS");
  bad += check<3, HOM3G_EULER, false>("3D Euler");
  bad += check<2, HOM3G_CNS, false>("2D CNS");
  bad += check<2, HOM3G_EULER, false>("2D Euler");
  return bad;
}
