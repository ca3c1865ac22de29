// Host-side check of the stencil kernels' tile map (mag_amd/csrc/kernels.hpp,
// tile_map): over the grid a stencil launch uses (4 tiles per 256-site block),
// every (direction, site block) must come out of exactly one block id, in
// both orders (X2_INNER false and true). Where the XCD slab walk applies, the
// block ids of XCD k (id mod 8 = k) must all fall in x3 in
// [k L3/8, (k+1) L3/8) and each tile's 256 sites must share x3. Prints one
// line per shape (the walk taken) and exits non-zero on the first violation.
#include <cstdio>
#include <vector>

#include "../../mag_amd/csrc/kernels.hpp"

using namespace mag;

template <bool X2>
static int check(const int L[4]) {
  const Geom g = make_geom(L);
  const int nbs = (g.V + BLOCK - 1) / BLOCK;
  const int nb = 4 * nbs;
  std::vector<int> hit((size_t)nb, 0);
  const int plane = L[0] * L[1], slice = plane * L[2];
  const bool slab =
      nb % 8 == 0 && L[3] % 8 == 0 && g.V % BLOCK == 0 && (X2 ? plane % BLOCK == 0 : slice % BLOCK == 0);
  for (int b = 0; b < nb; ++b) {
    int mu = -1, sblk = -1;
    tile_map<X2>(g, nb, b, mu, sblk);
    if (mu < 0 || mu > 3 || sblk < 0 || sblk >= nbs) {
      std::printf("FAIL %dx%dx%dx%d X2=%d: block %d -> (mu %d, sblk %d) out of range\n", L[0], L[1], L[2], L[3],
                  (int)X2, b, mu, sblk);
      return 1;
    }
    ++hit[(size_t)sblk * 4 + mu];
    if (slab) {
      const int h3 = L[3] / 8, xcd = b % 8;
      const int first = sblk * BLOCK, last = first + BLOCK - 1;
      const int x3a = first / slice, x3b = last / slice;
      if (x3a != x3b || x3a < xcd * h3 || x3a >= (xcd + 1) * h3) {
        std::printf("FAIL %dx%dx%dx%d X2=%d: block %d (XCD %d) covers x3 %d..%d\n", L[0], L[1], L[2], L[3],
                    (int)X2, b, xcd, x3a, x3b);
        return 1;
      }
    }
  }
  for (int t = 0; t < nb; ++t)
    if (hit[(size_t)t] != 1) {
      std::printf("FAIL %dx%dx%dx%d X2=%d: tile (sblk %d, mu %d) hit %d times\n", L[0], L[1], L[2], L[3], (int)X2,
                  t / 4, t % 4, hit[(size_t)t]);
      return 1;
    }
  std::printf("ok %dx%dx%dx%d X2=%d %s\n", L[0], L[1], L[2], L[3], (int)X2, slab ? "xcd-slab" : "lexicographic");
  return 0;
}

int main() {
  const int shapes[][4] = {{4, 4, 4, 4},   {6, 6, 6, 6},    {8, 8, 8, 8},    {12, 12, 12, 12}, {16, 16, 16, 16},
                           {24, 24, 24, 24}, {32, 32, 32, 32}, {16, 16, 4, 8},  {4, 6, 8, 10},    {4, 8, 4, 16},
                           {8, 4, 8, 8},   {24, 24, 24, 16}, {64, 64, 64, 8}, {4, 4, 4, 8},     {12, 12, 16, 24},
                           {8, 8, 4, 16}};
  int bad = 0;
  for (const auto& L : shapes) bad |= check<false>(L) | check<true>(L);
  return bad;
}
