// Lattice geometry and structure-of-arrays field views in HBM.
//
// Sites are ordered lexicographically, s = x0 + L0 (x1 + L1 (x2 + L2 x3)),
// the order of qlat's coordinate_from_index (used throughout the reference,
// e.g. field-hmc.h:349). Each field is structure-of-arrays in groups of W
// reals: the NC reals of a link are split into NC / W groups of W
// consecutive components (16 bytes: W = 4 in FP32, 2 in FP64; 8 bytes,
// W = 2, for 18-real FP32 fields, whose 18 is not a multiple of 4), and
// group g of direction mu holds those components for every site in turn:
//   real c of direction mu at site s  ->  p[((mu NC/W + c/W) V + s) W + c%W]
// A wavefront of 64 consecutive sites reads one group with one 16-byte load
// per lane, 1 KB contiguous: 3 loads per FP32 12-real link and 9 per FP64
// link instead of 12 and 18 single-component loads (the vector-memory data
// path, not the bytes, limits the stencil kernels). W = 1 (MAG_PACK = 0) is
// the plain component-major layout.
//   link fields   NC = 18 (full 3x3) or 12 (rows 0,1; row 2 = conj(r0 x r1))
//   algebra fields (momenta, forces, stout Q, Lambda)  NC = 8, H = sum h_a T_a
//   general matrix fields (stout C, Y, W, Z)          NC = 18
#pragma once
#include <type_traits>

#include "su3.hpp"

#define MAG_D __device__ __forceinline__

namespace mag {

#ifndef MAG_PACK
#define MAG_PACK 1
#endif

// Reals per group of a field with NC reals per link stored as TS.
template <class TS, int NC>
constexpr int pack_width() {
  if (!MAG_PACK) return 1;
  constexpr int w = 16 / (int)sizeof(TS);
  return NC % w == 0 ? w : (NC % 2 == 0 ? 2 : 1);
}
// Host-side width for a runtime real size and NC (same rule).
inline int pack_width_rt(int real_bytes, int nc) {
  if (!MAG_PACK) return 1;
  const int w = 16 / real_bytes;
  return nc % w == 0 ? w : (nc % 2 == 0 ? 2 : 1);
}
// Index of real c of direction mu at site s (W from pack_width).
MAG_HD size_t elem_index(int mu, int c, int s, int nc, int w, int V) {
  return ((size_t)(mu * (nc / w) + c / w) * (size_t)V + (size_t)s) * (size_t)w + (size_t)(c % w);
}

struct Geom {
  int L[4];
  int S[4];  // strides 1, L0, L0 L1, L0 L1 L2
  int V;
};

inline Geom make_geom(const int* L) {
  Geom g;
  int s = 1;
  for (int i = 0; i < 4; ++i) {
    g.L[i] = L[i];
    g.S[i] = s;
    s *= L[i];
  }
  g.V = s;
  return g;
}

struct Site {
  int s;
  int x[4];
};

MAG_HD Site site_of(const Geom& g, int s) {
  Site p;
  p.s = s;
  int r = s;
  p.x[0] = r % g.L[0]; r /= g.L[0];
  p.x[1] = r % g.L[1]; r /= g.L[1];
  p.x[2] = r % g.L[2];
  p.x[3] = r / g.L[2];
  return p;
}

// Index offset of a shift by d (|d| <= 2) along direction mu, with periodic wrap.
// Select a[i] from a 4-array without dynamic register indexing.
MAG_HD int sel4(const int* a, int i) { return i == 0 ? a[0] : (i == 1 ? a[1] : (i == 2 ? a[2] : a[3])); }

MAG_HD int off_rt(const Geom& g, const Site& p, int mu, int d) {
  const int L = sel4(g.L, mu), x = sel4(p.x, mu);
  int y = x + d;
  y = (y >= L) ? y - L : ((y < 0) ? y + L : y);
  return (y - x) * sel4(g.S, mu);
}

// Constraint class of link (x, mu) for block size mag (field-hmc.h:309-337):
// 0 free, 10 first, 1 middle, 100 last link of a coarse link.
MAG_HD int link_class_rt(const Site& p, int mu, int mag) {
  if (!mag) return 0;
  bool c = true;
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (i != mu) c = c && (p.x[i] % mag == 0);
  if (!c) return 0;
  const int r = sel4(p.x, mu) % mag;
  if (r == mag - 1) return 100;
  if (r == 0) return 10;
  return 1;
}

// ------------------------------------------------------------------ views
// Field pointers always address device global memory; the explicit address
// space keeps loads global (not flat) also when a view comes from a table in
// memory (batched launches).
template <class T>
using gp = __attribute__((address_space(1))) T*;
template <class T>
MAG_D gp<T> gptr(T* p) { return (gp<T>)p; }
template <class T>
MAG_D gp<const T> gptr(const T* p) { return (gp<const T>)p; }

// A group's base pointer p + (mu NC/W + g) V W is wavefront-uniform whenever
// mu is, and the site enters as a 32-bit unsigned offset, so the load/store
// takes the scalar-base + vector-offset form with one offset register shared
// by all groups of a site (no per-group 64-bit address arithmetic on the
// vector ALUs).
// Element at a 32-bit byte offset from a group base (groups hold < 2^28 sites).
template <class T>
using gref = __attribute__((address_space(1))) T&;
template <class T>
MAG_D gref<T> at(gp<T> base, unsigned s) {
  using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return *(gp<T>)((gp<C>)base + s * (unsigned)sizeof(T));
}

// One group of W reals as a vector (one load or store instruction per lane).
template <class TS, int W>
struct PackVec {
  typedef TS type __attribute__((ext_vector_type(W)));
};
template <class TS>
struct PackVec<TS, 1> {
  typedef TS type;
};

// Groups of NC reals per link in W-wide vectors; loads and stores of the
// reals [0, n) of direction mu at site s, through one vector access per
// group (NT: non-temporal).
template <class TS, int NC>
struct Packed {
  static constexpr int W = pack_width<TS, NC>();
  static constexpr int NG = NC / W;
  using VT = typename PackVec<TS, W>::type;
  MAG_D static gp<VT> group(TS* p, int V, int mu, int g) {
    return (gp<VT>)gptr(p + (size_t)(mu * NG + g) * (size_t)V * W);
  }
  template <int N, bool NT>
  MAG_D static void get(TS* p, int V, int mu, int s, TS r[N]) {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int g = 0; g < (N + W - 1) / W; ++g) {
      VT v;
      if constexpr (NT) v = __builtin_nontemporal_load(&at(group(p, V, mu, g), u));
      else v = at(group(p, V, mu, g), u);
      if constexpr (W == 1) {
        r[g] = v;
      } else {
#pragma unroll
        for (int k = 0; k < W; ++k)
          if (g * W + k < N) r[g * W + k] = v[k];
      }
    }
  }
  template <bool NT>
  MAG_D static void put(TS* p, int V, int mu, int s, const TS r[NC]) {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
      VT v;
      if constexpr (W == 1) {
        v = r[g];
      } else {
#pragma unroll
        for (int k = 0; k < W; ++k) v[k] = r[g * W + k];
      }
      if constexpr (NT) __builtin_nontemporal_store(v, &at(group(p, V, mu, g), u));
      else at(group(p, V, mu, g), u) = v;
    }
  }
};

template <class TS, int NC>
struct LinkView {
  TS* p;
  int V;
  using PK = Packed<TS, NC>;
  template <class TC, bool NT>
  MAG_D M3<TC> get(int mu, int s) const {
    TS r[NC];
    PK::template get<NC, NT>(p, V, mu, s, r);
    M3<TC> m;
#pragma unroll
    for (int k = 0; k < NC / 2; ++k) m.e[k] = {TC(r[2 * k]), TC(r[2 * k + 1])};
    if (NC == 12) reconstruct_row2(m);
    return m;
  }
  template <class TC, bool NT>
  MAG_D void put(int mu, int s, const M3<TC>& m) const {
    TS r[NC];
#pragma unroll
    for (int k = 0; k < NC / 2; ++k) {
      r[2 * k] = TS(m.e[k].r);
      r[2 * k + 1] = TS(m.e[k].i);
    }
    PK::template put<NT>(p, V, mu, s, r);
  }
  template <class TC>
  MAG_D M3<TC> load(int mu, int s) const { return get<TC, false>(mu, s); }
  // rows 0 and 1 only (row 2 zero), for a factor read only there
  template <class TC, bool NT = false>
  MAG_D M3<TC> load_r2(int mu, int s) const {
    TS r[12];
    PK::template get<12, NT>(p, V, mu, s, r);
    M3<TC> m;
#pragma unroll
    for (int k = 0; k < 6; ++k) m.e[k] = {TC(r[2 * k]), TC(r[2 * k + 1])};
#pragma unroll
    for (int k = 6; k < 9; ++k) m.e[k] = {TC(0), TC(0)};
    return m;
  }
  template <class TC>
  MAG_D M3<TC> load_r2_nt(int mu, int s) const { return load_r2<TC, true>(mu, s); }
  // A link known to be in SU(3): with R12 an 18-real link is read as rows 0
  // and 1 (two thirds of its groups) and row 2 rebuilt as conj(r0 x r1).
  template <class TC, bool R12>
  MAG_D M3<TC> load_su3(int mu, int s) const {
    if constexpr (R12 && NC == 18) {
      M3<TC> m = load_r2<TC>(mu, s);
      reconstruct_row2(m);
      return m;
    } else {
      return load<TC>(mu, s);
    }
  }
  template <class TC, bool R12>
  MAG_D M3<TC> load_su3_nt(int mu, int s) const {
    if constexpr (R12 && NC == 18) {
      M3<TC> m = load_r2_nt<TC>(mu, s);
      reconstruct_row2(m);
      return m;
    } else {
      return get<TC, true>(mu, s);
    }
  }
  template <class TC>
  MAG_D void store(int mu, int s, const M3<TC>& m) const { put<TC, false>(mu, s, m); }
  // Non-temporal forms for data a kernel touches once, at its own link
  // (read or written by this thread only): they do not displace the
  // neighbour links that the stencil reuses from L2.
  template <class TC>
  MAG_D M3<TC> load_nt(int mu, int s) const { return get<TC, true>(mu, s); }
  template <class TC>
  MAG_D void store_nt(int mu, int s, const M3<TC>& m) const { put<TC, true>(mu, s, m); }
};

template <class TS>
struct AlgView {
  TS* p;
  int V;
  using PK = Packed<TS, 8>;
  template <class TC, bool NT>
  MAG_D void get(int mu, int s, TC h[8]) const {
    TS r[8];
    PK::template get<8, NT>(p, V, mu, s, r);
#pragma unroll
    for (int a = 0; a < 8; ++a) h[a] = TC(r[a]);
  }
  template <class TC, bool NT>
  MAG_D void put(int mu, int s, const TC h[8]) const {
    TS r[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) r[a] = TS(h[a]);
    PK::template put<NT>(p, V, mu, s, r);
  }
  template <class TC>
  MAG_D void load(int mu, int s, TC h[8]) const { get<TC, false>(mu, s, h); }
  template <class TC>
  MAG_D void store(int mu, int s, const TC h[8]) const { put<TC, false>(mu, s, h); }
  template <class TC>
  MAG_D M3<TC> load_mat(int mu, int s) const {
    TC h[8];
    load(mu, s, h);
    return alg_to_mat(h);
  }
  template <class TC>
  MAG_D void load_nt(int mu, int s, TC h[8]) const { get<TC, true>(mu, s, h); }
  template <class TC>
  MAG_D void store_nt(int mu, int s, const TC h[8]) const { put<TC, true>(mu, s, h); }
  template <class TC>
  MAG_D M3<TC> load_mat_nt(int mu, int s) const {
    TC h[8];
    load_nt(mu, s, h);
    return alg_to_mat(h);
  }
};

template <class TS>
using MatView = LinkView<TS, 18>;

}  // namespace mag
