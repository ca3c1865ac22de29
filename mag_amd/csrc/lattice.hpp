// Lattice geometry and structure-of-arrays field views in HBM.
//
// Sites are ordered lexicographically, s = x0 + L0 (x1 + L1 (x2 + L2 x3)),
// the order of qlat's coordinate_from_index (used throughout the reference,
// e.g. field-hmc.h:349). Each field is component-major per direction:
//   real c of direction mu at site s  ->  p[(mu * NC + c) * V + s]
// so a wavefront of 64 consecutive sites reads 64 consecutive reals of one
// component (256 B in FP32): every load is fully coalesced.
//   link fields   NC = 18 (full 3x3) or 12 (rows 0,1; row 2 = conj(r0 x r1))
//   algebra fields (momenta, forces, stout Q, Lambda)  NC = 8, H = sum h_a T_a
//   general matrix fields (stout C, Y, W, Z)          NC = 18
#pragma once
#include <type_traits>

#include "su3.hpp"

#define MAG_D __device__ __forceinline__

namespace mag {

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

// Component c of direction mu at site s. The plane pointer p + (mu NC + c) V
// is wavefront-uniform whenever mu is, and the site enters as a 32-bit
// unsigned offset, so the load/store takes the scalar-base + vector-offset
// form with one offset register shared by all components of a site (no
// per-component 64-bit address arithmetic on the vector ALUs).
template <class T>
MAG_D gp<T> plane(T* p, size_t plane_index, int V) { return gptr(p + plane_index * (size_t)V); }
// Element at a 32-bit byte offset from a plane (planes hold < 2^30 reals).
template <class T>
using gref = __attribute__((address_space(1))) T&;
template <class T>
MAG_D gref<T> at(gp<T> base, unsigned s) {
  using C = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return *(gp<T>)((gp<C>)base + s * (unsigned)sizeof(T));
}

template <class TS, int NC>
struct LinkView {
  TS* p;
  int V;
  template <class TC>
  MAG_D M3<TC> load(int mu, int s) const {
    const unsigned u = (unsigned)s;
    M3<TC> m;
#pragma unroll
    for (int k = 0; k < NC / 2; ++k) {
      m.e[k].r = TC(at(plane(p, (size_t)(mu * NC + 2 * k), V), u));
      m.e[k].i = TC(at(plane(p, (size_t)(mu * NC + 2 * k + 1), V), u));
    }
    if (NC == 12) reconstruct_row2(m);
    return m;
  }
  // rows 0 and 1 only (row 2 zero), for a factor read only there
  template <class TC>
  MAG_D M3<TC> load_r2(int mu, int s) const {
    const unsigned u = (unsigned)s;
    M3<TC> m;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      m.e[k].r = TC(at(plane(p, (size_t)(mu * NC + 2 * k), V), u));
      m.e[k].i = TC(at(plane(p, (size_t)(mu * NC + 2 * k + 1), V), u));
    }
#pragma unroll
    for (int k = 6; k < 9; ++k) m.e[k] = {TC(0), TC(0)};
    return m;
  }
  template <class TC>
  MAG_D void store(int mu, int s, const M3<TC>& m) const {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int k = 0; k < NC / 2; ++k) {
      at(plane(p, (size_t)(mu * NC + 2 * k), V), u) = TS(m.e[k].r);
      at(plane(p, (size_t)(mu * NC + 2 * k + 1), V), u) = TS(m.e[k].i);
    }
  }
  // Non-temporal forms for data a kernel touches once, at its own link
  // (read or written by this thread only): they do not displace the
  // neighbour links that the stencil reuses from L2.
  template <class TC>
  MAG_D M3<TC> load_nt(int mu, int s) const {
    const unsigned u = (unsigned)s;
    M3<TC> m;
#pragma unroll
    for (int k = 0; k < NC / 2; ++k) {
      m.e[k].r = TC(__builtin_nontemporal_load(&at(plane(p, (size_t)(mu * NC + 2 * k), V), u)));
      m.e[k].i = TC(__builtin_nontemporal_load(&at(plane(p, (size_t)(mu * NC + 2 * k + 1), V), u)));
    }
    if (NC == 12) reconstruct_row2(m);
    return m;
  }
  template <class TC>
  MAG_D void store_nt(int mu, int s, const M3<TC>& m) const {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int k = 0; k < NC / 2; ++k) {
      __builtin_nontemporal_store(TS(m.e[k].r), &at(plane(p, (size_t)(mu * NC + 2 * k), V), u));
      __builtin_nontemporal_store(TS(m.e[k].i), &at(plane(p, (size_t)(mu * NC + 2 * k + 1), V), u));
    }
  }
};

template <class TS>
struct AlgView {
  TS* p;
  int V;
  template <class TC>
  MAG_D void load(int mu, int s, TC h[8]) const {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int a = 0; a < 8; ++a) h[a] = TC(at(plane(p, (size_t)(mu * 8 + a), V), u));
  }
  template <class TC>
  MAG_D void store(int mu, int s, const TC h[8]) const {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int a = 0; a < 8; ++a) at(plane(p, (size_t)(mu * 8 + a), V), u) = TS(h[a]);
  }
  template <class TC>
  MAG_D M3<TC> load_mat(int mu, int s) const {
    TC h[8];
    load(mu, s, h);
    return alg_to_mat(h);
  }
  template <class TC>
  MAG_D void load_nt(int mu, int s, TC h[8]) const {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int a = 0; a < 8; ++a) h[a] = TC(__builtin_nontemporal_load(&at(plane(p, (size_t)(mu * 8 + a), V), u)));
  }
  template <class TC>
  MAG_D void store_nt(int mu, int s, const TC h[8]) const {
    const unsigned u = (unsigned)s;
#pragma unroll
    for (int a = 0; a < 8; ++a) __builtin_nontemporal_store(TS(h[a]), &at(plane(p, (size_t)(mu * 8 + a), V), u));
  }
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
