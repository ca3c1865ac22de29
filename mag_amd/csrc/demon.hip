// demon_microcanonical (field-demon.h:25-126) on the device, in the
// reference's sequential order: links in site-index order, mu inner, two hits
// per link, each accepted hit changing the link before the next link's
// staples are formed. The chain is inherently serial (one demon energy, one
// RNG stream), so one wavefront runs it: the 64 lanes split the 168 staple
// paths of a link (6 plaquette, 18 rectangle, 72 chair, 72 twist; index
// tables of field-staple.h:27-135 with their defects kept), the four sums are
// reduced in a fixed lane order through LDS, and lane 0 makes the proposals
// and the window test. One launch per sweep.
#include <cstring>
#include <vector>

#include "kernels.hpp"
#include "runtime.hpp"

using namespace mag;

namespace {

constexpr int NPATH = 168;  // per direction
constexpr int PLEN = 5;

struct DemonTable {
  // dirs[mu][path][k]: 0..3 forward, 4..7 backward, -1 = end of path
  signed char dirs[4][NPATH][PLEN];
  unsigned char kind[NPATH];  // 0 plaquette, 1 rectangle, 2 chair, 3 twist
  // paths [kbeg[k], kbeg[k+1]) are of kind k (the same for every direction),
  // plaquettes of 3 links and the other kinds of 5
  short kbeg[5];
};

// Paths of get_staple_dagger (field-hmc.h:226-240), get_rectangular_dagger
// (field-hmc.h:137-201), init_chair_index / init_twist_index
// (field-staple.h:27-135).
void build_table(DemonTable& t) {
  std::memset(t.dirs, -1, sizeof t.dirs);
  const int D = 4;
  for (int mu = 0; mu < 4; ++mu) {
    int n = 0;
    auto put = [&](std::initializer_list<int> d, int kind) {
      int k = 0;
      for (int v : d) t.dirs[mu][n][k++] = (signed char)v;
      t.kind[n] = (unsigned char)kind;
      ++n;
    };
    for (int nu = 0; nu < D; ++nu) {
      if (nu == mu) continue;
      put({nu, mu, nu + D}, 0);
      put({nu + D, mu, nu}, 0);
    }
    for (int nu = 0; nu < D; ++nu) {
      if (nu == mu) continue;
      put({nu, mu, mu, nu + D, mu + D}, 1);
      put({nu + D, mu, mu, nu, mu + D}, 1);
      put({nu, nu, mu, nu + D, nu + D}, 1);
      put({nu + D, nu + D, mu, nu, nu}, 1);
      put({mu + D, nu, mu, mu, nu + D}, 1);
      put({mu + D, nu + D, mu, mu, nu}, 1);
    }
    for (int twist = 0; twist < 2; ++twist)
      for (int nu = 0; nu < D; ++nu) {
        if (nu == mu) continue;
        for (int la = 0; la < D; ++la) {
          if (la == nu || la == mu) continue;
          int b[4][4];
          if (!twist) {
            int c[4][4] = {{nu, la, la + D, nu + D}, {nu, la + D, la + D, nu},
                           {nu + D, la, la, nu + D}, {nu + D, la + D, la, nu}};
            std::memcpy(b, c, sizeof b);
          } else {
            int c[4][4] = {{nu, la, nu + D, la + D}, {nu, la + D, nu + D, la},
                           {nu + D, la, nu, la + D}, {nu + D, la + D, nu, la}};
            std::memcpy(b, c, sizeof b);
          }
          for (int k = 0; k < 4; ++k)
            for (int i = 1; i < 4; ++i) {
              int d[5], m = 0;
              for (int j = 0; j < 4; ++j) {
                if (j == i) d[m++] = mu;
                d[m++] = b[k][j];
              }
              t.dirs[mu][n][0] = (signed char)d[0];
              for (int j = 1; j < 5; ++j) t.dirs[mu][n][j] = (signed char)d[j];
              t.kind[n] = (unsigned char)(2 + twist);
              ++n;
            }
        }
      }
  }
  int k = 0;
  for (int q = 0; q < NPATH; ++q)
    while (k <= t.kind[q]) t.kbeg[k++] = (short)q;
  while (k <= 4) t.kbeg[k++] = (short)NPATH;
}

// The checkerboard kernel's assumptions about the table: kinds in order,
// plaquettes of 3 links, the rest of 5.
bool table_ok(const DemonTable& t) {
  for (int q = 0; q < NPATH; ++q) {
    const int kd = t.kind[q];
    if (q < t.kbeg[kd] || q >= t.kbeg[kd + 1]) return false;
    for (int mu = 0; mu < 4; ++mu) {
      int len = 0;
      while (len < PLEN && t.dirs[mu][q][len] >= 0) ++len;
      if (len != (kd == 0 ? 3 : 5)) return false;
    }
  }
  return true;
}

struct DemonState {
  double E[4];
  long long acc, rej;
};

__global__ __launch_bounds__(64) void k_demon_sweep(LinkView<double, 18> U, Geom g, const DemonTable* tab,
                                                    DemonState* st, double emax, uint64_t seed, int sweep) {
  __shared__ double part[72][64];
  __shared__ double tot[72];
  const int lane = threadIdx.x;
  double E[4] = {st->E[0], st->E[1], st->E[2], st->E[3]};
  long long acc = st->acc, rej = st->rej;
  for (int s = 0; s < g.V; ++s) {
    const Site p0 = site_of(g, s);
    for (int mu = 0; mu < 4; ++mu) {
      M3<double> sum[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) sum[k] = zero3<double>();
      for (int q = lane; q < NPATH; q += 64) {
        Site y = p0;
        M3<double> m = unit3<double>();
        for (int j = 0; j < PLEN; ++j) {
          const int d = tab->dirs[mu][q][j];
          if (d < 0) break;
          if (d < 4) {
            m = mul(m, U.template load<double>(d, y.s));
            y.s += off_rt(g, y, d, 1);
            y.x[0] = y.s % g.L[0];
            y.x[1] = (y.s / g.S[1]) % g.L[1];
            y.x[2] = (y.s / g.S[2]) % g.L[2];
            y.x[3] = y.s / g.S[3];
          } else {
            y.s += off_rt(g, y, d - 4, -1);
            y.x[0] = y.s % g.L[0];
            y.x[1] = (y.s / g.S[1]) % g.L[1];
            y.x[2] = (y.s / g.S[2]) % g.L[2];
            y.x[3] = y.s / g.S[3];
            m = mul_nd(m, U.template load<double>(d - 4, y.s));
          }
        }
        const int kd = tab->kind[q];
#pragma unroll
        for (int k = 0; k < 4; ++k)
          if (k == kd) add_to(sum[k], m);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int e = 0; e < 9; ++e) {
          part[k * 18 + 2 * e][lane] = sum[k].e[e].r;
          part[k * 18 + 2 * e + 1][lane] = sum[k].e[e].i;
        }
      __syncthreads();
      for (int comp = lane; comp < 72; comp += 64) {
        double a = 0.;
        for (int l = 0; l < 64; ++l) a += part[comp][l];
        tot[comp] = a;
      }
      __syncthreads();
      if (lane == 0) {
        M3<double> X[4];
        for (int k = 0; k < 4; ++k)
          for (int e = 0; e < 9; ++e) X[k].e[e] = {tot[k * 18 + 2 * e], tot[k * 18 + 2 * e + 1]};
        // plaquette and rectangle sums are used daggered (get_staple_dagger,
        // get_rectangular_dagger); chair and twist undaggered (field-staple.h:149,164)
        X[0] = dag(X[0]);
        X[1] = dag(X[1]);
        M3<double> v = U.template load<double>(mu, s);
        const uint32_t link = (uint32_t)s * 4u + (uint32_t)mu;
        for (int n = 0; n < 2; ++n) {
          const u32x4 w0 = rng_draw(seed, RNG_DEMON, (uint64_t)sweep, link, 3 * n);
          const u32x4 w1 = rng_draw(seed, RNG_DEMON, (uint64_t)sweep, link, 3 * n + 1);
          const u32x4 w2 = rng_draw(seed, RNG_DEMON, (uint64_t)sweep, link, 3 * n + 2);
          // random_su3_from_su2(0.9) (field-matrix.h:36-65)
          const double a0 = 0.9 + (1. - 0.9) * rng_u01(w0.x, w0.y);
          double a1 = -1. + 2. * rng_u01(w0.z, w0.w);
          double a2 = -1. + 2. * rng_u01(w1.x, w1.y);
          double a3 = -1. + 2. * rng_u01(w1.z, w1.w);
          const double nrm = sqrt((1. - a0 * a0) / (a1 * a1 + a2 * a2 + a3 * a3));
          a1 *= nrm;
          a2 *= nrm;
          a3 *= nrm;
          M3<double> m = unit3<double>();
          switch (w2.x % 3) {
            case 0: m(1, 1) = {a0, a3}; m(1, 2) = {a2, a1}; m(2, 1) = {-a2, a1}; m(2, 2) = {a0, -a3}; break;
            case 1: m(0, 0) = {a0, a3}; m(0, 2) = {a2, a1}; m(2, 0) = {-a2, a1}; m(2, 2) = {a0, -a3}; break;
            default: m(0, 0) = {a0, a3}; m(0, 1) = {a2, a1}; m(1, 0) = {-a2, a1}; m(1, 1) = {a0, -a3}; break;
          }
          const M3<double> mv = mul(m, v);
          double dS[4];
          bool ok = true;
          for (int k = 0; k < 4; ++k) {
            dS[k] = -retrace_mul(mv, X[k]) + retrace_mul(v, X[k]);
            ok = ok && (E[k] - dS[k] > -emax) && (E[k] - dS[k] < emax);
          }
          if (ok) {
            v = mv;
            for (int k = 0; k < 4; ++k) E[k] -= dS[k];
            ++acc;
          } else {
            ++rej;
          }
        }
        U.store(mu, s, v);
      }
      // make lane 0's link store visible to every lane's next loads
      __threadfence();
      __syncthreads();
    }
  }
  if (lane == 0) {
    for (int k = 0; k < 4; ++k) st->E[k] = E[k];
    st->acc = acc;
    st->rej = rej;
  }
}

// ------------------------------------------------------------ chair / twist staples
// chair_staple_dagger / twist_staple_dagger (field-staple.h:141-165) for every
// link: S(x, mu) = the sum of the 72 chair (kind 2) or twist (kind 3) path
// products of the table, undaggered as the reference returns it.
__global__ __launch_bounds__(BLOCK) void k_staple_table_sum(LinkView<double, 18> U, Geom g, const DemonTable* tab,
                                                            int kind, LinkView<double, 18> S) {
  const int s = blockIdx.x * BLOCK + threadIdx.x;
  const int mu = blockIdx.y;
  if (s >= g.V) return;
  const Site p0 = site_of(g, s);
  M3<double> acc = zero3<double>();
  for (int q = 0; q < NPATH; ++q) {
    if (tab->kind[q] != kind) continue;
    Site y = p0;
    M3<double> m = unit3<double>();
    for (int j = 0; j < PLEN; ++j) {
      const int d = tab->dirs[mu][q][j];
      if (d < 0) break;
      if (d < 4) {
        m = mul(m, U.template load<double>(d, y.s));
        y.s += off_rt(g, y, d, 1);
      } else {
        y.s += off_rt(g, y, d - 4, -1);
      }
      y.x[0] = y.s % g.L[0];
      y.x[1] = (y.s / g.S[1]) % g.L[1];
      y.x[2] = (y.s / g.S[2]) % g.L[2];
      y.x[3] = y.s / g.S[3];
      if (d >= 4) m = mul_nd(m, U.template load<double>(d - 4, y.s));
    }
    add_to(acc, m);
  }
  S.store(mu, s, acc);
}

// ------------------------------------------------------------ checkerboard demon
// demon_microcanonical with a demon per link (four energies per link, four
// MAG_REAL1 fields), updated in parallel: per direction MU, 128 classes
// (x_MU mod 2, then x_nu mod 4 for the transverse nu in increasing order).
// Every staple path of a link holds its MU-links at offsets with |d_nu| <= 2
// and d_MU in {-1, 0, 1}, so two links of one class never meet in a path:
// the updates of a class commute and each moves its own action change into
// its own demons. Eight lanes per class member; hit counts in per-member slots.
// The lanes sum one kind of path at a time (one accumulator; the finished
// sums wait in LDS), each lane the same paths in the same order as a single
// loop over all paths, so the sums are bitwise those of that loop.
struct DemonClassArgs {
  LinkView<double, 18> U;
  Geom g;
  const DemonTable* tab;
  double* E[4];   // per-link demon energies, [mu][site] each
  double* cnt;    // [thread][2]: accept, reject
  double emax;
  uint64_t seed;
  int sweep, a, cls;  // x_MU parity, transverse residues (base 4)
  int nthr;           // class members
};

// LPL lanes per class member, 64 / LPL members per wavefront: lane sub of a
// member sums paths sub, sub + LPL, ..., a butterfly over its LPL lanes
// (fixed order) reduces the four 3x3 sums, and its lane 0 makes the hits.
// Members are enumerated x0 fastest, so the lanes of one sub-index read
// neighbouring sites. Surplus lanes of the last wavefront repeat the last
// member and write nothing. Sweep of a 32^4 lattice on MI355X: 387 ms with
// 64 lanes per member (every lane's gather its own cache lines), 155 ms with
// 16, 142 ms with 8, 193 ms with 4 (too few wavefronts per class). A class
// launch at 32^4 is 1,024 one-wave blocks, one wave per SIMD, so the path
// loads are latency-bound: requesting a path's links together before its
// products (one round trip per path, not one per link) took the sweep from
// 107 to 79 ms; 16 lanes at 2 waves per SIMD gave 77.5 ms, and requesting
// the next path's links during this path's products 93 ms (AGPR copies).
constexpr int DEMON_LPL = 8;
// MAG_DEMON_READ12 = 1: the path links are read as rows 0 and 1 with row 2
// rebuilt (the gathers bound the class kernel; the links stay in SU(3)).
#ifndef MAG_DEMON_READ12
#define MAG_DEMON_READ12 1
#endif
// The paths of one kind, of LEN links each, that lane sub of a member sums
// (q = sub mod LPL, in increasing q, as one loop over all paths would give
// it): the links of a path are all requested before its products (sdir is
// the direction's path table in LDS).
template <int LPL, int LEN>
MAG_D void demon_kind_sum(const DemonClassArgs& A, const signed char* sdir, int qa, int qb, int sub, const int x[4],
                          M3<double>& X) {
  X = zero3<double>();
#pragma unroll 1
  for (int q = qa + ((sub - qa) % LPL + LPL) % LPL; q < qb; q += LPL) {
    int y[4] = {x[0], x[1], x[2], x[3]};
    M3<double> l[LEN];
    bool back[LEN];
#pragma unroll
    for (int j = 0; j < LEN; ++j) {
      const int d = sdir[q * PLEN + j];
      const int k = d & 3;
      back[j] = d >= 4;
      if (back[j]) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (i == k) y[i] = (y[i] == 0) ? A.g.L[i] - 1 : y[i] - 1;
      }
      const int sy = y[0] + A.g.S[1] * y[1] + A.g.S[2] * y[2] + A.g.S[3] * y[3];
      l[j] = MAG_DEMON_READ12 ? A.U.template load_r2<double>(k, sy) : A.U.template load<double>(k, sy);
      if (!back[j]) {
#pragma unroll
        for (int i = 0; i < 4; ++i)
          if (i == k) y[i] = (y[i] == A.g.L[i] - 1) ? 0 : y[i] + 1;
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // keep the path's loads ahead of its products
    M3<double> m;
#pragma unroll
    for (int j = 0; j < LEN; ++j) {
      if (MAG_DEMON_READ12) reconstruct_row2(l[j]);
      if (j == 0) m = back[0] ? dag(l[0]) : l[0];
      else m = back[j] ? mul_nd(m, l[j]) : mul(m, l[j]);
    }
    add_to(X, m);
  }
  // butterfly over the member's lanes: each ends with the full sum
#pragma unroll
  for (int off = LPL / 2; off >= 1; off >>= 1)
#pragma unroll
    for (int e = 0; e < 9; ++e) {
      X.e[e].r += __shfl_xor(X.e[e].r, off, 64);
      X.e[e].i += __shfl_xor(X.e[e].i, off, 64);
    }
}

template <int MU, int LPL>
__global__ __launch_bounds__(64) void k_demon_class(DemonClassArgs A) {
  constexpr int MPW = 64 / LPL;
  __shared__ signed char sdir[NPATH * PLEN];
  __shared__ double sum[MPW][4][18];  // a member's four staple sums, written by its lane 0
  const int lane = threadIdx.x;
  for (int i = lane; i < NPATH * PLEN; i += 64) sdir[i] = (&A.tab->dirs[MU][0][0])[i];
  __syncthreads();
  const int sub = lane % LPL, mem = lane / LPL;
  const int t0 = blockIdx.x * MPW + mem;
  const bool live = t0 < A.nthr;
  const int t = live ? t0 : A.nthr - 1;
  int x[4], r = t, c = A.cls;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int step = (i == MU) ? 2 : 4;
    const int h = r % (A.g.L[i] / step);
    r /= A.g.L[i] / step;
    if (i == MU) {
      x[i] = 2 * h + A.a;
    } else {
      x[i] = 4 * h + (c & 3);
      c >>= 2;
    }
  }
  const int s = x[0] + A.g.S[1] * x[1] + A.g.S[2] * x[2] + A.g.S[3] * x[3];
  // the four staple sums (same paths and kinds as the sequential demon), one
  // kind at a time
#pragma unroll 1
  for (int kd = 0; kd < 4; ++kd) {
    M3<double> X;
    if (kd == 0) demon_kind_sum<LPL, 3>(A, sdir, A.tab->kbeg[0], A.tab->kbeg[1], sub, x, X);
    else demon_kind_sum<LPL, 5>(A, sdir, A.tab->kbeg[kd], A.tab->kbeg[kd + 1], sub, x, X);
    if (sub == 0) {
#pragma unroll
      for (int e = 0; e < 9; ++e) {
        sum[mem][kd][2 * e] = X.e[e].r;
        sum[mem][kd][2 * e + 1] = X.e[e].i;
      }
    }
  }
  if (sub != 0 || !live) return;
  M3<double> X[4];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int e = 0; e < 9; ++e) X[k].e[e] = {sum[mem][k][2 * e], sum[mem][k][2 * e + 1]};
  // plaquette and rectangle sums are used daggered, chair and twist as they are
  X[0] = dag(X[0]);
  X[1] = dag(X[1]);
  double E[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) E[k] = __builtin_nontemporal_load(&A.E[k][(size_t)MU * A.g.V + s]);
  M3<double> v = A.U.template load_nt<double>(MU, s);
  const uint32_t link = (uint32_t)s * 4u + (uint32_t)MU;
  double acc = 0., rej = 0.;
#pragma unroll 1
  for (int n = 0; n < 2; ++n) {
    const u32x4 w0 = rng_draw(A.seed, RNG_DEMON, (uint64_t)A.sweep, link, 3 * n);
    const u32x4 w1 = rng_draw(A.seed, RNG_DEMON, (uint64_t)A.sweep, link, 3 * n + 1);
    const u32x4 w2 = rng_draw(A.seed, RNG_DEMON, (uint64_t)A.sweep, link, 3 * n + 2);
    const double a0 = 0.9 + (1. - 0.9) * rng_u01(w0.x, w0.y);
    double a1 = -1. + 2. * rng_u01(w0.z, w0.w);
    double a2 = -1. + 2. * rng_u01(w1.x, w1.y);
    double a3 = -1. + 2. * rng_u01(w1.z, w1.w);
    const double nrm = sqrt((1. - a0 * a0) / (a1 * a1 + a2 * a2 + a3 * a3));
    a1 *= nrm;
    a2 *= nrm;
    a3 *= nrm;
    M3<double> m = unit3<double>();
    switch (w2.x % 3) {
      case 0: m(1, 1) = {a0, a3}; m(1, 2) = {a2, a1}; m(2, 1) = {-a2, a1}; m(2, 2) = {a0, -a3}; break;
      case 1: m(0, 0) = {a0, a3}; m(0, 2) = {a2, a1}; m(2, 0) = {-a2, a1}; m(2, 2) = {a0, -a3}; break;
      default: m(0, 0) = {a0, a3}; m(0, 1) = {a2, a1}; m(1, 0) = {-a2, a1}; m(1, 1) = {a0, -a3}; break;
    }
    const M3<double> mv = mul(m, v);
    double dS[4];
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      dS[k] = -retrace_mul(mv, X[k]) + retrace_mul(v, X[k]);
      ok = ok && (E[k] - dS[k] > -A.emax) && (E[k] - dS[k] < A.emax);
    }
    if (ok) {
      v = mv;
#pragma unroll
      for (int k = 0; k < 4; ++k) E[k] -= dS[k];
      acc += 1.;
    } else {
      rej += 1.;
    }
  }
  A.U.store_nt(MU, s, v);
#pragma unroll
  for (int k = 0; k < 4; ++k) __builtin_nontemporal_store(E[k], &A.E[k][(size_t)MU * A.g.V + s]);
  A.cnt[(size_t)t * 2] += acc;
  A.cnt[(size_t)t * 2 + 1] += rej;
}

}  // namespace

extern "C" int mag_demon_microcanonical(mag_field U, int nsweeps, double E_max, uint64_t seed, double E[4],
                                        long acc_rej[2]) {
  if (!U || !E || !acc_rej || nsweeps < 0) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  if (U->kind != MAG_LINK18 || U->prec != MAG_F64)
    return fail(c, MAG_ERR_UNSUPPORTED, "the demon runs on FP64 18-real links");
  DemonTable* ht = new DemonTable();
  build_table(*ht);
  DemonTable* dt = nullptr;
  DemonState* ds = nullptr;
  int r = dev_alloc(c, (void**)&dt, sizeof(DemonTable));
  if (!r) r = dev_alloc(c, (void**)&ds, sizeof(DemonState));
  DemonState h{};
  for (int k = 0; k < 4; ++k) h.E[k] = E[k];
  if (!r && (hipMemcpyAsync(dt, ht, sizeof(DemonTable), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
             hipMemcpyAsync(ds, &h, sizeof h, hipMemcpyHostToDevice, c->stream) != hipSuccess))
    r = fail(c, MAG_ERR_DEVICE, "demon upload failed");
  for (int sw = 0; sw < nsweeps && !r; ++sw) {
    KTimer t(c, "demon");
    hipLaunchKernelGGL(k_demon_sweep, dim3(1), dim3(64), 0, c->stream, LinkView<double, 18>{(double*)U->p, U->g.V},
                       U->g, dt, ds, E_max, seed, sw);
    if (hipGetLastError() != hipSuccess) r = fail(c, MAG_ERR_DEVICE, "demon launch failed");
  }
  if (!r && (hipMemcpyAsync(&h, ds, sizeof h, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
             hipStreamSynchronize(c->stream) != hipSuccess))
    r = fail(c, MAG_ERR_DEVICE, "demon download failed");
  (void)hipStreamSynchronize(c->stream);
  dev_free(c, dt, dt ? sizeof(DemonTable) : 0);
  dev_free(c, ds, ds ? sizeof(DemonState) : 0);
  delete ht;
  if (r) return r;
  for (int k = 0; k < 4; ++k) E[k] = h.E[k];
  acc_rej[0] = (long)h.acc;
  acc_rej[1] = (long)h.rej;
  return MAG_OK;
}

extern "C" int mag_demon_checkerboard(mag_field U, mag_field Ep, mag_field Er, mag_field Ec, mag_field Et, int nsweeps,
                                      double E_max, uint64_t seed, long acc_rej[2]) {
  mag_field Ef[4] = {Ep, Er, Ec, Et};
  if (!U || nsweeps < 0) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  if (U->kind != MAG_LINK18 || U->prec != MAG_F64)
    return fail(c, MAG_ERR_UNSUPPORTED, "the demon runs on FP64 18-real links");
  for (int k = 0; k < 4; ++k) {
    if (!Ef[k] || Ef[k]->kind != MAG_REAL1 || Ef[k]->ctx != c)
      return fail(c, MAG_ERR_ARG, "demon energies are four MAG_REAL1 fields of the links' context");
    for (int i = 0; i < 4; ++i)
      if (Ef[k]->dims[i] != U->dims[i]) return fail(c, MAG_ERR_GEOMETRY, "demon fields differ in extents");
  }
  for (int i = 0; i < 4; ++i)
    if (U->g.L[i] % 4) return fail(c, MAG_ERR_GEOMETRY, "the checkerboard demon needs extents that are multiples of 4");
  if (c->rec) return fail(c, MAG_ERR_UNSUPPORTED, "the demon is not recordable");
  const int nthr = U->g.V / 128;
  int r;
  Workspace* ws = workspace(c, U->g, MAG_F64, 18);
  if ((r = ws->need_cnt())) return r;  // V doubles >= 2 per thread
  double* cnt = (double*)ws->Cnt;
  if ((r = fill_dev(c, cnt, 0., 2 * nthr))) return r;
  DemonTable* ht = new DemonTable();
  build_table(*ht);
  if (!table_ok(*ht)) {
    delete ht;
    return fail(c, MAG_ERR_STATE, "demon path table does not have the layout the class kernel assumes");
  }
  DemonTable* dt = nullptr;
  r = dev_alloc(c, (void**)&dt, sizeof(DemonTable));
  if (!r && hipMemcpyAsync(dt, ht, sizeof(DemonTable), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    r = fail(c, MAG_ERR_DEVICE, "demon table upload failed");
  DemonClassArgs A{};
  A.U = LinkView<double, 18>{(double*)U->p, U->g.V};
  A.g = U->g;
  A.tab = dt;
  for (int k = 0; k < 4; ++k) A.E[k] = (double*)Ef[k]->p;
  A.cnt = cnt;
  A.emax = E_max;
  A.seed = seed;
  constexpr int LPL = DEMON_LPL, MPW = 64 / LPL;
  void (*ks[4])(DemonClassArgs) = {k_demon_class<0, LPL>, k_demon_class<1, LPL>, k_demon_class<2, LPL>,
                                   k_demon_class<3, LPL>};
  A.nthr = nthr;
  const dim3 grid((nthr + MPW - 1) / MPW);
  for (int sw = 0; sw < nsweeps && !r; ++sw) {
    A.sweep = sw;
    for (int mu = 0; mu < 4 && !r; ++mu)
      for (int cl = 0; cl < 128 && !r; ++cl) {
        A.a = cl & 1;
        A.cls = cl >> 1;
        KTimer t(c, "demon_checkerboard");
        hipLaunchKernelGGL(ks[mu], grid, dim3(64), 0, c->stream, A);
        if (hipGetLastError() != hipSuccess) r = fail(c, MAG_ERR_DEVICE, "demon launch failed");
      }
  }
  std::vector<double> h((size_t)2 * nthr);
  if (!r && (hipMemcpyAsync(h.data(), cnt, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
                 hipSuccess ||
             hipStreamSynchronize(c->stream) != hipSuccess))
    r = fail(c, MAG_ERR_DEVICE, "demon download failed");
  (void)hipStreamSynchronize(c->stream);
  dev_free(c, dt, dt ? sizeof(DemonTable) : 0);
  delete ht;
  if (r) return r;
  if (acc_rej) {
    double a = 0., b = 0.;
    for (int t = 0; t < nthr; ++t) {
      a += h[(size_t)2 * t];
      b += h[(size_t)2 * t + 1];
    }
    acc_rej[0] = (long)a;
    acc_rej[1] = (long)b;
  }
  return MAG_OK;
}

namespace {
int staple_table_field(mag_field S, mag_field U, int kind) {
  if (!S || !U || S->ctx != U->ctx) return MAG_ERR_ARG;
  mag_ctx c = U->ctx;
  if (U->kind != MAG_LINK18 || U->prec != MAG_F64 || S->kind != MAG_LINK18 || S->prec != MAG_F64)
    return fail(c, MAG_ERR_UNSUPPORTED, "chair/twist staples take and give FP64 18-real fields");
  for (int i = 0; i < 4; ++i)
    if (S->dims[i] != U->dims[i]) return fail(c, MAG_ERR_ARG, "chair/twist staples: fields differ in shape");
  if (S->p == U->p) return fail(c, MAG_ERR_ARG, "chair/twist staples: output must not alias the links");
  if (c->rec) return fail(c, MAG_ERR_UNSUPPORTED, "chair/twist staples are not recordable");
  DemonTable* ht = new DemonTable();
  build_table(*ht);
  DemonTable* dt = nullptr;
  int r = dev_alloc(c, (void**)&dt, sizeof(DemonTable));
  if (!r && hipMemcpyAsync(dt, ht, sizeof(DemonTable), hipMemcpyHostToDevice, c->stream) != hipSuccess)
    r = fail(c, MAG_ERR_DEVICE, "staple table upload failed");
  if (!r) {
    KTimer t(c, "staple_table");
    hipLaunchKernelGGL(k_staple_table_sum, grid_links(U->g), dim3(BLOCK), 0, c->stream,
                       LinkView<double, 18>{(double*)U->p, U->g.V}, U->g, (const DemonTable*)dt, kind,
                       LinkView<double, 18>{(double*)S->p, U->g.V});
    if (hipGetLastError() != hipSuccess) r = fail(c, MAG_ERR_DEVICE, "staple table launch failed");
  }
  (void)hipStreamSynchronize(c->stream);
  dev_free(c, dt, dt ? sizeof(DemonTable) : 0);
  delete ht;
  return r;
}
}  // namespace

extern "C" int mag_chair_staple_dagger(mag_field S, mag_field U) { return staple_table_field(S, U, 2); }
extern "C" int mag_twist_staple_dagger(mag_field S, mag_field U) { return staple_table_field(S, U, 3); }
