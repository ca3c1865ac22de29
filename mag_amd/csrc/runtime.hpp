// Host-side runtime of libmag_amd.so: context, device fields, per-shape
// workspace, kernel timing, and the typed operation table Ops<TS, NC> whose
// members (defined in ops_impl.hpp) launch the gfx950 kernels. Ops is
// explicitly instantiated once per storage format in its own translation
// unit (inst_f64.hip: FP64 18-real links; inst_f32.hip: FP32 12-real links),
// so the heavy kernels compile in parallel and only once.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mag_amd.h"
#include "lattice.hpp"

struct mag_ctx_s;
struct mag_field_s;

namespace mag {

constexpr int RT_BLOCK = 256;

struct Workspace;
struct Recorder;

struct TimerRec {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
  double ms = 0.;
  long launches = 0;
};

// Device and pinned-host blocks of the chains a server retires, kept for the
// chains it admits next instead of freed: hipFree and hipHostFree wait for
// the whole device, which would stall every batch in flight each time a
// chain retires. Blocks are reused by exact size (a serving stream repeats
// a handful of lattice shapes); when an allocation fails the cached blocks
// are released and the allocation retried.
struct MemPool {
  std::multimap<size_t, void*> dev, host;
  size_t dev_cached = 0;
  ~MemPool() { release(); }
  void* take(std::multimap<size_t, void*>& m, size_t bytes) {
    auto it = m.find(bytes);
    if (it == m.end()) return nullptr;
    void* p = it->second;
    m.erase(it);
    return p;
  }
  void release() {
    for (auto& kv : dev) (void)hipFree(kv.second);
    for (auto& kv : host) (void)hipHostFree(kv.second);
    dev.clear();
    host.clear();
    dev_cached = 0;
  }
};

}  // namespace mag

struct mag_ctx_s {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  double* d_part = nullptr;
  size_t part_cap = 0;
  double* d_scalar = nullptr;
  double* h_scalar = nullptr;
  size_t bytes = 0;
  std::map<std::tuple<int, int, int, int, int, int>, std::unique_ptr<mag::Workspace>> ws;
  bool timing = false;
  std::string timing_family;  // when not empty, only this kernel family is timed
  std::map<std::string, mag::TimerRec> timers;
  std::vector<hipEvent_t> event_pool;
  mag::Recorder* rec = nullptr;  // set while a serving batch records its launches
  bool owns_stream = true;
  std::shared_ptr<mag::MemPool> pool;  // serving child contexts: frees go back to the server's pool
};

struct mag_field_s {
  mag_ctx ctx = nullptr;
  int dims[4] = {0, 0, 0, 0};
  int kind = 0, prec = 0;
  void* p = nullptr;
  size_t bytes = 0;
  mag::Geom g;
  bool owned = true;
};

namespace mag {

inline int fail(mag_ctx c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

#define HIPCHK(ctx, call)                                                                  \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return ::mag::fail(ctx, MAG_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define CHECK_LAUNCH(ctx) HIPCHK(ctx, hipGetLastError())

inline int ncomp(int kind) {
  return kind == MAG_LINK18 ? 18 : (kind == MAG_LINK12 ? 12 : (kind == MAG_ALG8 ? 8 : 1));
}
inline size_t real_size(int prec) { return prec == MAG_F64 ? 8 : 4; }

inline hipEvent_t take_event(mag_ctx c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  (void)hipEventCreate(&e);
  return e;
}

// Records HIP events around one kernel launch on the context's stream when
// timing is on (for every family, or only the one set by
// mag_ctx_set_timing_family); elapsed times are collected lazily by
// collect_timers.
struct KTimer {
  mag_ctx c;
  const char* name;
  bool on;
  hipEvent_t a = nullptr, b = nullptr;
  KTimer(mag_ctx c_, const char* n)
      : c(c_), name(n), on(c_->timing && (c_->timing_family.empty() || c_->timing_family == n)) {
    if (on) {
      a = take_event(c);
      b = take_event(c);
      (void)hipEventRecord(a, c->stream);
    }
  }
  ~KTimer() {
    if (on) {
      (void)hipEventRecord(b, c->stream);
      c->timers[name].pending.emplace_back(a, b);
    }
  }
};

inline void collect_timers(mag_ctx c) {
  for (auto& kv : c->timers) {
    for (auto& ev : kv.second.pending) {
      (void)hipEventSynchronize(ev.second);
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, ev.first, ev.second);
      kv.second.ms += ms;
      kv.second.launches++;
      c->event_pool.push_back(ev.first);
      c->event_pool.push_back(ev.second);
    }
    kv.second.pending.clear();
  }
}

inline int dev_alloc(mag_ctx c, void** p, size_t bytes) {
  *p = nullptr;
  if (c->pool && (*p = c->pool->take(c->pool->dev, bytes))) {
    c->pool->dev_cached -= bytes;
    c->bytes += bytes;
    return MAG_OK;
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess && c->pool && c->pool->dev_cached) {  // memory held by the pool: give it back, retry
    (void)hipGetLastError();
    c->pool->release();
    e = hipMalloc(p, bytes);
  }
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(c, MAG_ERR_OOM, "device allocation of " + std::to_string(bytes) + " bytes failed");
  }
  c->bytes += bytes;
  return MAG_OK;
}
inline void dev_free(mag_ctx c, void* p, size_t bytes) {
  if (!p) return;
  if (c->pool) {
    c->pool->dev.emplace(bytes, p);
    c->pool->dev_cached += bytes;
  } else {
    (void)hipFree(p);
  }
  c->bytes -= bytes;
}
inline int host_alloc(mag_ctx c, void** p, size_t bytes) {
  if (c->pool && (*p = c->pool->take(c->pool->host, bytes))) return MAG_OK;
  if (hipHostMalloc(p, bytes, hipHostMallocDefault) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(c, MAG_ERR_OOM, "pinned host allocation failed");
  }
  return MAG_OK;
}
inline void host_free(mag_ctx c, void* p, size_t bytes) {
  if (!p) return;
  if (c->pool) c->pool->host.emplace(bytes, p);
  else (void)hipHostFree(p);
}

inline int ensure_part(mag_ctx c, size_t n) {
  if (n <= c->part_cap) return MAG_OK;
  // recorded launches keep the buffer's address: it must not move then
  if (c->rec) return fail(c, MAG_ERR_STATE, "partial-sum buffer too small while recording");
  if (c->d_part) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    dev_free(c, c->d_part, c->part_cap * sizeof(double));
    c->d_part = nullptr;
    c->part_cap = 0;
  }
  const size_t cap = n < 4096 ? 4096 : n;
  int r = dev_alloc(c, (void**)&c->d_part, cap * sizeof(double));
  if (r) return r;
  c->part_cap = cap;
  return MAG_OK;
}

inline int fetch_scalars(mag_ctx c, int n) {
  HIPCHK(c, hipMemcpyAsync(c->h_scalar, c->d_scalar, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return MAG_OK;
}

inline dim3 grid_links(const Geom& g) { return dim3((g.V + RT_BLOCK - 1) / RT_BLOCK, 4, 1); }
// Stencil kernels: one 1D tile per (256-site block, mu), see stencil_tile().
inline dim3 grid_stencil(const Geom& g) { return dim3(4 * ((g.V + RT_BLOCK - 1) / RT_BLOCK), 1, 1); }

// ------------------------------------------------------------ workspace
// Scratch of one lattice shape and storage format, owned by the context and
// allocated on first use (never inside a timed trajectory once warm).
struct Workspace {
  mag_ctx c = nullptr;
  Geom g;
  int prec = 0, nc = 0;
  size_t rs = 0;
  std::vector<std::pair<void*, size_t>> allocs;
  void* Uaux = nullptr;  // link: FG auxiliary field
  void* F = nullptr;     // alg8: force (unfused paths)
  void* Z = nullptr;     // mat18: Z_0 of constrained smeared runs
  std::vector<void*> Us, Qs, Cs;  // stout levels: U_{k+1}, Q_k, C_k
  void* L[2] = {nullptr, nullptr};  // stout Lambda (alg8), ping-pong between levels
  void* W = nullptr;
  void* Uh[2] = {nullptr, nullptr};  // FP64 smeared fields for the Hamiltonian
  void* Utop = nullptr;  // FP64 12-real chains: the force's top smeared level, 18 reals
  void* Lam = nullptr;  // alg8: Lambda of the coarse/fine HMC (coarse shape)
  void* Eq = nullptr;   // mat18: e^{iQ} of the coarse/fine HMC (coarse shape)
  void* Cnt = nullptr;  // V doubles: per-thread hit counts of the local Metropolis
  void* Rec = nullptr;  // per-iteration records of the local Metropolis (grows)
  size_t rec_cap = 0;

  ~Workspace() {
    for (auto& a : allocs) dev_free(c, a.first, a.second);
  }
  int get(void** slot, size_t bytes) {
    if (*slot) return MAG_OK;
    int r = dev_alloc(c, slot, bytes);
    if (r == MAG_OK) allocs.emplace_back(*slot, bytes);
    return r;
  }
  size_t link_bytes() const { return (size_t)g.V * 4 * nc * rs; }
  size_t alg_bytes() const { return (size_t)g.V * 4 * 8 * rs; }
  size_t mat_bytes() const { return (size_t)g.V * 4 * 18 * rs; }
  size_t mat64_bytes() const { return (size_t)g.V * 4 * 18 * 8; }
  int need_md() {
    int r = get(&Uaux, link_bytes());
    if (r) return r;
    return get(&F, alg_bytes());
  }
  int need_stout(int n) {
    if ((int)Us.size() < n) {
      Us.resize(n, nullptr);
      Qs.resize(n, nullptr);
      Cs.resize(n, nullptr);
    }
    int r;
    for (int k = 0; k < n; ++k) {
      if ((r = get(&Us[k], link_bytes()))) return r;
      if ((r = get(&Qs[k], alg_bytes()))) return r;
      if ((r = get(&Cs[k], mat_bytes()))) return r;
    }
    if ((r = get(&L[0], alg_bytes()))) return r;
    if ((r = get(&L[1], alg_bytes()))) return r;
    return get(&W, mat_bytes());
  }
  int need_Z() { return get(&Z, mat_bytes()); }
  int need_top18() { return get(&Utop, mat64_bytes()); }
  int need_cnt() { return get(&Cnt, (size_t)g.V * sizeof(double)); }
  int need_rec(size_t n) {
    if (Rec && rec_cap >= n) return MAG_OK;
    if (Rec) {
      for (auto it = allocs.begin(); it != allocs.end(); ++it)
        if (it->first == Rec) {
          dev_free(c, it->first, it->second);
          allocs.erase(it);
          break;
        }
      Rec = nullptr;
    }
    int r = get(&Rec, n * sizeof(double));
    rec_cap = r ? 0 : n;
    return r;
  }
  int need_multi() {
    int r = get(&Lam, alg_bytes());
    if (r) return r;
    return get(&Eq, mat_bytes());
  }
  int need_h(int n) {
    int r = get(&Uh[0], mat64_bytes());
    if (r || n < 2) return r;
    return get(&Uh[1], mat64_bytes());
  }
};

inline Workspace* workspace(mag_ctx c, const Geom& g, int prec, int nc) {
  auto key = std::make_tuple(g.L[0], g.L[1], g.L[2], g.L[3], prec, nc);
  auto it = c->ws.find(key);
  if (it != c->ws.end()) return it->second.get();
  auto w = std::make_unique<Workspace>();
  w->c = c;
  w->g = g;
  w->prec = prec;
  w->nc = nc;
  w->rs = real_size(prec);
  Workspace* r = w.get();
  c->ws[key] = std::move(w);
  return r;
}

// ------------------------------------------------------------ typed ops
// Force epilogues (must match kernels.hpp ForceOut).
enum { OUT_F = 0, OUT_PADD = 1, OUT_UAUX = 2 };

template <class TS, int NC>
struct Ops {
  // gauge force of the (smeared) action on U: F, P += eps F, or Uaux = e^{i eps F} U
  static int force(mag_ctx c, const Geom& g, void* U, const mag_action& act, int out, void* F, void* P, void* Uaux,
                   double eps, Workspace* ws);
  static int link_update(mag_ctx c, const Geom& g, void* Uin, void* Uout, void* P, double eps, int mag);
  static int fg(mag_ctx c, const Geom& g, void* U, void* P, const mag_action& act, int steps, double dt,
                Workspace* ws);
  static int leapfrog(mag_ctx c, const Geom& g, void* U, void* P, const mag_action& act, int steps, double dt,
                      Workspace* ws);
  // plaquette / rectangle sums of U into scalar slots (slot, slot+1)
  static int action_sums(mag_ctx c, const Geom& g, void* U, bool rect, int slot);
  // the same for the n-level stout-smeared field, smeared in FP64
  static int smeared_action_sums(mag_ctx c, const Geom& g, void* U, double rho, int n, bool rect, int slot);
  static int smear(mag_ctx c, const Geom& g, void* Uout, void* U, double rho, int n, Workspace* ws);
  static int reunitarize(mag_ctx c, const Geom& g, void* U, int slot);
  static int haar(mag_ctx c, const Geom& g, void* U, uint64_t seed);
  static int unit(mag_ctx c, const Geom& g, void* U);
  static int wilson_loop(mag_ctx c, const Geom& g, void* U, const int r[3], int T, int slot, int* nperm);
  static int max_diff(mag_ctx c, const Geom& g, void* A, void* B, int slot);
};

// Dispatch on a link field's storage format.
template <class TS, int NC>
struct Tag {
  using T = TS;
  static constexpr int N = NC;
};

template <class F>
int dispatch_link(mag_ctx c, int prec, int kind, F&& f) {
  if (prec == MAG_F64 && kind == MAG_LINK18) return f(Tag<double, 18>{});
  if (prec == MAG_F32 && kind == MAG_LINK12) return f(Tag<float, 12>{});
  if (prec == MAG_F64 && kind == MAG_LINK12) return f(Tag<double, 12>{});
  return fail(c, MAG_ERR_UNSUPPORTED, "link fields are FP64 18- or 12-real, or FP32 12-real");
}

// ------------------------------------------------------------ common ops
// (common.hip: format-independent kernels)
int final_reduce(mag_ctx c, const double* part, int n, int stride, int op, int slot);
int kinetic_sum(mag_ctx c, const Geom& g, int prec, void* P, int mag, int slot);
// epoch_ptr (device memory) overrides epoch when set
int mom_refresh(mag_ctx c, const Geom& g, int prec, void* P, uint64_t seed, uint64_t epoch, double sig,
                const uint64_t* epoch_ptr = nullptr);
// kernel copy / fill (recordable; no runtime memcpy or memset nodes)
int copy_dev(mag_ctx c, void* dst, const void* src, size_t bytes);
int fill_dev(mag_ctx c, double* p, double v, int n);
int mom_update(mag_ctx c, const Geom& g, int prec, void* P, void* F, double eps);
int alg_scale(mag_ctx c, const Geom& g, int prec, void* P, double s);
int convert_links(mag_ctx c, const Geom& g, int prec_in, int kind_in, void* in, int prec_out, int kind_out, void* out);
int batch_expiQ(mag_ctx c, const double* dQ, double* dout, long n, int prec);
int batch_lambda(mag_ctx c, const double* dQ, const double* dS, const double* dU, double* dout, long n, int prec);
int batch_dexpiq(mag_ctx c, const double* dQ, const double* dDQ, double* dout, long n, int prec);
double host_uniform(uint64_t seed, uint32_t stream, uint64_t index);
// noalloc.hip: resolve the allocation-free scratch array of `device` (current)
int noalloc_setup(int device, std::string* err);
int constrained_plaq_sum(mag_ctx c, const Geom& g, int prec, int kind, void* U, int mag, int slot, double* norm);

inline double action_from_sums(const mag_action& act, double sp, double sr) {
  // field-hmc.h:632-640
  if (act.type == MAG_WILSON) return -sp * act.beta / 3.;
  return (-sp * (1. - 8. * act.c1) + (-sr) * act.c1) * act.beta / 3.;
}

}  // namespace mag
