// Continuous-batching HMC server. Independent chains (each with its own
// volume, action, beta, smearing, precision, seed and trajectory count)
// are submitted at any time. Every step():
//   1. admits pending chains in arrival order while their device footprint
//      fits the HBM budget (links, momenta, backup, workspace and RNG state
//      then stay resident until the chain retires);
//   2. groups running chains into shape classes (volume, precision, action
//      kind, smearing depth, constraint, integrator, MD steps) and splits
//      each class into batches of at most max_batch chains;
//   3. runs one trajectory of every chain: a batch's trajectory is recorded
//      once per chain through the launch layer and emitted as one launch per
//      kernel over all its chains (blockIdx.z = chain), optionally captured
//      in a HIP graph and replayed while the batch membership is unchanged;
//      batches go to a pool of streams, so batches of different classes
//      overlap on the device;
//   4. gathers each chain's FP64 scalars with one copy per batch, makes the
//      accept decisions on the host, restores rejected chains, records
//      per-trajectory completion times, retires finished, cancelled and
//      non-finite chains (freeing their memory for the next admissions).
// A chain's kernels and reductions are the same alone or in a batch, so its
// results do not depend on what it is batched with.
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "chain_internal.hpp"
#include "launch.hpp"

using namespace mag;

namespace {

constexpr int NSTREAMS = 4;

enum { ST_PENDING = 0, ST_RUNNING = 1, ST_DONE = 2, ST_FAILED = 3, ST_CANCELLED = 4 };

struct Entry {
  int id = 0;
  mag_serve_request req{};
  int state = ST_PENDING;
  mag_chain chain = nullptr;
  size_t footprint = 0;
  std::vector<mag_traj_result> results;
  std::vector<double> t_done;
  bool cancel = false;
};

using ClassKey = std::tuple<int, int, int, int, int, int, int, int, int, int, int>;

ClassKey class_of(const mag_serve_request& r) {
  return ClassKey(r.dims[0], r.dims[1], r.dims[2], r.dims[3], r.hmc.precision == MAG_F32 ? 1 : 0, r.act.type,
                  r.act.n_stout, r.act.mag, r.hmc.integrator, r.hmc.trajectory_length, r.hmc.reunitarize ? 1 : 0);
}

struct Batch {
  std::vector<Entry*> members;
  hipStream_t stream = nullptr;
  Recorder rec;
  void* tables = nullptr;
  size_t table_bytes = 0;
  double* gather = nullptr;  // S_COUNT doubles per member (device)
  double* h_gather = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool valid = false;
};

}  // namespace

struct mag_server_s {
  mag_ctx ctx = nullptr;
  size_t budget = 0, used = 0;
  int max_batch = 64;
  bool graphs = true;
  hipStream_t streams[NSTREAMS] = {};
  std::vector<std::unique_ptr<Entry>> entries;
  std::deque<Entry*> pending;
  std::map<std::pair<ClassKey, int>, std::unique_ptr<Batch>> batches;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  long trajectories = 0;
};

namespace {

double now_s(mag_server s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

void free_batch(mag_server s, Batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  if (b->tables) dev_free(s->ctx, b->tables, b->table_bytes);
  if (b->gather) dev_free(s->ctx, b->gather, b->members.size() * S_COUNT * sizeof(double));
  if (b->h_gather) (void)hipHostFree(b->h_gather);
  b->tables = nullptr;
  b->gather = nullptr;
  b->h_gather = nullptr;
  b->rec = Recorder();
  b->valid = false;
}

// Records the trajectory of every member, uploads the argument tables and
// (optionally) captures the emitted launches in a graph.
int build_batch(mag_server s, Batch* b) {
  mag_ctx x = s->ctx;
  const size_t n = b->members.size();
  int r;
  free_batch(s, b);
  if ((r = dev_alloc(x, (void**)&b->gather, n * S_COUNT * sizeof(double)))) return r;
  if (hipHostMalloc(&b->h_gather, n * S_COUNT * sizeof(double), hipHostMallocDefault) != hipSuccess)
    return fail(x, MAG_ERR_OOM, "pinned gather buffer");
  b->rec = Recorder();
  for (size_t i = 0; i < n; ++i) {
    mag_chain c = b->members[i]->chain;
    c->ctx->stream = b->stream;
    b->rec.begin_chain((int)i);
    c->ctx->rec = &b->rec;
    r = chain_traj_kernels(c, true);
    if (!r) r = copy_dev(c->ctx, b->gather + i * S_COUNT, c->ctx->d_scalar, S_COUNT * sizeof(double));
    c->ctx->rec = nullptr;
    if (r) return r;
    if (b->rec.mismatch || (i > 0 && b->rec.pos != b->rec.calls.size()))
      return fail(x, MAG_ERR_STATE, "chains of one class recorded different launch sequences");
  }
  // argument tables, one device block
  size_t total = 0;
  for (auto& c : b->rec.calls) total += (c.blob.size() + 255) / 256 * 256;
  if ((r = dev_alloc(x, &b->tables, total))) return r;
  b->table_bytes = total;
  std::vector<char> host(total);
  size_t off = 0;
  for (auto& c : b->rec.calls) {
    std::memcpy(host.data() + off, c.blob.data(), c.blob.size());
    c.table = (char*)b->tables + off;
    off += (c.blob.size() + 255) / 256 * 256;
  }
  HIPCHK(x, hipMemcpyAsync(b->tables, host.data(), total, hipMemcpyHostToDevice, b->stream));
  HIPCHK(x, hipStreamSynchronize(b->stream));
  if (s->graphs) {
    HIPCHK(x, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (auto& c : b->rec.calls) {
      void* args[] = {&c.table};
      (void)hipLaunchKernel(c.kb, dim3(c.grid.x, c.grid.y, (unsigned)n), dim3(BLOCK), args, 0, b->stream);
    }
    HIPCHK(x, hipStreamEndCapture(b->stream, &b->graph));
    HIPCHK(x, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  b->valid = true;
  return MAG_OK;
}

int launch_batch(mag_server s, Batch* b) {
  mag_ctx x = s->ctx;
  const size_t n = b->members.size();
  // trajectory index of every member (read by the momentum kernel)
  for (size_t i = 0; i < n; ++i) {
    mag_chain c = b->members[i]->chain;
    uint64_t e = (uint64_t)c->traj;
    HIPCHK(x, hipMemcpyAsync(c->d_epoch, &e, sizeof e, hipMemcpyHostToDevice, b->stream));
  }
  if (b->exec) {
    HIPCHK(x, hipGraphLaunch(b->exec, b->stream));
  } else {
    for (auto& c : b->rec.calls) {
      void* args[] = {&c.table};
      HIPCHK(x, hipLaunchKernel(c.kb, dim3(c.grid.x, c.grid.y, (unsigned)n), dim3(BLOCK), args, 0, b->stream));
    }
  }
  HIPCHK(x, hipMemcpyAsync(b->h_gather, b->gather, n * S_COUNT * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  return MAG_OK;
}

void retire(mag_server s, Entry* e, int state) {
  if (e->chain) {
    (void)hipStreamSynchronize(e->chain->ctx->stream);
    mag_chain_destroy(e->chain);
    e->chain = nullptr;
  }
  s->used -= e->footprint;
  e->state = state;
}

int admit(mag_server s) {
  while (!s->pending.empty()) {
    Entry* e = s->pending.front();
    if (e->cancel) {
      s->pending.pop_front();
      e->state = ST_CANCELLED;
      continue;
    }
    if (s->used + e->footprint > s->budget) break;  // arrival order: wait for memory
    s->pending.pop_front();
    mag_ctx cc = nullptr;
    int r = ctx_child_create(s->ctx, s->streams[0], &cc);
    if (r) return r;
    mag_chain ch = nullptr;
    r = mag_chain_create(cc, e->req.dims, &e->req.act, &e->req.hmc, &ch);
    if (r) {
      s->ctx->err = cc->err;
      mag_ctx_destroy(cc);
      e->state = ST_FAILED;
      continue;
    }
    ch->owns_ctx = true;
    e->chain = ch;
    if ((r = mag_haar_links(ch->U, e->req.links_seed))) return r;
    if (e->req.hmc.accept_mode != 0 && (r = chain_init_demon(ch))) return r;
    s->used += e->footprint;
    e->state = ST_RUNNING;
  }
  // admission work ran on stream 0; batches may run on any stream
  HIPCHK(s->ctx, hipStreamSynchronize(s->streams[0]));
  return MAG_OK;
}

}  // namespace

extern "C" {

int mag_server_footprint(const mag_serve_request* req, size_t* bytes) {
  if (!req || !bytes) return MAG_ERR_ARG;
  *bytes = chain_footprint(req->dims, req->act, req->hmc);
  return MAG_OK;
}

int mag_server_create(mag_ctx ctx, size_t hbm_budget, int max_batch, int use_graphs, mag_server* out) {
  if (!ctx || !out || max_batch < 1) return MAG_ERR_ARG;
  auto* s = new mag_server_s();
  s->ctx = ctx;
  s->budget = hbm_budget;
  s->max_batch = max_batch;
  s->graphs = use_graphs != 0;
  for (int i = 0; i < NSTREAMS; ++i)
    if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess) {
      mag_server_destroy(s);
      return fail(ctx, MAG_ERR_DEVICE, "stream creation failed");
    }
  *out = s;
  return MAG_OK;
}

int mag_server_destroy(mag_server s) {
  if (!s) return MAG_ERR_ARG;
  for (auto& st : s->streams)
    if (st) (void)hipStreamSynchronize(st);
  for (auto& kv : s->batches) free_batch(s, kv.second.get());
  for (auto& e : s->entries)
    if (e->chain) retire(s, e.get(), ST_CANCELLED);
  for (auto& st : s->streams)
    if (st) (void)hipStreamDestroy(st);
  delete s;
  return MAG_OK;
}

int mag_server_submit(mag_server s, const mag_serve_request* req, int* id) {
  if (!s || !req || !id || req->ntraj < 0) return MAG_ERR_ARG;
  const size_t fp = chain_footprint(req->dims, req->act, req->hmc);
  if (fp > s->budget) return fail(s->ctx, MAG_ERR_BUDGET, "chain footprint exceeds the server's HBM budget");
  for (int i = 0; i < 4; ++i)
    if (req->dims[i] < 4) return fail(s->ctx, MAG_ERR_GEOMETRY, "lattice extents must be >= 4");
  auto e = std::make_unique<Entry>();
  e->id = (int)s->entries.size();
  e->req = *req;
  e->footprint = fp;
  e->results.reserve((size_t)req->ntraj);
  *id = e->id;
  if (req->ntraj == 0) e->state = ST_DONE;
  else s->pending.push_back(e.get());
  s->entries.push_back(std::move(e));
  return MAG_OK;
}

int mag_server_cancel(mag_server s, int id) {
  if (!s || id < 0 || id >= (int)s->entries.size()) return MAG_ERR_ARG;
  s->entries[id]->cancel = true;
  return MAG_OK;
}

int mag_server_step(mag_server s, int* active, int* pending) {
  if (!s) return MAG_ERR_ARG;
  int r;
  if ((r = admit(s))) return r;
  // group running chains into shape classes and batches
  std::map<ClassKey, std::vector<Entry*>> classes;
  for (auto& e : s->entries)
    if (e->state == ST_RUNNING) classes[class_of(e->req)].push_back(e.get());
  std::map<std::pair<ClassKey, int>, std::unique_ptr<Batch>> next;
  int si = 0;
  std::vector<Batch*> order;
  for (auto& kv : classes) {
    const auto& v = kv.second;
    for (size_t b0 = 0, bi = 0; b0 < v.size(); b0 += s->max_batch, ++bi) {
      std::vector<Entry*> mem(v.begin() + b0, v.begin() + std::min(v.size(), b0 + (size_t)s->max_batch));
      auto key = std::make_pair(kv.first, (int)bi);
      auto it = s->batches.find(key);
      std::unique_ptr<Batch> b;
      if (it != s->batches.end() && it->second->members == mem) {
        b = std::move(it->second);
      } else {
        if (it != s->batches.end()) free_batch(s, it->second.get());
        b = std::make_unique<Batch>();
        b->members = mem;
      }
      b->stream = s->streams[si++ % NSTREAMS];
      order.push_back(b.get());
      next[key] = std::move(b);
    }
  }
  for (auto& kv : s->batches)
    if (kv.second) free_batch(s, kv.second.get());
  s->batches = std::move(next);
  // launch every batch (streams overlap), then wait
  for (Batch* b : order) {
    for (Entry* e : b->members) e->chain->ctx->stream = b->stream;
    if (!b->valid && (r = build_batch(s, b))) return r;
    if ((r = launch_batch(s, b))) return r;
  }
  for (auto& st : s->streams) HIPCHK(s->ctx, hipStreamSynchronize(st));
  const double t = now_s(s);
  // accept steps, results, retirement
  for (Batch* b : order) {
    bool changed = false;
    for (size_t i = 0; i < b->members.size(); ++i) {
      Entry* e = b->members[i];
      mag_traj_result res{};
      if ((r = chain_accept(e->chain, b->h_gather + i * S_COUNT, &res))) return r;
      e->results.push_back(res);
      e->t_done.push_back(t);
      ++s->trajectories;
      if (res.status != MAG_OK) {
        retire(s, e, ST_FAILED);
        changed = true;
      } else if (e->cancel) {
        retire(s, e, ST_CANCELLED);
        changed = true;
      } else if ((int)e->results.size() >= e->req.ntraj) {
        retire(s, e, ST_DONE);
        changed = true;
      }
    }
    if (changed) b->valid = false;
  }
  for (auto& st : s->streams) HIPCHK(s->ctx, hipStreamSynchronize(st));
  int na = 0, np = 0;
  for (auto& e : s->entries) {
    na += e->state == ST_RUNNING;
    np += e->state == ST_PENDING;
  }
  if (active) *active = na;
  if (pending) *pending = np;
  return MAG_OK;
}

int mag_server_status(mag_server s, int id, int* state, int* done_traj) {
  if (!s || id < 0 || id >= (int)s->entries.size()) return MAG_ERR_ARG;
  Entry* e = s->entries[id].get();
  if (state) *state = e->state;
  if (done_traj) *done_traj = (int)e->results.size();
  return MAG_OK;
}

int mag_server_results(mag_server s, int id, mag_traj_result* out, double* t_done, int max) {
  if (!s || id < 0 || id >= (int)s->entries.size() || max < 0) return MAG_ERR_ARG;
  Entry* e = s->entries[id].get();
  const int n = std::min<int>(max, (int)e->results.size());
  for (int i = 0; i < n; ++i) {
    if (out) out[i] = e->results[i];
    if (t_done) t_done[i] = e->t_done[i];
  }
  return MAG_OK;
}

int mag_server_info(mag_server s, size_t* used_bytes, long* trajectories, double* elapsed_s) {
  if (!s) return MAG_ERR_ARG;
  if (used_bytes) *used_bytes = s->used;
  if (trajectories) *trajectories = s->trajectories;
  if (elapsed_s) *elapsed_s = now_s(s);
  return MAG_OK;
}

}  // extern "C"
