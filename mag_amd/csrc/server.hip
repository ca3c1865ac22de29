// Continuous-batching HMC server. Independent chains (each with its own
// volume, action, beta, smearing, precision, seed and trajectory count)
// are submitted at any time and run asynchronously:
//   - admission: pending chains are admitted in arrival order while their
//     device footprint fits the HBM budget (links, momenta, backup,
//     workspace and RNG state then stay resident until the chain retires);
//   - batches: chains that are not on the device are grouped into shape
//     classes (volume, precision, action kind, smearing depth, constraint,
//     integrator, MD steps) and batches of at most max_batch chains. A
//     batch's trajectory is recorded once per chain through the launch layer
//     and emitted as one launch per kernel over all its chains
//     (blockIdx.z = chain), optionally captured in a HIP graph that is
//     replayed while the membership is unchanged;
//   - overlap: every batch is in flight on its own schedule and its
//     completion is an event, so a batch of small lattices runs its next
//     trajectory as soon as its last one finished. Batches of large lattices
//     (whose kernels fill the device by themselves) queue on one stream;
//     small batches go to the less loaded of two others, so they never wait
//     in a hardware queue behind a large trajectory. Admission, table
//     uploads and graph capture run on the context's own stream, and every
//     host->device copy of a launch comes from pinned memory, so launching
//     never blocks the host;
//   - memory: chains and batches draw device and pinned-host blocks from a
//     pool owned by the server and return them to it; nothing is freed
//     while the server runs (hipFree would wait for the whole device);
//   - completion: a finished batch's FP64 scalars (one copy per batch) feed
//     the accept decisions on the host; rejected chains are restored on the
//     batch's stream; finished, cancelled and non-finite chains retire, and
//     their memory returns to the budget once the device is done with it.
// mag_server_step() admits, launches every idle batch, then waits until at
// least one batch in flight has completed and processes all that have.
// A chain's kernels and reductions are the same alone or in a batch, so its
// results do not depend on what it is batched with or when.
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <thread>
#include <tuple>
#include <vector>

#include "chain_internal.hpp"
#include "launch.hpp"

using namespace mag;

namespace {

constexpr int NSTREAMS = 3;             // stream 0: large batches; 1, 2: small ones
constexpr double LARGE_SITES = 65536.;  // a batch of >= 16^4 sites in all fills the device

enum { ST_PENDING = 0, ST_RUNNING = 1, ST_DONE = 2, ST_FAILED = 3, ST_CANCELLED = 4 };

struct Batch;

struct Entry {
  int id = 0;
  mag_serve_request req{};
  int state = ST_PENDING;
  mag_chain chain = nullptr;
  size_t footprint = 0;
  std::vector<mag_traj_result> results;
  std::vector<double> t_done;
  bool cancel = false;
  Batch* batch = nullptr;  // batch the chain belongs to (null: not yet batched)
  // Recorded on the chain's stream after its accept step (which may restore
  // the links with a copy): its next trajectory, on whatever stream its batch
  // is launched next, waits for it.
  hipEvent_t ready = nullptr;
  bool ready_recorded = false;
};

using ClassKey = std::tuple<int, int, int, int, int, int, int, int, int, int, int>;

ClassKey class_of(const mag_serve_request& r) {
  const int fmt = r.hmc.precision == MAG_F32 ? 1 : (r.hmc.recon12 ? 2 : 0);  // link storage format
  return ClassKey(r.dims[0], r.dims[1], r.dims[2], r.dims[3], fmt, r.act.type,
                  r.act.n_stout, r.act.mag, r.hmc.integrator, r.hmc.trajectory_length, r.hmc.reunitarize ? 1 : 0);
}

// Relative device work of one trajectory of one chain (for stream balancing).
double traj_cost(const mag_serve_request& r) {
  const double V = (double)r.dims[0] * r.dims[1] * r.dims[2] * r.dims[3];
  const double per_force = (r.act.type == MAG_IWASAKI ? 4. : 1.) + 3. * r.act.n_stout;
  return V * (r.hmc.trajectory_length * (r.hmc.integrator == 0 ? 4. : 1.) * per_force + 4.);
}

struct Batch {
  ClassKey key;
  std::vector<Entry*> members;
  int sidx = 0;
  double cost = 0.;
  Recorder rec;
  void* tables = nullptr;
  size_t table_bytes = 0;
  double* gather = nullptr;  // S_COUNT doubles per member (device)
  size_t gather_n = 0;
  double* h_gather = nullptr;
  uint64_t* h_epoch = nullptr;  // pinned: trajectory index of each member
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipEvent_t done = nullptr;
  bool valid = false;     // recorded tables / graph match the members
  bool inflight = false;
};

struct Grave {
  mag_chain chain;
  hipEvent_t ev;
  size_t footprint;
};

}  // namespace

struct mag_server_s {
  mag_ctx ctx = nullptr;
  size_t budget = 0, used = 0;
  int max_batch = 64;
  bool graphs = true;
  hipStream_t streams[NSTREAMS] = {};
  hipStream_t aux = nullptr;  // the context's stream: admission, table uploads, graph capture
  std::shared_ptr<MemPool> pool = std::make_shared<MemPool>();
  mag_ctx mem = nullptr;  // pool-backed context for the batches' own buffers
  double load[NSTREAMS] = {};
  std::vector<std::unique_ptr<Entry>> entries;
  std::deque<Entry*> pending;
  std::vector<std::unique_ptr<Batch>> batches;
  std::vector<Grave> graves;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  long trajectories = 0;
  // host seconds spent in: admission, batch builds (record, upload, graph
  // capture), launches, waiting for completions, accept/retire processing
  double host_s[5] = {0., 0., 0., 0., 0.};
  long builds = 0;
};

namespace {

double now_s(mag_server s) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - s->t0).count();
}

void release_tables(mag_server s, Batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  dev_free(s->mem, b->tables, b->table_bytes);
  dev_free(s->mem, b->gather, b->gather_n * S_COUNT * sizeof(double));
  host_free(s->mem, b->h_gather, b->gather_n * S_COUNT * sizeof(double));
  host_free(s->mem, b->h_epoch, b->gather_n * sizeof(uint64_t));
  b->tables = nullptr;
  b->gather = nullptr;
  b->gather_n = 0;
  b->h_gather = nullptr;
  b->h_epoch = nullptr;
  b->rec = Recorder();
  b->valid = false;
}

void destroy_batch(mag_server s, Batch* b) {
  release_tables(s, b);
  if (b->done) (void)hipEventDestroy(b->done);
  b->done = nullptr;
}

// Records the trajectory of every member, uploads the argument tables and
// (optionally) captures the emitted launches in a graph, all on the
// auxiliary stream, which is then synchronised (no batch in flight waits).
int build_batch(mag_server s, Batch* b) {
  mag_ctx x = s->ctx;
  const size_t n = b->members.size();
  int r;
  release_tables(s, b);
  if (!b->done && hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess)
    return fail(x, MAG_ERR_DEVICE, "event creation failed");
  b->gather_n = n;
  if ((r = dev_alloc(s->mem, (void**)&b->gather, n * S_COUNT * sizeof(double))) ||
      (r = host_alloc(s->mem, (void**)&b->h_gather, n * S_COUNT * sizeof(double))) ||
      (r = host_alloc(s->mem, (void**)&b->h_epoch, n * sizeof(uint64_t))))
    return fail(x, r, s->mem->err);
  b->rec = Recorder();
  for (size_t i = 0; i < n; ++i) {
    mag_chain c = b->members[i]->chain;
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
  if ((r = dev_alloc(s->mem, &b->tables, total))) return fail(x, r, s->mem->err);
  b->table_bytes = total;
  std::vector<char> host(total);
  size_t off = 0;
  for (auto& c : b->rec.calls) {
    std::memcpy(host.data() + off, c.blob.data(), c.blob.size());
    c.table = (char*)b->tables + off;
    off += (c.blob.size() + 255) / 256 * 256;
  }
  HIPCHK(x, hipMemcpyAsync(b->tables, host.data(), total, hipMemcpyHostToDevice, s->aux));
  HIPCHK(x, hipStreamSynchronize(s->aux));
  if (s->graphs) {
    HIPCHK(x, hipStreamBeginCapture(s->aux, hipStreamCaptureModeThreadLocal));
    for (auto& c : b->rec.calls) {
      void* args[] = {&c.table};
      (void)hipLaunchKernel(c.kb, dim3(c.grid.x, c.grid.y, (unsigned)n), dim3(BLOCK), args, 0, s->aux);
    }
    HIPCHK(x, hipStreamEndCapture(s->aux, &b->graph));
    HIPCHK(x, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  }
  b->valid = true;
  return MAG_OK;
}

int launch_batch(mag_server s, Batch* b) {
  mag_ctx x = s->ctx;
  const size_t n = b->members.size();
  double sites = 0.;
  for (Entry* e : b->members)
    sites += (double)e->req.dims[0] * e->req.dims[1] * e->req.dims[2] * e->req.dims[3];
  int best = 0;
  if (sites < LARGE_SITES) best = (s->load[1] <= s->load[2]) ? 1 : 2;
  b->sidx = best;
  hipStream_t st = s->streams[best];
  for (Entry* e : b->members) {
    if (e->ready_recorded) HIPCHK(x, hipStreamWaitEvent(st, e->ready, 0));
    e->chain->ctx->stream = st;
  }
  // trajectory index of every member (read by the momentum kernel), from
  // pinned memory: the copies are asynchronous
  for (size_t i = 0; i < n; ++i) {
    mag_chain c = b->members[i]->chain;
    b->h_epoch[i] = (uint64_t)c->traj;
    HIPCHK(x, hipMemcpyAsync(c->d_epoch, b->h_epoch + i, sizeof(uint64_t), hipMemcpyHostToDevice, st));
  }
  if (b->exec) {
    HIPCHK(x, hipGraphLaunch(b->exec, st));
  } else {
    for (auto& c : b->rec.calls) {
      void* args[] = {&c.table};
      HIPCHK(x, hipLaunchKernel(c.kb, dim3(c.grid.x, c.grid.y, (unsigned)n), dim3(BLOCK), args, 0, st));
    }
  }
  HIPCHK(x, hipMemcpyAsync(b->h_gather, b->gather, n * S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(x, hipEventRecord(b->done, st));
  b->cost = 0.;
  for (Entry* e : b->members) b->cost += traj_cost(e->req);
  s->load[best] += b->cost;
  b->inflight = true;
  return MAG_OK;
}

// Retires a chain: its memory is released once the work already queued on
// its stream (a restore copy, say) has finished.
int retire(mag_server s, Entry* e, int state) {
  if (e->chain) {
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(ev, e->chain->ctx->stream) != hipSuccess)
      return fail(s->ctx, MAG_ERR_DEVICE, "retire event failed");
    s->graves.push_back({e->chain, ev, e->footprint});
    e->chain = nullptr;
    if (e->ready) (void)hipEventDestroy(e->ready);
    e->ready = nullptr;
    e->ready_recorded = false;
  } else {
    s->used -= e->footprint;
  }
  if (e->batch) {
    auto& m = e->batch->members;
    for (size_t i = 0; i < m.size(); ++i)
      if (m[i] == e) {
        m.erase(m.begin() + (long)i);
        break;
      }
    e->batch->valid = false;
    e->batch = nullptr;
  }
  e->state = state;
  return MAG_OK;
}

int reap(mag_server s, bool wait) {
  std::vector<Grave> keep;
  for (Grave& g : s->graves) {
    const hipError_t q = wait ? hipEventSynchronize(g.ev) : hipEventQuery(g.ev);
    if (q == hipErrorNotReady) {
      keep.push_back(g);
      continue;
    }
    if (q != hipSuccess) return fail(s->ctx, MAG_ERR_DEVICE, "retired chain's stream failed");
    (void)hipEventDestroy(g.ev);
    mag_chain_destroy(g.chain);
    s->used -= g.footprint;
  }
  s->graves.swap(keep);
  return MAG_OK;
}

int admit(mag_server s) {
  bool any = false;
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
    int r = ctx_child_create(s->ctx, s->aux, s->pool, &cc);
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
    if (hipEventCreateWithFlags(&e->ready, hipEventDisableTiming) != hipSuccess)
      return fail(s->ctx, MAG_ERR_DEVICE, "event creation failed");
    if ((r = mag_haar_links(ch->U, e->req.links_seed))) return r;
    if (e->req.hmc.accept_mode != 0 && (r = chain_init_demon(ch))) return r;
    s->used += e->footprint;
    e->state = ST_RUNNING;
    any = true;
  }
  if (any) HIPCHK(s->ctx, hipStreamSynchronize(s->aux));
  return MAG_OK;
}

// Accept steps of a completed batch, results, retirement.
int complete(mag_server s, Batch* b) {
  const double t = now_s(s);
  s->load[b->sidx] -= b->cost;
  b->inflight = false;
  const std::vector<Entry*> mem = b->members;
  for (size_t i = 0; i < mem.size(); ++i) {
    Entry* e = mem[i];
    mag_traj_result res{};
    int r;
    if ((r = chain_accept(e->chain, b->h_gather + i * S_COUNT, &res))) return r;
    e->results.push_back(res);
    e->t_done.push_back(t);
    ++s->trajectories;
    int st = -1;
    if (res.status != MAG_OK) st = ST_FAILED;
    else if (e->cancel) st = ST_CANCELLED;
    else if ((int)e->results.size() >= e->req.ntraj) st = ST_DONE;
    if (st >= 0) {
      if ((r = retire(s, e, st))) return r;
    } else {
      HIPCHK(s->ctx, hipEventRecord(e->ready, e->chain->ctx->stream));
      e->ready_recorded = true;
    }
  }
  return MAG_OK;
}

// Regroups the chains that are not on the device and launches every idle
// batch. Idle batches of one class are merged up to max_batch; chains not
// yet batched join the class's idle batches or open new ones.
int launch_idle(mag_server s) {
  int r;
  for (auto& e : s->entries)
    if (e->state == ST_RUNNING && e->cancel && !(e->batch && e->batch->inflight) &&
        (r = retire(s, e.get(), ST_CANCELLED)))
      return r;
  std::map<ClassKey, std::vector<Batch*>> idle;
  for (auto& b : s->batches)
    if (!b->inflight && !b->members.empty()) idle[b->key].push_back(b.get());
  std::map<ClassKey, std::vector<Entry*>> loose;
  for (auto& e : s->entries)
    if (e->state == ST_RUNNING && !e->batch) loose[class_of(e->req)].push_back(e.get());
  for (auto& kv : loose) {
    auto& bs = idle[kv.first];
    for (Entry* e : kv.second) {
      Batch* home = nullptr;
      for (Batch* b : bs)
        if ((int)b->members.size() < s->max_batch) {
          home = b;
          break;
        }
      if (!home) {
        s->batches.push_back(std::make_unique<Batch>());
        home = s->batches.back().get();
        home->key = kv.first;
        bs.push_back(home);
      }
      home->members.push_back(e);
      home->valid = false;
      e->batch = home;
    }
  }
  for (auto& kv : idle) {
    auto& bs = kv.second;
    for (size_t i = 0; i < bs.size(); ++i)
      for (size_t j = i + 1; j < bs.size(); ++j)
        if (!bs[j]->members.empty() && bs[i]->members.size() + bs[j]->members.size() <= (size_t)s->max_batch) {
          for (Entry* e : bs[j]->members) {
            bs[i]->members.push_back(e);
            e->batch = bs[i];
          }
          bs[j]->members.clear();
          bs[i]->valid = false;
        }
  }
  // drop empty idle batches, launch the rest
  std::vector<std::unique_ptr<Batch>> keep;
  for (auto& b : s->batches) {
    if (!b->inflight && b->members.empty()) {
      destroy_batch(s, b.get());
      continue;
    }
    keep.push_back(std::move(b));
  }
  s->batches.swap(keep);
  for (auto& b : s->batches) {
    if (b->inflight) continue;
    if (!b->valid) {
      const double t0 = now_s(s);
      r = build_batch(s, b.get());
      s->host_s[1] += now_s(s) - t0;
      s->host_s[2] -= now_s(s) - t0;  // counted as build, not launch
      ++s->builds;
      if (r) return r;
    }
    if ((r = launch_batch(s, b.get()))) return r;
  }
  return MAG_OK;
}

// Waits until at least one batch in flight has completed (if any is), then
// completes every batch that has.
int poll(mag_server s) {
  double tw = now_s(s);
  for (;;) {
    bool inflight = false, any = false;
    for (size_t i = 0; i < s->batches.size(); ++i) {
      Batch* b = s->batches[i].get();
      if (!b->inflight) continue;
      const hipError_t q = hipEventQuery(b->done);
      if (q == hipErrorNotReady) {
        inflight = true;
        continue;
      }
      if (q != hipSuccess) return fail(s->ctx, MAG_ERR_DEVICE, "batch failed on the device");
      const double tc = now_s(s);
      s->host_s[3] += tc - tw;
      int r;
      if ((r = complete(s, b))) return r;
      tw = now_s(s);
      s->host_s[4] += tw - tc;
      any = true;
    }
    if (any || !inflight) {
      s->host_s[3] += now_s(s) - tw;
      return MAG_OK;
    }
    std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
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
  s->aux = ctx->stream;
  if (ctx_child_create(ctx, s->aux, s->pool, &s->mem)) {
    mag_server_destroy(s);
    return MAG_ERR_OOM;
  }
  *out = s;
  return MAG_OK;
}

int mag_server_destroy(mag_server s) {
  if (!s) return MAG_ERR_ARG;
  for (auto& st : s->streams)
    if (st) (void)hipStreamSynchronize(st);
  if (s->aux) (void)hipStreamSynchronize(s->aux);
  for (auto& b : s->batches) destroy_batch(s, b.get());
  for (auto& e : s->entries) {
    e->batch = nullptr;
    if (e->chain) (void)retire(s, e.get(), ST_CANCELLED);
  }
  (void)reap(s, true);
  if (s->mem) mag_ctx_destroy(s->mem);
  s->batches.clear();
  s->pool.reset();  // the last chain context is gone: cached blocks are freed here
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
  double t = now_s(s);
  if ((r = reap(s, false)) || (r = admit(s))) return r;
  double t1 = now_s(s);
  s->host_s[0] += t1 - t;
  if ((r = launch_idle(s))) return r;
  t = now_s(s);
  s->host_s[2] += t - t1;
  if ((r = poll(s))) return r;
  int na = 0, np = 0;
  for (auto& e : s->entries) {
    na += e->state == ST_RUNNING;
    np += e->state == ST_PENDING;
  }
  if (na == 0 && np == 0 && (r = reap(s, true))) return r;
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

int mag_server_host_times(mag_server s, double seconds[5], long* builds) {
  if (!s || !seconds) return MAG_ERR_ARG;
  for (int i = 0; i < 5; ++i) seconds[i] = s->host_s[i];
  if (builds) *builds = s->builds;
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
