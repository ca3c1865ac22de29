"""Continuous-batching serving of independent HMC chains (C ABI mag_server_*).

Server wraps the library's scheduler; poisson_stream() draws the request mix
of the serving workload (512 chains, Poisson arrivals at 200/s, volumes
{4,6,8,12,16,24}^4, Wilson beta in [5.7, 6.2] or Iwasaki beta in [2.3, 2.7],
0-3 stout levels, 5-50 trajectories, own seed each); drive() submits requests
as their arrival times pass and steps the server, returning per-trajectory
completion times for latency statistics.
"""
import ctypes as C
import time

import numpy as np

from . import _lib
from ._lib import F32, F64, IWASAKI, WILSON, Action, HmcParams, ServeRequest, TrajResult

STATE = {0: "pending", 1: "running", 2: "done", 3: "failed", 4: "cancelled"}


def _check(code, ctx):
    if code != 0:
        m = _lib.lib().mag_last_error(ctx)
        raise _lib.MagError(code, m.decode() if m else "")


def request(dims, act_type=WILSON, beta=6.0, c1=0.0, n_stout=0, rho=0.1, ntraj=5, n_md=10, tau=1.0, seed=1,
            precision=F32, links_seed=None, accept_mode=0, demon_emax=0.0, num_forced_accept=0, recon12=False):
    a = Action(act_type, beta, c1 if act_type == IWASAKI else 0.0, 0, n_stout, rho)
    h = HmcParams(n_md, tau / n_md, 0, num_forced_accept, accept_mode, demon_emax, seed, 1, precision,
                  1 if (precision == F32 or recon12) else 0)
    return ServeRequest((C.c_int * 4)(*dims), a, h, ntraj, seed if links_seed is None else links_seed)


def footprint(req):
    b = C.c_size_t()
    _check(_lib.lib().mag_server_footprint(C.byref(req), C.byref(b)), None)
    return b.value


class Server:
    def __init__(self, ctx, hbm_budget=64 << 30, max_batch=64, graphs=True):
        self.ctx = ctx
        self._h = C.c_void_p()
        _check(_lib.lib().mag_server_create(ctx.handle, hbm_budget, max_batch, 1 if graphs else 0, C.byref(self._h)),
               ctx.handle)
        ctx.adopt(self)
        self.ntraj = {}

    def _release(self):
        if self._h:
            _lib.lib().mag_server_destroy(self._h)
        self._h = C.c_void_p()

    def close(self):
        self._release()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req):
        i = C.c_int()
        _check(_lib.lib().mag_server_submit(self._h, C.byref(req), C.byref(i)), self.ctx.handle)
        self.ntraj[i.value] = req.ntraj
        return i.value

    def cancel(self, cid):
        _check(_lib.lib().mag_server_cancel(self._h, cid), self.ctx.handle)

    def step(self):
        a, p = C.c_int(), C.c_int()
        _check(_lib.lib().mag_server_step(self._h, C.byref(a), C.byref(p)), self.ctx.handle)
        return a.value, p.value

    def status(self, cid):
        st, n = C.c_int(), C.c_int()
        _check(_lib.lib().mag_server_status(self._h, cid, C.byref(st), C.byref(n)), self.ctx.handle)
        return STATE[st.value], n.value

    def results(self, cid):
        n = self.ntraj[cid]
        out = (TrajResult * max(n, 1))()
        t = (C.c_double * max(n, 1))()
        _check(_lib.lib().mag_server_results(self._h, cid, out, t, n), self.ctx.handle)
        done = self.status(cid)[1]
        return [out[i] for i in range(done)], [t[i] for i in range(done)]

    def info(self):
        u, n, e = C.c_size_t(), C.c_long(), C.c_double()
        _check(_lib.lib().mag_server_info(self._h, C.byref(u), C.byref(n), C.byref(e)), self.ctx.handle)
        return u.value, n.value, e.value

    def host_times(self):
        """Host seconds in admission, batch builds, launches, waiting, accept/retire; and the build count."""
        t, n = (C.c_double * 5)(), C.c_long()
        _check(_lib.lib().mag_server_host_times(self._h, t, C.byref(n)), self.ctx.handle)
        return dict(zip(("admit", "build", "launch", "wait", "complete"), list(t))), n.value

    def run_until_idle(self, max_steps=100000):
        for _ in range(max_steps):
            a, p = self.step()
            if a == 0 and p == 0:
                return
        raise RuntimeError("server did not drain")


def poisson_stream(nchains=512, rate=200.0, seed=2026, volumes=(4, 6, 8, 12, 16, 24), traj_range=(5, 50),
                   precision=F32, n_md=10):
    """The serving workload's request mix: (arrival_s, request) pairs."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, size=nchains))
    out = []
    for i in range(nchains):
        L = int(rng.choice(volumes))
        wilson = rng.random() < 0.5
        beta = rng.uniform(5.7, 6.2) if wilson else rng.uniform(2.3, 2.7)
        ns = int(rng.integers(0, 4))
        ntraj = int(rng.integers(traj_range[0], traj_range[1] + 1))
        req = request((L,) * 4, WILSON if wilson else IWASAKI, beta, -0.331, ns, 0.1, ntraj, n_md=n_md,
                      seed=10_000 + i, precision=precision, num_forced_accept=ntraj)
        out.append((float(t[i]), req))
    return out


def drive(server, stream, time_scale=1.0, log=None, log_every=20.0):
    """Submit each request once its arrival time (scaled) has passed, stepping
    the server meanwhile; `log` (if given) receives a progress line every
    `log_every` seconds. Returns {id: (arrival_s, t_done list, results)}."""
    t0 = time.perf_counter()
    info0 = server.info()[2]
    ids, k = {}, 0
    server.extent = {}
    next_log = log_every
    while True:
        now = time.perf_counter() - t0
        if log is not None and now >= next_log:
            u, ntr, _ = server.info()
            log(f"{now:.0f} s: {k}/{len(stream)} submitted, {ntr} trajectories done, {u / 2**30:.1f} GiB in use")
            next_log += log_every
        while k < len(stream) and stream[k][0] * time_scale <= now:
            cid = server.submit(stream[k][1])
            ids[cid] = stream[k][0] * time_scale
            server.extent[cid] = stream[k][1].dims[0]
            k += 1
        a, p = server.step()
        if a == 0 and p == 0:
            if k == len(stream):
                break
            time.sleep(max(0.0, stream[k][0] * time_scale - (time.perf_counter() - t0)))
    out = {}
    for cid, arr in ids.items():
        res, t = server.results(cid)
        out[cid] = (arr, [x - info0 for x in t], res)
    return out, time.perf_counter() - t0


def latency_stats(out, extent=None):
    """Per-trajectory latency = gap between a chain's successive completions
    (the first measured from its arrival); chain latency = arrival -> done.
    With extent (chain id -> lattice extent), also the p50/p99 per-trajectory
    latency of each extent."""
    gaps, chain_lat, ntraj = [], [], 0
    by = {}
    for cid, (arr, t, res) in out.items():
        prev = arr
        for x in t:
            gaps.append(x - prev)
            if extent is not None:
                by.setdefault(extent[cid], []).append(x - prev)
            prev = x
        if t:
            chain_lat.append(t[-1] - arr)
        ntraj += len(t)
    g = np.array(gaps)
    st = {"trajectories": ntraj, "p50_traj_latency_ms": float(np.percentile(g, 50) * 1e3),
          "p99_traj_latency_ms": float(np.percentile(g, 99) * 1e3),
          "p50_chain_latency_s": float(np.percentile(chain_lat, 50)),
          "p99_chain_latency_s": float(np.percentile(chain_lat, 99))}
    if extent is not None:
        st["traj_latency_ms_by_extent"] = {
            f"{L}^4": [round(float(np.percentile(v, 50) * 1e3), 1), round(float(np.percentile(v, 99) * 1e3), 1)]
            for L, v in sorted(by.items())}
    return st
