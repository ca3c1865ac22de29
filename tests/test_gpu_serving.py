"""Serving: continuous batching of independent chains on one device.

A chain's per-trajectory results must be bitwise identical alone, in a mixed
batch, with and without graph replay, and through the single-chain API: its
kernels and reductions never depend on what it is batched with.
"""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import serving as S

pytestmark = pytest.mark.gpu


def mix():
    return [
        S.request((8, 8, 8, 8), M.WILSON, 5.9, ntraj=3, n_md=4, seed=11, precision=M.F32),
        S.request((8, 8, 8, 8), M.WILSON, 6.1, ntraj=4, n_md=4, seed=12, precision=M.F32),
        S.request((4, 4, 4, 4), M.IWASAKI, 2.5, -0.331, n_stout=1, ntraj=3, n_md=4, seed=13, precision=M.F32),
        S.request((4, 4, 4, 4), M.IWASAKI, 2.7, -0.331, n_stout=1, ntraj=2, n_md=4, seed=14, precision=M.F32),
        S.request((4, 4, 4, 8), M.IWASAKI, 2.6, -0.331, n_stout=2, ntraj=2, n_md=3, seed=15, precision=M.F64),
        S.request((8, 8, 8, 8), M.WILSON, 6.0, ntraj=3, n_md=4, seed=16, precision=M.F32),
    ]


def key(res):
    return [(r.h_old, r.h_new, r.accept, r.avg_plaq) for r in res]


def run_alone(ctx, req, graphs=True):
    srv = S.Server(ctx, graphs=graphs)
    cid = srv.submit(req)
    srv.run_until_idle()
    res, _ = srv.results(cid)
    srv.close()
    return key(res)


def test_alone_equals_mix_and_graph_equals_eager(ctx):
    reqs = mix()
    alone = [run_alone(ctx, r) for r in reqs]
    for graphs in (True, False):
        srv = S.Server(ctx, graphs=graphs, max_batch=2)
        ids = [srv.submit(r) for r in reqs]
        srv.run_until_idle()
        for cid, a in zip(ids, alone):
            res, t = srv.results(cid)
            assert srv.status(cid) == ("done", len(a))
            assert key(res) == a
            assert all(np.diff(t) >= 0)
        srv.close()


def test_c4_shaped_mix_alone_equals_mix(ctx):
    """Chains of the serving workload's larger shapes (12^4 to 24^4, Wilson and
    Iwasaki, 0-3 stout levels, FP32 and FP64, staggered arrivals): every
    chain's results are bitwise the same alone and in the mix."""
    reqs = [
        S.request((24, 24, 24, 24), M.IWASAKI, 2.55, -0.331, n_stout=3, ntraj=2, n_md=4, seed=51, precision=M.F32),
        S.request((12, 12, 12, 12), M.WILSON, 5.8, n_stout=0, ntraj=3, n_md=4, seed=52, precision=M.F32),
        S.request((16, 16, 16, 16), M.IWASAKI, 2.4, -0.331, n_stout=2, ntraj=2, n_md=4, seed=53, precision=M.F32),
        S.request((24, 24, 24, 24), M.WILSON, 6.1, n_stout=1, ntraj=2, n_md=4, seed=54, precision=M.F64),
        S.request((16, 16, 16, 16), M.IWASAKI, 2.7, -0.331, n_stout=3, ntraj=2, n_md=4, seed=55, precision=M.F64),
        S.request((12, 12, 12, 12), M.IWASAKI, 2.3, -0.331, n_stout=1, ntraj=3, n_md=4, seed=56, precision=M.F32),
        S.request((16, 16, 16, 16), M.IWASAKI, 2.45, -0.331, n_stout=2, ntraj=2, n_md=4, seed=57, precision=M.F32),
    ]
    alone = [run_alone(ctx, r) for r in reqs]
    srv = S.Server(ctx, hbm_budget=64 << 30)
    ids = [srv.submit(r) for r in reqs[:3]]
    srv.step()
    ids += [srv.submit(r) for r in reqs[3:5]]
    srv.step()
    ids += [srv.submit(r) for r in reqs[5:]]
    srv.run_until_idle()
    for cid, a, r in zip(ids, alone, reqs):
        res, _ = srv.results(cid)
        assert srv.status(cid) == ("done", r.ntraj)
        assert key(res) == a
        assert all(np.isfinite(x.h_new) for x in res)
    srv.close()


def test_server_matches_single_chain_api(ctx):
    req = S.request((8, 8, 8, 8), M.WILSON, 6.0, ntraj=3, n_md=5, seed=21, precision=M.F64)
    served = run_alone(ctx, req)
    arg = M.Arg_chmc(trajectory_length=5, dt=0.2, beta=6.0)
    ch = M.Chain(ctx, (8, 8, 8, 8), arg, seed=21, precision=M.F64)
    ch.haar(21)
    assert key(ch.run(3)) == served


def test_late_arrivals_join_running_batches(ctx):
    """Requests submitted while others run are admitted at the next step and
    give the same results as alone."""
    reqs = mix()
    alone = [run_alone(ctx, r) for r in reqs]
    srv = S.Server(ctx)
    ids = [srv.submit(reqs[0])]
    srv.step()
    ids.append(srv.submit(reqs[1]))
    srv.step()
    ids += [srv.submit(r) for r in reqs[2:]]
    srv.run_until_idle()
    for cid, a in zip(ids, alone):
        assert key(srv.results(cid)[0]) == a


def test_admission_by_hbm_footprint(ctx):
    reqs = [S.request((8, 8, 8, 8), M.WILSON, 6.0, ntraj=2, n_md=3, seed=30 + i, precision=M.F32) for i in range(5)]
    fp = S.footprint(reqs[0])
    srv = S.Server(ctx, hbm_budget=2 * fp + fp // 2)
    ids = [srv.submit(r) for r in reqs]
    peak = 0
    for _ in range(100):
        a, p = srv.step()
        peak = max(peak, a)
        used = srv.info()[0]
        assert used <= 2 * fp + fp // 2
        if a == 0 and p == 0:
            break
    assert peak == 2
    assert all(srv.status(i) == ("done", 2) for i in ids)
    with pytest.raises(M.MagError):
        S.Server(ctx, hbm_budget=fp // 2).submit(reqs[0])


def test_cancel(ctx):
    srv = S.Server(ctx)
    a = srv.submit(S.request((4, 4, 4, 4), M.WILSON, 6.0, ntraj=10, n_md=2, seed=41))
    b = srv.submit(S.request((4, 4, 4, 4), M.WILSON, 6.0, ntraj=3, n_md=2, seed=42))
    srv.step()
    srv.cancel(a)
    srv.run_until_idle()
    assert srv.status(a)[0] == "cancelled" and srv.status(a)[1] <= 2
    assert srv.status(b) == ("done", 3)


def test_poisson_stream_small(ctx):
    """A scaled-down serving stream: every chain completes with finite results."""
    stream = S.poisson_stream(nchains=24, rate=400.0, seed=3, volumes=(4, 6, 8), traj_range=(2, 4), n_md=4)
    srv = S.Server(ctx)
    out, wall = S.drive(srv, stream)
    st = S.latency_stats(out)
    assert st["trajectories"] == sum(r.ntraj for _, r in stream)
    for arr, t, res in out.values():
        assert all(np.isfinite(x.h_new) for x in res)
    assert st["p99_traj_latency_ms"] >= st["p50_traj_latency_ms"] > 0


@pytest.mark.parametrize("prec,recon12,n_stout", [(M.F32, False, 0), (M.F32, False, 3), (M.F64, False, 1),
                                                  (M.F64, False, 3), (M.F64, True, 0), (M.F64, True, 3)])
def test_footprint_covers_chain_memory(prec, recon12, n_stout):
    """Admission charges mag_server_footprint per chain: it must cover what a
    chain of that class allocates once every workspace is in use (one
    trajectory run), and not overstate it by much."""
    L = (8, 8, 8, 8)
    req = S.request(L, M.IWASAKI, 2.6, -0.331, n_stout=n_stout, ntraj=1, n_md=2, seed=5, precision=prec,
                    recon12=recon12)
    fp = S.footprint(req)
    c = M.Context(0)
    try:
        m0 = c.memory()
        arg = M.Arg_chmc(trajectory_length=2, dt=0.5, beta=2.6, gauge=M.Gauge(1, -0.331), n_stout=n_stout, rho=0.1)
        ch = M.Chain(c, L, arg, seed=5, precision=prec, recon12=recon12)
        ch.haar(5)
        ch.run(1)
        used = c.memory() - m0
        assert used <= fp, (used, fp)
        assert fp <= 1.2 * used, (used, fp)
    finally:
        c.close()
