"""The 32^4 mixed-precision workload as a parity case: Iwasaki (c1 = -0.331,
beta = 2.6) + 3-stout (rho = 0.1), FP32 12-real links, force-gradient
integrator with tau = 1 and 12 steps, FP64 Hamiltonian and accept with the
demon window held at +-0.5 % of the initial action. 100 trajectories from a
start thermalised by 30 forced-accept trajectories (Haar seed 33), with an
FP64 reversibility round trip (forward trajectory,
momentum flip, forward again, from the FP32 state widened to FP64 and
projected onto SU(3)) after
every 10: links back to 1e-10, Hamiltonian back to 1e-6 of a total of ~3e7.
Over the run: acceptance, and <exp(-dH)> = 1 within four standard errors."""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import F32

pytestmark = pytest.mark.gpu

L32 = (32, 32, 32, 32)


def test_c3_32x4_100_trajectories_with_reversibility(ctx):
    kw = dict(trajectory_length=12, dt=1.0 / 12, beta=2.6, gauge=M.Gauge(M.IWASAKI, -0.331), n_stout=3, rho=0.1)
    therm = M.Chain(ctx, L32, M.Arg_chmc(num_forced_accept_step=30, **kw), seed=33, precision=F32)
    therm.haar(33)
    therm.run(30)
    ch = M.Chain(ctx, L32, M.Arg_chmc(num_forced_accept_step=0, **kw), seed=34, precision=F32, accept_mode=2,
                 demon_emax=-0.005)
    ch.links().copy_from(therm.links())
    ch.reset_demon()
    ch.demon_state()
    res, rev = [], []
    for block in range(10):
        res += ch.run(10)
        dev, dh_rt = ch.reversibility()
        rev.append((dev, dh_rt))
        assert dev < 1e-10, (block, dev)
        assert abs(dh_rt) < 1e-6, (block, dh_rt)
    dh = np.array([r.dh for r in res])
    acc = np.array([r.accept for r in res], dtype=float)
    info = f"acceptance {acc.mean():.2f}, dH {dh.round(3).tolist()}, reversibility {rev}"
    assert np.all(np.isfinite(dh)), info
    assert acc.mean() > 0.5, info
    w = np.exp(-dh)
    assert abs(w.mean() - 1.0) < 4 * w.std(ddof=1) / np.sqrt(len(w)) + 1e-3, info
