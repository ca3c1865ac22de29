"""Local Metropolis (algGCtrnHeatBath, alg_gchb.h:107-387) on the device
against the oracle: the reference order (one lane, the same decisions) and
the checkerboard order (whose classes commute, so the device result does not
depend on thread order), with and without constrained pair updates; at 16^4
the strong-coupling plaquette beta/18 + beta^2/216, and at beta = 5.7 the
equilibrium plaquette of the device HMC chain."""
import numpy as np
import pytest

import mag_amd as M
from mag_amd import F64, LINK18

pytestmark = pytest.mark.gpu


def _run(ctx, orc, L, U0, mag, order, iters, nhits=2, beta=5.7, small=0.5, seed=7):
    f = M.Field(ctx, L, LINK18, F64).upload(U0)
    acc, p, cp = M.local_metropolis(f, beta, mag, iters, nhits, small, seed, order)
    V = U0.copy()
    oacc, op, ocp = orc.local_metropolis(L, V, beta, mag, iters, nhits, small, seed, order)
    return f.download(), acc, p, cp, V, oacc, op, ocp


@pytest.mark.parametrize("mag", [0, 2])
def test_reference_order_matches_oracle(ctx, orc, mag):
    L = (4, 4, 4, 4)
    U, acc, p, cp, V, oacc, op, ocp = _run(ctx, orc, L, orc.haar_links(L, 11 + mag), mag, M.ORDER_REFERENCE, 2)
    assert acc == list(oacc)
    assert np.abs(U - V).max() < 1e-12
    assert np.abs(p - op).max() < 1e-13 and np.abs(cp - ocp).max() < 1e-13


@pytest.mark.parametrize("mag", [0, 2])
def test_checkerboard_matches_oracle(ctx, orc, mag):
    L = (8, 8, 8, 8)
    U, acc, p, cp, V, oacc, op, ocp = _run(ctx, orc, L, orc.haar_links(L, 21 + mag), mag, M.ORDER_CHECKERBOARD, 3)
    assert acc == list(oacc)
    assert np.abs(U - V).max() < 1e-12
    assert np.abs(p - op).max() < 1e-13 and np.abs(cp - ocp).max() < 1e-13
    assert (acc[2] + acc[3] > 0) == (mag > 0)


def test_class_api(ctx):
    f = M.Field(ctx, (8, 8, 8, 8), LINK18, F64).haar(5)
    alg = M.AlgGCtrnHeatBath(f, 5.7, M.GchbArg(mag=2, numIter=2, nHits=3, small=0.5), seed=9)
    rows = alg.run()
    assert [r[0] for r in rows] == [0, 1]
    n, c, t = alg.accept_rate()
    assert 0 < n < 1 and 0 < c < 1 and 0 < t < 1
    with pytest.raises(M.MagError):
        M.local_metropolis(M.Field(ctx, (8, 8, 8, 9), LINK18, F64).unit(), 5.7, 0, 1, 1, 0.5, 1)
    with pytest.raises(M.MagError):  # the one-lane reference order is for small lattices
        M.local_metropolis(M.Field(ctx, (16, 16, 16, 32), LINK18, F64).unit(), 5.7, 0, 1, 1, 0.5, 1,
                           M.ORDER_REFERENCE)


def test_strong_coupling_16(ctx):
    """SU(3) strong coupling: <Re Tr U_p / 3> = beta/18 + beta^2/216 + O(beta^3)."""
    beta = 0.3
    f = M.Field(ctx, (16, 16, 16, 16), LINK18, F64).haar(31)
    _, p, _ = M.local_metropolis(f, beta, 0, 30, 4, 1.0, 32)
    ref = beta / 18 + beta ** 2 / 216
    assert abs(p[10:].mean() - ref) < 6e-4, (p[10:].mean(), ref)


def test_equilibrium_matches_hmc_8(ctx):
    """beta = 5.7 on 8^4: from the FP64 HMC chain's thermalised links the
    checkerboard local Metropolis keeps the chain's equilibrium plaquette."""
    L = (8, 8, 8, 8)
    arg = M.Arg_chmc(beta=5.7, trajectory_length=10, dt=0.1, num_forced_accept_step=10)
    ch = M.Chain(ctx, L, arg, seed=41)
    ch.haar(41)
    ch.run(40)
    hp = np.array([r.avg_plaq for r in ch.run(40)])
    f = M.Field(ctx, L, LINK18, F64).upload(ch.get_links())
    _, p, _ = M.local_metropolis(f, 5.7, 0, 40, 4, 0.5, 42)
    assert abs(p.mean() - hp.mean()) < 0.006, (p.mean(), hp.mean())
