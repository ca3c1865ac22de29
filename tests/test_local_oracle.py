"""Pins of the oracle's local Metropolis (algGCtrnHeatBath, alg_gchb.h:107-387).

  * proposals are SU(3) (unitary, det 1);
  * with a fixed staple sum sigma = k 1 the link samples exp((beta/3) k Re Tr U):
    its mean Re Tr U matches the SU(3) Weyl integral;
  * the hit bookkeeping follows the link classes (pairs for the constrained
    links below the last one of a coarse link, single updates otherwise);
  * both update orders (reference order and the 16 checkerboard classes),
    with and without constrained pair updates, reach the equilibrium
    plaquette of an independent algorithm (the force-gradient HMC chain)
    and the strong-coupling value beta/18 at small beta.
"""
import numpy as np
import pytest

from test_multi_oracle import _weyl_mean_retr

L4 = (4, 4, 4, 4)


def test_proposals_are_su3(orc):
    U = np.zeros((1, 1, 18))
    U[0, 0, [0, 8, 16]] = 1.0
    S = np.zeros((18,))
    u = U.reshape(18).copy()
    for it in range(50):
        orc.metropolis_link(u, S, 0.0, 4, 0.7, 11, it, 3)  # beta = 0: every hit accepted
    m = orc.as_cplx(u.reshape(1, 1, 18))[0, 0]
    assert np.abs(m @ m.conj().T - np.eye(3)).max() < 1e-12
    assert abs(np.linalg.det(m) - 1) < 1e-12
    assert np.abs(m - np.eye(3)).max() > 0.3


@pytest.mark.parametrize("beta", [0.5, 1.5])
def test_single_link_law(orc, beta):
    k = 6.0  # staple sum of a unit-link neighbourhood
    S = np.zeros(18)
    S[[0, 8, 16]] = k
    u = np.zeros(18)
    u[[0, 8, 16]] = 1.0
    acc = None
    vals = []
    for it in range(6000):
        acc = orc.metropolis_link(u, S, beta, 2, 0.8, 21, it, 0, acc)
        if it >= 500:
            vals.append(u[0] + u[8] + u[16])
    ref = _weyl_mean_retr(beta / 3.0 * k)
    assert abs(np.mean(vals) - ref) < 0.05, (np.mean(vals), ref)
    assert 0.2 < acc[0] / (acc[0] + acc[1]) < 0.95


def _classes(L, mag):
    single = pair = 0
    for x3 in range(L[3]):
        for x2 in range(L[2]):
            for x1 in range(L[1]):
                for x0 in range(L[0]):
                    x = (x0, x1, x2, x3)
                    for mu in range(4):
                        cons = mag > 0 and all(x[i] % mag == 0 for i in range(4) if i != mu)
                        if cons and x[mu] % mag != mag - 1:
                            pair += 1
                        else:
                            single += 1
    return single, pair


@pytest.mark.parametrize("mag,order", [(0, 0), (0, 1), (2, 0), (2, 1)])
def test_hit_bookkeeping(orc, mag, order):
    U = orc.haar_links(L4, 31)
    acc, plaq, cons = orc.local_metropolis(L4, U, 5.7, mag, 2, 3, 0.5, 32, order)
    single, pair = _classes(L4, mag)
    assert acc[0] + acc[1] == 2 * 3 * single
    assert acc[2] + acc[3] == 2 * 3 * pair
    assert abs(plaq[-1] - orc.avg_plaquette(L4, U)) < 1e-14
    if mag:
        assert abs(cons[-1] - orc.constrained_plaquette(L4, U, mag)) < 1e-14


def test_orders_differ_but_are_deterministic(orc):
    U0 = orc.haar_links(L4, 41)
    a, b, c = U0.copy(), U0.copy(), U0.copy()
    orc.local_metropolis(L4, a, 5.7, 0, 1, 2, 0.5, 42, 0)
    orc.local_metropolis(L4, b, 5.7, 0, 1, 2, 0.5, 42, 1)
    orc.local_metropolis(L4, c, 5.7, 0, 1, 2, 0.5, 42, 1)
    assert np.array_equal(b, c)
    assert np.abs(a - b).max() > 1e-3


def _hmc_chain(orc, L, beta, ntraj, seed):
    U = orc.haar_links(L, seed)
    h = orc.hmc(trajectory_length=10, dt=0.1, num_forced_accept=5, seed=seed + 1)
    act = orc.action(0, beta)
    _, res, _ = orc.run_hmc(L, U, act, h, 0, ntraj)
    return U, np.array([r.avg_plaq for r in res])


@pytest.mark.slow
def test_equilibrium_plaquette_matches_hmc(orc):
    """beta = 5.7 on 4^4: started from the HMC chain's equilibrated links,
    local Metropolis (both orders, mag 0 and 2) keeps the HMC chain's
    equilibrium plaquette: means over 80 sweeps against the mean over the
    second half of 100 HMC trajectories, tolerance 0.01 (about five
    combined standard errors; a biased update drifts away by more)."""
    beta = 5.7
    Ueq, hp = _hmc_chain(orc, L4, beta, 100, 51)
    hmc = hp[50:].mean()
    for mag, order in [(0, 0), (0, 1), (2, 0), (2, 1)]:
        U = Ueq.copy()
        _, plaq, _ = orc.local_metropolis(L4, U, beta, mag, 80, 4, 0.5, 53 + order + 10 * mag, order)
        assert abs(plaq.mean() - hmc) < 0.01, (mag, order, plaq.mean(), hmc)


def test_strong_coupling(orc):
    """<Re Tr U_p / 3> = beta/18 + O(beta^2) for SU(3); at beta = 0.3 the
    deviation is below 1e-3."""
    beta = 0.3
    U = orc.haar_links(L4, 61)
    _, plaq, _ = orc.local_metropolis(L4, U, beta, 0, 40, 4, 1.0, 62, 1)
    assert abs(plaq[10:].mean() - beta / 18) < 0.006, plaq[10:].mean()


# ------------------------------------------------------ checkerboard demon
def test_checkerboard_demon_conservation(orc):
    """Per-link demons: sum S_k + E_k conserved for the plaquette and
    rectangle sums, every demon inside its window."""
    L = (4, 4, 4, 8)
    U = orc.haar_links(L, 131)
    s0 = (-orc.total_plaq(L, U), -orc.total_rectangular(L, U))
    E, (acc, rej) = orc.demon_checkerboard(L, U, 2, 1.5, 132)
    s1 = (-orc.total_plaq(L, U), -orc.total_rectangular(L, U))
    assert acc > 0 and rej > 0
    assert abs(s1[0] + E[:, :, 0].sum() - s0[0]) < 1e-9
    assert abs(s1[1] + E[:, :, 1].sum() - s0[1]) < 1e-9
    assert np.abs(E).max() < 1.5
    # a second call continues from the demons it is given
    E2, _ = orc.demon_checkerboard(L, U, 1, 1.5, 133, E.copy())
    s2 = (-orc.total_plaq(L, U), -orc.total_rectangular(L, U))
    assert abs(s2[0] + E2[:, :, 0].sum() - s0[0]) < 1e-9


def test_checkerboard_demon_unbounded_window_equals_sequential(orc):
    """With an unbounded window every proposal is accepted, so each link ends
    as m2 m1 U whatever the visiting order: the checkerboard and the
    sequential demon give the same links, and the same plaquette and
    rectangle demon totals (minus the action change)."""
    L = (4, 4, 4, 4)
    U0 = orc.haar_links(L, 141)
    a, b = U0.copy(), U0.copy()
    E, (acc, rej) = orc.demon_checkerboard(L, a, 1, 1e300, 142)
    Es, (acc_s, rej_s) = orc.demon_microcanonical(L, b, 1, 1e300, 142)
    assert rej == 0 and rej_s == 0 and acc == acc_s == 2 * 4 * 256
    assert np.abs(a - b).max() < 1e-13
    assert abs(E[:, :, 0].sum() - Es[0]) < 1e-9 and abs(E[:, :, 1].sum() - Es[1]) < 1e-9
