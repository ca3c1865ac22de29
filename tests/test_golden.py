"""The oracle reproduces the committed golden fixtures (tests/golden/, written
by tests/golden/make_golden.py): scipy's expm vectors and the oracle's own
outputs on the seeded headline configurations."""
import os

import numpy as np

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


def test_expiq_matches_scipy_fixture(orc):
    d = load("expiq_scipy.npz")
    assert np.abs(orc.expiQ(d["Q"]) - d["expm"]).max() < 1e-14


def test_c1_fixture(orc):
    d = load("c1_wilson_8x4.npz")
    L = (8, 8, 8, 8)
    a = orc.action(0, 6.0)
    U = orc.haar_links(L, 1234)
    assert np.allclose([U.sum(), (U ** 2).sum()], d["links_checksum"], rtol=1e-14, atol=1e-12)
    assert abs(orc.total_plaq(L, U) - d["total_plaq"]) < 1e-12 * abs(d["total_plaq"])
    assert abs(orc.total_rectangular(L, U) - d["total_rect"]) < 1e-12 * abs(d["total_rect"])
    F = orc.get_force(L, U, a)
    assert np.abs(F[0] - d["force_site0"]).max() < 1e-13
    P = orc.init_momentum(L, 1234, 0)
    h0, _ = orc.hamiltonian(L, U, P, a)
    assert abs(h0 - d["h_old"]) < 1e-12 * abs(d["h_old"])
    orc.force_gradient_integrator(L, U, P, a, 8, 0.125, reunit=True)
    h1, _ = orc.hamiltonian(L, U, P, a)
    assert abs(h1 - d["h_new"]) < 1e-10 * abs(d["h_new"])
    assert np.abs(U[0] - d["links_after_site0"]).max() < 1e-10


def test_stout_fixture(orc):
    d = load("iwasaki_stout3_4x4.npz")
    L = (4, 4, 4, 4)
    a = orc.action(1, 2.6, -0.331, n_stout=3, rho=0.1)
    U = orc.haar_links(L, 7)
    assert np.abs(orc.get_force(L, U, a) - d["force"]).max() < 1e-12 * np.abs(d["force"]).max()
    assert abs(orc.gauge_action(L, U, a) - d["action"]) < 1e-12 * abs(d["action"])


def test_accept_sequence_fixture(orc):
    d = load("accept_seq_wilson_4x4.npz")
    L = (4, 4, 4, 4)
    U = orc.haar_links(L, 1234)
    h = orc.hmc(trajectory_length=3, dt=1 / 3, num_forced_accept=3, seed=99)
    _, res, _ = orc.run_hmc(L, U, orc.action(0, 5.7), h, 0, 20)
    assert [r.accept for r in res] == list(d["accept"])
    assert 0 < sum(d["accept"][3:]) < 17  # both outcomes occur
    assert np.allclose([r.dh for r in res], d["dh"], rtol=1e-6, atol=1e-9)
    assert [r.die_roll for r in res] == list(d["die"])
