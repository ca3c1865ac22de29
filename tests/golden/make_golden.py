"""Writes the golden fixtures under tests/golden/.

The reference (hummingtree/mag) holds no test vectors, so the fixtures are
produced here: expiQ inputs with scipy.linalg.expm outputs (an independent
library routine), and the FP64 CPU oracle's outputs on seeded configurations
(Philox Haar links / momenta, regenerable from the seeds). They pin the
oracle against regressions and give the GPU tests stored expectations at
the configurations of the headline runs (C1: 8^4 Wilson beta = 6.0 from Haar
seed 1234, one FG trajectory tau = 1, 8 steps).

Run:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import pyoracle as O  # noqa: E402


def expiq_vectors():
    rng = np.random.default_rng(20261202)
    T = O.generators()
    scales = np.repeat([1e-6, 1e-3, 0.05, 0.5, 2.0], 40)
    q = rng.normal(size=(len(scales), 8)) * scales[:, None]
    Q = np.einsum("na,aij->nij", q, T)
    E = np.array([sl.expm(1j * m) for m in Q])
    np.savez_compressed(os.path.join(HERE, "expiq_scipy.npz"), Q=Q, expm=E)


def c1_trajectory():
    L = (8, 8, 8, 8)
    a = O.action(0, 6.0)
    U = O.haar_links(L, 1234)
    P = O.init_momentum(L, 1234, 0)
    F = O.get_force(L, U, a)
    h0, parts0 = O.hamiltonian(L, U, P, a)
    dev = O.force_gradient_integrator(L, U, P, a, 8, 0.125, reunit=True)
    h1, parts1 = O.hamiltonian(L, U, P, a)
    np.savez_compressed(
        os.path.join(HERE, "c1_wilson_8x4.npz"),
        links_checksum=np.array([O.haar_links(L, 1234).sum(), (O.haar_links(L, 1234) ** 2).sum()]),
        total_plaq=O.total_plaq(L, O.haar_links(L, 1234)),
        total_rect=O.total_rectangular(L, O.haar_links(L, 1234)),
        force_site0=F[0], force_norm2=(F ** 2).sum(),
        h_old=h0, parts_old=np.array(parts0), h_new=h1, parts_new=np.array(parts1), reunit_dev=dev,
        links_after_site0=U[0], links_after_norm2=(U ** 2).sum(),
        wloop_221=O.avg_wilson_loop(L, O.haar_links(L, 1234), (2, 2, 1), 2))


def stout_force():
    L = (4, 4, 4, 4)
    a = O.action(1, 2.6, -0.331, n_stout=3, rho=0.1)
    U = O.haar_links(L, 7)
    F = O.get_force(L, U, a)
    S3 = O.stout_smear(L, U, 0.1, 3)
    np.savez_compressed(os.path.join(HERE, "iwasaki_stout3_4x4.npz"), force=F, smeared_site5=S3[5],
                        action=O.gauge_action(L, U, a))


def accept_sequence():
    L = (4, 4, 4, 4)
    a = O.action(0, 5.7)
    U = O.haar_links(L, 1234)
    h = O.hmc(trajectory_length=3, dt=1 / 3, num_forced_accept=3, seed=99)
    _, res, _ = O.run_hmc(L, U, a, h, 0, 20)
    np.savez_compressed(os.path.join(HERE, "accept_seq_wilson_4x4.npz"),
                        accept=np.array([r.accept for r in res]), h_old=np.array([r.h_old for r in res]),
                        dh=np.array([r.dh for r in res]), die=np.array([r.die_roll for r in res]),
                        plaq=np.array([r.avg_plaq for r in res]))


if __name__ == "__main__":
    expiq_vectors()
    c1_trajectory()
    stout_force()
    accept_sequence()
    print("golden fixtures written to", HERE)
